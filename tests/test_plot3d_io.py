"""PLOT3D, control-file and bc.dat formats (host code of the library; no GPU).

Pinned by the files the reference's examples hold (tests/golden): a solution file
written by magudi is read and written back byte for byte, as is a control-forcing file;
byte-swapped and multi-block variants are built from them with numpy."""
import os

import numpy as np
import pytest

from magudi_amd import cases, plot3d

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
QFILE = os.path.join(GOLD, "TestForcing.ic.q")
DAT = os.path.join(GOLD, "TestForcing.baseline.dat")
BC = os.path.join(GOLD, "AcousticMonopole.bc.dat")


def test_reference_solution_file_roundtrip(tmp_path):
    info = plot3d.detect(QFILE)
    assert info["n_grids"] == 1 and info["file_type"] == plot3d.SOLUTION
    assert info["sizes"][0] == (4, 4, 1) and info["native_endian"]
    q, aux = plot3d.read_solution(QFILE, nv=4)
    raw = open(QFILE, "rb").read()
    # record layout: [4|ngrid|4] [12|sizes|12] [32|aux|32] [640|5 x 16 doubles|640]
    head = np.frombuffer(raw[:40], dtype="<i4")
    assert list(head[:8]) == [4, 1, 4, 12, 4, 4, 1, 12]
    assert np.array_equal(aux, np.frombuffer(raw[36:68], dtype="<f8"))
    data = np.frombuffer(raw[76:76 + 640], dtype="<f8").reshape(5, 16)
    assert np.array_equal(q[:3], data[:3]) and np.array_equal(q[3], data[4])
    assert np.all(data[3] == 0.0)           # 2-D state: rho w slot unused
    out = tmp_path / "copy.q"
    plot3d.write_solution(str(out), (4, 4, 1), q, timestep=int(aux[0]), time=aux[3])
    if aux[1] == 0.0 and aux[2] == 0.0:
        assert open(out, "rb").read() == raw


def test_byte_swapped_and_multiblock(tmp_path):
    q, aux = plot3d.read_solution(QFILE, nv=4)
    raw = np.frombuffer(open(QFILE, "rb").read(), dtype=np.uint8)
    # big-endian copy: swap every 4-byte marker / int and every 8-byte double in place
    be = bytearray(raw.tobytes())
    def sw(off, n):
        be[off:off + n] = be[off:off + n][::-1]
    pos = 0
    while pos < len(be):
        m = int(np.frombuffer(raw[pos:pos + 4].tobytes(), "<i4")[0])
        sw(pos, 4)
        width = 4 if m in (4, 12) else 8
        for k in range(m // width):
            sw(pos + 4 + k * width, width)
        sw(pos + 4 + m, 4)
        pos += m + 8
    p = tmp_path / "be.q"
    p.write_bytes(bytes(be))
    info = plot3d.detect(str(p))
    assert not info["native_endian"]
    q2, aux2 = plot3d.read_solution(str(p), nv=4)
    assert np.array_equal(q2, q) and np.array_equal(aux2, aux)
    # two-block grid file written by hand; read the second block
    n0, n1 = (3, 2, 1), (4, 5, 1)
    x0 = np.arange(3 * 6, dtype="<f8")
    x1 = 100.0 + np.arange(3 * 20, dtype="<f8")
    ib1 = np.arange(20, dtype="<i4")
    rec = lambda b: np.int32(len(b)).tobytes() + b + np.int32(len(b)).tobytes()
    blob = rec(np.int32(2).tobytes()) + rec(np.array(n0 + n1, dtype="<i4").tobytes())
    blob += rec(x0.tobytes()) + rec(x1.tobytes() + ib1.tobytes())
    g = tmp_path / "two.xyz"
    g.write_bytes(blob)
    info = plot3d.detect(str(g))
    assert info["n_grids"] == 2 and info["sizes"] == [n0, n1]
    xyz, ib = plot3d.read_grid(str(g), block=1)
    assert np.array_equal(xyz.ravel(), x1) and np.array_equal(ib, ib1)


def test_grid_and_function_roundtrip(tmp_path):
    c = cases.acoustic_monopole(n=33, nsteps=1)
    ib = np.ones(c.npoints, dtype=np.intc)
    ib[5] = 0
    plot3d.write_grid(str(tmp_path / "g.xyz"), c.n, c.coords, ib)
    info = plot3d.detect(str(tmp_path / "g.xyz"))
    assert info["file_type"] == plot3d.GRID and info["has_iblank"]
    xyz, ib2 = plot3d.read_grid(str(tmp_path / "g.xyz"))
    assert np.array_equal(xyz[:2], c.coords) and np.all(xyz[2] == 0) and np.array_equal(ib2, ib)
    f = np.random.default_rng(0).standard_normal((2, c.npoints))
    plot3d.write_function(str(tmp_path / "m.f"), c.n, f)
    assert plot3d.detect(str(tmp_path / "m.f"))["n_scalars"] == 2
    assert np.array_equal(plot3d.read_function(str(tmp_path / "m.f")), f)
    with pytest.raises(Exception):
        plot3d.read_solution(str(tmp_path / "m.f"), nv=4)


def test_control_file_order_and_zaxpy(tmp_path):
    raw = np.fromfile(DAT, dtype="<f8")
    a = plot3d.read_control(DAT, npoints=16)           # forward stage order
    assert a.shape == (32, 16)
    assert np.array_equal(a[::-1].ravel(), raw)         # file holds the latest stage first
    plot3d.write_control(str(tmp_path / "c.dat"), a)
    assert open(tmp_path / "c.dat", "rb").read() == open(DAT, "rb").read()
    y = np.random.default_rng(1).standard_normal(a.shape)
    plot3d.write_control(str(tmp_path / "y.dat"), y)
    plot3d.zaxpy(str(tmp_path / "z.dat"), -0.25, DAT, str(tmp_path / "y.dat"))
    z = plot3d.read_control(str(tmp_path / "z.dat"), 16)
    assert np.array_equal(z, -0.25 * a + y)


def test_bc_table():
    rows = plot3d.read_bc(BC, params={"sponge.E": (0.2, 2.0)})
    assert [r.name for r in rows][:2] == ["farField.E", "farField.W"]
    types = [r.type for r in rows]
    assert types.count(cases.FAR_FIELD) == 4 and types.count(cases.SPONGE) == 4
    assert types[-2:] == [cases.COST_TARGET, cases.ACTUATOR]
    t = rows[-2]
    assert t.normal_dir == 0 and t.extent == (93, 109, 29, 173, 1, -1)
    assert rows[4].params[:2] == (0.2, 2.0)


def test_eight_byte_record_markers(tmp_path):
    """The reference's USE_EXTENDED_PLOT3D build frames records with 8-byte markers
    (include/config.h.in:34-40, src/PLOT3DFormat.c:22-25): the reference's solution file
    re-framed with 8-byte markers (numpy, little and big endian) reads the same; the
    writers produce that framing after set_record_markers(8)."""
    q, aux = plot3d.read_solution(QFILE, nv=4)
    raw = open(QFILE, "rb").read()
    # re-frame every record of the 4-byte file with 8-byte markers
    recs, pos = [], 0
    while pos < len(raw):
        m = int(np.frombuffer(raw[pos:pos + 4], dtype="<i4")[0])
        recs.append(raw[pos + 4:pos + 4 + m])
        pos += 8 + m
    for order in ("<", ">"):
        out = bytearray()
        for r in recs:
            mk = np.array([len(r)], dtype=order + "i8").tobytes()
            body = r
            if order == ">":      # swap every 4- or 8-byte item of the record body
                if len(r) in (4, 12):
                    body = np.frombuffer(r, dtype="<i4").astype(">i4").tobytes()
                else:
                    body = np.frombuffer(r, dtype="<f8").astype(">f8").tobytes()
            out += mk + body + mk
        f = tmp_path / f"wide{'le' if order == '<' else 'be'}.q"
        f.write_bytes(bytes(out))
        info = plot3d.detect(str(f))
        assert info["file_type"] == plot3d.SOLUTION and info["sizes"][0] == (4, 4, 1)
        assert info["native_endian"] == (order == "<")
        q2, aux2 = plot3d.read_solution(str(f), nv=4)
        assert np.array_equal(q2, q) and np.array_equal(aux2, aux)
    try:
        plot3d.set_record_markers(8)
        w = tmp_path / "written.q"
        plot3d.write_solution(str(w), (4, 4, 1), q, timestep=int(aux[0]), time=aux[3])
        head = np.frombuffer(w.read_bytes()[:16], dtype="<i8")
        assert head[0] == 4
        q3, _ = plot3d.read_solution(str(w), nv=4)
        assert np.array_equal(q3, q)
        if aux[1] == 0.0 and aux[2] == 0.0:
            assert w.read_bytes() == (tmp_path / "widele.q").read_bytes()
    finally:
        plot3d.set_record_markers(4)
    with pytest.raises(Exception):
        plot3d.set_record_markers(6)
