"""Fortran ISO_C_BINDING host side (magudi_amd/fortran) against the C ABI.

magudi's drivers are Fortran (bin/Forward.f90, bin/Adjoint.f90); the module
magudi_hip.f90 is what such a driver uses.  CPU checks: every bind(C) name is an
entry point of include/magudi_hip.h, and the bind(C) derived types have the C
structs' sizes; the magudi.inp reader.  GPU checks: a Fortran program that builds the
monopole in code prints the same J and sensitivity as the Python host path; the forward and
adjoint drivers, run on an example directory written with the library's PLOT3D writers
(magudi.inp, bc.dat, grid, initial condition, mollifiers), write J, the sensitivity and the
gradient file of the oracle's run of the same case, in the stage loop and in the whole-run
mode, and the forward driver with controller_switch reproduces the finite-difference
relation (J(A) - J(0)) / A = Grad0; the multiple-shooting drivers, run on such a directory with
per-segment initial conditions and a control forcing file, write the oracle's msforward /
msadjoint results (costs, mismatches, end snapshots, gradient inner products, IC gradients,
matching adjoint terminals and the control gradient file).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from magudi_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "magudi_amd", "fortran")
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang"
                                  if os.path.exists("/opt/rocm/bin/amdflang") else None)


def test_bind_names_are_declared():
    src = open(os.path.join(FDIR, "magudi_hip.f90")).read()
    names = set(re.findall(r'bind\(C,\s*name="(mgd_[a-z_]+)"\)', src))
    assert names, "no bind(C) interfaces found"
    missing = names - set(capi.declared_symbols())
    assert not missing, f"bound but not declared in include/magudi_hip.h: {missing}"


@pytest.mark.skipif(FC is None, reason="amdflang not available")
def test_struct_sizes_match(tmp_path):
    import ctypes as C
    prog = tmp_path / "sizes.f90"
    prog.write_text(
        "program sizes\n  use, intrinsic :: iso_c_binding\n  use magudi_hip\n"
        "  type(mgd_region_desc) :: d\n  type(mgd_patch_desc) :: p\n"
        "  print '(I0,1X,I0)', c_sizeof(d), c_sizeof(p)\nend program sizes\n")
    mod = tmp_path / "m"
    mod.mkdir()
    subprocess.check_call([FC, "-c", os.path.join(FDIR, "magudi_hip.f90"), "-J", str(mod),
                           "-o", str(tmp_path / "m.o")], cwd=tmp_path)
    libdir = os.path.dirname(capi.LIB_PATH)
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmagudi_hip.so not built")
    subprocess.check_call([FC, str(prog), str(tmp_path / "m.o"), "-I", str(mod),
                           "-L" + libdir, "-lmagudi_hip", "-Wl,-rpath," + libdir,
                           "-o", str(tmp_path / "sizes")], cwd=tmp_path)
    out = subprocess.check_output([str(tmp_path / "sizes")]).decode().split()
    assert int(out[0]) == C.sizeof(capi.RegionDesc)
    assert int(out[1]) == C.sizeof(capi.PatchDesc)


@pytest.mark.gpu
def test_fortran_driver_matches_python():
    exe = os.path.join(FDIR, "monopole_adjoint_hip")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (needs amdflang at build time)")
    from magudi_amd import cases
    from magudi_amd.solver import Region, Solver
    out = subprocess.run([exe, "41", "16", "4", "0.2"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = dict(line.split() for line in out.stdout.splitlines() if line.strip())
    c = cases.acoustic_monopole(n=41, nsteps=16, save_interval=4)
    c.dt = 0.2
    R = Region(c)
    R.states[0].conservedVariables = c.initial_state()
    sens, _, J = Solver(R).runAdjoint()
    np.testing.assert_allclose(float(vals["J"]), J[0], rtol=1e-12)
    np.testing.assert_allclose(float(vals["Grad0"]), sens[0], rtol=1e-10)


def _write_case_directory(d, c, prefix="Monopole", extra=""):
    """An example directory of case c in magudi's formats: magudi.inp, bc.dat, the PLOT3D
    grid, initial condition and mollifier files."""
    from magudi_amd import cases, plot3d
    names = {v: k for k, v in cases.PATCH_TYPE_NAMES.items()}
    n = c.n
    plot3d.write_grid(str(d / f"{prefix}.xyz"), n, c.coords)
    plot3d.write_solution(str(d / f"{prefix}.ic.q"), n, c.initial_state())
    plot3d.write_function(str(d / f"{prefix}.target_mollifier.f"), n, c.target_mollifier[None])
    plot3d.write_function(str(d / f"{prefix}.control_mollifier.f"), n, c.control_mollifier[None])
    rows = ["# Name Type Grid normDir iMin iMax jMin jMax kMin kMax"]
    for p in c.patches:
        rows.append(f"{p.name} {names[p.type]} 1 {p.normal_dir} " + " ".join(str(v) for v in p.extent))
    (d / "bc.dat").write_text("\n".join(rows) + "\n")
    sponge = [p.params[0] for p in c.patches if p.type == cases.SPONGE]
    src = c.sources[0]
    inp = f"""# magudi.inp of a generated monopole case
output_prefix = "{prefix}"
grid_file = "{prefix}.xyz"
initial_condition_file = "{prefix}.ic.q"
boundary_condition_file = "bc.dat"
control_mollifier_file = "{prefix}.control_mollifier.f"
target_mollifier_file = "{prefix}.target_mollifier.f"
include_viscous_terms = false
use_constant_CFL_mode = false
curvilinear_domain = false
enable_controller = true
enable_functional = true
cost_functional_type = "SOUND"
enable_adjoint_solver = true
defaults/discretization_scheme = "SBP 4-8"
time_step_size = {c.dt!r}
number_of_timesteps = {c.nsteps}
save_interval = {c.save_interval}
add_dissipation = true
dissipation_amount = {c.dissipation_amount!r}
composite_dissipation = false
defaults/sponge_amount = {sponge[0]!r}
number_of_acoustic_sources = 1
acoustic_source01/amplitude = {src.amplitude!r}
acoustic_source01/frequency = {src.frequency!r}
acoustic_source01/radius = {src.radius!r}
acoustic_source01/x = {src.location[0]!r}
acoustic_source01/y = {src.location[1]!r}
{extra}"""
    (d / "magudi.inp").write_text(inp)


def _read_scalar(path):
    return float(open(path).read().split()[0])


@pytest.mark.gpu
@pytest.mark.parametrize("whole_run", [False, True])
def test_fortran_drivers_on_an_example_directory(tmp_path, whole_run):
    """96 steps of the 41^2 monopole at dt 0.2 (the pulse crosses the annulus and reaches the
    sponges and the far-field boundaries): the drivers against the oracle."""
    import oracle_bind as ob
    from magudi_amd import cases, plot3d
    fwd, adj = os.path.join(FDIR, "forward_hip"), os.path.join(FDIR, "adjoint_hip")
    if not (os.path.exists(fwd) and os.path.exists(adj)):
        pytest.skip("Fortran drivers not built (needs amdflang at build time)")
    c = cases.acoustic_monopole(n=41, nsteps=96, save_interval=10)
    c.dt = 0.2
    _write_case_directory(tmp_path, c)
    flags = ["--whole-run"] if whole_run else []
    print(f"fortran drivers in {tmp_path} whole_run={whole_run}", flush=True)
    out = subprocess.run([adj, *flags], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    J = _read_scalar(tmp_path / "Monopole.forward_run.txt")
    sens = _read_scalar(tmp_path / "Monopole.adjoint_run.txt")
    na = len(c.patch_points(c.patches_of(cases.ACTUATOR)[0]))
    grad = plot3d.read_control(str(tmp_path / "Monopole.gradient_controlRegion.dat"), na)
    so, go, _, Jo = ob.Region(c).run_adjoint(c.initial_state(), c.dt, c.nsteps, c.save_interval)
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    assert abs(sens - so) <= 1e-10 * abs(so)
    assert np.abs(grad - go).max() <= 1e-10 * np.abs(go).max()
    # the forward driver with the control forcing A g along the gradient (ZAXPY, then
    # controller_switch = true): (J(A) - J(0)) / A against the sensitivity
    A = 1e-4
    plot3d.zaxpy(str(tmp_path / "Monopole.control_forcing_controlRegion.dat"), A,
                 str(tmp_path / "Monopole.gradient_controlRegion.dat"))
    with open(tmp_path / "magudi.inp", "a") as f:
        f.write("controller_switch = true\n")
    out = subprocess.run([fwd, *flags, "--output", "JA.txt"], cwd=tmp_path, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    JA = _read_scalar(tmp_path / "JA.txt")
    assert abs((JA - J) / A - sens) <= 1e-3 * abs(sens)


@pytest.mark.skipif(FC is None, reason="amdflang not available")
def test_input_reader(tmp_path):
    """magudi_hip_input: comments, quotes, defaults, the last definition of a key, bc.dat rows
    (InputHelper, src/InputHelperImpl.f90:142-269)."""
    (tmp_path / "t.inp").write_text(
        '# comment\noutput_prefix = "Case"   # trailing\nnumber_of_timesteps = 12\n'
        'time_step_size = 2.5e-3\nadd_dissipation = true\nnumber_of_timesteps = 16\n'
        "patches/a.b/sponge_amount = 0.25\n")
    (tmp_path / "bc.dat").write_text(
        "# Name Type Grid normDir iMin iMax jMin jMax kMin kMax\n"
        "  sponge.E  SPONGE  1  1  1 29  1 -1  1 -1\n\n  ctl ACTUATOR 1 0 3 7 -9 -2 1 -1 # c\n")
    prog = tmp_path / "reader.f90"
    prog.write_text(
        "program reader\n  use, intrinsic :: iso_c_binding\n  use magudi_hip_input\n"
        "  type(t_PatchRow), allocatable :: rows(:)\n"
        "  call parseInputFile('t.inp')\n"
        "  print '(A)', trim(getOption('output_prefix', 'x'))\n"
        "  print '(I0)', getOption('number_of_timesteps', 0)\n"
        "  print '(ES23.15E3)', getOption('time_step_size', 0.0_c_double)\n"
        "  print '(L1)', getOption('add_dissipation', .false.)\n"
        "  print '(ES23.15E3)', getOption('patches/a.b/sponge_amount', 1.0_c_double)\n"
        "  print '(ES23.15E3)', getOption('missing', 7.0_c_double)\n"
        "  call readBoundaryConditions('bc.dat', rows)\n"
        "  print '(I0,1X,A,1X,A,1X,I0,6(1X,I0))', size(rows), trim(rows(2)%name), "
        "trim(rows(2)%type), rows(2)%normalDirection, rows(2)%extent\n"
        "end program reader\n")
    mod = tmp_path / "m"
    mod.mkdir()
    subprocess.check_call([FC, "-c", os.path.join(FDIR, "magudi_hip_input.f90"), "-J", str(mod),
                           "-o", str(tmp_path / "in.o")], cwd=tmp_path)
    subprocess.check_call([FC, str(prog), str(tmp_path / "in.o"), "-I", str(mod),
                           "-o", str(tmp_path / "reader")], cwd=tmp_path)
    out = subprocess.check_output([str(tmp_path / "reader")], cwd=tmp_path).decode().split("\n")
    assert out[0] == "Case"
    assert out[1] == "16"
    assert float(out[2]) == 2.5e-3
    assert out[3] == "T"
    assert float(out[4]) == 0.25 and float(out[5]) == 7.0
    assert out[6].split() == ["2", "ctl", "ACTUATOR", "0", "3", "7", "-9", "-2", "1", "-1"]


@pytest.mark.gpu
@pytest.mark.parametrize("penalty", ["quadratic", "huber"])
def test_multiple_shooting_drivers_on_an_example_directory(tmp_path, penalty):
    """msforward_hip / msadjoint_hip (bin/msforward.f90, bin/msadjoint.f90) on a generated
    monopole directory with 3 segments of 8 steps, per-segment initial conditions
    <prefix>-<k>.ic.q (the oracle's states at t = 8 k dt, each but the first with a seeded
    momentum bump, times in the file headers) and a control forcing file: J, the segment
    costs and mismatches, the end snapshots, the total and per-segment gradient inner
    products, the IC gradients, the matching adjoint terminals and the control gradient
    file against the oracle's msforward / msadjoint (J 1e-12, gradients 1e-10)."""
    import ms_oracle as ms
    import oracle_bind as ob
    from magudi_amd import cases, plot3d
    fwd, adj = os.path.join(FDIR, "msforward_hip"), os.path.join(FDIR, "msadjoint_hip")
    if not (os.path.exists(fwd) and os.path.exists(adj)):
        pytest.skip("Fortran drivers not built (needs amdflang at build time)")
    nsplit, nts, K, w, h = 3, 8, 4, 20.0, 1e-3
    c = cases.acoustic_monopole(n=41, nsteps=nsplit * nts, save_interval=K)
    c.dt = 0.2
    extra = (f"time_splitting/number_of_segments = {nsplit}\n"
             f"time_splitting/segment_length = {nts}\n"
             "time_splitting/start_timestep = 0\n"
             f"time_splitting/penalty_weight = {w!r}\n"
             f'time_splitting/penalty_type = "{penalty}"\n'
             f"time_splitting/huber/threshold = {h!r}\n"
             "adjoint_nonzero_initial_condition = true\n"
             "controller_switch = true\n")
    _write_case_directory(tmp_path, c, extra=extra)
    O = ob.Region(c)
    na = O.n_actuator
    assert na > 0
    rng = np.random.default_rng(17)
    ctl = 1e-2 * rng.standard_normal((nsplit * 4 * nts, na))
    plot3d.write_control(str(tmp_path / "Monopole.control_forcing_controlRegion.dat"), ctl)
    x, y = c.coords[0], c.coords[1]
    nv = c.nd + 2
    q = c.initial_state()
    ics, t0s = [q], [0.0]
    for k in range(1, nsplit):
        _, q = O.run_forward(q, c.dt, nts, control=ms.control_slice(ctl, k - 1, nts),
                             t0=(k - 1) * nts * c.dt)
        x0, y0 = rng.uniform(-2, 2, size=2)
        bump = np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / 2.0)
        ic = q.copy()
        ic[1] = ic[1] + 1e-3 * bump
        ic[2] = ic[2] - 5e-4 * bump
        ics.append(ic)
        t0s.append(k * nts * c.dt)
    for k in range(nsplit):
        plot3d.write_solution(str(tmp_path / f"Monopole-{k}.ic.q"), c.n, ics[k],
                              timestep=k * nts, time=t0s[k])
    Jo, cost_o, L2_o, ends = ms.ms_forward(O, ics, t0s, c.dt, nts, w, control=ctl, kind=penalty,
                                           threshold=h)
    res = ms.ms_adjoint(O, ics, t0s, c.dt, nts, K, w, control=ctl, kind=penalty, threshold=h)
    print(f"msforward_hip / msadjoint_hip in {tmp_path} penalty={penalty}", flush=True)

    out = subprocess.run([fwd], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    J = _read_scalar(tmp_path / "Monopole.forward_run.txt")
    assert abs(J - Jo) <= 1e-12 * abs(Jo)
    rows = np.loadtxt(tmp_path / "Monopole.sub_J.txt", comments="#", ndmin=2)
    assert np.array_equal(rows[:, 0], np.arange(nsplit))
    np.testing.assert_allclose(rows[:, 1], cost_o, rtol=1e-12)
    assert rows[0, 2] == 0.0 and min(L2_o[1:]) > 0
    np.testing.assert_allclose(rows[1:, 2], L2_o[1:], rtol=1e-12)
    for k in range(nsplit):
        qT, aux = plot3d.read_solution(str(tmp_path / f"Monopole-{k}-{(k + 1) * nts:08d}.q"), nv)
        assert aux[0] == (k + 1) * nts and abs(aux[3] - (k + 1) * nts * c.dt) <= 1e-12
        assert np.abs(qT - ends[k]).max() <= 1e-12 * np.abs(ends[k]).max(), k

    out = subprocess.run([adj], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    total = _read_scalar(tmp_path / "Monopole.adjoint_run.txt")
    assert abs(total - res["total"]) <= 1e-10 * abs(res["total"])
    rows = np.loadtxt(tmp_path / "Monopole.sub_adjoint_run.txt", comments="#", ndmin=2)
    np.testing.assert_allclose(rows[:, 1], res["ctrl_ip"], rtol=1e-10)
    assert rows[0, 2] == 0.0
    np.testing.assert_allclose(rows[1:, 2], res["ic_ip"][1:], rtol=1e-10)
    for k in range(nsplit):
        g, _ = plot3d.read_solution(str(tmp_path / f"Monopole-{k}.ic.adjoint.q"), nv)
        go = res["ic_grad"][k]
        assert np.abs(g - go).max() <= 1e-10 * np.abs(go).max(), k
    for k in range(nsplit - 1):
        lamT, _ = plot3d.read_solution(
            str(tmp_path / f"Monopole-{k}-{(k + 1) * nts:08d}.adjoint.q"), nv)
        lo = res["scale"] * (ends[k] - ics[k + 1])
        assert np.abs(lamT - lo).max() <= 1e-10 * np.abs(lo).max(), k
    grad = plot3d.read_control(str(tmp_path / "Monopole.gradient_controlRegion.dat"), na)
    assert grad.shape == res["control_grad"].shape
    assert np.abs(grad - res["control_grad"]).max() <= 1e-10 * np.abs(res["control_grad"]).max()


@pytest.mark.parametrize("inp,driver,message", [
    ('output_prefix = "X"\n', "msforward_hip", "'initial_condition_file' (placeholder ok)"),
    ('initial_condition_file = "x"\n', "msforward_hip",
     "Required parameter 'time_splitting/number_of_segments' not found!"),
    ('initial_condition_file = "x"\nenable_adjoint_solver = true\n', "msadjoint_hip",
     "must contain 'adjoint_nonzero_initial_condition'"),
    ('initial_condition_file = "x"\ntime_splitting/number_of_segments = 2\n'
     "time_splitting/segment_length = 4\ntime_splitting/start_timestep = 0\n"
     'time_splitting/penalty_weight = 1.0\ntime_splitting/penalty_type = "l1"\n',
     "msforward_hip", "expected 'quadratic' or 'huber'"),
    ('initial_condition_file = "x"\ntime_splitting/number_of_segments = 2\n'
     "time_splitting/segment_length = 4\ntime_splitting/start_timestep = 0\n"
     "time_splitting/penalty_weight = 1.0\ntime_splitting/state_mollifier/enabled = true\n",
     "msforward_hip", "state_mollifier/enabled = true is not built"),
    ('output_prefix = "X"\n', "compute_norm_hip",
     "must contain 'enable_adjoint_solver' for compute_norm."),
    ('enable_adjoint_solver = true\n', "compute_norm_hip",
     "Required parameter 'time_splitting/number_of_segments' not found!"),
])
def test_multiple_shooting_driver_input_errors(tmp_path, inp, driver, message):
    """The drivers stop with the reference's messages (bin/msforward.f90:127,
    bin/msadjoint.f90:131-144, src/MSPenaltyImpl.f90:138, bin/compute_norm.f90:110-113) on missing or invalid time_splitting input, before any
    device work."""
    exe = os.path.join(FDIR, driver)
    if not os.path.exists(exe):
        pytest.skip("Fortran drivers not built (needs amdflang at build time)")
    (tmp_path / "magudi.inp").write_text(inp)
    out = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode != 0
    assert message in out.stdout + out.stderr


@pytest.mark.gpu
def test_compute_norm_driver_on_an_example_directory(tmp_path):
    """compute_norm_hip (bin/compute_norm.f90) on a generated monopole directory with 3
    segments of 8 steps: the actuator norm file holds H at the patch points (patch order)
    times the RK4 weight and dt of every stage (collectThermalActuatorNorm,
    src/ThermalActuatorImpl.f90:710-763), norm_ic.q state_controllability H in every
    variable, the layout the reference's optimiser reads, and --save_metrics the grid's
    Jacobian and metrics -- against the oracle's grid arrays, and the mismatch of
    number_of_timesteps with Nts * Nsplit stops with the reference's message."""
    import oracle_bind as ob
    from magudi_amd import cases, plot3d
    exe = os.path.join(FDIR, "compute_norm_hip")
    if not os.path.exists(exe):
        pytest.skip("Fortran drivers not built (needs amdflang at build time)")
    nsplit, nts, w = 3, 8, 2.5
    c = cases.acoustic_monopole(n=41, nsteps=nsplit * nts, save_interval=4)
    c.dt = 0.2
    extra = (f"time_splitting/number_of_segments = {nsplit}\n"
             f"time_splitting/segment_length = {nts}\n"
             "time_splitting/start_timestep = 0\n"
             f"time_splitting/state_controllability = {w!r}\n")
    _write_case_directory(tmp_path, c, extra=extra)
    print(f"compute_norm_hip in {tmp_path}", flush=True)
    out = subprocess.run([exe, "--save_metrics"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    O = ob.Region(c)
    H = O.grid(2)[0]
    pts = c.patch_points(c.patches_of(cases.ACTUATOR)[0])
    na = len(pts)
    norm = plot3d.read_control(str(tmp_path / "Monopole.norm_controlRegion.dat"), na)
    assert norm.shape == (4 * nsplit * nts, na)
    wts = [1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0]
    for s in range(norm.shape[0]):
        np.testing.assert_allclose(norm[s], H[pts] * (wts[s % 4] * c.dt), rtol=1e-14, atol=0.0)
    q, aux = plot3d.read_solution(str(tmp_path / "Monopole.norm_ic.q"), c.nd + 2)
    for v in range(c.nd + 2):
        np.testing.assert_allclose(q[v], w * H, rtol=1e-14, atol=0.0)
    rows = [ln.split() for ln in (tmp_path / "Monopole.layout.txt").read_text().splitlines()]
    assert rows[0] == ["#", "kind", "identifier", "size_in_real64"]
    assert rows[1] == ["actuator", "controlRegion", str(4 * nsplit * nts * na)]
    assert rows[2:] == [["ic", str(k), str(c.npoints * (c.nd + 2))] for k in range(1, nsplit)]
    jac = plot3d.read_function(str(tmp_path / "Monopole.Jacobian.f"))
    np.testing.assert_allclose(jac[0], O.grid(1)[0], rtol=1e-14, atol=0.0)
    with open(tmp_path / "magudi.inp", "a") as f:
        f.write(f"number_of_timesteps = {nsplit * nts + 1}\n")
    out = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode != 0
    assert "but compute_norm requires Nts * Nsplit" in out.stdout + out.stderr
