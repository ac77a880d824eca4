"""CPU-side checks of the C ABI: the library loads and exports every declared entry
point, and argument errors are reported (no GPU work is launched here)."""
import ctypes as C

import numpy as np
import pytest

from magudi_amd import capi, cases


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_device_count_without_gpu_is_safe():
    assert capi.device_count() >= 0


def test_region_create_rejects_bad_arguments():
    L = capi.lib()
    d = capi.RegionDesc()
    d.nd = 4
    h = C.c_void_p()
    coords = np.zeros(4)
    rc = L.mgd_region_create(C.byref(d), capi.dptr(coords), C.byref(h))
    assert rc == -1
    assert b"nd must be" in L.mgd_last_error()


def test_case_generators_shapes():
    c = cases.acoustic_monopole(n=41, nsteps=2)
    assert c.coords.shape == (2, 41 * 41)
    assert c.target.shape == (4, 41 * 41)
    act = c.patches_of(cases.ACTUATOR)[0]
    pts = c.patch_points(act)
    assert np.all(c.control_mollifier[pts] >= 0) and c.control_mollifier[pts].max() > 0.9
    k = cases.kolmogorov3d(n=8, nsteps=1)
    u = k.initial[1:4] / k.initial[0]
    assert abs(np.sqrt((u ** 2).sum(0).mean()) - 0.05) < 1e-12
