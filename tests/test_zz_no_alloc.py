"""The allocation-free inner product (mgd_inner_product_noalloc) against mgd_inner_product.

Both run the same kernels; the allocation-free entry point keeps the per-block partial
sums in the library's fixed device scratch array and writes the per-instance results to
a device array the caller owns.  The device array comes from the HIP runtime the library
itself links (libamdhip64), through ctypes.  Two calls in a row, the second on a smaller
region, must each equal the allocating entry point bit for bit; a region beyond the
scratch capacity is refused with an error, before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from magudi_amd import capi, cases

pytestmark = pytest.mark.gpu


def _hip():
    capi.lib()                      # loads libamdhip64 as the library's dependency
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _set_states(R, seed):
    rng = np.random.default_rng(seed)
    for i in range(R.n_instances):
        R.states[i].conservedVariables = rng.standard_normal((R.nv, R.npoints))
        R.states[i].adjointVariables = rng.standard_normal((R.nv, R.npoints))


def _noalloc(hip, R, dev, nbytes):
    # a stale value in the output array must not survive
    assert hip.hipMemset(dev, 0x7f, nbytes) == 0
    capi.check(capi.lib().mgd_inner_product_noalloc(R.handle, capi.CONSERVED, capi.ADJOINT_VARS, dev))
    R.sync()
    out = np.empty(R.n_instances)
    assert hip.hipMemcpy(out.ctypes.data, dev, 8 * R.n_instances, 2) == 0   # device to host
    return out


def _alloc(R):
    out = np.empty(R.n_instances)
    capi.check(capi.lib().mgd_inner_product(R.handle, capi.CONSERVED, capi.ADJOINT_VARS,
                                            capi.dptr(out)))
    return out


def test_noalloc_inner_product_matches_allocating_entry_point():
    from magudi_amd.solver import Region
    hip = _hip()
    dev = C.c_void_p()
    nbytes = 8 * 4
    assert hip.hipMalloc(C.byref(dev), nbytes) == 0
    try:
        big = Region(cases.acoustic_monopole(n=97, nsteps=1), n_instances=4)
        small = Region(cases.acoustic_monopole(n=41, nsteps=1), n_instances=2)
        _set_states(big, 1)
        _set_states(small, 2)
        for R in (big, small):
            print(f"noalloc inner product: {R.case.n[:2]} x {R.n_instances} instances", flush=True)
            got = _noalloc(hip, R, dev, nbytes)
            ref = _alloc(R)
            assert np.all(np.isfinite(ref))
            assert np.array_equal(got, ref), (got, ref)
        # the larger region again after the smaller one: the scratch holds the small
        # region's partial sums and must not leak into the result
        got = _noalloc(hip, big, dev, nbytes)
        assert np.array_equal(got, _alloc(big))
    finally:
        hip.hipFree(dev)


def test_noalloc_inner_product_rejects_regions_beyond_the_scratch():
    """A 1-D region of 2^26 + 1 points needs 2^18 + 1 partial sums, one more than the
    scratch holds: refused with an error, nothing launched."""
    from magudi_amd.solver import Region
    n = (1 << 26) + 1
    x = np.arange(n, dtype=np.float64) / n
    c = cases.Case(name="long1d", nd=1, n=(n, 1, 1), coords=x[None, :],
                   periodicity=(cases.PLANE, 0, 0), periodic_length=(1.0, 0, 0),
                   dissipation=False, target=cases.quiescent_state(1, n),
                   control_mollifier=None, target_mollifier=None, dt=1e-3, nsteps=1,
                   save_interval=1)
    R = Region(c)
    out = C.c_void_p(1)             # never written: the call fails before any launch
    rc = capi.lib().mgd_inner_product_noalloc(R.handle, capi.CONSERVED, capi.ADJOINT_VARS, out)
    assert rc == -1
    assert b"scratch" in capi.lib().mgd_last_error()
    R.close()
