"""Fortran ISO_C_BINDING host side (magudi_amd/fortran) against the C ABI.

magudi's drivers are Fortran (bin/Forward.f90, bin/Adjoint.f90); the module
magudi_hip.f90 is what such a driver uses.  CPU checks: every bind(C) name is an
entry point of include/magudi_hip.h, and the bind(C) derived types have the C
structs' sizes.  GPU check: the Fortran adjoint driver prints the same J and
sensitivity as the Python host path for the same monopole set-up.
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
    exe = os.path.join(FDIR, "adjoint_hip")
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
