"""A Fortran host (the reference's toolchain) calls the C-ABI through the
ISO_C_BINDING module frealign_amd/fortran/frealign_hip_mod.f90 and compares
the library's host geometry with the reference's own SEARCHANG / GETSYMMAT
(flang-compiled from the reference sources).  Needs no GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
REF = os.path.join(ROOT, "oracle", "_ref", "libfrealign_ref.so")
LIB = os.path.join(ROOT, "frealign_amd", "libfrealign_hip.so")
HOSTS = os.path.join(ROOT, "tests", "fortran")


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_host_matches_reference(tmp_path):
    if not os.path.exists(REF):
        pytest.skip("reference routines (oracle/_ref) not built")
    src = os.path.join(ROOT, "frealign_amd", "fortran")
    mod = tmp_path / "mod.o"
    r = subprocess.run([FLANG, "-O1", "-c", "-module-dir", str(tmp_path),
                        os.path.join(src, "frealign_hip_mod.f90"), "-o", str(mod)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exe = tmp_path / "fr_grid_check"
    r = subprocess.run([FLANG, "-O1", "-module-dir", str(tmp_path),
                        os.path.join(HOSTS, "fr_grid_check.f90"), str(mod), REF, LIB,
                        f"-Wl,-rpath,{os.path.dirname(REF)}:{os.path.dirname(LIB)}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "FR_GRID_CHECK mismatches: 0" in r.stdout, r.stdout[-2000:]


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_fortran_host_powell_matches_reference(tmp_path):
    """The library's VA04A called from Fortran reproduces the reference's
    flang-compiled VA04A bit for bit (final X, F and evaluation count)."""
    if not os.path.exists(REF):
        pytest.skip("reference routines (oracle/_ref) not built")
    src = os.path.join(ROOT, "frealign_amd", "fortran")
    mod = tmp_path / "mod.o"
    r = subprocess.run([FLANG, "-O1", "-c", "-module-dir", str(tmp_path),
                        os.path.join(src, "frealign_hip_mod.f90"), "-o", str(mod)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    exe = tmp_path / "fr_powell_check"
    r = subprocess.run([FLANG, "-O1", "-module-dir", str(tmp_path),
                        os.path.join(HOSTS, "fr_powell_check.f90"), str(mod), REF, LIB,
                        f"-Wl,-rpath,{os.path.dirname(REF)}:{os.path.dirname(LIB)}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "FR_POWELL_CHECK mismatches: 0" in r.stdout, r.stdout[-2000:]
