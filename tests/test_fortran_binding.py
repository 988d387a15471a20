"""The Fortran ISO_C_BINDING module (qpad_amd/fortran/qpad_hip_mod.f03) that
QPAD's side would use: it compiles with amdflang, its bind(C) types have the
C layout, and a Fortran program reaches the library and gets its error
messages back (CPU only: argument validation fails before any device call).
"""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import qpad_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "qpad_amd", "fortran", "qpad_hip_mod.f03")
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROGRAM = """
program t
  use qpad_hip_mod
  use, intrinsic :: iso_c_binding
  implicit none
  type(qpad_grid) :: g
  type(qpad_species_cfg) :: s
  type(qpad_beam_cfg) :: b
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  print '(a,i0)', 'abi ', qpad_abi_version()
  print '(a,i0)', 'grid ', c_sizeof(g)
  print '(a,i0)', 'species ', c_sizeof(s)
  print '(a,i0)', 'beam ', c_sizeof(b)
  g = qpad_grid(4, 16, 1, 0.1d0, 0.05d0, 1d-3, QPAD_BND_OPEN, 10, 1d-3, 1d-3, 10d0)
  rc = qpad_create(g, ctx)
  print '(a,i0)', 'rc ', rc
  print '(a,a)', 'msg ', qpad_error_message()
  call qpad_hip_check(rc, 'qpad_create', on_error)
contains
  subroutine on_error(msg)
    character(len=*), intent(in) :: msg
    print '(a,a)', 'err ', msg
  end subroutine
end program
"""


@pytest.mark.skipif(FC is None, reason="amdflang not in this image")
def test_fortran_module_binds_the_library(tmp_path):
    lib = L.load()
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run([FC, "-c", MOD, "-o", str(tmp_path / "mod.o"), "-J", str(tmp_path)], check=True,
                   cwd=tmp_path)
    (tmp_path / "t.f03").write_text(PROGRAM)
    subprocess.run([FC, "t.f03", "mod.o", "-I", str(tmp_path), f"-L{libdir}", "-lqpad_hip",
                    f"-Wl,-rpath,{libdir}", "-o", "t"], check=True, cwd=tmp_path)
    out = subprocess.run([str(tmp_path / "t")], check=True, capture_output=True, text=True).stdout
    kv = dict(line.strip().split(" ", 1) for line in out.splitlines() if line.strip())
    assert int(kv["abi"]) == lib.qpad_abi_version() == 2
    assert int(kv["grid"]) == C.sizeof(L.Grid)
    assert int(kv["species"]) == C.sizeof(L.SpeciesCfg)
    assert int(kv["beam"]) == C.sizeof(L.BeamCfg)
    assert int(kv["rc"]) == -1
    assert "nr" in kv["msg"]
    assert kv["err"].startswith("HIP: qpad_create: nr")


@pytest.mark.skipif(FC is None, reason="amdflang not in this image")
def test_fortran_module_declares_every_entry_point():
    import re
    text = open(MOD).read().lower()
    bound = set(re.findall(r'bind\(c, name="(qpad_[a-z0-9_]+)"\)', text))
    assert bound == set(L.SIGNATURES)


def test_fortran_driver_reports_library_errors_through_write_err(tmp_path):
    """The Fortran driver (qpad_amd/fortran/qpad_slice_driver.f03) stops
    through its write_err stand-in with the library's message when a call
    fails (CPU: the grid is refused before any device call)."""
    import fortran_driver
    from qpad_amd import configs
    if not os.path.exists(fortran_driver.DRIVER):
        pytest.skip("Fortran driver not built (no amdflang)")
    deck = configs.small(nr=4)
    with pytest.raises(subprocess.CalledProcessError) as e:
        fortran_driver.run(deck, str(tmp_path))
    assert e.value.returncode == 1
    assert e.value.stderr.strip().startswith("HIP: qpad_create: nr"), e.value.stderr
