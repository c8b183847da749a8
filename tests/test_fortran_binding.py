"""The Fortran side of the boundary: include/tcg_engine.f90 (the module a
maintainer adds to the reference's driver) compiles with the image's Fortran
compiler, names exactly the entry points of include/tcg.h, and a Fortran
program linked against the engine library calls through it (no GPU needed:
tcg_create reports the missing device with its status code)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from test_abi import declared_symbols

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "include", "tcg_engine.f90")
LIBDIR = os.path.join(ROOT, "william_ci_amd")
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

PROBE = """
program tcg_probe
  use tcg_engine
  implicit none
  type(c_ptr) :: ctx
  integer(c_int) :: rc, np
  real(c_double) :: L(3) = [18.0d0, 0.0d0, 0.0d0]
  integer(c_int) :: ncl(3) = [10, 1, 1]
  np = tcg_mlp_param_count()
  rc = tcg_create(1, L, ncl, 32, ctx)
  if (rc == TCG_OK) then
    rc = tcg_eri_build(ctx, 1, 0, c_null_ptr)        ! no basis yet: an error with a message
    print '(a,i0,a,i0,a,a)', 'params ', np, ' build ', rc, ' msg ', tcg_message(tcg_last_error(ctx))
    rc = tcg_destroy(ctx)
  else
    print '(a,i0,a,i0)', 'params ', np, ' create ', rc
  end if
end program tcg_probe
"""


def test_module_names_the_header_entry_points():
    text = open(MOD).read()
    names = set(re.findall(r'bind\(C, name="(tcg_\w+)"\)', text))
    assert names == set(declared_symbols())


@pytest.mark.skipif(not os.path.exists(FC), reason="no Fortran compiler")
@pytest.mark.skipif(not os.path.exists(os.path.join(LIBDIR, "libtcg.so")), reason="engine library not built")
def test_fortran_program_calls_the_engine(tmp_path):
    shutil.copy(MOD, tmp_path / "tcg_engine.f90")
    (tmp_path / "probe.f90").write_text(PROBE)
    subprocess.run([FC, "-c", "tcg_engine.f90"], cwd=tmp_path, check=True, capture_output=True, timeout=300)
    subprocess.run([FC, "probe.f90", "tcg_engine.o", f"-L{LIBDIR}", "-ltcg", f"-Wl,-rpath,{LIBDIR}", "-o", "probe"],
                   cwd=tmp_path, check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(tmp_path / "probe")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(os.path.join(LIBDIR, "libtcg.so"))
    m = re.search(r"params (\d+) (create|build) (-?\d+)(?: msg (.*))?", r.stdout)
    assert m, r.stdout
    assert int(m.group(1)) == lib.tcg_mlp_param_count()
    if m.group(2) == "create":
        assert int(m.group(3)) == -3                     # TCG_E_DEVICE: no GPU on this host
    else:
        assert int(m.group(3)) == -2 and "before tcg_set_basis" in m.group(4)
