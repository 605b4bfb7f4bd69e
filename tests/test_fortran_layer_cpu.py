"""The product's ISO_C_BINDING module compiles with the image's flang and
its derived types match the C structs of include/gfsph.h."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORTRAN = os.path.join(REPO, "deslizamientos-y-flujos_amd", "fortran")


@pytest.mark.skipif(shutil.which("flang") is None, reason="flang not available")
def test_gfsph_c_module_compiles_and_struct_sizes_match(tmp_path):
    prog = tmp_path / "sizes.f90"
    prog.write_text("""
program sizes
  use gfsph_c
  use, intrinsic :: iso_c_binding
  type(gfsph_params_t) :: p
  type(gfsph_dem_t) :: d
  print '(i0,1x,i0)', c_sizeof(p), c_sizeof(d)
end program sizes
""")
    subprocess.run(["flang", "-c", "-module-dir", str(tmp_path), os.path.join(FORTRAN, "gfsph_c.f90"),
                    "-o", str(tmp_path / "gfsph_c.o")], check=True)
    libdir = os.path.join(REPO, "deslizamientos-y-flujos_amd", "gfsph")
    subprocess.run(["flang", "-module-dir", str(tmp_path), str(prog), str(tmp_path / "gfsph_c.o"), "-o",
                    str(tmp_path / "sizes"), "-L" + libdir, "-lgfsph", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(tmp_path / "sizes")], capture_output=True, text=True, check=True).stdout.split()
    from gfsph import api
    import ctypes
    assert int(out[0]) == ctypes.sizeof(api.CParams)
    assert int(out[1]) == ctypes.sizeof(api.CDem)
