"""NetCDF classic output (gfsph_write_netcdf), read back by an independent
reader (scipy.io.netcdf_file).  No reference counterpart in this format
(parity unpinned against the reference; the file is pinned by the format's
published specification through that reader)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

from gfsph import GfsphError, write_netcdf
from gfsph.dem import DEM


def test_rasters_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    dem = DEM(Z=np.zeros((37, 53)), xmin=472168.9205068179, ymin=289080.60321233014, dx=10.0, dy=12.5)
    a = rng.normal(size=(37, 53))
    b = np.where(rng.random((37, 53)) < 0.2, -9999.0, rng.random((37, 53)))
    path = str(tmp_path / "h.nc")
    write_netcdf(path, dem, {"max_h": (a, "m", "maximum depth"), "p_exceed": (b, "1", "exceedance probability")},
                 title="test rasters", fill=-9999.0)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions == {"x": 53, "y": 37}
        assert f.Conventions == b"CF-1.8" and f.title == b"test rasters"
        np.testing.assert_array_equal(f.variables["x"][:], dem.xmin + np.arange(53) * dem.dx)
        np.testing.assert_array_equal(f.variables["y"][:], dem.ymin + np.arange(37) * dem.dy)
        assert f.variables["x"].units == b"m" and f.variables["x"].standard_name == b"projection_x_coordinate"
        v = f.variables["max_h"]
        assert v.dimensions == ("y", "x") and v.units == b"m" and v.long_name == b"maximum depth"
        assert v._FillValue == -9999.0
        np.testing.assert_array_equal(v[:], a)
        np.testing.assert_array_equal(f.variables["p_exceed"][:], b)


@pytest.mark.parametrize("name", ["", "2h", "x", "max-h"])
def test_invalid_names_are_rejected(name, tmp_path):
    dem = DEM(Z=np.zeros((2, 2)), xmin=0.0, ymin=0.0, dx=1.0, dy=1.0)
    with pytest.raises(GfsphError):
        write_netcdf(str(tmp_path / "bad.nc"), dem, {name: (np.zeros((2, 2)), "m", "")})
