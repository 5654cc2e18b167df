"""Reading a local OSCAR file into the VelocityField the advection samples.

The reference opens oscar_vel<year>.nc with xarray (velocity_fields.py:21-32)
and hands parcels the first depth's u, v with times in seconds since the first
level (particle_advecter.py:160-184).  The files here are written in OSCAR's
layout (time, depth, latitude 80N..80S decreasing, longitude 20..420E, CF time
units, NaN over land) and every value is checked against the array it was
written from: a wrong flip, axis order, depth or time unit fails.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from lagrangian_microbes_amd.velocity_fields import oscar_dataset, oscar_dataset_filename, read_oscar_file


def write_oscar(path, lons, lats, times, u, v, time_units="day since 1992-10-05 00:00:00", dtype="f4",
                fill=None):
    """u, v: [time, depth, latitude, longitude] in the file's own axis order."""
    with netcdf_file(path, "w", version=2) as nc:
        nc.createDimension("time", len(times))
        nc.createDimension("depth", u.shape[1])
        nc.createDimension("latitude", len(lats))
        nc.createDimension("longitude", len(lons))
        t = nc.createVariable("time", "i4", ("time",))
        t[:] = np.asarray(times, dtype=np.int32)
        t.units = time_units
        d = nc.createVariable("depth", "f4", ("depth",))
        d[:] = 15.0 + 10.0 * np.arange(u.shape[1], dtype=np.float32)
        la = nc.createVariable("latitude", "f8", ("latitude",))
        la[:] = lats
        lo = nc.createVariable("longitude", "f8", ("longitude",))
        lo[:] = lons
        for name, data in (("u", u), ("v", v)):
            x = nc.createVariable(name, dtype, ("time", "depth", "latitude", "longitude"))
            if fill is not None:
                x._FillValue = np.array(fill, dtype=dtype)
            x[:] = data.astype(dtype)


def oscar_like(nt=3, nz=2, ny=7, nx=10, seed=3):
    rng = np.random.default_rng(seed)
    lats = 36.0 - np.arange(ny) / 3.0           # decreasing, as OSCAR
    lons = 205.0 + np.arange(nx) / 3.0
    u = rng.normal(0.0, 0.2, (nt, nz, ny, nx)).astype(np.float32)
    v = rng.normal(0.0, 0.2, (nt, nz, ny, nx)).astype(np.float32)
    return lons, lats, u, v


def test_reads_oscar_layout(tmp_path):
    lons, lats, u, v = oscar_like()
    u[1, 0, 2, 3] = np.nan   # land in the depth that is read
    v[2, 0, 6, 0] = np.nan
    u[0, 1, 0, 0] = np.nan   # the second depth is never read
    path = str(tmp_path / oscar_dataset_filename(2017))
    write_oscar(path, lons, lats, [9000, 9005, 9010], u, v)
    f = oscar_dataset(2017, directory=str(tmp_path))
    assert f.nt == 3 and (f.ny, f.nx) == (7, 10)
    assert np.array_equal(f.times, [0.0, 5 * 86400.0, 10 * 86400.0])
    assert np.array_equal(f.lats, lats[::-1]) and np.array_equal(f.lons, lons)
    assert f.dx == pytest.approx(1 / 3, abs=1e-12) and f.dy == pytest.approx(1 / 3, abs=1e-12)
    eu = np.nan_to_num(u[:, 0, ::-1, :].astype(np.float64), nan=0.0)
    ev = np.nan_to_num(v[:, 0, ::-1, :].astype(np.float64), nan=0.0)
    assert np.array_equal(f.u, eu) and np.array_equal(f.v, ev)
    # the NaN nodes, at their flipped rows, are 0
    assert f.u[1, 7 - 1 - 2, 3] == 0.0 and f.v[2, 0, 0] == 0.0
    assert f.u[0, 6, 0] == np.float64(u[0, 0, 0, 0])  # the first depth, not the second
    assert read_oscar_file(path).u.tobytes() == f.u.tobytes()


def test_fill_values_become_zero(tmp_path):
    lons, lats, u, v = oscar_like(nt=2, nz=1, seed=5)
    u[0, 0, 1, 1] = -999.0
    path = str(tmp_path / "f.nc")
    write_oscar(path, lons, lats, [0, 1], u, v, fill=-999.0)
    f = read_oscar_file(path)
    assert f.u[0, 7 - 1 - 1, 1] == 0.0
    assert np.count_nonzero(f.u == 0.0) == 1


@pytest.mark.parametrize("units,scale", [("hours since 2000-01-01", 3600.0), ("seconds since 2000-01-01", 1.0),
                                         ("days since 1992-10-05", 86400.0)])
def test_time_units(tmp_path, units, scale):
    lons, lats, u, v = oscar_like(nt=2, nz=1)
    path = str(tmp_path / "t.nc")
    write_oscar(path, lons, lats, [100, 106], u, v, time_units=units)
    assert np.array_equal(read_oscar_file(path).times, [0.0, 6 * scale])


def test_steady_single_level(tmp_path):
    lons, lats, u, v = oscar_like(nt=1, nz=1)
    path = str(tmp_path / "s.nc")
    write_oscar(path, lons, lats, [0], u, v)
    f = read_oscar_file(path)
    assert f.nt == 1
    lev0, w = f.stage_schedule(0.0, 3600.0, 4)
    assert not lev0.any() and not w.any()


def test_bad_time_units_raise(tmp_path):
    lons, lats, u, v = oscar_like(nt=2, nz=1)
    path = str(tmp_path / "b.nc")
    write_oscar(path, lons, lats, [0, 1], u, v, time_units="fortnights since 2000-01-01")
    with pytest.raises(RuntimeError, match="time units"):
        read_oscar_file(path)


def test_netcdf4_file_is_refused(tmp_path):
    path = tmp_path / oscar_dataset_filename(2019)
    path.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(RuntimeError, match="NetCDF-4"):
        oscar_dataset(2019, directory=str(tmp_path))


def test_missing_file_raises(tmp_path):
    with pytest.raises(RuntimeError, match="OPeNDAP"):
        oscar_dataset(2018, directory=str(tmp_path))
