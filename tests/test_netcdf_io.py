"""The NetCDF writer/reader behind particle_data.nc and microbe_data.nc.

- CDF-2 files are read back by scipy.io.netcdf_file, an independent reader:
  names, dimensions, dtypes, coordinates, time units and data;
- the CDF-5 header of a tiny dataset equals the bytes assembled here by hand
  from the format grammar (8-byte counts, lengths, dimension ids, sizes and
  offsets; 4-byte tags and types; names and values padded to 4 bytes);
- a CDF-5 variable over 4 GiB (a sparse file) round-trips: its last rows
  lie more than 2^32 bytes past its start, and the next variable starts
  past 2^32;
- block writes of steps into columns, and the file-to-file column copy that
  makes particle_data.nc from microbe_data.nc;
- row writes of whole particles (the device-resident records' path), from
  arrays and from raw bytes, and their argument checks.
"""
import struct
from datetime import datetime, timedelta

import numpy as np
import pytest
from scipy.io import netcdf_file

from lagrangian_microbes_amd.netcdf_io import (ParticleDataset, copy_variables, open_particle_dataset,
                                               read_header, read_particle_dataset, write_particle_dataset)

START = datetime(2017, 1, 1)


def hourly(T):
    return [START + k * timedelta(hours=1) for k in range(T)]


def test_cdf2_read_by_scipy(tmp_path):
    rng = np.random.default_rng(1)
    lon = rng.random((9, 4)).astype(np.float32)
    sp = rng.integers(1, 4, (9, 4)).astype(np.int8)
    path = str(tmp_path / "m.nc")
    write_particle_dataset(path, {"longitude": lon, "species": sp}, hourly(4), 9)
    with netcdf_file(path, "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions == {"particle number": 9, "time": 4}
        assert f.variables["longitude"].dimensions == ("particle number", "time")
        assert f.variables["longitude"][:].dtype.newbyteorder("=") == np.float32
        assert np.array_equal(f.variables["longitude"][:], lon)
        assert np.array_equal(f.variables["species"][:], sp)
        assert f.variables["species"][:].dtype.itemsize == 1
        assert list(f.variables["particle number"][:]) == list(range(1, 10))
        assert list(f.variables["time"][:]) == [0, 1, 2, 3]
        assert f.variables["time"].units == b"hours since 2017-01-01 00:00:00"
    data, times, n = read_particle_dataset(path)
    assert n == 9 and times == hourly(4)
    assert np.array_equal(data["longitude"], lon) and np.array_equal(data["species"], sp)


def _name(s):
    b = s.encode()
    return struct.pack(">q", len(b)) + b + b"\0" * (-len(b) % 4)


def test_cdf5_header_bytes(tmp_path):
    path = str(tmp_path / "s.nc")
    sp = np.array([[1, 2, 3], [3, 2, 1]], dtype=np.int8)
    write_particle_dataset(path, {"species": sp}, hourly(3), 2, version=5)
    units, cal = b"hours since 2017-01-01 00:00:00", b"proleptic_gregorian"
    absent = struct.pack(">i", 0) + struct.pack(">q", 0)
    want = b"CDF\x05" + struct.pack(">q", 0)                               # magic, numrecs
    want += struct.pack(">i", 0x0A) + struct.pack(">q", 2)                  # dimension list
    want += _name("particle number") + struct.pack(">q", 2)
    want += _name("time") + struct.pack(">q", 3)
    want += absent                                                          # no global attributes
    want += struct.pack(">i", 0x0B) + struct.pack(">q", 3)                  # variable list
    want += _name("particle number") + struct.pack(">qq", 1, 0) + absent
    want += struct.pack(">i", 10) + struct.pack(">qq", 16, 4096)            # NC_INT64, vsize, begin
    want += _name("time") + struct.pack(">qq", 1, 1)
    want += struct.pack(">i", 0x0C) + struct.pack(">q", 2)
    want += _name("units") + struct.pack(">i", 2) + struct.pack(">q", len(units)) + units + b"\0" * (-len(units) % 4)
    want += _name("calendar") + struct.pack(">i", 2) + struct.pack(">q", len(cal)) + cal + b"\0" * (-len(cal) % 4)
    want += struct.pack(">i", 10) + struct.pack(">qq", 24, 8192)
    want += _name("species") + struct.pack(">qqq", 2, 0, 1) + absent
    want += struct.pack(">i", 1) + struct.pack(">qq", 8, 12288)             # NC_BYTE, 6 bytes padded to 8
    raw = open(path, "rb").read()
    assert raw[:len(want)] == want
    assert set(raw[len(want):4096]) <= {0}
    assert raw[4096:4112] == struct.pack(">qq", 1, 2)
    assert raw[8192:8216] == struct.pack(">qqq", 0, 1, 2)
    assert raw[12288:12294] == sp.tobytes()
    data, times, n = read_particle_dataset(path)
    assert np.array_equal(data["species"], sp) and times == hourly(3) and n == 2


def test_cdf5_variable_over_4gib(tmp_path):
    # 1,100,000 particles x 1,000 float32 steps = 4.4e9 bytes per variable: over the CDF-2 limit, so the
    # writer picks CDF-5 by itself; the file is sparse except for the rows written here
    n, T = 1_100_000, 1000
    path = str(tmp_path / "big.nc")
    with pytest.raises(ValueError, match="64-bit data"):
        ParticleDataset(path, n, hourly(T), {"longitude": np.float32}, version=2)
    rng = np.random.default_rng(2)
    rows = rng.random((3, T)).astype(np.float32)
    with ParticleDataset(path, n, hourly(T), {"longitude": np.float32, "latitude": np.float32}) as ds:
        assert ds.version == 5
        ds.vars["longitude"][-3:, :] = rows
        ds.vars["latitude"][-3:, :] = rows[::-1]
        ds.vars["latitude"][0, :] = rows[0]
    h = read_header(path)
    assert h["version"] == 5
    lb, ln = h["vars"]["longitude"]["span"]
    tb, tn = h["vars"]["latitude"]["span"]
    assert ln == tn == n * T * 4 > 2 ** 32 and tb > 2 ** 32 and tb >= lb + ln
    _, arrays = open_particle_dataset(path)
    assert np.array_equal(arrays["longitude"][-3:], rows) and np.array_equal(arrays["latitude"][-3:], rows[::-1])
    assert np.array_equal(arrays["latitude"][0], rows[0]) and not arrays["longitude"][0].any()
    assert list(arrays["time"][[0, -1]]) == [0, T - 1] and arrays["time"].dtype == np.dtype(">i8")


@pytest.mark.parametrize("version", [2, 5])
def test_block_writes_and_column_copy(tmp_path, version):
    rng = np.random.default_rng(3)
    n, T = 20_001, 37
    lon = rng.random((T, n)).astype(np.float32)   # step-major, as the device records
    sp = rng.integers(1, 4, (T, n)).astype(np.int8)
    m, p = str(tmp_path / "microbe.nc"), str(tmp_path / "particle.nc")
    with ParticleDataset(m, n, hourly(T), {"longitude": np.float32, "latitude": np.float32, "species": np.int8},
                         version=version) as ds:
        for t0, k in ((0, 10), (10, 10), (20, 16), (36, 1)):
            ds.write_steps("longitude", t0, lon[t0:t0 + k])
            ds.write_steps("latitude", t0, -lon[t0:t0 + k])
            ds.write_steps("species", t0, sp[t0:t0 + k])
        with pytest.raises(ValueError):
            ds.write_steps("species", 30, sp[:8])
    ParticleDataset(p, n, hourly(T), {"longitude": np.float32, "latitude": np.float32}, version=version).close()
    copy_variables(m, p, ["longitude", "latitude"])
    d, times, nn = read_particle_dataset(m)
    e, _, _ = read_particle_dataset(p)
    assert nn == n and times == hourly(T)
    assert np.array_equal(d["longitude"], lon.T) and np.array_equal(d["latitude"], -lon.T)
    assert np.array_equal(d["species"], sp.T)
    assert set(e) == {"longitude", "latitude"}
    assert np.array_equal(e["longitude"], lon.T) and np.array_equal(e["latitude"], -lon.T)
    if version == 2:
        with netcdf_file(p, "r", mmap=False) as f:
            assert np.array_equal(f.variables["latitude"][:], -lon.T)


@pytest.mark.parametrize("version", [2, 5])
def test_row_writes(tmp_path, version):
    rng = np.random.default_rng(4)
    n, T = 10_007, 13
    lon = rng.random((n, T)).astype(np.float32)   # particle-major, as the file holds it
    sp = rng.integers(1, 4, (n, T)).astype(np.int8)
    path = str(tmp_path / "rows.nc")
    with ParticleDataset(path, n, hourly(T), {"longitude": np.float32, "species": np.int8}, version=version,
                         threads=1) as ds:
        for p0 in range(0, n, 3000):
            ds.write_rows("longitude", p0, lon[p0:p0 + 3000])
            ds.write_rows("species", p0, sp[p0:p0 + 3000].view(np.uint8).reshape(-1))  # raw bytes
        with pytest.raises(ValueError):
            ds.write_rows("longitude", n - 1, lon[:2])        # past the last particle
        with pytest.raises(ValueError):
            ds.write_rows("longitude", 0, lon[0, :5])         # not whole rows
    d, times, nn = read_particle_dataset(path)
    assert nn == n and times == hourly(T)
    assert np.array_equal(d["longitude"], lon) and np.array_equal(d["species"], sp)
