"""The on-disk outputs: particle_data.nc and microbe_data.nc.

The reference builds an xarray Dataset with variables on the dimensions
("particle number", "time"), coordinates "particle number" = 1..N and
"time" = start + n*dt, and calls to_netcdf (particle_advecter.py:268-303,
interaction_simulator.py:62-77,119-122).  xarray, netCDF4 and h5py are not
installed here, so the same names, dimensions, dtypes and coordinates are
written as NetCDF-3 (64-bit offset) through scipy.io.netcdf_file, with the
time coordinate CF-encoded the way xarray encodes datetimes
("<unit> since <first time>", int32 counts).
"""
from datetime import datetime, timedelta

import numpy as np
from scipy.io import netcdf_file

_UNITS = (("days", 86400), ("hours", 3600), ("minutes", 60), ("seconds", 1))


def encode_times(times):
    """datetimes -> (int32 counts, CF units) using the coarsest exact unit."""
    if len(times) == 0:
        return np.zeros(0, np.int32), "hours since 1970-01-01 00:00:00"
    ref = times[0]
    secs = np.array([(t - ref) // timedelta(seconds=1) for t in times], dtype=np.int64)
    for unit, size in _UNITS:
        if np.all(secs % size == 0):
            vals = secs // size
            break
    if np.abs(vals).max(initial=0) >= 2 ** 31:
        raise ValueError("time axis does not fit NetCDF-3 int32")
    return vals.astype(np.int32), f"{unit} since {ref.strftime('%Y-%m-%d %H:%M:%S')}"


def decode_times(values, units):
    unit, _, ref = units.partition(" since ")
    size = dict(_UNITS)[unit]
    ref = datetime.strptime(ref.strip(), "%Y-%m-%d %H:%M:%S")
    return [ref + timedelta(seconds=int(v) * size) for v in np.asarray(values)]


def write_particle_dataset(path, variables, times, n_particles):
    """variables: {name: array [n_particles, len(times)]} (float32 or int8)."""
    with netcdf_file(path, "w", version=2) as f:
        f.createDimension("particle number", n_particles)
        f.createDimension("time", len(times))
        pn = f.createVariable("particle number", "i", ("particle number",))
        pn[:] = np.arange(1, n_particles + 1, dtype=np.int32)
        vals, units = encode_times(times)
        tv = f.createVariable("time", "i", ("time",))
        tv[:] = vals
        tv.units = units
        tv.calendar = "proleptic_gregorian"
        for name, arr in variables.items():
            arr = np.asarray(arr)
            code = {np.dtype(np.float32): "f", np.dtype(np.int8): "b", np.dtype(np.float64): "d"}[arr.dtype]
            var = f.createVariable(name, code, ("particle number", "time"))
            var[:] = arr


def read_particle_dataset(path, names=None):
    """Returns ({name: array [N, Nt]}, times, n_particles)."""
    with netcdf_file(path, "r", mmap=False) as f:
        out = {}
        for name, var in f.variables.items():
            if var.dimensions == ("particle number", "time") and (names is None or name in names):
                arr = np.array(var[:])
                out[name] = arr.astype(arr.dtype.newbyteorder("="))  # NetCDF stores big-endian
        tv = f.variables["time"]
        units = tv.units.decode() if isinstance(tv.units, bytes) else tv.units
        times = decode_times(tv[:], units)
        n = f.dimensions["particle number"]
    return out, times, int(n)
