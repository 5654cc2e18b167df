"""Velocity fields: the gridded u, v the advection samples.

The reference reads OSCAR surface currents over OPeNDAP
(velocity_fields.py:12-32) and hands parcels u[time][lat][lon] arrays on a
rectilinear lon/lat grid with times in seconds since the first level
(particle_advecter.py:163-180).  ``VelocityField`` is that contract as an
object; ``synthetic_fourier_field`` builds seeded eddy fields for offline runs;
``oscar_dataset`` keeps the reference name and reads a local OSCAR file
(NetCDF-3), but cannot fetch anything offline.
"""
import os

import numpy as np

M_PER_DEG = 1852.0 * 60.0  # metres of arc per degree


class VelocityField:
    """u, v in m/s on a uniform rectilinear lon/lat grid, with time levels.

    lons (nx), lats (ny): increasing (a decreasing axis, such as OSCAR's
    latitude, is flipped together with the data) and uniformly spaced.
    times: seconds, increasing, one entry per level; None or a single level
    means a steady field.  u, v: [nt, ny, nx] (or [ny, nx] when steady).
    """

    def __init__(self, lons, lats, u, v, times=None):
        lons = np.asarray(lons, dtype=np.float64)
        lats = np.asarray(lats, dtype=np.float64)
        u = np.asarray(u, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        if u.ndim == 2:
            u, v = u[None], v[None]
        if u.shape != v.shape or u.ndim != 3:
            raise ValueError("u and v must both be [nt, ny, nx]")
        if lons.ndim != 1 or lats.ndim != 1 or lons.size < 2 or lats.size < 2:
            raise ValueError("need at least 2 grid nodes along each axis")
        if lons[-1] < lons[0]:
            lons, u, v = lons[::-1], u[:, :, ::-1], v[:, :, ::-1]
        if lats[-1] < lats[0]:
            lats, u, v = lats[::-1], u[:, ::-1, :], v[:, ::-1, :]
        if u.shape[1:] != (lats.size, lons.size):
            raise ValueError(f"u is {u.shape[1:]} but the grid is ({lats.size}, {lons.size})")
        for name, ax in (("longitude", lons), ("latitude", lats)):
            d = np.diff(ax)
            if np.any(d <= 0) or np.max(np.abs(d - d.mean())) > 1e-9 * max(1.0, abs(d.mean())):
                raise ValueError(f"{name} axis must be uniformly spaced and increasing")
        times = np.zeros(1) if times is None else np.asarray(times, dtype=np.float64)
        if times.ndim != 1 or times.size != u.shape[0]:
            raise ValueError("one time per level is required")
        if times.size > 1 and np.any(np.diff(times) <= 0):
            raise ValueError("time levels must increase")
        self.lons, self.lats, self.times = lons, lats, times
        self.u = np.ascontiguousarray(u)
        self.v = np.ascontiguousarray(v)
        self.nt, self.ny, self.nx = u.shape
        self.x0, self.y0 = float(lons[0]), float(lats[0])
        self.dx = float((lons[-1] - lons[0]) / (self.nx - 1))
        self.dy = float((lats[-1] - lats[0]) / (self.ny - 1))

    @property
    def bounds(self):
        """(lon_min, lon_max, lat_min, lat_max) of the grid."""
        return float(self.lons[0]), float(self.lons[-1]), float(self.lats[0]), float(self.lats[-1])

    def interleaved(self):
        """[nt, ny, nx, 2] float64 with (u, v) adjacent per node (device layout)."""
        return np.ascontiguousarray(np.stack([self.u, self.v], axis=-1))

    def stage_schedule(self, t0, dt, nsteps):
        """Time levels and weights of the 2*nsteps+1 RK4 stage times.

        Stage k samples t0 + k*dt/2 (k = 0..2*nsteps).  Returns (lev0, w):
        the field there is level lev0 + w * (level lev0+1 - level lev0).  A time
        outside the levels raises, as parcels does without time
        extrapolation.
        """
        hh = 0.5 * float(dt)
        tau = float(t0) + np.arange(2 * int(nsteps) + 1, dtype=np.float64) * hh
        if self.nt == 1:
            return np.zeros(tau.size, np.int32), np.zeros(tau.size)
        lo, hi = self.times[0], self.times[-1]
        bad = (tau < lo) | (tau > hi)
        if np.any(bad):
            raise ValueError(f"time {tau[bad][0]} s lies outside the field's levels [{lo}, {hi}] s")
        lev = np.clip(np.searchsorted(self.times, tau, side="right") - 1, 0, self.nt - 2)
        w = (tau - self.times[lev]) / (self.times[lev + 1] - self.times[lev])
        return lev.astype(np.int32), w


def synthetic_fourier_field(lon_min, lon_max, lat_min, lat_max, spacing, n_modes, rms, seed, n_levels=1,
                            level_interval=None, wavelength_km=(50.0, 300.0)):
    """Seeded eddy field: each level a sum of n_modes random Fourier modes.

    Each level is built from a streamfunction
    psi = sum_m a_m sin(kx_m x + ky_m y + phi_m) on the local tangent plane of
    the grid centre (x east, y north, in metres, 1852*60 m per degree, x
    scaled by cos of the centre latitude), with u = -dpsi/dy, v = dpsi/dx:
    non-divergent, so particle density stays uniform on average.  Wavelengths
    are uniform in wavelength_km, directions and phases uniform on [0, 2 pi),
    amplitudes a_m = +-wavelength_m / (2 pi) with random sign, all from
    numpy.random.default_rng(seed), level after level.  Each level is scaled
    so that sqrt(mean(u^2 + v^2)) over its nodes equals rms (m/s).
    """
    nx = int(round((lon_max - lon_min) / spacing)) + 1
    ny = int(round((lat_max - lat_min) / spacing)) + 1
    lons = lon_min + spacing * np.arange(nx)
    lats = lat_min + spacing * np.arange(ny)
    lon_c, lat_c = 0.5 * (lon_min + lon_max), 0.5 * (lat_min + lat_max)
    X = (lons[None, :] - lon_c) * M_PER_DEG * np.cos(np.deg2rad(lat_c))
    Y = (lats[:, None] - lat_c) * M_PER_DEG
    rng = np.random.default_rng(seed)
    u = np.zeros((n_levels, ny, nx))
    v = np.zeros((n_levels, ny, nx))
    for lev in range(n_levels):
        lam = rng.uniform(wavelength_km[0], wavelength_km[1], n_modes) * 1e3
        theta = rng.uniform(0.0, 2.0 * np.pi, n_modes)
        phi = rng.uniform(0.0, 2.0 * np.pi, n_modes)
        sign = rng.choice([-1.0, 1.0], n_modes)
        k = 2.0 * np.pi / lam
        amp = sign * lam / (2.0 * np.pi)
        for m in range(n_modes):
            kx, ky = k[m] * np.cos(theta[m]), k[m] * np.sin(theta[m])
            c = np.cos(kx * X + ky * Y + phi[m])
            u[lev] -= amp[m] * ky * c
            v[lev] += amp[m] * kx * c
        s = rms / np.sqrt(np.mean(u[lev] ** 2 + v[lev] ** 2))
        u[lev] *= s
        v[lev] *= s
    times = None if n_levels == 1 else np.arange(n_levels, dtype=np.float64) * float(level_interval)
    return VelocityField(lons, lats, u, v, times)


def oscar_dataset_filename(year):
    return "oscar_vel" + str(year) + ".nc"


_TIME_UNIT_SECONDS = {"day": 86400.0, "hour": 3600.0, "minute": 60.0, "min": 60.0, "second": 1.0, "sec": 1.0,
                      "s": 1.0}


def _text(value):
    return value.decode("ascii", "replace") if isinstance(value, bytes) else str(value)


def _seconds_since_first(var):
    """CF time values ("<unit> since <epoch>") as seconds since the first one,
    the particle_advecter.py:170-173 conversion of the decoded datetimes."""
    units = _text(getattr(var, "units", "")).strip().lower()
    unit = units.split(" since ")[0].strip() if " since " in units else ""
    scale = _TIME_UNIT_SECONDS.get(unit[:-1] if unit.endswith("s") and unit[:-1] in _TIME_UNIT_SECONDS else unit)
    if scale is None:
        raise RuntimeError(f"time units {units!r} are not '<days|hours|minutes|seconds> since <epoch>'")
    t = np.ma.filled(np.ma.asarray(var[:], dtype=np.float64), np.nan).ravel()
    if t.size == 0 or not np.all(np.isfinite(t)):
        raise RuntimeError("the time axis is empty or has missing values")
    return (t - t[0]) * scale


def _axis_of(dims, names):
    for k, d in enumerate(dims):
        if _text(d).lower() in names:
            return k
    return None


def read_oscar_file(path):
    """The VelocityField parcels is handed from an OSCAR file.

    Follows particle_advecter.py:160-184: the first (nominal) depth only,
    u and v as [time, latitude, longitude], times in seconds since the first
    level.  What xarray and parcels do to the arrays on the way, restated:
    missing and fill values become NaN (xarray's CF decoding, scale_factor and
    add_offset applied), parcels' Field then sets NaN (land) to 0, and a
    decreasing latitude axis (OSCAR runs 80N to 80S) is flipped with the data.
    OSCAR's longitudes run 20 to 420 E on a uniform 1/3 degree grid and are
    kept as they are.

    Only NetCDF-3 files (classic or 64-bit offset) are readable here, through
    scipy.io.netcdf_file: there is no HDF5 reader for NetCDF-4.
    """
    with open(path, "rb") as fh:
        magic = fh.read(4)
    if magic[:3] != b"CDF" or magic[3:4] not in (b"\x01", b"\x02"):
        kind = "a NetCDF-4 (HDF5) file" if magic == b"\x89HDF" else "not a NetCDF-3 file"
        raise RuntimeError(f"{path} is {kind}: only NetCDF-3 (classic or 64-bit offset) is readable without "
                           "netCDF4/h5py; convert it with `nccopy -k classic` or pass a VelocityField")
    from scipy.io import netcdf_file

    with netcdf_file(path, "r", mmap=False, maskandscale=True) as nc:
        var = nc.variables
        missing = [k for k in ("time", "latitude", "longitude", "u", "v") if k not in var]
        if missing:
            raise RuntimeError(f"{path} has no variable(s) {missing} (OSCAR layout: time, depth, latitude, "
                               "longitude, u, v)")
        times = _seconds_since_first(var["time"])
        lats = np.asarray(var["latitude"][:], dtype=np.float64).ravel()
        lons = np.asarray(var["longitude"][:], dtype=np.float64).ravel()
        fields = []
        for name in ("u", "v"):
            x = var[name]
            dims = tuple(x.dimensions)
            data = np.ma.filled(np.ma.asarray(x[:], dtype=np.float64), np.nan)
            kt = _axis_of(dims, ("time",))
            ky = _axis_of(dims, ("latitude", "lat"))
            kx = _axis_of(dims, ("longitude", "lon"))
            kz = _axis_of(dims, ("depth", "z", "lev"))
            if None in (kt, ky, kx) or data.ndim != 3 + (kz is not None):
                raise RuntimeError(f"{name}{dims}: expected (time[, depth], latitude, longitude)")
            if kz is not None:
                data = np.take(data, 0, axis=kz)  # the nominal depth, particle_advecter.py:163-164
                kt, ky, kx = (k - (k > kz) for k in (kt, ky, kx))
            data = np.transpose(data, (kt, ky, kx))
            data[np.isnan(data)] = 0.0
            fields.append(data)
    return VelocityField(lons, lats, fields[0], fields[1], times if times.size > 1 else None)


def oscar_dataset(year, directory="."):
    """OSCAR surface currents of ``year`` from the local file
    ``oscar_vel{year}.nc`` in ``directory`` (the file velocity_fields.py:21-32
    caches and opens; the reference's working directory by default), as a
    VelocityField (``read_oscar_file``).  The reference's OPeNDAP download,
    used when that file is absent, needs the network: then this raises.
    """
    path = os.path.join(directory, oscar_dataset_filename(year))
    if not os.path.isfile(path):
        raise RuntimeError(
            f"OSCAR {year} is not available offline ({os.path.abspath(path)}): the reference downloads it over "
            "OPeNDAP; place the file there or pass a lagrangian_microbes_amd.VelocityField as velocity_field")
    return read_oscar_file(path)
