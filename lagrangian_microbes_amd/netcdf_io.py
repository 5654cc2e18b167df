"""The on-disk outputs: particle_data.nc and microbe_data.nc.

The reference builds an xarray Dataset with variables on the dimensions
("particle number", "time"), coordinates "particle number" = 1..N and
"time" = start + n*dt, and calls to_netcdf (particle_advecter.py:268-303,
interaction_simulator.py:62-77,119-122).  xarray, netCDF4 and h5py are not
installed here, so this module writes the same names, dimensions, dtypes and
coordinates in the NetCDF classic file format, with its own writer and
reader:

- CDF-2 (64-bit offset) when every variable is under 2 GiB, readable by any
  NetCDF-3 reader (scipy.io.netcdf_file included);
- CDF-5 (64-bit data) otherwise: 8-byte lengths and sizes, so a variable can
  hold more than 4 GiB (4M particles x 2,880 float32 steps = 46 GB).

Coordinates are encoded the way xarray encodes them for each format: the
time coordinate as "<unit> since <first time>" counts in the coarsest exact
unit, int32 in CDF-2 (NetCDF-3 has no 64-bit integers) and int64 in CDF-5,
like the particle numbers.

Files are written in place: ``ParticleDataset`` lays out the header and the
fixed-size variables and preallocates the file; ``write_steps`` stores a
block of steps ([k, N], one row per step, as the device records them) into
the variable's columns through a memory map, and ``write_rows`` a block of
whole particles (rows of the variable, one contiguous span) with pwrite.
The format grammar
followed is the NetCDF classic/64-bit-offset/64-bit-data format
specification (magic "CDF" + version byte, numrecs, tagged dimension,
attribute and variable lists, big-endian values padded to 4 bytes).
"""
import concurrent.futures
import mmap
import os
import struct
from datetime import datetime, timedelta

import numpy as np

_UNITS = (("days", 86400), ("hours", 3600), ("minutes", 60), ("seconds", 1))

# external type codes of the format and their big-endian numpy dtypes
NC_BYTE, NC_CHAR, NC_SHORT, NC_INT, NC_FLOAT, NC_DOUBLE = 1, 2, 3, 4, 5, 6
NC_UBYTE, NC_USHORT, NC_UINT, NC_INT64, NC_UINT64 = 7, 8, 9, 10, 11
_NC_DTYPE = {NC_BYTE: ">i1", NC_CHAR: "S1", NC_SHORT: ">i2", NC_INT: ">i4", NC_FLOAT: ">f4", NC_DOUBLE: ">f8",
             NC_UBYTE: ">u1", NC_USHORT: ">u2", NC_UINT: ">u4", NC_INT64: ">i8", NC_UINT64: ">u8"}
_CDF5_ONLY = {NC_UBYTE, NC_USHORT, NC_UINT, NC_INT64, NC_UINT64}
_TAG_DIM, _TAG_VAR, _TAG_ATT = 0x0A, 0x0B, 0x0C

# variables start on this boundary (the format only needs 4; a page keeps the
# memory-mapped columns aligned)
_ALIGN = 4096
# a CDF-2 variable stays below 2^31 bytes: sizes are 32-bit there, and some
# readers treat them as signed
CDF2_VAR_LIMIT = 2 ** 31 - 4


def nc_type_of(dtype):
    dt = np.dtype(dtype)
    for code, s in _NC_DTYPE.items():
        if code != NC_CHAR and np.dtype(s).newbyteorder("=") == dt.newbyteorder("="):
            return code
    raise TypeError(f"no NetCDF type for {dt}")


def encode_times(times, int64=False):
    """datetimes -> (integer counts, CF units) using the coarsest exact unit."""
    if len(times) == 0:
        return np.zeros(0, np.int64 if int64 else np.int32), "hours since 1970-01-01 00:00:00"
    ref = times[0]
    secs = np.array([(t - ref) // timedelta(seconds=1) for t in times], dtype=np.int64)
    for unit, size in _UNITS:
        if np.all(secs % size == 0):
            vals = secs // size
            break
    if not int64 and np.abs(vals).max(initial=0) >= 2 ** 31:
        raise ValueError("time axis does not fit int32")
    return vals.astype(np.int64 if int64 else np.int32), f"{unit} since {ref.strftime('%Y-%m-%d %H:%M:%S')}"


def decode_times(values, units):
    unit, _, ref = units.partition(" since ")
    size = dict(_UNITS)[unit]
    ref = datetime.strptime(ref.strip(), "%Y-%m-%d %H:%M:%S")
    return [ref + timedelta(seconds=int(v) * size) for v in np.asarray(values)]


# ---------------------------------------------------------------------------
# header encoding
# ---------------------------------------------------------------------------
class _Enc:
    """Big-endian header fields; lengths and sizes are 4 bytes in CDF-1/2 and
    8 bytes in CDF-5, offsets 8 bytes in CDF-2/5."""

    def __init__(self, version):
        self.version = version
        self.buf = bytearray()

    def i32(self, x):
        self.buf += struct.pack(">i", x)

    def non_neg(self, x):
        self.buf += struct.pack(">q" if self.version == 5 else ">i", x)

    def offset(self, x):
        self.buf += struct.pack(">q" if self.version >= 2 else ">i", x)

    def pad(self):
        self.buf += b"\0" * (-len(self.buf) % 4)

    def name(self, s):
        b = s.encode("utf-8")
        self.non_neg(len(b))
        self.buf += b
        self.pad()

    def tag_list(self, tag, n):
        if n == 0:  # ABSENT: a zero tag and a zero count
            self.i32(0)
            self.non_neg(0)
        else:
            self.i32(tag)
            self.non_neg(n)

    def atts(self, atts):
        self.tag_list(_TAG_ATT, len(atts))
        for k, v in atts.items():
            self.name(k)
            if isinstance(v, str):
                b = v.encode("utf-8")
                self.i32(NC_CHAR)
                self.non_neg(len(b))
                self.buf += b
            else:
                a = np.atleast_1d(np.asarray(v))
                code = nc_type_of(a.dtype)
                self.i32(code)
                self.non_neg(a.size)
                self.buf += a.astype(_NC_DTYPE[code]).tobytes()
            self.pad()


class _Var:
    def __init__(self, name, dims, dtype, atts=None):
        self.name = name
        self.dims = tuple(dims)
        self.code = nc_type_of(dtype)
        self.atts = dict(atts or {})
        self.shape = None
        self.nbytes = 0
        self.begin = 0


def _layout(version, dims, variables, gatts):
    """Header bytes and each variable's begin offset (fixed-size variables
    only, one after another from the first aligned offset past the header)."""
    dim_index = {d: k for k, (d, _) in enumerate(dims)}
    sizes = dict(dims)
    for v in variables:
        v.shape = tuple(sizes[d] for d in v.dims)
        v.nbytes = int(np.prod(v.shape, dtype=np.int64)) * np.dtype(_NC_DTYPE[v.code]).itemsize

    def encode(begins):
        e = _Enc(version)
        e.buf += b"CDF" + bytes([version])
        e.non_neg(0)  # numrecs: no record dimension
        e.tag_list(_TAG_DIM, len(dims))
        for d, n in dims:
            e.name(d)
            e.non_neg(n)
        e.atts(gatts)
        e.tag_list(_TAG_VAR, len(variables))
        for v, b in zip(variables, begins):
            e.name(v.name)
            e.non_neg(len(v.dims))
            for d in v.dims:
                e.non_neg(dim_index[d])
            e.atts(v.atts)
            e.i32(v.code)
            e.non_neg(v.nbytes + (-v.nbytes % 4))  # vsize: padded to 4 bytes
            e.offset(b)
        return bytes(e.buf)

    head = encode([0] * len(variables))  # offsets have a fixed width: the length is final
    pos = len(head) + (-len(head) % _ALIGN)
    for v in variables:
        v.begin = pos
        pos += v.nbytes + (-v.nbytes % 4)
        pos += -pos % _ALIGN
    return encode([v.begin for v in variables]), pos


class ParticleDataset:
    """particle_data.nc / microbe_data.nc being written.

    Variables ``names`` (dtype per name: float32 longitude/latitude, int8
    species) on ("particle number", "time"), with the coordinates
    "particle number" = 1..n_particles and "time" = ``times``.  The header and
    coordinates are written and the file preallocated at once; the data
    columns are filled through memory maps with ``write_steps``.  version:
    2 or 5, or None for CDF-2 when every variable is under 2 GiB, else CDF-5.
    """

    def __init__(self, path, n_particles, times, names, version=None, threads=8):
        n, T = int(n_particles), len(times)
        per_step = {k: np.dtype(v) for k, v in names.items()}
        biggest = max([n * T * dt.itemsize for dt in per_step.values()] + [8 * max(n, T)])
        if version is None:
            version = 2 if biggest <= CDF2_VAR_LIMIT else 5
        if version not in (2, 5):
            raise ValueError("version must be 2 (64-bit offset) or 5 (64-bit data)")
        if version == 2 and biggest > CDF2_VAR_LIMIT:
            raise ValueError(f"a variable of {biggest} bytes needs the 64-bit data format (version 5)")
        self.path, self.version, self.n, self.times = path, version, n, list(times)
        int64 = version == 5
        tvals, units = encode_times(self.times, int64=int64)
        dims = [("particle number", n), ("time", T)]
        coord_dt = np.int64 if int64 else np.int32
        variables = [_Var("particle number", ("particle number",), coord_dt),
                     _Var("time", ("time",), coord_dt, {"units": units, "calendar": "proleptic_gregorian"})]
        variables += [_Var(k, ("particle number", "time"), dt) for k, dt in per_step.items()]
        head, total = _layout(version, dims, variables, {})
        with open(path, "wb") as fh:
            fh.write(head)
            fh.truncate(total)
        self._fh = open(path, "r+b")
        self._mm = mmap.mmap(self._fh.fileno(), total) if total else None
        self.vars = {}
        for v in variables:
            if v.nbytes:
                self.vars[v.name] = np.ndarray(v.shape, dtype=_NC_DTYPE[v.code], buffer=self._mm, offset=v.begin)
            else:
                self.vars[v.name] = np.zeros(v.shape, dtype=_NC_DTYPE[v.code])
        self.vars["particle number"][:] = np.arange(1, n + 1, dtype=coord_dt)
        self.vars["time"][:] = tvals
        self.layout = {v.name: (v.begin, v.nbytes) for v in variables}
        self._pool = concurrent.futures.ThreadPoolExecutor(threads) if threads > 1 else None
        self._threads = threads

    def write_steps(self, name, t0, block):
        """block: [k, n] (step-major, native byte order) -> columns t0..t0+k-1
        of variable `name`."""
        self.write_block(t0, {name: block})

    def write_block(self, t0, blocks):
        """{name: [k, n] block} -> columns t0..t0+k-1 of each variable.  The
        (variable, particle tile) pieces go to the worker threads together
        (numpy's strided copies release the GIL)."""
        jobs = []
        tile = 8192
        for name, block in blocks.items():
            block = np.asarray(block)
            k = block.shape[0]
            if block.ndim != 2 or block.shape[1] != self.n:
                raise ValueError(f"block must be [k, {self.n}] (got {block.shape})")
            dst = self.vars[name]
            if t0 < 0 or t0 + k > dst.shape[1]:
                raise ValueError(f"steps {t0}..{t0 + k} outside the {dst.shape[1]}-step time axis")
            jobs += [(dst, block, i0, min(self.n, i0 + tile), k) for i0 in range(0, self.n, tile)]

        def part(job):
            dst, block, i0, i1, k = job
            dst[i0:i1, t0:t0 + k] = block[:, i0:i1].T

        if self._pool is None:
            for job in jobs:
                part(job)
        else:
            list(self._pool.map(part, jobs))

    def write_rows(self, name, p0, rows):
        """rows: [m, T] values (particle-major; converted to the file's
        big-endian type) or their raw bytes already in the file's byte order
        (a uint8 array) -> rows p0..p0+m-1 of variable `name`: one contiguous
        span of the file, written with pwrite (no memory map; the GIL is
        released while it runs)."""
        dst = self.vars[name]
        T, es = dst.shape[1], dst.dtype.itemsize
        arr = np.asarray(rows)
        if arr.dtype != np.uint8:
            arr = np.ascontiguousarray(arr, dtype=dst.dtype)
        data = memoryview(np.ascontiguousarray(arr)).cast("B")
        if T == 0 or len(data) % (T * es) or p0 < 0 or p0 + len(data) // (T * es) > self.n:
            raise ValueError(f"{len(data)} bytes are not whole rows of {name} from particle {p0}")
        begin, _ = self.layout[name]
        off, done = begin + p0 * T * es, 0
        while done < len(data):
            got = os.pwrite(self._fh.fileno(), data[done:done + (1 << 30)], off + done)
            if got <= 0:
                raise OSError(f"pwrite stopped with {len(data) - done} bytes left")
            done += got

    def close(self):
        if self._pool is not None:
            self._pool.shutdown()
            self._pool = None
        self.vars = {}
        if self._mm is not None:
            # no msync: the written pages are the file's page cache, which
            # later reads see and the kernel writes back (as after write())
            try:
                self._mm.close()
            except BufferError:  # a caller still holds a view: the map closes with it
                pass
            self._mm = None
        if self._fh is not None:
            self._fh.close()
            self._fh = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def write_particle_dataset(path, variables, times, n_particles, version=None):
    """variables: {name: array [n_particles, len(times)]} (float32 or int8)."""
    with ParticleDataset(path, n_particles, times, {k: np.asarray(v).dtype for k, v in variables.items()},
                         version) as ds:
        for name, arr in variables.items():
            ds.vars[name][:] = np.asarray(arr)


def copy_variables(src, dst, names):
    """Copies the data of variables `names` from file src to file dst, where
    both were laid out with the same dimensions and these variables (a
    kernel-side copy: the bytes never pass through Python)."""
    a, b = read_header(src), read_header(dst)
    with open(src, "rb") as fs, open(dst, "r+b") as fd:
        for name in names:
            (sb, sn), (db, dn) = a["vars"][name]["span"], b["vars"][name]["span"]
            if sn != dn:
                raise ValueError(f"{name}: {sn} bytes in {src}, {dn} in {dst}")
            left, so, do = sn, sb, db
            while left > 0:
                got = os.copy_file_range(fs.fileno(), fd.fileno(), left, so, do)
                if got <= 0:
                    raise OSError(f"copy_file_range stopped with {left} bytes left")
                left, so, do = left - got, so + got, do + got


# ---------------------------------------------------------------------------
# reading
# ---------------------------------------------------------------------------
class _Dec:
    def __init__(self, buf):
        self.buf, self.pos = buf, 0
        if bytes(buf[:3]) != b"CDF" or buf[3] not in (1, 2, 5):
            raise ValueError("not a NetCDF classic, 64-bit offset or 64-bit data file")
        self.version = buf[3]
        self.pos = 4

    def take(self, fmt):
        v = struct.unpack_from(fmt, self.buf, self.pos)[0]
        self.pos += struct.calcsize(fmt)
        return v

    def i32(self):
        return self.take(">i")

    def non_neg(self):
        return self.take(">q" if self.version == 5 else ">i")

    def offset(self):
        return self.take(">q" if self.version >= 2 else ">i")

    def skip_pad(self):
        self.pos += -self.pos % 4

    def name(self):
        n = self.non_neg()
        s = bytes(self.buf[self.pos:self.pos + n]).decode("utf-8")
        self.pos += n
        self.skip_pad()
        return s

    def tag_list(self, tag):
        t, n = self.i32(), self.non_neg()
        if t not in (0, tag) or (t == 0 and n != 0):
            raise ValueError(f"malformed header at byte {self.pos}")
        return n

    def atts(self):
        out = {}
        for _ in range(self.tag_list(_TAG_ATT)):
            k = self.name()
            code, n = self.i32(), self.non_neg()
            dt = np.dtype(_NC_DTYPE[code])
            raw = bytes(self.buf[self.pos:self.pos + n * dt.itemsize])
            self.pos += n * dt.itemsize
            self.skip_pad()
            out[k] = raw.decode("utf-8") if code == NC_CHAR else np.frombuffer(raw, dt)
        return out


def read_header(path):
    """Dimensions, global attributes and variables (dims, type, attributes,
    byte span) of a NetCDF classic, 64-bit offset or 64-bit data file without
    a record dimension."""
    with open(path, "rb") as fh:
        size = os.fstat(fh.fileno()).st_size
        head = fh.read(min(size, 1 << 20))
    d = _Dec(head)
    if d.non_neg() != 0:
        raise ValueError("record variables are not supported")
    dims = [(d.name(), d.non_neg()) for _ in range(d.tag_list(_TAG_DIM))]
    gatts = d.atts()
    variables = {}
    for _ in range(d.tag_list(_TAG_VAR)):
        name = d.name()
        ids = [d.non_neg() for _ in range(d.non_neg())]
        atts = d.atts()
        code = d.i32()
        d.non_neg()  # vsize
        begin = d.offset()
        shape = tuple(dims[i][1] for i in ids)
        nbytes = int(np.prod(shape, dtype=np.int64)) * np.dtype(_NC_DTYPE[code]).itemsize
        if code in _CDF5_ONLY and d.version != 5:
            raise ValueError(f"{name}: type {code} needs the 64-bit data format")
        variables[name] = {"dims": tuple(dims[i][0] for i in ids), "shape": shape, "dtype": _NC_DTYPE[code],
                           "atts": atts, "span": (begin, nbytes)}
    return {"version": d.version, "dims": dict(dims), "atts": gatts, "vars": variables}


def open_particle_dataset(path):
    """(header, {name: read-only memory-mapped array}) of a file written by
    ``ParticleDataset`` (or any fixed-size NetCDF classic-format file)."""
    h = read_header(path)
    out = {}
    for name, v in h["vars"].items():
        b, nb = v["span"]
        out[name] = (np.memmap(path, dtype=v["dtype"], mode="r", offset=b, shape=v["shape"]) if nb
                     else np.zeros(v["shape"], dtype=v["dtype"]))
    return h, out


def read_particle_dataset(path, names=None):
    """Returns ({name: array [N, Nt]}, times, n_particles)."""
    h, arrays = open_particle_dataset(path)
    out = {}
    for name, v in h["vars"].items():
        if v["dims"] == ("particle number", "time") and (names is None or name in names):
            a = arrays[name]
            out[name] = np.asarray(a).astype(a.dtype.newbyteorder("="))
    tv = h["vars"]["time"]
    times = decode_times(np.asarray(arrays["time"]), tv["atts"]["units"])
    return out, times, int(h["dims"]["particle number"])
