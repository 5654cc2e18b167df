"""Step records from the device to the host, off the compute path.

Two ways, chosen by the run's size:

- ``DeviceRecords``: a run whose records fit a device-memory budget keeps
  them in HBM ([T, n] arrays the chunks write in place), so nothing leaves
  the device while it runs; ``write_device_records`` then transposes each
  variable on the device into the file's particle-major rows and streams
  them out in large sequential writes;
- ``RecordStream``: larger runs move each chunk's records to the host while
  the run goes on, as below.

The reference collects every step's positions and species on the host
(particle_advecter.py:226-235,246-249; interaction_simulator.py:108-110) and
writes them once at the end (particle_advecter.py:262-303,
interaction_simulator.py:119-122).  ``RecordStream`` takes each chunk's
device records as the fused run produces them and moves them to the host
without stopping that run:

- a dedicated copy stream copies the chunk into a slot of a ring of pinned
  host buffers (at the default priority, like every stream here), after the
  work that wrote the records; ``submit`` returns the copy's event, which the
  run waits for before it writes that chunk buffer set again;
- a writer thread waits for each copy and hands the slot to a sink (the
  NetCDF writer), then frees the slot.

The ring is bounded (``ring_bytes``, at least two slots, one per chunk of
the run at most): the run only waits when every slot still waits for the
writer, so a run whose records fit the ring never waits, and a larger one
proceeds at the writer's pace instead of holding [N, steps] arrays in host
memory.
"""
import concurrent.futures
import queue
import threading
import time

import torch

DEFAULT_RING_BYTES = 8 << 30
DEFAULT_DEVICE_RECORD_BYTES = 64 << 30  # at most this much HBM (and a quarter of the device) holds a run's records
BLOCK_BYTES = 256 << 20                 # pinned host blocks of write_device_records


def device_record_budget(device, limit=None):
    """Bytes of records a run may keep on `device`: `limit` (default
    DEFAULT_DEVICE_RECORD_BYTES), at most a quarter of the device's memory
    and half of what is free now."""
    free, total = torch.cuda.mem_get_info(torch.device(device))
    cap = DEFAULT_DEVICE_RECORD_BYTES if limit is None else int(limit)
    return max(0, min(cap, total // 4, free // 2))


class DeviceRecords:
    """A run's step records resident on the device: longitude, latitude
    (float32) and species (int8) as [T, n] arrays (row t = step t, as the
    chunks write them) and the pair count of every step."""

    BYTES_PER_PARTICLE_STEP = 9

    def __init__(self, n, T, device):
        self.device = torch.device(device)
        self.n, self.T = int(n), int(T)
        self.longitude = torch.empty((self.T, self.n), dtype=torch.float32, device=self.device)
        self.latitude = torch.empty((self.T, self.n), dtype=torch.float32, device=self.device)
        self.species = torch.empty((self.T, self.n), dtype=torch.int8, device=self.device)
        self.npairs = torch.zeros(self.T, dtype=torch.int64, device=self.device)

    @classmethod
    def nbytes(cls, n, T):
        return cls.BYTES_PER_PARTICLE_STEP * int(n) * int(T) + 8 * int(T)

    @property
    def arrays(self):
        """(rec_x, rec_y, rec_s) for FusedRunner.run(records=...)."""
        return self.longitude, self.latitude, self.species


def write_device_records(arrays, files, block_bytes=None, threads=4):
    """Writes device arrays {name: [T, n] tensor (row t = step t)} into the
    particle-major variables of open ParticleDatasets: files = [(dataset,
    names it holds)].  Each array is cut into blocks of particles; a block
    is transposed on the device to [m, T] (the file's row order) and its
    values' bytes put in the file's big-endian order, then copied into
    one of two pinned host buffers and written with pwrite to every file
    that holds the variable, on worker threads, so the copy of the next
    block overlaps the writes of this one.  Returns the seconds spent
    waiting for copies and for writes.  block_bytes: size of each pinned
    buffer (default BLOCK_BYTES)."""
    block_bytes = BLOCK_BYTES if block_bytes is None else int(block_bytes)
    if not arrays:
        return 0.0, 0.0
    device = next(iter(arrays.values())).device
    bufs = [torch.empty(int(block_bytes), dtype=torch.uint8, pin_memory=True) for _ in range(2)]
    pending = [[], []]
    copy = torch.cuda.Stream(device)
    copy.wait_stream(torch.cuda.current_stream(device))  # the arrays as the caller left them
    wait_copy = wait_write = 0.0
    j = 0
    with concurrent.futures.ThreadPoolExecutor(threads) as pool:
        for name, X in arrays.items():
            T, n = X.shape
            es = X.element_size()
            targets = [ds for ds, names in files if name in names]
            if T == 0 or n == 0 or not targets:
                continue
            m = max(1, int(block_bytes) // (T * es))
            for p0 in range(0, n, m):
                p1 = min(n, p0 + m)
                slot = j % 2
                a = time.perf_counter()
                for f in pending[slot]:  # the buffer's previous block is written
                    f.result()
                wait_write += time.perf_counter() - a
                with torch.cuda.stream(copy):
                    blk = X[:, p0:p1].t().contiguous().view(-1).view(torch.uint8)
                    if es > 1:  # the file's values are big-endian: bytes reversed on the device
                        blk = blk.view(-1, es).flip(1).reshape(-1)
                    host = bufs[slot][:blk.numel()]
                    host.copy_(blk, non_blocking=True)
                    ev = torch.cuda.Event()
                    ev.record(copy)
                a = time.perf_counter()
                ev.synchronize()
                wait_copy += time.perf_counter() - a
                data = host.numpy()
                pending[slot] = [pool.submit(ds.write_rows, name, p0, data) for ds in targets]
                j += 1
        a = time.perf_counter()
        for f in pending[0] + pending[1]:
            f.result()
        wait_write += time.perf_counter() - a
    return wait_copy, wait_write


class RecordStream:
    """fields: {name: (dtype, per-step shape tuple)} of one chunk's records,
    e.g. {"longitude": (torch.float32, (n,)), "npairs": (torch.int64, ())};
    chunk: the most steps one submit carries; sink(t0, k, arrays) runs on the
    writer thread with host numpy views [k, ...] valid during the call;
    chunks: how many submits the run makes (no more slots are allocated)."""

    def __init__(self, fields, chunk, sink, device=None, chunks=None, ring_bytes=DEFAULT_RING_BYTES):
        self.device = torch.device(device or "cuda")
        self.fields = dict(fields)
        self.chunk = int(chunk)
        self.sink = sink
        per_step = sum(torch.empty((), dtype=dt).element_size() * _numel(shape) for dt, shape in self.fields.values())
        slot_bytes = max(1, per_step * self.chunk)
        nslots = max(2, int(ring_bytes) // slot_bytes)
        if chunks is not None:
            nslots = max(1, min(nslots, int(chunks)))
        self.slots = [{name: torch.empty((self.chunk,) + tuple(shape), dtype=dt, pin_memory=True)
                       for name, (dt, shape) in self.fields.items()} for _ in range(nslots)]
        self.copy = torch.cuda.Stream(self.device)
        self._free = queue.Queue()
        for s in self.slots:
            self._free.put(s)
        self._todo = queue.Queue()
        self._error = None
        self.wait_s = 0.0    # host time submit() spent waiting for a free slot
        self.write_s = 0.0   # writer-thread time in the sink
        self._thread = threading.Thread(target=self._drain, name="record-writer", daemon=True)
        self._thread.start()

    @property
    def ring_bytes(self):
        return sum(t.numel() * t.element_size() for s in self.slots for t in s.values())

    def submit(self, t0, k, tensors):
        """Queue steps t0..t0+k-1 (tensors: {name: device tensor [k, ...]},
        written by work on the current stream).  Returns the event after
        which those device tensors may be overwritten."""
        if self._error is not None:
            raise RuntimeError("record writer failed") from self._error
        if k > self.chunk:
            raise ValueError(f"{k} steps exceed the ring's slots of {self.chunk}")
        a = time.perf_counter()
        slot = self._free.get()
        self.wait_s += time.perf_counter() - a
        self.copy.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(self.copy):
            for name, t in tensors.items():
                slot[name][:k].copy_(t, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.copy)
        self._todo.put((slot, ev, int(t0), int(k)))
        return ev

    def _drain(self):
        while True:
            item = self._todo.get()
            if item is None:
                return
            slot, ev, t0, k = item
            try:
                if self._error is None:
                    ev.synchronize()
                    a = time.perf_counter()
                    self.sink(t0, k, {name: t[:k].numpy() for name, t in slot.items()})
                    self.write_s += time.perf_counter() - a
            except BaseException as e:  # reported by submit() / finish()
                self._error = e
            finally:
                self._free.put(slot)

    def finish(self):
        """Waits for every queued chunk to be written; raises the writer's error."""
        if self._thread is not None:
            self._todo.put(None)
            self._thread.join()
            self._thread = None
        if self._error is not None:
            raise RuntimeError("record writer failed") from self._error


def _numel(shape):
    n = 1
    for s in shape:
        n *= int(s)
    return n
