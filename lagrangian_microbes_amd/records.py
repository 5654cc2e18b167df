"""Step records from the device to the host, off the compute path.

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
import queue
import threading
import time

import torch

DEFAULT_RING_BYTES = 8 << 30


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
