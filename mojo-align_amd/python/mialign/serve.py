"""Served request loop over a resident index (continuous batching).

The C side (csrc/serve.hip) owns `slots` streams; `Server.submit` queues a
read batch; whenever a slot is free the server assembles one launch from
waiting work in deadline order (large batches cut into pieces, small ones
merged) and enqueues its copy-in, map kernel and copy-out on that slot's
stream, so requests arriving while others run overlap with them on the
device. Device memory the server holds stays within `hbm_budget` (0: no cap
of its own) and is only taken when the device's free HBM, read at that
moment, covers it. A request with `keys` set is one sample's batch: once
mapped it is sorted and duplicate-marked on its own (`is_dup` and `perm` in
its outputs), never together with other samples' reads.
`run_request_loop` drives Poisson-like arrivals through it and reports
throughput and p50/p99 batch latency; `run_plain` is the baseline of one
batch per launch, strictly one after another.
"""

from __future__ import annotations

import ctypes as C
import gc
import time
import weakref
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._lib import MA_ERR_NOMEM, Params, check, default_params, host_recs, lib, ptr, recs_struct
from .core import GpuIndex

DEFAULT_LAUNCH_READS = 128 << 10
DEFAULT_SIZE_WEIGHT = 0.5


def pinned_empty(shape, dtype) -> np.ndarray:
    """numpy array in page-locked host memory (freed with the array), so
    copies to and from the device run asynchronously."""
    dtype = np.dtype(dtype)
    nbytes = max(int(np.prod(shape, dtype=np.int64)) * dtype.itemsize, 1)
    p = C.c_void_p()
    check(lib().ma_host_alloc(nbytes, C.byref(p)), "ma_host_alloc")
    buf = (C.c_uint8 * nbytes).from_address(p.value)
    weakref.finalize(buf, lib().ma_host_free, p.value)
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape, dtype=np.int64))).reshape(shape)


def pinned_copy(a: np.ndarray) -> np.ndarray:
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out


def pinned_recs(n: int) -> Dict[str, np.ndarray]:
    """Merged-record arrays as consecutive slices of one pinned block (one
    device-to-host copy per request)."""
    proto = host_recs(0)
    block = pinned_empty((len(proto) * max(n, 1) * 4,), np.uint8)
    return {f: block[i * n * 4 : (i + 1) * n * 4].view(a.dtype) for i, (f, a) in enumerate(proto.items())}


@dataclass
class Request:
    bases: np.ndarray  # [n, stride] uint8
    lens: np.ndarray  # [n] uint32
    max_len: int
    encoding: int
    bs_r1: int = 0
    bs_r2: int = 0
    n_r1: Optional[int] = None
    arrival: float = 0.0  # seconds from the start of the run
    out: Optional[Dict[str, np.ndarray]] = None  # preallocated (pinned) merged records
    keys: int = 0  # 0 map only; 1 sort the sample SE; 2 PE (R1 rows then R2 rows, n_r1 = n / 2)
    markdup: bool = True

    @property
    def n(self) -> int:
        return int(self.lens.size)


@dataclass
class _Flight:
    req: Request
    out: Dict[str, np.ndarray]
    struct: object
    t_submit: float


class Server:
    def __init__(self, index: GpuIndex, slots: int = 2, hbm_budget: int = 0, max_len: int = 256,
                 launch_reads: Optional[int] = None, size_weight: float = DEFAULT_SIZE_WEIGHT,
                 persistent: bool = False, chunks_per_wave: int = 4):
        h = C.c_void_p()
        check(lib().ma_serve_open(index.handle, slots, int(hbm_budget), max_len, C.byref(h)),
              "ma_serve_open")
        self.handle = h.value
        self.index = index
        self.inflight: Dict[int, _Flight] = {}
        self._buf = (C.c_int64 * 256)()
        check(lib().ma_serve_tune(self.handle, int(launch_reads or DEFAULT_LAUNCH_READS), float(size_weight),
                                  int(persistent), int(chunks_per_wave)), "ma_serve_tune")

    def submit(self, req: Request, params: Optional[Params] = None) -> Optional[int]:
        """Ticket, or None when the request is not admitted now (its device
        memory does not fit: submit it again once others complete). A sorted
        request's outputs get "is_dup" and "perm" beside the records."""
        params = params or default_params(k=self.index.k)
        out = req.out if req.out is not None else pinned_recs(req.n)
        st = recs_struct(out)
        n_r1 = req.n if req.n_r1 is None else req.n_r1
        head = (self.handle, ptr(req.bases), req.bases.shape[1], ptr(req.lens), req.n, req.max_len, req.encoding,
                req.bs_r1, req.bs_r2, n_r1, C.byref(params), C.byref(st))
        if req.keys:
            for f in ("is_dup", "perm"):
                if out.get(f) is None or out[f].size < req.n:
                    out[f] = pinned_empty((req.n,), np.uint32)
            t = lib().ma_serve_submit_sorted(*head, int(req.keys), int(req.markdup), ptr(out["is_dup"]),
                                             ptr(out["perm"]))
        else:
            t = lib().ma_serve_submit(*head)
        if t == MA_ERR_NOMEM:
            return None
        check(int(t), "ma_serve_submit")
        self.inflight[int(t)] = _Flight(req, out, st, time.perf_counter())
        return int(t)

    def poll(self, ticket: int, wait: bool = False):
        """(done, records) — records is the merged record dict once done."""
        lat = C.c_double(0)
        r = check(lib().ma_serve_poll(self.handle, ticket, int(wait), C.byref(lat)), "ma_serve_poll")
        if r == 0:
            return False, None
        fl = self.inflight.pop(ticket)
        self.last_latency_ms = lat.value
        return True, fl.out

    def reap(self, wait: bool = False) -> Dict[int, Dict[str, np.ndarray]]:
        """Finished tickets since the last call -> their merged records."""
        buf = self._buf
        out: Dict[int, Dict[str, np.ndarray]] = {}
        while True:
            n = check(lib().ma_serve_reap(self.handle, buf, 256, int(wait and not out)), "ma_serve_reap")
            for i in range(n):
                fl = self.inflight.pop(int(buf[i]))
                out[int(buf[i])] = fl.out
            if n < 256:
                return out

    def set_targets(self, tid_of_cid: np.ndarray) -> None:
        """Target of each index contig for sorted requests (-1: none)."""
        t = np.ascontiguousarray(tid_of_cid, dtype=np.int32)
        check(lib().ma_serve_set_targets(self.handle, ptr(t), int(t.size)), "ma_serve_set_targets")

    def memory(self) -> Dict[str, int]:
        """Device bytes held, the budget (0 = none) and the device's free HBM now."""
        h, b, f = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        check(lib().ma_serve_memory(self.handle, C.byref(h), C.byref(b), C.byref(f)), "ma_serve_memory")
        return {"held": h.value, "budget": b.value, "free_hbm": f.value}

    def close(self) -> None:
        if self.handle:
            lib().ma_serve_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class LoopStats:
    reads: int
    wall_s: float
    latencies_ms: List[float] = field(default_factory=list)

    @property
    def reads_per_s(self) -> float:
        return self.reads / self.wall_s if self.wall_s > 0 else 0.0

    def pct(self, q: float) -> float:
        return float(np.percentile(np.array(self.latencies_ms), q)) if self.latencies_ms else 0.0


def run_request_loop(server: Server, requests: Sequence[Request], params: Optional[Params] = None,
                     keep_outputs: bool = False):
    """Submit requests at their arrival times (the server queues and merges
    them as slots and HBM free up); latency of a request = completion time -
    arrival time, queueing included."""
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    active: Dict[int, int] = {}  # ticket -> request index
    lat = [0.0] * len(requests)
    outs: Dict[int, Dict[str, np.ndarray]] = {}
    gc_was = gc.isenabled()
    gc.disable()  # no collector pauses inside the timed loop
    t0 = time.perf_counter()
    try:
        _drive(server, requests, params, order, active, lat, outs, keep_outputs, t0)
    finally:
        if gc_was:
            gc.enable()
    wall = time.perf_counter() - t0
    st = LoopStats(sum(r.n for r in requests), wall, lat)
    return (st, outs) if keep_outputs else st


def _drive(server, requests, params, order, active, lat, outs, keep_outputs, t0) -> None:
    """Submit due requests, collect finished ones, sleep briefly when idle. A
    request not admitted (its device memory does not fit now) is submitted
    again after the next completion, in arrival order."""
    pi = 0
    done = 0
    while done < len(requests):
        now = time.perf_counter() - t0
        held_back = False
        while pi < len(order) and requests[order[pi]].arrival <= now:
            i = order[pi]
            t = server.submit(requests[i], params)
            if t is None:
                if not active:
                    raise RuntimeError(f"request {i} ({requests[i].n} reads) not admitted with nothing in flight: "
                                       f"{server.memory()}")
                held_back = True
                break
            active[t] = i
            pi += 1
        finished = server.reap(wait=held_back)
        tnow = time.perf_counter() - t0
        for t, rec in finished.items():
            i = active.pop(t)
            lat[i] = (tnow - requests[i].arrival) * 1000.0
            if keep_outputs:
                outs[i] = rec
            done += 1
        if not active and pi < len(order):
            dt = requests[order[pi]].arrival - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(min(dt, 0.0005))


def run_plain(index: GpuIndex, requests: Sequence[Request], params: Optional[Params] = None,
              hbm_budget: int = 0, tid: Optional[np.ndarray] = None, **shape) -> LoopStats:
    """Baseline: one batch per launch, strictly one after another (a one-slot
    server, waiting for each batch before the next); latency = that batch's
    own copy-in + map (+ sort for a sorted request) + copy-out time."""
    srv = Server(index, slots=1, hbm_budget=hbm_budget, launch_reads=1 << 30, **shape)
    if tid is not None:
        srv.set_targets(tid)
    lat = []
    gc_was = gc.isenabled()
    gc.disable()
    try:
        big = max(requests, key=lambda r: r.n)  # sizes the launch buffer before timing
        srv.poll(srv.submit(Request(big.bases, big.lens, big.max_len, big.encoding, big.bs_r1, big.bs_r2,
                                    big.n_r1, out=big.out, keys=big.keys, markdup=big.markdup), params), wait=True)
        t0 = time.perf_counter()
        for r in requests:
            t = time.perf_counter()
            tk = srv.submit(r, params)
            if tk is None:
                raise RuntimeError("plain launch: request larger than the HBM budget")
            srv.poll(tk, wait=True)
            lat.append((time.perf_counter() - t) * 1000.0)
    finally:
        if gc_was:
            gc.enable()
        srv.close()
    return LoopStats(sum(r.n for r in requests), time.perf_counter() - t0, lat)
