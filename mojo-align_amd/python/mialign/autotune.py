"""Autotuner for the launch knobs of the map and sort kernels and the served
loop's launch size and slot count.

The reference hand-sets its batch knobs (fq2bam-meth/src/linear_gpu_locate.mojo:
1729-1751). Here every knob is a launch shape the library exposes
(ma_tune_set: radix tile size, reads per work-counter claim, persistent blocks
per CU, chunks per wave, the per-read-length-class split of uneven batches);
every value gives bit-identical records, order and duplicate flags (tests/
test_gpu_tune.py), so tuning is timing only. The search is coordinate descent
on the caller's own workload: each knob in turn over its candidates, the
others held at their best so far, each candidate timed with HIP events on the
stream the work runs on (best of `repeats`). Choices persist per device (its
name, CU count and the library version) in a JSON file under
$MIALIGN_TUNE_DIR (default ~/.cache/mialign) and load at the next run.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import re
import time
from pathlib import Path
from typing import Callable, Dict, Iterable, Optional

from ._lib import check, lib
from .core import device_info

# name -> (knob id, candidates in the order tried; the first is the built-in default)
KNOBS: Dict[str, tuple] = {
    "sort_tile_items": (0, (16, 8)),
    "map_max_claim": (1, (16, 8, 4, 32)),
    "map_blocks_per_cu": (2, (8, 6, 4, 12)),
    "map_chunks_per_wave": (3, (1, 2, 4)),
    "map_class_split": (4, (0, 1)),
}
MAP_KNOBS = ("map_class_split", "map_max_claim", "map_blocks_per_cu", "map_chunks_per_wave")
SORT_KNOBS = ("sort_tile_items",)


def get(name: str) -> int:
    v, d = C.c_int64(0), C.c_int64(0)
    check(lib().ma_tune_get(KNOBS[name][0], C.byref(v), C.byref(d)), "ma_tune_get")
    return int(v.value)


def default(name: str) -> int:
    v, d = C.c_int64(0), C.c_int64(0)
    check(lib().ma_tune_get(KNOBS[name][0], C.byref(v), C.byref(d)), "ma_tune_get")
    return int(d.value)


def set_knob(name: str, value: int) -> None:
    check(lib().ma_tune_set(KNOBS[name][0], int(value)), f"ma_tune_set({name}={value})")


def apply(values: Dict[str, int]) -> None:
    for k, v in values.items():
        if k in KNOBS:
            set_knob(k, v)


def reset() -> None:
    """Every knob back to its built-in default."""
    for k in KNOBS:
        set_knob(k, default(k))


def current() -> Dict[str, int]:
    return {k: get(k) for k in KNOBS}


def coordinate_descent(run: Callable[[], float], names: Iterable[str], repeats: int = 3,
                       candidates: Optional[Dict[str, tuple]] = None) -> dict:
    """Tune `names` for the device time `run()` returns (ms; it must leave the
    device idle). Returns the choices, the default and tuned times and every
    trial. The knobs are left at the chosen values."""
    names = list(names)
    cand = {k: (candidates or {}).get(k, KNOBS[k][1]) for k in names}

    def timed() -> float:
        return min(run() for _ in range(repeats))

    run()  # warm-up (buffers sized, code loaded)
    base = timed()
    best_t = base
    trials = []
    for k in names:
        keep = get(k)
        best_v = keep
        for v in cand[k]:
            if v == keep:
                continue
            set_knob(k, v)
            t = timed()
            trials.append({"knob": k, "value": v, "ms": round(t, 3)})
            if t < best_t * 0.99:  # a 1% margin keeps noise from flipping knobs
                best_t, best_v = t, v
        set_knob(k, best_v)
        trials.append({"knob": k, "value": keep, "ms": None, "kept": best_v == keep})
    return {"choices": {k: get(k) for k in names}, "default_ms": round(base, 3), "tuned_ms": round(best_t, 3),
            "trials": trials}


# ------------------------------------------------------------- persistence
def _device_key(device: int) -> str:
    info = device_info(device)
    name = re.sub(r"[^A-Za-z0-9]+", "_", info["name"]).strip("_") or "gpu"
    return f"{name}_{info['n_cu']}cu"


def cache_path(device: int = 0) -> Path:
    root = Path(os.environ.get("MIALIGN_TUNE_DIR", Path.home() / ".cache" / "mialign"))
    return root / f"tune_{_device_key(device)}.json"


def save(result: Dict[str, dict], device: int = 0, serve: Optional[dict] = None) -> Path:
    """Persist the current knob values (and served-loop choices) for this
    device; results and served-loop choices saved before for the same device
    and library are kept unless replaced."""
    p = cache_path(device)
    p.parent.mkdir(parents=True, exist_ok=True)
    old = load(device, apply_knobs=False) or {}
    results = dict(old.get("results", {}))
    results.update(result)
    doc = {"device": _device_key(device), "library": lib().ma_version().decode(), "saved": time.time(),
           "knobs": current(), "serve": serve if serve is not None else old.get("serve", {}), "results": results}
    tmp = p.with_suffix(".tmp")
    tmp.write_text(json.dumps(doc, indent=1))
    os.replace(tmp, p)
    return p


def load(device: int = 0, apply_knobs: bool = True) -> Optional[dict]:
    """The saved choices for this device and library version (applied unless
    apply_knobs is False), or None."""
    p = cache_path(device)
    try:
        doc = json.loads(p.read_text())
    except (OSError, ValueError):
        return None
    if doc.get("library") != lib().ma_version().decode() or doc.get("device") != _device_key(device):
        return None
    if apply_knobs:
        apply({k: int(v) for k, v in doc.get("knobs", {}).items()})
    return doc


# --------------------------------------------------------------- workloads
def tune_map(index, ws, stream, d_bases, stride: int, d_lens, n: int, max_len: int, out, *, encoding: int,
             bs_r1: int, params=None, repeats: int = 3, names=MAP_KNOBS) -> dict:
    """Map-kernel knobs on one resident batch (device arrays), timed with HIP
    events on `stream`."""
    from .core import Event, map_batch_dev

    ev0, ev1 = Event(), Event()

    def run() -> float:
        ev0.record(stream)
        map_batch_dev(index, ws, stream, d_bases, stride, d_lens, n, max_len, out, encoding=encoding,
                      bs_r1=bs_r1, params=params)
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_ms(ev1)

    return coordinate_descent(run, names, repeats)


def tune_sort(stream, n: int, d_coord, d_dhi, d_dlo, d_score, d_pair, d_is_dup, d_perm, d_dup_perm, d_scratch,
              scratch_bytes: int, repeats: int = 3) -> dict:
    """Radix tile size on one resident sort + markdup."""
    from .core import Event

    ev0, ev1 = Event(), Event()

    def run() -> float:
        ev0.record(stream)
        check(lib().ma_sort_markdup_dev(stream.handle, n, d_coord.ptr, d_dhi.ptr, d_dlo.ptr, d_score.ptr, d_pair.ptr,
                                        d_is_dup.ptr, d_perm.ptr, 1, d_dup_perm.ptr if d_dup_perm else None,
                                        d_scratch.ptr, scratch_bytes), "ma_sort_markdup_dev")
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_ms(ev1)

    return coordinate_descent(run, SORT_KNOBS, repeats)


def tune_serve(index, requests, run_loop: Callable, *, launch_reads=(65536, 131072, 262144), slots=(2, 3),
               max_p99_ratio: float = 2.0, min_throughput: float = 0.8) -> dict:
    """Served-loop launch size and slot count: `run_loop(launch_reads, slots)`
    serves `requests` and returns (reads/s, p99 ms) plus the plain loop's
    (reads/s, p99 ms); the choice is the lowest p99 among the settings that
    keep p99 <= max_p99_ratio x plain and throughput >= min_throughput x plain
    (the lowest p99 overall when none does)."""
    trials = []
    for lr in launch_reads:
        for sl in slots:
            thr, p99, plain_thr, plain_p99 = run_loop(lr, sl)
            trials.append({"launch_reads": lr, "slots": sl, "reads_per_s": round(thr, 1), "p99_ms": round(p99, 3),
                           "p99_ratio": round(p99 / plain_p99, 3), "throughput_ratio": round(thr / plain_thr, 3)})
    ok = [t for t in trials if t["p99_ratio"] <= max_p99_ratio and t["throughput_ratio"] >= min_throughput]
    best = min(ok or trials, key=lambda t: t["p99_ms"])
    return {"choice": {"launch_reads": best["launch_reads"], "slots": best["slots"]}, "trials": trials}
