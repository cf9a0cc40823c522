"""gpu_sort_markdup with the reference's signature, plus the SE key builder.

gpu_sort_markdup(api, n, coord_addr, dup_hi_addr, dup_lo_addr, score_addr,
pair_addr, is_dup_addr, perm_addr, do_markdup, dup_perm_addr) takes host
addresses of caller-owned arrays (u64 coord/dhi/dlo, u32 score/pair/is_dup/
perm/dup_perm), like fq2bam-meth/src/linear_gpu_sort.mojo:24-41, and writes
is_dup (original order), perm (stable coordinate order) and dup_perm (stable
(dhi, dlo) order; identity when do_markdup is False). The whole sort and the
duplicate marking run on the device (no per-pass host scan, no host merge).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .core import DeviceArray, DeviceRecs, Stream

_API_DEVICE = {"hip": 0, "amd": 0, "rocm": 0, "auto": 0, "gpu": 0}


def gpu_sort_markdup(api: str, n: int, coord_addr: int, dup_hi_addr: int, dup_lo_addr: int,
                     score_addr: int, pair_addr: int, is_dup_addr: int, perm_addr: int,
                     do_markdup: bool, dup_perm_addr: int = 0) -> None:
    if n <= 0:
        return
    dev = _API_DEVICE.get(str(api).lower().split(":")[0], 0)
    if ":" in str(api):
        dev = int(str(api).split(":")[1])
    check(
        lib().ma_sort_markdup(dev, int(n), coord_addr, dup_hi_addr, dup_lo_addr, score_addr, pair_addr,
                              is_dup_addr, perm_addr, int(bool(do_markdup)), dup_perm_addr or None),
        "gpu_sort_markdup",
    )


def sort_keys_se(recs: dict, lens: np.ndarray, tid_of_cid: np.ndarray, pair_base: int = 0,
                 device: int = 0):
    """SE sort keys (coord, dhi, dlo, score, pair) for merged records, computed
    on the device (linear_gpu_fm.mojo:371-405, 800-823)."""
    n = int(lens.size)
    st = Stream(device)
    d = DeviceRecs(n, device)
    raw = np.zeros(d.buf.nbytes, dtype=np.uint8)
    n1 = max(n, 1)
    for i, f in enumerate(("cid", "pos", "mapq", "flag", "sl", "sr", "nm")):
        raw[i * n1 * 4 : i * n1 * 4 + n * 4] = np.ascontiguousarray(recs[f]).view(np.uint8)
    d.buf.upload(raw)
    d_lens = DeviceArray.from_host(np.ascontiguousarray(lens, dtype=np.uint32), device)
    d_tid = DeviceArray.from_host(np.ascontiguousarray(tid_of_cid, dtype=np.int32), device)
    outs = [DeviceArray((n1,), dt, device) for dt in (np.uint64, np.uint64, np.uint64, np.uint32, np.uint32)]
    check(lib().ma_sort_keys_se_dev(st.handle, n, C.byref(d.struct), d_lens.ptr, d_tid.ptr, d_tid.shape[0],
                                    pair_base,
                                    *[o.ptr for o in outs]), "ma_sort_keys_se_dev")
    st.sync()
    return tuple(o.download()[:n] for o in outs)
