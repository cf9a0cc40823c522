"""ma_map_batch_dev_noalloc (no workspace; its work counter lives in the
library's static device scratch) against ma_map_batch_dev with a persistent
workspace, through the C ABI: two calls in a row on one stream, the second
with a smaller batch (the counter must be cleared between them), all ten
record fields of the merged, FW and RC outputs bit for bit."""

import ctypes as C

import numpy as np
import pytest

import mialign
from mialign import synth
from mialign._lib import MA_BS_C2T, REC_FIELDS, check, default_params, lib

pytestmark = pytest.mark.gpu


def _map(gix, st, reads, max_len, noalloc, ws=None):
    n = reads.n
    d_b = mialign.DeviceArray.from_host(np.ascontiguousarray(reads.bases))
    d_l = mialign.DeviceArray.from_host(np.ascontiguousarray(reads.lens, dtype=np.uint32))
    outs = [mialign.DeviceRecs(n) for _ in range(3)]
    prm = default_params(k=21)
    args = (d_b.ptr, reads.bases.shape[1], d_l.ptr, n, max_len, 0, MA_BS_C2T, MA_BS_C2T, n, C.byref(prm),
            C.byref(outs[0].struct), C.byref(outs[1].struct), C.byref(outs[2].struct))
    if noalloc:
        check(lib().ma_map_batch_dev_noalloc(gix.handle, st.handle, *args), "ma_map_batch_dev_noalloc")
    else:
        check(lib().ma_map_batch_dev(gix.handle, ws.handle, st.handle, *args), "ma_map_batch_dev")
    st.sync()
    return [o.download() for o in outs]


def test_noalloc_map_equals_workspace_map():
    print("test_noalloc_map_equals_workspace_map", flush=True)
    ref = synth.gen_reference(81, 2, 1_000_000)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    st = mialign.Stream(0)
    big = synth.gen_reads(82, ref, 60_000, 50, 150, max_len=150)
    small = synth.gen_reads(83, ref, 7_000, 50, 150, max_len=150)
    ws = mialign.Workspace(gix, 60_000, 150, persistent=True)
    for name, reads in (("60k reads", big), ("7k reads", small)):
        got = _map(gix, st, reads, 150, True)
        want = _map(gix, st, reads, 150, False, ws)
        for which, g, w in zip(("merged", "fw", "rc"), got, want):
            for f in REC_FIELDS:
                bad = np.flatnonzero(g[f] != w[f])
                assert bad.size == 0, f"{name} {which} {f}: {bad.size} reads differ (first {bad[:5]})"
        assert (got[0]["cid"] >= 0).mean() > 0.95
    gix.close()
