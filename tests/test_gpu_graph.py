"""The step's stage chain (map -> SE keys -> sort + markdup) captured as one
HIP graph of kernel nodes and replayed: every replay's records, keys,
coordinate order, duplicate order and flags equal the eager chain's, with
eager work, cleared outputs and synchronisations between the replays; with
the per-length-class split of uneven reads too (its partition kernels inside
the graph)."""

import ctypes as C

import numpy as np
import pytest

import mialign
from mialign import autotune, synth
from mialign._lib import MA_BS_C2T, MA_ENC_2BIT, REC_FIELDS, check, lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[test] {msg}", flush=True)


class Chain:
    def __init__(self, gix, packed, lens, max_len, tid):
        self.gix, self.n, self.max_len, self.stride = gix, lens.size, max_len, packed.shape[1]
        self.st = mialign.Stream(0)
        self.d_b = mialign.DeviceArray.from_host(packed)
        self.d_l = mialign.DeviceArray.from_host(lens)
        self.d_tid = mialign.DeviceArray.from_host(tid)
        self.n_tid = int(tid.size)
        self.recs = mialign.DeviceRecs(self.n)
        self.ws = mialign.Workspace(gix, self.n, max_len, persistent=True)
        self.keys = [mialign.DeviceArray((self.n,), dt) for dt in (np.uint64, np.uint64, np.uint64, np.uint32, np.uint32)]
        self.outs = [mialign.DeviceArray((self.n,), np.uint32) for _ in range(3)]
        self.sb = int(lib().ma_sort_scratch_bytes(self.n))
        self.scr = mialign.DeviceArray((self.sb,), np.uint8)

    def step(self):
        st = self.st
        mialign.map_batch_dev(self.gix, self.ws, st, self.d_b, self.stride, self.d_l, self.n, self.max_len, self.recs,
                              encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T)
        check(lib().ma_sort_keys_se_dev(st.handle, self.n, C.byref(self.recs.struct), self.d_l.ptr, self.d_tid.ptr,
                                        self.n_tid, 0, *[k.ptr for k in self.keys]), "keys")
        k = self.keys
        check(lib().ma_sort_markdup_dev(st.handle, self.n, k[0].ptr, k[1].ptr, k[2].ptr, k[3].ptr, k[4].ptr,
                                        self.outs[0].ptr, self.outs[1].ptr, 1, self.outs[2].ptr, self.scr.ptr,
                                        self.sb), "sort")

    def results(self):
        self.st.sync()
        r = self.recs.download()
        return [r[f] for f in REC_FIELDS] + [a.download() for a in self.keys + self.outs]

    def clear(self):
        self.recs.buf.upload(np.zeros(self.recs.buf.nbytes, np.uint8))
        for a in self.keys + self.outs:
            a.upload(np.zeros(a.shape, a.dtype))


@pytest.mark.parametrize("split", [0, 1])
def test_step_chain_graph_equals_eager(split):
    say(f"test_step_chain_graph_equals_eager split={split}")
    ref = synth.gen_reference(95, 3, 1_500_000, repeat_frac=0.03)
    dual = synth.dual_convert(ref)
    gix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21)
    reads = synth.gen_reads(96, ref, 120_000, 50, 250, max_len=250)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    tid = synth.tid_of_cid(dual.names)
    other = synth.gen_reads(97, ref, 5_000, 150, 150)
    autotune.apply({"map_class_split": split})
    try:
        ch = Chain(gix, packed, reads.lens, 250, tid)
        ch.step()
        eager = ch.results()
        assert (eager[0] >= 0).mean() > 0.9 and eager[-3].sum() > 0  # mapped, duplicates flagged
        with mialign.Graph.capture(ch.st) as cap:
            ch.step()
        g = cap.graph
        say(f"graph nodes: {g.n_nodes}")
        assert g.n_nodes >= 20
        for rep in range(3):
            ch.clear()
            mialign.map_batch(gix, other.bases, other.lens, 150, bs_r1=MA_BS_C2T)  # eager work between replays
            g.launch(ch.st)
            got = ch.results()
            for i, (a, b) in enumerate(zip(got, eager)):
                assert np.array_equal(a, b), f"replay {rep}: output {i} differs"
    finally:
        autotune.reset()
        gix.close()
