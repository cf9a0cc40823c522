"""Benchmark: reads/s of the seed -> extend -> sort path on one MI355X.

Headline workload: 150 bp WGBS reads against a 3.1 Gbp synthetic reference
(24 contigs, CpG depleted to ~1% of dinucleotides, 5% low-complexity repeat
blocks), dual C->T / G->A conversion (6.2 Gbp indexed), k=21 index built on
the device and resident in HBM; 10M single-end bisulfite reads (non-CpG C->T
p=0.99, CpG kept p=0.7, 0.5% substitutions, 5% planted duplicates), 2-bit
packed u8 input resident in HBM. One step = map every read (in-silico C->T,
seeds, vote, extend, rescue, high-occ tally, FW + RC, merge) + SE sort keys +
stable coordinate sort + duplicate marking, all on the device.

Prints one JSON line. `roofline` is for the dominant kernel (the map kernel),
priced with the algorithmic bytes of the oracle's counts on a sample;
`cpu_baseline` times the oracle (the C restatement of the reference's CPU
science, OpenMP over reads + the stable-sort markdup) on a bounded sample of
the same reads on this host, on every thread and on one. `parity_sample`
compares that sample's records; `sort_parity` the step's keys, coordinate
order and duplicate flags of every record with the oracle's. Extra sections:
`uneven_50_250bp` (50M reads of 50-250 bp in one persistent launch),
`serving` (the served request loop on the same index against plain
launches), `sort_stress` (200M sort + markdup records) and `config2_100mbp`
(a 100 Mbp reference, 10M x 150 bp, with the sections built on its records:
`indel_mode`, `bam_emit`, `pe_emit`, `fastq_ingest`, `fq2bam`).
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "mojo-align_amd" / "python"))
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

METRIC = "reads/s seed+extend+sort (bit-exact); served p50/p99 latency vs plain launch"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def serve_section(ix, packed, lens, L, n_req: int, slots: int, launch_reads: int, size_weight: float,
                  loads=(0.85, 1.0, 1.2), seeds=5, tid=None, samples: int = 64, tune: bool = False) -> dict:
    """Request loop: n_req batches (sizes lognormal, median 50k, clipped to
    [1k, 2M] reads) cut from the resident workload's reads (one pinned pool),
    each from one of `samples` samples (request i from sample i % samples),
    Poisson arrivals offered at each of `loads` x the plain loop's throughput.
    With a target table `tid`, every batch is also sorted and duplicate-marked
    on its own once mapped (SE keys, per sample: never over the device's mix).
    Admission: no budget of the server's own; it takes device memory only
    within the free HBM the driver reports at that moment.
    Plain = the same batches, one per launch, one after another (sorted too).
    With `tune`, launch_reads and slots are picked by autotune.tune_serve on
    arrival sequences of their own (not the measured ones) at the first load,
    and saved with the device's other choices.
    Each load runs `seeds` arrival sequences; its p50 / p99 are the medians of
    the sequences' own p50 / p99 and its throughput their mean (the
    per-sequence p99s are listed too).
    The headline figures are the first load's."""
    from mialign import serve
    from mialign._lib import MA_BS_C2T, MA_ENC_2BIT

    rng = np.random.default_rng(1003)
    n_pool = lens.size
    sizes = np.clip(np.exp(rng.normal(np.log(50_000), 1.0, n_req)), 1_000, min(2_000_000, n_pool)).astype(int)
    pool_b, pool_l = serve.pinned_copy(packed), serve.pinned_copy(lens)
    keys = 1 if tid is not None else 0
    reqs = []
    for sz in sizes:
        a = int(rng.integers(0, n_pool - sz + 1))
        reqs.append(serve.Request(pool_b[a : a + sz], pool_l[a : a + sz], L, MA_ENC_2BIT, bs_r1=MA_BS_C2T,
                                  out=serve.pinned_recs(int(sz)), keys=keys))
    budget = 0  # live free HBM only
    serve.run_plain(ix, reqs[:4], hbm_budget=budget, tid=tid)  # warm-up
    plain = serve.run_plain(ix, reqs, hbm_budget=budget, tid=tid)
    warm = [serve.Request(r.bases, r.lens, r.max_len, r.encoding, bs_r1=r.bs_r1, out=r.out, keys=r.keys)
            for r in reqs[:4]]

    def open_server(lr, sl):
        srv = serve.Server(ix, slots=sl, hbm_budget=budget, launch_reads=lr, size_weight=size_weight)
        if tid is not None:
            srv.set_targets(tid)
        serve.run_request_loop(srv, warm)  # same warm-up as the plain loop got
        return srv

    def arrivals(load, seed):
        arr = np.random.default_rng(seed)
        rate = load * plain.reads_per_s / float(sizes.mean())  # requests per second
        t = 0.0
        for r in reqs:
            t += float(arr.exponential(1.0 / rate))
            r.arrival = t

    tuning = None
    if tune:
        from mialign import autotune

        def trial(lr, sl):
            """The median p99 and the mean throughput of three arrival sequences
            (the p99 of one sequence of 512 requests is its 6th-worst latency)."""
            srv = open_server(lr, sl)
            got = []
            for k in range(3):
                arrivals(loads[0], 90_000 + 1000 * k)
                st = serve.run_request_loop(srv, reqs)
                got.append((st.reads_per_s, st.pct(99)))
            srv.close()
            thr, p99 = float(np.mean([g[0] for g in got])), float(np.median([g[1] for g in got]))
            log(f"serve tuning: launch_reads {lr} slots {sl}: p99 {p99:.1f} ms ({p99 / plain.pct(99):.2f} x plain)")
            return thr, p99, plain.reads_per_s, plain.pct(99)

        t0 = time.time()
        tuning = autotune.tune_serve(ix, reqs, trial, launch_reads=(131072, 262144, 524288), slots=(2, 4))
        tuning["search_s"] = round(time.time() - t0, 1)
        launch_reads, slots = tuning["choice"]["launch_reads"], tuning["choice"]["slots"]
        autotune.save({}, serve=tuning["choice"])
        log(f"serve tuning: chose {tuning['choice']} in {tuning['search_s']} s")
    srv = open_server(launch_reads, slots)
    sweep = []
    for load in loads:
        runs = []
        for k in range(seeds):
            arrivals(load, 2000 + int(load * 100) + 1000 * k)
            runs.append(serve.run_request_loop(srv, reqs))
        p50 = float(np.median([st.pct(50) for st in runs]))
        p99 = float(np.median([st.pct(99) for st in runs]))
        thr = float(np.mean([st.reads_per_s for st in runs]))
        sweep.append({
            "offered_load": load,
            "served_reads_per_s": round(thr, 1),
            "served_p50_ms": round(p50, 3),
            "served_p99_ms": round(p99, 3),
            "p99_ratio": round(p99 / plain.pct(99), 3) if plain.pct(99) else None,
            "throughput_ratio": round(thr / plain.reads_per_s, 3),
            "p99_ms_by_sequence": [round(st.pct(99), 3) for st in runs],
        })
        log(f"serving at {load:.2f}x plain: {sweep[-1]}")
    mem = srv.memory()
    srv.close()
    head = sweep[0]
    return {
        "requests": n_req,
        "samples": samples,
        "per_sample_sort_markdup": bool(keys),
        "reads": int(sizes.sum()),
        "batch_reads_median": int(np.median(sizes)),
        "offered_load": f"{loads[0]:.0%} of plain throughput, Poisson arrivals, median of {seeds} sequences",
        "served_reads_per_s": head["served_reads_per_s"],
        "served_p50_ms": head["served_p50_ms"],
        "served_p99_ms": head["served_p99_ms"],
        "plain_reads_per_s": round(plain.reads_per_s, 1),
        "plain_p50_ms": round(plain.pct(50), 3),
        "plain_p99_ms": round(plain.pct(99), 3),
        "p99_ratio": head["p99_ratio"],
        "throughput_ratio": head["throughput_ratio"],
        "load_sweep": sweep,
        "slots": slots,
        "launch_reads": launch_reads,
        "size_weight": size_weight,
        "tuning": tuning,
        "admission": "live free HBM (hipMemGetInfo when the server takes memory; 256 MiB headroom), no fixed budget",
        "device_bytes_held": mem["held"],
        "free_hbm_bytes_after": mem["free_hbm"],
    }


# HBM bytes one map launch moves at the default workload (100 Mbp x2 reference,
# 10M x 150 bp reads): rocprofv3 --pmc FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00, the
# factors read off copy_u64_kernel (a copy of known size) in the same runs.
MAP_TRAFFIC_PROFILED = {"genome_bp": 100_000_000, "reads": 10_000_000, "read_len": 150,
                        "bytes": 85.9e9, "kernel_ms": 38.9,
                        "source": "profiles/map_reads_kernel_pmc_100Mbp_10Mx150.md"}
SORT_BYTES_PER_RECORD = 170  # 2 x 13 B x 6 digit passes (45-bit key) + 13 B markdup read + 1 B flag


def calibration_copy(n: int = 10_000_000) -> None:
    """A plain device copy of known size for the counters' calibration: a
    one-pass radix sort of n u64 keys below 256 ends in its scratch buffer and
    copies the keys back with copy_u64_kernel (n x 8 bytes read and written)."""
    import mialign
    from mialign._lib import check, lib

    keys = mialign.DeviceArray.from_host(np.random.default_rng(5).integers(0, 256, n).astype(np.uint64))
    vals = mialign.DeviceArray((n,), np.uint32)
    sb = int(lib().ma_sort_scratch_bytes(n))
    scr = mialign.DeviceArray((sb,), np.uint8)
    st = mialign.Stream(0)
    check(lib().ma_radix_sort_u64_dev(st.handle, n, keys.ptr, vals.ptr, scr.ptr, sb), "ma_radix_sort_u64_dev")
    st.sync()
    log(f"calibration copy: {n * 8} bytes by copy_u64_kernel")


def sort_section(n: int, cpu_sample: int, with_cpu: bool) -> dict:
    """Sort + markdup stress: n records (key = contig<<40 | pos<<1 | strand over
    3.1 Gbp and 24 contigs, 12% in duplicate clusters of 2-8, u32 id, u8 nm),
    device resident; coord = dup key = key, score = 255 - nm, pair = id."""
    import mialign
    from mialign import synth
    from mialign._lib import check, lib

    t0 = time.time()
    key, rid, nm = synth.gen_sort_records(1004, n)
    score = (255 - nm.astype(np.uint32)).astype(np.uint32)
    log(f"sort stress: {n} records generated in {time.time() - t0:.1f}s")
    dev = 0
    st = mialign.Stream(dev)
    d_key = mialign.DeviceArray.from_host(key, dev)
    d_dlo = mialign.DeviceArray.from_host(np.zeros(n, np.uint64), dev)
    d_score = mialign.DeviceArray.from_host(score, dev)
    d_pair = mialign.DeviceArray.from_host(rid, dev)
    outs = {f: mialign.DeviceArray((n,), np.uint32, dev) for f in ("is_dup", "perm", "dup_perm")}
    sb = int(lib().ma_sort_scratch_bytes(n))
    scratch = mialign.DeviceArray((sb,), np.uint8, dev)
    ev_a, ev_b = mialign.Event(), mialign.Event()

    def run():
        check(lib().ma_sort_markdup_dev(st.handle, n, d_key.ptr, d_key.ptr, d_dlo.ptr, d_score.ptr, d_pair.ptr,
                                        outs["is_dup"].ptr, outs["perm"].ptr, 1, outs["dup_perm"].ptr,
                                        scratch.ptr, sb), "ma_sort_markdup_dev")

    run()
    st.sync()
    ms = []
    for _ in range(3):
        ev_a.record(st)
        run()
        ev_b.record(st)
        ev_b.synchronize()
        ms.append(ev_a.elapsed_ms(ev_b))
    t = float(np.mean(ms))
    n_dup = int(outs["is_dup"].download().sum())
    achieved = SORT_BYTES_PER_RECORD * n / (t / 1000.0) / 1e9
    out = {
        "workload": "sort + markdup stress: records over 3.1 Gbp / 24 contigs, 12% in duplicate clusters of 2-8",
        "records": n, "ms": round(t, 3), "records_per_s": round(n / t * 1000.0, 1), "duplicates": n_dup,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes_per_record": SORT_BYTES_PER_RECORD},
    }
    log(f"sort stress: {t:.2f} ms for {n} records ({n / t / 1e3:.1f} M records/s), {n_dup} duplicates")
    if with_cpu:
        from oracle import oracle as orc

        ns = min(cpu_sample, n)
        k2, r2, s2 = key[:ns].copy(), rid[:ns].copy(), score[:ns].copy()
        t0 = time.perf_counter()
        orc.sort_markdup(k2, k2, np.zeros(ns, np.uint64), s2, r2)
        cs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(ns / cs, 1), "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"first {ns} records, oracle stable sorts + markdup scan, {cs:.1f} s"}
        log(f"sort stress cpu baseline {ns / cs:.0f} records/s")
    return out


class StepRunner:
    """One bench step over resident reads: map (FW + RC + merge), SE sort keys,
    stable sorts + duplicate marking, all on one stream; the map kernel's time
    from HIP events on that stream."""

    def __init__(self, ix, packed, lens, L, tid, persistent=True):
        import ctypes as C
        import mialign
        from mialign._lib import lib

        self.C, self.mialign, self.lib = C, mialign, lib
        self.ix, self.L, self.n, self.stride = ix, L, lens.size, packed.shape[1]
        dev = 0
        self.stream = mialign.Stream(dev)
        self.d_bases = mialign.DeviceArray.from_host(packed, dev)
        self.d_lens = mialign.DeviceArray.from_host(lens, dev)
        self.d_tid = mialign.DeviceArray.from_host(tid, dev)
        self.n_tid = int(tid.size)
        self.recs = mialign.DeviceRecs(self.n, dev)
        self.ws = mialign.Workspace(ix, self.n, L, persistent=persistent)
        self.prm = mialign.default_params(k=21)
        self.keys = {f: mialign.DeviceArray((self.n,), dt, dev) for f, dt in
                     (("coord", np.uint64), ("dhi", np.uint64), ("dlo", np.uint64), ("score", np.uint32),
                      ("pair", np.uint32))}
        self.outs = {f: mialign.DeviceArray((self.n,), np.uint32, dev) for f in ("is_dup", "perm", "dup_perm")}
        self.scratch_bytes = int(lib().ma_sort_scratch_bytes(self.n))
        self.scratch = mialign.DeviceArray((self.scratch_bytes,), np.uint8, dev)
        self.ev_a, self.ev_b = mialign.Event(), mialign.Event()

    def step(self, timed_kernel: bool) -> None:
        from mialign._lib import MA_BS_C2T, MA_ENC_2BIT, check

        C, lib, n, st, keys, outs = self.C, self.lib, self.n, self.stream, self.keys, self.outs
        if timed_kernel:
            self.ev_a.record(st)
        self.mialign.map_batch_dev(self.ix, self.ws, st, self.d_bases, self.stride, self.d_lens, n, self.L,
                                   self.recs, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T, params=self.prm)
        if timed_kernel:
            self.ev_b.record(st)
        check(lib().ma_sort_keys_se_dev(st.handle, n, C.byref(self.recs.struct), self.d_lens.ptr, self.d_tid.ptr,
                                        self.n_tid, 0, keys["coord"].ptr, keys["dhi"].ptr, keys["dlo"].ptr,
                                        keys["score"].ptr, keys["pair"].ptr), "sort keys")
        check(lib().ma_sort_markdup_dev(st.handle, n, keys["coord"].ptr, keys["dhi"].ptr,
                                        keys["dlo"].ptr, keys["score"].ptr, keys["pair"].ptr,
                                        outs["is_dup"].ptr, outs["perm"].ptr, 1, outs["dup_perm"].ptr,
                                        self.scratch.ptr, self.scratch_bytes), "sort markdup")

    def capture(self):
        """One step (map, keys, sort + markdup) captured as a HIP graph."""
        with self.mialign.Graph.capture(self.stream) as cap:
            self.step(False)
        return cap.graph

    def run_graph(self, graph, steps: int, warmup: int, label: str):
        """Steps as graph replays: (reads/s, ms per step)."""
        for _ in range(warmup):
            graph.launch(self.stream)
        self.stream.sync()
        t_start = time.perf_counter()
        for _ in range(steps):
            graph.launch(self.stream)
        self.stream.sync()
        elapsed = time.perf_counter() - t_start
        log(f"{label}: {steps} graph replays, {1000.0 * elapsed / steps:.2f} ms/step")
        return self.n * steps / elapsed, 1000.0 * elapsed / steps

    def outputs(self):
        """The step's records and sort outputs (host copies), to compare runs."""
        self.stream.sync()
        r = self.recs.download()
        return [r[f] for f in FIELDS] + [self.outs[f].download() for f in ("is_dup", "perm", "dup_perm")]

    def tune(self, names, repeats: int = 3) -> dict:
        """Map knobs on this step's map launch and the radix tile on its sort."""
        from mialign import autotune
        from mialign._lib import MA_BS_C2T, MA_ENC_2BIT

        self.step(False)  # the sort is timed on this step's own keys
        self.stream.sync()
        mp = autotune.tune_map(self.ix, self.ws, self.stream, self.d_bases, self.stride, self.d_lens, self.n, self.L,
                               self.recs, encoding=MA_ENC_2BIT, bs_r1=MA_BS_C2T, params=self.prm, repeats=repeats,
                               names=names)
        k = self.keys
        so = autotune.tune_sort(self.stream, self.n, k["coord"], k["dhi"], k["dlo"], k["score"], k["pair"],
                                self.outs["is_dup"], self.outs["perm"], self.outs["dup_perm"], self.scratch,
                                self.scratch_bytes, repeats=repeats)
        return {"map": mp, "sort": so}

    def run(self, steps: int, warmup: int, label: str):
        """Returns (reads/s over the timed steps, ms per step, mean map-kernel ms)."""
        for i in range(warmup):
            self.step(False)
            self.stream.sync()
            log(f"{label}: warmup step {i + 1}/{warmup} done")
        kernel_ms = []
        self.stream.sync()
        t_start = time.perf_counter()
        for _ in range(steps):
            self.step(True)
            self.ev_b.synchronize()
            kernel_ms.append(self.ev_a.elapsed_ms(self.ev_b))
        self.stream.sync()
        elapsed = time.perf_counter() - t_start
        return self.n * steps / elapsed, 1000.0 * elapsed / steps, float(np.mean(kernel_ms))

    def close(self) -> None:
        for a in ("d_bases", "d_lens", "d_tid", "recs", "ws", "keys", "outs", "scratch"):
            setattr(self, a, None)


FIELDS = ("cid", "pos", "mapq", "flag", "sl", "sr", "nm")
ALL_FIELDS = FIELDS + ("iop", "iat", "ilen")


def pe_section(sr, reads, tid, blob, name_off, qual, with_cpu: bool) -> dict:
    """The step's records as pairs (read i and read n/2 + i): PE keys, sort +
    markdup and the PE record pack (flags, mate fields, tlen, pair-wise 0x400)
    on the device, timed with HIP events on their stream; the keys and the
    sort's order and duplicate flags equal the oracle's for every pair."""
    import ctypes as C

    import mialign
    from mialign import bam_emit
    from mialign._lib import check, lib
    from oracle import oracle as orc

    n = sr.n - (sr.n & 1)
    np_ = n // 2
    st = sr.stream
    keys = {f: mialign.DeviceArray((n,), dt) for f, dt in (("coord", np.uint64), ("dhi", np.uint64),
                                                          ("dlo", np.uint64), ("score", np.uint32),
                                                          ("pair", np.uint32))}
    perm, is_dup = mialign.DeviceArray((n,), np.uint32), mialign.DeviceArray((n,), np.uint32)
    sb = int(lib().ma_sort_scratch_bytes(n))
    scratch = mialign.DeviceArray((sb,), np.uint8)
    dreads = bam_emit.DeviceReads(blob, name_off, reads.bases[:n], reads.lens[:n], qual[:n])
    rg = b"mojo1"
    pk = bam_emit.BamPacker(n, dreads.out_bound(len(rg)))
    ka = [keys[f].ptr for f in ("coord", "dhi", "dlo", "score", "pair")]
    ev = [mialign.Event() for _ in range(4)]
    times = []
    for it in range(4):
        ev[0].record(st)
        check(lib().ma_sort_keys_pe_dev(st.handle, np_, C.byref(sr.recs.struct), sr.d_lens.ptr, sr.d_tid.ptr,
                                        int(tid.size), 0, *ka), "ma_sort_keys_pe_dev")
        ev[1].record(st)
        check(lib().ma_sort_markdup_dev(st.handle, n, *ka, is_dup.ptr, perm.ptr, 1, None, scratch.ptr, sb),
              "ma_sort_markdup_dev")
        ev[2].record(st)
        pk.pack_pe(st, np_, sr.recs, sr.d_tid, int(tid.size), dreads, rg, perm, is_dup, True)
        ev[3].record(st)
        ev[3].synchronize()
        if it:
            times.append([ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2]), ev[2].elapsed_ms(ev[3])])
    total = pk.total(n, st)
    k_ms, s_ms, p_ms = (float(np.median([t[i] for t in times])) for i in range(3))
    all_ms = k_ms + s_ms + p_ms
    out = {"workload": f"the step's {n} records as {np_} pairs (read i with read {np_} + i)",
           "pairs_per_s": round(np_ / (all_ms / 1000.0), 1), "keys_ms": round(k_ms, 3),
           "sort_markdup_ms": round(s_ms, 3), "pack_ms": round(p_ms, 3), "bam_stream_bytes": total,
           "pack_GBps_out": round(total / (p_ms / 1000.0) / 1e9, 1)}
    log(f"pe: {np_} pairs keys {k_ms:.2f} ms, sort+markdup {s_ms:.2f} ms, pack {p_ms:.2f} ms "
        f"({total / 1e9:.2f} GB of records)")
    if with_cpu:
        rec = {f: v[:n] for f, v in sr.recs.download().items()}
        want = orc.pe_fields(rec, reads.lens[:n], tid)
        same_keys = all(np.array_equal(keys[f].download(), want[f]) for f in ("coord", "dhi", "dlo", "score", "pair"))
        wd, wp, _ = orc.sort_markdup(want["coord"], want["dhi"], want["dlo"], want["score"], want["pair"])
        same_sort = bool(np.array_equal(perm.download(), wp) and np.array_equal(is_dup.download(), wd))
        out["parity"] = {"pairs": np_, "keys_equal": bool(same_keys), "order_and_dup_flags_equal": same_sort,
                         "duplicate_records": int(wd.sum())}
        log(f"pe: keys equal the oracle's {same_keys}; order and duplicate flags {same_sort}")
    return out


def indel_run(sr, steps: int = 3):
    """The reference's default extension (max_indel 4, the single-indel branch
    on) over the same step: reads/s of map + keys + sort + markdup, the map
    kernel's time and indel hits; returns (section dict, merged records). The
    gapless records and order are made again afterwards for what follows."""
    import mialign

    gapless = sr.prm
    sr.prm = mialign.default_params(k=21, max_indel=4)
    try:
        value, ms, map_ms = sr.run(steps, 1, "indel cap 4")
        merged = sr.recs.download()
    finally:
        sr.prm = gapless
        sr.step(False)
        sr.stream.sync()
    n = sr.n
    out = {"workload": f"the step's {n} reads with max_indel 4 (the reference default)", "max_indel": 4,
           "reads_per_s": round(value, 1), "ms_per_step": round(ms, 3), "map_kernel_ms": round(map_ms, 3),
           "indel_hits": int((merged["iop"] != 0).sum()), "mapped_frac": round(float((merged["cid"] >= 0).mean()), 4)}
    log(f"indel cap 4: {ms:.1f} ms/step, map {map_ms:.1f} ms, {value / 1e6:.1f} M reads/s, "
        f"{out['indel_hits']} indel hits")
    return out, merged


def indel_parity(out: dict, merged: dict, oix, reads, ns: int) -> None:
    """GPU == oracle (all ten record fields) at max_indel 4 on the first ns reads."""
    from oracle import oracle as orc

    ns = min(ns, reads.lens.size)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    want, _, _, _ = orc.map_batch(oix, orc.convert_reads(reads.bases[:ns], 1), reads.lens[:ns],
                                  orc.params(max_indel=4), nthreads=threads)
    ok = all(np.array_equal(merged[f][:ns], want[f]) for f in ALL_FIELDS)
    out["parity"] = {"reads": ns, "records_bit_exact": bool(ok), "oracle_indel_hits": int((want["iop"] != 0).sum())}
    log(f"indel cap 4: first {ns} reads equal the oracle's: {ok}")


def indel_section(sr, ix, dual, reads, with_cpu: bool, parity_reads: int) -> dict:
    """indel_run on the step, then its parity sample against the oracle."""
    from oracle import oracle as orc

    out, merged = indel_run(sr)
    if with_cpu:
        km, of, po = ix.export_dense()
        indel_parity(out, merged, orc.OracleIndex(km, of, po, dual.seq, dual.contig_off), reads, parity_reads)
    return out


def _subset_names(blob: np.ndarray, off: np.ndarray, idx: np.ndarray):
    """Names of reads idx (blob + offsets) in that order."""
    ln = (off[idx + 1] - off[idx]).astype(np.int64)
    new_off = np.zeros(idx.size + 1, np.int64)
    np.cumsum(ln, out=new_off[1:])
    src = np.repeat(off[idx] - new_off[:-1], ln) + np.arange(int(new_off[-1]), dtype=np.int64)
    return blob[src], new_off


# HBM bytes per 10M-record launch (150 bp reads, 100 Mbp workload) from FETCH_SIZE x 2 + WRITE_SIZE,
# profiles/bam_emit_and_fastq_ingest_10Mx150.md
BAM_TRAFFIC_PROFILED = {"reads": 10_000_000, "read_len": 150, "pack": 10.94e9, "bgzf_deflate": 19.2e9,
                        "fastq": 25.5e9, "source": "profiles/bam_emit_and_fastq_ingest_10Mx150.md"}


def fastq_section(reads, blob, name_off, qual, with_cpu: bool, cpu_records: int, parity_records: int) -> dict:
    """FASTQ ingest of the step's reads: their FASTQ text (names, bases, '+',
    qualities) resident in HBM, parsed on the device into base, C2T-converted
    and quality rows, lengths and names (ma_fastq_parse_dev); timed with HIP
    events. Algorithmic bytes = the text read once + the three row arrays, the
    two length arrays, the names and their offsets written."""
    import mialign
    from mialign import linear_fastq as lf
    from mialign import synth
    from mialign._lib import check, lib
    from mialign.serve import pinned_empty

    n, L = reads.lens.size, reads.bases.shape[1]
    t0 = time.time()
    text = synth.write_fastq(reads.bases, qual, reads.lens, blob, name_off)
    nbytes = int(text.size)
    d_text = mialign.DeviceArray((nbytes + 16,), np.uint8)
    check(lib().ma_memcpy_h2d(d_text.ptr, text.ctypes.data, nbytes, None), "h2d")
    batch = lf.DeviceFastqBatch(n, L, nbytes)
    parser = lf.FastqParser(nbytes, n)
    st = mialign.Stream(0)
    log(f"fastq: {nbytes / 1e9:.2f} GB of FASTQ text for {n} reads written and uploaded in {time.time() - t0:.1f}s")
    info = parser.parse(st, d_text.ptr, nbytes, True, batch, 1)
    ev0, ev1 = mialign.Event(), mialign.Event()
    ms = []
    for _ in range(5):
        ev0.record(st)
        check(lib().ma_fastq_parse_dev(st.handle, d_text.ptr, nbytes, 1, n, L, 1, batch.rows.ptr, batch.conv.ptr,
                                       batch.qual.ptr, batch.lens.ptr, batch.qual_lens.ptr, batch.names.ptr,
                                       batch.names_cap, batch.name_off.ptr, batch.info.ptr, parser.scratch.ptr,
                                       parser.scratch_bytes), "ma_fastq_parse_dev")
        ev1.record(st)
        ev1.synchronize()
        ms.append(ev0.elapsed_ms(ev1))
    dev_ms = float(np.median(ms))
    name_bytes = int(name_off[-1])
    alg = nbytes + 3 * n * L + 8 * n + name_bytes + 8 * (n + 1)
    achieved = alg / (dev_ms / 1000.0) / 1e9
    log(f"fastq: parse {dev_ms:.2f} ms = {n / dev_ms / 1e3:.1f}M records/s, {achieved:.0f} GB/s algorithmic")
    # size-independent round trip: the parsed batch equals the arrays the text was written from
    got = batch.download()
    conv = np.where((reads.bases == ord("C")) | (reads.bases == ord("c")), np.uint8(ord("T")), reads.bases)
    ok_all = bool(info.records == n and info.status == 0 and info.consumed == nbytes
                  and np.array_equal(got["rows"], reads.bases) and np.array_equal(got["conv"], conv)
                  and np.array_equal(got["qual"], qual) and np.array_equal(got["lens"], reads.lens)
                  and got["names"].tobytes() == blob[:name_bytes].tobytes()
                  and np.array_equal(got["name_off"], name_off))
    del got
    # the oracle on the text of the first records
    from oracle import oracle as orc

    m = min(parity_records, n)
    rec_end = _fastq_record_end(text, m)
    want = orc.fastq_parse(text[:rec_end].tobytes(), L, 1)
    ok_oracle = bool(want["n"] == m and np.array_equal(want["conv"], conv[:m]) and np.array_equal(want["qual"], qual[:m]))
    del want
    log(f"fastq: round trip of all {n} records exact: {ok_all}; first {m} equal the oracle's: {ok_oracle}")
    # PCIe-inclusive: pinned text copied in, then parsed
    pin = pinned_empty((nbytes,), np.uint8)
    pin[:] = text
    t_start = time.perf_counter()
    check(lib().ma_memcpy_h2d(d_text.ptr, pin.ctypes.data, nbytes, st.handle), "h2d")
    parser.parse(st, d_text.ptr, nbytes, True, batch, 1)
    pcie_s = time.perf_counter() - t_start
    cpu = None
    if with_cpu:
        ns = min(cpu_records, n)
        sample = text[: _fastq_record_end(text, ns)].tobytes()
        t0 = time.perf_counter()
        orc.fastq_parse(sample, L, 1)
        cs = time.perf_counter() - t0
        cpu = {"value": round(ns / cs, 1), "unit": "records/s", "cores": 1, "kind": "port",
               "sample": f"oracle FASTQ parse (the reference's one-thread reader restated) of the first {ns} "
                         f"records ({len(sample) / 1e6:.0f} MB), {cs:.2f} s"}
        log(f"fastq: cpu baseline {ns / cs:.0f} records/s")
    del pin, d_text, batch, parser, text
    return {
        "workload": f"FASTQ text of the step's {n} reads ({L} bp, Illumina-style names and qualities), "
                    "resident in HBM -> base / C2T / quality rows, lengths, names",
        "fastq_bytes": nbytes,
        "parse_ms": round(dev_ms, 3),
        "records_per_s": round(n / (dev_ms / 1000.0), 1),
        "GBps_text": round(nbytes / (dev_ms / 1000.0) / 1e9, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": "fq_* (count, lines, measure + rows, names)",
                     "algorithmic_bytes_per_read": round(alg / n, 1),
                     "traffic": _profiled(n, L, "fastq")},
        "pcie_inclusive_records_per_s": round(n / pcie_s, 1),
        "parity": {"all_records_round_trip": ok_all, "oracle_records": m, "oracle_equal": ok_oracle},
        "cpu_baseline": cpu,
    }


def fq2bam_section(ix, dual, reads, blob, name_off, qual, n: int) -> dict:
    """The whole product path on the first n reads: their FASTQ file (page
    cache) -> fastq_to_bam (device parse, map, keys, sort + markdup, pack,
    BGZF) -> a BAM file; wall time including the host's file reads and writes."""
    import tempfile

    from mialign import bam_emit, synth
    from mialign.fq2bam import fastq_to_bam
    from mialign.linear_index import LinearIndex

    n = min(n, reads.lens.size)
    d = tempfile.mkdtemp(prefix="mialign_fq2bam_")
    fq, fqz, out = os.path.join(d, "r1.fq"), os.path.join(d, "r1.fq.gz"), os.path.join(d, "out.bam")
    try:
        sub_off = name_off[: n + 1]
        text = synth.write_fastq(reads.bases[:n], qual[:n], reads.lens[:n], blob, sub_off)
        with open(fq, "wb") as fh:
            fh.write(memoryview(text))
        fq_bytes = int(text.size)
        del text
        lengths = [int(dual.contig_off[i + 1] - dual.contig_off[i]) for i in range(len(dual.names))]
        index = LinearIndex.from_gpu(ix, dual.names, lengths)
        fastq_to_bam(index, fq, out, "C2T", batch_reads=1 << 20)  # warm: allocations, file in page cache
        os.remove(out)  # the timed run writes a new file, as a run to a fresh path does
        t0 = time.perf_counter()
        stats = fastq_to_bam(index, fq, out, "C2T", batch_reads=1 << 20)
        wall = time.perf_counter() - t0
        bam_bytes = os.path.getsize(out)
        with open(out, "rb") as fh:
            digest = hashlib.sha256(fh.read()).hexdigest()
        os.remove(out)
        # the same reads as a BGZF-compressed FASTQ (bgzip framing, level 1), inflated on host threads
        with open(fq, "rb") as fh:
            z = bam_emit.bgzf_compress(np.frombuffer(fh.read(), np.uint8), 1)
        os.remove(fq)
        with open(fqz, "wb") as fh:
            fh.write(z)
        gz_bytes = len(z)
        del z
        t0 = time.perf_counter()
        stats_z = fastq_to_bam(index, fqz, out, "C2T", batch_reads=1 << 20)
        wall_z = time.perf_counter() - t0
        with open(out, "rb") as fh:
            same = hashlib.sha256(fh.read()).hexdigest() == digest
    finally:
        for f in (fq, fqz, out):
            if os.path.exists(f):
                os.remove(f)
        os.rmdir(d)
    log(f"fq2bam: {n} reads FASTQ ({fq_bytes / 1e9:.2f} GB) -> BAM ({bam_bytes / 1e9:.2f} GB) in {wall:.2f} s "
        f"= {n / wall / 1e6:.1f}M reads/s ({stats})")
    log(f"fq2bam: the same reads as BGZF FASTQ ({gz_bytes / 1e9:.2f} GB) in {wall_z:.2f} s = "
        f"{n / wall_z / 1e6:.1f}M reads/s; BAM identical to the plain-text run's: {same}")
    return {"workload": f"the first {n} reads as a FASTQ file -> coordinate-sorted, duplicate-marked BAM file",
            "reads": n, "fastq_bytes": fq_bytes, "bam_bytes": bam_bytes, "wall_s": round(wall, 3),
            "reads_per_s": round(n / wall, 1), "stages": stats,
            "bgzf_fastq": {"fastq_gz_bytes": gz_bytes, "wall_s": round(wall_z, 3),
                           "reads_per_s": round(n / wall_z, 1), "bam_identical": same,
                           "ingest_and_map_s": stats_z["ingest_and_map_s"]}}


def _profiled(n: int, read_len: int, what: str):
    """Profiled HBM bytes per launch when the run is the profiled workload, else None."""
    tp = BAM_TRAFFIC_PROFILED
    return tp[what] if (tp["reads"], tp["read_len"]) == (n, read_len) else None


def _fastq_record_end(text: np.ndarray, m: int) -> int:
    """Byte offset just past the m-th record of 4-line FASTQ text."""
    if m <= 0:
        return 0
    nl = np.flatnonzero(text[: min(text.size, 400 * m + 1024)] == 10)
    if nl.size < 4 * m:
        nl = np.flatnonzero(text == 10)
    return int(nl[4 * m - 1]) + 1


def _check_members(z: np.ndarray, stream: np.ndarray, k: int) -> bool:
    """The first k BGZF members of z inflate (zlib) to the stream's first k
    65280-byte blocks with matching CRC32 / ISIZE, and z ends with the EOF member."""
    import zlib

    pos, ok = 0, True
    for m in range(k):
        if pos + 18 > z.size:
            break
        bsize = int(z[pos + 16]) | int(z[pos + 17]) << 8
        mem = bytes(z[pos:pos + bsize + 1])
        body = zlib.decompress(mem[18:-8], -15)
        want = bytes(stream[65280 * m:65280 * (m + 1)])
        ok &= body == want and int.from_bytes(mem[-8:-4], "little") == zlib.crc32(want)
        ok &= int.from_bytes(mem[-4:], "little") == len(want)
        pos += bsize + 1
    eof = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
    return bool(ok and bytes(z[-28:]) == eof)


def bam_section(sr, reads, tid, n_tid: int, blob, name_off, qual, with_cpu: bool, cpu_records: int,
                parity_records: int) -> dict:
    """BAM emit of the step's records, resident in HBM: the device pack in
    coordinate order with the duplicate flag (ma_bam_pack_se_dev: sizes + scan +
    one wave per record), the stream's copy to pinned host memory and BGZF at
    level 1 on the host's threads (to /dev/null). The pack's time comes from
    HIP events on its stream; the pack kernel's algorithmic bytes are the
    record stream written plus the names, bases, qualities and the 56 bytes of
    per-record fields, permutation, flags and offsets read."""
    import ctypes as C

    import mialign
    from mialign import bam_emit, synth
    from mialign._lib import check, lib
    from mialign.serve import pinned_empty

    n, L = reads.lens.size, reads.bases.shape[1]
    t0 = time.time()
    dreads = bam_emit.DeviceReads(blob, name_off, reads.bases, reads.lens, qual)
    rg = b"mojo1"
    cap = dreads.out_bound(len(rg))
    pk = bam_emit.BamPacker(n, cap)
    st = sr.stream
    log(f"bam: names + qualities for {n} reads generated and uploaded in {time.time() - t0:.1f}s")
    ev0, ev1 = mialign.Event(), mialign.Event()

    def pack():
        pk.pack_se(st, n, sr.recs, sr.d_tid, n_tid, dreads, rg, sr.outs["perm"], sr.outs["is_dup"], True)

    pack()
    total = pk.total(n, st)
    pack_ms = []
    for _ in range(5):
        ev0.record(st)
        pack()
        ev1.record(st)
        ev1.synchronize()
        pack_ms.append(ev0.elapsed_ms(ev1))
    dev_ms = float(np.median(pack_ms))
    name_bytes = int(name_off[-1])
    in_bytes = name_bytes + 2 * int(reads.lens.astype(np.int64).sum()) + 56 * n
    achieved = (total + in_bytes) / (dev_ms / 1000.0) / 1e9
    log(f"bam: pack {dev_ms:.2f} ms for {n} records, {total / 1e9:.2f} GB stream, {achieved:.0f} GB/s algorithmic")

    # end to end: pack + copy to pinned host memory + BGZF level 1 (host threads)
    host = pinned_empty((total,), np.uint8)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t_start = time.perf_counter()
    pack()
    check(lib().ma_memcpy_d2h(host.ctypes.data, pk.d_out.ptr, total, st.handle), "d2h")
    st.sync()
    t_copy = time.perf_counter()
    w = bam_emit.BamWriter("/dev/null", ["c"], [1], threads=threads)
    w.write_raw(host.ctypes.data, total)
    z_bytes = w.close()
    t_end = time.perf_counter()
    emit_s = t_end - t_start
    log(f"bam: emit {emit_s:.2f} s (pack + copy {t_copy - t_start:.2f} s, BGZF {t_end - t_copy:.2f} s on "
        f"{threads} threads, {z_bytes / total:.3f} of the stream)")

    # end to end with BGZF on the device: pack + deflate members + copy of the members
    dz = bam_emit.DeviceBgzf(total)
    dz.compress(st, pk.d_out.ptr, total, 1, True)
    gz_total = dz.total(st)
    z_ms = []
    for _ in range(5):
        ev0.record(st)
        dz.compress(st, pk.d_out.ptr, total, 1, True)
        ev1.record(st)
        ev1.synchronize()
        z_ms.append(ev0.elapsed_ms(ev1))
    dz_ms = float(np.median(z_ms))
    zhost = pinned_empty((gz_total,), np.uint8)
    t_start = time.perf_counter()
    pack()
    dz.compress(st, pk.d_out.ptr, total, 1, True)
    gzt = dz.total(st)  # the members' length (one small copy and a sync)
    check(lib().ma_memcpy_d2h(zhost.ctypes.data, dz.d_dst.ptr, gzt, st.handle), "d2h")
    st.sync()
    with open("/dev/null", "wb") as fh:
        fh.write(memoryview(zhost))
    emit_dev_s = time.perf_counter() - t_start
    # the first 64 members and the EOF member of the device BGZF inflate to the stream
    dev_ok = _check_members(zhost, host, 64)
    log(f"bam: device BGZF {dz_ms:.2f} ms, {gz_total / total:.3f} of the stream; emit with device BGZF "
        f"{emit_dev_s:.3f} s; first members inflate to the stream: {dev_ok}")

    # parity: the first records of the device stream == the oracle's pack of the same records
    perm = sr.outs["perm"].download()
    is_dup = sr.outs["is_dup"].download()
    merged = sr.recs.download()
    m = min(parity_records, n)
    idx = perm[:m].astype(np.int64)
    sub_blob, sub_off = _subset_names(blob, name_off, idx)
    sub = {f: merged[f][idx] for f in FIELDS}
    from oracle import oracle as orc

    want = orc.bam_pack_se(sub, tid, sub_blob, sub_off, reads.bases[idx], reads.lens[idx], qual[idx], rg,
                           is_dup=is_dup[idx], apply_dup=True)
    ok = bytes(host[:len(want)]) == want
    log(f"bam: first {m} records of the coordinate-ordered stream equal the oracle's: {ok}")

    cpu = None
    if with_cpu:
        # the reference's host emit: one thread packs records (or_bam_pack_se on a
        # sample), then zlib level 1 members on a pool of threads (bgzf_block)
        from concurrent.futures import ThreadPoolExecutor

        ns = min(cpu_records, n)
        sub2 = {f: merged[f][:ns] for f in FIELDS}
        b2, o2 = _subset_names(blob, name_off, np.arange(ns))
        t0 = time.perf_counter()
        s2 = orc.bam_pack_se(sub2, tid, b2, o2, reads.bases[:ns], reads.lens[:ns], qual[:ns], rg)
        pack_s = time.perf_counter() - t0
        chunk = bytes(host[: min(total, 512 << 20)])
        blocks = [chunk[i:i + 65280] for i in range(0, len(chunk), 65280)]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            zs = list(ex.map(lambda b: orc.bgzf_block(b, 1), blocks))
        z_s = time.perf_counter() - t0
        per_read = total / n
        rate = 1.0 / (pack_s / ns + per_read * z_s / len(chunk))
        cpu = {"value": round(rate, 1), "unit": "reads/s", "cores": threads, "kind": "port",
               "sample": f"oracle pack of {ns} records on 1 thread ({ns / pack_s:.0f} records/s) + zlib level 1 "
                         f"BGZF of {len(chunk) >> 20} MiB of the stream on {threads} threads "
                         f"({len(chunk) / z_s / 1e6:.0f} MB/s); {sum(map(len, zs)) / len(chunk):.3f} of the stream"}
        del s2
        log(f"bam: cpu baseline {rate:.0f} reads/s")
    out = {
        "workload": f"the step's {n} records (150 bp SE, coordinate order, duplicates flagged) as BAM",
        "stream_bytes": total,
        "pack_ms": round(dev_ms, 3),
        "pack_records_per_s": round(n / (dev_ms / 1000.0), 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": "bam_pack_kernel (+ layout, sizes, scan)",
                     "algorithmic_bytes_per_read": round((total + in_bytes) / n, 1),
                     "traffic": _profiled(n, L, "pack")},
        "emit_s": round(emit_dev_s, 3),
        "emit_reads_per_s": round(n / emit_dev_s, 1),
        "emit": "pack + BGZF on the device + copy of the members to pinned host memory + write (/dev/null)",
        "bgzf_device": {"ms": round(dz_ms, 3), "ratio": round(gz_total / total, 4),
                        "GBps_in": round(total / (dz_ms / 1000.0) / 1e9, 1),
                        "traffic": _profiled(n, L, "bgzf_deflate"),
                        "first_members_inflate_to_stream": bool(dev_ok)},
        "emit_host_bgzf_s": round(emit_s, 3),
        "emit_host_bgzf_reads_per_s": round(n / emit_s, 1),
        "bgzf_host": {"level": 1, "threads": threads, "s": round(t_end - t_copy, 3),
                      "ratio": round(z_bytes / total, 4), "note": "zlib; members byte-identical to the reference's"},
        "parity_sample": {"records": m, "bytes_equal_oracle": bool(ok)},
        "cpu_baseline": cpu,
    }
    del host, zhost, dreads, pk, dz
    return out


def sr_enc():
    from mialign._lib import MA_ENC_2BIT

    return MA_ENC_2BIT


def sr_c2t():
    from mialign._lib import MA_BS_C2T

    return MA_BS_C2T


def cpu_info() -> dict:
    """CPU model name and logical CPU count of this host."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count()}


def map_bytes_per_read(cnt, n_sample: int, lens: np.ndarray) -> float:
    """Algorithmic bytes of one read's map: its 2-bit bases in, a 16-byte
    record out, and per pass (FW, plus RC where it runs) 24 bytes per seed
    probe (u64 key + u64 range), 4 bytes per posting visited (u32 position)
    and one 2-bit window of the read's length per reference window compared.
    probes / postings / windows are the oracle's counts over its sequential
    walk of the same reads; for uneven lengths the window term uses the
    sample's mean 2-bit read size."""
    w = float(np.mean((lens[:n_sample].astype(np.int64) + 3) // 4))
    return w + 16 + (24 * cnt.probes + 4 * cnt.postings + w * cnt.windows) / n_sample


def oracle_index(ix, dual):
    from oracle import oracle as orc

    t0 = time.time()
    km, of, po = ix.export_dense()
    oix = orc.OracleIndex(km, of, po, dual.seq, dual.contig_off)
    log(f"oracle index ready in {time.time() - t0:.1f}s")
    return oix


def cpu_map_baseline(oix, bases, lens, tid, ns, ns1, what):
    """Oracle on the first ns reads (OpenMP over reads on every host thread +
    SE keys + stable-sort markdup), and on the first ns1 reads with one thread;
    returns (baseline dict, oracle records, oracle counters)."""
    from oracle import oracle as orc

    conv = orc.convert_reads(bases[:ns], 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"cpu baseline ({what}): mapping {ns} reads on {threads} threads")
    t0 = time.perf_counter()
    want, _, _, cnt = orc.map_batch(oix, conv, lens[:ns], nthreads=threads)
    orc.sort_markdup(*orc.sort_keys_se(want, lens[:ns], tid))
    cpu_s = time.perf_counter() - t0
    ns1 = min(ns1, ns)
    t0 = time.perf_counter()
    w1, _, _, _ = orc.map_batch(oix, conv[:ns1], lens[:ns1], nthreads=1)
    orc.sort_markdup(*orc.sort_keys_se(w1, lens[:ns1], tid))
    cpu1_s = time.perf_counter() - t0
    cpu = {"value": round(ns / cpu_s, 1), "unit": "reads/s", "cores": threads, "kind": "port",
           "value_1core": round(ns1 / cpu1_s, 1), **cpu_info(),
           "sample": f"first {ns} reads of {what} (oracle map with OpenMP over reads on {threads} threads + SE "
                     f"keys + stable-sort markdup), {cpu_s:.1f} s; one thread: first {ns1} reads, {cpu1_s:.1f} s"}
    log(f"cpu baseline ({what}): {cpu['value']:.0f} reads/s on {threads} threads, "
        f"{cpu['value_1core']:.0f} reads/s on one")
    return cpu, want, cnt


def sort_parity(sr, merged, lens, tid) -> dict:
    """The step's keys, coordinate order and duplicate flags against the
    oracle's keys of the same records and its stable sorts + markdup scan,
    over every record of the step."""
    from oracle import oracle as orc

    t0 = time.time()
    keys = orc.sort_keys_se(merged, lens, tid)
    got = [sr.keys[f].download() for f in ("coord", "dhi", "dlo", "score", "pair")]
    keys_ok = all(np.array_equal(a, b) for a, b in zip(keys, got))
    d, p, _ = orc.sort_markdup(*keys)
    perm_ok = bool(np.array_equal(p, sr.outs["perm"].download()))
    dup_ok = bool(np.array_equal(d, sr.outs["is_dup"].download()))
    log(f"sort parity over {lens.size} records: keys {keys_ok}, perm {perm_ok}, is_dup {dup_ok} "
        f"({time.time() - t0:.1f}s)")
    return {"records": int(lens.size), "keys_bit_exact": bool(keys_ok), "perm_bit_exact": perm_ok,
            "is_dup_bit_exact": dup_ok, "duplicates": int(d.sum())}


def roofline_of(bpr: float, n: int, kernel_ms: float, traffic=None, **extra) -> dict:
    achieved = bpr * n / (kernel_ms / 1000.0) / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "map_reads_kernel",
           "kernel_ms": round(kernel_ms, 3), "algorithmic_bytes_per_read": round(bpr, 1)}
    out.update(extra)
    return out


class GenomeScale:
    """The 3.1 Gbp reference of the north-star, uneven-length and served
    workloads: 24 contigs, 1% CpG, 5% low-complexity repeat blocks, both
    conversions indexed (6.2 Gbp), k=21 index built on the device."""

    NAME = "3.1 Gbp synthetic (24 contigs, 1% CpG, 5% low-complexity repeat blocks) x2 conversions"

    def __init__(self, keep_dense: bool, genome_bp: int = 3_100_000_000):
        import mialign
        from mialign import synth

        t0 = time.time()
        self.ref = synth.gen_reference(31, 24, genome_bp, repeat_frac=0.05)
        self.dual = synth.dual_convert(self.ref)
        self.tid = synth.tid_of_cid(self.dual.names)
        log(f"genome scale: reference 3.1 Gbp x2 in {time.time() - t0:.1f}s")
        t0 = time.time()
        self.ix = mialign.GpuIndex.build(self.dual.seq, self.dual.contig_off, 21, 0, keep_dense=keep_dense)
        tb = time.time() - t0
        self.index = {"keys": self.ix.n_keys, "postings": self.ix.n_postings, "build_s": round(tb, 1),
                      "table": self.ix.hash_stats()}
        log(f"genome scale: index {self.ix.n_keys} keys, {self.ix.n_postings} postings in {tb:.1f}s, "
            f"{self.index['table']}")
        self._oix = None

    def oracle(self):
        if self._oix is None:
            self._oix = oracle_index(self.ix, self.dual)
        return self._oix

    def close(self) -> None:
        self._oix = None
        self.ix.close()


# HBM bytes one map launch moves on the headline workload (3.1 Gbp x2
# reference, 10M x 150 bp reads): rocprofv3 --pmc FETCH_SIZE x 2.00 +
# WRITE_SIZE x 1.00, the factors read off copy_u64_kernel (a copy of known
# size) in the same runs.
MAP_TRAFFIC_PROFILED_GS = {"reads": 10_000_000, "read_len": 150, "bytes": 160.4e9, "kernel_ms": 57.19,
                           "source": "profiles/map_reads_kernel_3100Mbp_10Mx150.md"}


def headline_section(args, gs: GenomeScale) -> dict:
    """North-star workload: 10M single-end 150 bp bisulfite reads against the
    3.1 Gbp x2 index; one step = map (FW + RC + merge) + SE keys + stable
    coordinate sort + duplicate marking, on one stream. Records of a sample
    equal the oracle's; keys, order and duplicate flags of every record equal
    the oracle's sort of the same records."""
    from mialign import synth

    n, L = args.reads, 150
    t0 = time.time()
    reads = synth.gen_reads(1102, gs.ref, n, L, L)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    log(f"headline: {n} x {L} bp reads generated + 2-bit packed in {time.time() - t0:.1f}s")
    sr = StepRunner(gs.ix, packed, reads.lens, L, gs.tid)
    tuning = None
    if not args.no_tune:
        from mialign import autotune

        autotune.reset()
        t0 = time.time()
        tuning = sr.tune(("map_max_claim", "map_blocks_per_cu", "map_chunks_per_wave"))
        tuning["search_s"] = round(time.time() - t0, 1)
        tuning["cache"] = str(autotune.save(tuning))
        log(f"headline: tuned {autotune.current()} in {tuning['search_s']} s: map {tuning['map']['default_ms']} -> "
            f"{tuning['map']['tuned_ms']} ms, sort {tuning['sort']['default_ms']} -> {tuning['sort']['tuned_ms']} ms")
    v, ms, map_ms = sr.run(args.steps, args.warmup, "headline 3.1 Gbp 150 bp")
    log(f"headline: {args.steps} steps, {ms:.1f} ms/step, map kernel {map_ms:.1f} ms, {v / 1e6:.2f} M reads/s")
    ref_out = sr.outputs()
    graph = None
    if not args.no_graph:
        g = sr.capture()
        gv, gms = sr.run_graph(g, args.steps, args.warmup, "headline graph")
        same = all(np.array_equal(a, b) for a, b in zip(sr.outputs(), ref_out))
        graph = {"nodes": g.n_nodes, "ms_per_step_graph": round(gms, 3), "ms_per_step_stream": round(ms, 3),
                 "reads_per_s_graph": round(gv, 1), "outputs_equal_stream": bool(same)}
        del g
    if tuning is not None:
        from mialign import autotune

        tuned = autotune.current()
        autotune.reset()
        dv, dms, dmap = sr.run(args.steps, 1, "headline default knobs")
        same = all(np.array_equal(a, b) for a, b in zip(sr.outputs(), ref_out))
        autotune.apply(tuned)
        tuning.update({"choices": tuned, "ms_per_step_default": round(dms, 3), "ms_per_step_tuned": round(ms, 3),
                       "map_kernel_ms_default": round(dmap, 3), "map_kernel_ms_tuned": round(map_ms, 3),
                       "outputs_equal_default": bool(same)})
        sr.step(False)  # the tuned step's outputs again
    merged = sr.recs.download()
    indel = indel_mer = None
    if not args.no_indel and not args.headline_only:
        indel, indel_mer = indel_run(sr, max(2, args.steps // 4))
    out = {"value": v, "ms_per_step": ms, "map_ms": map_ms, "reads": n, "read_len": L, "tuning": tuning,
           "graph": graph, "indel_mode": indel,
           "mapped_frac": round(float((merged["cid"] >= 0).mean()), 4),
           "duplicates_marked": int(sr.outs["is_dup"].download().sum()),
           "cpu_baseline": None, "roofline": None, "parity_sample": None, "sort_parity": None}
    if args.headline_only:
        sr.close()
        calibration_copy()
        return out
    out["sort_parity"] = sort_parity(sr, merged, reads.lens, gs.tid)
    sr.close()
    del packed
    if not args.no_cpu:
        ns = min(args.gs_cpu_sample, n)
        cpu, want, cnt = cpu_map_baseline(gs.oracle(), reads.bases, reads.lens, gs.tid, ns, args.cpu_1core_sample,
                                          "the 3.1 Gbp 150 bp workload")
        ok = all(np.array_equal(merged[f][:ns], want[f]) for f in FIELDS)
        bpr = map_bytes_per_read(cnt, ns, reads.lens)
        tp = MAP_TRAFFIC_PROFILED_GS
        traffic = tp["bytes"] if tp and (tp["reads"], tp["read_len"]) == (n, L) else None
        detail = None
        if traffic:
            detail = {"per_read": round(traffic / n, 1), "x_algorithmic": round(traffic / (bpr * n), 2),
                      "profiled_kernel_ms": tp["kernel_ms"], "source": tp["source"]}
        out["cpu_baseline"] = cpu
        out["parity_sample"] = {"reads": ns, "records_bit_exact": bool(ok), "fields": list(FIELDS)}
        out["roofline"] = roofline_of(bpr, n, map_ms, traffic, traffic_detail=detail,
                                      oracle_counts_per_read={"probes": round(cnt.probes / ns, 2),
                                                              "postings": round(cnt.postings / ns, 2),
                                                              "windows": round(cnt.windows / ns, 2),
                                                              "rc_pass_frac": round(cnt.reads_rc / ns, 3)})
        log(f"headline: sample parity {ok}, {bpr:.0f} B/read algorithmic, x{v / cpu['value']:.0f} the host")
        if indel is not None:
            indel_parity(indel, indel_mer, gs.oracle(), reads, args.indel_parity_sample)
    return out


def uneven_section(args, gs: GenomeScale) -> dict:
    """50M reads of 50-250 bp (uniform, as after adapter trimming) against the
    3.1 Gbp x2 index in one persistent map launch per step (+ keys, sort,
    markdup); a sample of the records against the oracle."""
    from mialign import synth

    n2 = args.gs_uneven_reads
    t0 = time.time()
    ur = synth.gen_reads(1103, gs.ref, n2, 50, 250, max_len=250)
    upacked = synth.pack_2bit(ur.bases, ur.lens)
    log(f"uneven: {n2} reads generated + packed in {time.time() - t0:.1f}s")
    sr = StepRunner(gs.ix, upacked, ur.lens, 250, gs.tid)
    tuning = None
    if not args.no_tune:
        from mialign import autotune

        before = autotune.current()
        t0 = time.time()
        mp = autotune.tune_map(sr.ix, sr.ws, sr.stream, sr.d_bases, sr.stride, sr.d_lens, sr.n, sr.L, sr.recs,
                               encoding=sr_enc(), bs_r1=sr_c2t(), params=sr.prm, repeats=2,
                               names=("map_class_split", "map_max_claim"))
        tuning = {"map": mp, "search_s": round(time.time() - t0, 1), "start": before}
        log(f"uneven: tuned {mp['choices']}: map {mp['default_ms']} -> {mp['tuned_ms']} ms")
    v, ms, map_ms = sr.run(max(1, args.steps // 2), 1, "uneven 50-250 bp")
    merged = sr.recs.download()
    if tuning is not None:
        from mialign import autotune

        tuned = autotune.current()
        autotune.apply(tuning["start"])
        dv, dms, dmap = sr.run(max(1, args.steps // 4), 1, "uneven default split/claim")
        same = all(np.array_equal(sr.recs.download()[f], merged[f]) for f in FIELDS)
        autotune.apply(tuned)
        tuning.update({"choices": tuned, "ms_per_step_default": round(dms, 3), "ms_per_step_tuned": round(ms, 3),
                       "map_kernel_ms_default": round(dmap, 3), "map_kernel_ms_tuned": round(map_ms, 3),
                       "records_equal_default": bool(same)})
        autotune.save({"uneven": tuning})
    sr.close()
    res = {"reads": n2, "read_len": "uniform 50-250", "mean_len": round(float(ur.lens.mean()), 1),
           "reads_per_s": round(v, 1), "ms_per_step": round(ms, 3), "map_kernel_ms": round(map_ms, 3),
           "mapped_frac": round(float((merged["cid"] >= 0).mean()), 4),
           "launch": "one persistent map launch over all reads per step (per read-length class when the "
                     "tuned class split is on)", "tuning": tuning}
    log(f"uneven 50-250 bp: {n2} reads, {ms:.1f} ms/step, map {map_ms:.1f} ms, {v / 1e6:.1f} M reads/s")
    if not args.no_cpu:
        ns = min(args.gs_uneven_cpu_sample, n2)
        cpu, want, cnt = cpu_map_baseline(gs.oracle(), ur.bases, ur.lens, gs.tid, ns,
                                          args.cpu_1core_sample // 2, "the 3.1 Gbp 50-250 bp workload")
        ok = all(np.array_equal(merged[f][:ns], want[f]) for f in FIELDS)
        bpr = map_bytes_per_read(cnt, ns, ur.lens)
        res["cpu_baseline"] = cpu
        res["x_cpu"] = round(v / cpu["value"], 1)
        res["parity_sample"] = {"reads": ns, "records_bit_exact": bool(ok), "fields": list(FIELDS)}
        res["roofline"] = roofline_of(bpr, n2, map_ms)
        log(f"uneven: sample parity {ok}, {bpr:.0f} B/read algorithmic")
        del want
    res["_packed"], res["_lens"] = upacked, ur.lens
    return res


def config2_section(args) -> dict:
    """The single-large-batch workload: 100 Mbp x2 reference (10 contigs),
    10M x 150 bp reads, the same step; then the sections that start from its
    records (indel mode, BAM emit, PE emit, FASTQ ingest, FASTQ -> BAM)."""
    import mialign
    from mialign import synth

    t0 = time.time()
    ref = synth.gen_reference(1001, args.contigs, args.genome)
    dual = synth.dual_convert(ref)
    log(f"config 2: reference {args.genome} bp x2 conversions in {time.time() - t0:.1f}s")
    t0 = time.time()
    ix = mialign.GpuIndex.build(dual.seq, dual.contig_off, 21, 0)
    log(f"config 2: index {ix.n_keys} keys, {ix.n_postings} postings in {time.time() - t0:.1f}s")
    n, L = args.reads, args.read_len
    reads = synth.gen_reads(1002, ref, n, L, L)
    packed = synth.pack_2bit(reads.bases, reads.lens)
    tid = synth.tid_of_cid(dual.names)
    sr = StepRunner(ix, packed, reads.lens, L, tid, persistent=args.launch == "persistent")
    value, ms_per_step, map_ms = sr.run(args.steps, args.warmup, "config 2")
    log(f"config 2: {ms_per_step:.1f} ms/step, map kernel {map_ms:.1f} ms, {value / 1e6:.2f} M reads/s")
    merged = sr.recs.download()
    out = {"workload": f"single large batch: {args.genome // 1_000_000} Mbp reference ({args.contigs} contigs, "
                       f"1% CpG) x2 conversions, {n} SE {L} bp bisulfite reads, 2-bit packed u8 input",
           "reads": n, "read_len": L, "reads_per_s": round(value, 1), "ms_per_step": round(ms_per_step, 3),
           "map_kernel_ms": round(map_ms, 3), "mapped_frac": round(float((merged["cid"] >= 0).mean()), 4)}
    if not args.no_cpu:
        out["sort_parity"] = sort_parity(sr, merged, reads.lens, tid)
    out["indel_mode"] = None if args.no_indel else indel_section(sr, ix, dual, reads, not args.no_cpu,
                                                                 args.indel_parity_sample)
    if not (args.no_bam and args.no_fastq) or args.fq2bam_reads > 0:
        blob, name_off, qual = synth.gen_names_quals(1004, reads.lens, L)
        if not args.no_bam:
            out["bam_emit"] = bam_section(sr, reads, tid, tid.size, blob, name_off, qual, not args.no_cpu,
                                          args.bam_cpu_sample, args.bam_parity_sample)
            out["pe_emit"] = pe_section(sr, reads, tid, blob, name_off, qual, not args.no_cpu)
        sr.close()
        if not args.no_fastq:
            out["fastq_ingest"] = fastq_section(reads, blob, name_off, qual, not args.no_cpu,
                                                args.fastq_cpu_sample, args.bam_parity_sample)
        if args.fq2bam_reads > 0:
            out["fq2bam"] = fq2bam_section(ix, dual, reads, blob, name_off, qual, args.fq2bam_reads)
        del blob, name_off, qual
    sr.close()
    if args.serve_100mbp and not args.no_serve:
        out["serving"] = serve_section(ix, packed, reads.lens, L, args.serve_requests, args.serve_slots,
                                       args.serve_launch_reads, args.serve_size_weight,
                                       tid=synth.tid_of_cid(dual.names))
    if not args.no_cpu:
        oix = oracle_index(ix, dual)
        ns = min(args.cpu_sample, n)
        cpu, want, cnt = cpu_map_baseline(oix, reads.bases, reads.lens, tid, ns, args.cpu_1core_sample,
                                          "the 100 Mbp workload")
        ok = all(np.array_equal(merged[f][:ns], want[f]) for f in FIELDS)
        bpr = map_bytes_per_read(cnt, ns, reads.lens)
        tp = MAP_TRAFFIC_PROFILED
        traffic = tp["bytes"] if (tp["genome_bp"], tp["reads"], tp["read_len"]) == (args.genome, n, L) else None
        out["cpu_baseline"] = cpu
        out["parity_sample"] = {"reads": ns, "records_bit_exact": bool(ok)}
        out["roofline"] = roofline_of(bpr, n, map_ms, traffic)
        del want, oix
    ix.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--reads", type=int, default=10_000_000, help="150 bp reads of the headline (3.1 Gbp) step")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline step (no checks, baselines or other sections; for profiling)")
    ap.add_argument("--genome", type=int, default=100_000_000, help="reference size of the config-2 section")
    ap.add_argument("--contigs", type=int, default=10)
    ap.add_argument("--c2-reads", type=int, default=10_000_000, help="150 bp reads of the config-2 section")
    ap.add_argument("--read-len", type=int, default=150)
    ap.add_argument("--cpu-sample", type=int, default=5_000_000)
    ap.add_argument("--cpu-1core-sample", type=int, default=100_000,
                    help="reads the one-thread CPU baseline maps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-100mbp", action="store_true", help="also serve on the 100 Mbp index")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--sort-records", type=int, default=200_000_000,
                    help="sort + markdup stress records (0 = skip)")
    ap.add_argument("--sort-cpu-sample", type=int, default=20_000_000)
    ap.add_argument("--sort-only", action="store_true", help="only the sort + markdup stress section")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the served loop on the 3.1 Gbp index (a pool of --serve-pool 50-250 bp reads)")
    ap.add_argument("--serve-pool", type=int, default=10_000_000)
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in launch knobs (no autotuning)")
    ap.add_argument("--knobs", default="", help='fixed launch knobs as JSON (e.g. {"map_max_claim": 8}); no tuning')
    ap.add_argument("--no-graph", action="store_true", help="skip the graph-replay timing of the headline step")
    ap.add_argument("--serve-slots", type=int, default=2)
    ap.add_argument("--serve-launch-reads", type=int, default=128 << 10)
    ap.add_argument("--serve-size-weight", type=float, default=0.5,
                    help="served order: arrival + weight x estimated stand-alone time (0 = arrival order)")
    ap.add_argument("--launch", choices=("persistent", "batch"), default="persistent",
                    help="map-kernel grid: persistent (8 blocks/CU + work counter) or sized to the batch")
    ap.add_argument("--no-config2", action="store_true", help="skip the 100 Mbp section and what builds on it")
    ap.add_argument("--config2-only", action="store_true", help="only the 100 Mbp section and what builds on it")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 50-250 bp launch (and serving on it)")
    ap.add_argument("--gs-genome", type=int, default=3_100_000_000,
                    help="reference size of the headline, uneven and served workloads")
    ap.add_argument("--gs-uneven-reads", type=int, default=50_000_000,
                    help="50-250 bp reads on the 3.1 Gbp index (one launch)")
    ap.add_argument("--gs-cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--gs-uneven-cpu-sample", type=int, default=200_000)
    ap.add_argument("--no-bam", action="store_true", help="skip the BAM emit section")
    ap.add_argument("--no-fastq", action="store_true", help="skip the FASTQ ingest section")
    ap.add_argument("--fastq-cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--fq2bam-reads", type=int, default=4_000_000,
                    help="reads through the whole FASTQ file -> BAM file path (0 = skip)")
    ap.add_argument("--bam-cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-indel", action="store_true", help="skip the max_indel 4 section")
    ap.add_argument("--indel-parity-sample", type=int, default=200_000)
    ap.add_argument("--bam-parity-sample", type=int, default=1_000_000)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(reads shard across GPUs with a replicated index; no multi-GPU run is made here)",
              file=sys.stderr)
        return 2
    if args.headline_only:
        args.no_cpu = args.no_serve = args.no_config2 = args.no_uneven = args.no_graph = True
        args.sort_records = 0
        args.no_tune = True
    if args.knobs:
        from mialign import autotune

        autotune.apply(json.loads(args.knobs))
        args.no_tune = True
        log(f"launch knobs {autotune.current()}")

    import mialign

    info = mialign.device_info(0)
    log(f"device {info['name']} cu={info['n_cu']} hbm={info['hbm_total'] / 2**30:.0f} GiB")
    if args.sort_only:
        out = sort_section(args.sort_records, args.sort_cpu_sample, not args.no_cpu)
        calibration_copy()
        print(json.dumps({"sort_stress": out}), flush=True)
        return 0

    if args.config2_only:
        a2 = argparse.Namespace(**vars(args))
        a2.reads = args.c2_reads
        out = config2_section(a2)
        calibration_copy()
        print(json.dumps({"config2_100mbp": out}), flush=True)
        return 0

    if args.serve_only:
        from mialign import synth

        gs = GenomeScale(keep_dense=False, genome_bp=args.gs_genome)
        ur = synth.gen_reads(1103, gs.ref, args.serve_pool, 50, 250, max_len=250)
        serving = serve_section(gs.ix, synth.pack_2bit(ur.bases, ur.lens), ur.lens, 250, args.serve_requests,
                                args.serve_slots, args.serve_launch_reads, args.serve_size_weight, tid=gs.tid,
                                tune=not args.no_tune)
        gs.close()
        print(json.dumps({"serving": serving}), flush=True)
        return 0

    gs = GenomeScale(keep_dense=not args.no_cpu, genome_bp=args.gs_genome)
    head = headline_section(args, gs)
    uneven = serving = None
    if not args.no_uneven:
        uneven = uneven_section(args, gs)
        upacked, ulens = uneven.pop("_packed"), uneven.pop("_lens")
        if not args.no_serve:
            serving = serve_section(gs.ix, upacked, ulens, 250, args.serve_requests, args.serve_slots,
                                    args.serve_launch_reads, args.serve_size_weight, tid=gs.tid,
                                    tune=not args.no_tune)
            log(f"serving (3.1 Gbp, 50-250 bp): {serving}")
        del upacked, ulens
    index_info = gs.index
    gs.close()
    del gs

    c2 = None
    if not args.no_config2:
        a2 = argparse.Namespace(**vars(args))
        a2.reads = args.c2_reads
        c2 = config2_section(a2)

    sort_stress = None
    if args.sort_records > 0:
        sort_stress = sort_section(args.sort_records, args.sort_cpu_sample, not args.no_cpu)

    line = {
        "metric": METRIC,
        "value": round(head["value"], 1),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"150 bp WGBS reads against a 3.1 Gbp synthetic reference: {GenomeScale.NAME}, "
                        f"{head['reads']} SE {head['read_len']} bp bisulfite reads (non-CpG C->T 0.99, CpG kept "
                        "0.7, 0.5% substitutions, 5% planted duplicates), 2-bit packed u8 input, k=21 index "
                        "resident in HBM; step = map + SE keys + coordinate sort + duplicate marking",
            "reads": head["reads"],
            "read_len": head["read_len"],
            "genome_bp": args.gs_genome,
            "k": 21,
            "index": index_info,
            "map_kernel_ms": round(head["map_ms"], 3),
            "mapped_frac": head["mapped_frac"],
            "duplicates_marked": head["duplicates_marked"],
        },
        "roofline": head["roofline"],
        "cpu_baseline": head["cpu_baseline"],
        "x_cpu": round(head["value"] / head["cpu_baseline"]["value"], 1) if head["cpu_baseline"] else None,
        "parity_sample": head["parity_sample"],
        "sort_parity": head["sort_parity"],
        "tuning": head["tuning"],
        "graph": head["graph"],
        "indel_mode": head["indel_mode"],
        "uneven_50_250bp": uneven,
        "serving": serving,
        "sort_stress": sort_stress,
        "config2_100mbp": c2,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
