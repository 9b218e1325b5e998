"""Benchmark of the batched alignment hot path on one MI355X.

Workload (default): the full-feature batch -- 1,000,000 pairs, query length
uniform 100-300 bp, target = query + 0-100 bp random flanks at 5% divergence,
0.5% N bases with N_PENALTY 1, reverse-complemented queries on half the pairs --
aligned LOCAL WITH_TB with second best (score, ends, starts, second best, CIGAR),
int32 scores, 4-bit packed on the device.  One step = one pass of the device
pipeline (pack, reverse/complement, shape sort, DP, traceback) over the whole
batch with the inputs already resident in HBM.

Prints one JSON line.  `value` is GCUPS = sum over pairs of Lq*Lt (unpadded
cells) / step time.  `roofline` prices the DP kernel against the integer-VALU
peak (N_CU * 64 lanes * clock; twice that when the batch ran a packed 16-bit
kernel, two cells per lane instruction) at the DP kernel's algorithmic int
ops per cell (15 for the default mode: recurrence 11 + second best 4); `traffic`
comes from rocprofv3 counter runs when given.  `cpu_baseline` times the CPU
oracle (an OpenMP restatement of the reference kernels) on a bounded prefix of
the same batch on this host and reports the GPU's agreement with it there.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic int ops per cell of the DP kernel: the recurrence with end
# tracking (substitution, diagonal add, 4-way max, gap open, two gap extends,
# two gap maxima, rank encode, running max) is 11; second best adds 4 (compare,
# select, max, running max); direction bits stored by the DP add 8 (four
# comparisons, four inserts).  With the checkpointed traceback the DP stores
# no direction bits (the traceback kernel recomputes the tiles it needs).
OPS_SCORE = 11
OPS_SECOND = 4
OPS_DIR = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def work_cells(batch, mode, k_band=32, ksw_w=-1):
    """Cells of the metric: sum of Lq*Lt, or for the tile-banded kernel the
    nominal band, 64 cells per (strip, query word) inside the band
    (src/kernels/banded.h:83-85: words [max(0, i - other + 1), min(band + i, QW))),
    or for KSW with a band half-width w the cells |j - i| <= w (w clamped per
    pair to Lq*match - gapo)/gape + 1 as the kernel does)."""
    if mode == "ksw" and ksw_w >= 0:
        Lq = batch.qlen.astype(np.int64)
        Lt = batch.tlen.astype(np.int64)
        w = np.minimum(ksw_w, np.maximum(1, (Lq * 1 - 6) // 1 + 1))
        tot = 0
        for i in range(int(Lt.max())):
            c = np.minimum(Lq, i + w + 1) - np.maximum(0, i - w)
            tot += int(np.sum(np.where(i < Lt, np.clip(c, 0, None), 0)))
        return tot
    if mode != "banded":
        return int(np.dot(batch.qlen.astype(np.int64), batch.tlen.astype(np.int64)))
    QW = (batch.qlen.astype(np.int64) + 7) // 8
    TW = (batch.tlen.astype(np.int64) + 7) // 8
    band = k_band >> 3
    other = TW - (QW - band)
    tot = 0
    for i in range(int(TW.max())):
        lo = np.maximum(0, i - other + 1)
        hi = np.minimum(band + i, QW)
        tot += int(np.sum(np.where(i < TW, np.clip(hi - lo, 0, None), 0)))
    return 64 * tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--mode", default=None,
                    choices=["local_tb_second", "local_tb", "local_score", "local_start", "global", "semi_global", "ksw",
                             "banded"],
                    help="default: local_tb_second for the full-feature batch, local_score for the uneven one")
    ap.add_argument("--workload", default="full", choices=["full", "uneven", "seed", "band"],
                    help="full: 1M pairs of 100-300 bp (+flanks); uneven: 500k pairs, log-normal lengths 50-5000 bp; "
                         "band: learned band vs full DP on 1-10 kbp pairs (train seed 5, held-out seed 6)")
    ap.add_argument("--band-pairs", type=int, default=200_000, help="pairs per set for --workload band")
    ap.add_argument("--pipeline", type=int, default=0,
                    help="also stream the batch in this many sub-batches through one storage and through two "
                         "alternating storages (H2D of one overlapping the other's compute); 0 = skip")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--check-widen", type=int, nargs="+", default=[128],
                    help="check-band widening(s) (diagonals each side) to evaluate; the first one is reported as value")
    ap.add_argument("--dispatch", type=int, default=0,
                    help="kernel dispatch policy (0 auto, 2 one lane per pair only, 3 one wave per pair)")
    ap.add_argument("--wf", type=int, nargs=3, default=None, metavar=("MIN_TW", "MIN_CELLS", "MAX_BLOCKS"),
                    help="one-wave-per-pair switch thresholds (gasal2_set_wavefront)")
    ap.add_argument("--compare-lane-only", action="store_true",
                    help="also time dispatch policy 2 (one lane per pair only) on the same batch")
    ap.add_argument("--autotune", type=int, default=0, metavar="PAIRS",
                    help="tune the lane / one-wave switch on a sample of this many pairs of the batch "
                         "(gasal2_amd.autotune) and time the batch with the winner")
    ap.add_argument("--tb-checkpoint", type=int, default=1, choices=[0, 1, 2],
                    help="LOCAL WITH_TB: checkpointed traceback on the lane path (1, default), also for one-wave "
                         "pairs (2), or direction words stored by the DP (0)")
    ap.add_argument("--tb-fused", type=int, default=0, choices=[0, 1],
                    help="checkpointed LOCAL WITH_TB: the separate traceback kernel walks every pair (0, default) "
                         "or the DP kernel walks each lane-path pair's traceback (1; the DP kernel time then "
                         "includes those walks)")
    ap.add_argument("--ksw-ext", type=int, nargs=2, default=None, metavar=("W", "ZDROP"),
                    help="KSW band half-width and z-drop (seed extension as bwa; default: the reference's none, 0)")
    ap.add_argument("--cpu-pairs", type=int, default=-1,
                    help="oracle sample for cpu_baseline: -1 = sized to about --cpu-seconds, 0 = skip")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES bytes and exit: the reference kernel whose "
                         "FETCH_SIZE / WRITE_SIZE counters (rocprofv3 --pmc) calibrate the DP kernels' byte counts")
    ap.add_argument("--traffic", default=None,
                    help="profile summary with a 'bytes_per_launch: N' line (measured HBM bytes per DP launch); "
                         "default: the committed one for the default workload and mode")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); pairs are independent and would be "
              "sharded across GPUs with no exchange -- see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        a = torch.empty(args.copy_calibration, dtype=torch.uint8, device="cuda")
        a.fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration}), flush=True)
        return

    import gasal2_amd as g

    if args.workload == "band":
        return band_main(args)

    t0 = time.time()
    if args.workload == "full":
        batch = g.synth_batch(2, 2, args.pairs)
        mode = args.mode or "local_tb_second"
        workload = ("full-feature batch: 1M pairs Lq~U{100..300}, target = query + 0-100 bp flanks, "
                    "5% divergence, 0.5% N (N_PENALTY 1), 50% reverse-complemented queries")
    elif args.workload == "seed":
        batch = g.synth_batch(4, 4, args.pairs if args.pairs != 1_000_000 else 2_000_000)
        mode = args.mode or "ksw"
        workload = ("seed extension: 2M pairs Lq, Lt ~ U{100..250}, t = q mutated 3%, h0 ~ U{10..60}; "
                    "KSW in the reference's form (no band, z-drop 0) or tile-banded LOCAL with k_band 32")
    else:
        batch = g.synth_batch(3, 3, args.pairs if args.pairs != 1_000_000 else 500_000)
        mode = args.mode or "local_score"
        workload = ("uneven-length batch: 500k pairs, Lq log-normal (median 400, clipped 50-5000 bp), "
                    "Lt = Lq * U[0.8,1.2], 10% divergence; one launch, per-pair lane/wave switch")
    kw, kz = args.ksw_ext if args.ksw_ext else (-1, 0)
    if args.ksw_ext:
        workload += f"; KSW band w={kw}, z-drop {kz}"
    cells = work_cells(batch, mode, ksw_w=kw)
    log(f"generated {batch.n} pairs, {cells / 1e9:.2f} G cells in {time.time() - t0:.1f}s")

    if mode == "local_tb_second":
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, secondBest=g.TRUE, isReverseComplement=True)
        ops = OPS_SCORE + OPS_SECOND + (0 if args.tb_checkpoint else OPS_DIR)
    elif mode == "local_tb":
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_TB, isReverseComplement=True)
        ops = OPS_SCORE + (0 if args.tb_checkpoint else OPS_DIR)
    elif mode == "local_score":
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITHOUT_START, isReverseComplement=True)
        ops = OPS_SCORE
    elif mode == "local_start":   # forward pass plus the reverse pass up to the start (counted once)
        p = g.Parameters(algo=g.LOCAL, start_pos=g.WITH_START, isReverseComplement=True)
        ops = OPS_SCORE
    elif mode == "global":
        p = g.Parameters(algo=g.GLOBAL, start_pos=g.WITH_START, isReverseComplement=True)
        ops = OPS_SCORE
    elif mode == "ksw":
        p = g.Parameters(algo=g.KSW, start_pos=g.WITHOUT_START)
        ops = OPS_SCORE
    elif mode == "banded":
        p = g.Parameters(algo=g.BANDED, start_pos=g.WITHOUT_START, k_band=32)
        ops = OPS_SCORE
    else:
        p = g.Parameters(algo=g.SEMI_GLOBAL, start_pos=g.WITH_START, isReverseComplement=True,
                         semiglobal_skipping_head=g.TARGET, semiglobal_skipping_tail=g.TARGET)
        ops = OPS_SCORE

    dev = g.device_info()
    peak_tops = dev["n_cu"] * 64 * dev["clock_mhz"] * 1e6 / 1e12
    a = g.Aligner(p, int(batch.qlen.max()), int(batch.tlen.max()), batch.n, scores=(1, 4, 6, 1), n_penalty=1)
    t0 = time.time()
    a.load(batch)
    log(f"uploaded in {time.time() - t0:.2f}s; device {dev}")

    L = g.lib()
    g.set_ksw_extension(kw, kz)
    L.gasal2_set_tb_checkpoint(args.tb_checkpoint)
    L.gasal2_set_tb_fused(args.tb_fused)
    if args.wf:
        L.gasal2_set_wavefront(*args.wf)
    tuned = None
    if args.autotune:
        from gasal2_amd import autotune
        rng = np.random.default_rng(7)
        sample = batch.subset(np.sort(rng.choice(batch.n, min(args.autotune, batch.n), replace=False)))
        t0 = time.time()
        r = autotune.tune(sample, p, n_penalty=1)
        L.gasal2_set_wavefront(*r["best"])
        tuned = {"sample_pairs": sample.n, "best": r["best"], "seconds": round(time.time() - t0, 1),
                 "candidates": len(r["timings"])}
        log(f"autotune on {sample.n} pairs: best {r['best']} of {len(r['timings'])} candidates "
            f"({tuned['seconds']} s)")

    dev_ms = []   # device time of each timed step (first to last event on the stream)

    def timed_steps(policy):
        L.gasal2_set_dispatch(policy)
        for _ in range(args.warmup):
            a.run_resident(timing=True)
            a.sync()
        dp, tb = [], []
        dev_ms.clear()
        a.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a.run_resident(timing=True)
            a.sync()
            dp.append(a.stage_ms(3))
            tb.append(a.stage_ms(4))
            dev_ms.append(a.stage_ms(0))
        a.sync()
        return time.perf_counter() - t0, dp, tb

    lane_only = None
    if args.compare_lane_only:
        w2, dp2, _ = timed_steps(2)
        lane_only = {"gcups": round(cells / (w2 / args.steps) / 1e9, 3), "ms_per_step": round(1e3 * w2 / args.steps, 3),
                     "dp_kernel_ms": round(float(np.mean(dp2)), 3)}
        log(f"one lane per pair only: {lane_only}")
    wall, dp_ms, tb_ms = timed_steps(args.dispatch)
    ms_per_step = 1e3 * wall / args.steps
    gcups = cells / (wall / args.steps) / 1e9
    dp_avg = float(np.mean(dp_ms))
    achieved_tops = cells * ops / (dp_avg * 1e-3) / 1e12
    # packed 16-bit kernels do two cells' ops per lane instruction: their peak is
    # twice the int32 one (only when no one-wave blocks shared the launch)
    paths = a.last_paths()
    packed = (paths & 1) and not (paths & 2) and mode != "ksw"
    peak_used = peak_tops * (2 if packed else 1)
    log(f"step {ms_per_step:.2f} ms ({gcups:.1f} GCUPS); device {np.mean(dev_ms):.2f} ms; DP kernel {dp_avg:.2f} ms; "
        f"traceback {np.mean(tb_ms):.2f} ms")

    # PCIe-inclusive rate of one batch through gasal_aln_async: host batch
    # already filled, timed from the call to results on the host
    qb, tb, n = a.fill(batch, per_pair=False)
    t0 = time.perf_counter()
    a.align_filled(qb, tb, n, poll_s=0)
    pcie_s = time.perf_counter() - t0
    res = a._collect(n, qb)
    L.gasal2_set_dispatch(0)
    log(f"gasal_aln_async incl. H2D/D2H: {pcie_s * 1e3:.1f} ms ({cells / pcie_s / 1e9:.1f} GCUPS)")

    traffic = None
    tpath = args.traffic
    if tpath is None and args.workload == "full" and mode == "local_tb_second" and args.pairs == 1_000_000:
        tpath = os.path.join(ROOT, "profiles", "dp_traffic_local_tb_second_1M.md")
    if tpath is None and args.workload == "seed" and mode == "ksw" and not args.ksw_ext and args.pairs == 1_000_000:
        tpath = os.path.join(ROOT, "profiles", "dp_traffic_ksw_seed_2M.md")
    if tpath and os.path.exists(tpath):
        with open(tpath) as f:
            for ln in f:
                if ln.startswith("bytes_per_launch:"):
                    traffic = int(ln.split(":")[1])

    cpu = None
    if args.cpu_pairs != 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        def oracle_on(m):
            sub = batch.subset(np.arange(m))
            t0 = time.perf_counter()
            r = O.align(sub, p.algo, start_pos=p.start_pos, second_best=bool(p.secondBest),
                        head=p.semiglobal_skipping_head, tail=p.semiglobal_skipping_tail, k_band=p.k_band,
                        is_rc=bool(p.isReverseComplement), n_penalty=1, n_threads=args.cpu_threads, with_cigar=False,
                        ksw_w=kw, ksw_zdrop=kz)
            return sub, r, time.perf_counter() - t0

        if args.cpu_pairs > 0:
            m = min(args.cpu_pairs, batch.n)
        else:   # pilot on 2000 pairs, then a prefix of about --cpu-seconds
            m0 = min(2000, batch.n)
            _, _, s0 = oracle_on(m0)
            m = int(min(batch.n, max(m0, m0 * args.cpu_seconds / max(s0, 1e-3))))
        sub, ref, cpu_s = oracle_on(m)
        sub_cells = work_cells(sub, mode, ksw_w=kw)
        # outputs the mode defines (reference src/res.cpp:8-73)
        keys = ["score"] if p.algo == g.GLOBAL else ["score", "qend", "tend"]
        if (p.algo == g.LOCAL and p.start_pos in (g.WITH_START, g.WITH_TB)) or \
                (p.algo == g.SEMI_GLOBAL and p.start_pos == g.WITH_START):
            keys += ["qstart", "tstart"]
        if p.secondBest and p.algo in (g.LOCAL, g.SEMI_GLOBAL):
            keys += ["s_score", "s_qend", "s_tend"]
        ok = np.ones(m, bool)
        for k in keys:
            ok &= res[k][:m] == ref[k]
        if p.start_pos == g.WITH_TB:
            ok &= res["n_ops"][:m] == ref["n_ops"]
        cpu = {"value": round(sub_cells / cpu_s / 1e9, 4), "unit": "GCUPS", "cores": args.cpu_threads,
               "kind": "port",
               "sample": f"first {m} pairs of the same batch, same mode, oracle OpenMP schedule(dynamic) "
                         f"({cpu_s:.1f}s)",
               "gpu_agreement_on_sample": float(ok.mean())}
        log(f"cpu oracle: {cpu}")

    pipe = pipeline_measure(g, batch, p, args.pipeline) if args.pipeline > 1 else None

    line = {
        "metric": "GCUPS + % int-VALU roofline",
        "value": round(gcups, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "mode": mode, "tb_checkpoint": args.tb_checkpoint, "tb_fused": args.tb_fused, "dispatch": args.dispatch, "wavefront_thresholds": args.wf, "pairs": batch.n, "cells": cells, "scores": [1, 4, 6, 1],
                   "device": dev["name"], "n_cu": dev["n_cu"], "clock_mhz": dev["clock_mhz"]},
        "roofline": ({"bound": "valu",
                      "kernel": "DP (one lane per pair)" if args.workload == "full" else "DP (one lane or one wave per pair)",
                      "ops_per_cell": ops, "packed_int16": bool(packed),
                      "achieved": round(achieved_tops, 3), "peak": round(peak_used, 3), "unit": "TOP/s",
                      "frac": round(achieved_tops / peak_used, 4), "traffic": traffic,
                      "dp_kernel_ms": round(dp_avg, 3)} if mode != "ksw" else
                     # KSW streams each pair's (H, E) row through HBM on every target row:
                     # priced against HBM with the measured traffic when the profile holds it,
                     # else the algorithmic upper bound (8 B per nominal cell, packed row)
                     {"bound": "hbm", "kernel": "k_ksw (one lane per pair)",
                      "bytes_per_cell": 8,
                      "achieved": round((traffic if traffic else cells * 8) / (dp_avg * 1e-3) / 1e9, 1),
                      "peak": 8000.0, "unit": "GB/s",
                      "frac": round((traffic if traffic else cells * 8) / (dp_avg * 1e-3) / 1e9 / 8000.0, 4),
                      "traffic": traffic, "achieved_from": "measured traffic" if traffic else "algorithmic bound",
                      "dp_kernel_ms": round(dp_avg, 3)}),
        "cpu_baseline": cpu,
        "pcie_inclusive_gcups": round(cells / pcie_s / 1e9, 3),
    }
    if lane_only:
        line["lane_only"] = lane_only
    if tuned:
        line["autotune"] = tuned
    if pipe:
        line["pipeline"] = pipe
    print(json.dumps(line), flush=True)
    a.close()


def pipeline_measure(g, batch, p, parts):
    """Whole batch through gasal_aln_async in `parts` sub-batches, host fill to
    results on the host: one storage (each batch waits for the previous one)
    vs two storages used alternately (batch k+1's fill and H2D overlap batch
    k's compute and D2H on the other storage's stream).  Outputs compared."""
    n = batch.n
    cuts = np.linspace(0, n, parts + 1).astype(int)
    subs = [batch.subset(np.arange(cuts[i], cuts[i + 1])) for i in range(parts)]
    mq, mt, mn = int(batch.qlen.max()), int(batch.tlen.max()), int(max(s.n for s in subs))
    keys = ("score", "qend", "tend", "qstart", "tstart")

    def run(nst):
        al = [g.Aligner(p, mq, mt, mn, scores=(1, 4, 6, 1), n_penalty=1) for _ in range(nst)]
        out = [None] * parts
        try:
            for a in al:   # warm: workspace and graph for the sub-batch shape
                qb, tb, m = a.fill(subs[0], per_pair=False)
                a.align_filled(qb, tb, m)
            t0 = time.perf_counter()
            pending = [None] * nst
            for k, sb in enumerate(subs):
                a = al[k % nst]
                if pending[k % nst] is not None:   # this storage's previous batch: wait, collect
                    j, qb0, m0 = pending[k % nst]
                    while g.gasal_is_aln_async_done(a.s) != 0:
                        pass
                    out[j] = {kk: a._collect(m0, qb0)[kk].copy() for kk in keys}
                qb, tb, m = a.fill(sb, per_pair=False)
                g.gasal_aln_async(a.s, qb, tb, m, a.params)
                pending[k % nst] = (k, qb, m)
            for i, a in enumerate(al):
                if pending[i] is not None:
                    j, qb0, m0 = pending[i]
                    while g.gasal_is_aln_async_done(a.s) != 0:
                        pass
                    out[j] = {kk: a._collect(m0, qb0)[kk].copy() for kk in keys}
            return time.perf_counter() - t0, out
        finally:
            for a in al:
                a.close()

    t1, o1 = run(1)
    t2, o2 = run(2)
    same = all(np.array_equal(o1[i][k], o2[i][k]) for i in range(parts) for k in keys)
    cells = float(np.dot(batch.qlen.astype(np.int64), batch.tlen.astype(np.int64)))
    res = {"parts": parts, "one_storage_s": round(t1, 4), "two_storages_s": round(t2, 4),
           "gain": round(t1 / t2, 3), "outputs_identical": bool(same),
           "gcups_one": round(cells / t1 / 1e9, 2), "gcups_two": round(cells / t2 / 1e9, 2)}
    log(f"pipeline: {res}")
    return res


def band_main(args):
    """Learned band vs plain full DP (LOCAL WITH_TB, scores 1/4/6/1) on
    long divergent pairs: a training set (seed 5) labelled by the full DP's
    traceback paths, a held-out set (seed 6) aligned three ways -- full DP,
    learned band + fallback, histogram-peak band + fallback -- each compared
    with the full DP on every output including the CIGAR.  GCUPS-equivalent =
    full sum of Lq*Lt over the path's wall time (features + network + banded
    DP + fallbacks for the banded paths)."""
    import torch
    import gasal2_amd as g
    from gasal2_amd import learned_band as LB
    n = args.band_pairs
    t0 = time.time()
    train = g.synth_batch(5, 5, n)
    held = g.synth_batch(5, 6, n)
    cells = float(np.dot(held.qlen.astype(np.int64), held.tlen.astype(np.int64)))
    log(f"generated 2 x {n} pairs in {time.time() - t0:.1f}s; held-out {cells / 1e12:.3f} T cells")
    # labels
    ref_tr, _, (dmin, dmax) = LB.full_dp(train, labels=True)
    feats, anchors, binws = [], [], []
    for s0 in range(0, n, 16384):
        sub = train.subset(np.arange(s0, min(n, s0 + 16384)))
        a = LB._aligner_for(sub, g.WITH_TB, (1, 4, 6, 1), None)
        try:
            a.load(sub)
            a.prep()
            f, an, bw = LB.features(a, sub.n)
            torch.cuda.synchronize()
            feats.append(f.clone())
            anchors.append(an.cpu().numpy())
            binws.append(bw.cpu().numpy())
        finally:
            a.close()
    feat = torch.cat(feats)
    An, Bn = np.concatenate(anchors), np.concatenate(binws)
    T, w = LB.make_targets(An, Bn, dmin, dmax, ref_tr["score"])
    net = LB.BandNet(seed=0)
    t0 = time.perf_counter()
    hist = LB.train(net, feat, T, w, epochs=args.epochs, batch=1024, lr=2e-3, log=log)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    _, _, Y = net.forward(feat)
    Yn = Y.cpu().numpy()
    mul, add = LB.calibrate_learned(Yn, An, Bn, dmin, dmax, ref_tr["score"])
    half = LB.calibrate_peak(feat.cpu().numpy(), An, Bn, train.qlen, train.tlen, dmin, dmax, ref_tr["score"])
    log(f"trained in {train_s:.1f}s (loss {hist[0]:.3f} -> {hist[-1]:.3f}); learned widen x{mul:.2f}+{add}; "
        f"peak half-width {half:.0f}")
    del ref_tr, feat
    # held-out
    ref, full_s, _ = LB.full_dp(held)
    log(f"full DP: {full_s:.2f}s ({cells / full_s / 1e9:.1f} GCUPS)")
    out = {}
    runs = []
    for cw in args.check_widen:
        runs += [(f"learned@{cw}" if cw != args.check_widen[0] else "learned", LB.learned_proposer(net, mul, add), cw),
                 (f"peak@{cw}" if cw != args.check_widen[0] else "peak", LB.peak_proposer(half), cw)]
    for name, prop, cw in runs:
        res, st = LB.band_align(held, prop, check_widen=cw)
        agree, bad = LB.agreement(res, ref)
        out[name] = {"check_widen": cw, "gcups_equiv": round(cells / st["total_s"] / 1e9, 2), "seconds": round(st["total_s"], 3),
                     "pairs_per_s": round(n / st["total_s"], 1), "hit_rate": round(st["hit_rate"], 5),
                     "final_agreement": agree, "band_cells_fraction": round(st["band_cells_fraction"], 5),
                     "fallback_pairs": st["fallback_pairs"], "flag_counts": st["flag_counts"],
                     "propose_s": round(st["propose_s"], 3), "banded_s": round(st["banded_s"], 3),
                     "fallback_s": round(st["fallback_s"], 3)}
        log(f"{name}: {out[name]}")
    line = {
        "metric": "learned-band vs full-DP pairs/s at 100% agreement",
        "value": out["learned"]["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": round(1e3 * out["learned"]["seconds"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32 (DP), bf16/fp32 (network)", "data": "synthetic",
        "config": {"workload": f"long divergent pairs: {n} train (seed 5) / {n} held-out (seed 6), Lq~U{{1000..10000}}, "
                               "10-15% divergence, Poisson(1/kbp) indel bursts U{5..50}",
                   "mode": "LOCAL WITH_TB", "scores": [1, 4, 6, 1], "held_out_cells": cells},
        "full_dp": {"gcups": round(cells / full_s / 1e9, 2), "seconds": round(full_s, 3),
                    "pairs_per_s": round(n / full_s, 1)},
        "learned": out["learned"], "peak_band": out["peak"],
        "other_check_widths": {k: v for k, v in out.items() if "@" in k},
        "speedup_vs_full": round(full_s / out["learned"]["seconds"], 3),
        "speedup_peak_vs_full": round(full_s / out["peak"]["seconds"], 3),
        "training": {"seconds": round(train_s, 2), "epochs": args.epochs, "loss_first": hist[0], "loss_last": hist[-1],
                     "widen_mul": mul, "widen_add": add, "peak_half_width": half},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
