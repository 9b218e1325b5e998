"""Benchmark of the batched alignment hot path on one MI355X.

Workload (default): the full-feature batch -- 1,000,000 pairs, query length
uniform 100-300 bp, target = query + 0-100 bp random flanks at 5% divergence,
0.5% N bases with N_PENALTY 1, reverse-complemented queries on half the pairs --
aligned LOCAL WITH_TB with second best (score, ends, starts, second best, CIGAR),
int32 scores, 4-bit packed on the device.  One step = one pass of the device
pipeline (pack, reverse/complement, shape sort, DP, traceback) over the whole
batch with the inputs already resident in HBM.

Prints one JSON line.  `value` is GCUPS = sum over pairs of Lq*Lt (unpadded
cells) / step time.  `roofline` prices the whole step against the
integer-VALU peak (N_CU * 64 lanes * clock; twice that when the batch ran a
packed 16-bit kernel, two cells per lane instruction) at the fixed
algorithmic ops per cell of the mode (16 with traceback, 11 without):
achieved = GCUPS * ops; `roofline.dp_kernel` prices the DP kernel alone (HIP
events on its stream) at the ops it executes; `traffic` is the DP kernel's
HBM bytes per launch from rocprofv3 counter runs.  `cpu_baseline` times the CPU
oracle (an OpenMP restatement of the reference kernels) on a bounded prefix of
the same batch on this host and reports the GPU's agreement with it there.

The default run also measures every other workload of the benchmark set under
`configs` (each with its own agreement check against the oracle): the 10k-pair
short-read batch, the full-feature batch's GLOBAL / SEMI_GLOBAL modes and its
packed-input form, steady state end to end (eight full-feature batches
streamed through the library's stream driver, first H2D to the last results
on the host; and FASTA files in, SAM out), the uneven-length batch, the seed-extension batch (KSW, both
forms, and tile-banded) and the learned band on long divergent pairs.
`--headline-only` skips them; `--workload` / `--mode` run one configuration.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Fixed algorithmic int ops per cell of the metric: the recurrence (score
# select 2, diagonal add 1, H max 2, gap open 1, E' 2, F' 2, running max 1) is
# 11; with traceback 16 (+ 2 source compares, 2 continuation compares, 1
# nibble insert).  N handling and position bookkeeping are not counted.
OPS_SCORE = 11
OPS_TB = 16
# What the DP kernel alone executes: second best adds 4 (compare, select, max,
# running max); direction bits stored by the DP add 8 (four comparisons, four
# inserts).  With the checkpointed traceback the DP stores no direction bits
# (the traceback kernel recomputes the tiles on the path).
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


def host_info(threads):
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": model, "threads": threads,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def default_cpu_threads():
    """The OpenMP default: OMP_NUM_THREADS when set (the host's CPU share),
    else every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


MODES = {
    # mode: (algo, start, second, head, tail, k_band)
    "local_tb_second": ("LOCAL", "WITH_TB", True, None, None, 0),
    "local_tb": ("LOCAL", "WITH_TB", False, None, None, 0),
    "local_score": ("LOCAL", "WITHOUT_START", False, None, None, 0),
    "local_start": ("LOCAL", "WITH_START", False, None, None, 0),
    "global": ("GLOBAL", "WITH_START", False, None, None, 0),
    "global_tb": ("GLOBAL", "WITH_TB", False, None, None, 0),
    "semi_global": ("SEMI_GLOBAL", "WITH_START", False, "TARGET", "TARGET", 0),
    "ksw": ("KSW", "WITHOUT_START", False, None, None, 0),
    "banded": ("BANDED", "WITHOUT_START", False, None, None, 32),
}


def params_for(g, mode, rc, packed=False):
    algo, start, second, head, tail, kb = MODES[mode]
    kw = dict(algo=getattr(g, algo), start_pos=getattr(g, start), secondBest=g.TRUE if second else g.FALSE,
              isReverseComplement=rc, isPacked=packed, k_band=kb)
    if head:
        kw.update(semiglobal_skipping_head=getattr(g, head), semiglobal_skipping_tail=getattr(g, tail))
    return g.Parameters(**kw)


def metric_ops(mode):
    return OPS_TB if MODES[mode][1] == "WITH_TB" else OPS_SCORE


def kernel_ops(mode, tb_checkpoint):
    algo, start, second = MODES[mode][:3]
    ops = OPS_SCORE + (OPS_SECOND if second else 0)
    if start == "WITH_TB" and (not tb_checkpoint or algo != "LOCAL"):   # checkpoints: LOCAL only
        ops += OPS_DIR
    return ops


def oracle_run(g, batch, p, m, threads, kw, kz, with_cigar=False):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    sub = batch.subset(np.arange(m)) if m < batch.n else batch
    t0 = time.perf_counter()
    r = O.align(sub, p.algo, start_pos=p.start_pos, second_best=bool(p.secondBest),
                head=p.semiglobal_skipping_head, tail=p.semiglobal_skipping_tail, k_band=p.k_band,
                is_rc=bool(p.isReverseComplement), n_penalty=1, n_threads=threads, with_cigar=with_cigar,
                ksw_w=kw, ksw_zdrop=kz)
    return sub, r, time.perf_counter() - t0


def agreement(g, p, res, ref, m, cigar_pairs=0, batch=None):
    """Share of the first m pairs whose every defined output equals the
    oracle's (reference src/res.cpp:8-73); CIGAR bytes in the reference layout
    on the first cigar_pairs pairs when the oracle ran with CIGARs."""
    keys = ["score"] if p.algo == g.GLOBAL else ["score", "qend", "tend"]
    if (p.algo == g.LOCAL and p.start_pos in (g.WITH_START, g.WITH_TB)) or \
            (p.algo == g.SEMI_GLOBAL and p.start_pos == g.WITH_START):
        keys += ["qstart", "tstart"]
    if p.secondBest and p.algo in (g.LOCAL, g.SEMI_GLOBAL):
        keys += ["s_score", "s_qend", "s_tend"]
    ok = np.ones(m, bool)
    for k in keys:
        ok &= res[k][:m] == ref[k]
    if p.start_pos == g.WITH_TB and p.algo in (g.LOCAL, g.GLOBAL):
        ok &= res["n_ops"][:m] == ref["n_ops"]
        if cigar_pairs and "cigars" in ref and batch is not None:
            ql = res["cigar_qlayout"]
            for k in range(min(cigar_pairs, m)):
                slot = 8 * ((int(batch.qlen[k]) + 7) // 8)
                nb = min(int(ref["n_ops"][k]), slot)
                o = int(batch.qoff[k])
                if bytes(ql[o:o + nb]) != ref["cigars"][k][:nb]:
                    ok[k] = False
    return float(ok.mean())


class Run:
    """One workload resident on the device in one storage."""

    def __init__(self, g, batch, p, cells, steps, warmup):
        self.g, self.batch, self.p, self.cells = g, batch, p, cells
        self.steps, self.warmup = steps, warmup
        self.a = g.Aligner(p, int(batch.qlen.max()), int(batch.tlen.max()), batch.n, scores=(1, 4, 6, 1),
                           n_penalty=1)
        self.a.load(batch)

    def timed(self, policy=0):
        L, a = self.g.lib(), self.a
        L.gasal2_set_dispatch(policy)
        try:
            for _ in range(self.warmup):
                a.run_resident(timing=True)
                a.sync()
            dp, tb, dev, rc = [], [], [], []
            a.sync()
            t0 = time.perf_counter()
            for _ in range(self.steps):
                a.run_resident(timing=True)
                a.sync()
                dp.append(a.stage_ms(3))
                tb.append(a.stage_ms(4))
                dev.append(a.stage_ms(0))
                rc.append(a.stage_ms(2))
            a.sync()
            wall = time.perf_counter() - t0
        finally:
            L.gasal2_set_dispatch(0)
        ms = 1e3 * wall / self.steps
        return {"gcups": round(self.cells / (wall / self.steps) / 1e9, 3), "ms_per_step": round(ms, 3),
                "device_ms": round(float(np.mean(dev)), 3), "dp_kernel_ms": round(float(np.mean(dp)), 3),
                "traceback_ms": round(float(np.mean(tb)), 3), "rc_ms": round(float(np.mean(rc)), 4),
                "paths": self.a.last_paths()}

    def rc_bytes(self):
        """Algorithmic HBM bytes of the reverse/complement stage: every padded
        4-bit word of a sequence whose op is nonzero read once and written
        once (1 byte per base)."""
        b = self.batch
        if not self.p.isReverseComplement:
            return 0
        qw = (b.qlen.astype(np.int64) + 7) // 8
        tw = (b.tlen.astype(np.int64) + 7) // 8
        return int(8 * (qw[b.qop != 0].sum() + tw[b.top != 0].sum()))

    def results(self):
        return self.a.download()

    def pcie_inclusive(self):
        """One batch through gasal_aln_async, host batch already filled,
        timed from the call to results on the host."""
        a = self.a
        qb, tb, n = a.fill(self.batch, per_pair=False)
        t0 = time.perf_counter()
        a.align_filled(qb, tb, n, poll_s=0)
        s = time.perf_counter() - t0
        return {"gcups": round(self.cells / s / 1e9, 3), "ms": round(1e3 * s, 3)}, a._collect(n, qb)

    def close(self):
        self.a.close()


def check(g, run, res, m, threads, kw=-1, kz=0, cigar_pairs=0):
    sub, ref, _ = oracle_run(g, run.batch, run.p, m, threads, kw, kz, with_cigar=cigar_pairs > 0)
    return agreement(g, run.p, res, ref, len(sub.qlen), cigar_pairs, sub)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--mode", default=None, choices=sorted(MODES),
                    help="one configuration: the mode (default local_tb_second for the full-feature batch, "
                         "local_score for the uneven one, ksw for the seed batch)")
    ap.add_argument("--workload", default=None, choices=["full", "uneven", "seed", "band"],
                    help="one configuration: full: 1M pairs of 100-300 bp (+flanks); uneven: 500k pairs, log-normal "
                         "lengths 50-5000 bp; seed: 2M seed extensions; band: learned band vs full DP on 1-10 kbp "
                         "pairs (train seed 5, held-out seed 6)")
    ap.add_argument("--headline-only", action="store_true",
                    help="measure the headline configuration only (no other workloads)")
    ap.add_argument("--packed", action="store_true", help="one configuration with isPacked=true host inputs")
    ap.add_argument("--band-pairs", type=int, default=200_000, help="pairs per set for the learned band")
    ap.add_argument("--band-chunk", type=int, default=0,
                    help="learned band: pairs per banded launch (0: planned from the bands and the free device memory)")
    ap.add_argument("--stream-batches", type=int, default=8,
                    help="steady state: full-feature batches streamed end to end (0: skip)")
    ap.add_argument("--stream-sub", type=int, default=1_000_000, help="steady state: pairs per sub-batch")
    ap.add_argument("--stream-storages", type=int, default=3, help="steady state: storages (host threads)")
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
    ap.add_argument("--tile-budget", type=float, default=None, metavar="GB",
                    help="traceback tile budget (gasal2_set_tile_budget; library default 64 GB)")
    ap.add_argument("--tb-checkpoint", type=int, default=1, choices=[0, 1, 2],
                    help="LOCAL WITH_TB: checkpointed traceback on the lane path (1, default), also for one-wave "
                         "pairs (2), or direction words stored by the DP (0)")
    ap.add_argument("--tb-fused", type=int, default=0, choices=[0, 1],
                    help="checkpointed LOCAL WITH_TB: the separate traceback kernel walks every pair (0, default) "
                         "or the DP kernel walks each lane-path pair's traceback (1; the DP kernel time then "
                         "includes those walks)")
    ap.add_argument("--pk16", type=int, default=1, choices=[0, 1, 2],
                    help="packed 16-bit kernels: 1 where values fit (default), 0 int32 only, 2 two pairs per lane "
                         "for every LOCAL mode (gasal2_set_pk16)")
    ap.add_argument("--ksw-ext", type=int, nargs=2, default=None, metavar=("W", "ZDROP"),
                    help="KSW band half-width and z-drop (seed extension as bwa; default: the reference's none, 0)")
    ap.add_argument("--cpu-pairs", type=int, default=-1,
                    help="oracle sample for cpu_baseline: -1 = sized to about --cpu-seconds, 0 = skip")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=default_cpu_threads())
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
        line = band_main(args)
        print(json.dumps(line), flush=True)
        return
    single = args.workload is not None or args.mode is not None or args.packed
    workload = args.workload or "full"
    t_start = time.time()

    t0 = time.time()
    if workload == "full":
        batch = g.synth_batch(2, 2, args.pairs)
        mode = args.mode or "local_tb_second"
        wdesc = ("full-feature batch: 1M pairs Lq~U{100..300}, target = query + 0-100 bp flanks, "
                 "5% divergence, 0.5% N (N_PENALTY 1), 50% reverse-complemented queries")
        rc = True
    elif workload == "seed":
        batch = g.synth_batch(4, 4, args.pairs if args.pairs != 1_000_000 else 2_000_000)
        mode = args.mode or "ksw"
        wdesc = ("seed extension: 2M pairs Lq, Lt ~ U{100..250}, t = q mutated 3%, h0 ~ U{10..60}; "
                 "KSW in the reference's form (no band, z-drop 0) or tile-banded LOCAL with k_band 32")
        rc = False
    else:
        batch = g.synth_batch(3, 3, args.pairs if args.pairs != 1_000_000 else 500_000)
        mode = args.mode or "local_score"
        wdesc = ("uneven-length batch: 500k pairs, Lq log-normal (median 400, clipped 50-5000 bp), "
                 "Lt = Lq * U[0.8,1.2], 10% divergence; one launch, per-pair lane/wave switch")
        rc = False
    kw, kz = args.ksw_ext if args.ksw_ext else (-1, 0)
    if args.ksw_ext:
        wdesc += f"; KSW band w={kw}, z-drop {kz}"
    if args.packed:
        wdesc += "; isPacked=true host inputs"
    cells = work_cells(batch, mode, ksw_w=kw)
    log(f"generated {batch.n} pairs, {cells / 1e9:.2f} G cells in {time.time() - t0:.1f}s")

    p = params_for(g, mode, rc, args.packed)
    ops = metric_ops(mode)
    kops = kernel_ops(mode, args.tb_checkpoint)

    dev = g.device_info()
    peak_tops = dev["n_cu"] * 64 * dev["clock_mhz"] * 1e6 / 1e12
    L = g.lib()
    g.set_ksw_extension(kw, kz)
    L.gasal2_set_tb_checkpoint(args.tb_checkpoint)
    if args.tile_budget is not None:
        L.gasal2_set_tile_budget(int(args.tile_budget * (1 << 30)))
    L.gasal2_set_tb_fused(args.tb_fused)
    L.gasal2_set_pk16(args.pk16)
    if args.wf:
        L.gasal2_set_wavefront(*args.wf)
    t0 = time.time()
    run = Run(g, batch, p, cells, args.steps, args.warmup)
    log(f"uploaded in {time.time() - t0:.2f}s; device {dev}")

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

    lane_only = run.timed(2) if args.compare_lane_only else None
    if lane_only:
        log(f"one lane per pair only: {lane_only}")
    t = run.timed(args.dispatch)
    gcups, ms_per_step, dp_avg = t["gcups"], t["ms_per_step"], t["dp_kernel_ms"]
    # packed 16-bit kernels do two cells' ops per lane instruction: their peak is
    # twice the int32 one (only when no one-wave blocks shared the launch)
    packed16 = bool((t["paths"] & 1) and not (t["paths"] & 2) and mode != "ksw")
    peak_used = peak_tops * (2 if packed16 else 1)
    log(f"step {ms_per_step:.2f} ms ({gcups:.1f} GCUPS); device {t['device_ms']:.2f} ms; DP kernel {dp_avg:.2f} ms; "
        f"traceback {t['traceback_ms']:.2f} ms")
    pcie, res = run.pcie_inclusive()
    log(f"gasal_aln_async incl. H2D/D2H: {pcie}")
    rcb = run.rc_bytes()
    rc = None
    if rcb and t["rc_ms"] > 0:
        rc_gbs = rcb / (t["rc_ms"] * 1e-3) / 1e9
        rc = {"kernel": "k_rc", "ms": t["rc_ms"], "algorithmic_bytes": rcb, "achieved": round(rc_gbs, 1),
              "peak": 8000.0, "unit": "GB/s", "frac": round(rc_gbs / 8000.0, 4)}
        log(f"reverse/complement: {rc}")

    traffic = None
    tpath = args.traffic
    if tpath is None and workload == "full" and mode == "local_tb_second" and args.pairs == 1_000_000 \
            and not args.packed:
        tpath = os.path.join(ROOT, "profiles", "dp_traffic_local_tb_second_1M.md")
    if tpath is None and workload == "seed" and mode == "ksw" and not args.ksw_ext and args.pairs == 1_000_000:
        tpath = os.path.join(ROOT, "profiles", "dp_traffic_ksw_seed_2M.md")
    if tpath and os.path.exists(tpath):
        with open(tpath) as f:
            for ln in f:
                if ln.startswith("bytes_per_launch:"):
                    traffic = int(ln.split(":")[1])

    cpu = None
    if args.cpu_pairs != 0:
        if args.cpu_pairs > 0:
            m = min(args.cpu_pairs, batch.n)
        else:   # pilot on 2000 pairs, then a prefix of about --cpu-seconds
            m0 = min(2000, batch.n)
            _, _, s0 = oracle_run(g, batch, p, m0, args.cpu_threads, kw, kz)
            m = int(min(batch.n, max(m0, m0 * args.cpu_seconds / max(s0, 1e-3))))
        sub, ref, cpu_s = oracle_run(g, batch, p, m, args.cpu_threads, kw, kz)
        sub_cells = work_cells(sub, mode, ksw_w=kw)
        cpu = {"value": round(sub_cells / cpu_s / 1e9, 4), "unit": "GCUPS", "cores": args.cpu_threads,
               "kind": "port",
               "sample": f"first {m} pairs of the same batch, same mode, oracle OpenMP schedule(dynamic) "
                         f"({cpu_s:.1f}s)",
               "gpu_agreement_on_sample": agreement(g, p, res, ref, len(sub.qlen)),
               "host": host_info(args.cpu_threads)}
        log(f"cpu oracle: {cpu}")
    run.close()

    if mode == "ksw":
        # KSW streams each pair's (H, E) row through HBM on every target row:
        # priced against HBM with the algorithmic bytes (8 B per nominal cell)
        roof = {"bound": "hbm", "kernel": "k_ksw (one lane per pair)", "bytes_per_cell": 8,
                "achieved": round(cells * 8 / (dp_avg * 1e-3) / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(cells * 8 / (dp_avg * 1e-3) / 1e9 / 8000.0, 4), "traffic": traffic,
                "dp_kernel_ms": dp_avg}
    else:
        ach = gcups * ops / 1e3
        dp_ach = cells * kops / (dp_avg * 1e-3) / 1e12
        roof = {"bound": "valu", "kernel": "whole device step (pack, reverse/complement, shape sort, DP, traceback)",
                "ops_per_cell": ops, "packed_int16": packed16,
                "achieved": round(ach, 3), "peak": round(peak_used, 3), "unit": "TOP/s",
                "frac": round(ach / peak_used, 4), "traffic": traffic,
                "dp_kernel": {"kernel": "DP (one lane per pair)" if workload == "full"
                              else "DP (one lane or one wave per pair)",
                              "ms": dp_avg, "ops_per_cell": kops, "achieved": round(dp_ach, 3),
                              "frac": round(dp_ach / peak_used, 4),
                              "note": "ops the DP kernel itself executes over its HIP-event duration"}}
    line = {
        "metric": "GCUPS + % int-VALU roofline",
        "value": gcups,
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": wdesc, "mode": mode, "tb_checkpoint": args.tb_checkpoint, "tb_fused": args.tb_fused,
                   "dispatch": args.dispatch, "pk16": args.pk16, "wavefront_thresholds": args.wf, "pairs": batch.n, "cells": cells,
                   "scores": [1, 4, 6, 1], "device": dev["name"], "n_cu": dev["n_cu"],
                   "clock_mhz": dev["clock_mhz"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "device_ms": t["device_ms"],
        "traceback_ms": t["traceback_ms"],
        "reverse_complement": rc,
        "pcie_inclusive": pcie,
    }
    if lane_only:
        line["lane_only"] = lane_only
    if tuned:
        line["autotune"] = tuned
    if not single and not args.headline_only:
        line["configs"] = other_configs(g, args, batch, peak_tops,
                                        res if workload == "full" and mode == "local_tb_second" and not args.packed
                                        else None)
    line["bench_seconds"] = round(time.time() - t_start, 1)
    print(json.dumps(line), flush=True)


def step_line(g, run, mode, peak_tops, threads, m, kw=-1, kz=0, cigar_pairs=0, policy=0):
    t = run.timed(policy)
    res = run.results()
    ops = metric_ops(mode)
    packed16 = bool((t["paths"] & 1) and not (t["paths"] & 2) and mode != "ksw")
    peak = peak_tops * (2 if packed16 else 1)
    out = {"gcups": t["gcups"], "ms_per_step": t["ms_per_step"], "dp_kernel_ms": t["dp_kernel_ms"],
           "traceback_ms": t["traceback_ms"], "ops_per_cell": ops, "packed_int16": packed16,
           "valu_frac": round(t["gcups"] * ops / 1e3 / peak, 4) if mode != "ksw" else None}
    if m:
        out["oracle_agreement"] = check(g, run, res, m, threads, kw, kz, cigar_pairs)
        out["agreement_pairs"] = min(m, run.batch.n)
    return out


def other_configs(g, args, full, peak_tops, headline_res=None):
    """Every other workload of the benchmark set, each with its agreement
    against the oracle (on all pairs where the oracle finishes in seconds,
    else on a prefix)."""
    L = g.lib()
    thr = args.cpu_threads
    out = {}

    def timed(name, fn):
        t0 = time.time()
        try:
            out[name] = fn()
        except Exception as e:   # a failed extra measurement is reported, not fatal to the headline
            out[name] = {"error": repr(e)}
        out[name]["seconds"] = round(time.time() - t0, 1)
        log(f"{name}: {out[name]}")

    # 10k short reads, LOCAL score + ends, every pair checked
    def c1():
        b = g.synth_batch(1, 1, 10_000)
        r = Run(g, b, params_for(g, "local_score", False), work_cells(b, "local_score"), 20, 3)
        try:
            return dict(step_line(g, r, "local_score", peak_tops, thr, b.n), pairs=b.n,
                        workload="10k pairs, q 150 bp, t = 50 + q + 50 (2% subst, 0.5% indels); LOCAL WITHOUT_START")
        finally:
            r.close()
    timed("short_reads_10k", c1)

    # full-feature batch: the other modes, and the headline with packed inputs
    fcells = work_cells(full, "global")
    for name, mode, packed in (("full_global_start", "global", False), ("full_semi_global_start", "semi_global", False),
                               ("full_local_tb_second_packed_inputs", "local_tb_second", True)):
        def c2(mode=mode, packed=packed):
            r = Run(g, full, params_for(g, mode, True, packed), fcells, 5, 1)
            try:
                o = step_line(g, r, mode, peak_tops, thr, 20_000, cigar_pairs=2000 if MODES[mode][1] == "WITH_TB" else 0)
                o["pcie_inclusive"] = r.pcie_inclusive()[0]
                return o
            finally:
                r.close()
        timed(name, c2)

    if args.stream_batches > 0 and headline_res is not None:
        sp = dict(batches=args.stream_batches, sub_pairs=args.stream_sub, storages=args.stream_storages, threads=thr)
        timed("full_steady_state_end_to_end",
              lambda: stream_measure(g, full, headline_res, params_for(g, "local_tb_second", True), **sp))
        timed("full_steady_state_end_to_end_packed_inputs",
              lambda: stream_measure(g, full, headline_res, params_for(g, "local_tb_second", True, True), **sp))

        def files():
            import tempfile
            with tempfile.TemporaryDirectory() as d:
                return files_measure(g, full, params_for(g, "local_tb_second", True), args.stream_sub,
                                     args.stream_storages, d)
        timed("full_fasta_to_sam", files)

    # uneven lengths: persistent queue with the lane/wave switch vs one lane per pair
    def c3():
        b = g.synth_batch(3, 3, 500_000)
        c = work_cells(b, "local_score")
        o = {}
        r = Run(g, b, params_for(g, "local_score", False), c, 5, 1)
        try:
            o["local_score"] = step_line(g, r, "local_score", peak_tops, thr, 3000)
            lo = r.timed(2)
            o["local_score_one_lane_per_pair"] = {"gcups": lo["gcups"], "ms_per_step": lo["ms_per_step"]}
        finally:
            r.close()
        # WITH_TB with checkpoints on the one-wave path too (gasal2_set_tb_checkpoint(2)):
        # this batch's long pairs store twice the bytes of direction words and
        # run about 15% faster
        L.gasal2_set_tb_checkpoint(2)
        r = Run(g, b, params_for(g, "local_tb", False), c, 3, 1)
        try:
            o["local_tb"] = step_line(g, r, "local_tb", peak_tops, thr, 2000, cigar_pairs=500)
            o["local_tb"]["tb_checkpoint"] = 2
        finally:
            r.close()
            L.gasal2_set_tb_checkpoint(args.tb_checkpoint)
        o["pairs"] = b.n
        o["cells"] = c
        return o
    timed("uneven_500k", c3)

    # seed extension: KSW reference form, KSW w 50 z-drop 100, tile-banded k_band 32
    def c4():
        b = g.synth_batch(4, 4, 2_000_000)
        o = {"pairs": b.n}
        for name, mode, ext in (("ksw_reference_form", "ksw", (-1, 0)), ("ksw_w50_zdrop100", "ksw", (50, 100)),
                                ("banded_k32", "banded", (-1, 0))):
            g.set_ksw_extension(*ext)
            c = work_cells(b, mode, ksw_w=ext[0])
            r = Run(g, b, params_for(g, mode, False), c, 3, 1)
            try:
                o[name] = step_line(g, r, mode, peak_tops, thr, 20_000, kw=ext[0], kz=ext[1])
                o[name]["cells"] = c
            finally:
                r.close()
        g.set_ksw_extension(-1, 0)
        return o
    timed("seed_extension_2M", c4)

    timed("learned_band", lambda: band_main(args))
    L.gasal2_set_dispatch(0)
    return out


def concat_batches(g, bs):
    """One Batch holding the pairs of several, in order."""
    qo, to = np.cumsum([0] + [len(b.q) for b in bs]), np.cumsum([0] + [len(b.t) for b in bs])
    cat = lambda f: np.concatenate([getattr(b, f) for b in bs])
    return g.Batch(cat("q"), cat("t"), np.concatenate([b.qoff + qo[i] for i, b in enumerate(bs)]),
                   np.concatenate([b.toff + to[i] for i, b in enumerate(bs)]), cat("qlen"), cat("tlen"),
                   cat("qop"), cat("top"))


def stream_measure(g, full, headline_res, p, batches, sub_pairs, storages, threads, reps=5, sample=96):
    """Steady state end to end through the library's stream driver
    (gasal2_stream_align, include/gasal2_stream.h): `batches` full-feature
    batches -- the headline batch, then the same generator with other seeds --
    streamed as sub-batches of `sub_pairs` pairs through `storages` storages,
    each driven by its own host thread (fill the pinned pages, gasal_aln_async,
    wait, copy the results out; the reference's example-program pattern,
    test_prog/test_prog.cpp:237-330, one stream per storage src/ctors.cpp:118)
    with the sub-batches' H2D and device programs in submission order, so one
    sub-batch's H2D and D2H overlap another's device work.  Clock: first
    gasal_aln_async (its H2D) to the last results on the host (also reported
    from the first host fill).  Warm storages, median of `reps` runs.  The
    headline batch's outputs are compared with its single-batch run on every
    pair, the other batches' with the oracle on their first `sample` pairs."""
    bs = [full] + [g.synth_batch(2, 20000 + i, full.n) for i in range(1, batches)]
    big = concat_batches(g, bs)
    ps = g.PairSet.from_batch(big, pack=bool(p.isPacked))
    st = g.Stream(p, sub_pairs, storages, int(big.qlen.max()), int(big.tlen.max()))
    runs = []
    try:
        r, _ = ps.align(p, n_penalty=1, stream=st)   # warm: workspace, graphs, first touch
        for _ in range(reps):
            r, stats = ps.align(p, n_penalty=1, stream=st)
            runs.append(stats)
    finally:
        st.close()
    keys = ["score", "qend", "tend", "qstart", "tstart", "s_score", "s_qend", "s_tend", "n_ops"]
    n = full.n
    same = all(np.array_equal(r[k][:n], headline_res[k]) for k in keys if k in r and k in headline_res)
    agree = []
    for i in range(1, batches):
        sub, ref, _ = oracle_run(g, bs[i], p, sample, threads, -1, 0)
        agree.append(agreement(g, p, {k: v[i * n:(i + 1) * n] for k, v in r.items()}, ref, len(sub.qlen)))
    secs = sorted(x["seconds"] for x in runs)
    med = secs[len(secs) // 2]
    best = min(runs, key=lambda x: abs(x["seconds"] - med))
    cells = best["cells"]
    res = {"pairs": big.n, "batches": batches, "sub_batch_pairs": sub_pairs, "storages": storages,
           "sub_batches": best["sub_batches"], "packed_inputs": bool(p.isPacked),
           "gcups": round(cells / med / 1e9, 1), "seconds": round(med, 4),
           "gcups_runs": [round(cells / x / 1e9, 1) for x in secs],
           "gcups_from_fill": round(cells / best["seconds_from_fill"] / 1e9, 1),
           "host_seconds": {k: round(best[k], 4) for k in ("fill_seconds", "launch_seconds", "wait_seconds",
                                                            "copy_seconds")},
           "headline_batch_identical_to_single_batch": bool(same),
           "other_batches_oracle_agreement": agree, "oracle_sample_pairs": sample,
           "clock": "first gasal_aln_async (H2D) to the last results on the host; median of runs"}
    log(f"stream: {res}")
    return res


def files_measure(g, full, p, sub_pairs, storages, tmpdir):
    """The host input path from files: the headline batch written as two
    FASTA files (op markers in the headers), read back by gasal2_pairs_read,
    aligned through the stream driver and written as SAM (gasal2_sam_write);
    each stage timed."""
    qp, tp, sp = (os.path.join(tmpdir, x) for x in ("q.fa", "t.fa", "out.sam"))
    for path, data, offs, lens, ops in ((qp, full.q, full.qoff, full.qlen, full.qop),
                                        (tp, full.t, full.toff, full.tlen, full.top)):
        with open(path, "wb") as f:
            mk = b"></+"
            raw = data.tobytes()
            f.write(b"".join(mk[o:o + 1] + b"p%d\n" % k + raw[a:a + l] + b"\n"
                             for k, (a, l, o) in enumerate(zip(offs.tolist(), lens.tolist(), ops.tolist()))))
    fasta_bytes = os.path.getsize(qp) + os.path.getsize(tp)
    t0 = time.perf_counter()
    ps = g.PairSet.read(qp, tp, names=True, pack=bool(p.isPacked))
    t_read = time.perf_counter() - t0
    st = g.Stream(p, sub_pairs, storages, int(full.qlen.max()), int(full.tlen.max()))
    try:
        cap = int(full.qlen.max() + full.tlen.max()) + 8
        ps.align(p, n_penalty=1, stream=st, cigar_cap=cap)   # warm
        t0 = time.perf_counter()
        r, stats = ps.align(p, n_penalty=1, stream=st, cigar_cap=cap)
        t_align = time.perf_counter() - t0
        t0 = time.perf_counter()
        ps.write_sam(sp, r)
        t_sam = time.perf_counter() - t0
    finally:
        st.close()
        ps.close()
    sam_bytes = os.path.getsize(sp)
    for x in (qp, tp, sp):
        os.remove(x)
    cells = stats["cells"]
    res = {"pairs": full.n, "fasta_bytes": fasta_bytes, "read_s": round(t_read, 3),
           "read_gb_s": round(fasta_bytes / t_read / 1e9, 2), "align_s": round(t_align, 4),
           "align_gcups": round(cells / t_align / 1e9, 1), "sam_bytes": sam_bytes, "sam_write_s": round(t_sam, 3),
           "sam_gb_s": round(sam_bytes / t_sam / 1e9, 2),
           "files_to_sam_gcups": round(cells / (t_read + t_align + t_sam) / 1e9, 1)}
    log(f"files: {res}")
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
        res, st = LB.band_align(held, prop, check_widen=cw, chunk=args.band_chunk or None, log=log)
        agree, bad = LB.agreement(res, ref)
        out[name] = {"check_widen": cw, "gcups_equiv": round(cells / st["total_s"] / 1e9, 2), "seconds": round(st["total_s"], 3),
                     "pairs_per_s": round(n / st["total_s"], 1), "hit_rate": round(st["hit_rate"], 5),
                     "final_agreement": agree, "band_cells_fraction": round(st["band_cells_fraction"], 5),
                     "fallback_pairs": st["fallback_pairs"], "flag_counts": st["flag_counts"],
                     "propose_s": round(st["propose_s"], 3), "banded_s": round(st["banded_s"], 3),
                     "fallback_s": round(st["fallback_s"], 3), "chunk_pairs": st["chunk_pairs"]}
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
                   "mode": "LOCAL WITH_TB", "scores": [1, 4, 6, 1], "held_out_cells": cells,
                   "band_chunk_pairs": args.band_chunk},
        "full_dp": {"gcups": round(cells / full_s / 1e9, 2), "seconds": round(full_s, 3),
                    "pairs_per_s": round(n / full_s, 1)},
        "learned": out["learned"], "peak_band": out["peak"],
        "other_check_widths": {k: v for k, v in out.items() if "@" in k},
        "speedup_vs_full": round(full_s / out["learned"]["seconds"], 3),
        "speedup_peak_vs_full": round(full_s / out["peak"]["seconds"], 3),
        "training": {"seconds": round(train_s, 2), "epochs": args.epochs, "loss_first": hist[0], "loss_last": hist[-1],
                     "widen_mul": mul, "widen_add": add, "peak_half_width": half},
    }
    return line


if __name__ == "__main__":
    main()
