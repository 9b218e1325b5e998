"""Learned band proposal for long pairs.

A small network reads the k-mer diagonal evidence of a pair and proposes a
diagonal band (centre, half-width); LOCAL alignment then runs inside that band
(gasal2_run_banded) and every pair the banded kernel flags (best cell or
traceback on the band edge, high score at the edge) is re-run with the full DP
(gasal2_run_subset).  The full DP is also the judge: `evaluate` reports the
pre-fallback hit rate and the final agreement with it.

Network: 68 -> 256 -> 256 -> 2, ReLU, bf16 weights and activations with fp32
accumulation, trained on the GPU by the library's own forward/backward/Adam
kernels (include/gasal2_band.h) on labels the full DP's traceback produces:
the diagonal range [dmin, dmax] of each pair's alignment path.  Outputs are in
fine-bin units around the feature anchor: y0 = (centre - anchor) / binw,
y1 = log(half-width / binw).

PyTorch only holds device buffers here; every computation is a HIP kernel of
the library.  No reference counterpart (the reference has one fixed tile band
per batch, src/kernels/banded.h:10-139).
"""
import ctypes as C
import math
import time

import numpy as np

from . import Aligner, Parameters, lib, LOCAL, WITH_TB, WITHOUT_START, FALSE

NFEAT = 68
HIDDEN = 256
MARGIN = 16        # label half-width slack (bases)
MIN_SHARE = 0.9    # a band must meet at least this share of the pair's k-mer hits
CHECK_WIDEN = 128  # check band: the proposed band widened by this many diagonals each side


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned band runs on the GPU")
    return torch


def _p(t):
    return C.c_void_p(t.data_ptr())


def _chk(r, what):
    if r != 0:
        raise RuntimeError(f"{what} failed with HIP error {r}")


class BandNet:
    """The band network's parameters and optimiser state in device memory."""

    SHAPES = [(HIDDEN, NFEAT), (HIDDEN,), (HIDDEN, HIDDEN), (HIDDEN,), (2, HIDDEN), (2,)]

    def __init__(self, seed=0):
        torch = _torch()
        g = np.random.default_rng(seed)
        ws = []
        for shp in self.SHAPES:
            if len(shp) == 2:   # He initialisation
                ws.append(g.normal(0.0, math.sqrt(2.0 / shp[1]), shp).astype(np.float32))
            else:
                ws.append(np.zeros(shp, np.float32))
        self.w = [torch.from_numpy(w).cuda() for w in ws]
        self.m = [torch.zeros_like(w) for w in self.w]
        self.v = [torch.zeros_like(w) for w in self.w]
        self.g = [torch.zeros_like(w) for w in self.w]
        self.wbf = [torch.zeros(w.shape, dtype=torch.int16, device="cuda") if w.dim() == 2 else None for w in self.w]
        self.step = 0
        self._refresh_bf16()

    def _refresh_bf16(self):
        L = lib()
        for w, b in zip(self.w, self.wbf):
            if b is not None:
                _chk(L.gasal2_lb_to_bf16(_p(w), _p(b), w.numel(), None), "gasal2_lb_to_bf16")

    def numpy_params(self):
        return [w.detach().cpu().numpy().astype(np.float64) for w in self.w]

    def load_numpy(self, params):
        torch = _torch()
        self.w = [torch.from_numpy(np.asarray(p, np.float32)).cuda() for p in params]
        self._refresh_bf16()

    def save(self, path):
        np.savez(path, *[w.detach().cpu().numpy() for w in self.w], step=np.array(self.step))

    @classmethod
    def load(cls, path):
        z = np.load(path)   # allow_pickle stays False
        net = cls()
        net.load_numpy([z[f"arr_{i}"] for i in range(6)])
        net.step = int(z["step"])
        return net

    def forward(self, X):
        """X: float32 [B, 68] on the device -> (H1, H2, Y)."""
        torch = _torch()
        L = lib()
        B = X.shape[0]
        H1 = torch.empty((B, HIDDEN), dtype=torch.float32, device="cuda")
        H2 = torch.empty((B, HIDDEN), dtype=torch.float32, device="cuda")
        Y = torch.empty((B, 2), dtype=torch.float32, device="cuda")
        W1, b1, W2, b2, W3, b3 = self.w
        _chk(L.gasal2_lb_dense_fwd(_p(X), B, NFEAT, _p(self.wbf[0]), _p(b1), HIDDEN, _p(H1), 1, None), "dense 1")
        _chk(L.gasal2_lb_dense_fwd(_p(H1), B, HIDDEN, _p(self.wbf[2]), _p(b2), HIDDEN, _p(H2), 1, None), "dense 2")
        _chk(L.gasal2_lb_dense_fwd(_p(H2), B, HIDDEN, _p(self.wbf[4]), _p(b3), 2, _p(Y), 0, None), "dense 3")
        return H1, H2, Y

    def gradients(self, X, T, wt):
        """Loss and gradients (into self.g) of the weighted MSE on one batch."""
        torch = _torch()
        L = lib()
        B = X.shape[0]
        H1, H2, Y = self.forward(X)
        dY = torch.empty_like(Y)
        loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        _chk(L.gasal2_lb_mse(_p(Y), _p(T), _p(wt), B, float(wt.sum().item()), _p(dY), _p(loss), None), "mse")
        dZ3 = torch.empty_like(Y)
        dH2 = torch.empty_like(H2)
        _chk(L.gasal2_lb_dense_bwd(_p(dY), _p(Y), 0, _p(H2), B, HIDDEN, _p(self.wbf[4]), 2, _p(dZ3), _p(dH2),
                                   _p(self.g[4]), _p(self.g[5]), None), "bwd 3")
        dZ2 = torch.empty_like(H2)
        dH1 = torch.empty_like(H1)
        _chk(L.gasal2_lb_dense_bwd(_p(dH2), _p(H2), 1, _p(H1), B, HIDDEN, _p(self.wbf[2]), HIDDEN, _p(dZ2), _p(dH1),
                                   _p(self.g[2]), _p(self.g[3]), None), "bwd 2")
        dZ1 = torch.empty_like(H1)
        _chk(L.gasal2_lb_dense_bwd(_p(dH1), _p(H1), 1, _p(X), B, NFEAT, _p(self.wbf[0]), HIDDEN, _p(dZ1), None,
                                   _p(self.g[0]), _p(self.g[1]), None), "bwd 1")
        return loss

    def train_step(self, X, T, wt, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        loss = self.gradients(X, T, wt)
        self.step += 1
        L = lib()
        for w, m, v, g, bf in zip(self.w, self.m, self.v, self.g, self.wbf):
            _chk(L.gasal2_lb_adam(_p(w), _p(m), _p(v), _p(g), w.numel(), lr, b1, b2, eps, self.step,
                                  _p(bf) if bf is not None else None, None), "adam")
        return loss


# ----------------------------------------------------------------- data ----
def _aligner_for(batch, start_pos, scores, n_penalty):
    p = Parameters(algo=LOCAL, start_pos=start_pos, secondBest=FALSE)
    a = Aligner(p, int(batch.qlen.max()), int(batch.tlen.max()), batch.n, scores=scores, n_penalty=n_penalty)
    return a


def features(a, n):
    """Features of the loaded and prepared batch on aligner `a`."""
    torch = _torch()
    L = lib()
    pq, pt, qoff, toff, qlen, tlen = a.device_views()
    feat = torch.empty((n, NFEAT), dtype=torch.float32, device="cuda")
    anchor = torch.empty(n, dtype=torch.int32, device="cuda")
    binw = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    _chk(L.gasal2_lb_features(C.c_void_p(pq), C.c_void_p(pt), C.c_void_p(qoff), C.c_void_p(toff), C.c_void_p(qlen),
                              C.c_void_p(tlen), n, _p(feat), _p(anchor), _p(binw), None), "features")
    return feat, anchor, binw


def length_order(batch):
    return np.argsort(batch.qlen.astype(np.int64) * batch.tlen.astype(np.int64), kind="stable")


def chunks_by_tb(batch, order, budget=24 << 30):
    """Consecutive chunks of `order` whose full-matrix direction tiles
    (groups x TW x 8QW x 256 B) stay under budget."""
    out, cur, mq, mt = [], [], 0, 0
    for k in order:
        q = (int(batch.qlen[k]) + 7) // 8
        t = (int(batch.tlen[k]) + 7) // 8
        nq, nt = max(mq, q), max(mt, t)
        if cur and ((len(cur) + 1 + 63) // 64) * nt * 8 * nq * 256 > budget:
            out.append(np.array(cur))
            cur, nq, nt = [], q, t
        cur.append(k)
        mq, mt = nq, nt
    if cur:
        out.append(np.array(cur))
    return out


def full_dp(batch, scores=(1, 4, 6, 1), n_penalty=None, labels=False, budget=96 << 30, log=None):
    """LOCAL WITH_TB full DP over the batch in length-sorted chunks on one
    storage (workspace allocated once and reused).  Returns (results dict over
    all pairs, seconds of the device program incl. its first allocation,
    [dmin, dmax] if labels)."""
    torch = _torch()
    L = lib()
    n = batch.n
    res = {k: np.zeros(n, np.int32) for k in ("score", "qend", "tend", "qstart", "tstart")}
    res["n_ops"] = np.zeros(n, np.uint32)
    res["cigars"] = [b""] * n
    dmin = np.zeros(n, np.int32)
    dmax = np.zeros(n, np.int32)
    secs = 0.0
    chunks = chunks_by_tb(batch, length_order(batch), budget)
    p = Parameters(algo=LOCAL, start_pos=WITH_TB, secondBest=FALSE)
    a = Aligner(p, int(batch.qlen.max()), int(batch.tlen.max()), max(len(c) for c in chunks), scores=scores,
                n_penalty=n_penalty)
    try:
        for ch in chunks:
            sub = batch.subset(ch)
            a.load(sub)
            a.sync()
            t0 = time.perf_counter()
            a.run_resident()
            a.sync()
            secs += time.perf_counter() - t0
            if labels:
                score, qend, tend, nops, cigar, cap = a.result_views()
                lo = torch.empty(sub.n, dtype=torch.int32, device="cuda")
                hi = torch.empty(sub.n, dtype=torch.int32, device="cuda")
                _chk(L.gasal2_lb_path_diag(C.c_void_p(cigar), cap, C.c_void_p(nops), C.c_void_p(qend), C.c_void_p(tend),
                                           sub.n, _p(lo), _p(hi), None), "path_diag")
                torch.cuda.synchronize()
                dmin[ch] = lo.cpu().numpy()
                dmax[ch] = hi.cpu().numpy()
            r = a.download()
            for k in ("score", "qend", "tend", "qstart", "tstart", "n_ops"):
                res[k][ch] = r[k]
            for i, k in enumerate(ch):
                res["cigars"][k] = a.cigar(i)
            if log:
                log(f"full DP chunk of {len(ch)} pairs done")
    finally:
        a.close()
    return res, secs, (dmin, dmax)


def make_targets(anchor, binw, dmin, dmax, score):
    """Network targets (fine-bin units) and weights (pairs with a path)."""
    c = 0.5 * (dmin.astype(np.float64) + dmax)
    hw = 0.5 * (dmax.astype(np.float64) - dmin) + MARGIN
    T = np.stack([(c - anchor) / binw, np.log(np.maximum(hw, 1.0) / binw)], 1).astype(np.float32)
    w = (score > 0).astype(np.float32)
    return T, w


def train(net, feat, T, w, epochs=6, batch=1024, lr=2e-3, seed=0, log=None):
    torch = _torch()
    n = feat.shape[0]
    g = np.random.default_rng(seed)
    Td = torch.from_numpy(T).cuda()
    wd = torch.from_numpy(w).cuda()
    hist = []
    for ep in range(epochs):
        perm = torch.from_numpy(g.permutation(n)).cuda()
        tot, cnt = 0.0, 0
        for s in range(0, n - batch + 1 if n >= batch else 1, batch):
            ix = perm[s:s + batch]
            if wd[ix].sum().item() <= 0:
                continue
            loss = net.train_step(feat[ix].contiguous(), Td[ix].contiguous(), wd[ix].contiguous(), lr=lr)
            tot += float(loss.item())
            cnt += 1
        hist.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.4f}")
    return hist


def calibrate_learned(Y, anchor, binw, dmin, dmax, score, coverage=0.995, add=16.0):
    """Smallest width multiplier so that `coverage` of the training paths
    (score > 0) fall inside the band with 8 bases to spare."""
    ok = score > 0
    c = anchor + binw * Y[:, 0]
    base = binw * np.exp(np.minimum(Y[:, 1], 20.0))
    need = np.maximum(c - dmin, dmax - c) + 8.0 - add
    ratio = need[ok] / np.maximum(base[ok], 1e-6)
    return max(1.0, float(np.quantile(ratio, coverage))) if ok.any() else 1.0, add


def calibrate_peak(feat, anchor, binw, qlen, tlen, dmin, dmax, score, coverage=0.995):
    """Fixed half-width (bases) around the histogram peak with the same coverage."""
    ok = score > 0
    f = feat[:, :64]
    bi = np.argmax(f, 1)
    c = anchor + binw * ((bi - 32) + 0.5)
    need = np.maximum(c - dmin, dmax - c) + 8.0
    return float(np.quantile(need[ok], coverage)) if ok.any() else 64.0


def _band_tb_bytes(qlen, tlen, dlo, dhi):
    """Band-compressed traceback words of each pair (the library's sizing in
    gasal2_run_banded): TW strips x the words any strip of the band touches,
    32 bytes per tile."""
    QW = (qlen.astype(np.int64) + 7) // 8
    TW = (tlen.astype(np.int64) + 7) // 8
    span = dhi.astype(np.int64) - dlo.astype(np.int64)
    bw = np.where(span < 0, 1, ((span + 14) >> 3) + 1)
    bw = np.clip(bw, 1, np.maximum(QW, 1))
    return TW * bw * 32


def _plan_chunks(order, est, budget, max_chunk):
    """Consecutive pieces of `order` whose estimated bytes stay under budget."""
    out, start, acc = [], 0, 0
    for i, k in enumerate(order):
        b = int(est[k])
        if i > start and (acc + b > budget or i - start >= max_chunk):
            out.append(order[start:i])
            start, acc = i, 0
        acc += b
    if start < len(order):
        out.append(order[start:])
    return out


def band_align(batch, propose, scores=(1, 4, 6, 1), n_penalty=None, start_pos=WITH_TB, check_widen=CHECK_WIDEN,
               chunk=None, tb_budget=None, sample=4096, log=None):
    """Banded LOCAL with full-DP fallback over the batch, in length-sorted
    chunks (one storage each).  Large chunks keep enough one-lane-per-pair
    waves in flight; the band-compressed traceback tiles of a chunk must fit
    `tb_budget` (default: 70% of the device memory free at the call).  With
    chunk=None the chunks are planned from the bands
    proposed for a sample of `sample` pairs (their mean share of the full
    tile grid, with a margin); a chunk whose proposed bands still exceed the
    budget is split.  An int `chunk` fixes the chunk size instead.
    `propose(a, n) -> (dlo, dhi, pre)` host arrays (pre: pairs sent to the
    full DP outright), called on the prepared chunk.  Returns (results over
    all pairs, stats with timings and flag counts)."""
    torch = _torch()
    n = batch.n
    res = {k: np.zeros(n, np.int32) for k in ("score", "qend", "tend", "qstart", "tstart")}
    res["n_ops"] = np.zeros(n, np.uint32)
    res["cigars"] = [b""] * n
    flags_all = np.zeros(n, np.uint8)
    band_cells = 0.0
    tp = tb_ = tf = 0.0
    order = length_order(batch)
    if tb_budget is None:
        tb_budget = int(0.7 * torch.cuda.mem_get_info()[0])
    QW = (batch.qlen.astype(np.int64) + 7) // 8
    TW = (batch.tlen.astype(np.int64) + 7) // 8
    if chunk is None:
        # band width relative to the query, from a sample's proposals
        rng = np.random.default_rng(0)
        smp = np.sort(rng.choice(n, min(sample, n), replace=False))
        sb = batch.subset(smp)
        a = _aligner_for(sb, start_pos, scores, n_penalty)
        try:
            a.load(sb)
            a.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            a.prep()
            dlo, dhi, _ = propose(a, sb.n)
            tp += time.perf_counter() - t0
        finally:
            a.close()
        # the mean share of the full tile grid, with a margin: chunks are sums
        # over many pairs, and one that still comes out over budget is split
        ratio = _band_tb_bytes(sb.qlen, sb.tlen, dlo, dhi) / np.maximum(TW[smp] * QW[smp] * 32, 1)
        r = min(1.0, 1.1 * float(ratio.mean()) + 1e-3)
        est = np.maximum(TW * np.ceil(r * QW) * 32, 32)
        # longest pairs first: the budget's last, partly filled chunk then
        # holds the shortest pairs, which cost least at its lower occupancy
        pieces = _plan_chunks(order[::-1], est, tb_budget, 1 << 20)
    else:
        pieces = [order[s0:s0 + chunk] for s0 in range(0, n, chunk)]
    sizes = []
    while pieces:
        ch = pieces.pop(0)
        sub = batch.subset(ch)
        a = _aligner_for(sub, start_pos, scores, n_penalty)
        try:
            a.load(sub)
            a.sync()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            a.prep()
            dlo, dhi, pre = propose(a, sub.n)
            t1 = time.perf_counter()
            need = int(_band_tb_bytes(sub.qlen, sub.tlen, dlo, dhi).sum()) if start_pos == WITH_TB else 0
            if need > tb_budget and len(ch) > 1:   # split and redo this chunk
                tp += t1 - t0
                h = len(ch) // 2
                pieces[:0] = [ch[:h], ch[h:]]
                continue
            flags = a.run_banded(dlo, dhi, check_widen=check_widen) | pre
            sizes.append(len(ch))
            t2 = time.perf_counter()
            a.run_subset(np.nonzero(flags)[0])
            a.sync()
            t3 = time.perf_counter()
            tp += t1 - t0
            tb_ += t2 - t1
            tf += t3 - t2
            r = a.download()
            for k in ("score", "qend", "tend", "qstart", "tstart"):
                if r.get(k) is not None:
                    res[k][ch] = r[k]
            if start_pos == WITH_TB:
                res["n_ops"][ch] = r["n_ops"]
                for i, k in enumerate(ch):
                    res["cigars"][k] = a.cigar(i)
            flags_all[ch] = flags
            band_cells += float(np.sum(np.clip(np.minimum(dhi.astype(np.int64) - dlo + 1, sub.qlen.astype(np.int64)), 0, None)
                                       * sub.tlen.astype(np.int64)))
        finally:
            a.close()
        if log:
            log(f"banded chunk of {len(ch)} pairs: {int(np.count_nonzero(flags))} fallbacks")
    fb = int(np.count_nonzero(flags_all))
    pc = batch.qlen.astype(np.int64) * batch.tlen.astype(np.int64)
    cells = float(pc.sum())
    stats = {"propose_s": tp, "banded_s": tb_, "fallback_s": tf, "total_s": tp + tb_ + tf,
             "fallback_pairs": fb, "hit_rate": 1.0 - fb / max(n, 1),
             "flag_counts": {str(b): int(np.sum((flags_all >> b) & 1)) for b in range(5)},
             "band_cells_fraction": band_cells / max(cells, 1.0),
             "fallback_cells_fraction": float(pc[flags_all != 0].sum()) / max(cells, 1.0),
             "chunk_pairs": sizes}
    return res, stats


def learned_proposer(net, mul, add, min_share=MIN_SHARE):
    def propose(a, n):
        torch = _torch()
        L = lib()
        feat, anchor, binw = features(a, n)
        _, _, Y = net.forward(feat)
        lo = torch.empty(n, dtype=torch.int32, device="cuda")
        hi = torch.empty(n, dtype=torch.int32, device="cuda")
        pre = torch.empty(n, dtype=torch.uint8, device="cuda")
        _chk(L.gasal2_lb_band_from_net(_p(Y), _p(feat), _p(anchor), _p(binw), n, float(mul), float(add),
                                       float(min_share), _p(lo), _p(hi), _p(pre), None), "band_from_net")
        torch.cuda.synchronize()
        return lo.cpu().numpy(), hi.cpu().numpy(), pre.cpu().numpy()
    return propose


def peak_proposer(half, frac=0.0, min_share=MIN_SHARE):
    def propose(a, n):
        torch = _torch()
        L = lib()
        feat, anchor, binw = features(a, n)
        _, _, _, _, qlen, tlen = a.device_views()
        lo = torch.empty(n, dtype=torch.int32, device="cuda")
        hi = torch.empty(n, dtype=torch.int32, device="cuda")
        pre = torch.empty(n, dtype=torch.uint8, device="cuda")
        _chk(L.gasal2_lb_band_peak(_p(feat), _p(anchor), _p(binw), C.c_void_p(qlen), C.c_void_p(tlen), n, float(half),
                                   float(frac), float(min_share), _p(lo), _p(hi), _p(pre), None), "band_peak")
        torch.cuda.synchronize()
        return lo.cpu().numpy(), hi.cpu().numpy(), pre.cpu().numpy()
    return propose


def agreement(res, ref, with_cigar=True):
    keys = ["score", "qend", "tend", "qstart", "tstart"]
    ok = np.ones(len(ref["score"]), bool)
    for k in keys:
        ok &= res[k] == ref[k]
    if with_cigar and "cigars" in res:
        ok &= res["n_ops"] == ref["n_ops"]
        ok &= np.array([x == y for x, y in zip(res["cigars"], ref["cigars"])])
    return float(ok.mean()), np.nonzero(~ok)[0]
