"""CPU references for the learned-band kernels (gasal2_amd/csrc/band.hip):
k-mer diagonal features, the bf16 dense layers with fp32 accumulation and
their backward pass, MSE, Adam, the band formulas and path diagonals.  Test
infrastructure only."""
import math

import numpy as np

K = 11
COARSE = 256
FINE = 64
NFEAT = 68
CODE2 = {1: 0, 3: 1, 4: 2, 7: 3}   # nibble -> 2-bit code (A, C, T, G)


def _nibbles(seq: bytes):
    return [b & 15 for b in seq]


def _kmers(nib, positions):
    out = {}
    for i in positions:
        code, ok = 0, True
        for t in range(K):
            c = CODE2.get(nib[i + t])
            if c is None:
                ok = False
                break
            code = (code << 2) | c
        if ok:
            out[i] = code
    return out


def features(q: bytes, t: bytes):
    """(feat[68] float32, anchor, binw) as k_features computes them."""
    Lq, Lt = len(q), len(t)
    nq, nt = Lq - K + 1, Lt - K + 1
    stride = (nq + 4095) // 4096 if nq > 4096 else 1
    qn, tn = _nibbles(q), _nibbles(t)
    first = {}
    for i, code in _kmers(qn, range(0, max(nq, 0), stride)).items():
        if code not in first or i < first[code]:
            first[code] = i
    hits = []
    for j, code in _kmers(tn, range(0, max(nt, 0))).items():
        i = first.get(code)
        if i is not None:
            hits.append(i - j)          # query column - target row
    span = Lq + Lt + 1
    coarse = np.zeros(COARSE, np.int64)
    for d in hits:
        coarse[((d + Lt) * COARSE) // span] += 1
    bi = int(np.argmax(coarse)) if coarse.max() > 0 else 0
    anchor = ((2 * bi + 1) * span) // (2 * COARSE) - Lt
    bw = (span + 1023) // 1024 if span > 4096 else 4
    fine = np.zeros(FINE, np.int64)
    for d in hits:
        off = d - anchor
        fb = (off // bw) + FINE // 2          # floor division, as the kernel's
        if 0 <= fb < FINE:
            fine[fb] += 1
    total = len(hits)
    inv = np.float32(1.0) / np.float32(max(total, 1))
    f = np.zeros(NFEAT, np.float32)
    f[:FINE] = fine.astype(np.float32) * inv
    f[64] = np.float32(math.log2(Lq + 1.0) / 14.0)
    f[65] = np.float32(math.log2(Lt + 1.0) / 14.0)
    f[66] = np.float32((Lt - Lq) / (Lt + Lq if Lt + Lq > 0 else 1))
    f[67] = np.float32(total / (min(Lq, Lt) + 1))
    return f, anchor, bw


# ------------------------------------------------------------------- MLP ----
def bf16_round(x):
    """float32 -> bf16 (round to nearest even) -> float32, as f2bf/bf2f."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF
    return ((u >> 16) << 16).astype(np.uint32).view(np.float32)


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFFFFFF) >> 16).astype(np.uint16)


def dense_fwd(X, Wf, b, relu):
    """Y = act(bf16(X) bf16(W)^T + b) with float64 accumulation (the kernel
    accumulates in fp32; tests allow for the difference)."""
    Z = bf16_round(X).astype(np.float64) @ bf16_round(Wf).astype(np.float64).T + b.astype(np.float64)
    return np.maximum(Z, 0.0) if relu else Z


def dense_bwd(dY, Y, relu, X, Wf):
    dZ = dY * (Y > 0) if relu else dY
    dX = dZ @ bf16_round(Wf).astype(np.float64)
    dW = dZ.T @ X.astype(np.float64)
    db = dZ.sum(0)
    return dZ, dX, dW, db


def mse(Yp, Tg, wt):
    ws = wt.sum()
    e = Yp - Tg
    loss = float((wt[:, None] * e * e).sum() / ws)
    dY = 2.0 * wt[:, None] * e / ws
    return loss, dY


def mlp_forward(params, X, exact=False):
    """params: (W1, b1, W2, b2, W3, b3); exact=True skips bf16 rounding
    (float64 throughout) for finite differences."""
    W1, b1, W2, b2, W3, b3 = params
    if exact:
        H1 = np.maximum(X @ W1.T + b1, 0.0)
        H2 = np.maximum(H1 @ W2.T + b2, 0.0)
        return H1, H2, H2 @ W3.T + b3
    H1 = dense_fwd(X, W1, b1, True)
    H2 = dense_fwd(H1.astype(np.float32), W2, b2, True)
    return H1, H2, dense_fwd(H2.astype(np.float32), W3, b3, False)


def mlp_grads(params, X, Tg, wt, exact=False):
    """Analytic gradients of the weighted MSE (the kernels' formulas)."""
    W1, b1, W2, b2, W3, b3 = params
    H1, H2, Y = mlp_forward(params, X, exact)
    loss, dY = mse(Y, Tg, wt)
    rw = (lambda w: w) if exact else (lambda w: bf16_round(w).astype(np.float64))
    dZ3 = dY
    dW3, db3 = dZ3.T @ H2, dZ3.sum(0)
    dH2 = dZ3 @ rw(W3)
    dZ2 = dH2 * (H2 > 0)
    dW2, db2 = dZ2.T @ H1, dZ2.sum(0)
    dH1 = dZ2 @ rw(W2)
    dZ1 = dH1 * (H1 > 0)
    dW1, db1 = dZ1.T @ X, dZ1.sum(0)
    return loss, (dW1, db1, dW2, db2, dW3, db3)


def adam(w, m, v, g, lr, b1, b2, eps, step):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1.0 / (1 - b1 ** step), 1.0 / (1 - b2 ** step)
    return w - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v


def hit_share(f, anchor, bw, lo, hi):
    share = 0.0
    for b in range(FINE):
        b0 = anchor + (b - FINE // 2) * bw
        if b0 + bw - 1 >= lo and b0 <= hi:
            share += float(f[b])
    return share


def band_from_net(Y, anchor, binw, mul, add):
    c = anchor + binw * Y[:, 0]
    h = binw * np.exp(np.minimum(Y[:, 1], 20.0)) * mul + add
    return np.floor(c - h).astype(np.int32), np.ceil(c + h).astype(np.int32)


def path_diag(cigar: bytes, qend, tend):
    d = qend - tend
    lo = hi = d
    for b in cigar:
        op, cnt = b & 3, b >> 2
        if op == 2:
            d += cnt
        elif op == 3:
            d -= cnt
        lo, hi = min(lo, d), max(hi, d)
    return lo, hi
