"""Learned-band references on the CPU: the backward formulas against finite
differences of the exact (float64, unrounded) network, the feature reference
against hand-built cases, and the path-diagonal walk."""
import numpy as np

import lbref as R


def _params(rng, scale=0.2):
    W1 = rng.normal(0, scale, (16, 8))
    b1 = rng.normal(0, 0.1, 16)
    W2 = rng.normal(0, scale, (16, 16))
    b2 = rng.normal(0, 0.1, 16)
    W3 = rng.normal(0, scale, (2, 16))
    b3 = rng.normal(0, 0.1, 2)
    return [W1, b1, W2, b2, W3, b3]


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(7)
    p = _params(rng)
    X = rng.normal(0, 1, (12, 8))
    T = rng.normal(0, 1, (12, 2))
    w = rng.uniform(0.5, 1.5, 12)
    _, grads = R.mlp_grads(p, X, T, w, exact=True)
    eps = 1e-6
    for pi in range(6):
        flat = p[pi].reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp, _ = R.mse(R.mlp_forward(p, X, exact=True)[2], T, w)
            flat[idx] = old - eps
            lm, _ = R.mse(R.mlp_forward(p, X, exact=True)[2], T, w)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = grads[pi].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (pi, idx, fd, an)


def test_planted_mistake_is_caught():
    """A sign slip in one backward term must fail the finite-difference check."""
    rng = np.random.default_rng(8)
    p = _params(rng)
    X = rng.normal(0, 1, (10, 8))
    T = rng.normal(0, 1, (10, 2))
    w = np.ones(10)
    _, g = R.mlp_grads(p, X, T, w, exact=True)
    bad = -g[2]   # planted: wrong sign on dW2
    eps = 1e-6
    flat = p[2].reshape(-1)
    old = flat[3]
    flat[3] = old + eps
    lp, _ = R.mse(R.mlp_forward(p, X, exact=True)[2], T, w)
    flat[3] = old - eps
    lm, _ = R.mse(R.mlp_forward(p, X, exact=True)[2], T, w)
    flat[3] = old
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - bad.reshape(-1)[3]) > 1e-4 or abs(fd) < 1e-9


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.140625], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert list(R.bf16_round(x)) == [1.0, 1.0, 1.0078125, -2.5, 3.140625]


def test_features_identical_sequences_put_all_hits_on_diagonal_zero():
    rng = np.random.default_rng(3)
    q = bytes(rng.choice(list(b"ACGT"), 2000).tolist())
    f, anchor, bw = R.features(q, q)
    nhits = 2000 - R.K + 1
    assert abs(f[67] * 2001 - nhits) < 1e-2
    # every hit is on diagonal 0, which falls in one fine bin
    assert abs(f[:64].sum() - 1.0) < 1e-6 and f[:64].max() == 1.0
    b = int(np.argmax(f[:64]))
    assert (b - 32) * bw <= -anchor < (b - 31) * bw


def test_features_shifted_copy_moves_the_diagonal():
    rng = np.random.default_rng(4)
    q = bytes(rng.choice(list(b"ACGT"), 3000).tolist())
    t = bytes(rng.choice(list(b"ACGT"), 300).tolist()) + q   # query base i sits at target row i + 300
    f, anchor, bw = R.features(q, t)
    b = int(np.argmax(f[:64]))
    d = anchor + (b - 32) * bw
    assert d <= -300 < d + bw                                  # diagonal = query column - target row


def test_features_n_breaks_kmers():
    q = b"ACGTACGTAN" * 50    # every 11-base window holds an N
    f, _, _ = R.features(q, q)
    assert f[67] == 0.0


def test_path_diag_walk():
    # end at (tend=20, qend=25): d=5; walking back: 10 M, 3 D (d += 3), 4 I (d -= 4)
    cg = bytes([(0 | (10 << 2)), (2 | (3 << 2)), (3 | (4 << 2))])
    assert R.path_diag(cg, 25, 20) == (4, 8)
