"""Independent formulations used to pin the oracle (tests only).

Nothing here mirrors the reference's strip/tile code: these are textbook
full-matrix dynamic programs over explicit alignment-automaton states, an
exhaustive enumerator of alignments for tiny inputs, and closed forms.
"""
import itertools

NEG = -10 ** 9


def nib(b):
    return b & 15


def sub_score(a, b, match, mismatch, nval):
    """Substitution of two bases (nibble equality; N involvement gives nval,
    or None meaning N is an ordinary base)."""
    ra, rb = nib(a), nib(b)
    if nval is not None and (ra == 0xE or rb == 0xE):
        return nval
    return match if ra == rb else -mismatch


def automaton_local(q, t, sc=(1, 4, 6, 1), nval=0):
    """Local alignment where every gap run is opened right after an aligned
    pair and gaps of different kinds never touch (states M, D, I).
    Returns (score, H) with H[i][j] = best score of an alignment ending at
    target i, query j (any state), floored at 0."""
    m, mm, go, ge = sc
    oe = go + ge
    Lt, Lq = len(t), len(q)
    M = [[NEG] * Lq for _ in range(Lt)]
    D = [[NEG] * Lq for _ in range(Lt)]
    I = [[NEG] * Lq for _ in range(Lt)]
    H = [[0] * Lq for _ in range(Lt)]
    best = 0
    for i in range(Lt):
        for j in range(Lq):
            s = sub_score(q[j], t[i], m, mm, nval)
            prev = 0
            if i > 0 and j > 0:
                prev = max(0, M[i - 1][j - 1], D[i - 1][j - 1], I[i - 1][j - 1])
            M[i][j] = prev + s
            if i > 0:
                D[i][j] = max(M[i - 1][j] - oe, D[i - 1][j] - ge)
            if j > 0:
                I[i][j] = max(M[i][j - 1] - oe, I[i][j - 1] - ge)
            H[i][j] = max(0, M[i][j], D[i][j], I[i][j])
            best = max(best, H[i][j])
    return best, H


def enumerate_local(q, t, sc=(1, 4, 6, 1)):
    """Exhaustive: every pair of substrings, every alignment path made of
    aligned pairs and gap runs, starting with an aligned pair, no two gap runs
    adjacent.  Exponential; for len <= 4."""
    m, mm, go, ge = sc

    def paths(a, b):
        # yields lists of ops over (a, b) that consume both fully
        if a == 0 and b == 0:
            yield []
            return
        if a > 0 and b > 0:
            for rest in paths(a - 1, b - 1):
                yield ["M"] + rest
        for k in range(1, a + 1):
            for rest in paths(a - k, b):
                if rest and rest[0] in "DI":
                    continue
                yield ["D"] * k + rest
        for k in range(1, b + 1):
            for rest in paths(a, b - k):
                if rest and rest[0] in "DI":
                    continue
                yield ["I"] * k + rest

    best = 0
    for ti, tj in itertools.combinations_with_replacement(range(len(t) + 1), 2):
        for qi, qj in itertools.combinations_with_replacement(range(len(q) + 1), 2):
            tt, qq = t[ti:tj], q[qi:qj]
            if not tt or not qq:
                continue
            for p in paths(len(tt), len(qq)):
                if p[0] != "M":
                    continue
                x = y = 0
                sc_ = 0
                prev = None
                for op in p:
                    if op == "M":
                        sc_ += m if nib(tt[x]) == nib(qq[y]) else -mm
                        x += 1
                        y += 1
                    elif op == "D":
                        sc_ -= (go + ge) if prev != "D" else ge
                        x += 1
                    else:
                        sc_ -= (go + ge) if prev != "I" else ge
                        y += 1
                    prev = op
                best = max(best, sc_)
    return best


def automaton_global(q, t, sc=(1, 4, 6, 1), nval=None):
    """Global alignment, gaps opened after aligned pairs, with the leading-gap
    costs the reference's boundary rows give: a leading query gap of k bases
    costs 0 for k == 1 and open + extend*(k-1) for k >= 2; a leading target gap
    of k bases costs open + extend*k."""
    m, mm, go, ge = sc
    oe = go + ge
    Lt, Lq = len(t), len(q)
    M = [[NEG] * Lq for _ in range(Lt)]
    D = [[NEG] * Lq for _ in range(Lt)]
    I = [[NEG] * Lq for _ in range(Lt)]
    for i in range(Lt):
        for j in range(Lq):
            s = sub_score(q[j], t[i], m, mm, nval)
            if i == 0 and j == 0:
                prev = 0
            elif i == 0:
                prev = 0 if j == 1 else -(go + ge * (j - 1))
            elif j == 0:
                prev = -(go + ge * i)
            else:
                prev = max(M[i - 1][j - 1], D[i - 1][j - 1], I[i - 1][j - 1])
            M[i][j] = prev + s
            if i > 0:
                D[i][j] = max(M[i - 1][j] - oe, D[i - 1][j] - ge)
            if j > 0:
                I[i][j] = max(M[i][j - 1] - oe, I[i][j - 1] - ge)
    return max(M[Lt - 1][Lq - 1], D[Lt - 1][Lq - 1], I[Lt - 1][Lq - 1])


def gotoh_local(q, t, sc=(1, 4, 6, 1), nval=0):
    """Textbook Smith-Waterman-Gotoh (gaps opened from H)."""
    m, mm, go, ge = sc
    oe = go + ge
    Lt, Lq = len(t), len(q)
    H = [[0] * (Lq + 1) for _ in range(Lt + 1)]
    E = [[NEG] * (Lq + 1) for _ in range(Lt + 1)]
    F = [[NEG] * (Lq + 1) for _ in range(Lt + 1)]
    best = 0
    for i in range(1, Lt + 1):
        for j in range(1, Lq + 1):
            E[i][j] = max(H[i - 1][j] - oe, E[i - 1][j] - ge)
            F[i][j] = max(H[i][j - 1] - oe, F[i][j - 1] - ge)
            H[i][j] = max(0, H[i - 1][j - 1] + sub_score(q[j - 1], t[i - 1], m, mm, nval), E[i][j], F[i][j])
            best = max(best, H[i][j])
    return best


def gotoh_semiglobal_target_target(q, t, sc=(1, 4, 6, 1), nval=0):
    """Textbook Gotoh, query aligned end to end inside the target (target head
    and tail free), with the reference's top boundary row H(-1, c) = 0 for
    c == 0 and -(open + extend*c) for c >= 1; score = max over target rows of the
    last query column."""
    m, mm, go, ge = sc
    oe = go + ge
    Lt, Lq = len(t), len(q)
    # index shift: row/col 0 is the boundary
    H = [[0] * (Lq + 1) for _ in range(Lt + 1)]
    E = [[NEG] * (Lq + 1) for _ in range(Lt + 1)]
    F = [[NEG] * (Lq + 1) for _ in range(Lt + 1)]
    for c in range(Lq):
        H[0][c + 1] = 0 if c == 0 else -(go + ge * c)
    for r in range(Lt + 1):
        H[r][0] = 0
    best = NEG
    for i in range(1, Lt + 1):
        for j in range(1, Lq + 1):
            E[i][j] = max(H[i - 1][j] - oe, E[i - 1][j] - ge)
            F[i][j] = max(H[i][j - 1] - oe, F[i][j - 1] - ge)
            H[i][j] = max(H[i - 1][j - 1] + sub_score(q[j - 1], t[i - 1], m, mm, nval), E[i][j], F[i][j])
        best = max(best, H[i][Lq])
    return best


def key_order_first(H, Lq_pad, rows, cols, pred):
    """First cell (row, col) in the key order (row // 8, col, row % 8) whose
    value satisfies pred."""
    for s in range((rows + 7) // 8):
        for c in range(cols):
            for r8 in range(8):
                r = 8 * s + r8
                if r < rows and pred(H[r][c]):
                    return r, c
    return None


def second_best_definition(H_key_cells):
    """second = max h(x) over cells x with h(x) < (running max up to and
    including x in key order); returns (value, index of first such cell)."""
    run = 0
    best2, where = 0, None
    for idx, h in enumerate(H_key_cells):
        run = max(run, h)
        if h < run and h > best2:
            best2, where = h, idx
    return best2, where


def nibble_string(words, n):
    s = "".join("%08x" % w for w in words)
    return s[:n]
