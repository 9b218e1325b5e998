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


# --------------------------------------------------------------------------
# Full-matrix formulations over nibble lists (one value per base, already
# padded the way the batch is).  Rows are target positions, columns query
# positions.  The boundary values encode the reference's documented rules as
# data; nothing below walks 8-row strips or packed words.

NEG16 = -32768
NONE_, QUERY_, TARGET_, BOTH_ = 0, 1, 2, 3


def t16(v):
    """Two's-complement truncation to 16 bits (the int16 row kept between
    8-row strips)."""
    v &= 0xFFFF
    return v - 0x10000 if v >= 0x8000 else v


def nibs(seq, pad_to=None, pad=0xE):
    out = [b & 15 for b in seq]
    if pad_to is not None:
        out += [pad] * (pad_to - len(out))
    return out


def sub_nib(a, b, match, mismatch, nval, n_value=0xE):
    """Substitution with the local N rule: a base whose nibble is N's scores
    nval, otherwise match / -mismatch by nibble equality."""
    if a == n_value or b == n_value:
        return nval
    return match if a == b else -mismatch


def semiglobal_matrix(qn, tn, head, sc=(1, 4, 6, 1), nval=0):
    """Classic Gotoh (gaps opened from H), every cell of the len(tn) x len(qn)
    matrix, no floor.  Boundaries by the head rule:
      query head free (QUERY, BOTH): the row above row 0 is H = 0, E = 0;
      otherwise H = 0 at column 0 and -(open + extend*c) at column c >= 1, E = -32768;
      target head free (TARGET, BOTH): left of column 0, H = 0 and diagonal 0;
      otherwise H(r, -1) = -(open + extend*r) and diagonal(r, 0) = 0 for r == 0,
      -(open + extend*r) for r >= 1;  F(r, -1) = -32768.
    The H and E of every row r = 8k - 1 reach row 8k truncated to 16 bits.
    Returns H (list of rows)."""
    m, mm, go, ge = sc
    oe = go + ge
    R, Cn = len(tn), len(qn)
    if head in (QUERY_, BOTH_):
        topH, topE = [0] * Cn, [0] * Cn
    else:
        topH, topE = [0] + [-(go + ge * c) for c in range(1, Cn)], [NEG16] * Cn
    if head in (TARGET_, BOTH_):
        leftH, leftD = [0] * R, [0] * R
    else:
        leftH = [-(go + ge * r) for r in range(R)]
        leftD = [0 if r == 0 else -(go + ge * r) for r in range(R)]
    H = [[0] * Cn for _ in range(R)]
    E = [[0] * Cn for _ in range(R)]
    F = [[0] * Cn for _ in range(R)]
    for r in range(R):
        for c in range(Cn):
            if r == 0:
                uh, ue = topH[c], topE[c]
            elif r % 8 == 0:
                uh, ue = t16(H[r - 1][c]), t16(E[r - 1][c])
            else:
                uh, ue = H[r - 1][c], E[r - 1][c]
            if c == 0:
                dg = leftD[r]
            elif r == 0:
                dg = topH[c - 1]
            elif r % 8 == 0:
                dg = t16(H[r - 1][c - 1])
            else:
                dg = H[r - 1][c - 1]
            lh, lf = (leftH[r], NEG16) if c == 0 else (H[r][c - 1], F[r][c - 1])
            F[r][c] = max(lh - oe, lf - ge)
            E[r][c] = max(uh - oe, ue - ge)
            H[r][c] = max(dg + sub_nib(qn[c], tn[r], m, mm, nval), E[r][c], F[r][c])
    return H


def _semi_top_init(head, c, sc):
    go, ge = sc[2], sc[3]
    if head in (QUERY_, BOTH_):
        return 0
    return 0 if c == 0 else -(go + ge * c)


def semiglobal_reference(q, t, head, tail, sc=(1, 4, 6, 1), nval=0, with_start=False):
    """Score, ends, second best and (with_start) starts of the reference's
    semi-global kernel, from semiglobal_matrix.  Outputs follow the kernel's
    documented bookkeeping (semiglobal_kernel_template.h:49-74,160-218,269-381):
      TARGET tail: over rows r < Lt of column Lq-1 in row order, the first
        strict maximum -> (score, target end); query end stays Lt;
        second = the largest value below the running maximum, same rows;
      QUERY tail: then over columns c of the last computed row (16-bit), c < Lq
        for the maximum, c < Lt for the second (columns never computed hold
        their initial value); a found column sets the target end to Lq;
      starts: the same DP on reversed copies (padding nibble 0) from 8-row strip
        TW - (end>>3) - 1 (one strip earlier when positive), stopping before a
        strip once the running maximum reaches the forward score; start =
        (length - 1) - reversed end."""
    qlen, tlen = len(q), len(t)
    QW, TW = (qlen + 7) // 8, (tlen + 7) // 8
    qn, tn = nibs(q, 8 * QW), nibs(t, 8 * TW)
    H = semiglobal_matrix(qn, tn, head, sc, nval)
    o = dict(score=NEG16, qend=tlen, tend=qlen, s_score=NEG16, s_qend=tlen, s_tend=qlen)
    if tail in (TARGET_, BOTH_):
        for r in range(8 * TW):
            h = H[r][qlen - 1]
            if h > o["score"] and r < tlen:
                o["tend"], o["score"] = r, h
            if h > o["s_score"] and h < o["score"] and r < tlen:
                o["s_tend"], o["s_score"] = r, h
    if tail in (QUERY_, BOTH_):
        for c in range(max(8 * QW, tlen, qlen)):
            v = t16(H[8 * TW - 1][c]) if c < 8 * QW else _semi_top_init(head, c, sc)
            if v > o["score"] and c < qlen:
                o["qend"], o["score"] = c, v
            if v > o["s_score"] and v < o["score"] and c < tlen:
                o["s_qend"], o["s_score"] = c, v
        if o["qend"] != tlen:
            o["tend"] = qlen
        if o["s_qend"] != tlen:
            o["s_tend"] = qlen
    if with_start:
        rq = nibs(q[::-1], 8 * QW, pad=0)
        rt = nibs(t[::-1], 8 * TW, pad=0)
        d = TW - ((o["tend"] >> 3) + 1)
        s0 = d - 1 if d > 0 else d
        rows = [rt[a] if 0 <= a < 8 * TW else 0 for a in range(8 * s0, 8 * TW)]
        Hr = semiglobal_matrix(rq, rows, head, sc, nval) if rows else []
        best, rend, qend = NEG16, 0, o["qend"]
        last = None
        for i in range(s0, TW):
            if not best < o["score"]:
                break
            last = i
            if tail in (TARGET_, BOTH_):
                for m8 in range(8):
                    a = 8 * i + m8
                    h = Hr[a - 8 * s0][qlen - 1]
                    if h > best and a < tlen:
                        rend, best = a, h
        if tail in (QUERY_, BOTH_):
            for c in range(max(8 * QW, tlen, qlen)):
                if last is not None and c < 8 * QW:
                    v = t16(Hr[8 * last + 7 - 8 * s0][c])
                else:
                    v = _semi_top_init(head, c, sc)
                if v > best and c < qlen:
                    qend, best = c, v
            if qend != tlen:
                rend = qlen
        o["tstart"] = (tlen - 1) - rend
        o["qstart"] = (qlen - 1) - qend
    return o


def banded_reference(q, t, k_band, sc=(1, 4, 6, 1), nval=0):
    """Smith-Waterman-Gotoh (gaps opened from H, floor 0) over the cells of the
    reference's tile band only (banded.h:35,83-85): target rows of 8-row strip i
    see query columns [8*lo, 8*hi) with band = k_band >> 3,
    other = TW - (QW - band), lo = max(0, i - other + 1), hi = min(band + i, QW).
    Cells outside the band read as H = E = 0; each row's band starts from a zero
    left boundary (H, F and the diagonal).  Returns (score, query end, target end)
    with the end = the first maximal cell in (strip, column, row) order."""
    m, mm, go, ge = sc
    oe = go + ge
    qlen, tlen = len(q), len(t)
    QW, TW = (qlen + 7) // 8, (tlen + 7) // 8
    qn, tn = nibs(q, 8 * QW), nibs(t, 8 * TW)
    band = k_band >> 3
    other = TW - (QW - band)
    lo = [max(0, i - other + 1) for i in range(TW)]
    hi = [min(band + i, QW) for i in range(TW)]
    H = [[0] * (8 * QW) for _ in range(8 * TW)]
    E = [[0] * (8 * QW) for _ in range(8 * TW)]
    F = [[0] * (8 * QW) for _ in range(8 * TW)]

    def inb(r, c):
        return r >= 0 and 8 * lo[r >> 3] <= c < 8 * hi[r >> 3]

    for r in range(8 * TW):
        s = r >> 3
        for c in range(8 * lo[s], 8 * hi[s]):
            cut = (lambda v: t16(v)) if r % 8 == 0 else (lambda v: v)
            uh = cut(H[r - 1][c]) if inb(r - 1, c) else 0
            ue = cut(E[r - 1][c]) if inb(r - 1, c) else 0
            if c == 8 * lo[s]:
                dg, lh, lf = 0, 0, 0
            else:
                dg = cut(H[r - 1][c - 1]) if inb(r - 1, c - 1) else 0
                lh, lf = H[r][c - 1], F[r][c - 1]
            F[r][c] = max(lh - oe, lf - ge)
            E[r][c] = max(uh - oe, ue - ge)
            H[r][c] = max(dg + sub_nib(qn[c], tn[r], m, mm, nval), F[r][c], E[r][c], 0)
    best, qe, te = 0, 0, 0
    for s in range(TW):
        for c in range(8 * lo[s], 8 * hi[s]):
            for r8 in range(8):
                h = H[8 * s + r8][c]
                if h > best:
                    best, qe, te = h, c, 8 * s + r8
    return best, qe, te


def ksw_unpruned_reference(q, t, h0, sc=(1, 4, 6, 1), nval=0):
    """bwa-style seed extension as a full matrix, for inputs where every cell
    stays positive (large h0), so that the reference's row pruning and its
    zero-row exit never act (ksw_kernel_template.h:77-197):
      H(-1,-1) = h0, H(-1, j) = max(0, h0 - open - extend*(j+1)),
      H(i, -1) = max(0, h0 - open - extend*(i+1)), E(0, j) = 0, F(i, 0) = 0;
      M = H(i-1, j-1) + s if H(i-1, j-1) else 0;  H = max(M, E, F);
      E(i+1, j) = max(E(i, j) - extend, M - open - extend, 0), F likewise along j.
    Row maximum: last column reaching it; best: first row with a larger row
    maximum; query-end score: largest H(i, Lq-1), last row reaching it.
    Result (clip penalty 5): (best, best column+1, best row+1) unless the
    query-end score is positive and above best - 5: (it, Lq, its row + 1)."""
    m, mm, go, ge = sc
    oe = go + ge
    qlen, tlen = len(q), len(t)
    qn, tn = nibs(q), nibs(t)
    prev = [max(0, h0 - go - ge * (j + 1)) for j in range(qlen)]   # H(i-1, j)
    prev_left = h0                                                    # H(i-1, -1)
    E = [0] * qlen
    best, bi, bj = h0, -1, -1
    gscore, gi = -1, -1
    for i in range(tlen):
        left = max(0, h0 - go - ge * (i + 1))
        f = 0
        rowm, rowj = 0, -1
        cur = [0] * qlen
        for j in range(qlen):
            d = prev_left if j == 0 else prev[j - 1]
            M = d + sub_nib(qn[j], tn[i], m, mm, nval) if d else 0
            h = max(M, E[j], f)
            cur[j] = h
            if h >= rowm:
                rowm, rowj = h, j
            E[j] = max(E[j] - ge, M - oe, 0)
            f = max(f - ge, M - oe, 0)
        if cur[qlen - 1] >= gscore:
            gscore, gi = cur[qlen - 1], i
        if rowm > best:
            best, bi, bj = rowm, i, rowj
        prev, prev_left = cur, left
    if gscore <= 0 or gscore <= best - 5:
        return best, bj + 1, bi + 1
    return gscore, qlen, gi + 1


def local_traceback_reference(q, t, sc=(1, 4, 6, 1), nval=0):
    """LOCAL WITH_TB from the alignment automaton: M (aligned pair), D (gap in
    the query, consumes target), I (gap in the target, consumes query), gaps
    opened from M, H = max(0, M, D, I).  Boundaries: everything left of column 0
    and above row 0 is 0, including the incoming gap values D(0, c) and I(r, 0).
    End: the first maximal cell in (strip, column, row) order.
    Traceback (the reference's reading of these matrices, get_tb.h:47-117):
      at an H cell the source is M (labelled 0 when the substitution is >= 0,
      else 1), else I (3), else D (2);  inside a gap the run continues while the
      gap value came from extending (ties count as extending), and the cell it
      was opened from is labelled 0;  ops 0, 1, 2 step the target back, 0, 1, 3
      the query;  a running label score (match, -mismatch, -(open+extend) on a
      run's first gap, -extend after) stops the walk on the cell where it
      equals the score.  Ops are run-length coded (count <= 63) in walk order;
      the last byte written keeps the previous cell's op.
    Returns dict(score, qend, tend, qstart, tstart, cigar=bytes)."""
    m, mm, go, ge = sc
    oe = go + ge
    qlen, tlen = len(q), len(t)
    QW, TW = (qlen + 7) // 8, (tlen + 7) // 8
    qn, tn = nibs(q, 8 * QW), nibs(t, 8 * TW)
    R, Cn = 8 * TW, 8 * QW
    S = [[sub_nib(qn[c], tn[r], m, mm, nval) for c in range(Cn)] for r in range(R)]
    M = [[0] * Cn for _ in range(R)]
    D = [[0] * Cn for _ in range(R)]
    I = [[0] * Cn for _ in range(R)]
    H = [[0] * Cn for _ in range(R)]
    for r in range(R):
        cut = t16 if r % 8 == 0 else (lambda v: v)
        for c in range(Cn):
            dg = cut(H[r - 1][c - 1]) if r > 0 and c > 0 else 0
            M[r][c] = dg + S[r][c]
            D[r][c] = max(M[r - 1][c] - oe, cut(D[r - 1][c]) - ge) if r > 0 else 0
            I[r][c] = max(M[r][c - 1] - oe, I[r][c - 1] - ge) if c > 0 else 0
            H[r][c] = max(0, M[r][c], D[r][c], I[r][c])
    best, te, qe = 0, 0, 0
    for s in range(TW):
        for c in range(Cn):
            for r8 in range(8):
                h = H[8 * s + r8][c]
                if h > best:
                    best, te, qe = h, 8 * s + r8, c
    out = []
    i, j = te, qe
    state, prev, count, acc = "H", 0, 0, 0
    while i >= 0 and j >= 0:
        if state == "H":
            h = H[i][j]
            if h == M[i][j]:
                op = 0 if S[i][j] >= 0 else 1
            elif h == I[i][j]:
                op = 3
            else:
                op = 2
        elif state == "I":
            op = 3 if not (M[i][j] - oe > I[i][j] - ge) else 0
        else:
            op = 2 if not (M[i][j] - oe > D[i][j] - ge) else 0
        if count < 63 and op == prev:
            count += 1
        else:
            if count:
                out.append(prev | count << 2)
            count = 1
        if op in (2, 3):
            acc -= oe if prev != op else ge
        else:
            acc += -mm if op == 1 else m
        if acc == best:
            break
        prev = op
        state = "I" if op == 3 else ("D" if op == 2 else "H")
        if op in (0, 1, 2):
            i -= 1
        if op in (0, 1, 3):
            j -= 1
    out.append(prev | count << 2)
    return dict(score=best, qend=qe, tend=te, qstart=j, tstart=i, cigar=bytes(out))
