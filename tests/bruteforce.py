"""Independent brute-force checkers for tiny inputs.

These enumerate every alignment path explicitly (no dynamic programming), so
they pin the oracle's optimal scores and end cells without restating its
recurrences.  A path is a start cell plus a move string over
D (diagonal), V (vertical: consume a row residue, gap in H) and
H (horizontal: consume a column residue, gap in V).  Rows index the row
sequence v, columns index the column sequence h.
"""


def run_cost(length, go, ge):
    """Cost of one maximal gap run of `length` moves in the Gotoh DP.

    The DP may re-open a gap right after a gap of the same direction, so a
    run of L moves costs the best split into pieces, each piece paying
    go + (len-1)*ge; that optimum is max(L*go, go + (L-1)*ge).
    """
    return max(length * go, go + (length - 1) * ge)


def path_score(h, v, i0, j0, moves, mat, go, ge):
    s, i, j = 0, i0, j0
    run, prev = 0, None
    for mv in moves:
        if mv != prev and run:
            s += run_cost(run, go, ge)
            run = 0
        if mv == "D":
            s += int(mat[int(v[i]) * 32 + int(h[j])])
            i += 1
            j += 1
        else:
            run += 1
            if mv == "V":
                i += 1
            else:
                j += 1
        prev = mv
    if run:
        s += run_cost(run, go, ge)
    return s, i, j


def path_stats(h, v, i0, j0, moves):
    matches, runs, prev = 0, 0, None
    i, j = i0, j0
    for mv in moves:
        if mv == "D":
            matches += int(v[i] == h[j])
            i += 1
            j += 1
        else:
            if mv != prev:
                runs += 1
            if mv == "V":
                i += 1
            else:
                j += 1
        prev = mv
    return {"matches": matches, "cols": len(moves), "gap_opens": runs}


def _paths(i, j, m, n, prefix, out):
    out.append((i, j, prefix))
    if i < m and j < n:
        _paths(i + 1, j + 1, m, n, prefix + "D", out)
    if i < m:
        _paths(i + 1, j, m, n, prefix + "V", out)
    if j < n:
        _paths(i, j + 1, m, n, prefix + "H", out)


def all_paths_from(i0, j0, m, n):
    out = []
    _paths(i0, j0, m, n, "", out)
    return out  # (end_i, end_j, moves)


def local_best_by_end(h, v, mat, go, ge):
    """best[(i, j)] = max score of a non-empty path ending at cell (i, j)."""
    m, n = len(v), len(h)
    best = {}
    for i0 in range(m + 1):
        for j0 in range(n + 1):
            for (i1, j1, mv) in all_paths_from(i0, j0, m, n):
                if not mv:
                    continue
                s, _, _ = path_score(h, v, i0, j0, mv, mat, go, ge)
                if s > best.get((i1, j1), -10**9):
                    best[(i1, j1)] = s
    return best


def local_first_max(h, v, mat, go, ge):
    """(score, v_end, h_end): first maximum in column-major order, > 0 only."""
    best = local_best_by_end(h, v, mat, go, ge)
    top, ci, cj = 0, 0, 0
    for j in range(len(h) + 1):
        for i in range(len(v) + 1):
            s = best.get((i, j), -10**9)
            if s > top:
                top, ci, cj = s, i, j
    return top, ci, cj


def global_best(h, v, mat, go, ge):
    m, n = len(v), len(h)
    top = None
    for (i1, j1, mv) in all_paths_from(0, 0, m, n):
        if i1 == m and j1 == n:
            s, _, _ = path_score(h, v, 0, 0, mv, mat, go, ge)
            top = s if top is None else max(top, s)
    return top


def global_optimal_paths(h, v, mat, go, ge):
    m, n = len(v), len(h)
    top = global_best(h, v, mat, go, ge)
    res = []
    for (i1, j1, mv) in all_paths_from(0, 0, m, n):
        if i1 == m and j1 == n:
            s, _, _ = path_score(h, v, 0, 0, mv, mat, go, ge)
            if s == top:
                res.append(mv)
    return top, res


def local_optimal_paths(h, v, mat, go, ge):
    """All (score, i0, j0, i1, j1, moves) of optimal non-empty local paths."""
    m, n = len(v), len(h)
    top, res = 0, []
    for i0 in range(m + 1):
        for j0 in range(n + 1):
            for (i1, j1, mv) in all_paths_from(i0, j0, m, n):
                if not mv:
                    continue
                s, _, _ = path_score(h, v, i0, j0, mv, mat, go, ge)
                if s > top:
                    top, res = s, []
                if s == top and s > 0:
                    res.append((i0, j0, i1, j1, mv))
    return top, res


# ---------------------------------------------------------------------------
# The traceback policy, restated over explicit paths.
#
# An annotated path spells every gap move as opening a DP gap run ("o") or
# extending one ("e"); an extension must follow a move of the same
# direction.  Scores: D adds the substitution score, an opening move adds go,
# an extending move adds ge.  The best annotated score of a move string is
# path_score's.  The policy the oracle and the kernels document --
# restart > diagonal > vertical > horizontal in the H state, and a tie
# between extending and opening a gap run keeps the extension -- selects,
# among the optimal annotated paths ending at the reported end cell, the one
# that is lexicographically greatest when read backward from the end under
# the order  stop > D > Ve > Vo > He > Ho  ("stop": the path begins here).
# This picks the path by comparing whole paths; it does not walk a DP.
# ---------------------------------------------------------------------------

RANK = {"stop": 6, "D": 5, "Ve": 4, "Vo": 3, "He": 2, "Ho": 1}


def _annotated(i, j, m, n, prev, prefix, out):
    out.append((i, j, prefix))
    if i < m and j < n:
        _annotated(i + 1, j + 1, m, n, "D", prefix + ("D",), out)
    if i < m:
        _annotated(i + 1, j, m, n, "V", prefix + ("Vo",), out)
        if prev == "V":
            _annotated(i + 1, j, m, n, "V", prefix + ("Ve",), out)
    if j < n:
        _annotated(i, j + 1, m, n, "H", prefix + ("Ho",), out)
        if prev == "H":
            _annotated(i, j + 1, m, n, "H", prefix + ("He",), out)


def annotated_paths_from(i0, j0, m, n):
    out = []
    _annotated(i0, j0, m, n, None, (), out)
    return out


def annotated_score(h, v, i0, j0, moves, mat, go, ge):
    s, i, j = 0, i0, j0
    for mv in moves:
        if mv == "D":
            s += int(mat[int(v[i]) * 32 + int(h[j])])
            i += 1
            j += 1
        else:
            s += go if mv[1] == "o" else ge
            if mv[0] == "V":
                i += 1
            else:
                j += 1
    return s


def _backward_key(moves):
    # read from the end; the path's beginning reads as "stop"
    return tuple(RANK[m] for m in reversed(moves)) + (RANK["stop"],)


def policy_path_local(h, v, mat, go, ge, end):
    """(i0, j0, moves) of the policy's optimal local path ending at `end`
    (the first-maximum end cell); moves are plain D/V/H."""
    m, n = len(v), len(h)
    cands, top = [], None
    for i0 in range(m + 1):
        for j0 in range(n + 1):
            for (i1, j1, mv) in annotated_paths_from(i0, j0, m, n):
                if not mv or (i1, j1) != end:
                    continue
                s = annotated_score(h, v, i0, j0, mv, mat, go, ge)
                if top is None or s > top:
                    top, cands = s, []
                if s == top:
                    cands.append((i0, j0, mv))
    best = max(cands, key=lambda c: _backward_key(c[2]))
    return best[0], best[1], "".join(x[0] for x in best[2]), top


def policy_path_global(h, v, mat, go, ge):
    m, n = len(v), len(h)
    cands, top = [], None
    for (i1, j1, mv) in annotated_paths_from(0, 0, m, n):
        if (i1, j1) != (m, n):
            continue
        s = annotated_score(h, v, 0, 0, mv, mat, go, ge)
        if top is None or s > top:
            top, cands = s, []
        if s == top:
            cands.append(mv)
    best = max(cands, key=_backward_key)
    return "".join(x[0] for x in best), top
