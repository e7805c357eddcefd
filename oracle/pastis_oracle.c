/*
 * pastis_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see the header).
 *
 * Each function restates one SeqAn routine that the reference calls; the
 * call site is cited on the function.  SeqAn is absent from the reference
 * checkout, so its behaviour is written here as explicit policies:
 *
 *   P1 local end cell: first maximum in SeqAn's visit order (columns over H,
 *      rows over V), i.e. ties go to the smaller H index, then the smaller
 *      V index; a cell only counts when its score is > 0.
 *   P2 traceback priority in the H state: diagonal > vertical (gap in H)
 *      > horizontal (gap in V); in a gap state a tie between extending and
 *      opening keeps the extension.  In local mode a cell whose raw value is
 *      <= 0 is a restart (the traceback stops there).
 *   P3 statistics as computeAlignmentStats: cols = alignment columns,
 *      matches = identical residues, gap_opens = maximal runs of gaps in
 *      either row; identity = (float)(100.0 * matches / cols).
 *   P4 X-drop: SeqAn's anti-diagonal gapped X-drop with its live-band
 *      update and "longest extension" endpoint rule (see or_xdrop_one).
 */
#include "pastis_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEG_INF (-(1 << 29))

/* inc/ScoreMat.hpp:104-131, order ARNDCQEGHILKMFPSTWYVBZX*J. */
static const int8_t BLOSUM62_25[25 * 25] = {
     4,-1,-2,-2, 0,-1,-1, 0,-2,-1,-1,-1,-1,-2,-1, 1, 0,-3,-2, 0,-2,-1, 0,-4,-1,
    -1, 5, 0,-2,-3, 1, 0,-2, 0,-3,-2, 2,-1,-3,-2,-1,-1,-3,-2,-3,-1, 0,-1,-4,-2,
    -2, 0, 6, 1,-3, 0, 0, 0, 1,-3,-3, 0,-2,-3,-2, 1, 0,-4,-2,-3, 3, 0,-1,-4,-3,
    -2,-2, 1, 6,-3, 0, 2,-1,-1,-3,-4,-1,-3,-3,-1, 0,-1,-4,-3,-3, 4, 1,-1,-4,-4,
     0,-3,-3,-3, 9,-3,-4,-3,-3,-1,-1,-3,-1,-2,-3,-1,-1,-2,-2,-1,-3,-3,-2,-4,-1,
    -1, 1, 0, 0,-3, 5, 2,-2, 0,-3,-2, 1, 0,-3,-1, 0,-1,-2,-1,-2, 0, 3,-1,-4,-2,
    -1, 0, 0, 2,-4, 2, 5,-2, 0,-3,-3, 1,-2,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4,-3,
     0,-2, 0,-1,-3,-2,-2, 6,-2,-4,-4,-2,-3,-3,-2, 0,-2,-2,-3,-3,-1,-2,-1,-4,-4,
    -2, 0, 1,-1,-3, 0, 0,-2, 8,-3,-3,-1,-2,-1,-2,-1,-2,-2, 2,-3, 0, 0,-1,-4,-3,
    -1,-3,-3,-3,-1,-3,-3,-4,-3, 4, 2,-3, 1, 0,-3,-2,-1,-3,-1, 3,-3,-3,-1,-4, 2,
    -1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-3,-1,-4, 4,
    -1, 2, 0,-1,-3, 1, 1,-2,-1,-3,-2, 5,-1,-3,-1, 0,-1,-3,-2,-2, 0, 1,-1,-4,-2,
    -1,-1,-2,-3,-1, 0,-2,-3,-2, 1, 2,-1, 5, 0,-2,-1,-1,-1,-1, 1,-3,-1,-1,-4, 2,
    -2,-3,-3,-3,-2,-3,-3,-3,-1, 0, 0,-3, 0, 6,-4,-2,-2, 1, 3,-1,-3,-3,-1,-4, 0,
    -1,-2,-2,-1,-3,-1,-1,-2,-2,-3,-3,-1,-2,-4, 7,-1,-1,-4,-3,-2,-2,-1,-2,-4,-3,
     1,-1, 1, 0,-1, 0, 0, 0,-1,-2,-2, 0,-1,-2,-1, 4, 1,-3,-2,-2, 0, 0, 0,-4,-2,
     0,-1, 0,-1,-1,-1,-1,-2,-2,-1,-1,-1,-1,-2,-1, 1, 5,-2,-2, 0,-1,-1, 0,-4,-1,
    -3,-3,-4,-4,-2,-2,-3,-2,-2,-3,-2,-3,-1, 1,-4,-3,-2,11, 2,-3,-4,-3,-2,-4,-2,
    -2,-2,-2,-3,-2,-1,-2,-3, 2,-1,-1,-2,-1, 3,-3,-2,-2, 2, 7,-1,-3,-2,-1,-4,-1,
     0,-3,-3,-3,-1,-2,-2,-3,-3, 3, 1,-2, 1,-1,-2,-2, 0,-3,-1, 4,-3,-2,-1,-4, 1,
    -2,-1, 3, 4,-3, 0, 1,-1, 0,-3,-4, 0,-3,-3,-2, 0,-1,-4,-3,-3, 4, 1,-1,-4,-4,
    -1, 0, 0, 1,-3, 3, 4,-2, 0,-3,-3, 1,-1,-3,-1, 0,-1,-3,-2,-2, 1, 4,-1,-4,-3,
     0,-1,-1,-1,-2,-1,-1,-1,-1,-1,-1,-1,-1,-1,-2, 0, 0,-2,-1,-1,-1,-1,-1,-4,-1,
    -4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4,-4, 1,-4,
    -1,-2,-3,-4,-1,-2,-3,-4,-3, 2, 4,-2, 2, 0,-3,-2,-1,-2,-1, 1,-4,-3,-1,-4, 4,
};

void or_blosum62_32(int8_t out[32 * 32])
{
    memset(out, 0, 32 * 32);
    for (int a = 0; a < 25; ++a)
        for (int b = 0; b < 25; ++b)
            out[a * 32 + b] = BLOSUM62_25[a * 25 + b];
}

/* trace byte: bits 0-1 = H-state source, bit 2 = E extended, bit 3 = F extended */
enum { SRC_NONE = 0, SRC_DIAG = 1, SRC_VERT = 2, SRC_HORZ = 3 };

static void traceback(const uint8_t *tb, int ld, const uint8_t *h,
                      const uint8_t *v, int i, int j, or_aln_t *out)
{
    /* Walk back from H state at (i, j); rows i index V, columns j index H. */
    int state = 0; /* 0 = H, 1 = E (horizontal), 2 = F (vertical) */
    int last = 0;  /* last emitted move: 0 none, 1 diag, 2 vert, 3 horz */
    int matches = 0, cols = 0, opens = 0;
    for (;;) {
        uint8_t t = tb[(size_t)j * ld + i];
        if (state == 0) {
            int src = t & 3;
            if (src == SRC_NONE)
                break;
            if (src == SRC_DIAG) {
                cols++;
                matches += (v[i - 1] == h[j - 1]);
                last = 1;
                --i;
                --j;
            } else if (src == SRC_VERT) {
                state = 2;
            } else {
                state = 1;
            }
        } else if (state == 1) {
            cols++;
            if (last != 3)
                opens++;
            last = 3;
            state = (t >> 2) & 1 ? 1 : 0;
            --j;
        } else {
            cols++;
            if (last != 2)
                opens++;
            last = 2;
            state = (t >> 3) & 1 ? 2 : 0;
            --i;
        }
    }
    out->v_beg = i;
    out->h_beg = j;
    out->matches = matches;
    out->cols = cols;
    out->gap_opens = opens;
}

/* seqan::localAlignment as called at SeqanFullAligner.cpp:231 (Gotoh local,
 * policies P1-P3). */
int or_sw_local(const uint8_t *h, int nh, const uint8_t *v, int nv,
                const int8_t *mat, int go, int ge, int with_stats,
                or_aln_t *out)
{
    memset(out, 0, sizeof(*out));
    int32_t *Hc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
    int32_t *E = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
    uint8_t *tb = NULL;
    int ld = nv + 1;
    if (!Hc || !E) {
        free(Hc);
        free(E);
        return -1;
    }
    if (with_stats) {
        tb = (uint8_t *)calloc((size_t)(nh + 1) * (size_t)ld, 1);
        if (!tb) {
            free(Hc);
            free(E);
            return -1;
        }
    }
    for (int i = 0; i <= nv; ++i) {
        Hc[i] = 0;
        E[i] = NEG_INF;
    }
    int32_t best = 0;
    int bi = 0, bj = 0;
    for (int j = 1; j <= nh; ++j) {
        int32_t hdiag = Hc[0];
        int32_t F = NEG_INF, hup = 0;
        const int8_t *mrow = mat + (size_t)h[j - 1];
        uint8_t *tcol = tb ? tb + (size_t)j * ld : NULL;
        for (int i = 1; i <= nv; ++i) {
            int eb, fb, src;
            int32_t ee = E[i] + ge, eo = Hc[i] + go;
            if (eo > ee) { E[i] = eo; eb = 0; } else { E[i] = ee; eb = 1; }
            int32_t fe = F + ge, fo = hup + go;
            if (fo > fe) { F = fo; fb = 0; } else { F = fe; fb = 1; }
            int32_t hv = hdiag + mrow[(size_t)v[i - 1] * 32];
            src = SRC_DIAG;
            if (F > hv) { hv = F; src = SRC_VERT; }
            if (E[i] > hv) { hv = E[i]; src = SRC_HORZ; }
            if (hv <= 0) { hv = 0; src = SRC_NONE; }
            hdiag = Hc[i];
            Hc[i] = hv;
            hup = hv;
            if (hv > best) { best = hv; bi = i; bj = j; }
            if (tcol)
                tcol[i] = (uint8_t)(src | (eb << 2) | (fb << 3));
        }
    }
    out->score = best;
    out->v_end = bi;
    out->h_end = bj;
    out->v_beg = bi;
    out->h_beg = bj;
    if (tb && best > 0)
        traceback(tb, ld, h, v, bi, bj, out);
    free(tb);
    free(Hc);
    free(E);
    return 0;
}

/* seqan::globalAlignment as called at SeqanXdropAligner.cpp:296 (Gotoh
 * global, no free end gaps; boundary rows follow the same recurrence). */
int or_nw_global(const uint8_t *h, int nh, const uint8_t *v, int nv,
                 const int8_t *mat, int go, int ge, int with_stats,
                 or_aln_t *out)
{
    memset(out, 0, sizeof(*out));
    int32_t *Hc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
    int32_t *E = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nv + 1));
    uint8_t *tb = NULL;
    int ld = nv + 1;
    if (!Hc || !E) {
        free(Hc);
        free(E);
        return -1;
    }
    if (with_stats) {
        tb = (uint8_t *)calloc((size_t)(nh + 1) * (size_t)ld, 1);
        if (!tb) {
            free(Hc);
            free(E);
            return -1;
        }
    }
    /* column 0: vertical gap run down the first column */
    Hc[0] = 0;
    E[0] = NEG_INF;
    {
        int32_t F = NEG_INF, hup = 0;
        for (int i = 1; i <= nv; ++i) {
            int32_t fe = F + ge, fo = hup + go;
            int fb;
            if (fo > fe) { F = fo; fb = 0; } else { F = fe; fb = 1; }
            Hc[i] = F;
            hup = F;
            E[i] = NEG_INF;
            if (tb)
                tb[i] = (uint8_t)(SRC_VERT | (fb << 3));
        }
    }
    int32_t e0 = NEG_INF; /* E(0, j) */
    for (int j = 1; j <= nh; ++j) {
        /* row 0: horizontal gap run along the first row */
        int eb0;
        int32_t ee0 = e0 + ge, eo0 = Hc[0] + go;
        if (eo0 > ee0) { e0 = eo0; eb0 = 0; } else { e0 = ee0; eb0 = 1; }
        int32_t hdiag = Hc[0];
        Hc[0] = e0;
        uint8_t *tcol = tb ? tb + (size_t)j * ld : NULL;
        if (tcol)
            tcol[0] = (uint8_t)(SRC_HORZ | (eb0 << 2));
        int32_t F = NEG_INF, hup = e0;
        const int8_t *mrow = mat + (size_t)h[j - 1];
        for (int i = 1; i <= nv; ++i) {
            int eb, fb, src;
            int32_t ee = E[i] + ge, eo = Hc[i] + go;
            if (eo > ee) { E[i] = eo; eb = 0; } else { E[i] = ee; eb = 1; }
            int32_t fe = F + ge, fo = hup + go;
            if (fo > fe) { F = fo; fb = 0; } else { F = fe; fb = 1; }
            int32_t hv = hdiag + mrow[(size_t)v[i - 1] * 32];
            src = SRC_DIAG;
            if (F > hv) { hv = F; src = SRC_VERT; }
            if (E[i] > hv) { hv = E[i]; src = SRC_HORZ; }
            hdiag = Hc[i];
            Hc[i] = hv;
            hup = hv;
            if (tcol)
                tcol[i] = (uint8_t)(src | (eb << 2) | (fb << 3));
        }
    }
    out->score = Hc[nv];
    out->v_beg = 0;
    out->h_beg = 0;
    out->v_end = nv;
    out->h_end = nh;
    if (tb) {
        /* global traceback ends at (0, 0); the trace there is SRC_NONE */
        tb[0] = SRC_NONE;
        traceback(tb, ld, h, v, nv, nh, out);
    }
    free(tb);
    free(Hc);
    free(E);
    return 0;
}

/* ------------------------------------------------------------------------
 * Gapped X-drop, one direction: SeqAn's _extendSeedGappedXDropOneDirection,
 * reached through extendSeed(seed, H, V, EXTEND_BOTH, score, X, GappedXDrop)
 * at SeqanXdropAligner.cpp:281-283.  The query (V) runs along the columns
 * of the anti-diagonal DP, the database (H) along the rows.  Values below
 * best - X become "undefined"; the live column range drops undefined edge
 * columns of the last two anti-diagonals, grows by one per step and is
 * clipped to the segment ends.  The endpoint is the "longest extension"
 * cell, not the best cell.
 * ---------------------------------------------------------------------- */
#define XD_UNDEF (INT_MIN / 4)

static int32_t or_xdrop_one(const uint8_t *qs, long qlen, const uint8_t *ds,
                            long dlen, int right, const int8_t *mat, int gap,
                            int X, long *out_col, long *out_row, int *updated,
                            uint64_t *cells)
{
    const long cols = qlen + 1, rows = dlen + 1;
    *updated = 0;
    *out_col = *out_row = 0;
    if (rows == 1 || cols == 1)
        return 0;
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(cols + 2));
    if (!buf)
        return XD_UNDEF;
    int32_t *a1 = buf, *a2 = buf + (cols + 2), *a3 = buf + 2 * (cols + 2);
    long l1 = 0, l2 = 1, l3 = 2;
    a2[0] = 0;
    if (-gap > X) {
        a3[0] = XD_UNDEF;
        a3[1] = XD_UNDEF;
    } else {
        a3[0] = gap;
        a3[1] = gap;
    }
    long minCol = 1, maxCol = 2;
    long off1 = 0, off2 = 0, off3 = 0;
    long d = 1;
    int32_t best = 0;
    while (minCol < maxCol) {
        ++d;
        int32_t *t = a1;
        a1 = a2;
        a2 = a3;
        a3 = t;
        l1 = l2;
        l2 = l3;
        off1 = off2;
        off2 = off3;
        off3 = minCol - 1;
        /* _initAntiDiag3 */
        l3 = maxCol + 1 - off3;
        a3[0] = XD_UNDEF;
        a3[maxCol - off3] = XD_UNDEF;
        if ((long)d * gap > (long)best - X) {
            if (off3 == 0)
                a3[0] = (int32_t)(d * gap);
            if (d - maxCol == 0)
                a3[maxCol - off3] = (int32_t)(d * gap);
        }
        int32_t adbest = (int32_t)(d * gap);
        for (long col = minCol; col < maxCol; ++col) {
            long i3 = col - off3, i2 = col - off2, i1 = col - off1;
            long qp, dp;
            if (right) {
                qp = col - 1;
                dp = d - col - 1;
            } else {
                qp = cols - 1 - col;
                dp = rows - 1 + col - d;
            }
            int32_t tmp = (a2[i2 - 1] > a2[i2] ? a2[i2 - 1] : a2[i2]) + gap;
            int32_t dg = a1[i1 - 1] + mat[(size_t)qs[qp] * 32 + ds[dp]];
            if (dg > tmp)
                tmp = dg;
            if (tmp < best - X) {
                a3[i3] = XD_UNDEF;
            } else {
                a3[i3] = tmp;
                if (tmp > adbest)
                    adbest = tmp;
            }
            ++*cells;
        }
        if (adbest > best)
            best = adbest;
        while (minCol - off3 < l3 && a3[minCol - off3] == XD_UNDEF &&
               minCol - off2 - 1 < l2 && a2[minCol - off2 - 1] == XD_UNDEF)
            ++minCol;
        while (maxCol - off3 > 0 && a3[maxCol - off3 - 1] == XD_UNDEF &&
               a2[maxCol - off2 - 1] == XD_UNDEF)
            --maxCol;
        ++maxCol;
        if (minCol < d + 2 - rows)
            minCol = d + 2 - rows;
        if (maxCol > cols)
            maxCol = cols;
    }
    /* longest extension: both ends, end of query, end of database, else
     * the best cell of the last live anti-diagonal */
    long lcol = l3 + off3 - 2;
    long lrow = d - lcol;
    int32_t ls = a3[lcol - off3];
    if (ls == XD_UNDEF) {
        if (a2[l2 - 2] != XD_UNDEF) {
            lcol = l2 + off2 - 2;
            lrow = d - 1 - lcol;
            ls = a2[lcol - off2];
        } else if (l2 > 2 && a2[l2 - 3] != XD_UNDEF) {
            lcol = l2 + off2 - 3;
            lrow = d - 1 - lcol;
            ls = a2[lcol - off2];
        }
    }
    if (ls == XD_UNDEF) {
        for (long i = 0; i < l1; ++i) {
            if (a1[i] > ls) {
                ls = a1[i];
                lcol = i + off1;
                lrow = d - 2 - lcol;
            }
        }
    }
    if (ls != XD_UNDEF) {
        *updated = 1;
        *out_col = lcol;
        *out_row = lrow;
    }
    free(buf);
    return ls;
}

/* extendSeed(..., EXTEND_BOTH, ..., GappedXDrop) on a simple seed
 * TSeed(beginH = col offset, beginV = row offset, k), SeqanXdropAligner.cpp:279-283. */
int or_xdrop_extend(const uint8_t *h, int nh, const uint8_t *v, int nv,
                    int beg_h, int beg_v, int k, const int8_t *mat, int gap,
                    int X, or_xdrop_t *out)
{
    memset(out, 0, sizeof(*out));
    if (beg_h < 0 || beg_v < 0 || k < 0 || beg_h + k > nh || beg_v + k > nv)
        return -2;
    int end_h = beg_h + k, end_v = beg_v + k;
    long c, r;
    int upd;
    /* left: query = V[0, beg_v), database = H[0, beg_h), both reversed */
    out->score_left = or_xdrop_one(v, beg_v, h, beg_h, 0, mat, gap, X, &c, &r,
                                   &upd, &out->cells);
    out->upd_left = upd;
    if (upd) {
        beg_h -= (int)r;
        beg_v -= (int)c;
    }
    /* right: query = V[end_v, nv), database = H[end_h, nh) */
    out->score_right = or_xdrop_one(v + end_v, nv - end_v, h + end_h,
                                    nh - end_h, 1, mat, gap, X, &c, &r, &upd,
                                    &out->cells);
    out->upd_right = upd;
    if (upd) {
        end_h += (int)r;
        end_v += (int)c;
    }
    out->beg_h = beg_h;
    out->beg_v = beg_v;
    out->end_h = end_h;
    out->end_v = end_v;
    return 0;
}

static float or_identity(int matches, int cols)
{
    /* AlignmentStats::alignmentIdentity is a float computed in double */
    return (float)(100.0 * (double)(float)matches / (double)(float)cols);
}

/* SeqanFullAligner.cpp:247-261 and SeqanXdropAligner.cpp:338-351. */
int or_predicate(const or_aln_t *a, int len_h, int len_v, double cov_thr,
                 int ani_thr, float *identity, float *score_aln)
{
    float ident = a->cols > 0 ? or_identity(a->matches, a->cols) : NAN;
    double amg = (double)(a->cols - a->gap_opens);
    double c1 = amg / len_h, c2 = amg / len_v;
    double cov = c1 > c2 ? c1 : c2;
    int pass = (cov >= cov_thr) && (ident >= (float)ani_thr);
    *identity = ident;
    *score_aln = pass ? ident / 100.0f : 0.0f;
    return pass;
}

int or_sfa_pair(const uint8_t *h, int nh, const uint8_t *v, int nv,
                const int8_t *mat, const or_params_t *p, or_pair_out_t *out)
{
    memset(out, 0, sizeof(*out));
    if (or_sw_local(h, nh, v, nv, mat, p->gap_open, p->gap_ext, 1, &out->aln))
        return -1;
    out->pass = or_predicate(&out->aln, nh, nv, p->cov_thr, p->ani_thr,
                             &out->identity, &out->score_aln);
    return 0;
}

/* SeqanXdropAligner.cpp:260-351: seeds, extension, global alignment of the
 * extended infixes, best identity (strict '>', so seed 0 keeps ties),
 * predicate on the full sequence lengths.  seeds = {first.first (row
 * offset), first.second (col offset), second.first, second.second}. */
int or_sxa_pair(const uint8_t *h, int nh, const uint8_t *v, int nv,
                const uint16_t seeds[4], const int8_t *mat,
                const or_params_t *p, or_pair_out_t *out)
{
    memset(out, 0, sizeof(*out));
    float best_ident = 0.0f;
    for (int cnt = 0; cnt < p->seed_cnt; ++cnt) {
        int row_off = cnt == 0 ? seeds[0] : seeds[2];
        int col_off = cnt == 0 ? seeds[1] : seeds[3];
        or_xdrop_t *x = &out->ext[cnt];
        int rc = or_xdrop_extend(h, nh, v, nv, col_off, row_off, p->seed_len,
                                 mat, p->xdrop_gap, p->xdrop, x);
        if (rc)
            return rc;
        or_aln_t a;
        if (or_nw_global(h + x->beg_h, x->end_h - x->beg_h, v + x->beg_v,
                         x->end_v - x->beg_v, mat, p->gap_open, p->gap_ext, 1,
                         &a))
            return -1;
        a.h_beg += x->beg_h;
        a.h_end += x->beg_h;
        a.v_beg += x->beg_v;
        a.v_end += x->beg_v;
        float ident = a.cols > 0 ? or_identity(a.matches, a.cols) : NAN;
        if (cnt == 0 || ident > best_ident) {
            out->aln = a;
            out->seed = cnt;
            best_ident = ident;
        }
    }
    out->pass = or_predicate(&out->aln, nh, nv, p->cov_thr, p->ani_thr,
                             &out->identity, &out->score_aln);
    return 0;
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

int or_sw_batch(const uint8_t *seqs, const uint64_t *offs, const uint32_t *pr,
                const uint32_t *pc, uint64_t n, const int8_t *mat, int go,
                int ge, int32_t *score, int32_t *v_end, int32_t *h_end,
                int nthreads)
{
    int err = 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(| : err)
    for (long long t = 0; t < (long long)n; ++t) {
        const uint8_t *v = seqs + offs[pr[t]];
        int nv = (int)(offs[pr[t] + 1] - offs[pr[t]]);
        const uint8_t *h = seqs + offs[pc[t]];
        int nh = (int)(offs[pc[t] + 1] - offs[pc[t]]);
        or_aln_t a;
        err |= or_sw_local(h, nh, v, nv, mat, go, ge, 0, &a) != 0;
        score[t] = a.score;
        v_end[t] = a.v_end;
        h_end[t] = a.h_end;
    }
    return err ? -1 : 0;
}

int or_sfa_batch(const uint8_t *seqs, const uint64_t *offs, const uint32_t *pr,
                 const uint32_t *pc, uint64_t n, const int8_t *mat,
                 const or_params_t *p, or_pair_out_t *out, int nthreads)
{
    int err = 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
    for (long long t = 0; t < (long long)n; ++t) {
        const uint8_t *v = seqs + offs[pr[t]];
        int nv = (int)(offs[pr[t] + 1] - offs[pr[t]]);
        const uint8_t *h = seqs + offs[pc[t]];
        int nh = (int)(offs[pc[t] + 1] - offs[pc[t]]);
        err |= or_sfa_pair(h, nh, v, nv, mat, p, &out[t]) != 0;
    }
    return err ? -1 : 0;
}

int or_sxa_batch(const uint8_t *seqs, const uint64_t *offs, const uint32_t *pr,
                 const uint32_t *pc, const uint16_t *seeds, uint64_t n,
                 const int8_t *mat, const or_params_t *p, or_pair_out_t *out,
                 int nthreads)
{
    int err = 0;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(| : err)
    for (long long t = 0; t < (long long)n; ++t) {
        const uint8_t *v = seqs + offs[pr[t]];
        int nv = (int)(offs[pr[t] + 1] - offs[pr[t]]);
        const uint8_t *h = seqs + offs[pc[t]];
        int nh = (int)(offs[pc[t] + 1] - offs[pc[t]]);
        err |= or_sxa_pair(h, nh, v, nv, seeds + 4 * t, mat, p, &out[t]) != 0;
    }
    return err ? -1 : 0;
}
