/*
 * kmer_oracle.c -- CPU oracle of the k-mer candidate generation.
 *
 * TEST INFRASTRUCTURE ONLY (see pastis_oracle.h).  Restates, for one block
 * of row sequences x column sequences on one process, what the reference
 * computes with CombBLAS (absent from the checkout):
 *   - the sequence-by-k-mer matrix A of generate_A / add_kmers
 *     (src/kmer/KmerOps.cpp:31-87): every k-mer at offsets 0, s, 2s, ...,
 *     len-k of a sequence, upper-cased with & 0xDF, coded in the alphabet's
 *     base (kcode = kcode*base + code, first residue most significant),
 *     skipped when a residue is not one of the 25 letters; its value is
 *     (cost = BLOSUM62 self score of the alphabet-mapped k-mer, offset);
 *   - C = A * A^T under KmerIntersectLoc (inc/sr.hpp:30-117): multiply gives
 *     count 1, first = (row offset, column offset), score = max of the two
 *     costs; add sums counts (ushort), keeps the accumulated first, takes
 *     the added product's first as second, and keeps the max score.
 *     Products of one (row, column) are added in the order of a
 *     column-by-column SpGEMM: k-mer code ascending, then the column
 *     sequence's occurrences (offset ascending), then the row sequence's.
 *     CombBLAS is absent, so this accumulation order is a stated policy
 *     (parity unpinned beyond it); a pair with one product keeps the
 *     default second = (0, 0) (inc/Types.hpp:170-171).
 *   - the symmetric pruning of the drivers in src/sim_search.cpp: strict
 *     upper triangle (c > r, :403-412 and :569-578 on one process) or the
 *     even/odd rule (:173-187), on global indices r + roff, c + coff.
 * Alphabets: inc/Alphabet.hpp:66-254 with the code assignment of
 * Alphabet::init (src/Alphabet.cpp:31-51).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pastis_oracle.h"

static const char kProtein[] = "ARNDCQEGHILKMFPSTWYVBZX*J";

/* diagonal of BLOSUM62 by letter (inc/ScoreMat.hpp:104-131) */
static int self_score_of(char c)
{
    static const int8_t diag[25] = {4, 5, 6, 6, 9, 5, 5, 6, 8, 4, 4, 5, 5,
                                    6, 7, 4, 5, 11, 7, 4, 4, 4, -1, 1, 4};
    const char *p = strchr(kProtein, c);
    return p && c ? diag[p - kProtein] : 0;
}

/* letter groups (first letter = representative) of each reduced alphabet */
static const char *const kGroups[6][8] = {
    {0},
    {"DENQ", "FWY", "ILMV", "KR", "ST", 0},                 /* murphy10 */
    {"EKQR", "IV", "LY", "AM", "HT", "DNS", "GP", 0},      /* dssp10 */
    {"AEFIKLMQRVW", 0},                                    /* gbmr10 */
    {"EKRQ", "DSN", "HC", "IV", "WYF", "LM", 0},           /* td10 */
    {"KREDQN", "ILV", "STA", 0},                           /* diamond */
};

int or_alph_tables(int alph, uint8_t code_of[128], int16_t self_of[128],
                   int *size)
{
    if (alph < 0 || alph > 5)
        return -1;
    char al_map[128];
    for (int i = 0; i < 128; ++i)
        al_map[i] = (char)i;
    for (int g = 0; kGroups[alph][g]; ++g) {
        const char *grp = kGroups[alph][g];
        for (const char *p = grp; *p; ++p)
            al_map[(int)*p] = grp[0];
    }
    memset(code_of, 255, 128);
    uint8_t next = 0;
    uint8_t rep_code[128];
    memset(rep_code, 255, sizeof(rep_code));
    for (const char *p = kProtein; *p; ++p) {
        const int d = al_map[(int)*p];
        if (rep_code[d] == 255)
            rep_code[d] = next++;
        code_of[(int)*p] = rep_code[d];
    }
    for (int i = 0; i < 128; ++i)
        self_of[i] = code_of[i] == 255 ? 0 : (int16_t)self_score_of(al_map[i]);
    *size = next;
    return 0;
}

typedef struct {
    uint64_t code;
    uint32_t seq;
    uint16_t off;
    int16_t cost;
} kocc_t;

static int cmp_occ(const void *a, const void *b)
{
    const kocc_t *x = (const kocc_t *)a, *y = (const kocc_t *)b;
    if (x->code != y->code)
        return x->code < y->code ? -1 : 1;
    if (x->seq != y->seq)
        return x->seq < y->seq ? -1 : 1;
    return x->off < y->off ? -1 : (x->off > y->off);
}

/* add_kmers over every sequence of one set */
static kocc_t *extract(const char *res, const uint64_t *offs, uint64_t n,
                       int k, int s, const uint8_t *code_of,
                       const int16_t *self_of, int base, uint64_t *nout)
{
    uint64_t cap = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t len = offs[i + 1] - offs[i];
        if (len >= (uint64_t)k)
            cap += (len - (uint64_t)k) / (uint64_t)s + 1;
    }
    kocc_t *o = (kocc_t *)malloc(sizeof(kocc_t) * (cap ? cap : 1));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t len = offs[i + 1] - offs[i];
        const char *q = res + offs[i];
        for (uint64_t p = 0; p + (uint64_t)k <= len; p += (uint64_t)s) {
            uint64_t code = 0;
            int cost = 0, ok = 1;
            for (int j = 0; j < k && ok; ++j) {
                const unsigned ch = (unsigned char)q[p + j] & 0xDFu;
                if (ch >= 128 || code_of[ch] == 255) {
                    ok = 0;
                    break;
                }
                code = code * (uint64_t)base + code_of[ch];
                cost += self_of[ch];
            }
            if (!ok)
                continue;
            o[m].code = code;
            o[m].seq = (uint32_t)i;
            o[m].off = (uint16_t)p;
            o[m].cost = (int16_t)cost;
            ++m;
        }
    }
    *nout = m;
    return o;
}

typedef struct {
    uint32_t c, r;
    uint64_t ord;
    uint16_t off_r, off_c;
    int16_t cost;
} prod_t;

static int cmp_prod(const void *a, const void *b)
{
    const prod_t *x = (const prod_t *)a, *y = (const prod_t *)b;
    if (x->c != y->c)
        return x->c < y->c ? -1 : 1;
    if (x->r != y->r)
        return x->r < y->r ? -1 : 1;
    return x->ord < y->ord ? -1 : (x->ord > y->ord);
}

static int keep_sym(int sym, uint64_t r, uint64_t c)
{
    if (sym == 1)   /* strict upper triangle */
        return c > r;
    if (sym == 2) { /* even/odd rule: PruneI drops when the lambda is true */
        const int ro = (int)(r & 1u), co = (int)(c & 1u);
        const int drop = (r == c) || (r > c ? (ro ^ co) : !(ro ^ co));
        return !drop;
    }
    return 1;
}

int or_kmer_candidates(const char *rres, const uint64_t *roffs, uint64_t nr,
                       const char *cres, const uint64_t *coffs, uint64_t nc,
                       int k, int stride, int alph, int sym, uint64_t roff,
                       uint64_t coff, or_cand_t **out, uint64_t *nout)
{
    uint8_t code_of[128];
    int16_t self_of[128];
    int base = 0;
    *out = NULL;
    *nout = 0;
    if (k < 1 || stride < 1 || or_alph_tables(alph, code_of, self_of, &base))
        return -1;
    if (!cres) {
        cres = rres;
        coffs = roffs;
        nc = nr;
    }
    uint64_t na = 0, nb = 0;
    kocc_t *ra = extract(rres, roffs, nr, k, stride, code_of, self_of, base,
                         &na);
    kocc_t *cb = extract(cres, coffs, nc, k, stride, code_of, self_of, base,
                         &nb);
    qsort(ra, na, sizeof(kocc_t), cmp_occ);
    qsort(cb, nb, sizeof(kocc_t), cmp_occ);
    /* products, code by code: column occurrences outer, row occurrences
     * inner */
    uint64_t np = 0, cap = 1024;
    prod_t *pr = (prod_t *)malloc(sizeof(prod_t) * cap);
    uint64_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (ra[i].code < cb[j].code) {
            ++i;
            continue;
        }
        if (cb[j].code < ra[i].code) {
            ++j;
            continue;
        }
        uint64_t i1 = i, j1 = j;
        while (i1 < na && ra[i1].code == ra[i].code)
            ++i1;
        while (j1 < nb && cb[j1].code == cb[j].code)
            ++j1;
        for (uint64_t b = j; b < j1; ++b)
            for (uint64_t a = i; a < i1; ++a) {
                if (!keep_sym(sym, ra[a].seq + roff, cb[b].seq + coff))
                    continue;
                if (np == cap) {
                    cap *= 2;
                    pr = (prod_t *)realloc(pr, sizeof(prod_t) * cap);
                }
                prod_t *p = &pr[np];
                p->c = cb[b].seq;
                p->r = ra[a].seq;
                p->ord = np;
                p->off_r = ra[a].off;
                p->off_c = cb[b].off;
                p->cost = ra[a].cost > cb[b].cost ? ra[a].cost : cb[b].cost;
                ++np;
            }
        i = i1;
        j = j1;
    }
    free(ra);
    free(cb);
    qsort(pr, np, sizeof(prod_t), cmp_prod);
    or_cand_t *o = (or_cand_t *)malloc(sizeof(or_cand_t) * (np ? np : 1));
    uint64_t m = 0;
    for (uint64_t a = 0; a < np;) {
        /* multiply: the first product */
        or_cand_t acc;
        memset(&acc, 0, sizeof(acc));
        acc.r = pr[a].r;
        acc.c = pr[a].c;
        acc.count = 1;
        acc.score = pr[a].cost;
        acc.seed[0] = pr[a].off_r;
        acc.seed[1] = pr[a].off_c;
        uint64_t b = a + 1;
        for (; b < np && pr[b].c == pr[a].c && pr[b].r == pr[a].r; ++b) {
            /* add(acc, multiply(product b)) */
            acc.count = (uint16_t)(acc.count + 1u);
            acc.seed[2] = pr[b].off_r;
            acc.seed[3] = pr[b].off_c;
            if (pr[b].cost > acc.score)
                acc.score = pr[b].cost;
        }
        o[m++] = acc;
        a = b;
    }
    free(pr);
    *out = o;
    *nout = m;
    return 0;
}

void or_free(void *p) { free(p); }
