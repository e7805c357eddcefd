/*
 * pastis_oracle.h -- CPU oracle for the batched pairwise alignment path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the alignment
 * semantics that the reference reaches through SeqAn (absent from the
 * reference checkout; SeqAn 2.x, branch "develop", unpinned).  It is the
 * checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * against it and never falls back to it.
 *
 * Conventions (mirroring the reference call sites):
 *   H = "horizontal" sequence = seqsc = the column sequence (lc),
 *       SeqanFullAligner.cpp:206-207, SeqanXdropAligner.cpp:224-225.
 *   V = "vertical"   sequence = seqsr = the row sequence (lr),
 *       SeqanFullAligner.cpp:204-205.
 *   The DP is filled column by column over H, rows over V (SeqAn order).
 *   Residues are codes 0..24 in the order "ARNDCQEGHILKMFPSTWYVBZX*J"
 *   (src/Alphabet.cpp:26); the score matrix is 32x32 int8, row-major,
 *   entries for codes >= 25 unused.
 *   Gap model: a gap of length L costs open + (L-1)*ext (both <= 0).
 */
#ifndef PASTIS_ORACLE_H
#define PASTIS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t score;
    int32_t v_beg, v_end;   /* row sequence (V) [beg, end) */
    int32_t h_beg, h_end;   /* column sequence (H) [beg, end) */
    int32_t matches;        /* identical residue columns */
    int32_t cols;           /* alignment columns (alignmentLength) */
    int32_t gap_opens;      /* maximal gap runs over both rows (numGapOpens) */
} or_aln_t;

typedef struct {
    int32_t beg_h, beg_v, end_h, end_v;  /* seed extents after extension */
    int32_t score_left, score_right;     /* longest-extension scores */
    int32_t upd_left, upd_right;         /* 1 if that side moved the seed */
    uint64_t cells;                      /* inner-loop cells evaluated */
} or_xdrop_t;

typedef struct {
    int32_t gap_open, gap_ext;   /* effective affine model for SW / NW */
    int32_t xdrop_gap;           /* linear gap of the X-drop extension */
    int32_t xdrop;               /* X */
    int32_t seed_len;            /* k */
    int32_t seed_cnt;            /* seeds tried per pair (1 or 2) */
    double  cov_thr;             /* aln_cov_thr, main.cpp:109 */
    int32_t ani_thr;             /* aln_ani_thr, main.cpp:110 */
} or_params_t;

typedef struct {
    or_aln_t aln;         /* stats of the winning seed / local alignment */
    float identity;       /* 100*matches/cols as SeqAn's float */
    float score_aln;      /* identity/100 when the predicate holds, else 0 */
    int32_t pass;
    int32_t seed;         /* winning seed index (xdrop path) */
    or_xdrop_t ext[2];    /* per-seed extension (xdrop path) */
} or_pair_out_t;

/* BLOSUM62 in the 25-letter order above, expanded to 32x32 (inc/ScoreMat.hpp:104-131). */
void or_blosum62_32(int8_t out[32 * 32]);

/* Gotoh local alignment, score and first-maximum end cell (column-major
 * over H, strict '>'); with_stats != 0 also runs the traceback and fills
 * begin, matches, cols and gap_opens.  Returns 0, or -1 on allocation failure. */
int or_sw_local(const uint8_t *h, int nh, const uint8_t *v, int nv,
                const int8_t *mat, int gap_open, int gap_ext, int with_stats,
                or_aln_t *out);

/* Gotoh global alignment with traceback statistics (no free end gaps). */
int or_nw_global(const uint8_t *h, int nh, const uint8_t *v, int nv,
                 const int8_t *mat, int gap_open, int gap_ext, int with_stats,
                 or_aln_t *out);

/* Gapped X-drop extension of a simple seed in both directions. */
int or_xdrop_extend(const uint8_t *h, int nh, const uint8_t *v, int nv,
                    int beg_h, int beg_v, int seed_len, const int8_t *mat,
                    int gap, int xdrop, or_xdrop_t *out);

/* Predicate and score_aln of SeqanFullAligner.cpp:247-261. */
int or_predicate(const or_aln_t *a, int len_h, int len_v, double cov_thr,
                 int ani_thr, float *identity, float *score_aln);

/* One --sfa pair: local alignment + statistics + predicate. */
int or_sfa_pair(const uint8_t *h, int nh, const uint8_t *v, int nv,
                const int8_t *mat, const or_params_t *p, or_pair_out_t *out);

/* One --sxa pair: per seed X-drop + global alignment of the infixes +
 * statistics; best identity (strict '>') wins; predicate on full lengths. */
int or_sxa_pair(const uint8_t *h, int nh, const uint8_t *v, int nv,
                const uint16_t seeds[4], const int8_t *mat,
                const or_params_t *p, or_pair_out_t *out);

/* Batched drivers over a packed sequence store (codes + offsets) and pair
 * index arrays; OpenMP over pairs.  nthreads <= 0 uses the default. */
int or_sw_batch(const uint8_t *seqs, const uint64_t *offs, const uint32_t *pr,
                const uint32_t *pc, uint64_t n, const int8_t *mat,
                int gap_open, int gap_ext, int32_t *score, int32_t *v_end,
                int32_t *h_end, int nthreads);
int or_sfa_batch(const uint8_t *seqs, const uint64_t *offs, const uint32_t *pr,
                 const uint32_t *pc, uint64_t n, const int8_t *mat,
                 const or_params_t *p, or_pair_out_t *out, int nthreads);
int or_sxa_batch(const uint8_t *seqs, const uint64_t *offs, const uint32_t *pr,
                 const uint32_t *pc, const uint16_t *seeds, uint64_t n,
                 const int8_t *mat, const or_params_t *p, or_pair_out_t *out,
                 int nthreads);

/* ---- k-mer candidates (kmer_oracle.c) ---------------------------------- */

/* One nonzero of the overlap matrix, laid out like the library's pa_pair:
 * count and score of CommonKmerLoc (ushort count wraps), seed[0..1] =
 * first.(row offset, column offset), seed[2..3] = second. */
typedef struct {
    uint32_t r, c;
    uint16_t count;
    int16_t score;
    uint16_t seed[4];
} or_cand_t;

/* Alphabet id 0 pdefault, 1 murphy10, 2 dssp10, 3 gbmr10, 4 td10, 5 diamond:
 * code_of[upper-case letter] (255 = not a k-mer letter), self score of the
 * mapped letter, and the alphabet size (the k-mer code's base). */
int or_alph_tables(int alph, uint8_t code_of[128], int16_t self_of[128],
                   int *size);

/* Candidates of row set x column set (cres NULL: the column set is the row
 * set), ASCII residues, k-mer length k, stride, alphabet, symmetric pruning
 * sym (0 none, 1 strict upper triangle, 2 even/odd) on global indices
 * r + roff, c + coff.  *out (free with or_free) holds *nout candidates in
 * column-major order (c, then r), before the k-mer threshold prune. */
int or_kmer_candidates(const char *rres, const uint64_t *roffs, uint64_t nr,
                       const char *cres, const uint64_t *coffs, uint64_t nc,
                       int k, int stride, int alph, int sym, uint64_t roff,
                       uint64_t coff, or_cand_t **out, uint64_t *nout);
void or_free(void *p);

/* ---- FASTA normalisation (fasta_oracle.c) ------------------------------ */

/* The sequences one process of the reference hands to the aligner for a
 * FASTA file of `size` bytes: *res_out / *offs_out (nseq + 1 entries) hold
 * the residues, *orig_out the record index of each sequence in the file.
 * Free the three arrays with or_free. */
int or_fasta_normalize(const char *file, uint64_t size, int k, char **res_out,
                       uint64_t **offs_out, uint64_t **orig_out,
                       uint64_t *nseq_out);

#ifdef __cplusplus
}
#endif
#endif
