/*
 * pastis_hip.h -- C ABI of the MI355X batched pairwise protein aligner.
 *
 * This is the device boundary behind the reference's aligner operator
 * pastis::PWAlign (inc/align/PWAlign.hpp:20-82).  Plain C types only: no
 * C++ exceptions, no torch or HIP types cross it.  Every call returns an
 * int status (PA_OK == 0); pa_strerror() names it and pa_last_error()
 * gives the context's last message.
 *
 * Sequence conventions follow the reference call sites:
 *   - the ROW sequence set is rseqs_ (SeqanFullAligner.hpp:27), indexed by
 *     the tuple's lr + bl_roffset; it is SeqAn's vertical sequence (seqsr);
 *   - the COLUMN sequence set is cseqs_ (SeqanFullAligner.hpp:28), indexed
 *     by lc + bl_coffset; it is SeqAn's horizontal sequence (seqsc).
 *   - residues are codes 0..24 in the order "ARNDCQEGHILKMFPSTWYVBZX*J"
 *     (src/Alphabet.cpp:26), or ASCII letters when loaded with
 *     PA_ENC_ASCII (case-insensitive, unknown letters become X);
 *   - a sequence holds at most 65535 residues (lengths are ushort in
 *     inc/FastaData.hpp:109-118).
 * Gap model of the DP kernels: a gap of length L costs
 *   gap_open + (L-1)*gap_ext   (both <= 0),
 * i.e. SeqAn's scoreGapOpen/scoreGapExtend.  pa_config_for_cli() maps the
 * reference's CLI values (-g, -e) onto that model (see PA_GAP_AS_CALLED).
 */
#ifndef PASTIS_HIP_H
#define PASTIS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PA_OK 0
#define PA_ERR_ARG 1        /* bad argument (null pointer, bad enum, ...) */
#define PA_ERR_HIP 2        /* HIP runtime error */
#define PA_ERR_NOMEM 3      /* device or host allocation failed */
#define PA_ERR_SEQ_LEN 4    /* a sequence is longer than 65535 */
#define PA_ERR_RESIDUE 5    /* a residue code is outside 0..24 */
#define PA_ERR_INDEX 6      /* a pair names a sequence that is not loaded */
#define PA_ERR_SEED 7       /* a seed does not fit inside its sequences */
#define PA_ERR_NOSEQ 8      /* sequences were not loaded */
#define PA_ERR_NODEV 9      /* no usable gfx950 device */
#define PA_ERR_LIMIT 10     /* a size limit of the entry point is exceeded */

/* ---- enums ----------------------------------------------------------- */
enum pa_mode { PA_MODE_SW = 0,     /* --sfa: full local Smith-Waterman */
               PA_MODE_XDROP = 1   /* --sxa: seed-and-extend X-drop    */ };
enum pa_side { PA_SIDE_ROW = 0, PA_SIDE_COL = 1, PA_SIDE_BOTH = 2 };
enum pa_enc { PA_ENC_CODES = 0, PA_ENC_ASCII = 1 };
enum pa_gap_model {
    /* SeqAn's Score(gapExtend, gapOpen) constructor receives the reference's
     * (gap_open, gap_ext) positionally (SeqanFullAligner.hpp:35-37), so the
     * CLI's -g is SeqAn's extension and -e its opening. */
    PA_GAP_AS_CALLED = 0,
    /* the CLI's documented meaning: -g opens, -e extends */
    PA_GAP_AS_DOCUMENTED = 1
};

/* result flags */
#define PA_FLAG_PASS 0x1     /* coverage and identity predicate held */
#define PA_FLAG_RERUN32 0x2  /* the int16 DP saturated; int32 re-run used */
#define PA_FLAG_EMPTY 0x4    /* local score 0: empty alignment */
#define PA_FLAG_NOTRACE 0x8  /* pa_align_graph only: the batch's trace codes
                                needed more memory than the graph was
                                captured for; the pair was not aligned */

typedef struct pa_config {
    int32_t mode;         /* enum pa_mode */
    int32_t gap_open;     /* DP gap model, see top of file */
    int32_t gap_ext;
    int32_t xdrop;        /* X of the X-drop extension (--sxa X) */
    int32_t xdrop_gap;    /* linear gap of the extension (SeqAn scoreGap) */
    int32_t seed_len;     /* k-mer length used as seed length (-k) */
    int32_t seed_cnt;     /* seeds tried per pair (--sc), 1 or 2 */
    int32_t ani_thr;      /* aln_ani_thr, src/main.cpp:110 */
    double cov_thr;       /* aln_cov_thr, src/main.cpp:109 */
    int32_t device;       /* HIP device ordinal */
    int32_t flags;        /* reserved, 0 */
    const int8_t *matrix; /* 32x32 row-major int8 scores, NULL = BLOSUM62 */
} pa_config;

/* One candidate pair.  r, c are indices into the loaded row / column sets
 * (the tuple's lr + bl_roffset and lc + bl_coffset).  count, kscore and
 * seed mirror CommonKmerLoc (inc/Types.hpp:163-215): seed[0..1] =
 * first.(row offset, col offset), seed[2..3] = second.(row, col). */
typedef struct pa_pair {
    uint32_t r, c;
    uint16_t count;
    int16_t kscore;
    uint16_t seed[4];
} pa_pair;

/* Full per-pair result of pa_align. */
typedef struct pa_result {
    int32_t score;           /* local score (SW) / global infix score (X-drop) */
    uint16_t r_beg, r_end;   /* aligned range in the row sequence [beg, end) */
    uint16_t c_beg, c_end;   /* aligned range in the column sequence */
    uint32_t matches;        /* AlignmentStats::numMatches */
    uint32_t cols;           /* AlignmentStats::alignmentLength */
    uint32_t gap_opens;      /* AlignmentStats::numGapOpens */
    float identity;          /* AlignmentStats::alignmentIdentity */
    float score_aln;         /* identity/100 if PA_FLAG_PASS else 0 */
    uint16_t flags;          /* PA_FLAG_* */
    uint16_t seed;           /* winning seed (X-drop mode) */
} pa_result;

/* Score and end cell of a local alignment (no traceback). */
typedef struct pa_score {
    int32_t score;
    uint16_t r_end, c_end;   /* exclusive ends; 0,0 when score is 0 */
} pa_score;

/* Seed extension of one seed. */
typedef struct pa_extent {
    uint16_t r_beg, r_end, c_beg, c_end;
} pa_extent;

/* Counters and timers of the calls made on a context so far. */
typedef struct pa_stats {
    uint64_t pairs;          /* pairs aligned */
    uint64_t dp_cells;       /* DP cells of the SW / global kernels */
    uint64_t xdrop_cells;    /* cells evaluated by the X-drop kernel */
    uint64_t rerun_pairs;    /* tasks re-run in int32 */
    double ms_sw;            /* device time of the SW kernels */
    double ms_xdrop;         /* device time of the X-drop kernels */
    double ms_nw;            /* device time of the global-alignment kernels */
    double ms_trace;         /* device time of the traceback kernels */
    double ms_total;         /* device time of whole calls (event to event) */
    uint64_t launches;       /* kernel launches */
    uint64_t graph_captures; /* graphs captured by pa_sw_scores_graph */
} pa_stats;

/* ---- lifetime -------------------------------------------------------- */

/* Replaces the aligner constructors SeqanFullAligner(gap_open, gap_ext)
 * (inc/align/SeqanFullAligner.hpp:35-37) and SeqanXdropAligner(gap_open,
 * gap_ext, seed_len, xdrop, seed_cnt) (inc/align/SeqanXdropAligner.hpp:38-42),
 * called from src/sim_search.cpp:119-124. */
int pa_create(const pa_config *cfg, struct pa_ctx **out);

/* Fills cfg from the reference's CLI values (-g, -e, -k, --sxa X, --sc) and a
 * gap model (enum pa_gap_model); mode is PA_MODE_SW when xdrop <= 0. */
int pa_config_for_cli(pa_config *cfg, int gap_open_cli, int gap_ext_cli,
                      int klength, int xdrop, int seed_count, int gap_model);

/* Replaces ~SeqanFullAligner / ~SeqanXdropAligner. */
void pa_destroy(struct pa_ctx *ctx);

const char *pa_strerror(int status);
const char *pa_last_error(struct pa_ctx *ctx);

/* ---- sequences ------------------------------------------------------- */

/* Coverage and identity thresholds of the predicate for later calls; the
 * reference passes them with every batch (params.aln_cov_thr, aln_ani_thr,
 * SeqanFullAligner.cpp:253-256). */
int pa_set_thresholds(struct pa_ctx *ctx, double cov_thr, int ani_thr);

/* Replaces construct_seqs / construct_seqs_bl
 * (src/align/SeqanFullAligner.cpp:28-168, SeqanXdropAligner.cpp:32-185):
 * uploads one sequence set once.  residues[offsets[i] .. offsets[i+1]) is
 * sequence i; offsets has nseq+1 entries.  PA_SIDE_BOTH makes the column set
 * the same as the row set (the diagonal-block case, SeqanFullAligner.cpp:82-87). */
int pa_load_seqs(struct pa_ctx *ctx, int side, const uint8_t *residues,
                 const uint64_t *offsets, uint64_t nseq, int encoding);

/* ---- the operator ---------------------------------------------------- */

/* Replaces PWAlign::aln_batch for both tuple types
 * (inc/align/PWAlign.hpp:45-63; SeqanFullAligner.cpp:173-266 in PA_MODE_SW,
 * SeqanXdropAligner.cpp:190-361 in PA_MODE_XDROP).  Host pointers; n == 0 is
 * legal (the beg == end batch of src/DistPWRunner.cpp:85-86,118-131).
 * Synchronous. */
int pa_align(struct pa_ctx *ctx, const pa_pair *pairs, uint64_t n,
             pa_result *out);

/* The same on device-resident pairs and results (pointers from
 * pa_device_alloc or hipMalloc on the context's device); stream may be NULL
 * for the context's own stream.  Returns after the work is queued; the
 * results are complete when the stream is.  The device entry points do not
 * return PA_ERR_INDEX / PA_ERR_SEED: a pair naming a sequence outside the
 * loaded sets, or (X-drop) whose seed does not fit inside its sequences, is
 * aligned as the empty alignment (score 0, empty extents, not passing), and
 * no kernel reads outside the sequence store for it. */
int pa_align_dev(struct pa_ctx *ctx, const pa_pair *d_pairs, uint64_t n,
                 pa_result *d_out, void *stream);

/* Score and end cell only: the localAlignment call of
 * SeqanFullAligner.cpp:231 without traceback.  Host and device variants. */
int pa_sw_scores(struct pa_ctx *ctx, const pa_pair *pairs, uint64_t n,
                 pa_score *out);
int pa_sw_scores_dev(struct pa_ctx *ctx, const pa_pair *d_pairs, uint64_t n,
                     pa_score *d_out, void *stream);

/* pa_sw_scores_dev as a replayed HIP graph: the first call with a given
 * (d_pairs, n, d_out, stream) captures the batch's kernels into a graph,
 * later calls replay it; up to 8 such graphs are kept (least recently used
 * dropped), so a loop over the batches of a resident pair array replays one
 * graph per batch.  The pairs' contents may change between calls.  Every
 * graph is recaptured after anything that moved or resized the context's
 * workspace or changed its sequence sets or tuning (a larger eager batch,
 * pa_load_seqs, pa_set_tuning), so eager calls may be mixed with replays.
 * The graphs share the context's workspace: replay them on one stream.
 * Replaces the per-batch launches behind pw_aln_batch's loop
 * (src/DistPWRunner.cpp:116-131) for score-only batches of equal size. */
int pa_sw_scores_graph(struct pa_ctx *ctx, const pa_pair *d_pairs, uint64_t n,
                       pa_score *d_out, void *stream);

/* pa_align_dev (the traced operator in the context's mode: --sfa local
 * alignment, or --sxa X-drop + global alignment, with traceback statistics
 * and the predicate) as a replayed HIP graph, with the caching rules of
 * pa_sw_scores_graph.  The first call with a given (d_pairs, n, d_out,
 * stream) aligns the batch eagerly, sizes the trace store from it (twice
 * the batch's trace codes, at most the context's trace memory) and
 * captures the batch's kernels with the trace store split into chunks on
 * the device; later calls replay the graph without any host
 * synchronisation.  A replayed batch whose trace codes do not fit in the
 * captured chunks is not aligned: its results carry PA_FLAG_NOTRACE (align
 * it with pa_align_dev).  Replaces the per-batch launches behind
 * pw_aln_batch's loop (src/DistPWRunner.cpp:116-131) for traced batches. */
int pa_align_graph(struct pa_ctx *ctx, const pa_pair *d_pairs, uint64_t n,
                   pa_result *d_out, void *stream);

/* pa_sw_scores_dev without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The call's workspace is a fixed 64 MiB device array of the
 * library (one per GPU, held until the process ends); it takes at most
 * 262144 pairs per call and loaded sequences of at most 4096 residues
 * (PA_ERR_LIMIT otherwise), runs no event timers, and calls on one GPU that
 * overlap in time race on that array.  Results equal pa_sw_scores_dev's. */
int pa_sw_scores_noalloc(struct pa_ctx *ctx, const pa_pair *d_pairs,
                         uint64_t n, pa_score *d_out, void *stream);

/* Seed extension only: extendSeed(..., GappedXDrop) of
 * SeqanXdropAligner.cpp:279-289 for seed `seed` (0 or 1) of every pair. */
int pa_xdrop_extents(struct pa_ctx *ctx, const pa_pair *pairs, uint64_t n,
                     int seed, pa_extent *out);

/* k-mer prefilter of src/DistPWRunner.cpp:48-54: keep[i] = 0 when
 * count <= ckthr || kscore <= mosthr, else 1.  Returns the kept count. */
int pa_kmer_prune(const pa_pair *pairs, uint64_t n, int ckthr, float mosthr,
                  uint8_t *keep, uint64_t *kept);

/* ---- k-mer candidate generation ---------------------------------------
 * Replaces the overlap matrix of the reference's similarity search: the
 * sequence-by-k-mer matrix A of generate_A (src/kmer/KmerOps.cpp:92-245)
 * multiplied by its transpose under the KmerIntersectLoc semiring
 * (inc/sr.hpp:30-117; src/sim_search.cpp:154-157, 533-534), computed for the
 * loaded row x column sets by a sort-based join on the GPU.  Each pair of
 * sequences sharing a k-mer gets count = shared k-mer occurrence pairs
 * (ushort, wraps), kscore = the largest BLOSUM62 self score of a shared
 * k-mer, seed[0..1] = (row, column) offsets of the first product and
 * seed[2..3] those of the last one ((0, 0) when count is 1), the products
 * being added in k-mer code order, then column offset, then row offset.
 * The symmetric pruning of the similarity search and the k-mer threshold
 * prune of src/DistPWRunner.cpp:48-54 are fused in. */
enum pa_alphabet {       /* --alph, inc/Alphabet.hpp:66-254 */
    PA_ALPH_PDEFAULT = 0, PA_ALPH_MURPHY10 = 1, PA_ALPH_DSSP10 = 2,
    PA_ALPH_GBMR10 = 3, PA_ALPH_TD10 = 4, PA_ALPH_DIAMOND = 5
};
enum pa_symmetric {
    PA_SYM_NONE = 0,     /* every pair, the diagonal included */
    PA_SYM_UPPER = 1,    /* global column > global row: mult_aln and the
                            partial blocks of mult_aln_bl_trg,
                            src/sim_search.cpp:403-412, 569-578 */
    PA_SYM_EVEN_ODD = 2  /* mult_aln_bl_idx, src/sim_search.cpp:173-187:
                            r != c, same parity below the diagonal,
                            different parity above it */
};

typedef struct pa_kmer_config {
    int32_t k;           /* k-mer length (-k); alphabet size ^ k < 2^31 */
    int32_t stride;      /* k-mer stride (--ks), >= 1 */
    int32_t alphabet;    /* enum pa_alphabet */
    int32_t symmetric;   /* enum pa_symmetric, on global indices */
    int32_t ckthr;       /* drop count <= ckthr (--ckthr) */
    float mosthr;        /* ... or kscore <= mosthr (--mosthr) */
    uint64_t row_base;   /* global index of loaded row sequence 0 */
    uint64_t col_base;   /* global index of loaded column sequence 0 */
} pa_kmer_config;

/* Candidates of all row sequences x column sequences [c_begin, c_end) as
 * pa_pair (r, c = loaded-set indices) in column-major order (c, then r).
 * *n_out receives their number.  With d_out NULL only the count is
 * computed; a second call with the same arguments (and no pa_load_seqs in
 * between) then copies the kept result into d_out (device memory, cap
 * entries) without recomputing; PA_ERR_LIMIT if it holds more than cap.
 * Synchronous. */
int pa_kmer_candidates(struct pa_ctx *ctx, const pa_kmer_config *kc,
                       uint64_t c_begin, uint64_t c_end, pa_pair *d_out,
                       uint64_t cap, uint64_t *n_out);

/* ---- input and output formats -----------------------------------------
 * FASTA normalisation: the sequences one process of the reference hands to
 * its aligners for a FASTA file (the one-rank reader of
 * src/DistFastaData.cpp:161-263 drops the file's last byte; the FastaData
 * constructor, src/FastaData.cpp:13-148, removes newlines inside sequences
 * and '*', turns U/u into T, drops sequences shorter than k -- taking the
 * previous sequence's last residue with a dropped record that is not the
 * last -- and get_sequence, inc/FastaData.hpp:105-129, reads lengths as
 * ushort).  Residues stay ASCII (load them with PA_ENC_ASCII). */
struct pa_fasta;
int pa_fasta_parse(const char *data, uint64_t size, int k,
                   struct pa_fasta **out);
/* counts: sequences, residues, sequences whose length wrapped past 65535 */
int pa_fasta_info(const struct pa_fasta *f, uint64_t *nseq, uint64_t *nres,
                  uint64_t *nwrapped);
const char *pa_fasta_residues(const struct pa_fasta *f);
const uint64_t *pa_fasta_offsets(const struct pa_fasta *f);   /* nseq + 1 */
const uint64_t *pa_fasta_records(const struct pa_fasta *f);   /* record
                                  index in the file of each sequence */
void pa_fasta_free(struct pa_fasta *f);
/* pa_fasta_parse + pa_load_seqs(PA_ENC_ASCII): construct_seqs from a FASTA
 * buffer (SeqanXdropAligner.cpp:32-105 over DistFastaData). */
int pa_load_fasta(struct pa_ctx *ctx, int side, const char *data,
                  uint64_t size, int k);

/* One nonzero of the similarity matrix: global row, column, score_aln. */
typedef struct pa_triple {
    uint64_t row, col;
    float score_aln;
    uint32_t pad;
} pa_triple;

/* The pairs whose alignment passed (score_aln != 0, the post-prune of
 * src/DistPWRunner.cpp:149-153) as triples (r + row_base, c + col_base,
 * score_aln) in the pairs' order, collected like src/sim_search.cpp:223-232.
 * d_out holds up to n triples; *n_out receives their number.  Synchronous. */
int pa_triples_dev(struct pa_ctx *ctx, const pa_pair *d_pairs,
                   const pa_result *d_res, uint64_t n, uint64_t row_base,
                   uint64_t col_base, pa_triple *d_out, uint64_t *n_out,
                   void *stream);
/* Writes the similarity matrix as Matrix Market (coordinate real general,
 * 1-based, column-major order, %g values), the output of
 * S.ParallelWriteMM(file, true) (src/sim_search.cpp:257-269). */
int pa_write_mm(const char *path, uint64_t nrows, uint64_t ncols,
                const pa_triple *triples, uint64_t n);

/* ---- learned prefilter ------------------------------------------------
 * Stands beside the k-mer prefilter of src/DistPWRunner.cpp:48-54 (which
 * prunes on count <= ckthr || score <= mosthr): an MLP 16 -> 64 -> 64 -> 1
 * over per-pair features predicts whether the exact alignment will pass the
 * coverage/identity predicate (SeqanXdropAligner.cpp:344-351,
 * SeqanFullAligner.cpp:253-260).  It has no counterpart in the reference.
 * Weights and activations are bf16, accumulation fp32.  Parameters are a flat
 * fp32 array: W1[16][64] b1[64] W2[64][64] b2[64] w3[64] b3 (in-major W).
 * Features are standardized as (f - mean) * istd before the first layer. */
#define PA_FILTER_NF 16
#define PA_FILTER_NPARAM 5313

typedef struct pa_train_config {
    uint32_t epochs;
    uint32_t batch;         /* samples per Adam step, 1..65536 */
    uint64_t seed;          /* shuffle seed of every epoch; 0 keeps order */
    float lr, beta1, beta2, eps;
    float pos_weight;       /* weight of positive samples in the BCE */
    int standardize;        /* 1: set mean/istd from these features first */
} pa_train_config;

/* Features of pairs over the loaded sequence sets (n x 16 fp32). */
int pa_filter_features(struct pa_ctx *ctx, const pa_pair *pairs, uint64_t n,
                       float *feats);
int pa_filter_features_dev(struct pa_ctx *ctx, const pa_pair *d_pairs,
                           uint64_t n, float *d_feats, void *stream);
/* Installs parameters (PA_FILTER_NPARAM floats) and the standardization
 * (16 + 16 floats; NULL keeps the current one, initially mean 0, istd 1);
 * resets the Adam state. */
int pa_filter_set_model(struct pa_ctx *ctx, const float *params,
                        const float *mean, const float *istd);
int pa_filter_get_model(struct pa_ctx *ctx, float *params, float *mean,
                        float *istd);
/* Mean weighted BCE over the n samples (one batch, in order) and its
 * gradient (PA_FILTER_NPARAM floats) at the current parameters; n <= 65536. */
int pa_filter_grad(struct pa_ctx *ctx, const float *feats,
                   const uint8_t *labels, uint64_t n, float pos_weight,
                   float *grad, float *loss);
/* Adam training on the GPU; epoch_loss (may be NULL) receives the mean loss
 * of every epoch. */
int pa_filter_train(struct pa_ctx *ctx, const float *feats,
                    const uint8_t *labels, uint64_t n,
                    const pa_train_config *tc, float *epoch_loss);
/* Logits of precomputed features. */
int pa_filter_logits(struct pa_ctx *ctx, const float *feats, uint64_t n,
                     float *logits);
/* Fused features + MLP on device pairs. */
int pa_filter_score_dev(struct pa_ctx *ctx, const pa_pair *d_pairs,
                        uint64_t n, float *d_logits, void *stream);
/* Keeps the pairs whose logit is >= tau (kept order is unspecified;
 * d_kept_index may be NULL); synchronizes to return *n_kept. */
int pa_filter_select_dev(struct pa_ctx *ctx, const pa_pair *d_pairs,
                         uint64_t n, float tau, pa_pair *d_kept,
                         uint32_t *d_kept_index, uint64_t *n_kept,
                         void *stream);

/* ---- device memory helpers (for the _dev entry points) --------------- */
int pa_device_alloc(struct pa_ctx *ctx, size_t bytes, void **ptr);
int pa_device_free(struct pa_ctx *ctx, void *ptr);
/* Page-locked host memory: copies between it and the device run at full
 * link speed and without a staging copy (HipAligner stages its batches
 * through it). */
int pa_host_alloc(struct pa_ctx *ctx, size_t bytes, void **ptr);
int pa_host_free(struct pa_ctx *ctx, void *ptr);
int pa_memcpy_h2d(struct pa_ctx *ctx, void *dst, const void *src, size_t bytes);
int pa_memcpy_d2h(struct pa_ctx *ctx, void *dst, const void *src, size_t bytes);
int pa_synchronize(struct pa_ctx *ctx);

/* Per kernel class: DP kernels are split by mode (sw = local score only,
 * swt = local traced, nwt = global traced), row class (c0..c3, with the
 * class's current kernel shape, e.g. l64r8 = 64 lanes x 8 rows per lane)
 * and arithmetic (i16 packed, i32); "rerun" classes are the int32 re-runs
 * of saturated tasks.  Example: "nwt.c3.l64r8.i16".  cells are DP cells of
 * the tasks the class ran. */
typedef struct pa_kernel_stat {
    char name[24];
    uint64_t launches;
    uint64_t cells;
    double ms;              /* summed device time of the class's launches */
} pa_kernel_stat;

/* ---- tuning ---------------------------------------------------------
 * Launch shapes chosen per context (no counterpart in the reference; the
 * in-library autotuner pa_autotune below searches them).  Results never
 * depend on them, only the speed does. */

/* Kernel shapes of the DP kernels: L lanes per task (64 / L tasks per wave)
 * and R rows per lane per half-strip, so one pass covers 2 * R * L rows;
 * longer tasks take several passes.  The R = 8 shapes keep 4-bit query
 * profiles; with a scoring matrix whose scores over the alphabet span more
 * than 16 values (or whose minimum is not negative) the context runs the
 * R = 4 shape with the same rows per pass in their place (pa_get_tuning
 * still returns the shapes that were set). */
enum pa_dp_shape {
    PA_SHAPE_L16R4 = 0,     /* 128 rows per pass */
    PA_SHAPE_L16R8 = 1,     /* 256 */
    PA_SHAPE_L64R4 = 2,     /* 512 */
    PA_SHAPE_L64R8 = 3,     /* 1024 */
    PA_SHAPE_L32R8 = 4,     /* 512 */
    PA_SHAPE_L8R8 = 5       /* 128 */
};

typedef struct pa_tuning {
    int32_t dp_wg_per_cu;   /* persistent DP workgroups (one wave) per CU,
                               1..32; 0 = as many as occupancy allows */
    int32_t xd_wg_per_cu;   /* X-drop workgroups (one wave) per CU, 0 = max */
    int32_t xd_ring;        /* X-drop LDS ring slots per lane: 16 or 32 */
    int32_t xd_coop;        /* 1: X-drop items go to the 16-lanes-per-item
                               kernel first (band up to 14 columns), the rest
                               to the one-lane-per-item kernels; 0: one lane
                               per item only */
    int32_t dp_edge[3];     /* DP row classes: a task whose row span has
                               qlen rows is in the first class c with
                               qlen <= dp_edge[c], class 3 takes the rest;
                               strictly increasing, 1..65535; default
                               128, 256, 512 */
    int32_t dp_shape[4];    /* pa_dp_shape of each row class; default
                               L8R8, L16R8, L32R8, L64R8 */
    int32_t tb_stage;       /* 1 (default): the traceback copies each lane
                               run of the trace it enters into LDS with wide
                               loads; 0: one byte load per step */
    int32_t xd_wave;        /* X-drop items whose band outgrows the
                               per-lane kernel's rings continue with 16
                               lanes per item: 1 (default) in a 128-column
                               window, then a 256-column one; 2: the
                               256-column window only; 0: none.  What is
                               left runs with global-memory rings, one lane
                               per item */
    int32_t tb_overlap;     /* 1: each class's tracebacks run on a second
                               stream beside the next classes' DP; 2: those
                               of all classes but the last beside the last
                               class's DP; 0 (default): all tracebacks
                               after the DP */
} pa_tuning;

int pa_get_tuning(struct pa_ctx *ctx, pa_tuning *out);
int pa_set_tuning(struct pa_ctx *ctx, const pa_tuning *t);

/* What pa_autotune times. */
enum pa_tune_target {
    PA_TUNE_SW_SCORES = 0,  /* pa_sw_scores_dev */
    PA_TUNE_ALIGN = 1       /* pa_align_dev in the context's mode */
};
#define PA_TUNE_FORCE 1u    /* time again even when the cache has a choice */

/* Autotuner (no counterpart in the reference): times the pairs at d_pairs
 * (device memory; at most `sample` of them, taken evenly spaced; 0 = a
 * default of 262144) under candidate tunings and installs the fastest on
 * the context.  The DP row-class edges, each class's kernel shape, the
 * persistent-grid size, the traceback staging and the X-drop kernel choice
 * are each chosen by the device time of the stage they affect.  The choice
 * is cached per device, target and alignment mode, so later contexts on the
 * same device (and HipAligner objects) install it without timing; `flags`
 * PA_TUNE_FORCE times again.  The caller's stream work is finished first;
 * the context's sequences must be loaded; its statistics (pa_get_stats) are
 * reset afterwards.  Sample buffers are allocated for the call.  *out
 * (optional) receives the installed tuning. */
int pa_autotune(struct pa_ctx *ctx, const pa_pair *d_pairs, uint64_t n,
                int target, uint64_t sample, uint32_t flags, pa_tuning *out);

/* ---- counters --------------------------------------------------------- */

/* Timers and counters of the calls made on a context, the device-side
 * counterpart of the reference's TimePod timers (inc/TraceUtils.hpp:25-217:
 * align|pre, align|cpu, align|post) and aln_pairs counters. */
int pa_get_stats(struct pa_ctx *ctx, pa_stats *out);
int pa_reset_stats(struct pa_ctx *ctx);
/* Fills up to cap entries; *n receives the number of classes with launches. */
int pa_get_kernel_stats(struct pa_ctx *ctx, pa_kernel_stat *out, int cap,
                        int *n);

#ifdef __cplusplus
}
#endif
#endif /* PASTIS_HIP_H */
