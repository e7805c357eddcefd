// k_kmer.hip -- k-mer candidate generation on the GPU (pa_kmer_candidates).
//
// The reference builds the candidate pairs as the sparse product C = A * A^T
// of the sequence-by-k-mer matrix A (generate_A, src/kmer/KmerOps.cpp:92-245)
// under the KmerIntersectLoc semiring (inc/sr.hpp:30-117) with CombBLAS.
// Here the same overlap matrix is a sort-based join:
//   1. every k-mer occurrence of a sequence set becomes (code, seq << 16 |
//      offset); codes of k-mers with a letter outside the alphabet are
//      marked invalid (add_kmers, KmerOps.cpp:31-87);
//   2. a stable radix sort by code groups equal k-mers, occurrences in
//      (sequence, offset) order inside a group; the sorted occurrences and
//      the group table of a set are kept for later calls;
//   3. for each k-mer group, the products are the (column occurrence, row
//      occurrence) pairs, column occurrences outer: the order in which a
//      column-by-column SpGEMM adds them;
//   4. a stable radix sort of the products by (column, row) lines up the
//      products of each pair in that order, and one pass over each run
//      folds them: count = products (ushort), score = max self score,
//      first = the first product's offsets, second = the last one's (the
//      default (0, 0) for a single product).
// The symmetric pruning of the similarity search (src/sim_search.cpp:
// 173-187, 403-412, 569-578) is applied to the products, the k-mer
// threshold prune of DistPWRunner.cpp:48-54 to the folded pairs.  Output:
// pa_pair in column-major order (column, then row).
//
// All of this is integer sort / scan / gather work bound by HBM traffic.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "pa_ctx.h"

using namespace pa;

namespace pa {

// The k-mer alphabets of inc/Alphabet.hpp:66-254: each letter of the
// 25-letter protein order maps to a representative; codes are given to
// representatives in the order they first appear while scanning that order
// (Alphabet::init, src/Alphabet.cpp:31-51).
struct KmerAlphabet {
    uint8_t code[kAlpha];    // alphabet code of each residue code
    int16_t self[kAlpha];    // BLOSUM62 self score of the representative
    int size;
};

struct KmerSet {            // sorted occurrences of one sequence set
    Buf keys, vals, gstart, gcode;
    uint64_t nocc = 0;      // valid occurrences
    uint64_t ngroups = 0;   // distinct codes
    uint64_t gen = ~0ull;   // sequence generation they were built for
    int k = 0, stride = 0, alph = -1;
};

struct KmerState {
    KmerSet rows, cols;
    Buf slots, keys_a, vals_a, pcount, poff, gscr, pkey, pval, pkey2, pval2,
        heads, runid, rstart, keep, kpos, tmp, cnt, cand;
    // the last call's result, reused by an identical call (the count / copy
    // protocol of pa_kmer_candidates)
    pa_kmer_config last{};
    uint64_t last_cb = 0, last_ce = 0, last_gen = ~0ull, last_n = 0;
    bool have_last = false;
};

void kmer_free(KmerState *k)
{
    if (!k)
        return;
    for (KmerSet *s : {&k->rows, &k->cols})
        for (Buf *b : {&s->keys, &s->vals, &s->gstart, &s->gcode})
            release(*b);
    for (Buf *b : {&k->slots, &k->keys_a, &k->vals_a, &k->pcount, &k->poff,
                   &k->gscr, &k->pkey, &k->pval, &k->pkey2, &k->pval2, &k->heads,
                   &k->runid, &k->rstart, &k->keep, &k->kpos, &k->tmp,
                   &k->cnt, &k->cand})
        release(*b);
    delete k;
}

}  // namespace pa

namespace {

const char kProtein[] = "ARNDCQEGHILKMFPSTWYVBZX*J";
const int8_t kSelfB62[kAlpha] = {4, 5, 6, 6, 9, 5, 5, 6, 8, 4, 4, 5, 5,
                                 6, 7, 4, 5, 11, 7, 4, 4, 4, -1, 1, 4};
// letter groups of the reduced alphabets; the first letter represents them
const char *const kGroups[6][8] = {
    {nullptr},
    {"DENQ", "FWY", "ILMV", "KR", "ST", nullptr},               // murphy10
    {"EKQR", "IV", "LY", "AM", "HT", "DNS", "GP", nullptr},     // dssp10
    {"AEFIKLMQRVW", nullptr},                                   // gbmr10
    {"EKRQ", "DSN", "HC", "IV", "WYF", "LM", nullptr},          // td10
    {"KREDQN", "ILV", "STA", nullptr},                          // diamond
};

bool make_alphabet(int id, KmerAlphabet &a)
{
    if (id < 0 || id > 5)
        return false;
    auto idx = [](char ch) { return (int)(std::strchr(kProtein, ch) - kProtein); };
    int rep[kAlpha];
    for (int i = 0; i < kAlpha; ++i)
        rep[i] = i;
    for (int g = 0; kGroups[id][g]; ++g)
        for (const char *p = kGroups[id][g]; *p; ++p)
            rep[idx(*p)] = idx(kGroups[id][g][0]);
    int code_of_rep[kAlpha];
    std::fill(code_of_rep, code_of_rep + kAlpha, -1);
    int next = 0;
    for (int i = 0; i < kAlpha; ++i) {
        if (code_of_rep[rep[i]] < 0)
            code_of_rep[rep[i]] = next++;
        a.code[i] = (uint8_t)code_of_rep[rep[i]];
        a.self[i] = kSelfB62[rep[i]];
    }
    a.size = next;
    return true;
}

// smallest b with 2^b - 1 >= v (v > 0)
int bits_for(uint64_t v)
{
    int b = 1;
    while (b < 64 && ((1ull << b) - 1ull) < v)
        ++b;
    return b;
}

constexpr int kTB = 256;

inline unsigned nblocks(uint64_t n, unsigned per = kTB)
{
    uint64_t b = (n + per - 1) / per;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

// Occurrence slot t -> (code or invalid, seq << 16 | offset).  slots[i] is
// the first slot of sequence i (exclusive scan of its k-mer positions).
__global__ void k_kmer_extract(const uint8_t *__restrict__ kc,
                               const uint64_t *__restrict__ seq_addr,
                               uint64_t base_addr,
                               const uint64_t *__restrict__ slots,
                               uint64_t nseq, uint64_t nslots, int k,
                               int stride, KmerAlphabet alph,
                               uint32_t invalid, uint32_t *__restrict__ keys,
                               uint64_t *__restrict__ vals,
                               unsigned long long *nbad)
{
    uint64_t bad = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         t < nslots; t += (uint64_t)gridDim.x * blockDim.x) {
        // sequence of the slot: last i with slots[i] <= t
        uint64_t lo = 0, hi = nseq - 1;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (slots[mid] <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        const uint64_t off = (t - slots[lo]) * (uint64_t)stride;
        const uint8_t *p = kc + (seq_addr[lo] - base_addr) + off;
        uint32_t code = 0;
        bool ok = true;
        for (int j = 0; j < k; ++j) {
            const int x = p[j];
            if (x >= kAlpha) {
                ok = false;
                break;
            }
            code = code * (uint32_t)alph.size + alph.code[x];
        }
        keys[t] = ok ? code : invalid;
        vals[t] = (lo << 16) | off;
        bad += ok ? 0 : 1;
    }
    if (bad)
        atomicAdd(nbad, (unsigned long long)bad);
}

// heads[i] = 1 where a new code group starts (i < n)
__global__ void k_group_heads(const uint32_t *__restrict__ keys, uint64_t n,
                              uint32_t *__restrict__ heads)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// group g (inclusive scan value - 1 at a head) -> start, code
__global__ void k_group_table(const uint32_t *__restrict__ keys,
                              const uint32_t *__restrict__ heads,
                              const uint64_t *__restrict__ incl, uint64_t n,
                              uint64_t *__restrict__ gstart,
                              uint32_t *__restrict__ gcode)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (heads[i]) {
            const uint64_t g = incl[i] - 1;
            gstart[g] = i;
            gcode[g] = keys[i];
        }
}

__device__ inline uint64_t lower_bound_u32(const uint32_t *a, uint64_t n,
                                           uint32_t v)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// first index in vals[lo, hi) whose sequence is >= s
__device__ inline uint64_t lower_bound_seq(const uint64_t *vals, uint64_t lo,
                                           uint64_t hi, uint64_t s)
{
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if ((vals[mid] >> 16) < s)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Per column group g: the column occurrences in [c_begin, c_end) and the
// row group with the same code; products = both counts multiplied.
struct JoinArgs {
    const uint64_t *cgstart;   // column groups (and their end at ngroups)
    const uint32_t *cgcode;
    const uint64_t *cvals;
    uint64_t cngroups, cnocc;
    const uint32_t *rkeys;     // row occurrences (sorted keys)
    uint64_t rnocc;
    const uint64_t *rgstart;   // row groups when the sets are shared
    bool shared;
    uint64_t c_begin, c_end;
    // per group: [ca, cb) columns, [ra, rb) rows
    uint64_t *ca, *ra;
    uint32_t *cn, *rn;
    uint64_t *pcount;
};

__global__ void k_join_count(JoinArgs a)
{
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         g < a.cngroups; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t gs = a.cgstart[g];
        const uint64_t ge = g + 1 < a.cngroups ? a.cgstart[g + 1] : a.cnocc;
        const uint64_t c0 = lower_bound_seq(a.cvals, gs, ge, a.c_begin);
        const uint64_t c1 = lower_bound_seq(a.cvals, c0, ge, a.c_end);
        uint64_t r0, r1;
        if (a.shared) {
            r0 = gs;
            r1 = ge;
        } else {
            const uint32_t code = a.cgcode[g];
            r0 = lower_bound_u32(a.rkeys, a.rnocc, code);
            r1 = lower_bound_u32(a.rkeys, a.rnocc, code + 1);
        }
        a.ca[g] = c0;
        a.ra[g] = r0;
        a.cn[g] = (uint32_t)(c1 - c0);
        a.rn[g] = (uint32_t)(r1 - r0);
        a.pcount[g] = (c1 - c0) * (r1 - r0);
    }
}

struct EmitArgs {
    const uint64_t *poff;     // exclusive scan of products per group
    uint64_t ngroups, nprod;
    const uint64_t *ca, *ra;
    const uint32_t *cn, *rn;
    const uint32_t *gcode;
    const uint64_t *cvals, *rvals;
    KmerAlphabet alph;
    int k;
    int sym;                  // pa_symmetric
    uint64_t row_base, col_base, c_begin;
    int rbits;
    uint64_t keymax;
    uint64_t *pkey, *pval;
    unsigned long long *ndrop;
};

// self score of a k-mer code (sum over its letters' representatives)
__device__ inline int code_cost(uint32_t code, const KmerAlphabet &al, int k)
{
    // representative self score per alphabet code: the first residue code
    // mapped to it
    int cost = 0;
    for (int j = 0; j < k; ++j) {
        const uint32_t d = code % (uint32_t)al.size;
        code /= (uint32_t)al.size;
        int s = 0;
        for (int x = 0; x < kAlpha; ++x)
            if (al.code[x] == d) {
                s = al.self[x];
                break;
            }
        cost += s;
    }
    return cost;
}

constexpr int kPerThread = 8;

// Products in the order of the scanned group offsets: thread t emits
// products [8t, 8t + 8); inside a group, column occurrence outer.
__global__ void k_emit_products(EmitArgs a)
{
    uint64_t drop = 0;
    const uint64_t t0 =
        ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kPerThread;
    if (t0 < a.nprod) {
        // group of product t0: last g with poff[g] <= t0 and a product
        uint64_t lo = 0, hi = a.ngroups - 1;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (a.poff[mid] <= t0)
                lo = mid;
            else
                hi = mid - 1;
        }
        uint64_t g = lo;
        uint64_t local = t0 - a.poff[g];
        int cost = -1;
        uint64_t cost_g = ~0ull;
        const uint64_t t1 = t0 + kPerThread < a.nprod ? t0 + kPerThread : a.nprod;
        for (uint64_t t = t0; t < t1; ++t) {
            while ((uint64_t)a.cn[g] * a.rn[g] <= local) {
                local = 0;
                ++g;
            }
            if (cost_g != g) {
                cost = code_cost(a.gcode[g], a.alph, a.k);
                cost_g = g;
            }
            const uint64_t bi = local / a.rn[g], ai = local % a.rn[g];
            const uint64_t cv = a.cvals[a.ca[g] + bi];
            const uint64_t rv = a.rvals[a.ra[g] + ai];
            const uint64_t c = cv >> 16, r = rv >> 16;
            const uint64_t R = r + a.row_base, C = c + a.col_base;
            bool keep = true;
            if (a.sym == PA_SYM_UPPER) {
                keep = C > R;
            } else if (a.sym == PA_SYM_EVEN_ODD) {
                const bool ro = R & 1, co = C & 1;
                keep = !(R == C || (R > C ? (ro ^ co) : !(ro ^ co)));
            }
            a.pkey[t] = keep ? ((c - a.c_begin) << a.rbits) | r : a.keymax;
            a.pval[t] = ((rv & 0xffffull) << 48) | ((cv & 0xffffull) << 32) |
                        (uint64_t)(uint16_t)(int16_t)cost;
            drop += keep ? 0 : 1;
            ++local;
        }
    }
    if (drop)
        atomicAdd(a.ndrop, (unsigned long long)drop);
}

__global__ void k_run_heads(const uint64_t *__restrict__ keys, uint64_t n,
                            uint32_t *__restrict__ heads)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_run_starts(const uint32_t *__restrict__ heads,
                             const uint64_t *__restrict__ incl, uint64_t n,
                             uint64_t *__restrict__ rstart)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (heads[i])
            rstart[incl[i] - 1] = i;
}

// Fold of one run (one candidate pair) and the k-mer threshold prune.
struct FoldArgs {
    const uint64_t *rstart;
    uint64_t nruns, nprod;
    const uint64_t *pkey, *pval;
    int rbits;
    uint64_t c_begin;
    int ckthr;
    float mosthr;
    pa_pair *cand;
    uint32_t *keep;
};

__global__ void k_fold_runs(FoldArgs a)
{
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
         g < a.nruns; g += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t s = a.rstart[g];
        const uint64_t e = g + 1 < a.nruns ? a.rstart[g + 1] : a.nprod;
        const uint64_t key = a.pkey[s];
        const uint64_t v0 = a.pval[s], vl = a.pval[e - 1];
        int score = (int16_t)(uint16_t)(v0 & 0xffffu);
        for (uint64_t i = s + 1; i < e; ++i) {
            const int x = (int16_t)(uint16_t)(a.pval[i] & 0xffffu);
            score = x > score ? x : score;
        }
        pa_pair p;
        p.r = (uint32_t)(key & ((1ull << a.rbits) - 1ull));
        p.c = (uint32_t)((key >> a.rbits) + a.c_begin);
        p.count = (uint16_t)(e - s);
        p.kscore = (int16_t)score;
        p.seed[0] = (uint16_t)(v0 >> 48);
        p.seed[1] = (uint16_t)(v0 >> 32);
        p.seed[2] = e - s > 1 ? (uint16_t)(vl >> 48) : 0;
        p.seed[3] = e - s > 1 ? (uint16_t)(vl >> 32) : 0;
        a.cand[g] = p;
        // DistPWRunner.cpp:48-54: count <= ckthr || score <= mosthr
        const bool prune = (int)p.count <= a.ckthr ||
                           (float)p.kscore <= a.mosthr;
        a.keep[g] = prune ? 0u : 1u;
    }
}

__global__ void k_compact(const pa_pair *__restrict__ cand,
                          const uint32_t *__restrict__ keep,
                          const uint64_t *__restrict__ pos, uint64_t n,
                          pa_pair *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        if (keep[i])
            out[pos[i]] = cand[i];
}

#define KCHK(ctx, x)                                                         \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess)                                                \
            return fail((ctx), PA_ERR_HIP,                                   \
                        std::string(#x) + ": " + hipGetErrorString(e_));     \
    } while (0)

// inclusive scan of u32 flags into u64
int scan_flags(pa_ctx *c, KmerState *K, const uint32_t *in, uint64_t *out,
               uint64_t n, bool inclusive, hipStream_t s)
{
    size_t tmp = 0;
    if (inclusive)
        KCHK(c, hipcub::DeviceScan::InclusiveSum(nullptr, tmp, in, out, n, s));
    else
        KCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s));
    int rc;
    if ((rc = grow(c, K->tmp, tmp)))
        return rc;
    if (inclusive)
        KCHK(c, hipcub::DeviceScan::InclusiveSum(K->tmp.p, tmp, in, out, n, s));
    else
        KCHK(c, hipcub::DeviceScan::ExclusiveSum(K->tmp.p, tmp, in, out, n, s));
    return PA_OK;
}

template <class T> int d2h_value(pa_ctx *c, T *dst, const T *src, hipStream_t s)
{
    KCHK(c, hipMemcpyAsync(dst, src, sizeof(T), hipMemcpyDeviceToHost, s));
    KCHK(c, hipStreamSynchronize(s));
    return PA_OK;
}

// Sorted k-mer occurrences and group table of one loaded set.
int build_set(pa_ctx *c, KmerState *K, KmerSet &S, bool rows,
              const pa_kmer_config &kc, const KmerAlphabet &al, int kbits,
              hipStream_t s)
{
    if (S.gen == c->seq_gen && S.k == kc.k && S.stride == kc.stride &&
        S.alph == kc.alphabet)
        return PA_OK;
    int rc;
    const std::vector<uint32_t> &len = rows ? c->rlen : c->clen;
    const uint64_t nseq = len.size();
    std::vector<uint64_t> slots(nseq + 1, 0);
    for (uint64_t i = 0; i < nseq; ++i) {
        const uint64_t L = len[i];
        const uint64_t np =
            L >= (uint64_t)kc.k ? (L - kc.k) / (uint64_t)kc.stride + 1 : 0;
        slots[i + 1] = slots[i] + np;
    }
    const uint64_t nslots = slots[nseq];
    S.k = kc.k;
    S.stride = kc.stride;
    S.alph = kc.alphabet;
    S.gen = c->seq_gen;
    S.nocc = S.ngroups = 0;
    if (nslots == 0)
        return PA_OK;
    if ((rc = grow(c, K->slots, (nseq + 1) * 8)) ||
        (rc = grow(c, K->keys_a, nslots * 4)) ||
        (rc = grow(c, K->vals_a, nslots * 8)) ||
        (rc = grow(c, S.keys, nslots * 4)) ||
        (rc = grow(c, S.vals, nslots * 8)) ||
        (rc = grow(c, K->cnt, 64)))
        return rc;
    KCHK(c, hipMemcpyAsync(K->slots.p, slots.data(), (nseq + 1) * 8,
                           hipMemcpyHostToDevice, s));
    KCHK(c, hipMemsetAsync(K->cnt.p, 0, 64, s));
    const uint32_t invalid = (uint32_t)((1ull << kbits) - 1ull);
    const Buf &kcb = rows ? c->rkc : c->ckc;
    const Buf &cb = rows ? c->rcodes : c->ccodes;
    hipLaunchKernelGGL(k_kmer_extract, dim3(nblocks(nslots)), dim3(kTB), 0, s,
                       P<uint8_t>(const_cast<Buf &>(kcb)),
                       P<uint64_t>(rows ? c->roffs : c->coffs),
                       (uint64_t)(uintptr_t)cb.p, P<uint64_t>(K->slots), nseq,
                       nslots, kc.k, kc.stride, al, invalid,
                       P<uint32_t>(K->keys_a), P<uint64_t>(K->vals_a),
                       (unsigned long long *)K->cnt.p);
    KCHK(c, hipGetLastError());
    size_t tmp = 0;
    KCHK(c, hipcub::DeviceRadixSort::SortPairs(
                 nullptr, tmp, P<uint32_t>(K->keys_a), P<uint32_t>(S.keys),
                 P<uint64_t>(K->vals_a), P<uint64_t>(S.vals), nslots, 0, kbits,
                 s));
    if ((rc = grow(c, K->tmp, tmp)))
        return rc;
    KCHK(c, hipcub::DeviceRadixSort::SortPairs(
                 K->tmp.p, tmp, P<uint32_t>(K->keys_a), P<uint32_t>(S.keys),
                 P<uint64_t>(K->vals_a), P<uint64_t>(S.vals), nslots, 0, kbits,
                 s));
    uint64_t nbad = 0;
    if ((rc = d2h_value(c, &nbad, P<uint64_t>(K->cnt), s)))
        return rc;
    S.nocc = nslots - nbad;
    if (S.nocc == 0)
        return PA_OK;
    // group table over the valid occurrences
    if ((rc = grow(c, K->heads, S.nocc * 4)) ||
        (rc = grow(c, K->runid, S.nocc * 8)))
        return rc;
    hipLaunchKernelGGL(k_group_heads, dim3(nblocks(S.nocc)), dim3(kTB), 0, s,
                       P<uint32_t>(S.keys), S.nocc, P<uint32_t>(K->heads));
    if ((rc = scan_flags(c, K, P<uint32_t>(K->heads), P<uint64_t>(K->runid),
                         S.nocc, true, s)))
        return rc;
    if ((rc = d2h_value(c, &S.ngroups, P<uint64_t>(K->runid) + S.nocc - 1, s)))
        return rc;
    if ((rc = grow(c, S.gstart, S.ngroups * 8)) ||
        (rc = grow(c, S.gcode, S.ngroups * 4)))
        return rc;
    hipLaunchKernelGGL(k_group_table, dim3(nblocks(S.nocc)), dim3(kTB), 0, s,
                       P<uint32_t>(S.keys), P<uint32_t>(K->heads),
                       P<uint64_t>(K->runid), S.nocc, P<uint64_t>(S.gstart),
                       P<uint32_t>(S.gcode));
    KCHK(c, hipGetLastError());
    return PA_OK;
}

bool same_cfg(const pa_kmer_config &a, const pa_kmer_config &b)
{
    return a.k == b.k && a.stride == b.stride && a.alphabet == b.alphabet &&
           a.symmetric == b.symmetric && a.ckthr == b.ckthr &&
           a.mosthr == b.mosthr && a.row_base == b.row_base &&
           a.col_base == b.col_base;
}

// Runs the join for columns [cb, ce) and leaves the candidates in K->cand
// (compacted), their number in *n.
int generate(pa_ctx *c, KmerState *K, const pa_kmer_config &kc, uint64_t cb,
             uint64_t ce, uint64_t *n, hipStream_t s)
{
    int rc;
    KmerAlphabet al{};
    if (!make_alphabet(kc.alphabet, al))
        return fail(c, PA_ERR_ARG, "unknown k-mer alphabet");
    uint64_t universe = 1;
    for (int j = 0; j < kc.k; ++j)
        universe *= (uint64_t)al.size;
    if (universe >= (1ull << 31))
        return fail(c, PA_ERR_ARG, "alphabet size ^ k must stay below 2^31");
    const int kbits = bits_for(universe);   // 2^kbits - 1 >= universe
    if ((rc = build_set(c, K, K->rows, true, kc, al, kbits, s)))
        return rc;
    const bool shared = c->shared_cols;
    KmerSet &C = shared ? K->rows : K->cols;
    if (!shared && (rc = build_set(c, K, K->cols, false, kc, al, kbits, s)))
        return rc;
    KmerSet &R = K->rows;
    *n = 0;
    if (C.ngroups == 0 || R.nocc == 0 || cb >= ce)
        return PA_OK;
    const uint64_t G = C.ngroups;
    if ((rc = grow(c, K->pcount, (G + 1) * 8)) ||
        (rc = grow(c, K->poff, (G + 1) * 8)) ||
        (rc = grow(c, K->gscr, G * 24)))
        return rc;
    // per group: first column / row occurrence (u64), their counts (u32)
    uint64_t *pcount = P<uint64_t>(K->pcount);
    uint64_t *ca = P<uint64_t>(K->gscr), *ra = ca + G;
    uint32_t *cn = (uint32_t *)(ra + G), *rn = cn + G;
    JoinArgs ja{P<uint64_t>(C.gstart), P<uint32_t>(C.gcode), P<uint64_t>(C.vals),
                G, C.nocc, P<uint32_t>(R.keys), R.nocc, P<uint64_t>(R.gstart),
                shared, cb, ce, ca, ra, cn, rn, pcount};
    hipLaunchKernelGGL(k_join_count, dim3(nblocks(G)), dim3(kTB), 0, s, ja);
    KCHK(c, hipGetLastError());
    {
        size_t tmp = 0;
        KCHK(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, pcount,
                                                 P<uint64_t>(K->poff), G + 1,
                                                 s));
        if ((rc = grow(c, K->tmp, tmp)))
            return rc;
        // pcount[G] = 0 so that poff[G] is the total
        KCHK(c, hipMemsetAsync(pcount + G, 0, 8, s));
        KCHK(c, hipcub::DeviceScan::ExclusiveSum(K->tmp.p, tmp, pcount,
                                                 P<uint64_t>(K->poff), G + 1,
                                                 s));
    }
    uint64_t nprod = 0;
    if ((rc = d2h_value(c, &nprod, P<uint64_t>(K->poff) + G, s)))
        return rc;
    if (nprod == 0)
        return PA_OK;
    const uint64_t ncols_blk = ce - cb;
    const uint64_t nrows = c->rlen.size();
    const int rbits = bits_for(nrows);
    const int cbits = bits_for(ncols_blk);
    if (rbits + cbits > 63)
        return fail(c, PA_ERR_LIMIT, "row x column block too large");
    const uint64_t keymax = (1ull << (rbits + cbits)) - 1ull;   // unused key
    if ((rc = grow(c, K->pkey, nprod * 8)) || (rc = grow(c, K->pval, nprod * 8)) ||
        (rc = grow(c, K->pkey2, nprod * 8)) ||
        (rc = grow(c, K->pval2, nprod * 8)) || (rc = grow(c, K->cnt, 64)))
        return rc;
    KCHK(c, hipMemsetAsync(K->cnt.p, 0, 64, s));
    EmitArgs ea{P<uint64_t>(K->poff), G, nprod, ca, ra, cn, rn,
                P<uint32_t>(C.gcode), P<uint64_t>(C.vals), P<uint64_t>(R.vals),
                al, kc.k, kc.symmetric, kc.row_base, kc.col_base, cb, rbits,
                keymax, P<uint64_t>(K->pkey), P<uint64_t>(K->pval),
                (unsigned long long *)K->cnt.p};
    hipLaunchKernelGGL(k_emit_products,
                       dim3(nblocks((nprod + kPerThread - 1) / kPerThread)),
                       dim3(kTB), 0, s, ea);
    KCHK(c, hipGetLastError());
    {
        size_t tmp = 0;
        KCHK(c, hipcub::DeviceRadixSort::SortPairs(
                     nullptr, tmp, P<uint64_t>(K->pkey), P<uint64_t>(K->pkey2),
                     P<uint64_t>(K->pval), P<uint64_t>(K->pval2), nprod, 0,
                     rbits + cbits, s));
        if ((rc = grow(c, K->tmp, tmp)))
            return rc;
        KCHK(c, hipcub::DeviceRadixSort::SortPairs(
                     K->tmp.p, tmp, P<uint64_t>(K->pkey), P<uint64_t>(K->pkey2),
                     P<uint64_t>(K->pval), P<uint64_t>(K->pval2), nprod, 0,
                     rbits + cbits, s));
    }
    uint64_t ndrop = 0;
    if ((rc = d2h_value(c, &ndrop, P<uint64_t>(K->cnt), s)))
        return rc;
    const uint64_t np = nprod - ndrop;   // dropped products sort last
    if (np == 0)
        return PA_OK;
    if ((rc = grow(c, K->heads, np * 4)) || (rc = grow(c, K->runid, np * 8)))
        return rc;
    hipLaunchKernelGGL(k_run_heads, dim3(nblocks(np)), dim3(kTB), 0, s,
                       P<uint64_t>(K->pkey2), np, P<uint32_t>(K->heads));
    if ((rc = scan_flags(c, K, P<uint32_t>(K->heads), P<uint64_t>(K->runid),
                         np, true, s)))
        return rc;
    uint64_t nruns = 0;
    if ((rc = d2h_value(c, &nruns, P<uint64_t>(K->runid) + np - 1, s)))
        return rc;
    if ((rc = grow(c, K->rstart, nruns * 8)) ||
        (rc = grow(c, K->keep, nruns * 4 + 8)) ||
        (rc = grow(c, K->kpos, (nruns + 1) * 8)) ||
        (rc = grow(c, K->cand, nruns * sizeof(pa_pair))))
        return rc;
    hipLaunchKernelGGL(k_run_starts, dim3(nblocks(np)), dim3(kTB), 0, s,
                       P<uint32_t>(K->heads), P<uint64_t>(K->runid), np,
                       P<uint64_t>(K->rstart));
    // folded pairs go to pkey (free now) before the compaction
    if ((rc = grow(c, K->pkey, nruns * sizeof(pa_pair))))
        return rc;
    FoldArgs fa{P<uint64_t>(K->rstart), nruns, np, P<uint64_t>(K->pkey2),
                P<uint64_t>(K->pval2), rbits, cb, kc.ckthr, kc.mosthr,
                P<pa_pair>(K->pkey), P<uint32_t>(K->keep)};
    hipLaunchKernelGGL(k_fold_runs, dim3(nblocks(nruns)), dim3(kTB), 0, s, fa);
    KCHK(c, hipGetLastError());
    if ((rc = scan_flags(c, K, P<uint32_t>(K->keep), P<uint64_t>(K->kpos),
                         nruns, false, s)))
        return rc;
    uint64_t last = 0;
    uint32_t lastkeep = 0;
    if ((rc = d2h_value(c, &last, P<uint64_t>(K->kpos) + nruns - 1, s)) ||
        (rc = d2h_value(c, &lastkeep, P<uint32_t>(K->keep) + nruns - 1, s)))
        return rc;
    const uint64_t nout = last + lastkeep;
    hipLaunchKernelGGL(k_compact, dim3(nblocks(nruns)), dim3(kTB), 0, s,
                       P<pa_pair>(K->pkey), P<uint32_t>(K->keep),
                       P<uint64_t>(K->kpos), nruns, P<pa_pair>(K->cand));
    KCHK(c, hipGetLastError());
    KCHK(c, hipStreamSynchronize(s));
    *n = nout;
    return PA_OK;
}

}  // namespace

extern "C" int pa_kmer_candidates(pa_ctx *c, const pa_kmer_config *kc,
                                  uint64_t c_begin, uint64_t c_end,
                                  pa_pair *d_out, uint64_t cap,
                                  uint64_t *n_out)
{
    if (!c || !kc || !n_out)
        return PA_ERR_ARG;
    *n_out = 0;
    if (!c->have_rows || !c->have_cols)
        return fail(c, PA_ERR_NOSEQ, "load row and column sequences first");
    if (kc->k < 1 || kc->stride < 1 || kc->symmetric < PA_SYM_NONE ||
        kc->symmetric > PA_SYM_EVEN_ODD)
        return fail(c, PA_ERR_ARG, "bad k-mer configuration");
    if (c_end > c->clen.size())
        c_end = c->clen.size();
    if (c_begin > c_end)
        c_begin = c_end;
    if (c->rlen.size() > 0xffffffffull || c->clen.size() > 0xffffffffull)
        return fail(c, PA_ERR_LIMIT, "more than 2^32 sequences");
    hipSetDevice(c->device);
    if (!c->kmer)
        c->kmer = new KmerState();
    KmerState *K = c->kmer;
    hipStream_t s = c->st;
    int rc;
    if (!(K->have_last && same_cfg(K->last, *kc) && K->last_cb == c_begin &&
          K->last_ce == c_end && K->last_gen == c->seq_gen)) {
        K->have_last = false;
        uint64_t n = 0;
        if ((rc = generate(c, K, *kc, c_begin, c_end, &n, s)))
            return rc;
        K->last = *kc;
        K->last_cb = c_begin;
        K->last_ce = c_end;
        K->last_n = n;
        K->last_gen = c->seq_gen;
        K->have_last = true;
    }
    *n_out = K->last_n;
    if (!d_out)
        return PA_OK;
    if (K->last_n > cap)
        return fail(c, PA_ERR_LIMIT,
                    std::to_string(K->last_n) + " candidates exceed the "
                    "output capacity " + std::to_string(cap));
    if (K->last_n)
        KCHK(c, hipMemcpyAsync(d_out, K->cand.p, K->last_n * sizeof(pa_pair),
                               hipMemcpyDeviceToDevice, s));
    KCHK(c, hipStreamSynchronize(s));
    return PA_OK;
}
