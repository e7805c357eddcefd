// k_misc.hip -- task construction, binning and result kernels.
#include "k_misc.h"

namespace pa {
namespace {

// Adds v summed over the block to *dst with one atomic per block (many
// same-address atomics from every wavefront serialise).  Every thread of the
// block calls it; blockDim.x is a multiple of 64, at most 1024.
__device__ inline void block_atomic_add(unsigned long long *dst, uint64_t v)
{
    __shared__ uint64_t part[16];
    for (int off = 32; off >= 1; off >>= 1)
        v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t sum = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
            sum += part[w];
        if (sum)
            atomicAdd(dst, (unsigned long long)sum);
    }
    __syncthreads();
}

// grid of a grid-stride loop over n items: at most 2048 blocks of 256
inline int stride_blocks(uint64_t n)
{
    const uint64_t b = (n + 255) / 256;
    return (int)(b < 2048 ? (b > 0 ? b : 1) : 2048);
}

__global__ void k_accumulate(const uint32_t *src, unsigned long long *dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && *src)
        atomicAdd(dst, (unsigned long long)*src);
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Pair -> local-alignment task: rows = row sequence (SeqAn V), columns =
// column sequence (SeqAn H), as localAlignment(seqsc, seqsr) at
// SeqanFullAligner.cpp:231.
// A pair naming a sequence outside the loaded sets (nr rows, nc columns)
// becomes an empty task: its result is the empty alignment and no kernel
// reads residues for it.
__global__ void k_build_sw_tasks(const pa_pair *__restrict__ pairs, uint64_t n,
                                 const uint64_t *__restrict__ roffs,
                                 const uint64_t *__restrict__ coffs,
                                 uint64_t nr, uint64_t nc,
                                 DpTask *__restrict__ tasks,
                                 unsigned long long *cells)
{
    uint64_t c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const pa_pair p = pairs[i];
        const bool ok = p.r < nr && p.c < nc;
        const uint32_t r = ok ? p.r : 0, cc = ok ? p.c : 0;
        DpTask t;
        t.q_off = roffs[r];
        t.qlen = ok ? (uint32_t)(roffs[r + 1] - roffs[r]) : 0u;
        t.t_off = coffs[cc];
        t.tlen = ok ? (uint32_t)(coffs[cc + 1] - coffs[cc]) : 0u;
        t.id = (uint32_t)i;
        t.flags = 0;
        tasks[i] = t;
        c += (uint64_t)t.qlen * t.tlen;
    }
    block_atomic_add(cells, c);
}

// Pair x seed -> extension task (TSeed(col offset, row offset, k),
// SeqanXdropAligner.cpp:273-280).  A pair naming a sequence outside the
// loaded sets, or a seed that does not fit inside its sequences, gets
// kXdInvalid: the extension kernels then write empty extents without reading
// residues, so its infix alignment is empty and it does not pass.
__global__ void k_build_xd_tasks(const pa_pair *__restrict__ pairs, uint64_t n,
                                 int seed_cnt, int seed_len,
                                 const uint64_t *__restrict__ roffs,
                                 const uint64_t *__restrict__ coffs,
                                 uint64_t nr, uint64_t nc,
                                 XdTask *__restrict__ tasks)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n * (uint64_t)seed_cnt)
        return;
    const uint64_t i = k / seed_cnt;
    const int s = (int)(k % seed_cnt);
    const pa_pair p = pairs[i];
    const bool ok = p.r < nr && p.c < nc;
    const uint32_t r = ok ? p.r : 0, c = ok ? p.c : 0;
    XdTask t;
    t.v_off = roffs[r];
    t.nv = ok ? (uint32_t)(roffs[r + 1] - roffs[r]) : 0u;
    t.h_off = coffs[c];
    t.nh = ok ? (uint32_t)(coffs[c + 1] - coffs[c]) : 0u;
    t.beg_v = s == 0 ? p.seed[0] : p.seed[2];
    t.beg_h = s == 0 ? p.seed[1] : p.seed[3];
    t.id = (uint32_t)(i * 2 + s);
    const bool fits = ok && t.beg_v + (uint32_t)seed_len <= t.nv &&
                      t.beg_h + (uint32_t)seed_len <= t.nh;
    t.pad = fits ? 0u : kXdInvalid;
    tasks[k] = t;
}

// Extension -> global alignment of the extended infixes
// (globalAlignment(seqsc_ex, seqsr_ex), SeqanXdropAligner.cpp:284-296).
__global__ void k_build_nw_tasks(const XdTask *__restrict__ xt, uint64_t nt,
                                 const XdOut *__restrict__ xo,
                                 DpTask *__restrict__ tasks, ResBound rows,
                                 ResBound cols, unsigned long long *cells,
                                 unsigned long long *xcells)
{
    uint64_t c = 0, xc = 0;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nt;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const XdTask x = xt[k];
        const XdOut o = xo[x.id];
        DpTask t;
        t.q_off = x.v_off + (uint64_t)o.beg_v;
        t.qlen = (uint32_t)(o.end_v - o.beg_v);
        t.t_off = x.h_off + (uint64_t)o.beg_h;
        t.tlen = (uint32_t)(o.end_h - o.beg_h);
        t.id = x.id;
        t.flags = 0;
        if (t.qlen && t.tlen) {
            // residue bound of the two spans (traced int16 holds 4*score+3)
            const uint64_t iq = t.q_off - rows.base, it = t.t_off - cols.base;
            const uint32_t uq = rows.pre[iq + t.qlen] - rows.pre[iq];
            const uint32_t ut = cols.pre[it + t.tlen] - cols.pre[it];
            const uint32_t ub = uq < ut ? uq : ut;
            if (4ull * ub + 3ull < 32760ull)
                t.flags |= kTaskBounded;
        }
        tasks[k] = t;
        c += (uint64_t)t.qlen * t.tlen;
        xc += (uint64_t)o.cells_left + o.cells_right;
    }
    block_atomic_add(cells, c);
    block_atomic_add(xcells, xc);
}

__device__ inline bool global_i16_ok(uint32_t q, uint32_t t, uint32_t flags,
                                     int maxs, int open, int ext)
{
    // traced int16 kernels hold 4*value + tag in [-32000, 32767]: the
    // largest score is at most maxs * min(q, t) and at most the residue
    // bound (kTaskBounded); the smallest is the all-gap path's
    const int gmax = open > ext ? open : ext;
    const int mn = q < t ? (int)q : (int)t;
    const long ub = (long)maxs * mn;
    const long lb = 2L * open + (long)(q + t) * gmax + (ext < open ? ext : open);
    return (4 * ub + 3 < 32760 || (flags & kTaskBounded)) && 4 * lb > -31000;
}

// Wavefront steps of a task in its class's kernel, all passes.
__device__ inline uint32_t task_steps(uint32_t qlen, uint32_t tlen,
                                      const DpShapes &sh)
{
    const int R = rows_per_lane(qlen, sh), L = lanes_per_task(qlen, sh);
    const uint32_t rows = 2u * R * L;
    const uint32_t npass = (qlen + rows - 1) / rows;
    const uint32_t nl = trace_last_lanes(qlen, R, L);
    return (npass - 1) * (tlen + 2u * L - 1u) + tlen + 2u * (nl - 1u) + 1u;
}

// Bin and bucket of a task (-1: empty task).  Buckets are 1/8 of an octave
// of the step count wide (16 .. 2^23 steps), the longest in bucket 0.
__device__ inline int bin_key(const DpTask &t, const DpShapes &sh, int local,
                              int maxs, int open, int ext)
{
    if (t.qlen == 0 || t.tlen == 0)
        return -1;
    const int rc = dp_class(t.qlen, sh);
    const bool i16 =
        local ? true
              : global_i16_ok(t.qlen, t.tlen, t.flags, maxs, open, ext);
    const int bin = 2 * rc + (i16 ? 0 : 1);
    // 8 * log2(steps / 16) from the leading bit and the next three
    const uint32_t x = task_steps(t.qlen, t.tlen, sh) >> 4;
    int b = 0;
    if (x) {
        const int e = 31 - __builtin_clz(x);
        const int f = (int)((e >= 3 ? x >> (e - 3) : x << (3 - e)) & 7u);
        b = 8 * e + f;
    }
    const int bucket = kBuckets - 1 - (b < kBuckets - 1 ? b : kBuckets - 1);
    return bin * kBuckets + bucket;
}

// Pass 1: histogram of the (bin, bucket) keys: LDS counters per block, then
// one global atomic per key that occurs in the block.
__global__ __launch_bounds__(256) void k_bin_hist(
    const DpTask *__restrict__ tasks, uint32_t a, uint32_t b,
    const uint32_t *rng, DpShapes sh, int local, int maxs, int open, int ext,
    uint32_t *hist)
{
    if (rng) {   // task range from the device (chunks sized on the GPU)
        a = rng[0];
        b = rng[1];
    }
    __shared__ uint32_t h[kBins * kBuckets];
    for (int k = threadIdx.x; k < kBins * kBuckets; k += blockDim.x)
        h[k] = 0;
    __syncthreads();
    const uint32_t i = a + blockIdx.x * blockDim.x + threadIdx.x;
    const int key = i < b ? bin_key(tasks[i], sh, local, maxs, open, ext) : -1;
    if (key >= 0)
        atomicAdd(&h[key], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < kBins * kBuckets; k += blockDim.x)
        if (h[k])
            atomicAdd(&hist[k], h[k]);
}

// Pass 2 (one block of 64 threads): fill = exclusive scan of the buckets of
// each bin, counts = bin totals.
__global__ void k_bin_scan(const uint32_t *hist, uint32_t *fill,
                           uint32_t *counts)
{
    const int bin = threadIdx.x;
    if (bin >= kBins)
        return;
    uint32_t run = 0;
    for (int k = 0; k < kBuckets; ++k) {
        fill[bin * kBuckets + k] = run;
        run += hist[bin * kBuckets + k];
    }
    counts[bin] = run;
}

// Pass 3: each task goes to its bin's list at its bucket's next slot.  The
// block counts its keys in LDS (the LDS atomic's old value is the task's
// rank), reserves one range per key with a global atomic, and scatters.
__global__ __launch_bounds__(256) void k_bin(
    const DpTask *__restrict__ tasks, uint32_t a, uint32_t b,
    const uint32_t *rng, DpShapes sh, int local, int maxs, int open, int ext,
    uint32_t *__restrict__ lists, uint32_t cap, uint32_t *fill,
    unsigned long long *bin_cells)
{
    if (rng) {
        a = rng[0];
        b = rng[1];
    }
    __shared__ uint32_t h[kBins * kBuckets], base[kBins * kBuckets];
    __shared__ unsigned long long cc[kBins];
    for (int k = threadIdx.x; k < kBins * kBuckets; k += blockDim.x)
        h[k] = 0;
    if (threadIdx.x < kBins)
        cc[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t i = a + blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    uint64_t cells = 0;
    if (i < b) {
        const DpTask t = tasks[i];
        key = bin_key(t, sh, local, maxs, open, ext);
        cells = (uint64_t)t.qlen * t.tlen;
    }
    uint32_t rank = 0;
    if (key >= 0) {
        rank = atomicAdd(&h[key], 1u);
        atomicAdd(&cc[key / kBuckets], (unsigned long long)cells);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kBins * kBuckets; k += blockDim.x)
        base[k] = h[k] ? atomicAdd(&fill[k], h[k]) : 0u;
    if (threadIdx.x < kBins && cc[threadIdx.x])
        atomicAdd(&bin_cells[threadIdx.x], cc[threadIdx.x]);
    __syncthreads();
    if (key >= 0)
        lists[(size_t)(key / kBuckets) * cap + base[key] + rank] = i;
}

__global__ void k_trace_sizes(const DpTask *__restrict__ tasks, uint64_t n,
                              DpShapes sh, uint64_t *__restrict__ sizes)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const DpTask t = tasks[i];
        sizes[i] = (t.qlen && t.tlen)
                       ? trace_bytes(t.qlen, t.tlen, rows_per_lane(t.qlen, sh),
                                     lanes_per_task(t.qlen, sh))
                       : 0;
    }
}

// Empty tasks (a zero-length span) never reach a DP kernel.
__global__ void k_empty_out(const DpTask *__restrict__ tasks, uint32_t a,
                            uint32_t b, const uint32_t *rng, int local,
                            int open, int ext, DpOut *out)
{
    if (rng) {
        a = rng[0];
        b = rng[1];
    }
    const uint32_t i = a + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b)
        return;
    const DpTask t = tasks[i];
    if (t.qlen && t.tlen)
        return;
    DpOut o = {};
    if (!local) {
        const int L = (int)(t.qlen + t.tlen);
        const int gmax = open > ext ? open : ext;
        o.score = L ? open + (L - 1) * gmax : 0;
        o.q_end = (int)t.qlen;
        o.t_end = (int)t.tlen;
        o.cols = L;
        o.gap_opens = L ? 1 : 0;
    }
    out[i] = o;
}

__global__ void k_mark(const uint32_t *list, const uint32_t *count,
                       uint8_t *flags, uint8_t bit)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < *count)
        flags[list[k]] |= bit;
}

__device__ inline float identity_of(uint32_t matches, uint32_t cols)
{
    return (float)(100.0 * (double)(float)matches / (double)(float)cols);
}

// SeqanFullAligner.cpp:247-261 / SeqanXdropAligner.cpp:338-351
__device__ inline bool predicate(uint32_t cols, uint32_t gap_opens,
                                 uint32_t matches, int len_c, int len_r,
                                 double cov_thr, int ani_thr, float *ident)
{
    if (cols == 0) {
        *ident = __int_as_float(0x7fc00000);
        return false;
    }
    const float id = identity_of(matches, cols);
    *ident = id;
    const double amg = (double)((int)cols - (int)gap_opens);
    const double c1 = amg / len_c, c2 = amg / len_r;
    const double cov = c1 > c2 ? c1 : c2;
    return cov >= cov_thr && id >= (float)ani_thr;
}

__global__ void k_finalize_sw(const DpTask *__restrict__ tasks,
                              const DpOut *__restrict__ dp,
                              const uint8_t *__restrict__ tflags, uint64_t n,
                              double cov_thr, int ani_thr,
                              const uint32_t *err, pa_result *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (err && *err) {   // the trace did not fit the captured chunks
        pa_result r = {};
        r.flags = PA_FLAG_NOTRACE;
        out[i] = r;
        return;
    }
    const DpTask t = tasks[i];
    const DpOut o = dp[i];
    pa_result r = {};
    r.score = o.score;
    r.r_beg = (uint16_t)o.q_beg;
    r.r_end = (uint16_t)o.q_end;
    r.c_beg = (uint16_t)o.t_beg;
    r.c_end = (uint16_t)o.t_end;
    r.matches = (uint32_t)o.matches;
    r.cols = (uint32_t)o.cols;
    r.gap_opens = (uint32_t)o.gap_opens;
    float id;
    const bool pass = predicate(r.cols, r.gap_opens, r.matches, (int)t.tlen,
                                (int)t.qlen, cov_thr, ani_thr, &id);
    r.identity = id;
    r.score_aln = pass ? id / 100.0f : 0.0f;
    r.flags = (uint16_t)((pass ? PA_FLAG_PASS : 0) |
                         ((tflags[i] & 1) ? PA_FLAG_RERUN32 : 0) |
                         (o.score == 0 ? PA_FLAG_EMPTY : 0));
    r.seed = 0;
    out[i] = r;
}

// Best of the seeds by identity, strict '>' (SeqanXdropAligner.cpp:303-323),
// predicate on the full sequence lengths (:338-351).
__global__ void k_finalize_xd(const XdTask *__restrict__ xt,
                              const XdOut *__restrict__ xo,
                              const DpOut *__restrict__ dp,
                              const uint8_t *__restrict__ tflags, uint64_t n,
                              int seed_cnt, double cov_thr, int ani_thr,
                              const uint32_t *err, pa_result *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (err && *err) {   // the trace did not fit the captured chunks
        pa_result r = {};
        r.flags = PA_FLAG_NOTRACE;
        out[i] = r;
        return;
    }
    int best = 0;
    float best_id = 0.0f;
    for (int s = 0; s < seed_cnt; ++s) {
        const DpOut o = dp[i * seed_cnt + s];
        const float id = o.cols ? identity_of((uint32_t)o.matches,
                                              (uint32_t)o.cols)
                                : __int_as_float(0x7fc00000);
        if (s == 0 || id > best_id) {
            best = s;
            best_id = id;
        }
    }
    const uint64_t k = i * seed_cnt + best;
    const DpOut o = dp[k];
    const XdTask x = xt[i * seed_cnt + best];
    const XdOut e = xo[i * 2 + best];
    pa_result r = {};
    r.score = o.score;
    r.r_beg = (uint16_t)e.beg_v;
    r.r_end = (uint16_t)e.end_v;
    r.c_beg = (uint16_t)e.beg_h;
    r.c_end = (uint16_t)e.end_h;
    r.matches = (uint32_t)o.matches;
    r.cols = (uint32_t)o.cols;
    r.gap_opens = (uint32_t)o.gap_opens;
    float id;
    const bool pass = predicate(r.cols, r.gap_opens, r.matches, (int)x.nh,
                                (int)x.nv, cov_thr, ani_thr, &id);
    r.identity = id;
    r.score_aln = pass ? id / 100.0f : 0.0f;
    r.flags = (uint16_t)((pass ? PA_FLAG_PASS : 0) |
                         ((tflags[k] & 1) ? PA_FLAG_RERUN32 : 0));
    r.seed = (uint16_t)best;
    out[i] = r;
}

__global__ void k_scores_out(const DpOut *__restrict__ dp, uint64_t n,
                             pa_score *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const DpOut o = dp[i];
    pa_score s;
    s.score = o.score;
    s.r_end = (uint16_t)o.q_end;
    s.c_end = (uint16_t)o.t_end;
    out[i] = s;
}

__global__ void k_extents_out(const XdOut *__restrict__ xo, uint64_t n,
                              int seed, pa_extent *__restrict__ out)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const XdOut e = xo[i * 2 + seed];
    pa_extent x;
    x.r_beg = (uint16_t)e.beg_v;
    x.r_end = (uint16_t)e.end_v;
    x.c_beg = (uint16_t)e.beg_h;
    x.c_end = (uint16_t)e.end_h;
    out[i] = x;
}

inline unsigned blocks(uint64_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

hipError_t fill_u32(uint32_t *p, uint32_t v, uint32_t n, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(n, 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t accumulate(const uint32_t *src, unsigned long long *dst,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_accumulate, dim3(1), dim3(64), 0, s, src, dst);
    return hipGetLastError();
}

hipError_t build_sw_tasks(const pa_pair *pairs, uint64_t n,
                          const uint64_t *roffs, const uint64_t *coffs,
                          uint64_t nr, uint64_t nc, DpTask *tasks,
                          unsigned long long *cells, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_build_sw_tasks, dim3(stride_blocks(n)), dim3(256), 0, s,
                       pairs, n, roffs, coffs, nr, nc, tasks, cells);
    return hipGetLastError();
}

hipError_t build_xd_tasks(const pa_pair *pairs, uint64_t n, int seed_cnt,
                          int seed_len, const uint64_t *roffs,
                          const uint64_t *coffs, uint64_t nr, uint64_t nc,
                          XdTask *tasks, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_build_xd_tasks, dim3(blocks(n * seed_cnt, 256)),
                       dim3(256), 0, s, pairs, n, seed_cnt, seed_len, roffs,
                       coffs, nr, nc, tasks);
    return hipGetLastError();
}

hipError_t build_nw_tasks(const XdTask *xt, uint64_t nt, const XdOut *xo,
                          DpTask *tasks, ResBound rows, ResBound cols,
                          unsigned long long *cells,
                          unsigned long long *xcells, hipStream_t s)
{
    if (!nt)
        return hipSuccess;
    hipLaunchKernelGGL(k_build_nw_tasks, dim3(stride_blocks(nt)), dim3(256), 0, s,
                       xt, nt, xo, tasks, rows, cols, cells, xcells);
    return hipGetLastError();
}

hipError_t bin_tasks(const DpTask *tasks, uint32_t a, uint32_t b,
                     const DpShapes &sh, int local, int maxs, int open,
                     int ext, uint32_t *lists,
                     uint32_t cap, uint32_t *counts, uint32_t *hist,
                     uint32_t *fill, unsigned long long *bin_cells,
                     hipStream_t s, const uint32_t *rng)
{
    if (b <= a)
        return hipSuccess;
    hipLaunchKernelGGL(k_bin_hist, dim3(blocks(b - a, 256)), dim3(256), 0, s,
                       tasks, a, b, rng, sh, local, maxs, open, ext, hist);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(64), 0, s, hist, fill,
                       counts);
    hipLaunchKernelGGL(k_bin, dim3(blocks(b - a, 256)), dim3(256), 0, s, tasks,
                       a, b, rng, sh, local, maxs, open, ext, lists, cap, fill,
                       bin_cells);
    return hipGetLastError();
}

hipError_t trace_sizes(const DpTask *tasks, uint64_t n, const DpShapes &sh,
                       uint64_t *sizes, hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_trace_sizes, dim3(blocks(n, 256)), dim3(256), 0, s,
                       tasks, n, sh, sizes);
    return hipGetLastError();
}

hipError_t empty_out(const DpTask *tasks, uint32_t a, uint32_t b, int local,
                     int open, int ext, DpOut *out, hipStream_t s,
                     const uint32_t *rng)
{
    if (b <= a)
        return hipSuccess;
    hipLaunchKernelGGL(k_empty_out, dim3(blocks(b - a, 256)), dim3(256), 0, s,
                       tasks, a, b, rng, local, open, ext, out);
    return hipGetLastError();
}

// Chunks of tasks [0, n) whose trace codes fit in `cap` bytes, from the
// exclusive scan toffs (n + 1 entries): chunk k is tasks [a_k, b_k) with the
// largest b_k such that toffs[b_k] - toffs[a_k] <= cap, and its base offset
// toffs[a_k].  Chunks past the end are empty.  *err is set when a single
// task needs more than `cap` or the tasks need more than nchunks chunks.
__global__ void k_trace_chunks(const uint64_t *__restrict__ toffs, uint32_t n,
                               uint64_t cap, int nchunks, uint32_t *ranges,
                               uint64_t *bases, uint32_t *err)
{
    if (blockIdx.x != 0 || threadIdx.x != 0)
        return;
    uint32_t a = 0;
    for (int k = 0; k < nchunks; ++k) {
        uint32_t b = a;
        if (a < n) {
            uint32_t lo = a, hi = n;   // largest b in [a, n] that fits
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo + 1) / 2;
                if (toffs[mid] - toffs[a] <= cap)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            b = lo;
            if (b == a) {   // one task larger than the store
                *err = 1u;
                b = a = n;
            }
        }
        ranges[2 * k] = a;
        ranges[2 * k + 1] = b;
        bases[k] = toffs[a];
        a = b;
    }
    if (a < n)
        *err = 1u;
}

hipError_t trace_chunks(const uint64_t *toffs, uint32_t n, uint64_t cap,
                        int nchunks, uint32_t *ranges, uint64_t *bases,
                        uint32_t *err, hipStream_t s)
{
    hipLaunchKernelGGL(k_trace_chunks, dim3(1), dim3(64), 0, s, toffs, n, cap,
                       nchunks, ranges, bases, err);
    return hipGetLastError();
}

hipError_t mark_flags(const uint32_t *list, const uint32_t *count,
                      uint32_t max_n, uint8_t *flags, uint8_t bit,
                      hipStream_t s)
{
    if (!max_n)
        return hipSuccess;
    hipLaunchKernelGGL(k_mark, dim3(blocks(max_n, 256)), dim3(256), 0, s, list,
                       count, flags, bit);
    return hipGetLastError();
}

hipError_t finalize_sw(const DpTask *tasks, const DpOut *dp,
                       const uint8_t *tflags, uint64_t n, double cov_thr,
                       int ani_thr, pa_result *out, hipStream_t s,
                       const uint32_t *err)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_finalize_sw, dim3(blocks(n, 256)), dim3(256), 0, s,
                       tasks, dp, tflags, n, cov_thr, ani_thr, err, out);
    return hipGetLastError();
}

hipError_t finalize_xd(const XdTask *xt, const XdOut *xo, const DpOut *dp,
                       const uint8_t *tflags,
                       uint64_t n, int seed_cnt, double cov_thr, int ani_thr,
                       pa_result *out, hipStream_t s, const uint32_t *err)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_finalize_xd, dim3(blocks(n, 256)), dim3(256), 0, s,
                       xt, xo, dp, tflags, n, seed_cnt, cov_thr,
                       ani_thr, err, out);
    return hipGetLastError();
}

hipError_t scores_out(const DpOut *dp, uint64_t n, pa_score *out,
                      hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_scores_out, dim3(blocks(n, 256)), dim3(256), 0, s, dp,
                       n, out);
    return hipGetLastError();
}

hipError_t extents_out(const XdOut *xo, uint64_t n, int seed, pa_extent *out,
                       hipStream_t s)
{
    if (!n)
        return hipSuccess;
    hipLaunchKernelGGL(k_extents_out, dim3(blocks(n, 256)), dim3(256), 0, s, xo,
                       n, seed, out);
    return hipGetLastError();
}

}  // namespace pa
