// k_xdrop_coop.hip -- gapped X-drop extension, one work item per 16-lane row
// of a wavefront (four items per wave).
//
// Same algorithm and results as xd_dir in k_xdrop.hip (the restatement of
// SeqAn's GappedXDrop one-direction extension used at
// src/align/SeqanXdropAligner.cpp:281-283), mapped onto lanes instead of
// loops:
//   - the cells written on anti-diagonal d are the columns [off, maxCol]
//     (off = minCol - 1); column c lives in lane c mod 16 of the item's row,
//     so as long as that range is at most 16 columns wide every column has
//     its own lane and the three live anti-diagonals are three registers;
//   - the left neighbour (column c-1) is lane c-1 mod 16: one DPP row_ror:1
//     move; the anti-diagonal maximum is four DPP rotations;
//   - the new band limits (leading / trailing undefined columns) come from a
//     ballot of the per-column conditions, read in column order.
// An item whose written range would exceed 16 columns stops and goes to the
// overflow list; the per-lane kernels of k_xdrop.hip recompute it from
// scratch.  A row whose item ends takes the next item of the queue at once,
// so the four rows of a wave stay busy whatever the item lengths.
#include <limits.h>

#include "pa_internal.h"

namespace pa {
namespace {

constexpr int32_t kUndef = INT_MIN / 4;
constexpr int G = 16;   // lanes per item
constexpr uint32_t kChunk = 32;   // items claimed per queue atomic

// lane (j - k) mod 16 of the same row
template <int K> __device__ inline int ror(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x120 + K, 0xf, 0xf, false);
}

__device__ inline int row_max(int v)
{
    v = max(v, ror<1>(v));
    v = max(v, ror<2>(v));
    v = max(v, ror<4>(v));
    v = max(v, ror<8>(v));
    return v;
}

__device__ inline int row_min(int v)
{
    v = min(v, ror<1>(v));
    v = min(v, ror<2>(v));
    v = min(v, ror<4>(v));
    v = min(v, ror<8>(v));
    return v;
}

// 16 bits of a ballot belonging to this lane's row
__device__ inline uint32_t row_bits(uint64_t b, int row)
{
    return (uint32_t)(b >> (16 * row)) & 0xffffu;
}

// bit k of the result is bit (k + r) mod 16 of m
__device__ inline uint32_t rot16(uint32_t m, int r)
{
    r &= 15;
    return ((m >> r) | (m << (16 - r))) & 0xffffu;
}

struct CoopArgs {
    const XdTask *tasks;
    const uint32_t *list;   // work items k = 2 * task + dir, NULL = 0..count
    const uint32_t *count;
    uint32_t *queue;
    const int8_t *mat;
    int gap, X, seed_len;
    XdOut *out;
    uint32_t *ovf_list, *ovf_count;
};

enum : int { kNeed = 0, kRun = 1, kDone = 2 };

__global__ __launch_bounds__(256) void k_xdrop_coop(CoopArgs a)
{
    __shared__ int8_t mat[32 * 32];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x)
        mat[i] = a.mat[i];
    __syncthreads();
    const uint32_t n = *a.count;
    const int lane = threadIdx.x & 63;
    const int j = lane & 15;
    const int row = lane >> 4;
    const int gap = a.gap, X = a.X;

    // row-uniform state of the current item
    int state = kNeed;
    uint32_t k = 0;
    const uint8_t PA_GLOBAL *qs = nullptr, *ds = nullptr;
    int cols = 0, rows = 0, d = 0, minCol = 0, maxCol = 0, best = 0;
    int off1 = 0, off2 = 0, off3 = 0, l1 = 0, l2 = 0, l3 = 0;
    bool right = false;
    bool invalid = false;   // kXdInvalid task: empty extents
    uint32_t nc = 0;
    // this lane's values on anti-diagonals d-2, d-1, d
    int A1 = kUndef, A2 = kUndef, A3 = kUndef;
    // items claimed by this wave and not yet handed out: [cbase, cend)
    uint32_t cbase = 0, cend = 0;

    for (;;) {
        // ---- hand out items to rows that need one.  The wave claims
        // kChunk items per atomic and hands them out in row order; all
        // lanes run this block together (wave-uniform control flow).
        {
            uint64_t need = __ballot(state == kNeed && j == 0);
            while (need) {
                if (cbase >= cend) {
                    uint32_t b = atomicAdd(a.queue, lane == 0 ? kChunk : 0u);
                    b = __builtin_amdgcn_readfirstlane(b);
                    cbase = b;
                    cend = b + kChunk;
                }
                const uint32_t cnt = (uint32_t)__popcll(need);
                const uint32_t take = min(cnt, cend - cbase);
                const uint32_t rank = (uint32_t)__popcll(
                    need & ((1ull << (16 * row)) - 1ull));
                const bool mine = ((need >> (16 * row)) & 1ull) != 0 &&
                                  rank < take;
                if (mine) {
                    const uint32_t idx = cbase + rank;
                    if (idx >= n) {
                        state = kDone;
                    } else {
                        k = a.list ? glb(a.list)[idx] : idx;
                        const XdTask T = a.tasks[k >> 1];
                        const uint8_t PA_GLOBAL *h = glb_addr<const uint8_t>(T.h_off);
                        const uint8_t PA_GLOBAL *v = glb_addr<const uint8_t>(T.v_off);
                        const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
                        right = (k & 1) != 0;
                        if (!right) {   // query V[0, beg_v), database H[0, beg_h)
                            qs = v;
                            cols = beg_v + 1;
                            ds = h;
                            rows = beg_h + 1;
                        } else {        // suffixes after the seed
                            qs = v + beg_v + a.seed_len;
                            cols = (int)T.nv - beg_v - a.seed_len + 1;
                            ds = h + beg_h + a.seed_len;
                            rows = (int)T.nh - beg_h - a.seed_len + 1;
                        }
                        // anti-diagonal 0: column 0 = 0; anti-diagonal 1:
                        // columns 0 and 1 = gap (undefined when -gap > X)
                        const int g1 = -gap > X ? kUndef : gap;
                        A2 = j == 0 ? 0 : kUndef;
                        A3 = j <= 1 ? g1 : kUndef;
                        d = 1;
                        minCol = 1;
                        maxCol = 2;
                        off1 = off2 = off3 = 0;
                        l1 = 0;
                        l2 = 1;
                        l3 = 2;
                        best = 0;
                        nc = 0;
                        state = kRun;
                        if ((T.pad & kXdInvalid) || rows <= 1 || cols <= 1)
                            state = 3;   // finished without an update
                        invalid = (T.pad & kXdInvalid) != 0;
                    }
                }
                cbase += take;
                for (uint32_t t = 0; t < take; ++t)
                    need &= need - 1;   // the rows served
            }
        }
        if (!__ballot(state != kDone))
            break;

        // ---- one anti-diagonal for running rows whose band is not empty
        const bool run = state == kRun && minCol < maxCol;
        // the written range of the new anti-diagonal is [minCol-1, maxCol]
        const bool ovf = run && maxCol - minCol + 2 > G;
        const bool step = run && !ovf;
        int nd = d + 1;
        int noff3 = minCol - 1;
        int nl3 = maxCol + 1 - noff3;
        // this lane's column on the new anti-diagonal
        const int c = noff3 + ((j - noff3) & 15);
        const bool inr = c <= maxCol;
        const bool comp = inr && c >= minCol && c < maxCol;
        // left neighbours: column c-1 on anti-diagonal d (A3) and d-1 (A2);
        // column c on d is this lane's A3 (the written ranges nest, so these
        // registers hold exactly those columns wherever they are read)
        const int lA3 = ror<1>(A3);
        const int lA2 = ror<1>(A2);
        int v = kUndef;
        {
            const int lim = best - X;
            const bool ld = comp && step;
            const int qp = right ? c - 1 : cols - 1 - c;
            const int dp = right ? nd - c - 1 : rows - 1 + c - nd;
            // rows without a live item read a byte of the score table
            const uint8_t PA_GLOBAL *qa =
                ld ? qs + qp : glb((const uint8_t *)a.mat);
            const uint8_t PA_GLOBAL *da =
                ld ? ds + dp : glb((const uint8_t *)a.mat);
            const int sc = mat[(int)(*qa & 31) * 32 + (*da & 31)];
            int tmp = max(lA3, A3) + gap;
            tmp = max(tmp, lA2 + sc);
            const int cv = tmp < lim ? kUndef : tmp;
            const bool edge_ok = (long)nd * gap > (long)best - X;
            const int ev_lo = edge_ok && noff3 == 0 ? nd * gap : kUndef;
            const int ev_hi = edge_ok && nd == maxCol ? nd * gap : kUndef;
            v = c == noff3 ? ev_lo : (c == maxCol ? ev_hi : (comp ? cv : kUndef));
        }
        // anti-diagonal best: d*gap or better defined computed cells
        const int adb = row_max(comp ? v : INT_MIN);
        // minCol: first column c >= minCol (column order) where the run of
        // cells undefined on the new anti-diagonal and (column c-1) on d
        // stops; column c-1 lies on d's written range [off3, off3+l3) or
        // the run stops there
        const bool stop = !(v == kUndef && c - 1 - off3 < l3 && lA3 == kUndef);
        const uint32_t mlo = row_bits(__ballot(inr && c >= minCol && stop), row);
        // maxCol: last c in [off, maxCol-1] defined on the new anti-diagonal
        // or on d
        const bool keep = v != kUndef || A3 != kUndef;
        const uint32_t mhi = row_bits(__ballot(inr && c < maxCol && keep), row);
        if (step) {
            const uint32_t olo = rot16(mlo, noff3);  // bit i = column off3 + i
            const uint32_t ohi = rot16(mhi, noff3);
            // lowest set bit at or above minCol - off3 = 1
            const uint32_t lo1 = olo & ~1u;
            const int nmin = lo1 ? noff3 + __builtin_ctz(lo1) : maxCol + 1;
            const int nmax = ohi ? noff3 + 31 - __builtin_clz(ohi) + 1 : noff3;
            best = max(best, max(adb, nd * gap));
            nc += (uint32_t)(maxCol - minCol);
            A1 = A2;
            A2 = A3;
            A3 = v;
            d = nd;
            off1 = off2;
            off2 = off3;
            off3 = noff3;
            l1 = l2;
            l2 = l3;
            l3 = nl3;
            minCol = max(nmin, nd + 2 - rows);
            maxCol = min(nmax + 1, cols);
        }
        if (ovf) {
            state = 4;   // overflow: recomputed by the per-lane kernels
        }

        // ---- finish rows whose band emptied (or that never started)
        const bool fin = (state == kRun && minCol >= maxCol) || state == 3 ||
                         state == 4;
        if (__ballot(fin)) {
            // longest extension: the last computed column of the last
            // anti-diagonal, then the end of d-1, else the best of d-2
            const int rb = row * 16;
            int lcol = l3 + off3 - 2;
            int lrow = d - lcol;
            int ls = __shfl(A3, rb + (lcol & 15));
            const int c2a = l2 + off2 - 2, c2b = l2 + off2 - 3;
            const int v2a = __shfl(A2, rb + (c2a & 15));
            const int v2b = __shfl(A2, rb + (c2b & 15));
            if (ls == kUndef) {
                if (v2a != kUndef) {
                    lcol = c2a;
                    lrow = d - 1 - lcol;
                    ls = v2a;
                } else if (l2 > 2 && v2b != kUndef) {
                    lcol = c2b;
                    lrow = d - 1 - lcol;
                    ls = v2b;
                }
            }
            // first maximum of anti-diagonal d-2 over columns [off1, off1+l1)
            const int c1 = off1 + ((j - off1) & 15);
            const bool in1 = c1 < off1 + l1;
            const int m1 = row_max(in1 ? A1 : INT_MIN);
            const int cm = row_min(in1 && A1 == m1 ? c1 : INT_MAX);
            if (ls == kUndef && l1 > 0 && m1 > ls) {
                ls = m1;
                lcol = cm;
                lrow = d - 2 - lcol;
            }
            if (fin && j == 0) {
                const XdTask T = a.tasks[k >> 1];
                XdOut PA_GLOBAL &o = glb(a.out)[T.id];
                const bool upd = state != 3 && ls != kUndef;
                const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
                const uint16_t fl = state == 4 ? 1 : 0;
                if (state == 4) {
                    const uint32_t pos = atomicAdd(a.ovf_count, 1u);
                    a.ovf_list[pos] = k;
                }
                if (!right) {
                    o.beg_h = invalid ? 0 : (upd ? beg_h - lrow : beg_h);
                    o.beg_v = invalid ? 0 : (upd ? beg_v - lcol : beg_v);
                    o.score_left = state == 3 ? 0 : ls;
                    o.cells_left = nc;
                    o.flags_left = fl;
                } else {
                    o.end_h = invalid ? 0
                              : (upd ? beg_h + a.seed_len + lrow
                                     : beg_h + a.seed_len);
                    o.end_v = invalid ? 0
                              : (upd ? beg_v + a.seed_len + lcol
                                     : beg_v + a.seed_len);
                    o.score_right = state == 3 ? 0 : ls;
                    o.cells_right = nc;
                    o.flags_right = fl;
                }
            }
            state = fin ? kNeed : state;
        }
    }
}

}  // namespace

hipError_t launch_xdrop_coop(const XdTask *tasks, const uint32_t *list,
                             const uint32_t *count, uint32_t *queue,
                             const int8_t *mat, int gap, int xdrop,
                             int seed_len, XdOut *out, uint32_t *ovf_list,
                             uint32_t *ovf_count, int grid, hipStream_t s)
{
    CoopArgs a{tasks, list, count, queue, mat, gap, xdrop, seed_len,
               out, ovf_list, ovf_count};
    hipLaunchKernelGGL(k_xdrop_coop, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

int xdrop_coop_occupancy()
{
    static int cache = 0;
    if (!cache) {
        int n = 0;
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_xdrop_coop, 256, 0);
        cache = n > 0 ? n : 1;
    }
    return cache;
}

}  // namespace pa
