// k_xdrop.hip -- gapped X-drop seed extension (SeqAn's GappedXDrop on a
// simple seed, as called at src/align/SeqanXdropAligner.cpp:281-283).
//
// One extension task (one seed of one pair, both directions) per lane.  Each
// lane keeps the three live anti-diagonals in its own slice of a global
// scratch array (cached in L1/L2); the band update, the clipping at the
// segment ends and the "longest extension" end point follow the oracle's
// or_xdrop_one step for step, so the extents and the evaluated-cell count are
// bit-identical.  A band wider than the slice marks the task for a re-run
// with a larger slice.
#include <limits.h>

#include "pa_internal.h"

namespace pa {
namespace {

constexpr int32_t kUndef = INT_MIN / 4;

struct OneDir {
    int32_t score;
    int32_t col, row;
    bool upd;
    bool ovf;
};

__device__ OneDir xd_one(const uint8_t *qs, int qlen, const uint8_t *ds,
                         int dlen, bool right, const int8_t *mat, int gap,
                         int X, int32_t *buf, uint32_t cap, uint32_t *cells)
{
    OneDir res = {0, 0, 0, false, false};
    const int cols = qlen + 1, rows = dlen + 1;
    if (rows == 1 || cols == 1)
        return res;
    int32_t *a1 = buf, *a2 = buf + cap, *a3 = buf + 2 * cap;
    int l1 = 0, l2 = 1, l3 = 2;
    a2[0] = 0;
    if (-gap > X) {
        a3[0] = kUndef;
        a3[1] = kUndef;
    } else {
        a3[0] = gap;
        a3[1] = gap;
    }
    int minCol = 1, maxCol = 2;
    int off1 = 0, off2 = 0, off3 = 0;
    int d = 1;
    int32_t best = 0;
    uint32_t nc = 0;
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
        l3 = maxCol + 1 - off3;
        if ((uint32_t)l3 > cap) {
            res.ovf = true;
            return res;
        }
        a3[0] = kUndef;
        a3[maxCol - off3] = kUndef;
        if ((long)d * gap > (long)best - X) {
            if (off3 == 0)
                a3[0] = d * gap;
            if (d - maxCol == 0)
                a3[maxCol - off3] = d * gap;
        }
        int32_t adbest = d * gap;
        const int32_t lim = best - X;
        for (int col = minCol; col < maxCol; ++col) {
            const int i3 = col - off3, i2 = col - off2, i1 = col - off1;
            int qp, dp;
            if (right) {
                qp = col - 1;
                dp = d - col - 1;
            } else {
                qp = cols - 1 - col;
                dp = rows - 1 + col - d;
            }
            const int32_t l = a2[i2 - 1], u = a2[i2];
            int32_t tmp = (l > u ? l : u) + gap;
            const int32_t dg = a1[i1 - 1] + mat[(int)qs[qp] * 32 + ds[dp]];
            tmp = dg > tmp ? dg : tmp;
            if (tmp < lim) {
                a3[i3] = kUndef;
            } else {
                a3[i3] = tmp;
                adbest = tmp > adbest ? tmp : adbest;
            }
        }
        nc += (uint32_t)(maxCol - minCol);
        best = adbest > best ? adbest : best;
        while (minCol - off3 < l3 && a3[minCol - off3] == kUndef &&
               minCol - off2 - 1 < l2 && a2[minCol - off2 - 1] == kUndef)
            ++minCol;
        while (maxCol - off3 > 0 && a3[maxCol - off3 - 1] == kUndef &&
               a2[maxCol - off2 - 1] == kUndef)
            --maxCol;
        ++maxCol;
        if (minCol < d + 2 - rows)
            minCol = d + 2 - rows;
        if (maxCol > cols)
            maxCol = cols;
    }
    *cells += nc;
    int lcol = l3 + off3 - 2;
    int lrow = d - lcol;
    int32_t ls = a3[lcol - off3];
    if (ls == kUndef) {
        if (a2[l2 - 2] != kUndef) {
            lcol = l2 + off2 - 2;
            lrow = d - 1 - lcol;
            ls = a2[lcol - off2];
        } else if (l2 > 2 && a2[l2 - 3] != kUndef) {
            lcol = l2 + off2 - 3;
            lrow = d - 1 - lcol;
            ls = a2[lcol - off2];
        }
    }
    if (ls == kUndef) {
        for (int i = 0; i < l1; ++i) {
            if (a1[i] > ls) {
                ls = a1[i];
                lcol = i + off1;
                lrow = d - 2 - lcol;
            }
        }
    }
    res.score = ls;
    if (ls != kUndef) {
        res.upd = true;
        res.col = lcol;
        res.row = lrow;
    }
    return res;
}

__global__ __launch_bounds__(256) void k_xdrop(
    const XdTask *__restrict__ tasks, const uint32_t *__restrict__ list,
    const uint32_t *count, uint32_t *queue, const uint8_t *__restrict__ codes,
    const int8_t *__restrict__ gmat, int gap, int X, int32_t *scratch,
    uint32_t cap, XdOut *out, uint32_t *ovf_list, uint32_t *ovf_count,
    int seed_len)
{
    __shared__ int8_t mat[32 * 32];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x)
        mat[i] = gmat[i];
    __syncthreads();
    const uint32_t n = *count;
    const int lane = threadIdx.x & 63;
    int32_t *buf = scratch + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) *
                                 (size_t)(3 * cap);
    for (;;) {
        // all lanes join the atomic (lane 0 adds 64) and the loop body has
        // no divergent continue, so the wave re-enters the loop together
        uint32_t base = atomicAdd(queue, lane == 0 ? 64u : 0u);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= n)
            break;
        const uint32_t k = base + (uint32_t)lane;
        if (k < n) {
        const uint32_t tid = list ? list[k] : k;
        const XdTask T = tasks[tid];
        const uint8_t *h = (const uint8_t *)(uintptr_t)T.h_off;
        const uint8_t *v = (const uint8_t *)(uintptr_t)T.v_off;
        int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
        int end_h = beg_h + seed_len, end_v = beg_v + seed_len;
        uint32_t cells = 0;
        XdOut o = {};
        // left: query = V[0, beg_v), database = H[0, beg_h), reversed
        OneDir L = xd_one(v, beg_v, h, beg_h, false, mat, gap, X, buf, cap,
                          &cells);
        OneDir Rr = {0, 0, 0, false, false};
        if (!L.ovf)
            Rr = xd_one(v + end_v, (int)T.nv - end_v, h + end_h,
                        (int)T.nh - end_h, true, mat, gap, X, buf, cap,
                        &cells);
        if (L.ovf || Rr.ovf) {
            o.flags = 1;
            const uint32_t pos = atomicAdd(ovf_count, 1u);  // no wave op follows
            ovf_list[pos] = tid;
        } else {
            if (L.upd) {
                beg_h -= L.row;
                beg_v -= L.col;
            }
            if (Rr.upd) {
                end_h += Rr.row;
                end_v += Rr.col;
            }
            o.beg_h = beg_h;
            o.beg_v = beg_v;
            o.end_h = end_h;
            o.end_v = end_v;
            o.score_left = L.score;
            o.score_right = Rr.score;
            o.cells = cells;
        }
        out[T.id] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-ring variant.  Each lane keeps the three live anti-diagonals in its own
// column ring of kRing slots in LDS: the value of column c on an
// anti-diagonal sits in slot c mod kRing of that anti-diagonal's ring.  The
// algorithm only reads columns inside the range written on that anti-diagonal
// ([off, maxCol]), so the ring is exact as long as that range is at most kRing
// columns wide; a wider band marks the task for the global-scratch kernel.
// Ring layout: [3 anti-diagonals][kRing slots][64 lanes] of int32.
// ---------------------------------------------------------------------------
constexpr int kRing = 32;

struct Ring {
    int32_t *base;  // LDS, this lane's column of the [3][kRing][64] array
    __device__ inline int32_t &at(int r, int col) const
    {
        return base[(r * kRing + (col & (kRing - 1))) * 64];
    }
};

__device__ OneDir xd_one_ring(const uint8_t *qs, int qlen, const uint8_t *ds,
                              int dlen, bool right, const int8_t *mat, int gap,
                              int X, const Ring &A, uint32_t *cells)
{
    OneDir res = {0, 0, 0, false, false};
    const int cols = qlen + 1, rows = dlen + 1;
    if (rows == 1 || cols == 1)
        return res;
    int r1 = 0, r2 = 1, r3 = 2;   // ring of anti-diagonals d-2, d-1, d
    int l1 = 0, l2 = 1, l3 = 2;
    A.at(r2, 0) = 0;
    if (-gap > X) {
        A.at(r3, 0) = kUndef;
        A.at(r3, 1) = kUndef;
    } else {
        A.at(r3, 0) = gap;
        A.at(r3, 1) = gap;
    }
    int minCol = 1, maxCol = 2;
    int off1 = 0, off2 = 0, off3 = 0;
    int d = 1;
    int32_t best = 0;
    uint32_t nc = 0;
    while (minCol < maxCol) {
        ++d;
        const int t = r1;
        r1 = r2;
        r2 = r3;
        r3 = t;
        l1 = l2;
        l2 = l3;
        off1 = off2;
        off2 = off3;
        off3 = minCol - 1;
        l3 = maxCol + 1 - off3;
        if (l3 > kRing) {
            res.ovf = true;
            return res;
        }
        A.at(r3, off3) = kUndef;
        A.at(r3, maxCol) = kUndef;
        if ((long)d * gap > (long)best - X) {
            if (off3 == 0)
                A.at(r3, off3) = d * gap;
            if (d - maxCol == 0)
                A.at(r3, maxCol) = d * gap;
        }
        int32_t adbest = d * gap;
        const int32_t lim = best - X;
        for (int col = minCol; col < maxCol; ++col) {
            int qp, dp;
            if (right) {
                qp = col - 1;
                dp = d - col - 1;
            } else {
                qp = cols - 1 - col;
                dp = rows - 1 + col - d;
            }
            const int32_t l = A.at(r2, col - 1), u = A.at(r2, col);
            int32_t tmp = (l > u ? l : u) + gap;
            const int32_t dg =
                A.at(r1, col - 1) + mat[(int)qs[qp] * 32 + ds[dp]];
            tmp = dg > tmp ? dg : tmp;
            if (tmp < lim) {
                A.at(r3, col) = kUndef;
            } else {
                A.at(r3, col) = tmp;
                adbest = tmp > adbest ? tmp : adbest;
            }
        }
        nc += (uint32_t)(maxCol - minCol);
        best = adbest > best ? adbest : best;
        while (minCol - off3 < l3 && A.at(r3, minCol) == kUndef &&
               minCol - off2 - 1 < l2 && A.at(r2, minCol - 1) == kUndef)
            ++minCol;
        while (maxCol - off3 > 0 && A.at(r3, maxCol - 1) == kUndef &&
               A.at(r2, maxCol - 1) == kUndef)
            --maxCol;
        ++maxCol;
        if (minCol < d + 2 - rows)
            minCol = d + 2 - rows;
        if (maxCol > cols)
            maxCol = cols;
    }
    *cells += nc;
    int lcol = l3 + off3 - 2;
    int lrow = d - lcol;
    int32_t ls = A.at(r3, lcol);
    if (ls == kUndef) {
        if (A.at(r2, l2 + off2 - 2) != kUndef) {
            lcol = l2 + off2 - 2;
            lrow = d - 1 - lcol;
            ls = A.at(r2, lcol);
        } else if (l2 > 2 && A.at(r2, l2 + off2 - 3) != kUndef) {
            lcol = l2 + off2 - 3;
            lrow = d - 1 - lcol;
            ls = A.at(r2, lcol);
        }
    }
    if (ls == kUndef) {
        for (int i = 0; i < l1; ++i) {
            const int32_t v = A.at(r1, off1 + i);
            if (v > ls) {
                ls = v;
                lcol = i + off1;
                lrow = d - 2 - lcol;
            }
        }
    }
    res.score = ls;
    if (ls != kUndef) {
        res.upd = true;
        res.col = lcol;
        res.row = lrow;
    }
    return res;
}

__global__ __launch_bounds__(64) void k_xdrop_ring(
    const XdTask *__restrict__ tasks, const uint32_t *count, uint32_t *queue,
    const int8_t *__restrict__ gmat, int gap, int X, XdOut *out,
    uint32_t *ovf_list, uint32_t *ovf_count, int seed_len)
{
    __shared__ int8_t mat[32 * 32];
    __shared__ int32_t ring[3 * kRing * 64];
    for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x)
        mat[i] = gmat[i];
    __syncthreads();
    const uint32_t n = *count;
    const int lane = threadIdx.x;
    const Ring A{ring + lane};
    for (;;) {
        uint32_t base = atomicAdd(queue, lane == 0 ? 64u : 0u);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= n)
            break;
        const uint32_t k = base + (uint32_t)lane;
        if (k < n) {
            const XdTask T = tasks[k];
            const uint8_t *h = (const uint8_t *)(uintptr_t)T.h_off;
            const uint8_t *v = (const uint8_t *)(uintptr_t)T.v_off;
            int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
            int end_h = beg_h + seed_len, end_v = beg_v + seed_len;
            uint32_t cells = 0;
            XdOut o = {};
            OneDir L = xd_one_ring(v, beg_v, h, beg_h, false, mat, gap, X, A,
                                   &cells);
            OneDir Rr = {0, 0, 0, false, false};
            if (!L.ovf)
                Rr = xd_one_ring(v + end_v, (int)T.nv - end_v, h + end_h,
                                 (int)T.nh - end_h, true, mat, gap, X, A,
                                 &cells);
            if (L.ovf || Rr.ovf) {
                o.flags = 1;
                const uint32_t pos = atomicAdd(ovf_count, 1u);
                ovf_list[pos] = k;
            } else {
                if (L.upd) {
                    beg_h -= L.row;
                    beg_v -= L.col;
                }
                if (Rr.upd) {
                    end_h += Rr.row;
                    end_v += Rr.col;
                }
                o.beg_h = beg_h;
                o.beg_v = beg_v;
                o.end_h = end_h;
                o.end_v = end_v;
                o.score_left = L.score;
                o.score_right = Rr.score;
                o.cells = cells;
            }
            out[T.id] = o;
        }
    }
}

}  // namespace

hipError_t launch_xdrop_k(const XdTask *tasks, const uint32_t *list,
                          const uint32_t *count, uint32_t *queue,
                          const uint8_t *codes, const int8_t *mat, int gap,
                          int xdrop, int32_t *scratch, uint32_t cap,
                          XdOut *out, uint32_t *ovf_list, uint32_t *ovf_count,
                          int seed_len, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_xdrop, dim3(grid), dim3(256), 0, s, tasks, list,
                       count, queue, codes, mat, gap, xdrop, scratch, cap, out,
                       ovf_list, ovf_count, seed_len);
    return hipGetLastError();
}

hipError_t launch_xdrop_ring(const XdTask *tasks, const uint32_t *count,
                             uint32_t *queue, const int8_t *mat, int gap,
                             int xdrop, XdOut *out, uint32_t *ovf_list,
                             uint32_t *ovf_count, int seed_len, int grid,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_xdrop_ring, dim3(grid), dim3(64), 0, s, tasks, count,
                       queue, mat, gap, xdrop, out, ovf_list, ovf_count,
                       seed_len);
    return hipGetLastError();
}

}  // namespace pa
