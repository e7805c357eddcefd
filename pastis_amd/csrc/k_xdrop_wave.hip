// k_xdrop_wave.hip -- gapped X-drop seed extension with one wave per work
// item, for the items whose band outgrows the per-lane kernel's LDS rings
// (k_xdrop.hip).  Same algorithm, step for step, as k_xdrop_pl and the
// oracle's or_xdrop_one (SeqAn's one-direction GappedXDrop as called at
// src/align/SeqanXdropAligner.cpp:281-283), so extents and evaluated-cell
// counts are bit-identical.
//
// The cells of one anti-diagonal depend only on the two before it, so the
// 64 lanes of a wave evaluate 64 consecutive columns of the band at once,
// chunk after chunk.  The three live anti-diagonals sit in LDS rings of
// kCap columns indexed by absolute column (the band of an item never spans
// more than kCap columns here; a wider one is handed on to the global-ring
// kernel through the overflow list); the residues of both segments are
// staged in LDS when they fit.  The band update is a pair of ballots (the
// first column from the left and the last from the right that stay live)
// and the anti-diagonal's best score a wave maximum.  An item is claimed
// with one atomic per wave, so long extensions spread over the whole GPU.
#include <limits.h>

#include "pa_internal.h"

namespace pa {
namespace {

constexpr int32_t kUndef = INT_MIN / 4;
constexpr int kCap = 1024;      // columns of one anti-diagonal ring
constexpr int kResCap = 4096;   // residues of a segment staged in LDS

struct WaveArgs {
    const XdTask *tasks;
    const uint32_t *list;   // work items (k = 2 * task + dir)
    const uint32_t *count;
    uint32_t *queue;
    const int8_t *mat;
    int gap, X, seed_len;
    XdOut *out;
    uint32_t *ovf_list, *ovf_count;
};

__device__ inline int32_t wave_max(int32_t v)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const int32_t o = __shfl_xor(v, m);
        v = o > v ? o : v;
    }
    return v;
}

// Writes the half of XdOut that item k (direction k & 1) owns.
__device__ inline void wave_write(const WaveArgs &a, uint32_t k, int32_t score,
                                  int col, int row, bool upd, uint32_t nc,
                                  uint16_t flags)
{
    const XdTask T = a.tasks[k >> 1];
    XdOut PA_GLOBAL &o = glb(a.out)[T.id];
    const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
    if ((k & 1) == 0) {
        o.beg_h = upd ? beg_h - row : beg_h;
        o.beg_v = upd ? beg_v - col : beg_v;
        o.score_left = score;
        o.cells_left = nc;
        o.flags_left = flags;
    } else {
        const int end_h = beg_h + a.seed_len, end_v = beg_v + a.seed_len;
        o.end_h = upd ? end_h + row : end_h;
        o.end_v = upd ? end_v + col : end_v;
        o.score_right = score;
        o.cells_right = nc;
        o.flags_right = flags;
    }
}

__global__ __launch_bounds__(64) void k_xdrop_wave(WaveArgs a)
{
    __shared__ int8_t mat[32 * 32];
    __shared__ int32_t ring[3][kCap];
    __shared__ uint8_t qres[kResCap], dres[kResCap];
    const int lane = threadIdx.x;
    for (int i = lane; i < 32 * 32; i += 64)
        mat[i] = a.mat[i];
    const uint32_t n = *a.count;
    const int gap = a.gap, X = a.X;
    constexpr int M = kCap - 1;
    for (;;) {
        // every lane takes part in the claim (lanes other than 0 add 0)
        uint32_t idx = atomicAdd(a.queue, lane == 0 ? 1u : 0u);
        idx = __builtin_amdgcn_readfirstlane(idx);
        if (idx >= n)
            break;
        const uint32_t k = glb(a.list)[idx];
        const XdTask T = a.tasks[k >> 1];
        const uint8_t PA_GLOBAL *h = glb_addr<const uint8_t>(T.h_off);
        const uint8_t PA_GLOBAL *v = glb_addr<const uint8_t>(T.v_off);
        const int beg_h = (int)T.beg_h, beg_v = (int)T.beg_v;
        const bool right = (k & 1) != 0;
        // segments as in k_xdrop_pl: column c of the query is qb[qstep * c],
        // row r of the database db[dstep * r], both 1-based
        int cols, rows, qstep, dstep;
        const uint8_t PA_GLOBAL *qb, *db;
        if (!right) {
            cols = beg_v + 1;
            rows = beg_h + 1;
            qb = v + (cols - 1);
            qstep = -1;
            db = h + (rows - 1);
            dstep = -1;
        } else {
            cols = (int)T.nv - beg_v - a.seed_len + 1;
            rows = (int)T.nh - beg_h - a.seed_len + 1;
            qb = v + beg_v + a.seed_len - 1;
            qstep = 1;
            db = h + beg_h + a.seed_len - 1;
            dstep = 1;
        }
        // the per-lane kernel has finished every item that reaches here
        // with a non-trivial band, so both segments are non-empty
        const bool staged = cols - 1 <= kResCap && rows - 1 <= kResCap;
        __syncthreads();   // the previous item's last reads are done
        if (staged) {
            for (int c = 1 + lane; c < cols; c += 64)
                qres[c - 1] = qb[qstep * c];
            for (int r = 1 + lane; r < rows; r += 64)
                dres[r - 1] = db[dstep * r];
        }
        // item_init: anti-diagonals 0 and 1
        int r1 = 0, r2 = 1, r3 = 2, l1 = 0, l2 = 1, l3 = 2;
        int off1 = 0, off2 = 0, off3 = 0;
        int d = 1, minCol = 1, maxCol = 2, best = 0;
        uint32_t nc = 0;
        if (lane == 0) {
            ring[r2][0] = 0;
            ring[r3][0] = -gap > X ? kUndef : gap;
            ring[r3][1] = -gap > X ? kUndef : gap;
        }
        bool overflow = false;
        for (;;) {
            __syncthreads();
            // one anti-diagonal (item_step)
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
            if (l3 > kCap) {
                overflow = true;
                break;
            }
            const int maxCol0 = maxCol;
            if (lane == 0) {
                ring[r3][off3 & M] = kUndef;
                ring[r3][maxCol0 & M] = kUndef;
                if ((long)d * gap > (long)best - X) {
                    if (off3 == 0)
                        ring[r3][0] = d * gap;
                    if (d - maxCol0 == 0)
                        ring[r3][maxCol0 & M] = d * gap;
                }
            }
            int32_t adbest = d * gap;
            const int32_t lim = best - X;
            for (int base = minCol; base < maxCol0; base += 64) {
                const int col = base + lane;
                if (col < maxCol0) {
                    const int32_t u = ring[r2][col & M];
                    const int32_t l = ring[r2][(col - 1) & M];
                    int32_t tmp = (l > u ? l : u) + gap;
                    const int row = d - col;
                    const int qr = staged ? (int)qres[col - 1]
                                          : (int)qb[qstep * col];
                    const int dr = staged ? (int)dres[row - 1]
                                          : (int)db[dstep * row];
                    const int32_t dg =
                        ring[r1][(col - 1) & M] + mat[qr * 32 + dr];
                    tmp = dg > tmp ? dg : tmp;
                    if (tmp < lim) {
                        ring[r3][col & M] = kUndef;
                    } else {
                        ring[r3][col & M] = tmp;
                        adbest = tmp > adbest ? tmp : adbest;
                    }
                }
            }
            nc += (uint32_t)(maxCol0 - minCol);
            adbest = wave_max(adbest);
            best = adbest > best ? adbest : best;
            __syncthreads();
            // new minCol: the first column c >= minCol that does not satisfy
            // (c in A3's range, A3[c] undefined, c-1 in A2's range, A2[c-1]
            // undefined); maxCol0 + 1 when all of [minCol, maxCol0] do
            int nmin = maxCol0 + 1;
            for (int base = minCol; base <= maxCol0; base += 64) {
                const int c = base + lane;
                const bool dead = c - off3 < l3 &&
                                  ring[r3][c & M] == kUndef &&
                                  c - off2 - 1 < l2 &&
                                  ring[r2][(c - 1) & M] == kUndef;
                const uint64_t live = __ballot(c <= maxCol0 && !dead);
                if (live) {
                    nmin = base + __ffsll((unsigned long long)live) - 1;
                    break;
                }
            }
            // new maxCol: one past the last column x in [off3, maxCol0) with
            // A3[x] or A2[x] defined (off3 when none), plus one
            int nmax = off3;
            for (int top = maxCol0 - 1; top >= off3; top -= 64) {
                const int x = top - lane;
                const bool alive = x >= off3 &&
                                   !(ring[r3][x & M] == kUndef &&
                                     ring[r2][x & M] == kUndef);
                const uint64_t live = __ballot(alive);
                if (live) {
                    nmax = top - (__ffsll((unsigned long long)live) - 1) + 1;
                    break;
                }
            }
            ++nmax;
            if (nmin < d + 2 - rows)
                nmin = d + 2 - rows;
            if (nmax > cols)
                nmax = cols;
            minCol = nmin;
            maxCol = nmax;
            if (minCol >= maxCol)
                break;
        }
        if (overflow) {
            if (lane == 0) {
                const uint32_t pos = atomicAdd(a.ovf_count, 1u);
                a.ovf_list[pos] = k;
                wave_write(a, k, 0, 0, 0, false, 0, 1);
            }
            continue;
        }
        __syncthreads();
        // item_end: the "longest extension" end point
        int lcol = l3 + off3 - 2;
        int lrow = d - lcol;
        int32_t ls = ring[r3][lcol & M];
        if (ls == kUndef) {
            if (ring[r2][(l2 + off2 - 2) & M] != kUndef) {
                lcol = l2 + off2 - 2;
                lrow = d - 1 - lcol;
                ls = ring[r2][lcol & M];
            } else if (l2 > 2 && ring[r2][(l2 + off2 - 3) & M] != kUndef) {
                lcol = l2 + off2 - 3;
                lrow = d - 1 - lcol;
                ls = ring[r2][lcol & M];
            }
        }
        if (ls == kUndef) {
            // the first maximum above kUndef of anti-diagonal d - 2: per
            // lane over its columns, then the largest value and, among
            // equal values, the smallest column
            int32_t bv = kUndef;
            int bi = INT_MAX;
            for (int i = lane; i < l1; i += 64) {
                const int32_t x = ring[r1][(off1 + i) & M];
                if (x > bv) {
                    bv = x;
                    bi = i;
                }
            }
#pragma unroll
            for (int m = 1; m < 64; m <<= 1) {
                const int32_t ov = __shfl_xor(bv, m);
                const int oi = __shfl_xor(bi, m);
                if (ov > bv || (ov == bv && oi < bi)) {
                    bv = ov;
                    bi = oi;
                }
            }
            if (bv > kUndef) {
                ls = bv;
                lcol = bi + off1;
                lrow = d - 2 - lcol;
            }
        }
        if (lane == 0)
            wave_write(a, k, ls, lcol, lrow, ls != kUndef, nc, 0);
    }
}

}  // namespace

hipError_t launch_xdrop_wave(const XdTask *tasks, const uint32_t *list,
                             const uint32_t *count, uint32_t *queue,
                             const int8_t *mat, int gap, int xdrop,
                             int seed_len, XdOut *out, uint32_t *ovf_list,
                             uint32_t *ovf_count, int grid, hipStream_t s)
{
    WaveArgs a{tasks, list, count, queue, mat, gap, xdrop, seed_len,
               out, ovf_list, ovf_count};
    hipLaunchKernelGGL(k_xdrop_wave, dim3(grid), dim3(64), 0, s, a);
    return hipGetLastError();
}

int xdrop_wave_occupancy()
{
    static int cache = 0;
    if (!cache) {
        int n = 0;
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_xdrop_wave, 64, 0);
        cache = n > 0 ? n : 1;
    }
    return cache;
}

}  // namespace pa
