// k_dp.hip -- wavefront dynamic-programming kernels for gfx950.
//
// One DP task per L lanes of a wavefront (L = 64, 32, 16 or 8).  The row
// sequence (SeqAn's vertical sequence, seqsr) is split over the task's
// lanes in strips of 2R rows: the upper R rows ("lo") and the lower R rows
// ("hi") of a lane share one register per row, packed int16 lo/hi, and the
// hi half runs one column behind the lo half so both halves advance with
// the same instruction.  Lane l works on column t-2l (lo) and t-2l-1 (hi) at
// step t; the H and F values of its last row go to lane l+1 through a DPP
// shift.  Columns (the horizontal sequence, seqsc) stream through; a pass
// covers 2RL rows and longer row sequences take several passes linked by a
// boundary row in global memory.
//
// The substitution scores come from a query profile in LDS: for each residue
// code c and each row, score(row residue, c) as an int8.
//
// Recurrences (Gotoh; rows i over the row sequence, columns j over the
// column sequence; gap of length L costs open + (L-1)*ext):
//   E(i,j) = max(E(i,j-1) + ext, H(i,j-1) + open)        ties keep extension
//   F(i,j) = max(F(i-1,j) + ext, H(i-1,j) + open)        ties keep extension
//   H(i,j) = max(H(i-1,j-1) + s, F(i,j), E(i,j) [, 0 local])
//            ties: restart (local) > diagonal > F (vertical) > E (horizontal)
// Local mode reports the first maximum in column-major order (smaller column,
// then smaller row), as SeqAn's column-wise fill with a strict '>' update.
//
// Traced kernels work on 4*value + tag so that one max picks the value and
// the predecessor at once (tags: 3 restart (local) or diagonal between equal
// residues (global), 2 diagonal, 1 F, 0 E); they store
// a 4-bit code per cell (bits 0-1 source, bit 2 F extended, bit 3 E
// extended) for the traceback.
#include "pa_internal.h"

#include <type_traits>

namespace pa {
namespace {

typedef short s2v __attribute__((ext_vector_type(2)));

// v_perm selectors that turn byte 1 (byte 3) of lw and of hw into two
// sign-extended int16 halves: byte 8 / 9 of a selector repeats the sign bit
// of byte 1 / 3 of the low source, 10 / 11 of the high source.
constexpr uint32_t kSelB1 = 0x0a050801u;
constexpr uint32_t kSelB3 = 0x0b070903u;
// bytes 0 and 2 of the low source, then bytes 0 and 2 of the high source
constexpr uint32_t kSelEven = 0x06040200u;

// Value of lane - 1 of the same task (L lanes per task, tasks on aligned
// lane groups); the first lane of a task (fl) takes `old`.  DPP wave_shr:1
// for 64- and 32-lane tasks, row_shr:1 (within 16-lane rows) for 16- and
// 8-lane tasks; a lane whose DPP source lane has no value keeps the "old"
// operand, which covers the first lane of a 64- or 16-lane task, and a
// select covers the first lane of a 32- or 8-lane task, whose DPP source is
// the last lane of the task before it.
template <int L> __device__ inline int lane_up(int old, int src, bool fl)
{
    constexpr int CTRL = L >= 32 ? 0x138 : 0x111;
    const int v = __builtin_amdgcn_update_dpp(old, src, CTRL, 0xf, 0xf, false);
    if constexpr (L == 32 || L == 8)
        return fl ? old : v;
    else
        return v;
}

template <bool I16> struct Ops;

// Two int16 lanes in one VGPR: x = lo half-strip row, y = hi half-strip row.
template <> struct Ops<true> {
    typedef s2v V;
    typedef uint32_t Key;
    __device__ static inline V mk(int lo, int hi)
    {
        V v;
        v.x = (short)lo;
        v.y = (short)hi;
        return v;
    }
    __device__ static inline V splat(int x) { return mk(x, x); }
    __device__ static inline V add(V a, V b)
    {
        return __builtin_elementwise_add_sat(a, b);
    }
    __device__ static inline V max(V a, V b)
    {
        return __builtin_elementwise_max(a, b);
    }
    __device__ static inline uint32_t u(V a)
    {
        return __builtin_bit_cast(uint32_t, a);
    }
    __device__ static inline V v(uint32_t x) { return __builtin_bit_cast(V, x); }
    __device__ static inline V andm(V a, int m)
    {
        uint32_t mm = ((uint32_t)m & 0xffffu) * 0x10001u;
        return v(u(a) & mm);
    }
    __device__ static inline V orm(V a, int m)
    {
        uint32_t mm = ((uint32_t)m & 0xffffu) * 0x10001u;
        return v(u(a) | mm);
    }
    // a | 1 in both halves: sets the F predecessor tag of a traced F value
    // whose low two bits are 00 or 01
    __device__ static inline V fmark(V a) { return v(u(a) | 0x00010001u); }
    __device__ static inline int lo(V a) { return a.x; }
    __device__ static inline int hi(V a) { return a.y; }
    // profile rows 4w..4w+3 of the lo strip (bytes of lw) and of the hi
    // strip (bytes of hw), sign-extended to int16 pairs
    template <int R>
    __device__ static inline void unpack(const uint32_t *lw, const uint32_t *hw,
                                         V *s)
    {
#pragma unroll
        for (int w = 0; w < R / 4; ++w) {
            const uint32_t l = lw[w], h = hw[w], l8 = l << 8, h8 = h << 8;
            s[4 * w + 0] = v(__builtin_amdgcn_perm(h8, l8, kSelB1));
            s[4 * w + 1] = v(__builtin_amdgcn_perm(h, l, kSelB1));
            s[4 * w + 2] = v(__builtin_amdgcn_perm(h8, l8, kSelB3));
            s[4 * w + 3] = v(__builtin_amdgcn_perm(h, l, kSelB3));
        }
    }
    // 4-bit profile words of 8 rows (lw: lo strip, hw: hi strip; row r in
    // byte r % 4, high nibble for r >= 4) as int16 pairs n * MUL + base
    template <int MUL>
    __device__ static inline void unpack_nib(uint32_t lw, uint32_t hw,
                                             int base, V *s)
    {
        typedef unsigned short u2v __attribute__((ext_vector_type(2)));
        const V B = splat(base);
        // the multiplier as an opaque register value: n * MUL + base is then
        // one packed multiply-add (v_pk_mad_u16) instead of a shift and an add
        int mul = MUL;
        if (MUL != 1)
            PA_OPAQUE(mul);
        const V M = splat(mul);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            // [byte k of lw, 0, byte k of hw, 0]
            const uint32_t sel = (uint32_t)k | 0x0c00u |
                                 ((uint32_t)(4 + k) << 16) | 0x0c000000u;
            const uint32_t p = __builtin_amdgcn_perm(hw, lw, sel);
            const V n0 = v(p & 0x000f000fu);
            const V n1 = __builtin_bit_cast(
                V, __builtin_bit_cast(u2v, p) >> (unsigned short)4);
            s[k] = MUL == 1 ? add(n0, B) : n0 * M + B;
            s[k + 4] = MUL == 1 ? add(n1, B) : n1 * M + B;
        }
    }
    // new row-above input: lo <- hi of lane-1, hi <- own lo (lane_up).
    // The first lane of a task takes `first` (the boundary value) in the
    // high half.
    template <int L> __device__ static inline V shuffle_up(V out, int first,
                                                           bool fl)
    {
        uint32_t o = u(out);
        uint32_t sh = (uint32_t)lane_up<L>((int)((uint32_t)first << 16),
                                           (int)o, fl);
        return v(__builtin_amdgcn_alignbit(o, sh, 16));
    }
    __device__ static inline V set_lo(V a, int x)
    {
        a.x = (short)x;
        return a;
    }
    // replace the halves selected by (ilo, ihi) of old with those of nw
    __device__ static inline V blend(V old, V nw, bool ilo, bool ihi)
    {
        uint32_t m = (ilo ? 0x0000ffffu : 0u) | (ihi ? 0xffff0000u : 0u);
        return v((u(nw) & m) | (u(old) & ~m));
    }
    // F-extension tag (bit 0 of fp) into bit 0, E-extension tag (bit 1 of
    // ep) kept in bit 1; higher bits are don't-care
    __device__ static inline V efbits(V fp, V ep)
    {
        return v((u(fp) & 0x00010001u) | (u(ep) & 0xfffefffeu));
    }
    // Trace record of R rows x 2 halves.  Code nibble of a cell: bits 0-1
    // predecessor tag, bit 2 F extended, bit 3 E extended.  Row r, half h
    // sits in byte 4*(r/4) + 2*(r%2) + h, high nibble when (r/2) is odd.
    template <int R>
    __device__ static inline void pack(const V *hp, const V *ef, uint32_t *w)
    {
        uint32_t y[R / 2];
#pragma unroll
        for (int q = 0; q < R / 2; ++q) {
            const uint32_t px =
                __builtin_amdgcn_perm(u(ef[2 * q + 1]), u(ef[2 * q]), kSelEven);
            const uint32_t ph =
                __builtin_amdgcn_perm(u(hp[2 * q + 1]), u(hp[2 * q]), kSelEven);
            y[q] = (ph & 0x03030303u) | ((px << 2) & 0xfcfcfcfcu);
        }
#pragma unroll
        for (int k = 0; k < R / 4; ++k)
            w[k] = (y[2 * k] & 0x0f0f0f0fu) | ((y[2 * k + 1] << 4) & 0xf0f0f0f0u);
    }
    __device__ static inline Key key(int score, int col)
    {
        return ((uint32_t)score << 16) | (uint32_t)(0xffff - col);
    }
    __device__ static inline int key_score(Key k) { return (int)(k >> 16); }
    __device__ static inline int key_col(Key k)
    {
        return 0xffff - (int)(k & 0xffffu);
    }
    static constexpr int kNeg = -16384;
    static constexpr int kNegTraced = -32000;
};

struct I2 {
    int x, y;
};

// Two int32 values per row (the exact re-run and the long global tasks).
template <> struct Ops<false> {
    typedef I2 V;
    typedef uint64_t Key;
    __device__ static inline V mk(int lo, int hi) { return I2{lo, hi}; }
    __device__ static inline V splat(int x) { return I2{x, x}; }
    __device__ static inline V add(V a, V b) { return I2{a.x + b.x, a.y + b.y}; }
    __device__ static inline V max(V a, V b)
    {
        return I2{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y};
    }
    __device__ static inline V andm(V a, int m) { return I2{a.x & m, a.y & m}; }
    __device__ static inline V orm(V a, int m) { return I2{a.x | m, a.y | m}; }
    __device__ static inline V fmark(V a) { return I2{a.x | 1, a.y | 1}; }
    __device__ static inline int lo(V a) { return a.x; }
    __device__ static inline int hi(V a) { return a.y; }
    template <int R>
    __device__ static inline void unpack(const uint32_t *lw, const uint32_t *hw,
                                         V *s)
    {
#pragma unroll
        for (int r = 0; r < R; ++r)
            s[r] = I2{(int)(int8_t)(lw[r / 4] >> (8 * (r % 4))),
                      (int)(int8_t)(hw[r / 4] >> (8 * (r % 4)))};
    }
    template <int MUL>
    __device__ static inline void unpack_nib(uint32_t lw, uint32_t hw,
                                             int base, V *s)
    {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int sh = 8 * (r % 4) + 4 * (r / 4);
            s[r] = I2{(int)((lw >> sh) & 15u) * MUL + base,
                      (int)((hw >> sh) & 15u) * MUL + base};
        }
    }
    template <int L> __device__ static inline V shuffle_up(V out, int first,
                                                           bool fl)
    {
        return I2{lane_up<L>(first, out.y, fl), out.x};
    }
    __device__ static inline V set_lo(V a, int x)
    {
        a.x = x;
        return a;
    }
    __device__ static inline V blend(V old, V nw, bool ilo, bool ihi)
    {
        return I2{ilo ? nw.x : old.x, ihi ? nw.y : old.y};
    }
    __device__ static inline V efbits(V fp, V ep)
    {
        return I2{(fp.x & 1) | (ep.x & ~1), (fp.y & 1) | (ep.y & ~1)};
    }
    // same record layout as the int16 kernels
    template <int R>
    __device__ static inline void pack(const V *hp, const V *ef, uint32_t *w)
    {
#pragma unroll
        for (int k = 0; k < R / 4; ++k)
            w[k] = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t clo = ((uint32_t)hp[r].x & 3u) |
                                 (((uint32_t)ef[r].x & 3u) << 2);
            const uint32_t chi = ((uint32_t)hp[r].y & 3u) |
                                 (((uint32_t)ef[r].y & 3u) << 2);
            const int sh = 8 * (2 * (r % 2)) + 4 * ((r / 2) % 2);
            w[r / 4] |= (clo << sh) | (chi << (sh + 8));
        }
    }
    __device__ static inline Key key(int score, int col)
    {
        return ((uint64_t)(uint32_t)score << 16) | (uint64_t)(0xffff - col);
    }
    __device__ static inline int key_score(Key k) { return (int)(k >> 16); }
    __device__ static inline int key_col(Key k)
    {
        return 0xffff - (int)(k & 0xffffu);
    }
    static constexpr int kNeg = -(1 << 28);
    static constexpr int kNegTraced = -(1 << 28);
};

// One trace record (R bytes) to global memory.
__device__ inline void store_rec(uint8_t PA_GLOBAL *p, const uint32_t *w,
                                 int nw)
{
    if (nw == 1) {
        *(uint32_t PA_GLOBAL *)p = w[0];
    } else {
        typedef uint32_t v2u __attribute__((ext_vector_type(2)));
        v2u x;
        x.x = w[0];
        x.y = w[1];
        *(v2u PA_GLOBAL *)p = x;
    }
}

template <int R> struct ProfWords;
template <> struct ProfWords<4> {
    __device__ static inline void load(const uint8_t *p, uint32_t *w)
    {
        w[0] = *(const uint32_t *)p;
    }
};
template <> struct ProfWords<8> {
    __device__ static inline void load(const uint8_t *p, uint32_t *w)
    {
        uint2 x = *(const uint2 *)p;
        w[0] = x.x;
        w[1] = x.y;
    }
};

// maximum over the L lanes of a task (L = 64: the wave)
template <int L> __device__ inline uint64_t group_max_u64(uint64_t v)
{
    for (int off = L / 2; off >= 1; off >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, off);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), off);
        uint64_t o = ((uint64_t)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

// L lanes per task: 64 (one task per wave), or 32, 16 or 8 (G = 64 / L tasks
// per wave on aligned lane groups, in lockstep; the wave runs as many passes
// as its longest task needs, and each slot keeps its own boundary row).
//
// Kernels with 8 rows per lane keep 4-bit query profiles (rows_nib), half
// the LDS of byte profiles, and are compiled for three waves per SIMD, so
// that twelve one-wave workgroups fit a CU.
template <int R, int L, bool LOCAL, bool TRACE, bool I16>
__global__ __launch_bounds__(64)
PA_WAVES_PER_EU(rows_nib(R) && I16 ? 3 : 1)
void k_dp(DpArgs a)
{
    typedef Ops<I16> O;
    typedef typename O::V V;
    typedef typename O::Key Key;
    static_assert(L == 64 || L == 32 || L == 16 || L == 8, "lanes per task");
    static_assert(R == 4 || R == 8, "rows per lane");
    constexpr bool NIB = rows_nib(R);
    constexpr int G = 64 / L;                 // tasks per wave
    constexpr int ROWS = 2 * R * L;           // rows per pass
    // 32-bit words of one residue code's profile: 4 rows (bytes) or 8 rows
    // (nibbles) per word
    constexpr int PW = NIB ? ROWS / 8 : ROWS / 4;
    // columns per residue block: one per lane (L = 64) or four per lane
    constexpr int RB = L == 64 ? 64 : 4 * L;
    constexpr int SH = TRACE ? 2 : 0;
    constexpr int NEG = TRACE ? O::kNegTraced : O::kNeg;
    HIP_DYNAMIC_SHARED(uint32_t, prof32)
    const int lane = threadIdx.x;
    const int gl = lane % L;     // lane within the task
    const int grp = lane / L;    // task slot of the wave
    const bool fl = gl == 0;     // first lane of its task
    // query profile of this lane's task
    uint32_t *gprof32 = prof32 + grp * (kProfChars * PW);
    const uint8_t *prof = (const uint8_t *)gprof32;
    // profile matrix of this mode in global memory (1 KB, cached): the score,
    // or 4*score+2 when traced, as int8, -128 outside the alphabet (row 31:
    // padding rows).  Global traces tag a diagonal step between equal
    // residues 3 instead of 2 (4*score+3): the traceback counts identities
    // from the tags, and diagonal still wins every tie against F and E.  The
    // LDS holds the query profiles only.  4-bit profiles (NIB) hold score -
    // nib_min (0 outside the alphabet, the lowest score: padding rows stay
    // below every real cell) and tag every diagonal 2.
    const int nbase = TRACE ? 4 * a.nib_min + 2 : a.nib_min;
    const uint32_t count = *a.count;
    const int open = a.gap_open * (1 << SH);
    const int ext = a.gap_ext * (1 << SH);
    const int gmax = open > ext ? open : ext;
    const V OPEN = O::splat(open);
    // Traced: E values carry the extension tag in bit 1 (+2), F values are
    // carried with bit 0 set (fmark), so extending adds the plain 4*ext.
    const V EXT = O::splat(ext);
    const V EXTE = O::splat(TRACE ? ext + 2 : ext);
    // F value of "no cell": NEG, tagged like every carried F when traced
    constexpr int NEGF = TRACE ? NEG + 1 : NEG;
    // boundary row of this task slot (multi-pass tasks)
    int32_t PA_GLOBAL *bnd =
        glb(a.bnd) + ((size_t)blockIdx.x * G + (size_t)grp) * a.bnd_stride;

    // tasks claimed by this wave and not yet started: [cbase, cend)
    uint32_t cbase = 0, cend = 0;
    const uint32_t chunk = a.chunk;
    for (;;) {
        // One queue atomic claims `chunk` tasks.  Every lane takes part in
        // the atomic (only lane 0 adds): a lane-0-only branch ahead of the
        // readfirstlane lets the compiler rotate the loop so that the other
        // lanes re-enter it without lane 0.
        if (cbase >= cend) {
            uint32_t b = atomicAdd(a.queue, lane == 0 ? chunk * G : 0u);
            b = __builtin_amdgcn_readfirstlane(b);
            cbase = b;
            cend = b + chunk * G;
        }
        const uint32_t k0 = cbase;
        cbase += G;
        if (k0 >= count)
            break;
        // task slot grp takes task k0 + grp (none past the end)
        const uint32_t k = k0 + (uint32_t)grp;
        const bool valid = k < count;
        const uint32_t tid = glb(a.list)[valid ? k : k0];
        DpTask T = a.tasks[tid];
        if (!valid)
            T.qlen = T.tlen = 0;
        const uint8_t PA_GLOBAL *qs = glb_addr<const uint8_t>(T.q_off);
        const uint8_t PA_GLOBAL *ts = glb_addr<const uint8_t>(T.t_off);
        const int qlen = (int)T.qlen, tlen = (int)T.tlen;
        const int npass_own = (qlen + ROWS - 1) / ROWS;
        int npass = npass_own;   // wave-uniform: the most passes of the slots
        if (G > 1) {
#pragma unroll
            for (int q = 1; q < G; ++q)
                npass = max(npass, __shfl(npass_own, q * L));
            npass = max(npass, __shfl(npass_own, 0));
        }
        uint8_t PA_GLOBAL *tr = nullptr;
        if (TRACE)
            tr = glb(a.trace) +
                 (glb(a.trace_off)[tid] -
                  (a.trace_base ? *glb(a.trace_base) : 0ull));

        Key bk_lo = O::key(0, 0), bk_hi = O::key(0, 0);
        int bp_lo = 0, bp_hi = 0;
        V saved[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            saved[r] = O::splat(0);
        int gscore = 0;

        for (int p = 0; p < npass; ++p) {
            const int i0 = p * ROWS;
            __syncthreads();
            // query profile of rows [i0, i0 + ROWS): each lane writes its own
            // 2R rows, four rows (one dword per residue code) at a time,
            // gathering bytes of four matrix rows with v_perm; 4-bit
            // profiles put rows 4-7 of a strip in the high nibbles of rows
            // 0-3's bytes (one dword per strip and code)
            uint32_t xs[NIB ? kProfChars : 1];
#pragma unroll
            for (int g = 0; g < R / 2; ++g) {
                const int row0 = i0 + 2 * R * gl + 4 * g;
                int qc[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    qc[k] = row0 + k < qlen ? (int)qs[row0 + k] : 31;
                uint32_t mr[4][8];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    // generic pointer: vector-struct loads from the global
                    // address space do not compile
                    const uint4 *src = (const uint4 *)(a.pmat + qc[k] * 8);
                    const uint4 x0 = src[0], x1 = src[1];
                    mr[k][0] = x0.x; mr[k][1] = x0.y; mr[k][2] = x0.z;
                    mr[k][3] = x0.w; mr[k][4] = x1.x; mr[k][5] = x1.y;
                    mr[k][6] = x1.z; mr[k][7] = x1.w;
                }
                uint32_t *dst = gprof32 + (2 * R * gl + 4 * g) / 4;
#pragma unroll
                for (int c = 0; c < kProfChars; ++c) {
                    const int wd = c >> 2, sh = c & 3;
                    const uint32_t sel =
                        (uint32_t)sh | ((uint32_t)(4 + sh) << 8) | 0x0c0c0000u;
                    const uint32_t x01 =
                        __builtin_amdgcn_perm(mr[1][wd], mr[0][wd], sel);
                    const uint32_t x23 =
                        __builtin_amdgcn_perm(mr[3][wd], mr[2][wd], sel);
                    const uint32_t x =
                        __builtin_amdgcn_perm(x23, x01, 0x05040100u);
                    if constexpr (NIB) {
                        if ((g & 1) == 0)
                            xs[c] = x;
                        else
                            gprof32[c * PW + 2 * gl + (g >> 1)] =
                                xs[c] | (x << 4);
                    } else {
                        dst[c * PW] = x;
                    }
                }
            }
            if (p > 0)
                __threadfence();
            __syncthreads();

            const int rlo = i0 + 2 * R * gl;
            const bool active = rlo < qlen;
            int nact = (qlen - i0 + 2 * R - 1) / (2 * R);
            nact = nact > L ? L : (nact < 0 ? 0 : nact);
            V H[R], E[R], Ho[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int ilo = rlo + r + 1, ihi = rlo + R + r + 1;
                const int hlo = LOCAL ? 0 : open + (ilo - 1) * gmax;
                const int hhi = LOCAL ? 0 : open + (ihi - 1) * gmax;
                H[r] = O::mk(hlo, hhi);
                E[r] = O::splat(NEG);
                Ho[r] = O::add(H[r], OPEN);
            }
            V outH = H[R - 1];
            V outF = LOCAL ? O::splat(NEGF)
                           : (TRACE ? O::fmark(H[R - 1]) : H[R - 1]);
            // inputs of step 1: row above at column t (up) and t-1 (diag)
            int bh0, bh1, bf1;
            if (p == 0) {
                bh0 = 0;
                bh1 = LOCAL ? 0 : open;
                bf1 = NEGF;
            } else {
                bh0 = bnd[0];
                bh1 = bnd[2];
                bf1 = bnd[3];
            }
            V upH = O::template shuffle_up<L>(outH, bh1, fl);
            V upF = O::template shuffle_up<L>(outF, bf1, fl);
            V dg = O::splat(NEG);
            dg = gl == 0 ? O::set_lo(dg, bh0) : dg;
            int c_lo = kSentinel, c_hi = kSentinel;
            const int own_last = nact > 0 ? tlen + 2 * (nact - 1) + 1 : 0;
            int last_step = own_last;   // wave-uniform: longest slot
            if (G > 1) {
#pragma unroll
                for (int q = 1; q < G; ++q)
                    last_step = max(last_step, __shfl(own_last, q * L));
                last_step = max(last_step, __shfl(own_last, 0));
            }
            // cell (qlen, tlen) for the global score
            const int qrow = qlen - 1 - i0;
            const int cap_lane = qrow / (2 * R);
            const int cap_half = (qrow % (2 * R)) / R;
            const int cap_r = qrow % R;
            // values of the slot's task, in scalar registers when one task
            // fills the wave, so that the per-step tests on them are scalar
            auto uni = [](int x) {
                if constexpr (L == 64)
                    return __builtin_amdgcn_readfirstlane(x);
                else
                    return x;
            };
            const int cap_step = uni(tlen + 2 * cap_lane + cap_half);
            const bool last_pass = p == uni(npass_own) - 1;   // this slot's
            const bool more_passes = p + 1 < uni(npass_own);

            // trace record of (step 1, this lane); the next step's record
            // follows it, except after the last step of a segment, which
            // skips the other lanes' runs of that segment
            const int t_first = 2 * gl + 1;
            const int t_last = active ? tlen + 1 + 2 * gl : 0;
            uint8_t PA_GLOBAL *trp =
                TRACE && active ? tr + trace_rec(qlen, tlen, R, L, p, 1, gl) * R
                                : tr;
            const uint32_t pass_lanes =
                TRACE ? (p + 1 < npass_own ? (uint32_t)L : (uint32_t)nact) : 1u;
            uint32_t seg_skip = (pass_lanes - 1u) * kTraceSeg * R;
            if (L == 64)   // one task per wave: keep it in a scalar register
                seg_skip = __builtin_amdgcn_readfirstlane(seg_skip);
            // Column residues and (in later passes) the boundary row reach
            // a task's first lane through v_readlane / a shuffle from
            // registers that are filled a block ahead: lane l holds column
            // base + l (64-lane tasks) or columns base + 4l .. base + 4l + 3
            // (narrower tasks) of a block of RB residues, and column base + l
            // of a block of L boundary values, with base = 1 (mod block).
            // The blocks switch at steps t = 0 (mod RB or L) outside the step
            // loop, so no step waits on global memory.  The boundary row of
            // this pass is written 2L - 1 steps after the column is read
            // (the last lane lags that far), and the look-ahead only reads
            // columns ahead of the step, so it never reads a value of this
            // pass.
            auto res_at = [&](int j) {   // residue of column j (1-based)
                return j <= tlen ? (int)ts[j - 1] : kSentinel;
            };
            auto res_blk = [&](int base) {   // block of RB columns from base
                if (L == 64)
                    return res_at(base + lane);
                const int c = base + 4 * gl;
                return res_at(c) | (res_at(c + 1) << 8) |
                       (res_at(c + 2) << 16) | (res_at(c + 3) << 24);
            };
            int rc_cur = res_blk(1), rc_nxt = res_blk(RB + 1);
            int bh_cur = 0, bf_cur = NEGF, bh_nxt = 0, bf_nxt = NEGF;
            if (p > 0) {
                if (gl + 1 <= tlen) {
                    bh_cur = bnd[2 * (gl + 1)];
                    bf_cur = bnd[2 * (gl + 1) + 1];
                }
                if (L + gl + 1 <= tlen) {
                    bh_nxt = bnd[2 * (L + gl + 1)];
                    bf_nxt = bnd[2 * (L + gl + 1) + 1];
                }
            }
            // Column residues of step tn from those of step tn-1: lane 0
            // takes column tn, the others take lane-1's hi residue of the
            // previous step (two columns back for them); no lane-dependent
            // branch surrounds the DPP move.
            auto next_res = [&](int tn) {
                const int ri = (tn - 1) & (RB - 1);
                int c0;
                if (L == 64) {
                    c0 = __builtin_amdgcn_readlane(rc_cur, ri);
                } else {
                    const int w = __shfl(rc_cur, grp * L + (ri >> 2));
                    c0 = (w >> (8 * (ri & 3))) & 0xff;
                }
                // the first lane of a task has no source lane and keeps the
                // DPP "old" operand, column tn's residue
                const int nc = lane_up<L>(c0, c_hi, fl);
                c_hi = c_lo;
                c_lo = nc;
            };
            // profile words of the current residues: R bytes of the lo and
            // the hi strip, or one word of nibbles each
            auto load_prof = [&](uint32_t *lw, uint32_t *hw) {
                if constexpr (NIB) {
                    lw[0] = gprof32[c_lo * PW + 2 * gl];
                    hw[0] = gprof32[c_hi * PW + 2 * gl + 1];
                } else {
                    ProfWords<R>::load(prof + c_lo * ROWS + 2 * R * gl, lw);
                    ProfWords<R>::load(prof + c_hi * ROWS + 2 * R * gl + R,
                                       hw);
                }
            };
            // residues and profile words of step 1
            next_res(1);
            uint32_t lwA[R / 4], hwA[R / 4], lwB[R / 4], hwB[R / 4];
            load_prof(lwA, hwA);
            // The step loop, specialised for the first pass (boundary row
            // from the gap model) and later passes (boundary row from the
            // previous pass).  Steps run in blocks of L between the block
            // switches, two at a time with the profile words of the next
            // step loaded into the other buffer, so that no register copy at
            // the loop's back edge waits for a load.
            auto run_steps = [&](auto first_pass) {
                constexpr bool FIRST = decltype(first_pass)::value;
                // one step t: lw / hw hold its profile words; the words of
                // step t+1 are loaded into lwn / hwn
                auto step = [&](int t, const uint32_t *lw, const uint32_t *hw,
                                uint32_t *lwn, uint32_t *hwn) {
                    next_res(t + 1);
                    load_prof(lwn, hwn);
                    int nbh = 0, nbf = NEGF;  // boundary for lane 0 at step t+1
                    {
                        const int jn = t + 1;
                        if constexpr (FIRST) {
                            nbh = LOCAL ? 0 : open + (jn - 1) * gmax;
                        } else {
                            const int bi = (jn - 1) & (L - 1);
                            if (L == 64) {
                                if (jn <= tlen) {
                                    nbh = __builtin_amdgcn_readlane(bh_cur, bi);
                                    nbf = __builtin_amdgcn_readlane(bf_cur, bi);
                                }
                            } else {
                                const int xh = __shfl(bh_cur, grp * L + bi);
                                const int xf = __shfl(bf_cur, grp * L + bi);
                                nbh = jn <= tlen ? xh : nbh;
                                nbf = jn <= tlen ? xf : nbf;
                            }
                        }
                    }
                    const int jlo = t - 2 * gl;
                    const int jhi = jlo - 1;
                    // columns 1 .. tlen + 1 of the lo half: steps t_first ..
                    // t_last of this lane
                    if (t >= t_first && t <= t_last) {
                        V s[R];
                        if constexpr (NIB)
                            O::template unpack_nib<TRACE ? 4 : 1>(lw[0], hw[0],
                                                                  nbase, s);
                        else
                            O::template unpack<R>(lw, hw, s);
                        V hdiag = dg, hup = upH, fup = upF;
                        V hoUp = O::add(upH, OPEN);
                        V cm = O::splat(0);
                        V hps[R], efs[R];
    #pragma unroll
                        for (int r = 0; r < R; ++r) {
                            if (!TRACE) {
                                V e = O::max(O::add(E[r], EXT), Ho[r]);
                                V f = O::max(O::add(fup, EXT), hoUp);
                                V hd = O::add(hdiag, s[r]);
                                V h = O::max(O::max(hd, e), f);
                                if (LOCAL)
                                    h = O::max(h, O::splat(0));
                                hdiag = H[r];
                                H[r] = h;
                                E[r] = e;
                                Ho[r] = O::add(h, OPEN);
                                hoUp = Ho[r];
                                hup = h;
                                fup = f;
                                if (LOCAL)
                                    cm = O::max(cm, h);
                            } else {
                                // tags: 3 restart or equal-residue
                                // diagonal, 2 diagonal, 1 F, 0 E
                                V ep = O::max(O::add(E[r], EXTE), Ho[r]);
                                V e = O::andm(ep, ~3);
                                // fup carries tag 01 and hoUp tag 00, so fp
                                // ends in 01 (extended) or 00 (opened)
                                V fp = O::max(O::add(fup, EXT), hoUp);
                                V f1 = O::fmark(fp);
                                V hd = O::add(hdiag, s[r]);
                                V hp = O::max(O::max(hd, f1), e);
                                if (LOCAL)
                                    hp = O::max(hp, O::splat(3));
                                V h = O::andm(hp, ~3);
                                hps[r] = hp;
                                efs[r] = O::efbits(fp, ep);
                                hdiag = H[r];
                                H[r] = h;
                                E[r] = e;
                                Ho[r] = O::add(h, OPEN);
                                hoUp = Ho[r];
                                hup = h;
                                fup = f1;
                                if (LOCAL)
                                    cm = O::max(cm, h);
                            }
                        }
                        outH = hup;
                        outF = fup;
                        if (TRACE) {
                            uint32_t w[R / 4];
                            O::template pack<R>(hps, efs, w);
                            store_rec(trp, w, R / 4);
                        }
                        if (LOCAL) {
                            const bool vlo = jlo <= tlen;
                            const bool vhi = jhi >= 1 && jhi <= tlen;
                            const Key klo = vlo ? O::key(O::lo(cm) >> SH, jlo) : (Key)0;
                            const Key khi = vhi ? O::key(O::hi(cm) >> SH, jhi) : (Key)0;
                            const bool ilo = klo > bk_lo, ihi = khi > bk_hi;
                            bk_lo = ilo ? klo : bk_lo;
                            bk_hi = ihi ? khi : bk_hi;
                            bp_lo = ilo ? p : bp_lo;
                            bp_hi = ihi ? p : bp_hi;
                            if (ilo || ihi) {
    #pragma unroll
                                for (int r = 0; r < R; ++r)
                                    saved[r] = O::blend(saved[r], H[r], ilo, ihi);
                            }
                        } else if (last_pass && t == cap_step) {
                            if (gl == cap_lane) {
    #pragma unroll
                                for (int r = 0; r < R; ++r)
                                    if (r == cap_r)
                                        gscore = cap_half ? O::hi(H[r])
                                                          : O::lo(H[r]);
                            }
                        }
                        if (more_passes) {
                            if (gl == L - 1 && jhi >= 0 && jhi <= tlen) {
                                bnd[2 * jhi] = O::hi(outH);
                                bnd[2 * jhi + 1] = O::hi(outF);
                            }
                        }
                    }
                    dg = upH;
                    upH = O::template shuffle_up<L>(outH, nbh, fl);
                    upF = O::template shuffle_up<L>(outF, nbf, fl);
                    if (TRACE) {
                        trp += R;
                        if (t % (int)kTraceSeg == 0)   // wave-uniform
                            trp += seg_skip;
                    }
                };
                for (int tb = 0; tb <= last_step; tb += L) {
                    if (tb > 0) {
                        // block switch before step tb (tb = 0 mod L): the
                        // residues every RB steps, the boundary row every L
                        if (tb % RB == 0) {
                            rc_cur = rc_nxt;
                            rc_nxt = res_blk(tb + RB + 1);
                        }
                        if constexpr (!FIRST) {
                            bh_cur = bh_nxt;
                            bf_cur = bf_nxt;
                            const int jb = tb + 1 + L + gl;
                            bh_nxt = jb <= tlen ? bnd[2 * jb] : 0;
                            bf_nxt = jb <= tlen ? bnd[2 * jb + 1] : NEGF;
                        }
                    }
                    const int te = min(last_step, tb + L - 1);
                    int t = tb > 0 ? tb : 1;
                    for (; t < te; t += 2) {
                        step(t, lwA, hwA, lwB, hwB);
                        step(t + 1, lwB, hwB, lwA, hwA);
                    }
                    if (t == te) {
                        step(t, lwA, hwA, lwB, hwB);
#pragma unroll
                        for (int k = 0; k < R / 4; ++k) {
                            lwA[k] = lwB[k];
                            hwA[k] = hwB[k];
                        }
                    }
                }
            };
            if (p == 0)
                run_steps(std::true_type{});
            else
                run_steps(std::false_type{});
        }

        // ---- results ----------------------------------------------------
        DpOut o = {};
        if (LOCAL) {
            // first row of each half whose saved value equals its best score
            uint64_t best = 0;
            for (int hf = 0; hf < 2; ++hf) {
                const Key bk = hf ? bk_hi : bk_lo;
                const int bp = hf ? bp_hi : bp_lo;
                const int sc = O::key_score(bk);
                if (sc <= 0)
                    continue;
                const int col = O::key_col(bk);
                const int rbase = bp * ROWS + 2 * R * gl + hf * R;
                int row = -1;
#pragma unroll
                for (int r = R - 1; r >= 0; --r) {
                    const int vv = hf ? O::hi(saved[r]) : O::lo(saved[r]);
                    if ((vv >> SH) == sc && rbase + r < qlen)
                        row = rbase + r;
                }
                if (row < 0)
                    continue;
                const uint64_t ck = ((uint64_t)(uint32_t)sc << 40) |
                                    ((uint64_t)(0xffff - col) << 24) |
                                    (uint64_t)(0xffffff - row);
                best = ck > best ? ck : best;
            }
            best = group_max_u64<L>(best);
            const int sc = (int)(best >> 40);
            if (sc > 0) {
                o.score = sc;
                o.t_end = 0xffff - (int)((best >> 24) & 0xffff);
                o.q_end = 0xffffff - (int)(best & 0xffffff) + 1;
            }
            if (I16) {
                // a saturated task is appended by its slot's first lane (no
                // wave operation depends on the atomic)
                const int lim = TRACE ? 8190 : 32766;
                if (valid && sc >= lim && gl == 0) {
                    const uint32_t pos = atomicAdd(a.sat_count, 1u);
                    a.sat_list[pos] = tid;
                }
            }
        } else {
            // the capturing lane of the slot's last pass holds H(qlen, tlen)
            const int qrow = qlen - 1 - (npass_own - 1) * ROWS;
            const int cl = qrow / (2 * R);
            int g = __shfl(gscore, grp * L + (cl & (L - 1)));
            o.score = g >> SH;
            o.q_end = qlen;
            o.t_end = tlen;
        }
        // all lanes of a slot hold the same result; every lane stores it
        // (same address, same bytes) so no lane-0 branch closes the loop
        o.q_beg = o.q_end;
        o.t_beg = o.t_end;
        if (valid)
            a.out[tid] = o;
    }
}

}  // namespace
}  // namespace pa

// ---------------------------------------------------------------------------
// Traceback over the stored 4-bit codes, one task per thread.  Walks back
// from the end cell (local) or (qlen, tlen) (global) under the same priority
// the DP used, and counts what computeAlignmentStats counts: columns,
// identical residue pairs and maximal gap runs in either row.
// ---------------------------------------------------------------------------
namespace pa {
namespace {


// R and L are fixed per launch (one launch per DP class, whose kernel
// shape wrote the trace), so the layout arithmetic divides by constants.
// Global traces of byte-profile kernels mark a diagonal step between equal
// residues with tag 3 (restart does not occur there), so LOCAL = false needs
// no residue loads; those of 4-bit-profile kernels (rows_nib) tag every
// diagonal 2 and the walk compares the residues, as local walks do.
//
// STAGE: the walk reads the trace one lane run at a time: the kTraceSeg
// records of one lane in one segment (8R contiguous bytes, 8R-aligned) are
// loaded with wide loads into the thread's slot in LDS when the walk enters
// them, and the codes are read from there.  Diagonal and horizontal moves
// step back one record inside the run, vertical moves stay in the record,
// so a run serves several steps.  Without STAGE each step loads its byte
// from global memory.

// Residues of one sequence as the walk reads them (backwards, mostly one
// position per diagonal step): the aligned 8-byte word holding the last one
// read is kept, so a word load serves several steps.  The sequence store is
// allocated with slack behind its last residue.
struct ResWin {
    const uint8_t PA_GLOBAL *base;
    uintptr_t word = 1;   // address of the word held (1: none)
    uint64_t v = 0;
    __device__ inline int at(int i)
    {
        const uintptr_t p = (uintptr_t)(base + i);
        const uintptr_t wa = p & ~(uintptr_t)7;
        if (wa != word) {
            word = wa;
            v = *(const uint64_t PA_GLOBAL *)wa;
        }
        return (int)((v >> (8u * (uint32_t)(p & 7))) & 0xffu);
    }
};

template <int R, int L, bool LOCAL, bool STAGE>
__global__ __launch_bounds__(kTbThreads) void k_traceback(TbArgs a)
{
    constexpr int RUNW = 2 * R;   // 32-bit words of one lane run (8R bytes)
    // one run per thread; the odd stride spreads threads over the banks
    __shared__ uint32_t stage[STAGE ? kTbThreads * (RUNW + 1) : 1];
    uint32_t *run = stage + (STAGE ? threadIdx.x * (RUNW + 1) : 0);
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *a.count)
        return;
    const uint32_t tid = glb(a.list)[k];
    const DpTask T = a.tasks[tid];
    DpOut o = a.out[tid];
    const uint8_t PA_GLOBAL *qs = glb_addr<const uint8_t>(T.q_off);
    const uint8_t PA_GLOBAL *ts = glb_addr<const uint8_t>(T.t_off);
    const int tlen = (int)T.tlen;
    const uint8_t *trg =
        a.trace + (glb(a.trace_off)[tid] -
                   (a.trace_base ? *glb(a.trace_base) : 0ull));
    // Position of row i-1 in the trace layout: pass p, lane l, half h, row r
    // of the half; kept up to date as the walk moves up, so a step costs a
    // few adds.  Byte offsets within one task's trace fit in 32 bits.
    constexpr int ROWS = 2 * R * L;
    constexpr uint32_t RUN = kTraceSeg * R;   // one lane's run of a segment
    const uint32_t qlen = T.qlen;
    const uint32_t full = qlen ? (qlen - 1) / ROWS : 0;
    const uint32_t last_lanes = qlen ? trace_last_lanes(qlen, R, L) : 0;
    const uint32_t pass_bytes =
        trace_seg_steps((uint32_t)tlen + 2u * L - 1u) * (uint32_t)L * R;
    int p = 0, l = 0, h = 0, r = 0;
    uint32_t pbase = 0, seg_bytes = 0;   // of pass p
    auto set_pass = [&]() {
        pbase = (uint32_t)p * pass_bytes;
        seg_bytes = ((uint32_t)p < full ? (uint32_t)L : last_lanes) * RUN;
    };
    int i = o.q_end, j = o.t_end;
    if (i > 0) {
        const int i0 = i - 1, rr = i0 % ROWS, q = rr % (2 * R);
        p = i0 / ROWS;
        l = rr / (2 * R);
        h = q / R;
        r = q % R;
        set_pass();
    }
    // moves the position to row i-2 (the walk's next row)
    auto up = [&]() {
        if (--r < 0) {
            r = R - 1;
            if (--h < 0) {
                h = 1;
                if (--l < 0) {
                    l = L - 1;
                    if (--p >= 0)
                        set_pass();
                }
            }
        }
    };
    uint32_t cur_run = 0xffffffffu;   // byte offset of the staged run
    ResWin qwin{qs}, twin{ts};
    int matches = 0, cols = 0, runs = 0;
    int last = 0;   // 1 diagonal, 2 vertical, 3 horizontal
    int state = 0;  // 0 H, 1 E (horizontal), 2 F (vertical)
    if (LOCAL && o.score <= 0) {
        i = j = 0;
    } else {
        for (;;) {
            if (i == 0 || j == 0) {
                if (LOCAL)
                    break;
                // global boundary: a single gap run to the origin
                if (i > 0) {
                    cols += i;
                    runs += last != 2;
                    i = 0;
                } else if (j > 0) {
                    cols += j;
                    runs += last != 3;
                    j = 0;
                }
                break;
            }
            // record of (step t, lane l); row r, half h of a record sits in
            // byte 4*(r/4) + 2*(r%2) + h, high nibble when (r/2) is odd
            const uint32_t t1 = (uint32_t)(j + 2 * l + h) - 1u;
            const uint32_t run_off =
                pbase + (t1 / kTraceSeg) * seg_bytes + (uint32_t)l * RUN;
            const uint32_t bi = (t1 % kTraceSeg) * R + 4u * (r >> 2) +
                                2u * (r & 1) + (uint32_t)h;
            uint32_t byte;
            if (STAGE) {
                if (run_off != cur_run) {
                    cur_run = run_off;
                    // generic pointer: vector-struct loads from the global
                    // address space do not compile
                    const uint4 *src = (const uint4 *)(trg + run_off);
#pragma unroll
                    for (int q = 0; q < RUNW / 4; ++q) {
                        const uint4 x = src[q];
                        run[4 * q + 0] = x.x;
                        run[4 * q + 1] = x.y;
                        run[4 * q + 2] = x.z;
                        run[4 * q + 3] = x.w;
                    }
                }
                byte = (run[bi >> 2] >> (8u * (bi & 3u))) & 0xffu;
            } else {
                byte = glb(trg)[run_off + bi];
            }
            const int c = (int)((r & 2) ? (byte >> 4) : (byte & 15u));
            if (state == 0) {
                const int src = c & 3;
                if (LOCAL && src == 3)
                    break;  // restart
                if (src >= 2) {
                    ++cols;
                    if (LOCAL || rows_nib(R))
                        matches += qwin.at(i - 1) == twin.at(j - 1);
                    else
                        matches += src == 3;
                    last = 1;
                    --i;
                    up();
                    --j;
                } else {
                    state = src == 1 ? 2 : 1;
                }
            } else if (state == 1) {
                ++cols;
                runs += last != 3;
                last = 3;
                state = (c >> 3) & 1 ? 1 : 0;   // E extended
                --j;
            } else {
                ++cols;
                runs += last != 2;
                last = 2;
                state = (c >> 2) & 1 ? 2 : 0;   // F extended
                --i;
                up();
            }
        }
    }
    o.q_beg = i;
    o.t_beg = j;
    o.matches = matches;
    o.cols = cols;
    o.gap_opens = runs;
    a.out[tid] = o;
}

template <int R, int L, bool LC, bool T, bool I>
hipError_t launch_one(const DpArgs &a, int grid, hipStream_t s)
{
    // G task profiles of 2RL rows each (128R rows in all), bytes or nibbles
    const size_t lds = (size_t)kProfChars * (rows_nib(R) ? 64 : 128) * R;
    hipLaunchKernelGGL((k_dp<R, L, LC, T, I>), dim3(grid), dim3(64), lds, s,
                       a);
    return hipGetLastError();
}

template <int R, int L, bool LC, bool T, bool I>
int occupancy_one()
{
    const size_t lds = (size_t)kProfChars * (rows_nib(R) ? 64 : 128) * R;
    int n = 0;
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_dp<R, L, LC, T, I>, 64,
                                                 lds);
    return n > 0 ? n : 1;
}

template <int R, int L>
int occupancy_c(bool local, bool trace, bool i16)
{
    if (local) {
        if (trace)
            return i16 ? occupancy_one<R, L, true, true, true>()
                       : occupancy_one<R, L, true, true, false>();
        return i16 ? occupancy_one<R, L, true, false, true>()
                   : occupancy_one<R, L, true, false, false>();
    }
    return i16 ? occupancy_one<R, L, false, true, true>()
               : occupancy_one<R, L, false, true, false>();
}

template <int R, int L>
hipError_t launch_c(bool local, bool trace, bool i16, const DpArgs &a,
                    int grid, hipStream_t s)
{
    if (local) {
        if (trace)
            return i16 ? launch_one<R, L, true, true, true>(a, grid, s)
                       : launch_one<R, L, true, true, false>(a, grid, s);
        return i16 ? launch_one<R, L, true, false, true>(a, grid, s)
                   : launch_one<R, L, true, false, false>(a, grid, s);
    }
    return i16 ? launch_one<R, L, false, true, true>(a, grid, s)
               : launch_one<R, L, false, true, false>(a, grid, s);
}

template <int R, int L>
hipError_t launch_traceback_c(const TbArgs &a, bool stage, int grid,
                              hipStream_t s)
{
    const dim3 g(grid), b(kTbThreads);
    if (a.local) {
        if (stage)
            hipLaunchKernelGGL((k_traceback<R, L, true, true>), g, b, 0, s, a);
        else
            hipLaunchKernelGGL((k_traceback<R, L, true, false>), g, b, 0, s, a);
    } else {
        if (stage)
            hipLaunchKernelGGL((k_traceback<R, L, false, true>), g, b, 0, s, a);
        else
            hipLaunchKernelGGL((k_traceback<R, L, false, false>), g, b, 0, s,
                               a);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_dp(int shape, bool local, bool trace, bool i16,
                     const DpArgs &a, int grid, hipStream_t s)
{
    switch (shape) {
    case kShapeL16R4: return launch_c<4, 16>(local, trace, i16, a, grid, s);
    case kShapeL16R8: return launch_c<8, 16>(local, trace, i16, a, grid, s);
    case kShapeL64R4: return launch_c<4, 64>(local, trace, i16, a, grid, s);
    case kShapeL64R8: return launch_c<8, 64>(local, trace, i16, a, grid, s);
    case kShapeL32R8: return launch_c<8, 32>(local, trace, i16, a, grid, s);
    case kShapeL8R8: return launch_c<8, 8>(local, trace, i16, a, grid, s);
    default: return hipErrorInvalidValue;
    }
}

int dp_occupancy(int shape, bool local, bool trace, bool i16)
{
    static int cache[kShapes][8] = {};
    const int vi = (local ? 4 : 0) + (trace ? 2 : 0) + (i16 ? 1 : 0);
    if (shape < 0 || shape >= kShapes)
        return 1;
    if (!cache[shape][vi])
        cache[shape][vi] =
            shape == kShapeL16R4   ? occupancy_c<4, 16>(local, trace, i16)
            : shape == kShapeL16R8 ? occupancy_c<8, 16>(local, trace, i16)
            : shape == kShapeL64R4 ? occupancy_c<4, 64>(local, trace, i16)
            : shape == kShapeL64R8 ? occupancy_c<8, 64>(local, trace, i16)
            : shape == kShapeL32R8 ? occupancy_c<8, 32>(local, trace, i16)
                                   : occupancy_c<8, 8>(local, trace, i16);
    return cache[shape][vi];
}

hipError_t launch_traceback(const TbArgs &a, int shape, bool stage, int grid,
                            hipStream_t s)
{
    switch (shape) {
    case kShapeL16R4: return launch_traceback_c<4, 16>(a, stage, grid, s);
    case kShapeL16R8: return launch_traceback_c<8, 16>(a, stage, grid, s);
    case kShapeL64R4: return launch_traceback_c<4, 64>(a, stage, grid, s);
    case kShapeL64R8: return launch_traceback_c<8, 64>(a, stage, grid, s);
    case kShapeL32R8: return launch_traceback_c<8, 32>(a, stage, grid, s);
    case kShapeL8R8: return launch_traceback_c<8, 8>(a, stage, grid, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pa
