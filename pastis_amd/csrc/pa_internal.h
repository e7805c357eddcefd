// pa_internal.h -- structures shared by the HIP kernels and the host side of
// libpastis_hip.so.  Not part of the public ABI (include/pastis_hip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pastis_hip.h"

namespace pa {

// Pointers into global memory.  Pointers that reach a kernel inside an
// argument struct or as a stored address are generic, and generic accesses
// compile to flat instructions, which also count against the LDS wait
// counter; casting them to the global address space gives global_load /
// global_store instead.
#ifndef PA_GLOBAL   // the CPU emulation build of the tests defines it empty
#define PA_GLOBAL __attribute__((address_space(1)))
#endif
// Waves per SIMD a kernel's register allocation must allow.
#ifndef PA_WAVES_PER_EU   // the CPU emulation build of the tests: empty
#define PA_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif
// Makes a wave-uniform integer opaque to the optimiser (no instruction): it
// is then kept in a scalar register and not folded into the code using it.
#ifndef PA_OPAQUE   // the CPU emulation build: a general register
#define PA_OPAQUE(x) asm volatile("" : "+s"(x))
#endif
template <class T> __device__ inline T PA_GLOBAL *glb(T *p)
{
    return (T PA_GLOBAL *)p;
}
template <class T> __device__ inline T PA_GLOBAL *glb_addr(uint64_t a)
{
    return (T PA_GLOBAL *)(uintptr_t)a;
}

// Residue codes 0..24 are real letters; code 25 is the sentinel used for the
// DP's column 0 and for columns past the end of the streamed sequence.  Its
// profile entries are -128, strictly below any real score.
constexpr int kAlpha = 25;
constexpr int kSentinel = 25;
constexpr int kProfChars = 26;

// One DP task: the row (V, lanes) and column (H, streamed) sequence spans.
struct DpTask {
    uint64_t q_off;   // first residue of the row span in the sequence store
    uint64_t t_off;   // first residue of the column span
    uint32_t qlen;    // rows
    uint32_t tlen;    // columns
    uint32_t id;      // output slot
    uint32_t flags;
};
static_assert(sizeof(DpTask) == 32, "DpTask layout");

// Output of a DP task: score and end cell (1-based DP coordinates, i.e. the
// exclusive 0-based end), plus traceback statistics when traced.
struct DpOut {
    int32_t score;
    int32_t q_end, t_end;     // rows / columns consumed at the end cell
    int32_t q_beg, t_beg;     // traceback start (filled by the traceback)
    int32_t matches, cols, gap_opens;
};
static_assert(sizeof(DpOut) == 32, "DpOut layout");

// Task flags
constexpr uint32_t kTaskSat = 1u;  // int16 saturated, needs the int32 re-run

// Profile matrices: 32x32 int8, row = row-sequence code, column = column
// code, -128 outside the 25x25 alphabet.  The DP kernels build their query
// profiles from the one of their mode: the score (untraced), 4*score+2
// (traced local) or 4*score+2, +1 on the diagonal (traced global: a
// diagonal step between equal residues carries tag 3).  Byte offsets inside
// the device copy of the scoring matrix, after the plain 32x32 table.
//
// The kernels with 8 rows per lane (rows_nib) hold 4-bit query profiles:
// kPmatNib gives score - nib_min in 0..15 (0 outside the alphabet), and
// the kernel adds nib_min (untraced) or 4*nib_min + 2 (traced) back; their
// global traces tag every diagonal step 2 and the traceback compares the
// residues for identities.  They need a matrix whose scores over the 25x25
// alphabet span at most 16 values with a negative minimum (BLOSUM62: -4 ..
// 11); with any other matrix the context runs their 4-row counterparts.
constexpr size_t kPmatUntraced = 1024, kPmatLocal = 2048, kPmatGlobal = 3072,
                 kPmatNib = 4096;
constexpr size_t kPmatBytes = 5120;
// Global task whose scores are bounded by the residue bound below (no cell
// exceeds min over the two spans of the sum of each residue's best score),
// small enough for the traced int16 kernels.
constexpr uint32_t kTaskBounded = 2u;

// Residue bound of a sequence store: pre[k] = sum over codes[0..k) of
// max(0, best score of that code against any code), as uint32 with
// wrap-around (differences over one sequence's span stay exact).
struct ResBound {
    uint64_t base;          // device address of the store's first residue
    const uint32_t *pre;    // prefix sums, one more entry than residues
};

// Arguments of the wavefront DP kernels.
struct DpArgs {
    const uint8_t *codes;      // sequence store
    const DpTask *tasks;
    const uint32_t *list;      // task indices of this launch
    const uint32_t *count;     // device: number of entries in list
    uint32_t *queue;           // device: work-queue head (zeroed before launch)
    const int8_t *mat;         // 32x32 scores
    const uint32_t *pmat;      // profile matrix of the mode (kPmat* below)
    DpOut *out;
    uint8_t *trace;            // trace store (traced kernels)
    const uint64_t *trace_off; // per task byte offset into trace
    const uint64_t *trace_base;// device: offset of the chunk's first task
                               // (chunks sized on the GPU), or null
    int32_t *bnd;              // per-workgroup boundary rows (multi-pass)
    uint32_t bnd_stride;       // int32 entries per workgroup
    uint32_t *sat_list;        // tasks that saturated (int16 kernels)
    uint32_t *sat_count;
    int32_t gap_open, gap_ext;
    uint32_t chunk;            // tasks claimed per queue atomic (>= 1)
    int32_t nib_min;           // 4-bit profiles: the score of nibble 0
};

// Arguments of the traceback kernel.
struct TbArgs {
    const uint8_t *codes;
    const DpTask *tasks;
    const uint32_t *list;
    const uint32_t *count;
    const uint8_t *trace;
    const uint64_t *trace_off;
    const uint64_t *trace_base;   // as DpArgs::trace_base
    DpOut *out;
    int local;
};

// DP kernel shapes: L lanes per task (G = 64 / L tasks per wave, one per
// L-lane group) and R rows per lane per half, so a pass covers 2RL rows and
// longer tasks take several passes linked by a boundary row.  Every shape
// runs any row count; the kernels are compiled for these six.
constexpr int kShapes = 6;
constexpr int kShapeL16R4 = 0, kShapeL16R8 = 1, kShapeL64R4 = 2,
              kShapeL64R8 = 3, kShapeL32R8 = 4, kShapeL8R8 = 5;
__host__ __device__ constexpr int shape_rows(int sh)
{
    return sh == kShapeL16R4 || sh == kShapeL64R4 ? 4 : 8;
}
__host__ __device__ constexpr int shape_lanes(int sh)
{
    return sh == kShapeL64R4 || sh == kShapeL64R8 ? 64
           : sh == kShapeL32R8                    ? 32
           : sh == kShapeL8R8                     ? 8
                                                  : 16;
}
// task slots (boundary rows) of a one-wave DP workgroup
__host__ __device__ constexpr int shape_slots(int sh)
{
    return 64 / shape_lanes(sh);
}
// 4-bit query profiles: the shapes with 8 rows per lane
__host__ __device__ constexpr bool rows_nib(int R) { return R == 8; }
// the shape that replaces a 4-bit-profile shape when the matrix does not
// fit 4 bits (same rows per pass where one exists)
__host__ __device__ constexpr int shape_byte_fallback(int sh)
{
    return sh == kShapeL16R8   ? kShapeL16R4
           : sh == kShapeL64R8 ? kShapeL64R4
           : sh == kShapeL32R8 ? kShapeL64R4
           : sh == kShapeL8R8  ? kShapeL16R4
                               : sh;
}

// DP row classes: a task's class is the first c with qlen <= edge[c] (class
// 3 takes the rest), and each class has its kernel shape.  Narrow groups
// keep more of a wave busy on short tasks (the wavefront's fill and drain
// take 2L steps per pass and the last pass rounds the rows up to 2RL); wide
// ones cost fewer boundary rows, profile builds and loop overhead per cell
// on long tasks.  The defaults (kDefaultShapes) are
//   0: qlen <= 128: 16 lanes x 4 rows, 128 rows per pass (one pass);
//   1: qlen <= 256: 16 lanes x 8 rows, 256 rows per pass (one pass);
//   2: qlen <= 512: 64 lanes x 4 rows, 512 rows per pass (one pass);
//   3: longer:      64 lanes x 8 rows, 1024 rows per pass.
struct DpShapes {
    uint32_t edge[3];
    uint8_t shape[4];
};
constexpr DpShapes kDefaultShapes = {{128, 256, 512},
                                     {kShapeL8R8, kShapeL16R8, kShapeL32R8,
                                      kShapeL64R8}};

__host__ __device__ inline int dp_class(uint32_t qlen, const DpShapes &sh)
{
    return qlen <= sh.edge[0] ? 0
           : qlen <= sh.edge[1] ? 1
           : qlen <= sh.edge[2] ? 2
                                : 3;
}
__host__ __device__ inline int rows_per_lane(uint32_t qlen, const DpShapes &sh)
{
    return shape_rows(sh.shape[dp_class(qlen, sh)]);
}
__host__ __device__ inline int lanes_per_task(uint32_t qlen, const DpShapes &sh)
{
    return shape_lanes(sh.shape[dp_class(qlen, sh)]);
}

// Trace layout of a task traced with R rows per lane and L lanes: pass after
// pass; a pass holds one record of R bytes (2R 4-bit codes) per (step,
// active lane).  Full passes have L active lanes and tlen + 2L - 1 steps;
// the last pass has nact lanes and tlen + 2*(nact-1) + 1 steps.  Steps are
// grouped in segments of kTraceSeg: a segment holds, lane after lane, the
// kTraceSeg consecutive records of each lane, so a traceback that walks
// back along a lane reads neighbouring bytes (8R contiguous bytes per lane
// and segment) while one step of the DP still writes L records in a few
// cache lines.  A pass is padded to whole segments.
constexpr uint32_t kTraceSeg = 8;

// Threads (one traceback walk each) per workgroup of the traceback kernels.
constexpr int kTbThreads = 256;

__host__ __device__ inline uint32_t trace_last_lanes(uint32_t qlen, int R,
                                                     int L)
{
    const uint32_t rows = 2u * (uint32_t)R * (uint32_t)L;
    const uint32_t rem = qlen - ((qlen - 1) / rows) * rows;
    return (rem + 2u * R - 1u) / (2u * R);
}

// steps of a pass rounded up to whole segments
__host__ __device__ inline uint32_t trace_seg_steps(uint32_t steps)
{
    return (steps + kTraceSeg - 1u) / kTraceSeg * kTraceSeg;
}

__host__ __device__ inline uint64_t trace_bytes(uint32_t qlen, uint32_t tlen,
                                                int R, int L)
{
    const uint32_t rows = 2u * (uint32_t)R * (uint32_t)L;
    const uint32_t full = (qlen - 1) / rows;
    const uint32_t nl = trace_last_lanes(qlen, R, L);
    return ((uint64_t)full * trace_seg_steps(tlen + 2u * L - 1u) * L +
            (uint64_t)trace_seg_steps(tlen + 2u * (nl - 1u) + 1u) * nl) *
           (uint64_t)R;
}

// Record index (in units of R bytes) of (pass p, step t >= 1, lane).
__host__ __device__ inline uint64_t trace_rec(uint32_t qlen, uint32_t tlen,
                                              int R, int L, uint32_t p,
                                              uint32_t t, uint32_t lane)
{
    const uint32_t rows = 2u * (uint32_t)R * (uint32_t)L;
    const uint32_t full = (qlen - 1) / rows;
    const uint32_t nact = p < full ? (uint32_t)L : trace_last_lanes(qlen, R, L);
    const uint32_t seg = (t - 1u) / kTraceSeg, off = (t - 1u) % kTraceSeg;
    return (uint64_t)p * trace_seg_steps(tlen + 2u * L - 1u) * L +
           (uint64_t)seg * nact * kTraceSeg + lane * kTraceSeg + off;
}

// X-drop extension task: one seed of one pair.
struct XdTask {
    uint64_t h_off, v_off;     // column (H) and row (V) sequence starts
    uint32_t nh, nv;           // lengths
    uint32_t beg_h, beg_v;     // seed start
    uint32_t id;               // output slot
    uint32_t pad;              // kXdInvalid: no extension, empty extents
};
static_assert(sizeof(XdTask) == 40, "XdTask layout");
// The pair names a sequence that is not loaded or its seed does not fit.
constexpr uint32_t kXdInvalid = 1u;

// Seed extension result; the left direction writes beg_*, score_left,
// cells_left, flags_left and the right direction the other half.
struct XdOut {
    int32_t beg_h, beg_v, end_h, end_v;
    int32_t score_left, score_right;
    uint32_t cells_left, cells_right;
    uint16_t flags_left, flags_right;  // 1 band too wide, 2 budget exhausted
    uint32_t pad;
};
static_assert(sizeof(XdOut) == 40, "XdOut layout");

// Launchers (k_dp.hip, k_xdrop.hip, k_misc.hip).
hipError_t launch_dp(int shape, bool local, bool trace, bool i16,
                     const DpArgs &a, int grid, hipStream_t s);
// Traceback of tasks traced by the DP kernels of one shape; stage: copy each
// lane run (8R bytes) the walk enters into LDS with wide loads instead of
// loading one byte per step.
hipError_t launch_traceback(const TbArgs &a, int shape, bool stage, int grid,
                            hipStream_t s);
// X-drop with one 16-lane row of a wave per work item (k_xdrop_coop.hip);
// items whose band outgrows 16 columns are appended to ovf_list.
hipError_t launch_xdrop_coop(const XdTask *tasks, const uint32_t *list,
                             const uint32_t *count, uint32_t *queue,
                             const int8_t *mat, int gap, int xdrop,
                             int seed_len, XdOut *out, uint32_t *ovf_list,
                             uint32_t *ovf_count, int grid, hipStream_t s);
int xdrop_coop_occupancy();
// X-drop with 16 lanes per work item (k_xdrop_wave.hip), for the items whose
// band outgrew the per-lane kernel's LDS rings: a 128-column window, or
// 256 columns with `wide`; items whose band outgrows the window are appended
// to ovf_list.
hipError_t launch_xdrop_wave(const XdTask *tasks, const uint32_t *list,
                             const uint32_t *count, uint32_t *queue,
                             const int8_t *mat, int gap, int xdrop,
                             int seed_len, XdOut *out, uint32_t *ovf_list,
                             uint32_t *ovf_count, int grid, int wide,
                             hipStream_t s);
int xdrop_wave_occupancy(int wide);

// Workgroups of one DP kernel variant (of a kernel shape) that fit on one
// CU (occupancy).
int dp_occupancy(int shape, bool local, bool trace, bool i16);
// Same for the LDS-ring X-drop kernel with `ring` slots per lane (the
// variant launch_xdrop picks for these X-drop and gap values).
int xdrop_occupancy(int ring, int xdrop, int gap);

hipError_t launch_xdrop(const XdTask *tasks, const uint32_t *list,
                        const uint32_t *count, uint32_t *queue,
                        const int8_t *mat, int gap, int xdrop, int seed_len,
                        int32_t *scratch, int cap, XdOut *out,
                        uint32_t *ovf_list, uint32_t *ovf_count, int grid,
                        int ring, hipStream_t s);

}  // namespace pa
