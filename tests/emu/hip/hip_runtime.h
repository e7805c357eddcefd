// CPU emulation of the HIP subset used by pastis_amd/csrc -- TEST ONLY.
//
// Compiles the library's kernel and host sources for the CPU so that hangs,
// out-of-bounds accesses (under AddressSanitizer) and parity bugs can be
// found without a GPU.  Each workgroup runs as blockDim.x std::threads; the
// wave64 collectives the kernels use (DPP wave_shr:1, readfirstlane, shfl,
// ballot) exchange values through per-wave slots between two barriers, which
// matches the hardware as long as every lane of the wave reaches the
// collective (true for these kernels: collectives sit outside divergent
// code).  Workgroups run one after another.  Never linked into the product.
#pragma once

#include <atomic>
#include <cmath>
#include <barrier>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __launch_bounds__(...)
#define __shared__ static
#define __constant__ static const
#define PA_GLOBAL   // host memory has one address space
#define PA_WAVES_PER_EU(n)   // no register allocation to constrain
#define PA_OPAQUE(x) asm volatile("" : "+r"(x))
#define __forceinline__ inline
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu::dyn_shared();
#define __align__(n) __attribute__((aligned(n)))
// a device variable is a host variable here: its "device address" is its
// address
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint2 {
    unsigned x, y;
};
inline uint2 make_uint2(unsigned a, unsigned b) { return {a, b}; }
struct uint4 {
    unsigned x, y, z, w;
};

typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
constexpr hipError_t hipErrorInvalidValue = 1;
constexpr hipError_t hipErrorOutOfMemory = 2;
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
enum hipMemcpyKind {
    hipMemcpyHostToDevice,
    hipMemcpyDeviceToHost,
    hipMemcpyDeviceToDevice,
    hipMemcpyHostToHost
};
constexpr unsigned hipStreamNonBlocking = 1;
struct hipDeviceProp_t {
    char gcnArchName[256];
    int multiProcessorCount;
};

namespace emu {
extern thread_local dim3 tIdx, bIdx, bDim, gDim;
void *dyn_shared();
uint64_t wave_exchange(uint64_t v, int src_lane, bool xor_mode);
uint64_t wave_ballot(bool p);
void block_sync();
void launch_impl(dim3 g, dim3 b, size_t shmem, void (*body)(void *),
                 void *arg);

template <class F> void trampoline(void *p) { (*(F *)p)(); }

template <class K, class... A>
void launch(K kernel, dim3 g, dim3 b, size_t shmem, A... args)
{
    auto f = [=]() { kernel(args...); };
    launch_impl(g, b, shmem, &trampoline<decltype(f)>, (void *)&f);
}
}  // namespace emu

#define threadIdx emu::tIdx
#define blockIdx emu::bIdx
#define blockDim emu::bDim
#define gridDim emu::gDim

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
    emu::launch(kernel, dim3(grid), dim3(block), (size_t)(shmem), __VA_ARGS__)

inline void __syncthreads() { emu::block_sync(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { __threadfence(); }

template <class T> T atomicAdd(T *p, T v)
{
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}

template <class T> inline T emu_bits_in(uint64_t x)
{
    T r;
    std::memcpy(&r, &x, sizeof(T));
    return r;
}
template <class T> inline uint64_t emu_bits_out(T v)
{
    uint64_t x = 0;
    std::memcpy(&x, &v, sizeof(T));
    return x;
}
template <class T> T __shfl(T v, int src)
{
    return emu_bits_in<T>(emu::wave_exchange(emu_bits_out(v), src & 63, false));
}
template <class T> T __shfl_xor(T v, int m)
{
    return emu_bits_in<T>(emu::wave_exchange(emu_bits_out(v), m, true));
}
inline uint64_t __ballot(bool p) { return emu::wave_ballot(p); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline float __uint_as_float(unsigned x)
{
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}
inline unsigned __float_as_uint(float f)
{
    unsigned x;
    std::memcpy(&x, &f, 4);
    return x;
}
// host fp32 arithmetic rounds to nearest even; volatile keeps each
// operation rounded on its own (no contraction)
inline float __fadd_rn(float a, float b) { volatile float r = a + b; return r; }
inline float __fsub_rn(float a, float b) { volatile float r = a - b; return r; }
inline float __fmul_rn(float a, float b) { volatile float r = a * b; return r; }
inline float __fdiv_rn(float a, float b) { volatile float r = a / b; return r; }
inline float __fsqrt_rn(float a) { return std::sqrt(a); }
inline float __int_as_float(int x)
{
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// DPP patterns the kernels use: wave_shr:1 (0x138), row_shr:1 (0x111) and
// row_ror:k
// (0x121..0x12f, lane j of a 16-lane row reads lane (j - k) mod 16).
inline int emu_update_dpp(int old, int src, int ctrl, int, int, bool)
{
    const int lane = (int)(emu::tIdx.x & 63);
    if (ctrl == 0x138) {
        uint64_t v = emu::wave_exchange((uint32_t)src, lane - 1, false);
        return lane == 0 ? old : (int)(uint32_t)v;
    }
    if (ctrl == 0x111) {   // row_shr:1, the first lane of a row keeps old
        uint64_t v = emu::wave_exchange((uint32_t)src, lane - 1, false);
        return (lane & 15) == 0 ? old : (int)(uint32_t)v;
    }
    if (ctrl > 0x120 && ctrl <= 0x12f) {
        const int k = ctrl - 0x120;
        const int srcl = (lane & ~15) | ((lane - k) & 15);
        return (int)(uint32_t)emu::wave_exchange((uint32_t)src, srcl, false);
    }
    std::abort();
}
#define __builtin_amdgcn_update_dpp emu_update_dpp
// DPP move without an "old" operand (used for patterns where every lane has
// a source)
inline int emu_mov_dpp(int src, int ctrl, int rm, int bm, bool bc)
{
    return emu_update_dpp(0, src, ctrl, rm, bm, bc);
}
#define __builtin_amdgcn_mov_dpp emu_mov_dpp
inline int max(int a, int b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline unsigned emu_readfirstlane(unsigned v)
{
    return (unsigned)emu::wave_exchange(v, 0, false);
}
#define __builtin_amdgcn_readfirstlane emu_readfirstlane
inline int emu_readlane(int v, int idx)
{
    return (int)(uint32_t)emu::wave_exchange((uint32_t)v, idx & 63, false);
}
#define __builtin_amdgcn_readlane emu_readlane
inline unsigned emu_perm(unsigned s0, unsigned s1, unsigned sel)
{
    const uint64_t src = ((uint64_t)s0 << 32) | s1;
    unsigned out = 0;
    for (int i = 0; i < 4; ++i) {
        const unsigned s = (sel >> (8 * i)) & 0xff;
        unsigned b;
        if (s < 8)
            b = (unsigned)(src >> (8 * s)) & 0xff;
        else if (s < 12)
            b = ((src >> (16 * (s - 8) + 15)) & 1) ? 0xff : 0;
        else if (s == 12)
            b = 0;
        else
            b = 0xff;
        out |= b << (8 * i);
    }
    return out;
}
#define __builtin_amdgcn_perm emu_perm
inline unsigned emu_alignbit(unsigned a, unsigned b, unsigned s)
{
    return (unsigned)((((uint64_t)a << 32) | b) >> (s & 31));
}
#define __builtin_amdgcn_alignbit emu_alignbit

// runtime
hipError_t hipMalloc(void **p, size_t n);
template <class T> inline hipError_t hipMalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, n);
}
hipError_t hipFree(void *p);
#define hipHostMallocDefault 0u
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind,
                          hipStream_t);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipStreamSynchronize(hipStream_t);
hipError_t hipStreamCreateWithFlags(hipStream_t *, unsigned);
hipError_t hipStreamDestroy(hipStream_t);
hipError_t hipEventCreate(hipEvent_t *);
#define hipEventDisableTiming 2u
hipError_t hipEventCreateWithFlags(hipEvent_t *, unsigned);
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned);
hipError_t hipEventRecord(hipEvent_t, hipStream_t);
hipError_t hipEventSynchronize(hipEvent_t);
hipError_t hipEventElapsedTime(float *, hipEvent_t, hipEvent_t);
hipError_t hipEventDestroy(hipEvent_t);
hipError_t hipGetDeviceCount(int *);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *, int);
// graph capture is not emulated: the calls report failure
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum { hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return 1; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    *g = nullptr;
    return 1;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *,
                                      void *, size_t)
{
    return 1;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return 1; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return 0; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return 0; }
template <class F>
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t)
{
    *n = 2;
    return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void **p, const void *sym)
{
    *p = const_cast<void *>(sym);
    return hipSuccess;
}
hipError_t hipSetDevice(int);
hipError_t hipMemGetInfo(size_t *, size_t *);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t);
