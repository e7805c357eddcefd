// CPU emulation runtime for tests/emu/hip/hip_runtime.h -- TEST ONLY.
#include "hip/hip_runtime.h"

#include <cstdio>
#include <mutex>

namespace emu {
thread_local dim3 tIdx, bIdx, bDim, gDim;

namespace {
struct Wave {
    uint64_t slot[64];
    std::unique_ptr<std::barrier<>> bar;
};
std::vector<uint8_t> g_dyn;
std::vector<Wave> g_waves;
std::unique_ptr<std::barrier<>> g_block;
std::mutex g_launch;
}  // namespace

void *dyn_shared() { return g_dyn.data(); }

uint64_t wave_exchange(uint64_t v, int src, bool xor_mode)
{
    const int lane = (int)(tIdx.x & 63), w = (int)(tIdx.x / 64);
    Wave &W = g_waves[w];
    W.slot[lane] = v;
    W.bar->arrive_and_wait();
    const int s = xor_mode ? (lane ^ src) : src;
    const uint64_t r = (s >= 0 && s < 64) ? W.slot[s] : 0;
    W.bar->arrive_and_wait();
    return r;
}

uint64_t wave_ballot(bool p)
{
    const int lane = (int)(tIdx.x & 63), w = (int)(tIdx.x / 64);
    Wave &W = g_waves[w];
    W.slot[lane] = p ? 1 : 0;
    W.bar->arrive_and_wait();
    uint64_t m = 0;
    for (int i = 0; i < 64; ++i)
        m |= (W.slot[i] & 1ull) << i;
    W.bar->arrive_and_wait();
    return m;
}

void block_sync() { g_block->arrive_and_wait(); }

void launch_impl(dim3 g, dim3 b, size_t shmem, void (*body)(void *), void *arg)
{
    std::lock_guard<std::mutex> lk(g_launch);
    if (b.x % 64 != 0) {
        std::fprintf(stderr, "emu: block size %u is not a multiple of 64\n", b.x);
        std::abort();
    }
    const unsigned nw = b.x / 64;
    for (unsigned blk = 0; blk < g.x; ++blk) {
        g_dyn.assign(shmem + 16, 0xa5);
        g_waves.clear();
        g_waves.resize(nw);
        for (auto &w : g_waves)
            w.bar = std::make_unique<std::barrier<>>(64);
        g_block = std::make_unique<std::barrier<>>(b.x);
        std::vector<std::thread> th;
        th.reserve(b.x);
        for (unsigned t = 0; t < b.x; ++t)
            th.emplace_back([=]() {
                tIdx = dim3(t);
                bIdx = dim3(blk);
                bDim = b;
                gDim = g;
                body(arg);
            });
        for (auto &x : th)
            x.join();
    }
}
}  // namespace emu

hipError_t hipMalloc(void **p, size_t n)
{
    *p = std::malloc(n ? n : 1);
    if (!*p)
        return hipErrorOutOfMemory;
    std::memset(*p, 0xa5, n);
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned)
{
    return hipMalloc(p, n);
}
hipError_t hipHostFree(void *p)
{
    return hipFree(p);
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k,
                          hipStream_t)
{
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n)
{
    std::memset(d, v, n);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = (hipStream_t)1;
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = (hipEvent_t)1;
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned)
{
    return hipEventCreate(e);
}
// kernels run to completion at launch, so stream order is program order
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned)
{
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.f;
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int)
{
    std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
    p->multiProcessorCount = 2;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipMemGetInfo(size_t *f, size_t *t)
{
    *f = (size_t)16 << 30;
    *t = (size_t)32 << 30;
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
