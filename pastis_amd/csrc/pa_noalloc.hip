// pa_noalloc.hip -- pa_sw_scores_noalloc: the score-only local alignment of
// pa_sw_scores_dev with its workspace in a fixed device array, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid allocation).
//
// The workspace of one call -- DP tasks and outputs, task flags, bin and
// saturation lists, queue counters and the boundary rows of multi-pass
// tasks -- lives in `scratch`, a file-scope device array of kScratchBytes
// (64 MiB) that every GPU the library runs on holds until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it is a
// stand-in, not a device address): its device address is looked up once per
// device with hipGetSymbolAddress, when the first context on that device is
// created, and only that address is used.  Calls on one GPU share the array:
// the counters and flags that must start at zero are cleared by fill kernels
// at the start of every call, and calls that overlap on one GPU race.
#include <mutex>

#include "k_misc.h"
#include "pa_ctx.h"

using namespace pa;

namespace {

// Limits of one call.
constexpr uint64_t kMaxPairs = 262144;     // pairs per call
constexpr uint32_t kMaxLen = 4096;         // longest loaded sequence
constexpr int kMaxSlots = 1024;            // boundary rows of kMaxLen

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Regions of the scratch array (byte offsets), sized for kMaxPairs pairs.
constexpr size_t kCtrWords = 2048;         // >= the counter block's words
constexpr size_t kOffTasks = 0;
constexpr size_t kOffOut = align256(kOffTasks + kMaxPairs * sizeof(DpTask));
constexpr size_t kOffFlags = align256(kOffOut + kMaxPairs * sizeof(DpOut));
constexpr size_t kOffLists = align256(kOffFlags + kMaxPairs);
constexpr size_t kOffSat =
    align256(kOffLists + (size_t)kBins * kMaxPairs * sizeof(uint32_t));
constexpr size_t kOffCtr =
    align256(kOffSat + (size_t)kClasses * kMaxPairs * sizeof(uint32_t));
constexpr size_t kOffBnd = align256(kOffCtr + kCtrWords * sizeof(uint32_t));
constexpr size_t kBndBytes =
    (size_t)kMaxSlots * 2u * (kMaxLen + 2u) * sizeof(int32_t);
constexpr size_t kScratchBytes = align256(kOffBnd + kBndBytes);
static_assert(kScratchBytes <= (size_t)64 << 20, "scratch stays modest");

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};   // device address of scratch per device

// The context's workspace buffers viewed onto the scratch regions for the
// duration of one call, with event timers off and the DP grid capped so that
// the boundary rows fit; everything is restored on return.
class ScratchView {
public:
    ScratchView(pa_ctx *c, uint8_t *base) : c_(c)
    {
        Buf *bufs[] = {&c->tasks, &c->dpout, &c->tflags, &c->lists, &c->sat,
                       &c->ctr, &c->bnd};
        const size_t offs[] = {kOffTasks, kOffOut, kOffFlags, kOffLists,
                               kOffSat, kOffCtr, kOffBnd};
        const size_t ends[] = {kOffOut, kOffFlags, kOffLists, kOffSat,
                               kOffCtr, kOffBnd, kScratchBytes};
        for (int i = 0; i < 7; ++i) {
            saved_[i] = *bufs[i];
            bufs[i]->p = base + offs[i];
            bufs[i]->cap = ends[i] - offs[i];
        }
        quiet_ = c->quiet;
        slot_cap_ = c->slot_cap;
        c->quiet = true;
        // the DP grids hold no more task slots than the boundary-row region
        // has rows of the loaded sets' longest column sequence
        const size_t row = 2u * ((size_t)c->max_clen + 2u) * sizeof(int32_t);
        c->slot_cap = kBndBytes / row;
    }
    ~ScratchView()
    {
        Buf *bufs[] = {&c_->tasks, &c_->dpout, &c_->tflags, &c_->lists,
                       &c_->sat, &c_->ctr, &c_->bnd};
        for (int i = 0; i < 7; ++i)
            *bufs[i] = saved_[i];
        c_->quiet = quiet_;
        c_->slot_cap = slot_cap_;
    }

private:
    pa_ctx *c_;
    Buf saved_[7];
    bool quiet_;
    uint64_t slot_cap_;
};

}  // namespace

void pa::noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device])
        return;
    void *p = nullptr;
    // the device is current (pa_create set it)
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess)
        g_scratch[device] = p;
}

extern "C" int pa_sw_scores_noalloc(pa_ctx *c, const pa_pair *d_pairs,
                                    uint64_t n, pa_score *d_out, void *stream)
{
    if (!c)
        return PA_ERR_ARG;
    if (!c->have_rows || !c->have_cols)
        return fail(c, PA_ERR_NOSEQ, "load row and column sequences first");
    if (n == 0)
        return PA_OK;
    if (!d_pairs || !d_out)
        return fail(c, PA_ERR_ARG, "null pointer");
    if (n > kMaxPairs)
        return fail(c, PA_ERR_LIMIT,
                    "pa_sw_scores_noalloc takes at most 262144 pairs per call");
    if (c->max_rlen > kMaxLen || c->max_clen > kMaxLen)
        return fail(c, PA_ERR_LIMIT,
                    "pa_sw_scores_noalloc needs loaded sequences of at most "
                    "4096 residues");
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (c->device >= 0 && c->device < kMaxDevices)
            base = g_scratch[c->device];
    }
    if (!base)
        return fail(c, PA_ERR_HIP, "scratch array of this device not resolved");
    if (hipSetDevice(c->device) != hipSuccess)
        return fail(c, PA_ERR_HIP, "hipSetDevice");
    ScratchView view(c, (uint8_t *)base);
    return sw_scores_impl(c, d_pairs, n, d_out,
                          stream ? (hipStream_t)stream : c->st);
}
