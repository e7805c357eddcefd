// pa_tune.hip -- pa_autotune: picks the launch shapes of a context (pa_tuning)
// by timing a sample of the caller's own pairs, and caches the choice per
// device, target and alignment mode.
//
// Each knob is chosen by the device time of the stage it affects (event
// timers of the context: per-class DP kernel time, traceback time, X-drop
// time), coordinate by coordinate, starting from the installed tuning:
//   1. X-drop kernel (one lane per item / 16 lanes per item first), its LDS
//      ring size and its workgroups per CU, by X-drop time;
//   2. DP row-class edges, by DP + traceback time;
//   3. the kernel shape of each row class, by that class's DP time plus the
//      traceback time (a shape also fixes the trace layout);
//   4. persistent DP workgroups per CU, by DP time;
//   5. traceback staging, by traceback time;
//   6. tracebacks beside the next classes' DP, beside the last class's DP
//      or after all of it, by the wall time of the call;
//   7. X-drop items outgrowing the per-lane rings: the band kernels' window
//      chain or global rings, by X-drop time.
// A candidate replaces the incumbent only when that time is lower by more
// than kMinGain, so timing noise does not flip choices, and the wall time
// of the whole call is not higher by more than kMinGain.  A sample can mislead on
// heavy-tailed length mixes (a few long tasks set a kernel's time), so when
// the sample is smaller than the caller's set, the choice is checked against
// the starting tuning on up to kValidate evenly spaced pairs of the whole
// set and kept only if the whole call is faster there too.  Results never depend on the
// tuning; only the speed does.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "k_misc.h"
#include "pa_ctx.h"

using namespace pa;

namespace {

constexpr double kMinGain = 0.01;
constexpr uint64_t kDefaultSample = 262144;
constexpr uint64_t kValidate = 1u << 23;
constexpr int kReps = 2;

struct Times {
    double dp = 0, cls[kClasses] = {0, 0, 0, 0}, trace = 0, xdrop = 0;
    double wall = 0;   // host time of the whole call, synchronised
};

std::mutex g_mu;
std::map<std::tuple<int, int, int>, pa_tuning> g_cache;

__global__ void k_gather_pairs(const pa_pair *__restrict__ src, uint64_t n,
                               uint64_t m, pa_pair *__restrict__ dst)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m)
        dst[i] = src[i * n / m];
}

class Tuner {
public:
    Tuner(pa_ctx *c, const pa_pair *pairs, uint64_t n, int target, void *out)
        : c_(c), pairs_(pairs), n_(n), target_(target), out_(out)
    {
    }

    // Device times of one candidate: a warm-up call (allocations, first
    // launches), then the minimum of kReps timed calls per stage.
    int measure(const pa_tuning &t, Times *best)
    {
        int rc = pa_set_tuning(c_, &t);
        if (rc)
            return rc;
        if ((rc = run()))
            return rc;
        Times b;
        b.dp = b.trace = b.xdrop = b.wall = INFINITY;
        for (int k = 0; k < kClasses; ++k)
            b.cls[k] = INFINITY;
        for (int r = 0; r < kReps; ++r) {
            if ((rc = pa_reset_stats(c_)))
                return rc;
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = run()))
                return rc;
            b.wall = std::min(
                b.wall, std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0)
                            .count());
            pa_stats st;
            if ((rc = pa_get_stats(c_, &st)))
                return rc;
            b.dp = std::min(b.dp, st.ms_sw + st.ms_nw);
            b.trace = std::min(b.trace, st.ms_trace);
            b.xdrop = std::min(b.xdrop, st.ms_xdrop);
            for (int k = 0; k < kClasses; ++k) {
                double ms = 0;
                for (int m = 0; m < 3; ++m)
                    ms += c_->kstat[m][2 * k].ms + c_->kstat[m][2 * k + 1].ms +
                          c_->kstat[m][kBins + k].ms;
                b.cls[k] = std::min(b.cls[k], ms);
            }
        }
        *best = b;
        ++runs_;
        return PA_OK;
    }

    int runs() const { return runs_; }

private:
    int run()
    {
        int rc = target_ == PA_TUNE_SW_SCORES
                     ? pa_sw_scores_dev(c_, pairs_, n_, (pa_score *)out_,
                                        nullptr)
                     : pa_align_dev(c_, pairs_, n_, (pa_result *)out_,
                                    nullptr);
        if (rc)
            return rc;
        return pa_synchronize(c_);
    }

    pa_ctx *c_;
    const pa_pair *pairs_;
    uint64_t n_;
    int target_;
    void *out_;
    int runs_ = 0;
};

bool faster(double cand, double inc) { return cand < inc * (1.0 - kMinGain); }

// A candidate wins when the stage it affects is faster and the whole call is
// not slower (a shape can speed up its class's kernels and still lengthen
// the call, through the trace store it needs or the walks it overlaps).
bool better(double cand, double inc, const Times &c, const Times &i)
{
    return faster(cand, inc) && c.wall <= i.wall * (1.0 + kMinGain);
}

// The coordinate search; `t` starts as the installed tuning and ends as the
// choice.
int search(Tuner &tu, bool xdrop, pa_tuning &t)
{
    int rc;
    Times cur;
    if ((rc = tu.measure(t, &cur)))
        return rc;

    if (xdrop) {
        // 1. X-drop kernel, ring and grid
        for (int coop = 0; coop < 2; ++coop)
            for (int ring : {16, 32})
                for (int wg : {0, 8}) {
                    if (coop == t.xd_coop && ring == t.xd_ring &&
                        wg == t.xd_wg_per_cu)
                        continue;
                    pa_tuning cand = t;
                    cand.xd_coop = coop;
                    cand.xd_ring = ring;
                    cand.xd_wg_per_cu = wg;
                    Times tm;
                    if ((rc = tu.measure(cand, &tm)))
                        return rc;
                    if (better(tm.xdrop, cur.xdrop, tm, cur)) {
                        t = cand;
                        cur = tm;
                    }
                }
    }

    // 2. row-class edges
    static const int32_t edges[][3] = {
        {128, 256, 512}, {96, 192, 384}, {128, 256, 768}, {128, 384, 768},
        {64, 256, 512},  {128, 192, 512}};
    for (const auto &e : edges) {
        if (std::equal(e, e + 3, t.dp_edge))
            continue;
        pa_tuning cand = t;
        std::copy(e, e + 3, cand.dp_edge);
        Times tm;
        if ((rc = tu.measure(cand, &tm)))
            return rc;
        if (better(tm.dp + tm.trace, cur.dp + cur.trace, tm, cur)) {
            t = cand;
            cur = tm;
        }
    }

    // 3. the kernel shape of each class
    for (int k = 0; k < kClasses; ++k)
        for (int sh = 0; sh < kShapes; ++sh) {
            if (sh == t.dp_shape[k])
                continue;
            pa_tuning cand = t;
            cand.dp_shape[k] = sh;
            Times tm;
            if ((rc = tu.measure(cand, &tm)))
                return rc;
            if (better(tm.cls[k] + tm.trace, cur.cls[k] + cur.trace, tm,
                       cur)) {
                t = cand;
                cur = tm;
            }
        }

    // 4. persistent DP grid
    for (int wg : {0, 16, 8}) {
        if (wg == t.dp_wg_per_cu)
            continue;
        pa_tuning cand = t;
        cand.dp_wg_per_cu = wg;
        Times tm;
        if ((rc = tu.measure(cand, &tm)))
            return rc;
        if (better(tm.dp, cur.dp, tm, cur)) {
            t = cand;
            cur = tm;
        }
    }

    // 5. traceback staging
    if (cur.trace > 0) {
        pa_tuning cand = t;
        cand.tb_stage = 1 - t.tb_stage;
        Times tm;
        if ((rc = tu.measure(cand, &tm)))
            return rc;
        if (better(tm.trace, cur.trace, tm, cur)) {
            t = cand;
            cur = tm;
        }
    }

    // 6. tracebacks beside the DP or after it, by wall time
    if (cur.trace > 0) {
        const int start_ov = t.tb_overlap;
        for (int ov = 0; ov < 3; ++ov) {
            if (ov == start_ov)
                continue;
            pa_tuning cand = t;
            cand.tb_overlap = ov;
            Times tm;
            if ((rc = tu.measure(cand, &tm)))
                return rc;
            if (faster(tm.wall, cur.wall)) {
                t = cand;
                cur = tm;
            }
        }
    }

    // 7. wide X-drop bands
    if (xdrop) {
        for (int wv = 0; wv < 3; ++wv) {
            if (wv == t.xd_wave)
                continue;
            pa_tuning cand = t;
            cand.xd_wave = wv;
            Times tm;
            if ((rc = tu.measure(cand, &tm)))
                return rc;
            if (better(tm.xdrop, cur.xdrop, tm, cur)) {
                t = cand;
                cur = tm;
            }
        }
    }
    return PA_OK;
}

// Evenly spaced pairs [0, m) of src[0, n) into dst.
hipError_t gather(const pa_pair *src, uint64_t n, uint64_t m, pa_pair *dst,
                  hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_pairs, dim3((unsigned)((m + 255) / 256)),
                       dim3(256), 0, s, src, n, m, dst);
    return hipStreamSynchronize(s);
}

}  // namespace

extern "C" int pa_autotune(pa_ctx *c, const pa_pair *d_pairs, uint64_t n,
                           int target, uint64_t sample, uint32_t flags,
                           pa_tuning *out)
{
    if (!c)
        return PA_ERR_ARG;
    if (target != PA_TUNE_SW_SCORES && target != PA_TUNE_ALIGN)
        return fail(c, PA_ERR_ARG, "target must be a pa_tune_target");
    if (!c->have_rows || !c->have_cols)
        return fail(c, PA_ERR_NOSEQ, "load row and column sequences first");
    if (n == 0 || !d_pairs)
        return fail(c, PA_ERR_ARG, "pa_autotune needs pairs to time");
    if (hipSetDevice(c->device) != hipSuccess)
        return fail(c, PA_ERR_HIP, "hipSetDevice");
    const auto key = std::make_tuple(c->device, target, (int)c->cfg.mode);
    if (!(flags & PA_TUNE_FORCE)) {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_cache.find(key);
        if (it != g_cache.end()) {
            int rc = pa_set_tuning(c, &it->second);
            if (!rc && out)
                *out = it->second;
            return rc;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->st));
    const uint64_t m = std::min(n, sample ? sample : kDefaultSample);
    const uint64_t mv = m < n ? std::min(n, std::max(m, kValidate)) : 0;
    const uint64_t cap = std::max(m, mv);
    const size_t res_size = target == PA_TUNE_SW_SCORES ? sizeof(pa_score)
                                                        : sizeof(pa_result);
    pa_pair *dp = nullptr;
    void *dout = nullptr;
    if (hipMalloc(&dp, cap * sizeof(pa_pair)) != hipSuccess ||
        hipMalloc(&dout, cap * res_size) != hipSuccess) {
        hipFree(dp);
        return fail(c, PA_ERR_NOMEM, "pa_autotune sample buffers");
    }
    const pa_tuning start = c->tune;
    pa_tuning t = start;
    int rc = PA_OK;
    hipError_t e = gather(d_pairs, n, m, dp, c->st);
    const bool xdrop = target == PA_TUNE_ALIGN && c->cfg.mode == PA_MODE_XDROP;
    if (e != hipSuccess) {
        rc = fail(c, PA_ERR_HIP,
                  std::string("pa_autotune gather: ") + hipGetErrorString(e));
    } else {
        Tuner tu(c, dp, m, target, dout);
        rc = search(tu, xdrop, t);
    }
    if (!rc && mv && std::memcmp(&t, &start, sizeof(t)) != 0) {
        // the choice against the starting tuning on a larger share of the set
        if ((e = gather(d_pairs, n, mv, dp, c->st)) != hipSuccess) {
            rc = fail(c, PA_ERR_HIP, std::string("pa_autotune gather: ") +
                                         hipGetErrorString(e));
        } else {
            Tuner tv(c, dp, mv, target, dout);
            Times a, b;
            if (!(rc = tv.measure(start, &a)) && !(rc = tv.measure(t, &b)) &&
                !faster(b.wall, a.wall))
                t = start;
        }
    }
    hipFree(dp);
    hipFree(dout);
    if (rc) {
        pa_set_tuning(c, &start);
        return rc;
    }
    if ((rc = pa_set_tuning(c, &t)))
        return rc;
    pa_reset_stats(c);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = t;
    }
    if (out)
        *out = t;
    return PA_OK;
}
