// C ABI of libdespeckcl_amd.so: argument checks, device workspace, host<->device copies.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_goldstein.h"
#include "dkx_nlinsar.h"
#include "dkx_nlsar.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

namespace dkx {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

// Grow-only device workspace per device, reused across calls so that repeated calls (and
// the benchmark's timed loop) do not allocate.  Calls on one device are serialised by the
// mutex of that device's entry.
// Device copies of NL-SAR constant blobs (weight LUTs, Gaussian taps), keyed by content, so
// that calls alternating between parameter sets do not re-upload them.
struct ConstSlot {
    float* ptr = nullptr;
    size_t cap = 0;
    std::vector<float> host;        // content, and the source of the last upload
    hipEvent_t uploaded = nullptr;  // completion of the last upload from `host`
    unsigned long long used = 0;    // call counter at the last use (LRU)
};
constexpr int CONST_SLOTS = 4;

struct DeviceWorkspace {
    std::mutex mu;
    void* ptr = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;       // the library's own stream (host-array entry points)
    ConstSlot consts[CONST_SLOTS];
    unsigned long long calls = 0;
    // completion of the last call that used this workspace, and its stream: a call on
    // another stream waits for it, so calls sharing the workspace never overlap on the GPU
    hipEvent_t done = nullptr;
    hipStream_t done_stream = nullptr;
    bool have_done = false;
};

// make `stream` wait for the previous user of the workspace
static int workspace_acquire(DeviceWorkspace* w, hipStream_t stream)
{
    if (w->have_done && w->done_stream != stream) DKX_HIP_CHECK(hipStreamWaitEvent(stream, w->done, 0));
    return 0;
}

static int workspace_release(DeviceWorkspace* w, hipStream_t stream)
{
    if (!w->done) DKX_HIP_CHECK(hipEventCreateWithFlags(&w->done, hipEventDisableTiming));
    DKX_HIP_CHECK(hipEventRecord(w->done, stream));
    w->done_stream = stream;
    w->have_done = true;
    return 0;
}

// Device copy of a constants blob for a call on `stream`, which must already wait for the
// workspace's previous user (workspace_acquire).  Every earlier call on this device is
// ordered before that point, so a slot holding other content may be overwritten by an upload
// enqueued on `stream`: no kernel still reads it, and nothing waits on the host for the device.
static int set_constants(DeviceWorkspace* w, const std::vector<float>& blob, hipStream_t stream,
                         const float** dev)
{
    w->calls++;
    ConstSlot* pick = nullptr;
    for (ConstSlot& c : w->consts)
        if (c.ptr && c.host == blob) pick = &c;
    if (pick) {
        pick->used = w->calls;
        *dev = pick->ptr;
        return 0;
    }
    pick = &w->consts[0];
    for (ConstSlot& c : w->consts)
        if (c.used < pick->used) pick = &c;
    // the slot's host copy may still be the source of an upload in flight
    if (pick->uploaded) DKX_HIP_CHECK(hipEventSynchronize(pick->uploaded));
    else DKX_HIP_CHECK(hipEventCreateWithFlags(&pick->uploaded, hipEventDisableTiming));
    const size_t need = std::max<size_t>(blob.size(), 1);
    if (pick->cap < need) {
        // an allocation change: the previous buffer may still be read by the last call
        if (pick->ptr) {
            if (w->have_done) DKX_HIP_CHECK(hipEventSynchronize(w->done));
            DKX_HIP_CHECK(hipFree(pick->ptr));
        }
        pick->ptr = nullptr;
        pick->cap = 0;
        pick->host.clear();
        DKX_HIP_CHECK(hipMalloc(&pick->ptr, need * sizeof(float)));
        pick->cap = need;
    }
    pick->host = blob;
    pick->used = w->calls;
    DKX_HIP_CHECK(hipMemcpyAsync(pick->ptr, pick->host.data(), blob.size() * sizeof(float),
                                 hipMemcpyHostToDevice, stream));
    DKX_HIP_CHECK(hipEventRecord(pick->uploaded, stream));
    *dev = pick->ptr;
    return 0;
}

static std::mutex g_ws_mu;
static std::map<int, DeviceWorkspace*> g_ws;

static DeviceWorkspace* workspace_for_current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto it = g_ws.find(dev);
    if (it != g_ws.end()) return it->second;
    DeviceWorkspace* w = new DeviceWorkspace();
    g_ws[dev] = w;
    return w;
}

static int ensure(DeviceWorkspace* w, size_t bytes)
{
    if (w->bytes >= bytes) return 0;
    if (w->ptr) {
        // every call that used the old buffer is ordered before the last one's event
        if (w->have_done) DKX_HIP_CHECK(hipEventSynchronize(w->done));
        DKX_HIP_CHECK(hipFree(w->ptr));
        w->ptr = nullptr;
        w->bytes = 0;
    }
    hipError_t e = hipMalloc(&w->ptr, bytes);
    if (e != hipSuccess) {
        w->ptr = nullptr;
        set_error("device workspace allocation of " + std::to_string(bytes) + " bytes failed: " +
                  hipGetErrorString(e));
        return DKX_ERESOURCE;
    }
    w->bytes = bytes;
    return 0;
}

static int ensure_stream(DeviceWorkspace* w)
{
    if (!w->stream) DKX_HIP_CHECK(hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking));
    return 0;
}

// NL-SAR: chosen per pass (NLSAR_TILE_AUTO); NL-InSAR: chosen per image (NLINSAR_TILE_AUTO)
static std::atomic<int> g_tile[2] = {{NLSAR_TILE_AUTO}, {NLINSAR_TILE_AUTO}};
static std::atomic<int> g_nli_store{NLINSAR_STORE_AUTO};   // NLINSAR_STORE_* of single calls

static std::atomic<bool> g_timing{false};
static std::mutex g_timing_mu;
struct Mark { int kind; hipEvent_t a, b; };
static std::vector<Mark> g_marks;
static std::vector<std::pair<int, float>> g_times;

bool timing_on() { return g_timing.load(); }

void timing_push(int kind, hipEvent_t start, hipEvent_t stop)
{
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_marks.push_back(Mark{kind, start, stop});
}

// resolve the events recorded by the last call (after it synchronised)
static void timing_collect()
{
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_times.clear();
    for (const Mark& m : g_marks) {
        float ms = -1.f;
        if (hipEventSynchronize(m.b) == hipSuccess && hipEventElapsedTime(&ms, m.a, m.b) != hipSuccess) ms = -1.f;
        g_times.emplace_back(m.kind, ms);
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    g_marks.clear();
}

int check_stats(const dkx_nlsar_stats* stats, int n_stats, const int* P, int nP, const int* S,
                int nS, std::vector<NlsarStatsView>& views, int* last_unused)
{
    // The reference looks statistics up with operator[] (nlsar_filter_sub_image.cpp:278):
    // a missing (patch, scale) entry would silently filter with empty statistics, so it is
    // rejected here, as is a key given twice (a std::map holds each key once).  Entries for
    // other keys are accepted: the reference builds ENL maps for them too (:224-228), which
    // matters for the selection only when the collection's last key is one of them.
    if (!stats || n_stats < nP * nS || n_stats > 1 << 20) {
        set_error("nlsar: the statistics must hold an entry for every (patch size, scale size) pair");
        return DKX_ESTATS;
    }
    views.clear();
    int lastp = 0, lasts = 0;
    for (int i = 0; i < n_stats; i++) {
        const dkx_nlsar_stats& s = stats[i];
        for (int j = 0; j < i; j++)
            if (stats[j].patch_size == s.patch_size && stats[j].scale_size == s.scale_size) {
                set_error("nlsar: statistics key (" + std::to_string(s.patch_size) + ", " +
                          std::to_string(s.scale_size) + ") appears twice");
                return DKX_ESTATS;
            }
        // nlsar::params::operator<, parameters.h:37-40
        if (i == 0 || s.patch_size > lastp || (s.patch_size == lastp && s.scale_size > lasts)) {
            lastp = s.patch_size;
            lasts = s.scale_size;
        }
    }
    bool last_in = false;
    for (int a = 0; a < nP; a++)
        for (int b = 0; b < nS; b++) {
            if (P[a] == lastp && S[b] == lasts) last_in = true;
            const dkx_nlsar_stats* e = nullptr;
            for (int i = 0; i < n_stats; i++)
                if (stats[i].patch_size == P[a] && stats[i].scale_size == S[b]) e = &stats[i];
            if (!e) {
                set_error("nlsar: no statistics for (patch " + std::to_string(P[a]) + ", scale " +
                          std::to_string(S[b]) + ")");
                return DKX_ESTATS;
            }
            if (e->lut_size < 2 || e->lut_size > 65536 || !e->quantilles || !e->chi2cdf_inv ||
                !(e->dissims_max >= e->dissims_min)) {
                set_error("nlsar: invalid statistics entry");
                return DKX_ESTATS;
            }
            for (int k = 0; k < e->lut_size; k++)
                if (!(e->quantilles[k] >= 0.0f && e->quantilles[k] < 1.0f)) {
                    set_error("nlsar: quantile LUT values must lie in [0, 1)");
                    return DKX_ESTATS;
                }
            views.push_back(NlsarStatsView{e->patch_size, e->scale_size, e->lut_size, e->dissims_min,
                                           e->dissims_max, e->quantilles, e->chi2cdf_inv});
        }
    if (last_unused) *last_unused = last_in ? 0 : 1;
    return 0;
}

// Everything that can be checked without a device: argument checks, statistics, plan.
// all_keys: compute every key's statistics (the ENL / alpha maps and the undefined-selection
// count need them); the plan depends on it (the wide pass variant serves the selected key only)
int nlsar_prepare(NlsarCall& call, int dim, int all_keys, int height, int width, int sw,
                         const int* P, int nP, const int* S, int nS, const dkx_nlsar_stats* stats,
                         int n_stats, float h, float c)
{
    if (!P || !S) { set_error("nlsar: null size list"); return DKX_EINVAL; }
    if (nP < 1 || nS < 1) { set_error("nlsar: empty patch or scale size list"); return DKX_EINVAL; }
    int last_unused = 0;
    int rc = check_stats(stats, n_stats, P, nP, S, nS, call.views, &last_unused);
    if (rc) return rc;
    // the pass variant: the caller's setting, else the autotuner's record for this parameter
    // set and image size, else chosen per pass
    int tile = g_tile[0].load();
    if (tile == NLSAR_TILE_AUTO) tile = tuned_tile(0, all_keys ? 1 : 0, sw, P, nP, S, nS, (double)height * width);
    if (tile < 0) tile = NLSAR_TILE_AUTO;
    call.pp = NlsarParams{dim, sw, P, nP, S, nS, call.views.data(), (int)call.views.size(), h, c,
                          tile, all_keys, last_unused};
    rc = nlsar_common(call.pp, &call.cm);
    if (rc) return rc;
    return nlsar_image_plan(call.cm, height, width, -1, -1, &call.ip);
}

// a1 / a2 / phi -> ref / pha / coh for InSAR input, cov -> cov_out for covariance input
static int nlsar_device_impl(NlsarCall& call, const float* a1, const float* a2, const float* phi,
                             const float* cov, float* ref, float* pha, float* coh, float* cov_out,
                             float* enl_maps, float* alpha_maps, unsigned long long* ub,
                             hipStream_t stream, DeviceWorkspace* w)
{
    if ((enl_maps || alpha_maps || ub) && !call.pp.all_keys) {
        set_error("nlsar: internal error, diagnostics without every key's statistics");
        return DKX_EINVAL;
    }
    std::vector<float> blob;
    int rc = nlsar_build_constants(call.pp, call.cm, blob);
    if (rc) return rc;
    rc = ensure(w, call.ip.bytes);
    if (rc) return rc;
    rc = workspace_acquire(w, stream);
    if (rc) return rc;
    const float* d_const = nullptr;
    rc = set_constants(w, blob, stream, &d_const);
    if (rc) return rc;
    NlsarImageDesc d;
    nlsar_image_desc(call.ip, (char*)w->ptr, a1, a2, phi, cov, ref, pha, coh, cov_out, enl_maps, alpha_maps, &d);
    rc = nlsar_enqueue(call.pp, call.cm, &d, 1, nullptr, d_const, stream);
    if (rc) return rc;
    if (ub) DKX_HIP_CHECK(hipMemcpyAsync(ub, d.ub_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    rc = workspace_release(w, stream);
    if (rc) return rc;
    if (ub) DKX_HIP_CHECK(hipStreamSynchronize(stream));
    return 0;
}

struct NlinsarCall {
    NlinsarProblem pb;
    NlinsarPlan plan;
};

int nlinsar_problem(int height, int width, int sw, int p, int niter, int lmin, int tile,
                    NlinsarProblem* pb)
{
    // nlinsar.cpp:65-78: h from the patch size (tabulated for 3 and 7, else the seeded
    // restatement of the reference's Monte-Carlo quantile, nlinsar_simu.cpp)
    if (p < 1 || p % 2 == 0 || p > 2 * NLI_MAX_PSH + 1) {
        set_error("nlinsar: patch sizes must be odd, up to 21 (the weights kernels are compiled for those)");
        return DKX_EINVAL;
    }
    const float h_theo = nlinsar_h_theo(p);
    const int patch_area = p * p;
    const float h_para = h_theo * patch_area;
    const float T_para = (float)(2.0 / h_para * M_PI / 4 * patch_area);
    *pb = NlinsarProblem{height, width, sw, p, niter, lmin, h_para, T_para, 80, tile};
    return 0;
}

static int nlinsar_prepare(NlinsarCall& call, int height, int width, int sw, int p, int niter,
                           int lmin)
{
    // the tile shape: the caller's setting, else the autotuner's record for these parameters
    // and image size, else the built-in rule
    int tile = g_tile[1].load();
    if (tile == NLINSAR_TILE_AUTO) tile = tuned_tile(1, 0, sw, &p, 1, nullptr, 0, (double)height * width);
    if (tile < 0) tile = NLINSAR_TILE_AUTO;
    int rc = nlinsar_problem(height, width, sw, p, niter, lmin, tile, &call.pb);
    if (rc) return rc;
    call.pb.store = g_nli_store.load();
    return nlinsar_plan(call.pb, &call.plan);
}

static int nlinsar_device_impl(NlinsarCall& call, const float* a1, const float* a2,
                               const float* phi, float* ref, float* pha, float* coh,
                               unsigned long long* nsm, hipStream_t stream, DeviceWorkspace* w)
{
    int rc = ensure(w, call.plan.ws_bytes);
    if (rc) return rc;
    rc = workspace_acquire(w, stream);
    if (rc) return rc;
    rc = nlinsar_run(call.pb, call.plan, a1, a2, phi, ref, pha, coh, w->ptr, nsm, stream);
    if (rc) return rc;
    return workspace_release(w, stream);
}

static int goldstein_prepare(GoldsteinParams& gp, GoldsteinPlan& pl, int height, int width,
                             int patch_size, int overlap, float alpha)
{
    if (!std::isfinite(alpha)) {
        set_error("goldstein: alpha must be finite");
        return DKX_EINVAL;
    }
    gp = GoldsteinParams{height, width, patch_size, overlap, alpha};
    return goldstein_plan(gp, &pl);
}

static int goldstein_device_impl(const GoldsteinParams& gp, const GoldsteinPlan& pl,
                                 const float* a1, const float* a2, const float* phi, float* ref,
                                 float* pha, hipStream_t stream, DeviceWorkspace* w)
{
    int rc = ensure(w, std::max<size_t>(pl.ws_bytes, 256));
    if (rc) return rc;
    rc = workspace_acquire(w, stream);
    if (rc) return rc;
    rc = goldstein_run(gp, pl, a1, a2, phi, ref, pha, w->ptr, stream);
    if (rc) return rc;
    return workspace_release(w, stream);
}

// host arrays: copy in, run on the library stream, copy out
struct HostStage {
    void* ptr = nullptr;
    ~HostStage() { if (ptr) (void)hipFree(ptr); }
};

// Times run(): prep() then run() once as a warm-up, then reps times prep() and run(), the
// runs alone between two events on `stream`; ms: their mean.
template <class Prep, class Run>
static int time_runs(hipStream_t stream, int reps, Prep prep, Run run, float* ms)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DKX_HIP_CHECK(hipEventCreate(&e0));
    DKX_HIP_CHECK(hipEventCreate(&e1));
    int rc = prep();
    if (!rc) rc = run();
    float total = 0.f;
    for (int r = 0; r < reps && !rc; r++) {
        rc = prep();
        if (rc) break;
        if (hipEventRecord(e0, stream) != hipSuccess) { rc = DKX_EHIP; break; }
        rc = run();
        if (rc) break;
        if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = DKX_EHIP; break; }
        float t = 0.f;
        if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) { rc = DKX_EHIP; break; }
        total += t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == DKX_EHIP) set_error("autotune: event timing failed");
    *ms = total / (float)std::max(reps, 1);
    return rc;
}

}  // namespace dkx

using namespace dkx;

extern "C" {

const char* dkx_last_error(void) { return dkx::last_error(); }
int dkx_version(void) { return 1; }

int dkx_set_tile(int filter, int tile_index)
{
    return dkx::guarded<int>([&]() -> int {
        if (filter < 0 || filter > 1 || tile_index < -1 || tile_index > 2) {
            set_error("dkx_set_tile: invalid filter or tile index");
            return DKX_EINVAL;
        }
        g_tile[filter].store(tile_index);
        return 0;
    });
}

int dkx_get_tile(int filter) { return (filter >= 0 && filter <= 1) ? g_tile[filter].load() : -1; }

int dkx_set_nlinsar_weight_store(int mode)
{
    return dkx::guarded<int>([&]() -> int {
        if (mode < NLINSAR_STORE_AUTO || mode > NLINSAR_STORE_NONE) {
            set_error("dkx_set_nlinsar_weight_store: mode must be -1, 0, 1 or 2");
            return DKX_EINVAL;
        }
        g_nli_store.store(mode);
        return 0;
    });
}

int dkx_get_nlinsar_weight_store(void) { return g_nli_store.load(); }

int dkx_set_timing(int enabled)
{
    g_timing.store(enabled != 0);
    return 0;
}

int dkx_stage_times(int* kinds, float* ms, int max_entries)
{
    return dkx::guarded<int>([&]() -> int {
        timing_collect();
        std::lock_guard<std::mutex> lk(g_timing_mu);
        const int n = (int)g_times.size();
        for (int i = 0; i < n && i < max_entries; i++) {
            if (kinds) kinds[i] = g_times[i].first;
            if (ms) ms[i] = g_times[i].second;
        }
        return n;
    });
}

int dkx_nlsar_insar_device(const float* a1, const float* a2, const float* phi, float* ref,
                           float* pha, float* coh, int height, int width, int sw, const int* P,
                           int nP, const int* S, int nS, const dkx_nlsar_stats* stats, int n_stats,
                           float h, float c, float* enl_maps, float* alpha_maps,
                           unsigned long long* ub, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlsar: null array argument"); return DKX_EINVAL; }
        NlsarCall call;
        const int diag = (enl_maps || alpha_maps || ub) ? 1 : 0;
        int rc = nlsar_prepare(call, 0, diag, height, width, sw, P, nP, S, nS, stats, n_stats, h, c);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        return nlsar_device_impl(call, a1, a2, phi, nullptr, ref, pha, coh, nullptr, enl_maps,
                                 alpha_maps, ub, (hipStream_t)stream, w);
    });
}

int dkx_nlsar_insar(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                    float* coh, int height, int width, int sw, const int* P, int nP, const int* S,
                    int nS, const dkx_nlsar_stats* stats, int n_stats, float h, float c)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlsar: null array argument"); return DKX_EINVAL; }
        NlsarCall call;
        int rc = nlsar_prepare(call, 0, 0, height, width, sw, P, nP, S, nS, stats, n_stats, h, c);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        rc = ensure_stream(w);
        if (rc) return rc;
        const size_t n = (size_t)height * width;
        HostStage st;
        DKX_HIP_CHECK(hipMalloc(&st.ptr, 6 * n * sizeof(float)));
        float* d = (float*)st.ptr;
        DKX_HIP_CHECK(hipMemcpyAsync(d, a1, n * sizeof(float), hipMemcpyHostToDevice, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(d + n, a2, n * sizeof(float), hipMemcpyHostToDevice, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(d + 2 * n, phi, n * sizeof(float), hipMemcpyHostToDevice, w->stream));
        rc = nlsar_device_impl(call, d, d + n, d + 2 * n, nullptr, d + 3 * n, d + 4 * n, d + 5 * n,
                               nullptr, nullptr, nullptr, nullptr, w->stream, w);
        if (rc) { (void)hipStreamSynchronize(w->stream); return rc; }
        DKX_HIP_CHECK(hipMemcpyAsync(ref, d + 3 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(pha, d + 4 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(coh, d + 5 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipStreamSynchronize(w->stream));
        return 0;
    });
}

int dkx_nlsar_covmat_device(const float* cov, float* cov_out, int height, int width, int dim,
                            int sw, const int* P, int nP, const int* S, int nS,
                            const dkx_nlsar_stats* stats, int n_stats, float h, float c,
                            float* enl_maps, float* alpha_maps, unsigned long long* ub,
                            void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        if (!cov || !cov_out) { set_error("nlsar: null array argument"); return DKX_EINVAL; }
        if (dim != 2 && dim != 3) {
            set_error("nlsar: covariance matrices of dimension 2 or 3 are supported");
            return dim == 1 ? DKX_EUNSUPPORTED : DKX_EINVAL;
        }
        NlsarCall call;
        const int diag = (enl_maps || alpha_maps || ub) ? 1 : 0;
        int rc = nlsar_prepare(call, dim, diag, height, width, sw, P, nP, S, nS, stats, n_stats, h, c);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        return nlsar_device_impl(call, nullptr, nullptr, nullptr, cov, nullptr, nullptr, nullptr,
                                 cov_out, enl_maps, alpha_maps, ub, (hipStream_t)stream, w);
    });
}

int dkx_nlsar_covmat(const float* cov, float* cov_out, int height, int width, int dim, int sw,
                     const int* P, int nP, const int* S, int nS, const dkx_nlsar_stats* stats,
                     int n_stats, float h, float c)
{
    return dkx::guarded<int>([&]() -> int {
        if (!cov || !cov_out) { set_error("nlsar: null array argument"); return DKX_EINVAL; }
        if (dim != 2 && dim != 3) {
            set_error("nlsar: covariance matrices of dimension 2 or 3 are supported");
            return dim == 1 ? DKX_EUNSUPPORTED : DKX_EINVAL;
        }
        NlsarCall call;
        int rc = nlsar_prepare(call, dim, 0, height, width, sw, P, nP, S, nS, stats, n_stats, h, c);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        rc = ensure_stream(w);
        if (rc) return rc;
        const size_t n = (size_t)2 * dim * dim * height * width;
        HostStage st;
        DKX_HIP_CHECK(hipMalloc(&st.ptr, 2 * n * sizeof(float)));
        float* d = (float*)st.ptr;
        DKX_HIP_CHECK(hipMemcpyAsync(d, cov, n * sizeof(float), hipMemcpyHostToDevice, w->stream));
        rc = nlsar_device_impl(call, nullptr, nullptr, nullptr, d, nullptr, nullptr, nullptr, d + n,
                               nullptr, nullptr, nullptr, w->stream, w);
        if (rc) { (void)hipStreamSynchronize(w->stream); return rc; }
        DKX_HIP_CHECK(hipMemcpyAsync(cov_out, d + n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipStreamSynchronize(w->stream));
        return 0;
    });
}

int dkx_nlinsar_device(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                       float* coh, int height, int width, int sw, int p, int niter, int lmin,
                       unsigned long long* nsm, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlinsar: null array argument"); return DKX_EINVAL; }
        NlinsarCall call;
        int rc = nlinsar_prepare(call, height, width, sw, p, niter, lmin);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        return nlinsar_device_impl(call, a1, a2, phi, ref, pha, coh, nsm, (hipStream_t)stream, w);
    });
}

int dkx_nlsar_insar_device_autotune(const float* a1, const float* a2, const float* phi, float* ref,
                                    float* pha, float* coh, int height, int width, int sw, const int* P,
                                    int nP, const int* S, int nS, const dkx_nlsar_stats* stats,
                                    int n_stats, float h, float c, int every_key, int reps, void* stream_,
                                    int* tile_out, float* ms_out)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlsar: null array argument"); return DKX_EINVAL; }
        NlsarCall call;
        int rc = nlsar_prepare(call, 0, every_key ? 1 : 0, height, width, sw, P, nP, S, nS, stats, n_stats, h, c);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        const hipStream_t stream = (hipStream_t)stream_;
        reps = std::max(reps, 1);
        float ms_t[NLSAR_NTILES];
        for (int t = 0; t < NLSAR_NTILES; t++) {
            call.pp.tile = t;
            call.cm.tile = t;
            float ms = 0.f;
            {
                std::lock_guard<std::mutex> lk(w->mu);
                rc = time_runs(stream, reps, [] { return 0; },
                               [&] {
                                   return nlsar_device_impl(call, a1, a2, phi, nullptr, ref, pha, coh, nullptr,
                                                            nullptr, nullptr, nullptr, stream, w);
                               },
                               &ms);
            }
            if (rc) return rc;
            if (ms_out) ms_out[t] = ms;
            ms_t[t] = ms;
        }
        const int best = nlsar_pick_variant(ms_t);
        record_tuned_tile(0, every_key ? 1 : 0, sw, P, nP, S, nS, (double)height * width, best);
        if (tile_out) *tile_out = best;
        return 0;
    });
}

int dkx_nlinsar_device_autotune(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                                float* coh, int height, int width, int sw, int p, int niter, int lmin,
                                int reps, void* stream_, int* tile_out, float* ms_out)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlinsar: null array argument"); return DKX_EINVAL; }
        NlinsarCall call;
        int rc = nlinsar_prepare(call, height, width, sw, p, niter, lmin);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        const hipStream_t stream = (hipStream_t)stream_;
        reps = std::max(reps, 1);
        // the first-iteration estimates, restored before every run (the runs overwrite them)
        const size_t n = (size_t)height * width;
        float* init = nullptr;
        if (hipMalloc(&init, 3 * n * sizeof(float)) != hipSuccess) {
            (void)hipGetLastError();
            set_error("nlinsar autotune: allocation of the estimate copy failed");
            return DKX_ERESOURCE;
        }
        float* io[3] = {ref, pha, coh};
        rc = 0;
        for (int k = 0; k < 3 && !rc; k++)
            if (hipMemcpyAsync(init + k * n, io[k], n * sizeof(float), hipMemcpyDeviceToDevice, stream) != hipSuccess)
                rc = DKX_EHIP;
        auto restore = [&]() -> int {
            for (int k = 0; k < 3; k++)
                DKX_HIP_CHECK(hipMemcpyAsync(io[k], init + k * n, n * sizeof(float), hipMemcpyDeviceToDevice, stream));
            return 0;
        };
        int best = -1;
        float best_ms = 0.f;
        for (int t = 0; t < 3 && !rc; t++) {
            NlinsarCall ct = call;
            ct.pb.tile = t;
            if (nlinsar_plan(ct.pb, &ct.plan)) {
                if (ms_out) ms_out[t] = -1.f;   // a shape these parameters do not allow
                continue;
            }
            float ms = -1.f;
            {
                std::lock_guard<std::mutex> lk(w->mu);
                rc = time_runs(stream, reps, restore,
                               [&] { return nlinsar_device_impl(ct, a1, a2, phi, ref, pha, coh, nullptr, stream, w); },
                               &ms);
            }
            if (rc) break;
            if (ms_out) ms_out[t] = ms;
            if (ms >= 0.f && (best < 0 || ms < best_ms)) {
                best = t;
                best_ms = ms;
            }
        }
        if (!rc && best < 0) {
            set_error("nlinsar autotune: no tile shape runs with these parameters");
            rc = DKX_EUNSUPPORTED;
        }
        (void)hipStreamSynchronize(stream);
        (void)hipFree(init);
        if (rc) return rc;
        record_tuned_tile(1, 0, sw, &p, 1, nullptr, 0, (double)height * width, best);
        if (tile_out) *tile_out = best;
        return 0;
    });
}

int dkx_nlinsar(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                float* coh, int height, int width, int sw, int p, int niter, int lmin)
{
    return dkx::guarded<int>([&]() -> int {
        if (!a1 || !a2 || !phi || !ref || !pha || !coh) { set_error("nlinsar: null array argument"); return DKX_EINVAL; }
        NlinsarCall call;
        int rc = nlinsar_prepare(call, height, width, sw, p, niter, lmin);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        rc = ensure_stream(w);
        if (rc) return rc;
        const size_t n = (size_t)height * width;
        HostStage st;
        DKX_HIP_CHECK(hipMalloc(&st.ptr, 6 * n * sizeof(float)));
        float* d = (float*)st.ptr;
        const float* src[6] = {a1, a2, phi, ref, pha, coh};
        for (int k = 0; k < 6; k++)
            DKX_HIP_CHECK(hipMemcpyAsync(d + k * n, src[k], n * sizeof(float), hipMemcpyHostToDevice, w->stream));
        rc = nlinsar_device_impl(call, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, nullptr,
                                 w->stream, w);
        if (rc) { (void)hipStreamSynchronize(w->stream); return rc; }
        DKX_HIP_CHECK(hipMemcpyAsync(ref, d + 3 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(pha, d + 4 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(coh, d + 5 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipStreamSynchronize(w->stream));
        return 0;
    });
}

int dkx_goldstein_device(const float* a1, const float* a2, const float* phi, float* ref,
                         float* pha, float* coh, int height, int width, int patch_size,
                         int overlap, float alpha, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        (void)coh;   // not written, as in the reference
        if (!a1 || !a2 || !phi || !ref || !pha) { set_error("goldstein: null array argument"); return DKX_EINVAL; }
        GoldsteinParams gp;
        GoldsteinPlan pl;
        int rc = goldstein_prepare(gp, pl, height, width, patch_size, overlap, alpha);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        return goldstein_device_impl(gp, pl, a1, a2, phi, ref, pha, (hipStream_t)stream, w);
    });
}

int dkx_goldstein(const float* a1, const float* a2, const float* phi, float* ref, float* pha,
                  float* coh, int height, int width, int patch_size, int overlap, float alpha)
{
    return dkx::guarded<int>([&]() -> int {
        (void)coh;   // not written, as in the reference
        if (!a1 || !a2 || !phi || !ref || !pha) { set_error("goldstein: null array argument"); return DKX_EINVAL; }
        GoldsteinParams gp;
        GoldsteinPlan pl;
        int rc = goldstein_prepare(gp, pl, height, width, patch_size, overlap, alpha);
        if (rc) return rc;
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) { set_error("no HIP device"); return DKX_EHIP; }
        std::lock_guard<std::mutex> lk(w->mu);
        rc = ensure_stream(w);
        if (rc) return rc;
        const size_t n = (size_t)height * width;
        HostStage st;
        DKX_HIP_CHECK(hipMalloc(&st.ptr, 5 * n * sizeof(float)));
        float* d = (float*)st.ptr;
        const float* src[3] = {a1, a2, phi};
        for (int k = 0; k < 3; k++)
            DKX_HIP_CHECK(hipMemcpyAsync(d + k * n, src[k], n * sizeof(float), hipMemcpyHostToDevice, w->stream));
        rc = goldstein_device_impl(gp, pl, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, w->stream, w);
        if (rc) { (void)hipStreamSynchronize(w->stream); return rc; }
        DKX_HIP_CHECK(hipMemcpyAsync(ref, d + 3 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipMemcpyAsync(pha, d + 4 * n, n * sizeof(float), hipMemcpyDeviceToHost, w->stream));
        DKX_HIP_CHECK(hipStreamSynchronize(w->stream));
        return 0;
    });
}

void dkx_release_workspace(void)
{
    dkx::guarded_void([&] {
        DeviceWorkspace* w = workspace_for_current_device();
        if (!w) return;
        std::lock_guard<std::mutex> lk(w->mu);
        if (w->ptr) {
            if (w->have_done) (void)hipEventSynchronize(w->done);
            (void)hipFree(w->ptr);
        }
        w->ptr = nullptr;
        w->bytes = 0;
    });
}

}  // extern "C"
