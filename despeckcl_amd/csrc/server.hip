// Serving: a request loop around the filters.  Requests of mixed filter type and image size
// arrive at any time; a scheduler thread admits them against a fixed HBM budget (one arena,
// sub-allocated per request, so serving never calls hipMalloc), batches the waiting requests
// of each filter into one launch per stage (continuous batching: a new batch forms as soon as
// a lane of that filter and memory are free, while earlier work still runs; the NL-SAR and
// NL-InSAR lanes run on streams of their own, so their work overlaps), and records each
// request's completion time from a host callback on its stream.  The weight LUTs and
// Gaussian taps stay resident on the device for the server's lifetime.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlinsar.h"
#include "dkx_nlsar.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace dkx {

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device allocation (offsets aligned to 256 bytes).
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
    std::map<size_t, size_t> free_;   // offset -> size
    void reset(char* b, size_t c)
    {
        base = b;
        cap = c;
        free_.clear();
        free_[0] = c;
    }
    bool alloc(size_t n, size_t* off)
    {
        n = (n + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const size_t rest = it->second - n, o = it->first;
            free_.erase(it);
            if (rest) free_[o + n] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = (n + 255) / 256 * 256;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Request {
    long long id;
    dkx_request r;
    double t_arrive = 0, t_done = -1;
    int status = 0;   // 0 waiting, 1 on the GPU, 2 done, < 0 failed
    size_t off = 0, bytes = 0;
};

struct Server;

// One unit of GPU work: an NL-SAR or an NL-InSAR batch.
struct Flight {
    Server* srv = nullptr;
    int kind = 0, lane = 0;
    std::vector<Request*> reqs;
    std::atomic<bool> done{false};
    std::atomic<double> t_done{0.0};
};

constexpr int SAR_LANES = 2;
constexpr int NLI_LANES = 2;
constexpr int MAX_BATCH = 256;
constexpr int MAX_NLI_BATCH = 128;
// an NL-InSAR batch stops growing once it holds this many pixels (one 4096 x 4096 image), so
// that small requests do not wait behind a long batch while the other lane is free
constexpr long long NLI_BATCH_PIXELS = 1ll << 24;

struct Server {
    // NL-SAR parameters (owned copies) and resident constants
    std::vector<int> P, S;
    std::vector<std::vector<float>> lut_q, lut_c;
    std::vector<NlsarStatsView> views;
    NlsarParams pp{};
    NlsarCommon cm{};
    float* consts = nullptr;
    // NL-InSAR parameters
    NlinsarProblem nl_pb{};
    // device memory and streams
    Arena arena;
    void* arena_ptr = nullptr;
    hipStream_t sar_stream[SAR_LANES] = {};
    hipStream_t nli_stream[NLI_LANES] = {};
    NlsarImageDesc* pinned_table[SAR_LANES] = {};   // staging of a lane's batch table
    NlsarImageDesc* dev_table[SAR_LANES] = {};      // the table itself (MAX_BATCH entries)
    unsigned int* dev_queues[SAR_LANES] = {};       // its persistent passes' tile counters
    unsigned char* nli_pinned[NLI_LANES] = {};     // staging of a lane's NL-InSAR tables
    char* nli_table[NLI_LANES] = {};               // the tables (MAX_NLI_BATCH images)
    Flight* sar_busy[SAR_LANES] = {};
    Flight* nli_busy[NLI_LANES] = {};
    // queues
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> incoming;
    std::deque<Request*> wait_sar, wait_nli;
    std::unordered_map<long long, Request*> reqs;
    long long next_id = 1;
    std::atomic<bool> stop{false};
    std::atomic<long long> n_done{0};
    std::thread sched;
    std::string sched_error;
    // counters
    long long batches = 0, batched_images = 0;
    long long nli_batches = 0, nli_batched_images = 0;
};

// Completion callback on a flight's stream.  Once `done` is set the scheduler may retire and
// delete the flight, so nothing of it is touched afterwards (the server outlives its streams).
static void host_done(void* p)
{
    Flight* f = (Flight*)p;
    Server* s = f->srv;
    f->t_done.store(now_s());
    f->done.store(true);
    s->cv.notify_all();
}

static void fail_request(Server* s, Request* r, const std::string& msg)
{
    r->status = -1;
    r->t_done = now_s();
    s->sched_error = msg;
    s->n_done.fetch_add(1);
}

// retire finished flights: free their memory, stamp their requests
static void retire(Server* s)
{
    auto fin = [&](Flight*& f) {
        if (!f || !f->done.load()) return;
        for (Request* r : f->reqs) {
            r->t_done = f->t_done.load();
            r->status = 2;
            s->arena.release(r->off, r->bytes);
            s->n_done.fetch_add(1);
        }
        delete f;
        f = nullptr;
    };
    for (int i = 0; i < SAR_LANES; i++) fin(s->sar_busy[i]);
    for (int i = 0; i < NLI_LANES; i++) fin(s->nli_busy[i]);
}

static void launch_sar(Server* s, int lane)
{
    if (s->wait_sar.empty()) return;
    Flight* f = new Flight();
    f->srv = s;
    f->kind = 0;
    f->lane = lane;
    std::vector<NlsarImageDesc> descs;
    while (!s->wait_sar.empty() && (int)f->reqs.size() < MAX_BATCH) {
        Request* r = s->wait_sar.front();
        NlsarImagePlan ip;
        if (nlsar_image_plan(s->cm, r->r.height, r->r.width, -1, -1, &ip)) {
            s->wait_sar.pop_front();
            fail_request(s, r, last_error());
            continue;
        }
        size_t off;
        if (!s->arena.alloc(ip.bytes, &off)) break;   // wait for memory to free up
        s->wait_sar.pop_front();
        r->off = off;
        r->bytes = ip.bytes;
        r->status = 1;
        NlsarImageDesc d;
        nlsar_image_desc(ip, s->arena.base + off, r->r.ampl_master, r->r.ampl_slave, r->r.phase, nullptr,
                         r->r.ref_filt, r->r.phase_filt, r->r.coh_filt, nullptr, nullptr, nullptr, &d);
        descs.push_back(d);
        f->reqs.push_back(r);
    }
    if (f->reqs.empty()) {
        delete f;
        return;
    }
    const int n = (int)descs.size();
    int rows_c, rows_a, blocks, mcw, maw;
    nlsar_batch_layout(descs.data(), n, &rows_c, &rows_a, &blocks, &mcw, &maw);
    // the pass variant: the autotuner's record for this batch's pixel count, else per pass
    double px = 0;
    for (const NlsarImageDesc& d : descs) px += (double)d.H * d.W;
    int tile = tuned_tile(0, 0, s->pp.sw, s->P.data(), (int)s->P.size(), s->S.data(), (int)s->S.size(), px);
    if (tile < 0) tile = NLSAR_TILE_AUTO;
    s->pp.tile = tile;
    s->cm.tile = tile;
    // the lane's table was reserved at creation; its previous batch has finished (the lane
    // was free) and the stream orders the copy after it
    const size_t table_bytes = sizeof(NlsarImageDesc) * (size_t)n;
    hipStream_t st = s->sar_stream[lane];
    int rc = 0;
    memcpy(s->pinned_table[lane], descs.data(), table_bytes);
    NlsarImageDesc* d_table = s->dev_table[lane];
    if (hipMemcpyAsync(d_table, s->pinned_table[lane], table_bytes, hipMemcpyHostToDevice, st) != hipSuccess) rc = -5;
    if (!rc) rc = nlsar_enqueue(s->pp, s->cm, descs.data(), n, d_table, s->consts, st, s->dev_queues[lane]);
    if (!rc && hipLaunchHostFunc(st, host_done, f) != hipSuccess) rc = -5;
    if (rc) {
        // nothing reached the GPU that still needs the memory once the stream drains
        (void)hipStreamSynchronize(st);
        for (Request* r : f->reqs) {
            s->arena.release(r->off, r->bytes);
            fail_request(s, r, last_error());
        }
        delete f;
        return;
    }
    s->batches++;
    s->batched_images += n;
    s->sar_busy[lane] = f;
}

static void launch_nli(Server* s, int lane)
{
    if (s->wait_nli.empty()) return;
    Flight* f = new Flight();
    f->srv = s;
    f->kind = 1;
    f->lane = lane;
    std::vector<int> H, W;
    long long px = 0;
    while (!s->wait_nli.empty() && (int)f->reqs.size() < MAX_NLI_BATCH && px < NLI_BATCH_PIXELS) {
        Request* r = s->wait_nli.front();
        size_t bytes = 0;
        if (nlinsar_batch_image_bytes(s->nl_pb, r->r.height, r->r.width, &bytes)) {
            s->wait_nli.pop_front();
            fail_request(s, r, last_error());
            continue;
        }
        size_t off;
        if (!s->arena.alloc(bytes, &off)) break;   // wait for memory to free up
        s->wait_nli.pop_front();
        r->off = off;
        r->bytes = bytes;
        r->status = 1;
        f->reqs.push_back(r);
        H.push_back(r->r.height);
        W.push_back(r->r.width);
        px += (long long)r->r.height * r->r.width;
    }
    if (f->reqs.empty()) {
        delete f;
        return;
    }
    const int n = (int)f->reqs.size();
    hipStream_t st = s->nli_stream[lane];
    std::vector<NlinsarPlan> plans;
    NlinsarBatchLaunch L;
    int rc = nlinsar_batch_plan(s->nl_pb, H.data(), W.data(), n, plans, &L);
    std::vector<NlinsarBatchItem> items(n);
    for (int i = 0; i < n && !rc; i++) {
        const Request* r = f->reqs[i];
        if (plans[i].ws_bytes > r->bytes) {
            set_error("serve: NL-InSAR workspace bound exceeded");
            rc = -1;
            break;
        }
        items[i] = NlinsarBatchItem{r->r.ampl_master, r->r.ampl_slave, r->r.phase, r->r.ref_filt, r->r.phase_filt,
                                    r->r.coh_filt, r->r.height, r->r.width, s->arena.base + r->off};
    }
    std::vector<unsigned char> host;
    if (!rc) rc = nlinsar_batch_tables(s->nl_pb, plans, items.data(), L, s->nli_table[lane], host);
    // the lane's tables were reserved at creation; its previous batch has finished (the lane
    // was free) and the stream orders the copy after it
    if (!rc) {
        memcpy(s->nli_pinned[lane], host.data(), host.size());
        if (hipMemcpyAsync(s->nli_table[lane], s->nli_pinned[lane], host.size(), hipMemcpyHostToDevice, st) !=
            hipSuccess)
            rc = -5;
    }
    // every request starts from zero estimates, like the Python API
    if (!rc) rc = nlinsar_batch_enqueue(s->nl_pb, L, s->nli_table[lane], st, 1);
    if (!rc && hipLaunchHostFunc(st, host_done, f) != hipSuccess) rc = -5;
    if (rc) {
        (void)hipStreamSynchronize(st);
        for (Request* r : f->reqs) {
            s->arena.release(r->off, r->bytes);
            fail_request(s, r, last_error());
        }
        delete f;
        return;
    }
    s->nli_batches++;
    s->nli_batched_images += n;
    s->nli_busy[lane] = f;
}

static void scheduler(Server* s)
{
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        while (!s->incoming.empty()) {
            Request* r = s->incoming.front();
            s->incoming.pop_front();
            (r->r.kind == 0 ? s->wait_sar : s->wait_nli).push_back(r);
        }
        retire(s);
        for (int i = 0; i < SAR_LANES; i++)
            if (!s->sar_busy[i]) launch_sar(s, i);
        for (int i = 0; i < NLI_LANES; i++)
            if (!s->nli_busy[i]) launch_nli(s, i);
        // a request larger than the whole budget can never be admitted
        auto reject_oversized = [&](std::deque<Request*>& q, bool sar) {
            bool idle = true;
            for (int i = 0; i < SAR_LANES; i++) idle &= s->sar_busy[i] == nullptr;
            for (int i = 0; i < NLI_LANES; i++) idle &= s->nli_busy[i] == nullptr;
            if (!idle || q.empty()) return;
            Request* r = q.front();
            q.pop_front();
            fail_request(s, r, sar ? "serve: NL-SAR request exceeds the HBM budget"
                                   : "serve: NL-InSAR request exceeds the HBM budget");
        };
        reject_oversized(s->wait_sar, true);
        reject_oversized(s->wait_nli, false);
        if (s->stop.load() && s->incoming.empty() && s->wait_sar.empty() && s->wait_nli.empty()) {
            bool busy = false;
            for (int i = 0; i < SAR_LANES; i++) busy |= s->sar_busy[i] != nullptr;
            for (int i = 0; i < NLI_LANES; i++) busy |= s->nli_busy[i] != nullptr;
            if (!busy) break;
        }
        // completions notify without the lock, so wake up periodically as well
        s->cv.wait_for(lk, std::chrono::microseconds(200));
    }
}

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_server_create(unsigned long long hbm_budget_bytes, int search_window_size, const int* patch_sizes,
                      int n_patch_sizes, const int* scale_sizes, int n_scale_sizes, const dkx_nlsar_stats* stats,
                      int n_stats, float h_param, float c_param, int nl_search_window_size, int nl_patch_size,
                      int nl_niter, int nl_lmin, dkx_server* out)
{
    return dkx::guarded<int>([&]() -> int {
        if (!out) { set_error("serve: null handle pointer"); return DKX_EINVAL; }
        if (!patch_sizes || !scale_sizes || n_patch_sizes < 1 || n_scale_sizes < 1) {
            set_error("serve: empty patch or scale size list");
            return DKX_EINVAL;
        }
        std::vector<NlsarStatsView> views;
        int last_unused = 0;
        int rc = check_stats(stats, n_stats, patch_sizes, n_patch_sizes, scale_sizes, n_scale_sizes, views,
                             &last_unused);
        if (rc) return rc;
        NlinsarProblem nl_pb;
        rc = nlinsar_problem(0, 0, nl_search_window_size, nl_patch_size, nl_niter, nl_lmin, NLINSAR_TILE_AUTO,
                             &nl_pb);
        if (rc) return rc;
        nl_pb.store = dkx_get_nlinsar_weight_store();
        {
            // the parameter checks of a plan, on a one-pixel image
            NlinsarProblem probe = nl_pb;
            probe.H = probe.W = 1;
            NlinsarPlan plan;
            rc = nlinsar_plan(probe, &plan);
            if (rc) return rc;
        }
        Server* s = new Server();
        s->P.assign(patch_sizes, patch_sizes + n_patch_sizes);
        s->S.assign(scale_sizes, scale_sizes + n_scale_sizes);
        for (const NlsarStatsView& v : views) {
            s->lut_q.emplace_back(v.quantilles, v.quantilles + v.lut_size);
            s->lut_c.emplace_back(v.chi2cdf_inv, v.chi2cdf_inv + v.lut_size);
        }
        for (size_t k = 0; k < views.size(); k++) {
            NlsarStatsView v = views[k];
            v.quantilles = s->lut_q[k].data();
            v.chi2cdf_inv = s->lut_c[k].data();
            s->views.push_back(v);
        }
        s->pp = NlsarParams{0, search_window_size, s->P.data(), n_patch_sizes, s->S.data(), n_scale_sizes,
                            s->views.data(), (int)s->views.size(), h_param, c_param, NLSAR_TILE_AUTO, 0,
                            last_unused};
        rc = nlsar_common(s->pp, &s->cm);
        std::vector<float> blob;
        if (!rc) rc = nlsar_build_constants(s->pp, s->cm, blob);
        if (rc) {
            delete s;
            return rc;
        }
        s->nl_pb = nl_pb;
        auto cleanup = [&](const std::string& msg) {
            set_error(msg);
            if (s->consts) (void)hipFree(s->consts);
            if (s->arena_ptr) (void)hipFree(s->arena_ptr);
            for (int i = 0; i < SAR_LANES; i++) {
                if (s->sar_stream[i]) (void)hipStreamDestroy(s->sar_stream[i]);
                if (s->pinned_table[i]) (void)hipHostFree(s->pinned_table[i]);
                if (s->dev_table[i]) (void)hipFree(s->dev_table[i]);
                if (s->dev_queues[i]) (void)hipFree(s->dev_queues[i]);
            }
            for (int i = 0; i < NLI_LANES; i++) {
                if (s->nli_stream[i]) (void)hipStreamDestroy(s->nli_stream[i]);
                if (s->nli_pinned[i]) (void)hipHostFree(s->nli_pinned[i]);
                if (s->nli_table[i]) (void)hipFree(s->nli_table[i]);
            }
            delete s;
            return DKX_ERESOURCE;
        };
        if (hipMalloc(&s->consts, blob.size() * sizeof(float)) != hipSuccess) return cleanup("serve: constants allocation failed");
        if (hipMemcpy(s->consts, blob.data(), blob.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup("serve: constants upload failed");
        if (hipMalloc(&s->arena_ptr, (size_t)hbm_budget_bytes) != hipSuccess)
            return cleanup("serve: the HBM budget of " + std::to_string(hbm_budget_bytes) + " bytes cannot be allocated");
        s->arena.reset((char*)s->arena_ptr, (size_t)hbm_budget_bytes);
        for (int i = 0; i < SAR_LANES; i++) {
            if (hipStreamCreateWithFlags(&s->sar_stream[i], hipStreamNonBlocking) != hipSuccess) return cleanup("serve: stream creation failed");
            if (hipHostMalloc((void**)&s->pinned_table[i], sizeof(NlsarImageDesc) * MAX_BATCH, 0) != hipSuccess)
                return cleanup("serve: pinned staging allocation failed");
            if (hipMalloc((void**)&s->dev_table[i], sizeof(NlsarImageDesc) * MAX_BATCH) != hipSuccess)
                return cleanup("serve: batch table allocation failed");
            if (hipMalloc((void**)&s->dev_queues[i], NLSAR_QUEUE_BYTES) != hipSuccess)
                return cleanup("serve: batch table allocation failed");
        }
        const size_t nli_tables = nlinsar_batch_table_bytes(MAX_NLI_BATCH);
        for (int i = 0; i < NLI_LANES; i++) {
            if (hipStreamCreateWithFlags(&s->nli_stream[i], hipStreamNonBlocking) != hipSuccess) return cleanup("serve: stream creation failed");
            if (hipHostMalloc((void**)&s->nli_pinned[i], nli_tables, 0) != hipSuccess)
                return cleanup("serve: pinned staging allocation failed");
            if (hipMalloc((void**)&s->nli_table[i], nli_tables) != hipSuccess)
                return cleanup("serve: batch table allocation failed");
        }
        s->sched = std::thread(scheduler, s);
        *out = reinterpret_cast<dkx_server>(s);
        return 0;
    });
}

int dkx_server_submit(dkx_server h, const dkx_request* req, double arrival_time_s, long long* id)
{
    return dkx::guarded<int>([&]() -> int {
        Server* s = reinterpret_cast<Server*>(h);
        if (!s || !req) { set_error("serve: null argument"); return DKX_EINVAL; }
        if ((req->kind != 0 && req->kind != 1) || req->height < 1 || req->width < 1 || !req->ampl_master ||
            !req->ampl_slave || !req->phase || !req->ref_filt || !req->phase_filt || !req->coh_filt) {
            set_error("serve: invalid request");
            return DKX_EINVAL;
        }
        Request* r = new Request();
        r->r = *req;
        r->t_arrive = arrival_time_s > 0 ? arrival_time_s : now_s();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->stop.load()) {
                delete r;
                set_error("serve: server is shutting down");
                return DKX_EINVAL;
            }
            r->id = s->next_id++;
            s->reqs[r->id] = r;
            s->incoming.push_back(r);
        }
        s->cv.notify_all();
        if (id) *id = r->id;
        return 0;
    });
}

// 1 done (latency in ms from arrival to completion), 0 still pending, < 0 failed or unknown
int dkx_server_poll(dkx_server h, long long id, double* latency_ms)
{
    return dkx::guarded<int>([&]() -> int {
        Server* s = reinterpret_cast<Server*>(h);
        if (!s) { set_error("serve: null handle"); return DKX_EINVAL; }
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) { set_error("serve: unknown request"); return DKX_EINVAL; }
        Request* r = it->second;
        if (r->status == 2) {
            if (latency_ms) *latency_ms = (r->t_done - r->t_arrive) * 1e3;
            return 1;
        }
        if (r->status < 0) {
            set_error(s->sched_error.empty() ? "serve: request failed" : s->sched_error);
            return DKX_EHIP;
        }
        return 0;
    });
}

int dkx_server_wait(dkx_server h, long long id, double* latency_ms)
{
    return dkx::guarded<int>([&]() -> int {
        while (true) {
            const int rc = dkx_server_poll(h, id, latency_ms);
            if (rc != 0) return rc == 1 ? 0 : rc;
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    });
}

int dkx_server_run_trace(dkx_server h, const dkx_request* reqs, const double* arrival_s, int n, double* latency_ms,
                         double* makespan_s)
{
    return dkx::guarded<int>([&]() -> int {
        Server* s = reinterpret_cast<Server*>(h);
        if (!s || !reqs || !arrival_s || n < 0) { set_error("serve: null argument"); return DKX_EINVAL; }
        std::vector<long long> ids(n);
        const double t0 = now_s();
        for (int i = 0; i < n; i++) {
            const double ta = t0 + arrival_s[i];
            while (true) {
                const double dt = ta - now_s();
                if (dt <= 0) break;
                if (dt > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(dt - 1e-4));
            }
            const int rc = dkx_server_submit(h, &reqs[i], ta, &ids[i]);
            if (rc) return rc;
        }
        int first_err = 0;
        double last = t0;
        for (int i = 0; i < n; i++) {
            double lat = 0;
            const int rc = dkx_server_wait(h, ids[i], &lat);
            if (rc && !first_err) first_err = rc;
            if (latency_ms) latency_ms[i] = rc ? -1.0 : lat;
            last = std::max(last, t0 + arrival_s[i] + lat * 1e-3);
        }
        if (makespan_s) *makespan_s = last - t0;
        return first_err;
    });
}

int dkx_server_nlinsar_counters(dkx_server h, long long* batches, long long* batched_images)
{
    return dkx::guarded<int>([&]() -> int {
        Server* s = reinterpret_cast<Server*>(h);
        if (!s) { set_error("serve: null handle"); return DKX_EINVAL; }
        std::lock_guard<std::mutex> lk(s->mu);
        if (batches) *batches = s->nli_batches;
        if (batched_images) *batched_images = s->nli_batched_images;
        return 0;
    });
}

int dkx_server_counters(dkx_server h, long long* batches, long long* batched_images)
{
    return dkx::guarded<int>([&]() -> int {
        Server* s = reinterpret_cast<Server*>(h);
        if (!s) { set_error("serve: null handle"); return DKX_EINVAL; }
        std::lock_guard<std::mutex> lk(s->mu);
        if (batches) *batches = s->batches;
        if (batched_images) *batched_images = s->batched_images;
        return 0;
    });
}

void dkx_server_destroy(dkx_server h)
{
    dkx::guarded_void([&] {
        Server* s = reinterpret_cast<Server*>(h);
        if (!s) return;
        s->stop.store(true);
        s->cv.notify_all();
        if (s->sched.joinable()) s->sched.join();
        for (int i = 0; i < SAR_LANES; i++) {
            (void)hipStreamSynchronize(s->sar_stream[i]);
            (void)hipStreamDestroy(s->sar_stream[i]);
            (void)hipHostFree(s->pinned_table[i]);
            (void)hipFree(s->dev_table[i]);
            (void)hipFree(s->dev_queues[i]);
        }
        for (int i = 0; i < NLI_LANES; i++) {
            (void)hipStreamSynchronize(s->nli_stream[i]);
            (void)hipStreamDestroy(s->nli_stream[i]);
            (void)hipHostFree(s->nli_pinned[i]);
            (void)hipFree(s->nli_table[i]);
        }
        for (auto& kv : s->reqs) delete kv.second;
        (void)hipFree(s->arena_ptr);
        (void)hipFree(s->consts);
        delete s;
    });
}

}  // extern "C"
