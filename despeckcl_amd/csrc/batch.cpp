// Batched NL-SAR: many interferograms of uneven size that share the filter parameters,
// filtered with one launch per stage over the tiles of all images.  A batch object owns its
// device workspace (the grids of every image), the device table of image descriptors, the
// device constants (weight LUTs, Gaussian taps) and, once captured, a HIP graph of the whole
// stage sequence, replayed while neither the images nor the parameters change.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlsar.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace dkx {

struct NlsarBatch {
    // parameters (owned copies)
    std::vector<int> P, S;
    std::vector<std::vector<float>> lut_q, lut_c;
    std::vector<NlsarStatsView> views;
    NlsarParams pp{};
    NlsarCommon cm{};
    int tile = -1;               // -1: the autotuner's choice for this batch shape, else fixed
    int tile_used = NLSAR_TILE_AUTO;
    int diagnostics = 0;         // statistics of every key (undefined-selection counts)
    // images
    std::vector<NlsarImageDesc> descs;
    size_t ws_bytes = 0;
    // device state
    void* ws = nullptr;
    size_t ws_cap = 0;
    float* consts = nullptr;
    size_t consts_cap = 0;
    NlsarImageDesc* table = nullptr;   // lives in ws
    unsigned int* queues = nullptr;    // the persistent passes' tile counters, in ws
    hipEvent_t done = nullptr;         // completion of the last enqueue
    bool pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
};

static double batch_pixels(const NlsarBatch& b)
{
    double px = 0;
    for (const NlsarImageDesc& d : b.descs) px += (double)d.H * d.W;
    return px;
}

static int batch_wait(NlsarBatch* b)
{
    if (b->pending) DKX_HIP_CHECK(hipEventSynchronize(b->done));
    b->pending = false;
    return 0;
}

static int drop_graph(NlsarBatch* b)
{
    if (b->exec) DKX_HIP_CHECK(hipGraphExecDestroy(b->exec));
    if (b->graph) DKX_HIP_CHECK(hipGraphDestroy(b->graph));
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_stream = nullptr;
    return 0;
}

static int choose_tile(NlsarBatch* b)
{
    int t = b->tile;
    if (b->pp.dim != 0) t = NLSAR_TILE_AUTO;   // covariance input: one pass variant per dimension
    else if (t < 0) {
        // the autotuner's record for this batch shape, else per pass as the single calls
        t = tuned_tile(0, b->diagnostics, b->pp.sw, b->P.data(), (int)b->P.size(), b->S.data(), (int)b->S.size(),
                       batch_pixels(*b));
        if (t < 0) t = NLSAR_TILE_AUTO;
    }
    b->tile_used = t;
    b->pp.tile = t;
    b->cm.tile = t;
    return 0;
}

static int batch_enqueue(NlsarBatch* b, hipStream_t stream)
{
    if (b->descs.empty()) { set_error("nlsar batch: no images set"); return DKX_EINVAL; }
    int rc = nlsar_enqueue(b->pp, b->cm, b->descs.data(), (int)b->descs.size(), b->table, b->consts, stream,
                           b->queues);
    if (rc) return rc;
    return 0;
}

static int batch_mark(NlsarBatch* b, hipStream_t stream)
{
    if (!b->done) DKX_HIP_CHECK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    DKX_HIP_CHECK(hipEventRecord(b->done, stream));
    b->pending = true;
    return 0;
}

// The images of a batch: InSAR triples (dim 0) or covariance matrices of one dimension.
struct BatchImages {
    const dkx_nlsar_image* insar = nullptr;
    const dkx_nlsar_covmat_image* cov = nullptr;
    int n = 0, dim = 0;
    int height(int i) const { return insar ? insar[i].height : cov[i].height; }
    int width(int i) const { return insar ? insar[i].width : cov[i].width; }
};

// dkx_nlsar_batch_set / _set_covmat
static int batch_set(NlsarBatch* b, const BatchImages& im, int sw, const int* P, int nP, const int* S, int nS,
                     const dkx_nlsar_stats* stats, int n_stats, float hpar, float cpar)
{
        const int n = im.n;
        if (!b || (!im.insar && !im.cov) || n < 1) { set_error("nlsar batch: no images"); return DKX_EINVAL; }
        if (!P || !S || nP < 1 || nS < 1) { set_error("nlsar batch: empty patch or scale size list"); return DKX_EINVAL; }
        if (im.cov && im.dim != 2 && im.dim != 3) {
            set_error("nlsar batch: covariance matrices of dimension 2 or 3 are supported");
            return im.dim == 1 ? DKX_EUNSUPPORTED : DKX_EINVAL;
        }
        for (int i = 0; i < n; i++) {
            const bool null_array = im.insar ? (!im.insar[i].ampl_master || !im.insar[i].ampl_slave || !im.insar[i].phase ||
                                                !im.insar[i].ref_filt || !im.insar[i].phase_filt || !im.insar[i].coh_filt)
                                             : (!im.cov[i].covmat || !im.cov[i].covmat_filt);
            if (null_array) {
                set_error("nlsar batch: null array in image " + std::to_string(i));
                return DKX_EINVAL;
            }
        }
        std::vector<NlsarStatsView> views;
        int last_unused = 0;
        int rc = check_stats(stats, n_stats, P, nP, S, nS, views, &last_unused);
        if (rc) return rc;
        // everything that can fail without the device is checked before touching the batch
        NlsarBatch nb;
        nb.P.assign(P, P + nP);
        nb.S.assign(S, S + nS);
        for (const NlsarStatsView& v : views) {
            nb.lut_q.emplace_back(v.quantilles, v.quantilles + v.lut_size);
            nb.lut_c.emplace_back(v.chi2cdf_inv, v.chi2cdf_inv + v.lut_size);
        }
        for (size_t k = 0; k < views.size(); k++) {
            NlsarStatsView v = views[k];
            v.quantilles = nb.lut_q[k].data();
            v.chi2cdf_inv = nb.lut_c[k].data();
            nb.views.push_back(v);
        }
        nb.pp = NlsarParams{im.cov ? im.dim : 0, sw, nb.P.data(), nP, nb.S.data(), nS, nb.views.data(), (int)nb.views.size(),
                            hpar, cpar, 1, b->diagnostics, last_unused};
        rc = nlsar_common(nb.pp, &nb.cm);
        if (rc) return rc;
        std::vector<NlsarImagePlan> plans(n);
        // the table first, then the tile counters
        const size_t qoff = (sizeof(NlsarImageDesc) * (size_t)n + 255) / 256 * 256;
        size_t off = qoff + (NLSAR_QUEUE_BYTES + 255) / 256 * 256;
        std::vector<size_t> offs(n);
        for (int i = 0; i < n; i++) {
            rc = nlsar_image_plan(nb.cm, im.height(i), im.width(i), -1, -1, &plans[i]);
            if (rc) return rc;
            offs[i] = off;
            off += plans[i].bytes;
        }
        std::vector<float> blob;
        rc = nlsar_build_constants(nb.pp, nb.cm, blob);
        if (rc) return rc;

        // device side: wait for the batch's last run, grow the workspace, upload
        rc = batch_wait(b);
        if (rc) return rc;
        rc = drop_graph(b);
        if (rc) return rc;
        if (b->ws_cap < off) {
            if (b->ws) DKX_HIP_CHECK(hipFree(b->ws));
            b->ws = nullptr;
            b->ws_cap = 0;
            hipError_t e = hipMalloc(&b->ws, off);
            if (e != hipSuccess) {
                b->ws = nullptr;
                set_error("nlsar batch: workspace allocation of " + std::to_string(off) + " bytes failed: " +
                          hipGetErrorString(e));
                return DKX_ERESOURCE;
            }
            b->ws_cap = off;
        }
        if (b->consts_cap < blob.size()) {
            if (b->consts) DKX_HIP_CHECK(hipFree(b->consts));
            b->consts = nullptr;
            b->consts_cap = 0;
            DKX_HIP_CHECK(hipMalloc(&b->consts, blob.size() * sizeof(float)));
            b->consts_cap = blob.size();
        }
        DKX_HIP_CHECK(hipMemcpy(b->consts, blob.data(), blob.size() * sizeof(float), hipMemcpyHostToDevice));
        nb.descs.resize(n);
        for (int i = 0; i < n; i++) {
            char* ws_i = (char*)b->ws + offs[i];
            if (im.insar) {
                const dkx_nlsar_image& x = im.insar[i];
                nlsar_image_desc(plans[i], ws_i, x.ampl_master, x.ampl_slave, x.phase, nullptr, x.ref_filt, x.phase_filt,
                                 x.coh_filt, nullptr, nullptr, nullptr, &nb.descs[i]);
            } else {
                const dkx_nlsar_covmat_image& x = im.cov[i];
                nlsar_image_desc(plans[i], ws_i, nullptr, nullptr, nullptr, x.covmat, nullptr, nullptr, nullptr,
                                 x.covmat_filt, nullptr, nullptr, &nb.descs[i]);
            }
        }
        int rows_c, rows_a, blocks, mcw, maw;
        nlsar_batch_layout(nb.descs.data(), n, &rows_c, &rows_a, &blocks, &mcw, &maw);
        DKX_HIP_CHECK(hipMemcpy(b->ws, nb.descs.data(), sizeof(NlsarImageDesc) * (size_t)n, hipMemcpyHostToDevice));

        b->P = std::move(nb.P);
        b->S = std::move(nb.S);
        b->lut_q = std::move(nb.lut_q);
        b->lut_c = std::move(nb.lut_c);
        b->views = std::move(nb.views);
        b->pp = nb.pp;
        b->pp.P = b->P.data();
        b->pp.S = b->S.data();
        b->pp.stats = b->views.data();
        b->cm = nb.cm;
        b->descs = std::move(nb.descs);
        b->ws_bytes = off;
        b->table = (NlsarImageDesc*)b->ws;
        b->queues = (unsigned int*)((char*)b->ws + qoff);
        return choose_tile(b);
}

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_nlsar_batch_create(dkx_nlsar_batch* out)
{
    return dkx::guarded<int>([&]() -> int {
        if (!out) { set_error("nlsar batch: null handle pointer"); return DKX_EINVAL; }
        *out = reinterpret_cast<dkx_nlsar_batch>(new NlsarBatch());
        return 0;
    });
}

void dkx_nlsar_batch_destroy(dkx_nlsar_batch h)
{
    dkx::guarded_void([&] {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b) return;
        (void)batch_wait(b);
        (void)drop_graph(b);
        if (b->done) (void)hipEventDestroy(b->done);
        if (b->ws) (void)hipFree(b->ws);
        if (b->consts) (void)hipFree(b->consts);
        delete b;
    });
}

int dkx_nlsar_batch_set(dkx_nlsar_batch h, const dkx_nlsar_image* images, int n, int sw, const int* P,
                        int nP, const int* S, int nS, const dkx_nlsar_stats* stats, int n_stats,
                        float hpar, float cpar)
{
    return dkx::guarded<int>([&]() -> int {
        BatchImages im;
        im.insar = images;
        im.n = n;
        return batch_set(reinterpret_cast<NlsarBatch*>(h), im, sw, P, nP, S, nS, stats, n_stats, hpar, cpar);
    });
}

int dkx_nlsar_batch_set_covmat(dkx_nlsar_batch h, const dkx_nlsar_covmat_image* images, int n, int dim, int sw,
                               const int* P, int nP, const int* S, int nS, const dkx_nlsar_stats* stats,
                               int n_stats, float hpar, float cpar)
{
    return dkx::guarded<int>([&]() -> int {
        BatchImages im;
        im.cov = images;
        im.n = n;
        im.dim = dim;
        return batch_set(reinterpret_cast<NlsarBatch*>(h), im, sw, P, nP, S, nS, stats, n_stats, hpar, cpar);
    });
}

int dkx_nlsar_batch_set_tile(dkx_nlsar_batch h, int tile)
{
    return dkx::guarded<int>([&]() -> int {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b || tile < -1 || tile >= NLSAR_NTILES) { set_error("nlsar batch: invalid tile"); return DKX_EINVAL; }
        int rc = batch_wait(b);
        if (rc) return rc;
        rc = drop_graph(b);
        if (rc) return rc;
        b->tile = tile;
        if (!b->descs.empty()) return choose_tile(b);
        return 0;
    });
}

int dkx_nlsar_batch_set_diagnostics(dkx_nlsar_batch h, int on)
{
    return dkx::guarded<int>([&]() -> int {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b) { set_error("nlsar batch: null handle"); return DKX_EINVAL; }
        int rc = batch_wait(b);
        if (rc) return rc;
        rc = drop_graph(b);
        if (rc) return rc;
        b->diagnostics = on ? 1 : 0;
        b->pp.all_keys = b->diagnostics;
        return 0;
    });
}

int dkx_nlsar_batch_get_tile(dkx_nlsar_batch h)
{
    NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
    return b ? b->tile_used : -1;
}

int dkx_nlsar_batch_enqueue(dkx_nlsar_batch h, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b) { set_error("nlsar batch: null handle"); return DKX_EINVAL; }
        int rc = batch_enqueue(b, (hipStream_t)stream);
        if (rc) return rc;
        return batch_mark(b, (hipStream_t)stream);
    });
}

int dkx_nlsar_batch_graph_launch(dkx_nlsar_batch h, void* stream_)
{
    return dkx::guarded<int>([&]() -> int {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b) { set_error("nlsar batch: null handle"); return DKX_EINVAL; }
        hipStream_t stream = (hipStream_t)stream_;
        if (!stream) { set_error("nlsar batch: graph launches need a created stream"); return DKX_EINVAL; }
        if (!b->exec) {
            DKX_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            int rc = batch_enqueue(b, stream);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                set_error(std::string("nlsar batch: graph capture failed: ") + hipGetErrorString(e));
                return DKX_EHIP;
            }
            b->graph = g;
            DKX_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_stream = stream;
        }
        DKX_HIP_CHECK(hipGraphLaunch(b->exec, stream));
        return batch_mark(b, stream);
    });
}

int dkx_nlsar_batch_ub_pixels(dkx_nlsar_batch h, unsigned long long* counts)
{
    return dkx::guarded<int>([&]() -> int {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b || !counts) { set_error("nlsar batch: null argument"); return DKX_EINVAL; }
        if (!b->diagnostics) {
            set_error("nlsar batch: undefined-selection counts need the statistics of every key "
                      "(dkx_nlsar_batch_set_diagnostics)");
            return DKX_EINVAL;
        }
        int rc = batch_wait(b);
        if (rc) return rc;
        for (size_t i = 0; i < b->descs.size(); i++)
            DKX_HIP_CHECK(hipMemcpy(&counts[i], b->descs[i].ub_count, sizeof(unsigned long long), hipMemcpyDeviceToHost));
        return 0;
    });
}

int dkx_nlsar_batch_autotune(dkx_nlsar_batch h, void* stream_, int reps, int* tile_out, float* ms_out)
{
    return dkx::guarded<int>([&]() -> int {
        NlsarBatch* b = reinterpret_cast<NlsarBatch*>(h);
        if (!b || b->descs.empty()) { set_error("nlsar batch: no images set"); return DKX_EINVAL; }
        if (b->pp.dim != 0) {
            // covariance input: one pass variant per dimension, nothing to choose or record
            if (tile_out) *tile_out = -1;
            if (ms_out)
                for (int t = 0; t < NLSAR_NTILES; t++) ms_out[t] = 0.f;
            return 0;
        }
        hipStream_t stream = (hipStream_t)stream_;
        reps = std::max(reps, 1);
        int rc = batch_wait(b);
        if (rc) return rc;
        rc = drop_graph(b);
        if (rc) return rc;
        hipEvent_t e0, e1;
        DKX_HIP_CHECK(hipEventCreate(&e0));
        DKX_HIP_CHECK(hipEventCreate(&e1));
        float ms_t[NLSAR_NTILES];
        for (int t = 0; t < NLSAR_NTILES; t++) {
            b->pp.tile = t;
            b->cm.tile = t;
            rc = batch_enqueue(b, stream);   // warm-up
            if (rc) break;
            DKX_HIP_CHECK(hipEventRecord(e0, stream));
            for (int r = 0; r < reps && !rc; r++) rc = batch_enqueue(b, stream);
            if (rc) break;
            DKX_HIP_CHECK(hipEventRecord(e1, stream));
            DKX_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            DKX_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (ms_out) ms_out[t] = ms;
            ms_t[t] = ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        const int best = nlsar_pick_variant(ms_t);
        record_tuned_tile(0, b->diagnostics, b->pp.sw, b->P.data(), (int)b->P.size(), b->S.data(), (int)b->S.size(),
                          batch_pixels(*b), best);
        if (tile_out) *tile_out = best;
        b->tile = -1;
        return choose_tile(b);
    });
}

}  // extern "C"
