// Batched NL-InSAR: many interferograms of uneven size that share the filter parameters,
// filtered with one launch per stage over all images (the weights on a persistent grid that
// drains a tile queue per XCD).  A batch object owns the device workspace of every image, the
// device tables of image descriptors and, once captured, a HIP graph of the whole stage
// sequence, replayed while neither the images nor the parameters change.
#include "../../include/despeckcl_amd.h"

#include "dkx_common.h"
#include "dkx_nlinsar.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace dkx {

struct NlinsarBatch {
    NlinsarProblem pb{};
    std::vector<NlinsarPlan> plans;
    NlinsarBatchLaunch L{};
    std::vector<size_t> ws_off;   // each image's workspace, relative to ws
    size_t tables_off = 0;
    // device state
    void* ws = nullptr;
    size_t ws_cap = 0;
    hipEvent_t done = nullptr;    // completion of the last enqueue
    bool pending = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int nb_wait(NlinsarBatch* b)
{
    if (b->pending) DKX_HIP_CHECK(hipEventSynchronize(b->done));
    b->pending = false;
    return 0;
}

static int nb_drop_graph(NlinsarBatch* b)
{
    if (b->exec) DKX_HIP_CHECK(hipGraphExecDestroy(b->exec));
    if (b->graph) DKX_HIP_CHECK(hipGraphDestroy(b->graph));
    b->exec = nullptr;
    b->graph = nullptr;
    return 0;
}

static int nb_mark(NlinsarBatch* b, hipStream_t stream)
{
    if (!b->done) DKX_HIP_CHECK(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    DKX_HIP_CHECK(hipEventRecord(b->done, stream));
    b->pending = true;
    return 0;
}

static int nb_enqueue(NlinsarBatch* b, hipStream_t stream)
{
    if (b->plans.empty()) { set_error("nlinsar batch: no images set"); return DKX_EINVAL; }
    return nlinsar_batch_enqueue(b->pb, b->L, (char*)b->ws + b->tables_off, stream);
}

}  // namespace dkx

using namespace dkx;

extern "C" {

int dkx_nlinsar_batch_create(dkx_nlinsar_batch* out)
{
    return dkx::guarded<int>([&]() -> int {
        if (!out) { set_error("nlinsar batch: null handle pointer"); return DKX_EINVAL; }
        *out = reinterpret_cast<dkx_nlinsar_batch>(new NlinsarBatch());
        return 0;
    });
}

void dkx_nlinsar_batch_destroy(dkx_nlinsar_batch h)
{
    dkx::guarded_void([&] {
        NlinsarBatch* b = reinterpret_cast<NlinsarBatch*>(h);
        if (!b) return;
        (void)nb_wait(b);
        (void)nb_drop_graph(b);
        if (b->done) (void)hipEventDestroy(b->done);
        if (b->ws) (void)hipFree(b->ws);
        delete b;
    });
}

int dkx_nlinsar_batch_set(dkx_nlinsar_batch h, const dkx_nlinsar_image* images, int n, int sw, int p,
                          int niter, int lmin)
{
    return dkx::guarded<int>([&]() -> int {
        NlinsarBatch* b = reinterpret_cast<NlinsarBatch*>(h);
        if (!b || !images || n < 1) { set_error("nlinsar batch: no images"); return DKX_EINVAL; }
        std::vector<int> H(n), W(n);
        for (int i = 0; i < n; i++) {
            const dkx_nlinsar_image& im = images[i];
            if (!im.ampl_master || !im.ampl_slave || !im.phase || !im.ref_filt || !im.phase_filt || !im.coh_filt) {
                set_error("nlinsar batch: null array in image " + std::to_string(i));
                return DKX_EINVAL;
            }
            H[i] = im.height;
            W[i] = im.width;
        }
        // everything that can fail without the device is checked before touching the batch
        // the tile shape: dkx_set_tile's, else the autotuner's record for the batch's pixel
        // count, else the built-in rule
        double px = 0;
        for (int i = 0; i < n; i++) px += (double)H[i] * W[i];
        int tile = dkx_get_tile(1);
        if (tile == NLINSAR_TILE_AUTO) tile = tuned_tile(1, 0, sw, &p, 1, nullptr, 0, px);
        if (tile < 0) tile = NLINSAR_TILE_AUTO;
        NlinsarProblem pb;
        int rc = nlinsar_problem(0, 0, sw, p, niter, lmin, tile, &pb);
        if (rc) return rc;
        pb.store = dkx_get_nlinsar_weight_store();
        std::vector<NlinsarPlan> plans;
        NlinsarBatchLaunch L;
        rc = nlinsar_batch_plan(pb, H.data(), W.data(), n, plans, &L);
        if (rc) return rc;
        size_t off = (L.bytes + 255) / 256 * 256;   // the tables first
        std::vector<size_t> offs(n);
        for (int i = 0; i < n; i++) {
            offs[i] = off;
            off += (plans[i].ws_bytes + 255) / 256 * 256;
        }

        // device side: wait for the batch's last run, grow the workspace, upload the tables
        rc = nb_wait(b);
        if (rc) return rc;
        rc = nb_drop_graph(b);
        if (rc) return rc;
        if (b->ws_cap < off) {
            if (b->ws) DKX_HIP_CHECK(hipFree(b->ws));
            b->ws = nullptr;
            b->ws_cap = 0;
            const hipError_t e = hipMalloc(&b->ws, off);
            if (e != hipSuccess) {
                b->ws = nullptr;
                (void)hipGetLastError();
                set_error("nlinsar batch: workspace allocation of " + std::to_string(off) + " bytes failed: " +
                          hipGetErrorString(e));
                return DKX_ERESOURCE;
            }
            b->ws_cap = off;
        }
        std::vector<NlinsarBatchItem> items(n);
        for (int i = 0; i < n; i++) {
            const dkx_nlinsar_image& im = images[i];
            items[i] = NlinsarBatchItem{im.ampl_master, im.ampl_slave, im.phase, im.ref_filt, im.phase_filt,
                                        im.coh_filt, im.height, im.width, (char*)b->ws + offs[i]};
        }
        std::vector<unsigned char> host;
        rc = nlinsar_batch_tables(pb, plans, items.data(), L, (char*)b->ws, host);
        if (rc) return rc;
        DKX_HIP_CHECK(hipMemcpy(b->ws, host.data(), host.size(), hipMemcpyHostToDevice));
        b->pb = pb;
        b->plans = std::move(plans);
        b->L = L;
        b->ws_off = std::move(offs);
        b->tables_off = 0;
        return 0;
    });
}

int dkx_nlinsar_batch_enqueue(dkx_nlinsar_batch h, void* stream)
{
    return dkx::guarded<int>([&]() -> int {
        NlinsarBatch* b = reinterpret_cast<NlinsarBatch*>(h);
        if (!b) { set_error("nlinsar batch: null handle"); return DKX_EINVAL; }
        int rc = nb_enqueue(b, (hipStream_t)stream);
        if (rc) return rc;
        return nb_mark(b, (hipStream_t)stream);
    });
}

int dkx_nlinsar_batch_graph_launch(dkx_nlinsar_batch h, void* stream_)
{
    return dkx::guarded<int>([&]() -> int {
        NlinsarBatch* b = reinterpret_cast<NlinsarBatch*>(h);
        if (!b) { set_error("nlinsar batch: null handle"); return DKX_EINVAL; }
        hipStream_t stream = (hipStream_t)stream_;
        if (!stream) { set_error("nlinsar batch: graph launches need a created stream"); return DKX_EINVAL; }
        if (!b->exec) {
            DKX_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            const int rc = nb_enqueue(b, stream);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) {
                set_error(std::string("nlinsar batch: graph capture failed: ") + hipGetErrorString(e));
                return DKX_EHIP;
            }
            b->graph = g;
            DKX_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        DKX_HIP_CHECK(hipGraphLaunch(b->exec, stream));
        return nb_mark(b, stream);
    });
}

int dkx_nlinsar_batch_smoothed(dkx_nlinsar_batch h, unsigned long long* counts)
{
    return dkx::guarded<int>([&]() -> int {
        NlinsarBatch* b = reinterpret_cast<NlinsarBatch*>(h);
        if (!b || !counts) { set_error("nlinsar batch: null argument"); return DKX_EINVAL; }
        int rc = nb_wait(b);
        if (rc) return rc;
        for (size_t i = 0; i < b->plans.size(); i++) {
            unsigned int c = 0;
            DKX_HIP_CHECK(hipMemcpy(&c, (char*)b->ws + b->ws_off[i] + b->plans[i].off_cnt + sizeof(unsigned int),
                                    sizeof(unsigned int), hipMemcpyDeviceToHost));
            counts[i] = c;
        }
        return 0;
    });
}

}  // extern "C"
