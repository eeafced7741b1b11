// Batched projection: many (volume, geometry, detector) projections of uneven size in one
// persistent kernel, captured once into a hipGraph and replayed. The per-view FP64 setup of every
// projection, the queue reset and the persistent kernel are the nodes of the graph.
#include "autotune.h"
#include "ctlproj.h"
#include "kernels.h"
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <new>
#include <vector>

using namespace ctlrt;

namespace {

struct BatchInst {
    const ctlproj_volume* vol = nullptr;
    std::vector<double> pmats;
    uint32_t V = 0, M = 0, C = 0, R = 0;
    ctlproj_settings s{};
    float* out = nullptr;
    EffectiveVolume ev;
    double* d_pmats = nullptr;
    ctlk::RayGeom* d_geom = nullptr;
    ~BatchInst()
    {
        if (d_pmats) (void)hipFree(d_pmats);
        if (d_geom) (void)hipFree(d_geom);
    }
};

} // namespace

struct ctlproj_batch {
    ctlproj_context* ctx = nullptr;
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<BatchInst>> insts;
    bool finalized = false;
    ctlk::Storage st = ctlk::Storage::F32;
    ctlk::Layout ly = ctlk::Layout::Interleaved;
    int interp = 1;
    ctlk::LaunchShape shape;
    bool staged = false; // every item can use the LDS-staged kernel
    uint32_t grid = 0;
    std::vector<ctlk::BatchItem> items;
    ctlk::BatchItem* d_items = nullptr;
    size_t items_cap = 0;
    ctlk::BatchQueue* d_q = nullptr;
    unsigned long long total_tiles = 0;
    unsigned long long* d_count = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool timed = false;
    float tuned_ms = 0.0f;
};

namespace {

ctlk::RayParams rayParams(const BatchInst& in)
{
    ctlk::RayParams rp;
    rp.channels = in.C;
    rp.rows = in.R;
    rp.modules = in.M;
    rp.views = in.V;
    rp.rpp_x = in.s.rays_per_pixel[0];
    rp.rpp_y = in.s.rays_per_pixel[1];
    rp.interpolate = in.s.interpolate ? 1 : 0;
    return rp;
}

// Item table for a launch shape; returns the total number of tiles.
ctlproj_status buildItems(ctlproj_batch* b, const ctlk::LaunchShape& sh)
{
    b->items.clear();
    unsigned long long begin = 0;
    for (const auto& in : b->insts) {
        ctlk::BatchItem it{};
        it.vol = in->ev.view;
        it.geom = in->d_geom;
        it.out = in->out;
        it.rp = rayParams(*in);
        it.kc = ctlk::makeCastConst(&in->ev.view, in->ev.geom, it.rp, sh);
        it.tile_begin = begin;
        begin += (unsigned long long)it.kc.tiles_per_vm * in->V * in->M;
        b->items.push_back(it);
    }
    b->total_tiles = begin;
    if (b->items.size() > b->items_cap) {
        if (b->d_items) (void)hipFree(b->d_items);
        b->d_items = nullptr;
        b->items_cap = 0;
        hipError_t e = hipMalloc(&b->d_items, b->items.size() * sizeof(ctlk::BatchItem));
        if (e != hipSuccess) return hipFail(e, "hipMalloc(batch items)");
        b->items_cap = b->items.size();
    }
    hipError_t e = hipMemcpyAsync(b->d_items, b->items.data(), b->items.size() * sizeof(ctlk::BatchItem),
                                  hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "batch item upload");
}

ctlproj_status enqueueSetups(ctlproj_batch* b)
{
    for (const auto& in : b->insts) {
        hipError_t e = ctlk::launchSetup(in->d_pmats, in->V, in->M, in->C, in->R, in->ev.geom,
                                         in->s.interpolate ? 1 : 0, in->d_geom, b->stream);
        if (e != hipSuccess) return hipFail(e, "setup kernel");
    }
    return CTLPROJ_OK;
}

// LDS box of the batch launch: 0 selects the unstaged kernel.
uint32_t ldsTexels(const ctlproj_batch* b, const ctlk::LaunchShape& sh)
{
    return b->staged && sh.chunk ? sh.lds_texels : 0u;
}

// One timed launch over every `stride`-th run of ctlk::kSampleRun consecutive tiles (autotuner;
// the whole runs that fit, or every tile with stride 1).
float timedSample(ctlproj_batch* b, const ctlk::LaunchShape& sh, uint32_t stride)
{
    if (buildItems(b, sh) != CTLPROJ_OK) return -1.0f;
    uint32_t grid = 0;
    if (ctlk::batchGrid(b->st, b->ly, b->interp, sh.unroll, sh.blocks_per_cu, ldsTexels(b, sh), &grid) != hipSuccess)
        return -1.0f;
    const unsigned long long span = (unsigned long long)stride * ctlk::kSampleRun;
    if (stride > 1 && b->total_tiles < span) stride = 1;
    const unsigned long long n =
        stride == 1 ? b->total_tiles : b->total_tiles / ((unsigned long long)stride * ctlk::kSampleRun) * ctlk::kSampleRun;
    hipError_t e = ctlk::launchQueueInit(b->d_q, n, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->e0, b->stream);
    if (e == hipSuccess)
        e = ctlk::launchBatch(b->d_items, uint32_t(b->items.size()), b->st, b->ly, b->interp, sh.unroll,
                              sh.blocks_per_cu, ldsTexels(b, sh), grid, b->d_q, stride, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b->e1);
    float ms = -1.0f;
    if (e == hipSuccess && hipEventElapsedTime(&ms, b->e0, b->e1) != hipSuccess) ms = -1.0f;
    return e == hipSuccess ? ms : -1.0f;
}

} // namespace

extern "C" {

ctlproj_status ctlproj_batch_create(ctlproj_context* ctx, ctlproj_batch** out)
{
    if (!ctx || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    auto* b = new (std::nothrow) ctlproj_batch;
    if (!b) return fail(CTLPROJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    b->ctx = ctx;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&b->e0)) != hipSuccess || (e = hipEventCreate(&b->e1)) != hipSuccess ||
        (e = hipMalloc(&b->d_q, sizeof(ctlk::BatchQueue))) != hipSuccess ||
        (e = hipMalloc(&b->d_count, sizeof(unsigned long long))) != hipSuccess) {
        ctlproj_batch_destroy(b);
        return hipFail(e, "batch resources");
    }
    *out = b;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_destroy(ctlproj_batch* b)
{
    if (!b) return CTLPROJ_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->insts.clear();
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_q) (void)hipFree(b->d_q);
    if (b->d_count) (void)hipFree(b->d_count);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_add(ctlproj_batch* b, const ctlproj_volume* vol, const double* pmats,
                                 uint32_t views, uint32_t modules, uint32_t channels, uint32_t rows,
                                 const ctlproj_settings* settings, float* device_out,
                                 size_t out_count, uint32_t* index)
{
    if (!b) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (b->finalized) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batch is finalized");
    if (!vol || !vol->d_data) return fail(CTLPROJ_ERR_NO_DATA, "no or contradictory data in volume object");
    if (vol->ctx != b->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume belongs to another context");
    if (largeVolume(vol)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batches hold volumes below 4 GiB only");
    if (views == 0 || modules == 0 || channels == 0 || rows == 0 || !pmats)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "every batch entry needs views and a detector");
    if (out_count != size_t(views) * modules * rows * channels || !device_out)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "device output must hold views*modules*rows*channels floats");
    ctlproj_settings s;
    if (settings) s = *settings;
    else ctlproj_settings_default(&s);
    ctlproj_status st = validateSettings(s);
    if (st != CTLPROJ_OK) return st;
    for (size_t i = 0, n = size_t(views) * modules * 12; i < n; ++i)
        if (!std::isfinite(pmats[i])) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "non-finite projection matrix entry");
    if (!b->insts.empty()) {
        const auto& f = *b->insts.front();
        if (vol->view.storage != f.vol->view.storage || vol->view.layout != f.vol->view.layout ||
            (s.interpolate != 0) != (f.s.interpolate != 0))
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT,
                        "a batch holds one storage type and one interpolation mode");
    }
    auto in = std::make_unique<BatchInst>();
    in->vol = vol;
    in->pmats.assign(pmats, pmats + size_t(views) * modules * 12);
    in->V = views;
    in->M = modules;
    in->C = channels;
    in->R = rows;
    in->s = s;
    in->out = device_out;
    if (index) *index = uint32_t(b->insts.size());
    b->insts.push_back(std::move(in));
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_finalize(ctlproj_batch* b, int autotune)
{
    if (!b) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (b->finalized) return CTLPROJ_OK;
    if (b->insts.empty()) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "empty batch");
    if (injected(b->ctx)) return CTLPROJ_ERR_INJECTED;
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    const auto& first = *b->insts.front();
    b->st = first.vol->view.storage;
    b->ly = first.vol->view.layout;
    b->interp = first.s.interpolate ? 1 : 0;
    for (auto& in : b->insts) {
        ctlproj_status st = effectiveVolume(in->vol, in->s, b->stream, in->ev);
        if (st != CTLPROJ_OK) return st;
        if (in->ev.view.large) // upsampled to 4 GiB or more
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batches hold volumes below 4 GiB only");
        const size_t n = in->pmats.size();
        if ((e = hipMalloc(&in->d_pmats, n * sizeof(double))) != hipSuccess ||
            (e = hipMalloc(&in->d_geom, size_t(in->V) * in->M * sizeof(ctlk::RayGeom))) != hipSuccess ||
            (e = hipMemcpyAsync(in->d_pmats, in->pmats.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                b->stream)) != hipSuccess)
            return hipFail(e, "batch geometry upload");
    }
    ctlproj_status st = enqueueSetups(b);
    if (st != CTLPROJ_OK) return st;
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return hipFail(e, "batch setup");

    // launch shape: tuned on a sample of this batch's own tiles, or the table / default
    const ctlk::Storage s0 = b->st;
    b->staged = !b->insts.empty();
    for (const auto& in : b->insts) b->staged = b->staged && stagedWorkload(in->s, s0, b->ly);
    const std::string key = ctltune::key(s0, b->ly, b->interp, first.C, first.R, b->staged) + "_batch";
    ctlk::LaunchShape sh = ctlk::defaultShape(b->staged);
    bool have = false;
    if (autotune) {
        if ((st = buildItems(b, sh)) != CTLPROJ_OK) return st;
        const uint32_t stride = uint32_t(std::max<unsigned long long>(1, b->total_tiles / 65536));
        ctltune::Choice best;
        have = ctltune::tune(key, [&](const ctlk::LaunchShape& c) { return timedSample(b, c, stride); }, &best,
                             b->staged);
        if (have) {
            sh = best.shape;
            b->tuned_ms = best.ms;
        }
    }
    if (!have) ctltune::lookup(key, &sh);
    b->shape = sh;
    if ((st = buildItems(b, sh)) != CTLPROJ_OK) return st;
    if ((e = ctlk::batchGrid(b->st, b->ly, b->interp, sh.unroll, sh.blocks_per_cu, ldsTexels(b, sh), &b->grid)) !=
        hipSuccess)
        return hipFail(e, "occupancy query");

    // capture: per-view setups, queue reset, persistent kernel (kernels only, no memset nodes)
    if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return hipFail(e, "hipStreamBeginCapture");
    st = enqueueSetups(b);
    hipError_t ec = hipSuccess;
    if (st == CTLPROJ_OK) ec = ctlk::launchQueueInit(b->d_q, b->total_tiles, b->stream);
    if (st == CTLPROJ_OK && ec == hipSuccess)
        ec = ctlk::launchBatch(b->d_items, uint32_t(b->items.size()), b->st, b->ly, b->interp, sh.unroll,
                               sh.blocks_per_cu, ldsTexels(b, sh), b->grid, b->d_q, 1, b->stream);
    e = hipStreamEndCapture(b->stream, &b->graph);
    if (st != CTLPROJ_OK) return st;
    if (ec != hipSuccess) return hipFail(ec, "batch capture");
    if (e != hipSuccess) return hipFail(e, "hipStreamEndCapture");
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
        return hipFail(e, "hipGraphInstantiate");
    b->finalized = true;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_launch(ctlproj_batch* b)
{
    if (!b) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (!b->finalized) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "batch is not finalized");
    if (injected(b->ctx)) return CTLPROJ_ERR_INJECTED;
    hipError_t e = hipEventRecord(b->e0, b->stream);
    if (e == hipSuccess) e = hipGraphLaunch(b->exec, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->e1, b->stream);
    if (e != hipSuccess) return hipFail(e, "hipGraphLaunch");
    b->timed = true;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_synchronize(ctlproj_batch* b)
{
    if (!b) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batch is NULL");
    hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "hipStreamSynchronize");
}

ctlproj_status ctlproj_batch_last_ms(ctlproj_batch* b, float* ms)
{
    if (!b || !ms) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->timed) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "the batch has not been launched");
    hipError_t e = hipEventSynchronize(b->e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, b->e0, b->e1);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "hipEventElapsedTime");
}

ctlproj_status ctlproj_batch_info(const ctlproj_batch* b, uint32_t* tile_w, uint32_t* tile_h,
                                  uint32_t* unroll, uint32_t* grid, uint64_t* tiles)
{
    if (!b) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "batch is NULL");
    if (!b->finalized) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "batch is not finalized");
    if (tile_w) *tile_w = b->shape.wave_w;
    if (tile_h) *tile_h = b->shape.wave_h;
    if (unroll) *unroll = b->shape.unroll;
    if (grid) *grid = b->grid;
    if (tiles) *tiles = b->total_tiles;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_shape(const ctlproj_batch* b, ctlproj_settings* settings)
{
    if (!b || !settings) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->finalized) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "batch is not finalized");
    settings->tile_w = b->shape.wave_w;
    settings->tile_h = b->shape.wave_h;
    settings->unroll = b->shape.unroll;
    settings->blocks_per_cu = b->shape.blocks_per_cu;
    const uint32_t box = ldsTexels(b, b->shape);
    settings->lds_chunk = box ? int32_t(b->shape.chunk) : -1;
    settings->lds_texels = box;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_batch_count_samples(ctlproj_batch* b, uint64_t* n_samples)
{
    if (!b || !n_samples) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->finalized) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "batch is not finalized");
    hipError_t e = hipMemsetAsync(b->d_count, 0, sizeof(unsigned long long), b->stream);
    for (const auto& in : b->insts)
        if (e == hipSuccess) e = ctlk::launchCount(in->ev.geom, in->d_geom, rayParams(*in), b->d_count, b->stream);
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, b->d_count, sizeof h, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hipFail(e, "count kernel");
    *n_samples = h;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_autotune_table(char* buf, size_t cap, size_t* needed)
{
    const std::string s = ctltune::exportTable();
    if (needed) *needed = s.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, s.size());
        std::copy(s.begin(), s.begin() + long(n), buf);
        buf[n] = '\0';
    }
    return CTLPROJ_OK;
}

} // extern "C"
