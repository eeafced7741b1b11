// Serving loop: resident volumes, admission by HBM footprint, continuous batching.
//
// Requests wait in arrival order until their device footprint (output + P-matrices + ray
// geometry) fits into the server's arena; admission uploads the P-matrices and runs the FP64
// per-view setup, then splits the request into chunks of views. A scheduler thread forms
// persistent-kernel launches from the ready chunks of all admitted requests (any volume, any
// detector size, one storage type and mode per launch) on three launch slots, each with its own
// stream: scan launches queue one behind the other on the device and leave part of it free, and
// single-view requests (DRRs) take any free slot, so that they run beside a scan launch. A completion thread retires launches, copies finished
// requests to the caller's host arrays, releases their arena blocks and wakes waiters.
#include "arena.h"
#include "ctlproj.h"
#include "kernels.h"
#include "runtime.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

using namespace ctlrt;
using Clock = std::chrono::steady_clock;

namespace {

// Launches in flight: at most kSlots - 1 of them carry scan chunks, so that a single-view request
// (DRR) always finds a slot.
constexpr uint32_t kSlots = 3;
// A launch with scan chunks uses the device's workgroup capacity less 1/kScanFree of it, so that the
// workgroups of a DRR launch behind it start at once; scan launches run one after another on the
// device (each waits on the event of the one before), so that the free part stays free.
constexpr uint32_t kScanFree = 8;
constexpr uint32_t kSlotItems = 4096;

using ctlsrv::Arena;
using ctlsrv::alignUp;
using ctlsrv::kAlign;

enum class State { Pending, Admitted, Copying, Done, Failed };

struct Request {
    uint64_t id = 0;
    const ctlproj_volume* vol = nullptr;
    std::vector<double> pmats;
    uint32_t V = 0, M = 0, C = 0, R = 0;
    ctlproj_settings s{};
    float* host_out = nullptr;
    Clock::time_point t_submit, t_done;
    State state = State::Pending;
    ctlproj_status status = CTLPROJ_OK;
    size_t off = 0, bytes = 0;
    float* d_out = nullptr;
    double* d_pmats = nullptr;
    ctlk::RayGeom* d_geom = nullptr;
    ctlk::VolumeGeom vg{};
    uint32_t chunks_total = 0, chunks_done = 0;
    uint32_t inflight = 0;      // chunks in launches not yet retired
    bool release_when_idle = false;
    hipEvent_t prep_done = nullptr, copy_done = nullptr;
    size_t outBytes() const { return size_t(V) * M * R * C * sizeof(float); }
};

struct Chunk {
    Request* r;
    uint32_t v0, nv;
};

struct Slot {
    hipStream_t stream = nullptr;
    ctlk::BatchItem* d_items = nullptr;
    ctlk::BatchItem* h_items = nullptr;
    ctlk::BatchQueue* d_q = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Chunk> chunks;
    bool busy = false;
    bool drr = false; // single-view requests only
};

using GroupKey = std::tuple<int, int, int, int>; // storage, layout, interpolate, LDS-staged

double percentile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * double(v.size() - 1);
    const size_t i = size_t(pos);
    const double f = pos - double(i);
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

} // namespace

struct ctlproj_server {
    ctlproj_context* ctx = nullptr;
    char* arena_base = nullptr;
    Arena arena;
    uint64_t max_launch_pixels = 0;
    hipStream_t prep = nullptr, copy = nullptr;
    Slot slots[kSlots];
    uint32_t next_slot = 0;
    std::map<GroupKey, uint32_t> grid_cache;

    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> pending;
    std::deque<Chunk> ready;
    std::vector<Request*> copying;
    uint64_t outstanding = 0;

    // statistics since the last reset
    uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0, views = 0, launches = 0;
    uint32_t cu_masked_slots = 0, max_scans_in_flight = 0;
    uint64_t drr_beside_scan = 0;
    std::vector<double> lat_single, lat_multi;
    Clock::time_point first_submit{}, last_done{};
    bool any_submit = false;

    std::thread scheduler, completer;
};

namespace {

void releaseLocked(ctlproj_server* S, Request* r)
{
    if (r->bytes) S->arena.release(r->off, r->bytes);
    r->bytes = 0;
    r->release_when_idle = false;
}

void finishLocked(ctlproj_server* S, Request* r, ctlproj_status st)
{
    r->t_done = Clock::now();
    r->status = st;
    r->state = st == CTLPROJ_OK ? State::Done : State::Failed;
    // a failed request's other chunks may still run in a launch: keep its block until they retire
    if (r->inflight == 0) releaseLocked(S, r);
    else r->release_when_idle = true;
    if (st != CTLPROJ_OK)
        for (auto it = S->ready.begin(); it != S->ready.end();)
            it = (it->r == r) ? S->ready.erase(it) : std::next(it);
    if (r->prep_done) (void)hipEventDestroy(r->prep_done);
    if (r->copy_done) (void)hipEventDestroy(r->copy_done);
    r->prep_done = r->copy_done = nullptr;
    r->pmats.clear();
    r->pmats.shrink_to_fit();
    if (st == CTLPROJ_OK) {
        ++S->completed;
        S->views += r->V;
        const double ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
        (r->V == 1 ? S->lat_single : S->lat_multi).push_back(ms);
        S->last_done = r->t_done;
    } else {
        ++S->failed;
    }
    --S->outstanding;
    S->cv_done.notify_all();
    S->cv_sched.notify_all();
}

// Uploads geometry and runs the per-view FP64 setup of an admitted request (prep stream).
ctlproj_status prepare(ctlproj_server* S, Request* r)
{
    const size_t out_b = alignUp(r->outBytes());
    const size_t pm_b = alignUp(r->pmats.size() * sizeof(double));
    r->d_out = reinterpret_cast<float*>(S->arena_base + r->off);
    r->d_pmats = reinterpret_cast<double*>(S->arena_base + r->off + out_b);
    r->d_geom = reinterpret_cast<ctlk::RayGeom*>(S->arena_base + r->off + out_b + pm_b);
    for (int k = 0; k < 3; ++k) {
        r->vg.n[k] = r->vol->n[k];
        r->vg.voxel_mm[k] = r->vol->voxel_mm[k];
        r->vg.offset_mm[k] = r->vol->offset_mm[k];
    }
    r->vg.increment_mm = std::min(std::min(r->vg.voxel_mm[0], r->vg.voxel_mm[1]), r->vg.voxel_mm[2]) *
                         r->s.ray_sampling;
    hipError_t e = hipEventCreateWithFlags(&r->prep_done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->copy_done, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipMemcpyAsync(r->d_pmats, r->pmats.data(), r->pmats.size() * sizeof(double),
                           hipMemcpyHostToDevice, S->prep);
    if (e == hipSuccess)
        e = ctlk::launchSetup(r->d_pmats, r->V, r->M, r->C, r->R, r->vg, r->s.interpolate ? 1 : 0,
                              r->d_geom, S->prep);
    if (e == hipSuccess) e = hipEventRecord(r->prep_done, S->prep);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "request setup");
}

GroupKey groupOf(const Request* r)
{
    return GroupKey(int(r->vol->view.storage), int(r->vol->view.layout), r->s.interpolate ? 1 : 0,
                    stagedWorkload(r->s, r->vol->view.storage, r->vol->view.layout) ? 1 : 0);
}

// Admission in arrival order; a request that does not fit lets later, smaller ones pass.
void admitLocked(ctlproj_server* S)
{
    for (auto it = S->pending.begin(); it != S->pending.end();) {
        Request* r = *it;
        const size_t need = alignUp(r->outBytes()) + alignUp(r->pmats.size() * sizeof(double)) +
                            alignUp(size_t(r->V) * r->M * sizeof(ctlk::RayGeom));
        size_t off = 0;
        if (!S->arena.alloc(need, &off)) {
            ++it;
            continue;
        }
        r->off = off;
        r->bytes = need;
        ctlproj_status st = prepare(S, r);
        it = S->pending.erase(it);
        if (st != CTLPROJ_OK) {
            finishLocked(S, r, st);
            continue;
        }
        r->state = State::Admitted;
        // chunks of views, each at most max_launch_pixels / 4 pixels, so long scans interleave
        const uint64_t per_view = uint64_t(r->M) * r->R * r->C;
        const uint32_t cv = uint32_t(std::max<uint64_t>(1, (S->max_launch_pixels / 4) / std::max<uint64_t>(1, per_view)));
        for (uint32_t v0 = 0; v0 < r->V; v0 += cv) {
            S->ready.push_back(Chunk{ r, v0, std::min(cv, r->V - v0) });
            ++r->chunks_total;
        }
    }
}

// Retires the chunks of a finished (or failed) launch.
void retireLocked(ctlproj_server* S, std::vector<Chunk>& chunks)
{
    for (auto& c : chunks) {
        --c.r->inflight;
        if (c.r->release_when_idle && c.r->inflight == 0) releaseLocked(S, c.r);
    }
    chunks.clear();
}

void failChunksLocked(ctlproj_server* S, std::vector<Chunk>& chunks, ctlproj_status st)
{
    for (auto& c : chunks)
        if (c.r->state == State::Admitted) finishLocked(S, c.r, st);
    retireLocked(S, chunks);
}

// Forms one launch from the ready chunks of the first ready chunk's group (with drr_only, of the
// first single-view request's group, taking single-view requests only).
void launchLocked(ctlproj_server* S, Slot& slot, bool drr_only)
{
    auto first = S->ready.begin();
    if (drr_only)
        while (first != S->ready.end() && first->r->V != 1) ++first;
    if (first == S->ready.end()) return;
    const GroupKey key = groupOf(first->r);
    // groups of trilinear requests with one ray per pixel run the LDS-staged kernel, the others
    // the unstaged kernel
    const bool staged = std::get<3>(key) != 0;
    ctlk::LaunchShape shape = ctlk::defaultShape(staged);
    // unstaged launches keep four workgroups per CU (the kernel's registers would allow eight): a
    // DRR launch then finds room beside a running scan launch. 8 x 256^3 FP16 phantoms, 80% 512^2
    // DRRs at 200 req/s: single-view p50 2.5 ms against 4.8 ms uncapped, views/s within 1%
    if (!staged) shape.blocks_per_cu = 4;
    const uint32_t lds_texels = staged ? shape.lds_texels : 0u;
    uint64_t pixels = 0;
    unsigned long long tiles = 0;
    uint32_t n = 0;
    slot.chunks.clear();
    // single-view requests (DRRs) form a launch of their own, so that they finish in well under a
    // millisecond instead of with the chunks of long scans; without any, the chunks go in arrival
    // order
    for (int pass = 0; pass < (drr_only ? 1 : 2) && !(pass == 1 && n > 0); ++pass)
        for (auto it = S->ready.begin(); it != S->ready.end() && n < kSlotItems;) {
            Request* r = it->r;
            const uint64_t px = uint64_t(it->nv) * r->M * r->R * r->C;
            if ((pass == 0 && r->V != 1) || groupOf(r) != key ||
                (n > 0 && pixels + px > S->max_launch_pixels)) {
                ++it;
                continue;
            }
            ctlk::BatchItem& b = slot.h_items[n];
            b.vol = r->vol->view;
            b.geom = r->d_geom + size_t(it->v0) * r->M;
            b.out = r->d_out + size_t(it->v0) * r->M * r->R * r->C;
            b.rp.channels = r->C;
            b.rp.rows = r->R;
            b.rp.modules = r->M;
            b.rp.views = it->nv;
            b.rp.rpp_x = r->s.rays_per_pixel[0];
            b.rp.rpp_y = r->s.rays_per_pixel[1];
            b.rp.interpolate = r->s.interpolate ? 1 : 0;
            b.kc = ctlk::makeCastConst(&b.vol, r->vg, b.rp, shape);
            b.tile_begin = tiles;
            tiles += (unsigned long long)b.kc.tiles_per_vm * it->nv * r->M;
            pixels += px;
            slot.chunks.push_back(*it);
            ++r->inflight;
            ++n;
            it = S->ready.erase(it);
        }
    uint32_t& grid = S->grid_cache[key];
    hipError_t e = hipSuccess;
    if (grid == 0)
        e = ctlk::batchGrid(ctlk::Storage(std::get<0>(key)), ctlk::Layout(std::get<1>(key)), std::get<2>(key),
                            shape.unroll, shape.blocks_per_cu, lds_texels, &grid);
    if (e == hipSuccess)
        e = hipMemcpyAsync(slot.d_items, slot.h_items, n * sizeof(ctlk::BatchItem), hipMemcpyHostToDevice,
                           slot.stream);
    Request* last = nullptr;
    for (const auto& c : slot.chunks)
        if (e == hipSuccess && c.r != last) {
            e = hipStreamWaitEvent(slot.stream, c.r->prep_done, 0);
            last = c.r;
        }
    bool drr = true;
    for (const auto& c : slot.chunks) drr = drr && c.r->V == 1;
    const uint32_t launch_grid = drr ? grid : std::max<uint32_t>(1, grid - grid / kScanFree);
    if (!drr)
        for (const Slot& other : S->slots)
            if (e == hipSuccess && &other != &slot && other.busy && !other.drr)
                e = hipStreamWaitEvent(slot.stream, other.done, 0);
    if (e == hipSuccess) e = ctlk::launchQueueInit(slot.d_q, tiles, slot.stream);
    if (e == hipSuccess)
        e = ctlk::launchBatch(slot.d_items, n, ctlk::Storage(std::get<0>(key)), ctlk::Layout(std::get<1>(key)),
                              std::get<2>(key), shape.unroll, shape.blocks_per_cu, lds_texels, launch_grid, slot.d_q,
                              1, slot.stream);
    if (e == hipSuccess) e = hipEventRecord(slot.done, slot.stream);
    if (e != hipSuccess) {
        failChunksLocked(S, slot.chunks, hipFail(e, "serving launch"));
        return;
    }
    slot.busy = true;
    slot.drr = drr;
    ++S->launches;
    if (!drr) {
        uint32_t scans = 0;
        for (const Slot& sl : S->slots) scans += (sl.busy && !sl.drr) ? 1u : 0u;
        S->max_scans_in_flight = std::max(S->max_scans_in_flight, scans);
    }
}

void schedulerMain(ctlproj_server* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (S->stop && S->outstanding == 0) break;
        admitLocked(S);
        bool launched = false;
        if (!S->ready.empty()) {
            bool drr_ready = false;
            for (const Chunk& c : S->ready) drr_ready = drr_ready || c.r->V == 1;
            uint32_t scans = 0;
            for (const Slot& sl : S->slots) scans += (sl.busy && !sl.drr) ? 1u : 0u;
            // DRRs take any free slot; scan launches at most kSlots - 1
            if (drr_ready || scans + 1 < kSlots)
                for (uint32_t k = 0; k < kSlots; ++k) {
                    Slot& slot = S->slots[(S->next_slot + k) % kSlots];
                    if (slot.busy) continue;
                    launchLocked(S, slot, drr_ready);
                    S->next_slot = (S->next_slot + k + 1) % kSlots;
                    launched = slot.busy;
                    break;
                }
        }
        if (launched) S->cv_done.notify_all(); // wake the completion thread
        else S->cv_sched.wait_for(lk, std::chrono::microseconds(500));
    }
}

void completerMain(ctlproj_server* S)
{
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    std::vector<Request*> to_copy;
    for (;;) {
        bool progress = false;
        for (Slot& slot : S->slots) {
            if (!slot.busy) continue;
            const hipError_t q = hipEventQuery(slot.done);
            if (q == hipErrorNotReady) continue;
            progress = true;
            slot.busy = false;
            if (q == hipSuccess && slot.drr)
                for (const Slot& other : S->slots)
                    if (other.busy && !other.drr && hipEventQuery(other.done) == hipErrorNotReady) {
                        ++S->drr_beside_scan;
                        break;
                    }
            if (q != hipSuccess) {
                failChunksLocked(S, slot.chunks, hipFail(q, "serving kernel"));
                continue;
            }
            for (const auto& c : slot.chunks) {
                Request* r = c.r;
                if (r->state == State::Admitted && ++r->chunks_done == r->chunks_total) {
                    r->state = State::Copying;
                    to_copy.push_back(r);
                }
            }
            retireLocked(S, slot.chunks);
            S->cv_sched.notify_all();
        }
        if (!to_copy.empty()) {
            // a copy into pageable memory can block: issue it without holding the lock
            lk.unlock();
            std::vector<std::pair<Request*, hipError_t>> issued;
            for (Request* r : to_copy) {
                hipError_t e = hipMemcpyAsync(r->host_out, r->d_out, r->outBytes(), hipMemcpyDeviceToHost, S->copy);
                if (e == hipSuccess) e = hipEventRecord(r->copy_done, S->copy);
                issued.emplace_back(r, e);
            }
            lk.lock();
            for (auto& ie : issued) {
                if (ie.second != hipSuccess) finishLocked(S, ie.first, hipFail(ie.second, "result copy"));
                else S->copying.push_back(ie.first);
            }
            to_copy.clear();
        }
        for (auto it = S->copying.begin(); it != S->copying.end();) {
            const hipError_t q = hipEventQuery((*it)->copy_done);
            if (q == hipErrorNotReady) {
                ++it;
                continue;
            }
            progress = true;
            finishLocked(S, *it, q == hipSuccess ? CTLPROJ_OK : hipFail(q, "result copy"));
            it = S->copying.erase(it);
        }
        if (S->stop && S->outstanding == 0) break;
        if (!progress) S->cv_done.wait_for(lk, std::chrono::microseconds(50));
    }
}

} // namespace

extern "C" {

ctlproj_status ctlproj_server_create(ctlproj_context* ctx, size_t hbm_budget, uint64_t max_launch_pixels,
                                     ctlproj_server** out)
{
    if (!ctx || !out || hbm_budget < (size_t(1) << 20))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad arguments (budget must be >= 1 MiB)");
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    auto* S = new (std::nothrow) ctlproj_server;
    if (!S) return fail(CTLPROJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    S->ctx = ctx;
    S->max_launch_pixels = max_launch_pixels ? max_launch_pixels : (uint64_t(16) << 20);
    hbm_budget = hbm_budget / kAlign * kAlign;
    if ((e = hipMalloc(reinterpret_cast<void**>(&S->arena_base), hbm_budget)) != hipSuccess) {
        delete S;
        return hipFail(e, "hipMalloc(server arena)");
    }
    S->arena.init(hbm_budget);
    if ((e = hipStreamCreateWithFlags(&S->prep, hipStreamNonBlocking)) == hipSuccess)
        e = hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking);
    // The slot streams carry a CU mask of the whole device: HIP gives such a stream a hardware
    // queue of its own, whereas plain streams share a pool of GPU_MAX_HW_QUEUES queues (4 by
    // default), and a DRR launch on a stream that shares a scan launch's queue waits behind it.
    // (Such streams also order themselves against the null stream, which the server never uses.)
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, S->ctx->device);
    const std::vector<uint32_t> all_cus((uint32_t(cus) + 31) / 32, ~0u);
    for (Slot& sl : S->slots) {
        if (e == hipSuccess) {
            if (hipExtStreamCreateWithCUMask(&sl.stream, uint32_t(all_cus.size()), all_cus.data()) == hipSuccess) {
                ++S->cu_masked_slots;
            } else {
                (void)hipGetLastError();
                sl.stream = nullptr;
                e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
            }
        }
        if (e == hipSuccess) e = hipMalloc(&sl.d_items, kSlotItems * sizeof(ctlk::BatchItem));
        if (e == hipSuccess) e = hipHostMalloc(&sl.h_items, kSlotItems * sizeof(ctlk::BatchItem), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&sl.d_q, sizeof(ctlk::BatchQueue));
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        S->stop = true;
        ctlproj_server_destroy(S);
        return hipFail(e, "server resources");
    }
    S->scheduler = std::thread(schedulerMain, S);
    S->completer = std::thread(completerMain, S);
    *out = S;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_server_destroy(ctlproj_server* S)
{
    if (!S) return CTLPROJ_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_sched.notify_all();
        S->cv_done.notify_all();
    }
    if (S->scheduler.joinable()) S->scheduler.join();
    if (S->completer.joinable()) S->completer.join();
    (void)hipSetDevice(S->ctx->device);
    for (Slot& sl : S->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.d_items) (void)hipFree(sl.d_items);
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        if (sl.d_q) (void)hipFree(sl.d_q);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (S->prep) (void)hipStreamSynchronize(S->prep);
    if (S->copy) (void)hipStreamSynchronize(S->copy);
    if (S->prep) (void)hipStreamDestroy(S->prep);
    if (S->copy) (void)hipStreamDestroy(S->copy);
    if (S->arena_base) (void)hipFree(S->arena_base);
    delete S;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_server_submit(ctlproj_server* S, const ctlproj_volume* vol, const double* pmats,
                                     uint32_t views, uint32_t modules, uint32_t channels, uint32_t rows,
                                     const ctlproj_settings* settings, float* host_out, size_t out_count,
                                     uint64_t* request_id)
{
    if (!S) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "server is NULL");
    if (!vol || !vol->d_data) return fail(CTLPROJ_ERR_NO_DATA, "no or contradictory data in volume object");
    if (vol->ctx != S->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume belongs to another context");
    if (largeVolume(vol)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "the server projects volumes below 4 GiB only");
    if (views == 0 || modules == 0 || channels == 0 || rows == 0 || !pmats || !host_out)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "a request needs views, a detector and an output");
    if (out_count != size_t(views) * modules * rows * channels)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output must hold views*modules*rows*channels floats");
    ctlproj_settings s;
    if (settings) s = *settings;
    else ctlproj_settings_default(&s);
    ctlproj_status st = validateSettings(s);
    if (st != CTLPROJ_OK) return st;
    if (s.interpolate && s.volume_upsampling != 1)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "the server projects resident volumes as they are (upsampling 1)");
    for (size_t i = 0, n = size_t(views) * modules * 12; i < n; ++i)
        if (!std::isfinite(pmats[i])) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "non-finite projection matrix entry");
    auto r = std::make_unique<Request>();
    r->vol = vol;
    r->pmats.assign(pmats, pmats + size_t(views) * modules * 12);
    r->V = views;
    r->M = modules;
    r->C = channels;
    r->R = rows;
    r->s = s;
    r->host_out = host_out;
    const size_t need = alignUp(r->outBytes()) + alignUp(r->pmats.size() * sizeof(double)) +
                        alignUp(size_t(views) * modules * sizeof(ctlk::RayGeom));
    std::lock_guard<std::mutex> lk(S->mu);
    if (need > S->arena.size()) {
        ++S->rejected;
        return fail(CTLPROJ_ERR_OVER_BUDGET, "request needs " + std::to_string(need) +
                                                 " bytes of HBM, the server's budget is " +
                                                 std::to_string(S->arena.size()));
    }
    if (S->stop) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "server is stopping");
    r->id = S->next_id++;
    r->t_submit = Clock::now();
    if (!S->any_submit) {
        S->first_submit = r->t_submit;
        S->any_submit = true;
    }
    ++S->submitted;
    ++S->outstanding;
    if (request_id) *request_id = r->id;
    S->pending.push_back(r.get());
    S->reqs.emplace(r->id, std::move(r));
    S->cv_sched.notify_all();
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_server_wait(ctlproj_server* S, uint64_t id, double timeout_s, double* latency_ms)
{
    if (!S) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "unknown request id");
    Request* r = it->second.get();
    auto done = [&] { return r->state == State::Done || r->state == State::Failed; };
    if (timeout_s < 0) S->cv_done.wait(lk, done);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done))
        return fail(CTLPROJ_ERR_TIMEOUT, "request still running");
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->status != CTLPROJ_OK) return fail(r->status, "request failed on the device");
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_server_drain(ctlproj_server* S)
{
    if (!S) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->outstanding == 0; });
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_server_get_stats(ctlproj_server* S, ctlproj_server_stats* out)
{
    if (!S || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    ctlproj_server_stats st{};
    st.submitted = S->submitted;
    st.completed = S->completed;
    st.failed = S->failed;
    st.rejected = S->rejected;
    st.views = S->views;
    st.launches = S->launches;
    std::vector<double> all(S->lat_single);
    all.insert(all.end(), S->lat_multi.begin(), S->lat_multi.end());
    st.p50_ms = percentile(all, 0.50);
    st.p99_ms = percentile(all, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double v : all) {
        sum += v;
        mx = std::max(mx, v);
    }
    st.mean_ms = all.empty() ? 0.0 : sum / double(all.size());
    st.max_ms = mx;
    st.p50_single_ms = percentile(S->lat_single, 0.50);
    st.p99_single_ms = percentile(S->lat_single, 0.99);
    st.p50_multi_ms = percentile(S->lat_multi, 0.50);
    st.p99_multi_ms = percentile(S->lat_multi, 0.99);
    st.span_s = (S->any_submit && S->completed) ? std::chrono::duration<double>(S->last_done - S->first_submit).count() : 0.0;
    st.slots = kSlots;
    st.cu_masked_slots = S->cu_masked_slots;
    st.max_scan_launches_in_flight = S->max_scans_in_flight;
    st.drr_launches_beside_scan = S->drr_beside_scan;
    *out = st;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_server_reset_stats(ctlproj_server* S)
{
    if (!S) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "server is NULL");
    std::lock_guard<std::mutex> lk(S->mu);
    S->submitted = S->completed = S->failed = S->rejected = S->views = S->launches = 0;
    S->max_scans_in_flight = 0;
    S->drr_beside_scan = 0;
    S->lat_single.clear();
    S->lat_multi.clear();
    S->any_submit = false;
    // forget finished requests
    for (auto it = S->reqs.begin(); it != S->reqs.end();) {
        if (it->second->state == State::Done || it->second->state == State::Failed) it = S->reqs.erase(it);
        else ++it;
    }
    return CTLPROJ_OK;
}

} // extern "C"
