// C ABI of the forward projector: environment, resident volumes, projector, geometry, phantoms.
//
// Replaces the OpenCL environment of CTL (OpenCLConfig, openclconfig.cpp:15-171; ClFileLoader;
// PinnedBuf*) and RayCasterProjector::project (raycasterprojector.cpp:139-375). Code objects are
// compiled ahead of time for gfx950, so there is no runtime kernel build.
#include "autotune.h"
#include "ctlproj.h"
#include "geometry.h"
#include "kernels.h"
#include "phantom.h"
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <limits>
#include <new>

namespace ctlrt {

ctlproj_status hipFail(hipError_t e, const char* what)
{
    fail(CTLPROJ_ERR_DEVICE, std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CTLPROJ_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return CTLPROJ_ERR_NO_DEVICE;
    return CTLPROJ_ERR_DEVICE;
}

bool injected(ctlproj_context* ctx)
{
    int cur = ctx->inject.load();
    while (cur > 0)
        if (ctx->inject.compare_exchange_weak(cur, cur - 1)) {
            fail(CTLPROJ_ERR_INJECTED, "injected failure");
            return true;
        }
    return false;
}

EffectiveVolume::~EffectiveVolume()
{
    if (tmp) (void)hipFree(tmp);
}

ctlproj_status validateSettings(const ctlproj_settings& s)
{
    if (s.rays_per_pixel[0] < 1 || s.rays_per_pixel[1] < 1 || s.rays_per_pixel[0] > 64 ||
        s.rays_per_pixel[1] > 64)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "rays per pixel must be in [1, 64]");
    // A step far below a voxel makes every ray loop (box diagonal / step) times; refuse steps
    // below 1/1000 voxel instead of launching a kernel that would run for minutes.
    if (!std::isfinite(s.ray_sampling) || s.ray_sampling < 1e-3f)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "ray sampling must be finite and >= 0.001");
    if (s.volume_upsampling > 16)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume upsampling factor must be <= 16");
    if ((s.tile_w || s.tile_h) && s.tile_w * s.tile_h != 64)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "tile_w * tile_h must be 64 (one wavefront)");
    if (s.unroll > 4) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "unroll must be 0 to 4");
    if (s.blocks_per_cu > 8)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "blocks_per_cu must be 0..8");
    if (s.lds_chunk != 0 && s.lds_chunk != -1 && (s.lds_chunk < 16 || s.lds_chunk > 1024))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "lds_chunk must be 0, -1 or 16..1024");
    if (s.lds_texels != 0 && (s.lds_texels < 1024 || s.lds_texels > ctlk::kMaxLdsTexels))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "lds_texels must be 0 or 1024.." + std::to_string(ctlk::kMaxLdsTexels));
    return CTLPROJ_OK;
}

bool stagedWorkload(const ctlproj_settings& s, ctlk::Storage st, ctlk::Layout ly)
{
    // FP16/BF16 sampling is bound by its arithmetic, not by the gathers, and staging only adds
    // work: it stages on request (lds_chunk > 0), FP32 by default
    return ly == ctlk::Layout::Interleaved && s.interpolate && s.rays_per_pixel[0] == 1 &&
           s.rays_per_pixel[1] == 1 && (st == ctlk::Storage::F32 ? s.lds_chunk >= 0 : s.lds_chunk > 0);
}

ctlk::LaunchShape shapeFor(const ctlproj_settings& s, ctlk::Storage st, ctlk::Layout ly,
                           uint32_t channels, uint32_t rows)
{
    // explicit settings win; otherwise the autotuner's choice for this workload class, if any,
    // else the default shape of the kernel that applies
    const bool staged = stagedWorkload(s, st, ly);
    ctlk::LaunchShape sh = ctlk::defaultShape(staged);
    const bool tuned = ctltune::lookup(ctltune::key(st, ly, s.interpolate ? 1 : 0, channels, rows, staged), &sh);
    if (s.tile_w && s.tile_h) {
        sh.wave_w = s.tile_w;
        sh.wave_h = s.tile_h;
    }
    // narrow modules stack the wavefronts (the autotuner's candidates include stacked shapes)
    if (!tuned) sh = ctlk::narrowModuleShape(sh, channels, staged);
    if (s.unroll) sh.unroll = s.unroll;
    if (s.blocks_per_cu) sh.blocks_per_cu = s.blocks_per_cu;
    if (s.lds_chunk) sh.chunk = s.lds_chunk < 0 ? 0u : uint32_t(s.lds_chunk);
    if (s.lds_texels) sh.lds_texels = s.lds_texels & ~127u;
    return sh;
}

namespace {
// Padded arrays below this size are one buffer of 32-bit byte offsets; larger ones are read with
// 64-bit addresses (VolumeView::large).
constexpr uint64_t kLargeBytes = (1ull << 32) - 16;

// Limits of the gather addressing: 24-bit multiplies of texel rows and slices, 32-bit byte
// offsets below 4 GiB, and above it 32-bit texel row indices in the interleaved layout.
ctlproj_status checkPaddedSize(const uint32_t n[3], uint32_t f, ctlk::Storage st, ctlk::Layout ly,
                               size_t* bytes)
{
    const uint64_t px = uint64_t(n[0]) * f + 2, py = uint64_t(n[1]) * f + 2, pz = uint64_t(n[2]) * f + 2;
    const uint64_t eb = ctlk::texelBytes(st, ly);
    if (px * py >= (1ull << 24) || pz >= (1ull << 24))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume slice too large ((X+2)*(Y+2) must be < 2^24 voxels)");
    const uint64_t b = px * py * pz * eb;
    if (b >= kLargeBytes) {
        if (ly != ctlk::Layout::Interleaved)
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volumes of 4 GiB or more need the interleaved layout");
        if (py * pz >= (1ull << 32))
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume has too many texel rows ((Y+2)*(Z+2) must be < 2^32)");
    }
    *bytes = size_t(b);
    return CTLPROJ_OK;
}

ctlk::VolumeView makeView(void* d, const uint32_t n[3], ctlk::Storage st, ctlk::Layout ly, size_t bytes)
{
    ctlk::VolumeView v{};
    v.data = d;
    for (int k = 0; k < 3; ++k) v.n[k] = n[k];
    v.pitch_y = n[0] + 2;
    v.pitch_z = (n[0] + 2) * (n[1] + 2);
    v.storage = st;
    v.layout = ly;
    v.row_bytes = uint32_t((n[0] + 2) * ctlk::texelBytes(st, ly));
    v.rows = (n[1] + 2) * (n[2] + 2);
    // CTLPROJ_VOLUME_WIDE=1 reads every interleaved volume as volumes of 4 GiB or more are read
    // (the same kernels, at sizes that small tests reach)
    const char* wide = std::getenv("CTLPROJ_VOLUME_WIDE");
    const bool force = wide && std::strcmp(wide, "1") == 0 && ly == ctlk::Layout::Interleaved;
    v.large = (bytes >= kLargeBytes || force) ? 1u : 0u;
    // a large volume's buffer covers its first 4 GiB: the probes' streams and nothing else
    v.bytes = uint32_t(std::min<uint64_t>(bytes, kLargeBytes));
    return v;
}
} // namespace

bool largeVolume(const ctlproj_volume* vol) { return vol && vol->view.large; }

ctlproj_status effectiveVolume(const ctlproj_volume* vol, const ctlproj_settings& s,
                               hipStream_t stream, EffectiveVolume& ev)
{
    // non-interpolating mode forces factor 1 (raycasterprojector.cpp:148-151)
    const uint32_t f = s.interpolate ? s.volume_upsampling : 1u;
    ctlk::VolumeGeom& g = ev.geom;
    for (int k = 0; k < 3; ++k) g.offset_mm[k] = vol->offset_mm[k];
    if (f == 1) { // VolumeSpecs::upSampleVolume case 1 (raycasterprojector.cpp:611-624)
        ev.view = vol->view;
        for (int k = 0; k < 3; ++k) {
            g.n[k] = vol->n[k];
            g.voxel_mm[k] = vol->voxel_mm[k];
        }
    } else if (f == 0) { // case 0: one voxel holding the mean (:592-610)
        const uint32_t one[3] = { 1, 1, 1 };
        size_t bytes = 0;
        ctlproj_status st = checkPaddedSize(one, 1, vol->view.storage, vol->view.layout, &bytes);
        if (st != CTLPROJ_OK) return st;
        const size_t sum_at = (bytes + 7) / 8 * 8;
        hipError_t e = hipMalloc(&ev.tmp, sum_at + sizeof(double));
        if (e != hipSuccess) return hipFail(e, "hipMalloc(mean volume)");
        double* d_sum = reinterpret_cast<double*>(static_cast<char*>(ev.tmp) + sum_at);
        if ((e = hipMemsetAsync(d_sum, 0, sizeof(double), stream)) != hipSuccess)
            return hipFail(e, "hipMemsetAsync");
        if ((e = ctlk::launchSum(vol->view, d_sum, stream)) != hipSuccess) return hipFail(e, "sum kernel");
        const double count = double(vol->n[0]) * vol->n[1] * vol->n[2];
        if ((e = ctlk::launchFillMean(d_sum, count, vol->view.storage, vol->view.layout, ev.tmp,
                                      stream)) != hipSuccess)
            return hipFail(e, "fill kernel");
        ev.view = makeView(ev.tmp, one, vol->view.storage, vol->view.layout, bytes);
        for (int k = 0; k < 3; ++k) {
            g.n[k] = 1;
            g.voxel_mm[k] = vol->voxel_mm[k] * float(vol->n[k]);
        }
    } else { // case f > 1: replicate each voxel f^3 times (:625-659), on the device
        size_t bytes = 0;
        ctlproj_status st = checkPaddedSize(vol->n, f, vol->view.storage, vol->view.layout, &bytes);
        if (st != CTLPROJ_OK) return st;
        hipError_t e = hipMalloc(&ev.tmp, bytes);
        if (e != hipSuccess) return hipFail(e, "hipMalloc(upsampled volume)");
        if ((e = ctlk::launchUpsample(vol->view, f, ev.tmp, stream)) != hipSuccess)
            return hipFail(e, "upsample kernel");
        uint32_t nn[3];
        for (int k = 0; k < 3; ++k) {
            nn[k] = vol->n[k] * f;
            g.n[k] = nn[k];
            g.voxel_mm[k] = vol->voxel_mm[k] / float(f);
        }
        ev.view = makeView(ev.tmp, nn, vol->view.storage, vol->view.layout, bytes);
    }
    // increment_mm = min(voxel size) * raySampling (raycasterprojector.cpp:178-179)
    const float smallest = std::min(std::min(g.voxel_mm[0], g.voxel_mm[1]), g.voxel_mm[2]);
    g.increment_mm = smallest * s.ray_sampling;
    return CTLPROJ_OK;
}

// Shared creation path of resident volumes from a device float array.
ctlproj_status createVolume(ctlproj_context* ctx, const float* d_src, const uint32_t dims[3],
                            const float voxel_mm[3], const float offset_mm[3], ctlk::Storage st,
                            ctlproj_volume** out)
{
    // Interleaved texels by default; CTLPROJ_VOLUME_LAYOUT=plain selects one voxel per texel.
    const char* env = std::getenv("CTLPROJ_VOLUME_LAYOUT");
    const ctlk::Layout ly = (env && std::strcmp(env, "plain") == 0) ? ctlk::Layout::Plain
                                                                    : ctlk::Layout::Interleaved;
    size_t bytes = 0;
    ctlproj_status s = checkPaddedSize(dims, 1, st, ly, &bytes);
    if (s != CTLPROJ_OK) return s;
    ctlproj_volume* v = new (std::nothrow) ctlproj_volume;
    if (!v) return fail(CTLPROJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    v->ctx = ctx;
    hipError_t e = hipMalloc(&v->d_data, bytes);
    if (e != hipSuccess) {
        delete v;
        return hipFail(e, "hipMalloc(volume)");
    }
    v->bytes = bytes;
    if ((e = ctlk::launchPack(d_src, dims, 1, st, ly, v->d_data, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
        (void)hipFree(v->d_data);
        delete v;
        return hipFail(e, "volume pack kernel");
    }
    for (int k = 0; k < 3; ++k) {
        v->n[k] = dims[k];
        v->voxel_mm[k] = voxel_mm[k];
        v->offset_mm[k] = offset_mm[k];
    }
    v->view = makeView(v->d_data, dims, st, ly, bytes);
    // The occupancy table lets the ray casters skip chunks of samples that read only zero voxels
    // (their sums are unchanged). CTLPROJ_EMPTY_SKIP=0 leaves it out: every sample is then taken.
    const char* skip = std::getenv("CTLPROJ_EMPTY_SKIP");
    if (!(skip && std::strcmp(skip, "0") == 0)) {
        if ((e = hipMalloc(reinterpret_cast<void**>(&v->d_occ), ctlk::occBytes(dims))) != hipSuccess ||
            (e = ctlk::launchOccupancy(v->view, v->d_occ, ctx->stream)) != hipSuccess ||
            (e = hipStreamSynchronize(ctx->stream)) != hipSuccess) {
            if (v->d_occ) (void)hipFree(v->d_occ);
            (void)hipFree(v->d_data);
            delete v;
            return hipFail(e, "volume occupancy table");
        }
        v->view.occ = v->d_occ;
        for (int k = 0; k < 3; ++k) v->view.occ_n[k] = ctlk::occBricks(dims[k]);
    }
    *out = v;
    return CTLPROJ_OK;
}

} // namespace ctlrt

using namespace ctlrt;

namespace {

bool validDims(const uint32_t* d) { return d && d[0] && d[1] && d[2]; }

bool validVoxel(const float* v)
{
    for (int k = 0; k < 3; ++k)
        if (!(std::isfinite(v[k]) && v[k] > 0.0f)) return false;
    return true;
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }

ctlproj_status setDevice(ctlproj_context* ctx)
{
    hipError_t e = hipSetDevice(ctx->device);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "hipSetDevice");
}

template <class T> ctlproj_status ensureDevice(T*& p, size_t& cap, size_t count)
{
    if (count <= cap && p) return CTLPROJ_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) return hipFail(e, "hipMalloc");
    cap = count;
    return CTLPROJ_OK;
}

ctlk::RayParams rayParams(const ctlproj_projector* p, uint32_t views)
{
    ctlk::RayParams rp;
    rp.channels = p->C;
    rp.rows = p->R;
    rp.modules = p->M;
    rp.views = views;
    rp.rpp_x = p->s.rays_per_pixel[0];
    rp.rpp_y = p->s.rays_per_pixel[1];
    rp.interpolate = p->s.interpolate ? 1 : 0;
    return rp;
}

// Upload P-matrices and run the FP64 per-view setup for the effective volume.
ctlproj_status prepareGeometry(ctlproj_projector* p, const EffectiveVolume& ev)
{
    const size_t nm = size_t(p->V) * p->M;
    ctlproj_status st = ensureDevice(p->d_geom, p->geom_cap, nm);
    if (st != CTLPROJ_OK) return st;
    hipError_t e = ctlk::launchSetup(p->d_pmats, p->V, p->M, p->C, p->R, ev.geom,
                                     p->s.interpolate ? 1 : 0, p->d_geom, p->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "setup kernel");
}

ctlproj_status checkProject(ctlproj_projector* p, const ctlproj_volume* vol, const void* out,
                            size_t out_count, bool need_out = true)
{
    if (!p) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    if (!p->configured) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "project() called before configure()");
    if (!vol || !vol->d_data)
        return fail(CTLPROJ_ERR_NO_DATA, "no or contradictory data in volume object");
    if (vol->ctx != p->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume belongs to another context");
    const size_t need = size_t(p->V) * p->M * p->R * p->C;
    if (out_count != need)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT,
                    "output holds " + std::to_string(out_count) + " values, projection needs " +
                        std::to_string(need));
    if (need_out && need && !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output pointer is NULL");
    if (injected(p->ctx)) return CTLPROJ_ERR_INJECTED;
    return CTLPROJ_OK;
}

ctlproj_status launchCast(ctlproj_projector* p, const EffectiveVolume& ev, uint32_t v0, uint32_t nv,
                          float* d_out, int accumulate = 0, const ctlk::RayGeom* geom = nullptr,
                          unsigned long long* tally = nullptr)
{
    const ctlk::RayParams rp = rayParams(p, nv);
    const ctlk::LaunchShape sh = shapeFor(p->s, ev.view.storage, ev.view.layout, p->C, p->R);
    hipError_t e = ctlk::launchRaycast(ev.view, ev.geom, (geom ? geom : p->d_geom) + size_t(v0) * p->M, rp, sh,
                                       d_out, p->stream, accumulate, tally);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "raycast kernel");
}

} // namespace

extern "C" {

void ctlproj_settings_default(ctlproj_settings* s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->rays_per_pixel[0] = s->rays_per_pixel[1] = 1;
    s->ray_sampling = 0.3f;
    s->volume_upsampling = 1;
    s->interpolate = 1;
}

ctlproj_status ctlproj_settings_optimized_for(const float voxel_mm[3], double pixel_width_mm,
                                              double pixel_height_mm, ctlproj_settings* out)
{
    if (!voxel_mm || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (pixel_width_mm == 0.0 || pixel_height_mm == 0.0)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "pixel dimensions are singular");
    ctlproj_settings_default(out);
    // the preset ray grid whose aspect is closest to the pixel's (first one wins on ties), laid
    // with its longer side along the pixel's longer side
    double ratio = pixel_width_mm / pixel_height_mm;
    const bool wide = ratio > 1.0;
    if (wide) ratio = 1.0 / ratio;
    static const uint32_t presets[5][2] = { { 1, 2 }, { 1, 3 }, { 1, 4 }, { 2, 3 }, { 3, 4 } };
    double best = std::fabs(ratio - 1.0);
    for (const auto& r : presets) {
        const double dev = std::fabs(ratio - double(r[0]) / double(r[1]));
        if (dev < best) {
            best = dev;
            out->rays_per_pixel[0] = wide ? r[1] : r[0];
            out->rays_per_pixel[1] = wide ? r[0] : r[1];
        }
    }
    const double voxel = double(std::min(voxel_mm[0], std::min(voxel_mm[1], voxel_mm[2])));
    const double sub_pixel = std::min(pixel_width_mm / out->rays_per_pixel[0],
                                      pixel_height_mm / out->rays_per_pixel[1]);
    out->volume_upsampling = std::max(uint32_t(voxel / sub_pixel), 1u); // voxels larger than a sub-pixel
    const uint32_t more = std::max(uint32_t(sub_pixel / voxel), 1u);    // voxels smaller than a sub-pixel
    out->rays_per_pixel[0] *= more;
    out->rays_per_pixel[1] *= more;
    return CTLPROJ_OK;
}

const char* ctlproj_last_error(void) { return lastError(); }

ctlproj_status ctlproj_device_count(int* count)
{
    if (!count) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) n = 0;
    else if (e != hipSuccess) return hipFail(e, "hipGetDeviceCount");
    *count = n;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_context_create(int device, ctlproj_context** out)
{
    if (!out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(CTLPROJ_ERR_NO_DEVICE, "no HIP device available");
    // invalid device IDs throw in the reference (raycasterprojector.cpp:434-461)
    if (device < 0 || device >= n)
        return fail(CTLPROJ_ERR_NO_DEVICE, "device ID is not available. ID = " + std::to_string(device) +
                                               ", number of devices = " + std::to_string(n));
    auto* ctx = new (std::nothrow) ctlproj_context;
    if (!ctx) return fail(CTLPROJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    ctx->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete ctx;
        return hipFail(e, "context stream");
    }
    // the device's static scratch (allocation-free projection): its address, once per device
    if ((e = ctlk::initScratch(device)) != hipSuccess) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return hipFail(e, "hipGetSymbolAddress(scratch)");
    }
    *out = ctx;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_context_destroy(ctlproj_context* ctx)
{
    if (!ctx) return CTLPROJ_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_debug_inject_failure(ctlproj_context* ctx, int count)
{
    if (!ctx || count < 0) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad arguments");
    ctx->inject.store(count);
    return CTLPROJ_OK;
}

// ---- volumes ---------------------------------------------------------------------------------

ctlproj_status ctlproj_volume_create(ctlproj_context* ctx, const float* host, const uint32_t dims[3],
                                     const float voxel_mm[3], const float offset_mm[3],
                                     ctlproj_dtype storage, ctlproj_volume** out)
{
    if (!ctx || !out || !voxel_mm || !offset_mm) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (!host || !validDims(dims)) return fail(CTLPROJ_ERR_NO_DATA, "no or contradictory data in volume object");
    if (!validVoxel(voxel_mm) || !finite3(offset_mm))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "voxel size must be positive and finite, offset finite");
    if (storage < CTLPROJ_F32 || storage > CTLPROJ_BF16) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad storage type");
    ctlproj_status st = setDevice(ctx);
    if (st != CTLPROJ_OK) return st;
    if (injected(ctx)) return CTLPROJ_ERR_INJECTED;
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t count = size_t(dims[0]) * dims[1] * dims[2];
    float* d_src = nullptr;
    hipError_t e = hipMalloc(&d_src, count * sizeof(float));
    if (e != hipSuccess) return hipFail(e, "hipMalloc(volume upload)");
    e = hipMemcpyAsync(d_src, host, count * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(d_src);
        return hipFail(e, "hipMemcpyAsync(volume upload)");
    }
    st = createVolume(ctx, d_src, dims, voxel_mm, offset_mm, ctlk::Storage(storage), out);
    (void)hipFree(d_src);
    return st;
}

ctlproj_status ctlproj_volume_create_from_device(ctlproj_context* ctx, const float* device_ptr,
                                                 const uint32_t dims[3], const float voxel_mm[3],
                                                 const float offset_mm[3], ctlproj_dtype storage,
                                                 ctlproj_volume** out)
{
    if (!ctx || !out || !voxel_mm || !offset_mm) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (!device_ptr || !validDims(dims)) return fail(CTLPROJ_ERR_NO_DATA, "no or contradictory data in volume object");
    if (!validVoxel(voxel_mm) || !finite3(offset_mm))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "voxel size must be positive and finite, offset finite");
    if (storage < CTLPROJ_F32 || storage > CTLPROJ_BF16) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad storage type");
    ctlproj_status st = setDevice(ctx);
    if (st != CTLPROJ_OK) return st;
    if (injected(ctx)) return CTLPROJ_ERR_INJECTED;
    std::lock_guard<std::mutex> lk(ctx->mu);
    // the caller's array may be written by work on other streams: order after the device
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hipFail(e, "hipDeviceSynchronize");
    return createVolume(ctx, device_ptr, dims, voxel_mm, offset_mm, ctlk::Storage(storage), out);
}

ctlproj_status ctlproj_volume_destroy(ctlproj_volume* vol)
{
    if (!vol) return CTLPROJ_OK;
    (void)hipSetDevice(vol->ctx->device);
    if (vol->d_data) (void)hipFree(vol->d_data);
    if (vol->d_occ) (void)hipFree(vol->d_occ);
    if (vol->d_brick) (void)hipFree(vol->d_brick);
    delete vol;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_volume_bytes(const ctlproj_volume* vol, size_t* bytes)
{
    if (!vol || !bytes) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *bytes = vol->bytes;
    return CTLPROJ_OK;
}

// ---- projector -------------------------------------------------------------------------------

ctlproj_status ctlproj_projector_create(ctlproj_context* ctx, const ctlproj_settings* settings,
                                        ctlproj_projector** out)
{
    if (!ctx || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    ctlproj_settings s;
    if (settings) s = *settings;
    else ctlproj_settings_default(&s);
    ctlproj_status st = validateSettings(s);
    if (st != CTLPROJ_OK) return st;
    if ((st = setDevice(ctx)) != CTLPROJ_OK) return st;
    auto* p = new (std::nothrow) ctlproj_projector;
    if (!p) return fail(CTLPROJ_ERR_OUT_OF_MEMORY, "host allocation failed");
    p->ctx = ctx;
    p->s = s;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&p->ev_start)) != hipSuccess || (e = hipEventCreate(&p->ev_stop)) != hipSuccess ||
        (e = hipMalloc(&p->d_count, sizeof(unsigned long long))) != hipSuccess) {
        ctlproj_projector_destroy(p);
        return hipFail(e, "projector resources");
    }
    *out = p;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_projector_destroy(ctlproj_projector* p)
{
    if (!p) return CTLPROJ_OK;
    (void)hipSetDevice(p->ctx->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->copy_stream) (void)hipStreamSynchronize(p->copy_stream);
    if (p->d_pmats) (void)hipFree(p->d_pmats);
    if (p->d_geom) (void)hipFree(p->d_geom);
    if (p->d_out) (void)hipFree(p->d_out);
    if (p->d_count) (void)hipFree(p->d_count);
    for (float* h : p->h_stage)
        if (h) (void)hipHostFree(h);
    if (p->ev_start) (void)hipEventDestroy(p->ev_start);
    if (p->ev_stop) (void)hipEventDestroy(p->ev_stop);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (p->copy_stream) (void)hipStreamDestroy(p->copy_stream);
    delete p;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_projector_set_settings(ctlproj_projector* p, const ctlproj_settings* s)
{
    if (!p || !s) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    ctlproj_status st = validateSettings(*s);
    if (st != CTLPROJ_OK) return st;
    p->s = *s;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_projector_get_settings(const ctlproj_projector* p, ctlproj_settings* s)
{
    if (!p || !s) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *s = p->s;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_projector_configure(ctlproj_projector* p, const double* pmats, uint32_t views,
                                           uint32_t modules, uint32_t channels, uint32_t rows)
{
    if (!p) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    if (modules == 0 || channels == 0 || rows == 0)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "view dimensions must be non-zero");
    if (views && !pmats) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "pmats is NULL");
    const size_t n = size_t(views) * modules * 12;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(pmats[i])) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "non-finite projection matrix entry");
    ctlproj_status st = setDevice(p->ctx);
    if (st != CTLPROJ_OK) return st;
    p->configured = false;
    p->pmats.assign(pmats, pmats + n);
    p->V = views;
    p->M = modules;
    p->C = channels;
    p->R = rows;
    if (views) {
        if ((st = ensureDevice(p->d_pmats, p->d_pmats_cap, n)) != CTLPROJ_OK) return st;
        hipError_t e = hipMemcpyAsync(p->d_pmats, p->pmats.data(), n * sizeof(double),
                                      hipMemcpyHostToDevice, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) return hipFail(e, "P-matrix upload");
    }
    p->configured = true;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_project_device(ctlproj_projector* p, const ctlproj_volume* vol,
                                      float* device_out, size_t out_count)
{
    ctlproj_status st = checkProject(p, vol, device_out, out_count);
    if (st != CTLPROJ_OK) return st;
    if (p->V == 0) return CTLPROJ_OK;
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    EffectiveVolume ev;
    if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
    if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;
    hipError_t e = hipEventRecord(p->ev_start, p->stream);
    if (e != hipSuccess) return hipFail(e, "hipEventRecord");
    if ((st = launchCast(p, ev, 0, p->V, device_out)) != CTLPROJ_OK) return st;
    if ((e = hipEventRecord(p->ev_stop, p->stream)) != hipSuccess) return hipFail(e, "hipEventRecord");
    p->timed = true;
    if (ev.tmp) { // the temporary volume must outlive the kernel
        if ((e = hipStreamSynchronize(p->stream)) != hipSuccess) return hipFail(e, "hipStreamSynchronize");
    }
    return CTLPROJ_OK;
}

static_assert(ctlk::kScratchRayGeoms == CTLPROJ_NOALLOC_MAX_VIEW_MODULES,
              "the header's limit is the scratch's capacity");

ctlproj_status ctlproj_project_device_noalloc(ctlproj_projector* p, const ctlproj_volume* vol,
                                              float* device_out, size_t out_count)
{
    ctlproj_status st = checkProject(p, vol, device_out, out_count);
    if (st != CTLPROJ_OK) return st;
    if (p->V == 0) return CTLPROJ_OK;
    // upsampling factors other than 1 build a temporary volume (an allocation)
    const uint32_t f = p->s.interpolate ? p->s.volume_upsampling : 1u;
    if (f != 1)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT,
                    "the allocation-free projection needs volume upsampling factor 1 (got " + std::to_string(f) + ")");
    const size_t nm = size_t(p->V) * p->M;
    if (nm > ctlk::kScratchRayGeoms)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT,
                    "the allocation-free projection takes at most " + std::to_string(ctlk::kScratchRayGeoms) +
                        " (view, module) pairs, got " + std::to_string(nm));
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    auto* geom = static_cast<ctlk::RayGeom*>(ctlk::scratchPtr(p->ctx->device));
    if (!geom) return fail(CTLPROJ_ERR_DEVICE, "the device scratch has no address (context not initialised)");
    EffectiveVolume ev;
    if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
    // every record the launch reads is written by the setup kernel first, on the same stream
    hipError_t e = ctlk::launchSetup(p->d_pmats, p->V, p->M, p->C, p->R, ev.geom, p->s.interpolate ? 1 : 0, geom,
                                     p->stream);
    if (e != hipSuccess) return hipFail(e, "setup kernel");
    if ((e = hipEventRecord(p->ev_start, p->stream)) != hipSuccess) return hipFail(e, "hipEventRecord");
    if ((st = launchCast(p, ev, 0, p->V, device_out, 0, geom)) != CTLPROJ_OK) return st;
    if ((e = hipEventRecord(p->ev_stop, p->stream)) != hipSuccess) return hipFail(e, "hipEventRecord");
    p->timed = true;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_project_composite_device(ctlproj_projector* p, const ctlproj_volume* const* vols,
                                                uint32_t n, float* device_out, size_t out_count)
{
    if (!vols || n == 0) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "no volumes");
    ctlproj_status st;
    for (uint32_t k = 0; k < n; ++k)
        if ((st = checkProject(p, vols[k], device_out, out_count)) != CTLPROJ_OK) return st;
    if (p->V == 0) return CTLPROJ_OK;
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    hipError_t e = hipEventRecord(p->ev_start, p->stream);
    if (e != hipSuccess) return hipFail(e, "hipEventRecord");
    // one stream: the sub-volumes' projections add up in launch order, deterministically
    for (uint32_t k = 0; k < n; ++k) {
        EffectiveVolume ev;
        if ((st = effectiveVolume(vols[k], p->s, p->stream, ev)) != CTLPROJ_OK) return st;
        if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;
        if ((st = launchCast(p, ev, 0, p->V, device_out, k > 0 ? 1 : 0)) != CTLPROJ_OK) return st;
        if (ev.tmp && (e = hipStreamSynchronize(p->stream)) != hipSuccess) // temporary outlives the kernel
            return hipFail(e, "hipStreamSynchronize");
    }
    if ((e = hipEventRecord(p->ev_stop, p->stream)) != hipSuccess) return hipFail(e, "hipEventRecord");
    p->timed = true;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_project_composite(ctlproj_projector* p, const ctlproj_volume* const* vols, uint32_t n,
                                         float* out, size_t out_count)
{
    if (!p) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    if (!out && out_count) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output pointer is NULL");
    ctlproj_status st = setDevice(p->ctx);
    if (st != CTLPROJ_OK) return st;
    const size_t need = size_t(p->V) * p->M * p->R * p->C;
    if (need && out_count == need && (st = ensureDevice(p->d_out, p->out_cap, need)) != CTLPROJ_OK) return st;
    if ((st = ctlproj_project_composite_device(p, vols, n, p->d_out, out_count)) != CTLPROJ_OK) return st;
    if (need == 0) return CTLPROJ_OK;
    hipError_t e = hipMemcpyAsync(out, p->d_out, need * sizeof(float), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "D2H copy");
}

ctlproj_status ctlproj_average_projections(ctlproj_context* ctx, const float* device_in, uint32_t k,
                                           size_t n, int intensity_domain, float* device_out)
{
    if (!ctx || (k && n && (!device_in || !device_out))) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (k == 0) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "no projections to average");
    ctlproj_status st = setDevice(ctx);
    if (st != CTLPROJ_OK) return st;
    if (injected(ctx)) return CTLPROJ_ERR_INJECTED;
    hipError_t e = ctlk::launchAverage(device_in, k, n, intensity_domain ? 1 : 0, device_out, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "average kernel");
}

ctlproj_status ctlproj_projector_synchronize(ctlproj_projector* p)
{
    if (!p) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    hipError_t e = hipStreamSynchronize(p->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "hipStreamSynchronize");
}

ctlproj_status ctlproj_projector_last_kernel_ms(ctlproj_projector* p, float* ms)
{
    if (!p || !ms) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!p->timed) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "no kernel has been timed yet");
    hipError_t e = hipEventSynchronize(p->ev_stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, p->ev_start, p->ev_stop);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "hipEventElapsedTime");
}

namespace {
// Copy of a staged chunk into the caller's array, split over up to four threads when it is large:
// a freshly allocated array takes a page fault per page on first touch, and the faults of one
// thread bound the copy (360 views of 1024 x 768 into a new array: 97-99 ms end to end with one
// thread, 96 ms with four).
void hostCopy(void* dst, const void* src, size_t bytes)
{
    const size_t kMinSplit = size_t(4) << 20;
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t parts = bytes < kMinSplit ? 1 : std::min<size_t>(4, hw > 1 ? hw : 1);
    if (parts <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = ((bytes / parts) + 4095) & ~size_t(4095);
    char* d = static_cast<char*>(dst);
    const char* s = static_cast<const char*>(src);
    std::vector<std::thread> helpers;
    for (size_t t = 1; t < parts; ++t) {
        const size_t a = std::min(bytes, part * t), b = std::min(bytes, part * (t + 1));
        if (a >= b) continue;
        try {
            helpers.emplace_back([=]() { std::memcpy(d + a, s + a, b - a); });
        } catch (...) { // no thread: copy this part here (no exception crosses the C ABI)
            std::memcpy(d + a, s + a, b - a);
        }
    }
    std::memcpy(d, s, std::min(bytes, part));
    for (auto& h : helpers) h.join();
}
} // namespace

ctlproj_status ctlproj_project(ctlproj_projector* p, const ctlproj_volume* vol, float* out,
                               size_t out_count, ctlproj_progress_fn progress, void* user)
{
    ctlproj_status st = checkProject(p, vol, out, out_count);
    if (st != CTLPROJ_OK) return st;
    if (p->V == 0) return CTLPROJ_OK;
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    EffectiveVolume ev;
    if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
    if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;

    const size_t per_view = size_t(p->M) * p->R * p->C;
    if ((st = ensureDevice(p->d_out, p->out_cap, per_view * p->V)) != CTLPROJ_OK) return st;
    // Views go out in chunks of about 32 MiB: the device-to-host copy and the host copy into
    // the caller's array of one chunk overlap the ray casting of the next ones.
    const uint32_t chunk = uint32_t(std::max<size_t>(1, std::min<size_t>(p->V, (size_t(32) << 20) / (per_view * 4 + 1))));
    const size_t stage = size_t(chunk) * per_view;
    hipError_t e;
    if (stage > p->stage_cap) {
        for (float*& h : p->h_stage) {
            if (h) (void)hipHostFree(h);
            h = nullptr;
        }
        p->stage_cap = 0;
        for (float*& h : p->h_stage)
            if ((e = hipHostMalloc(&h, stage * sizeof(float), hipHostMallocDefault)) != hipSuccess)
                return hipFail(e, "hipHostMalloc(staging)");
        p->stage_cap = stage;
    }
    const uint32_t n_chunks = (p->V + chunk - 1) / chunk;
    std::vector<hipEvent_t> done(n_chunks, nullptr), copied(n_chunks, nullptr);
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(p->stream);
        (void)hipStreamSynchronize(p->copy_stream);
        for (auto ev2 : done) if (ev2) (void)hipEventDestroy(ev2);
        for (auto ev2 : copied) if (ev2) (void)hipEventDestroy(ev2);
    };
    if ((e = hipEventRecord(p->ev_start, p->stream)) != hipSuccess) { cleanup(); return hipFail(e, "hipEventRecord"); }
    for (uint32_t k = 0; k < n_chunks; ++k) {
        const uint32_t v0 = k * chunk, nv = std::min(chunk, p->V - v0);
        if ((st = launchCast(p, ev, v0, nv, p->d_out + size_t(v0) * per_view)) != CTLPROJ_OK) { cleanup(); return st; }
        if ((e = hipEventCreateWithFlags(&done[k], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventRecord(done[k], p->stream)) != hipSuccess) { cleanup(); return hipFail(e, "hipEventRecord"); }
    }
    if ((e = hipEventRecord(p->ev_stop, p->stream)) != hipSuccess) { cleanup(); return hipFail(e, "hipEventRecord"); }
    p->timed = true;
    auto drain = [&](uint32_t k) -> ctlproj_status {
        const uint32_t v0 = k * chunk, nv = std::min(chunk, p->V - v0);
        hipError_t e2 = hipEventSynchronize(copied[k]);
        if (e2 != hipSuccess) return hipFail(e2, "D2H copy");
        hostCopy(out + size_t(v0) * per_view, p->h_stage[k % 2], size_t(nv) * per_view * sizeof(float));
        if (progress)
            for (uint32_t v = v0; v < v0 + nv; ++v) progress(int(v), user);
        return CTLPROJ_OK;
    };
    for (uint32_t k = 0; k < n_chunks; ++k) {
        const uint32_t v0 = k * chunk, nv = std::min(chunk, p->V - v0);
        // staging buffer k%2 was last used by chunk k-2, drained below before this point
        if ((e = hipStreamWaitEvent(p->copy_stream, done[k], 0)) != hipSuccess ||
            (e = hipMemcpyAsync(p->h_stage[k % 2], p->d_out + size_t(v0) * per_view,
                                size_t(nv) * per_view * sizeof(float), hipMemcpyDeviceToHost,
                                p->copy_stream)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&copied[k], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventRecord(copied[k], p->copy_stream)) != hipSuccess) {
            cleanup();
            return hipFail(e, "D2H copy");
        }
        if (k >= 1 && (st = drain(k - 1)) != CTLPROJ_OK) { cleanup(); return st; }
    }
    st = drain(n_chunks - 1);
    cleanup();
    return st;
}

ctlproj_status ctlproj_project_host_volume(ctlproj_projector* p, const float* host_volume,
                                           const uint32_t dims[3], const float voxel_mm[3],
                                           const float offset_mm[3], float* out, size_t out_count)
{
    if (!p) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    ctlproj_volume* vol = nullptr;
    ctlproj_status st = ctlproj_volume_create(p->ctx, host_volume, dims, voxel_mm, offset_mm, CTLPROJ_F32, &vol);
    if (st != CTLPROJ_OK) return st;
    st = ctlproj_project(p, vol, out, out_count, nullptr, nullptr);
    ctlproj_volume_destroy(vol);
    return st;
}

ctlproj_status ctlproj_projector_count_samples(ctlproj_projector* p, const ctlproj_volume* vol,
                                               uint64_t* n_samples)
{
    if (!n_samples) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "n_samples is NULL");
    ctlproj_status st =
        checkProject(p, vol, nullptr, p ? size_t(p->V) * p->M * p->R * p->C : 0, /*need_out=*/false);
    if (st != CTLPROJ_OK) return st;
    *n_samples = 0;
    if (p->V == 0) return CTLPROJ_OK;
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    EffectiveVolume ev;
    if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
    if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;
    hipError_t e = hipMemsetAsync(p->d_count, 0, sizeof(unsigned long long), p->stream);
    if (e == hipSuccess) e = ctlk::launchCount(ev.geom, p->d_geom, rayParams(p, p->V), p->d_count, p->stream);
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, p->d_count, sizeof h, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return hipFail(e, "count kernel");
    *n_samples = h;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_projector_count_taken(ctlproj_projector* p, const ctlproj_volume* vol, float* device_out,
                                            size_t out_count, uint64_t* n_taken)
{
    if (!n_taken) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "n_taken is NULL");
    ctlproj_status st = checkProject(p, vol, device_out, out_count);
    if (st != CTLPROJ_OK) return st;
    *n_taken = 0;
    if (p->V == 0) return CTLPROJ_OK;
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    EffectiveVolume ev;
    if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
    if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;
    hipError_t e = hipMemsetAsync(p->d_count, 0, sizeof(unsigned long long), p->stream);
    if (e != hipSuccess) return hipFail(e, "hipMemsetAsync");
    if ((st = launchCast(p, ev, 0, p->V, device_out, 0, nullptr, p->d_count)) != CTLPROJ_OK) return st;
    unsigned long long h = 0;
    e = hipMemcpyAsync(&h, p->d_count, sizeof h, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return hipFail(e, "sample tally");
    *n_taken = h;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_projector_autotune(ctlproj_projector* p, const ctlproj_volume* vol,
                                         float* best_ms)
{
    ctlproj_status st = checkProject(p, vol, nullptr, p ? size_t(p->V) * p->M * p->R * p->C : 0, false);
    if (st != CTLPROJ_OK) return st;
    if (p->V == 0) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "no views to tune on");
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    EffectiveVolume ev;
    if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
    if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;
    // time on up to 24 evenly spaced views, gathered into a contiguous geometry block
    const uint32_t nv = std::min<uint32_t>(p->V, 24);
    const size_t per_view = size_t(p->M) * p->R * p->C;
    ctlk::RayGeom* d_g = nullptr;
    float* d_o = nullptr;
    hipError_t e = hipMalloc(&d_g, size_t(nv) * p->M * sizeof(ctlk::RayGeom));
    if (e == hipSuccess) e = hipMalloc(&d_o, size_t(nv) * per_view * sizeof(float));
    for (uint32_t k = 0; k < nv && e == hipSuccess; ++k) {
        const uint32_t v = uint32_t(uint64_t(k) * p->V / nv);
        e = hipMemcpyAsync(d_g + size_t(k) * p->M, p->d_geom + size_t(v) * p->M, p->M * sizeof(ctlk::RayGeom),
                           hipMemcpyDeviceToDevice, p->stream);
    }
    if (e != hipSuccess) {
        if (d_g) (void)hipFree(d_g);
        if (d_o) (void)hipFree(d_o);
        return hipFail(e, "autotune buffers");
    }
    const ctlk::RayParams rp = rayParams(p, nv);
    auto run = [&](const ctlk::LaunchShape& sh) -> float {
        hipError_t r = hipEventRecord(p->ev_start, p->stream);
        if (r == hipSuccess) r = ctlk::launchRaycast(ev.view, ev.geom, d_g, rp, sh, d_o, p->stream);
        if (r == hipSuccess) r = hipEventRecord(p->ev_stop, p->stream);
        if (r == hipSuccess) r = hipEventSynchronize(p->ev_stop);
        float ms = -1.0f;
        if (r == hipSuccess && hipEventElapsedTime(&ms, p->ev_start, p->ev_stop) != hipSuccess) ms = -1.0f;
        return r == hipSuccess ? ms : -1.0f;
    };
    ctltune::Choice best;
    const bool staged = stagedWorkload(p->s, ev.view.storage, ev.view.layout);
    const bool ok = ctltune::tune(ctltune::key(ev.view.storage, ev.view.layout, p->s.interpolate ? 1 : 0,
                                               p->C, p->R, staged),
                                  run, &best, staged, p->C);
    (void)hipFree(d_g);
    (void)hipFree(d_o);
    if (!ok) return fail(CTLPROJ_ERR_DEVICE, "autotuning failed: no candidate ran");
    if (best_ms) *best_ms = best.ms;
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_decode_geometry(const double* pmats, uint32_t views, uint32_t modules,
                                       uint32_t channels, uint32_t rows, int reference, double a,
                                       double b, double* source_pos, double* source_rot,
                                       double* pixel_size, double* skew, double* module_pos,
                                       double* module_rot)
{
    if (!pmats || !source_pos || !source_rot || !pixel_size || !skew || !module_pos || !module_rot)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (modules == 0 || channels == 0 || rows == 0) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "empty detector");
    if (reference < CTLPROJ_DECODE_PIXEL_SIZE || reference > CTLPROJ_DECODE_SOURCE_DETECTOR_DISTANCE)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "invalid value for the physical dimension");
    if (!(a > 0.0) || (reference == CTLPROJ_DECODE_PIXEL_SIZE && !(b > 0.0)))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "reference lengths must be positive");
    std::vector<ctlgeo::PMat> P(modules);
    ctlgeo::DecodedView d;
    for (uint32_t v = 0; v < views; ++v) {
        for (uint32_t m = 0; m < modules; ++m)
            std::memcpy(P[m].p, pmats + (size_t(v) * modules + m) * 12, 12 * sizeof(double));
        if (!ctlgeo::decodeView(P.data(), modules, channels, rows, reference, a, b, &d))
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "singular projection matrix in view " + std::to_string(v));
        std::memcpy(source_pos + 3 * size_t(v), d.source.v, 3 * sizeof(double));
        std::memcpy(source_rot + 9 * size_t(v), d.source_rotation.m, 9 * sizeof(double));
        pixel_size[2 * size_t(v)] = d.pixel_w;
        pixel_size[2 * size_t(v) + 1] = d.pixel_h;
        skew[v] = d.skew;
        for (uint32_t m = 0; m < modules; ++m) {
            const size_t k = size_t(v) * modules + m;
            std::memcpy(module_pos + 3 * k, d.modules[m].position.v, 3 * sizeof(double));
            std::memcpy(module_rot + 9 * k, d.modules[m].rotation.m, 9 * sizeof(double));
        }
    }
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_encode_generic(const double* source_pos, const double* module_pos,
                                      const double* module_rot, const double* pixel_size,
                                      const double* skew, uint32_t views, uint32_t modules,
                                      uint32_t channels, uint32_t rows, double* out, size_t out_count)
{
    if (!source_pos || !module_pos || !module_rot || !pixel_size || !out)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (out_count != size_t(views) * modules * 12)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output must hold views*modules*12 values");
    if (modules == 0 || channels == 0 || rows == 0) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "empty detector");
    std::vector<ctlgeo::Location> locs(modules);
    std::vector<ctlgeo::PMat> P(modules);
    for (uint32_t v = 0; v < views; ++v) {
        ctlgeo::Vec3 src;
        std::memcpy(src.v, source_pos + 3 * size_t(v), 3 * sizeof(double));
        for (uint32_t m = 0; m < modules; ++m) {
            const size_t k = size_t(v) * modules + m;
            std::memcpy(locs[m].position.v, module_pos + 3 * k, 3 * sizeof(double));
            std::memcpy(locs[m].rotation.m, module_rot + 9 * k, 9 * sizeof(double));
        }
        const double pw = pixel_size[2 * size_t(v)], ph = pixel_size[2 * size_t(v) + 1];
        if (!(pw > 0.0) || !(ph > 0.0)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "pixel size must be positive");
        ctlgeo::encodeModules(src, locs.data(), modules, channels, rows, pw, ph, skew ? skew[v] : 0.0, P.data());
        for (uint32_t m = 0; m < modules; ++m)
            std::memcpy(out + (size_t(v) * modules + m) * 12, P[m].p, 12 * sizeof(double));
    }
    return CTLPROJ_OK;
}

// ---- gather-ceiling probes (diagnostics) -----------------------------------------------------

ctlproj_status ctlproj_diag_gather_probe(ctlproj_projector* p, const ctlproj_volume* vol, int mode,
                                         float* device_out, size_t out_count, float* ms, double* bytes)
{
    if (!ms || !bytes) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (mode != CTLPROJ_PROBE_GATHER && mode != CTLPROJ_PROBE_STREAM && mode != CTLPROJ_PROBE_LDS &&
        mode != CTLPROJ_PROBE_LDS_PAIRS)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "unknown probe mode");
    ctlproj_status st = CTLPROJ_OK;
    hipError_t e = hipSuccess;
    if (mode == CTLPROJ_PROBE_GATHER) {
        if ((st = checkProject(p, vol, device_out, out_count)) != CTLPROJ_OK) return st;
        if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
        if (p->V == 0) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "no views to probe");
        if (!p->s.interpolate || p->s.rays_per_pixel[0] != 1 || p->s.rays_per_pixel[1] != 1 ||
            p->s.volume_upsampling != 1 || vol->view.layout != ctlk::Layout::Interleaved || vol->view.large)
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT,
                        "the gather probe needs trilinear, one ray per pixel, no upsampling and an "
                        "interleaved volume below 4 GiB");
        EffectiveVolume ev;
        if ((st = effectiveVolume(vol, p->s, p->stream, ev)) != CTLPROJ_OK) return st;
        if ((st = prepareGeometry(p, ev)) != CTLPROJ_OK) return st;
        const ctlk::RayParams rp = rayParams(p, p->V);
        unsigned long long n = 0;
        e = hipMemsetAsync(p->d_count, 0, sizeof(unsigned long long), p->stream);
        if (e == hipSuccess) e = ctlk::launchCount(ev.geom, p->d_geom, rp, p->d_count, p->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(&n, p->d_count, sizeof n, hipMemcpyDeviceToHost, p->stream);
        if (e == hipSuccess) e = hipEventRecord(p->ev_start, p->stream);
        if (e == hipSuccess) e = ctlk::launchProbeGather(ev.view, ev.geom, p->d_geom, rp, device_out, p->stream);
        if (e == hipSuccess) e = hipEventRecord(p->ev_stop, p->stream);
        if (e == hipSuccess) e = hipEventSynchronize(p->ev_stop);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, p->ev_start, p->ev_stop);
        if (e != hipSuccess) return hipFail(e, "gather probe");
        *bytes = double(n) * (ev.view.storage == ctlk::Storage::F32 ? 32.0 : 16.0);
        return CTLPROJ_OK;
    }
    if (!p || !vol || !vol->d_data || !device_out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (vol->ctx != p->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume belongs to another context");
    if ((st = setDevice(p->ctx)) != CTLPROJ_OK) return st;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipDeviceGetAttribute");
    const uint32_t grid = uint32_t(cus) * 8, iters = 16384;
    if (out_count < size_t(grid) * 256)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "the stream and LDS probes need 2048 floats per CU of output");
    if (mode == CTLPROJ_PROBE_STREAM && vol->view.bytes < 16384)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "the stream probe needs a volume of 16 KiB or more");
    e = hipEventRecord(p->ev_start, p->stream);
    if (e == hipSuccess)
        e = mode == CTLPROJ_PROBE_STREAM
                ? ctlk::launchProbeStream(vol->view.data, grid, iters, device_out, p->stream)
                : ctlk::launchProbeLds(grid, iters, mode == CTLPROJ_PROBE_LDS_PAIRS, device_out, p->stream);
    if (e == hipSuccess) e = hipEventRecord(p->ev_stop, p->stream);
    if (e == hipSuccess) e = hipEventSynchronize(p->ev_stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, p->ev_start, p->ev_stop);
    if (e != hipSuccess) return hipFail(e, mode == CTLPROJ_PROBE_STREAM ? "stream probe" : "LDS probe");
    // the pair probe reads 15 of every 16 elements
    *bytes = double(grid) * 256.0 * iters * 16.0 * (mode == CTLPROJ_PROBE_LDS_PAIRS ? 15.0 / 16.0 : 1.0);
    return CTLPROJ_OK;
}

// ---- geometry --------------------------------------------------------------------------------

static bool toDetector(const ctlproj_detector* d, ctlgeo::Detector& out)
{
    if (!d || d->pixels[0] == 0 || d->pixels[1] == 0 || !(d->pixel_mm[0] > 0) || !(d->pixel_mm[1] > 0))
        return false;
    out.channels = d->pixels[0];
    out.rows = d->pixels[1];
    out.pixel_w = d->pixel_mm[0];
    out.pixel_h = d->pixel_mm[1];
    out.cylindrical = d->cylindrical != 0;
    out.modules = out.cylindrical ? d->modules : 1;
    if (out.modules == 0) return false;
    out.angulation = d->angulation;
    out.spacing = d->module_spacing;
    out.skew = d->skew;
    return true;
}

ctlproj_status ctlproj_encode_trajectory(const ctlproj_gantry* gantry, const ctlproj_detector* det,
                                         const ctlproj_trajectory* traj, uint32_t views,
                                         double* out, size_t out_count)
{
    return ctlproj_encode_trajectory_shifted(gantry, det, traj, views, nullptr, out, out_count);
}

ctlproj_status ctlproj_encode_trajectory_shifted(const ctlproj_gantry* gantry, const ctlproj_detector* det,
                                                 const ctlproj_trajectory* traj, uint32_t views,
                                                 const double source_shift[3], double* out, size_t out_count)
{
    ctlgeo::Detector d;
    if (!gantry || !traj || !toDetector(det, d)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad system description");
    if (out_count != size_t(views) * d.modules * 12 || (views && !out))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output must hold views*modules*12 doubles");
    ctlgeo::Gantry g;
    g.type = gantry->type == CTLPROJ_TUBULAR ? ctlgeo::Gantry::Tubular : ctlgeo::Gantry::CArm;
    g.carm_span = gantry->carm_span;
    g.sdd = gantry->sdd;
    g.sid = gantry->sid;
    g.rotation_angle = gantry->rotation;
    g.pitch_position = gantry->pitch;
    g.tilt = gantry->tilt;
    ctlgeo::Trajectory t;
    if (traj->type < 0 || traj->type > 4) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "unknown trajectory");
    t.type = ctlgeo::Trajectory::Type(traj->type);
    for (int i = 0; i < 6; ++i) t.p[i] = traj->p[i];
    std::vector<ctlgeo::PMat> pm(d.modules);
    ctlgeo::Vec3 focal{ { 0, 0, 0 } };
    if (source_shift)
        for (int k = 0; k < 3; ++k) focal[k] = source_shift[k];
    for (uint32_t v = 0; v < views; ++v) {
        if (!ctlgeo::prepareView(t, v, views, d, g))
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "trajectory is not applicable to the gantry type");
        ctlgeo::encodeSingleView(g, d, focal, pm.data());
        for (uint32_t m = 0; m < d.modules; ++m)
            std::memcpy(out + (size_t(v) * d.modules + m) * 12, pm[m].p, 12 * sizeof(double));
    }
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_encode_carm_poses(double carm_span, const ctlproj_detector* det,
                                         const double* positions, const double* rotations,
                                         uint32_t views, double* out, size_t out_count)
{
    return ctlproj_encode_carm_poses_shifted(carm_span, det, positions, rotations, views, nullptr, out,
                                             out_count);
}

ctlproj_status ctlproj_encode_carm_poses_shifted(double carm_span, const ctlproj_detector* det,
                                                 const double* positions, const double* rotations,
                                                 uint32_t views, const double source_shift[3], double* out,
                                                 size_t out_count)
{
    ctlgeo::Detector d;
    if (!toDetector(det, d)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad detector description");
    if (views && (!positions || !rotations || !out))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (out_count != size_t(views) * d.modules * 12)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output must hold views*modules*12 doubles");
    ctlgeo::Gantry g;
    g.type = ctlgeo::Gantry::CArm;
    g.carm_span = carm_span;
    std::vector<ctlgeo::PMat> pm(d.modules);
    ctlgeo::Vec3 focal{ { 0, 0, 0 } };
    if (source_shift)
        for (int k = 0; k < 3; ++k) focal[k] = source_shift[k];
    for (uint32_t v = 0; v < views; ++v) {
        for (int k = 0; k < 3; ++k) g.carm_location.position[k] = positions[3 * v + k];
        for (int k = 0; k < 9; ++k) g.carm_location.rotation.m[k] = rotations[9 * v + k];
        ctlgeo::encodeSingleView(g, d, focal, pm.data());
        for (uint32_t m = 0; m < d.modules; ++m)
            std::memcpy(out + (size_t(v) * d.modules + m) * 12, pm[m].p, 12 * sizeof(double));
    }
    return CTLPROJ_OK;
}

// ---- phantoms --------------------------------------------------------------------------------

ctlproj_status ctlproj_phantom_ellipsoids(const uint32_t dims[3], const float voxel_mm[3],
                                          uint32_t n_ellipsoids, uint64_t seed,
                                          float semi_axis_min_mm, float semi_axis_max_mm,
                                          float mu_min, float mu_max, float background_mu,
                                          float* out)
{
    if (!validDims(dims) || !voxel_mm || !validVoxel(voxel_mm) || !out)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "bad phantom arguments");
    ctlphantom::PhantomSpec s;
    for (int k = 0; k < 3; ++k) {
        s.dims[k] = dims[k];
        s.voxel_mm[k] = voxel_mm[k];
    }
    s.n_ellipsoids = n_ellipsoids;
    s.seed = seed;
    s.semi_axis_min_mm = semi_axis_min_mm;
    s.semi_axis_max_mm = semi_axis_max_mm;
    s.mu_min = mu_min;
    s.mu_max = mu_max;
    s.background_mu = background_mu;
    ctlphantom::ellipsoids(s, out);
    return CTLPROJ_OK;
}

} // extern "C"
