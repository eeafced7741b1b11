// Device-side parts of the projector extensions that call the projection many times
// (modules/src/projectors/spectraleffectsextension.cpp, dynamicprojectorextension.cpp): the
// projections stay in device memory and are combined there by one kernel.
#include "ctlproj.h"
#include "runtime.h"

#include <cmath>
#include <vector>

using namespace ctlrt;

namespace {

ctlproj_status begin(ctlproj_context* ctx)
{
    if (!ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "context is NULL");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    if (injected(ctx)) return CTLPROJ_ERR_INJECTED;
    return CTLPROJ_OK;
}

ctlproj_status finish(hipError_t e, hipStream_t s, const char* what)
{
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, what);
}

} // namespace

ctlproj_status ctlproj_spectral_extinction(ctlproj_context* ctx, const float* device_material_projs,
                                           uint32_t n_materials, uint32_t views, size_t pixels_per_view,
                                           const float* mu, const double* intensity, const float* response,
                                           uint32_t n_bins, const double* total_intensity, float* device_out)
{
    ctlproj_status st = begin(ctx);
    if (st != CTLPROJ_OK) return st;
    if (n_materials == 0 || n_materials > ctlk::kMaxMaterials)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "1 to 8 materials are supported");
    if (views == 0 || pixels_per_view == 0) return CTLPROJ_OK;
    if (!device_material_projs || !device_out || !total_intensity || (n_bins && (!mu || !intensity)))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    // bins whose intensities sum to zero over the views are skipped (spectraleffectsextension.cpp:
    // 337-341, qFuzzyIsNull); intensities and I0 enter the transforms as float (singleviewdata.cpp:
    // 137, 152)
    std::vector<float> m, in, r, tot(views);
    uint32_t used = 0;
    for (uint32_t b = 0; b < n_bins; ++b) {
        double sum = 0.0;
        for (uint32_t v = 0; v < views; ++v) sum += intensity[size_t(b) * views + v];
        if (std::fabs(sum) <= 1e-12) continue;
        for (uint32_t k = 0; k < n_materials; ++k) m.push_back(mu[size_t(b) * n_materials + k]);
        for (uint32_t v = 0; v < views; ++v) in.push_back(float(intensity[size_t(b) * views + v]));
        r.push_back(response ? response[b] : 1.0f);
        ++used;
    }
    for (uint32_t v = 0; v < views; ++v) tot[v] = float(total_intensity[v]);
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    DeviceBuffer d_mu, d_in, d_r, d_tot;
    if ((st = deviceUpload(d_mu, m.data(), m.size() * sizeof(float), s, "spectral constants")) != CTLPROJ_OK ||
        (st = deviceUpload(d_in, in.data(), in.size() * sizeof(float), s, "spectral constants")) != CTLPROJ_OK ||
        (st = deviceUpload(d_r, r.data(), r.size() * sizeof(float), s, "spectral constants")) != CTLPROJ_OK ||
        (st = deviceUpload(d_tot, tot.data(), tot.size() * sizeof(float), s, "spectral constants")) != CTLPROJ_OK)
        return st;
    // the material projections may have been written on other streams
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = ctlk::launchSpectral(device_material_projs, n_materials, views, pixels_per_view, d_mu.as<float>(),
                                 d_in.as<float>(), d_r.as<float>(), used, d_tot.as<float>(), device_out, s);
    return finish(e, s, "spectral kernel");
}

ctlproj_status ctlproj_combine_views(ctlproj_context* ctx, const float* device_a, const float* coeff_a,
                                     const float* device_b, const float* coeff_b, uint32_t views,
                                     size_t pixels_per_view, float* device_out)
{
    ctlproj_status st = begin(ctx);
    if (st != CTLPROJ_OK) return st;
    if (views == 0 || pixels_per_view == 0) return CTLPROJ_OK;
    if (!device_a || !coeff_a || !device_out || (device_b && !coeff_b))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    hipStream_t s = ctx->stream;
    DeviceBuffer d_ca, d_cb;
    if ((st = deviceUpload(d_ca, coeff_a, views * sizeof(float), s, "view coefficients")) != CTLPROJ_OK) return st;
    if (device_b && (st = deviceUpload(d_cb, coeff_b, views * sizeof(float), s, "view coefficients")) != CTLPROJ_OK)
        return st;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = ctlk::launchCombineViews(device_a, d_ca.as<float>(), device_b, device_b ? d_cb.as<float>() : nullptr,
                                     views, pixels_per_view, device_out, s);
    return finish(e, s, "combine kernel");
}

namespace {

// Projections of n volumes with the configured projector, one after another into d_proj.
ctlproj_status projectAll(ctlproj_projector* p, const ctlproj_volume* const* vols, uint32_t n, float* d_proj,
                          size_t per_proj)
{
    for (uint32_t k = 0; k < n; ++k) {
        ctlproj_status st = ctlproj_project_device(p, vols[k], d_proj + k * per_proj, per_proj);
        if (st != CTLPROJ_OK) return st;
    }
    return ctlproj_projector_synchronize(p);
}

// Copy of a result to the caller's array on the projector's own (non-blocking) stream: a copy on
// the legacy null stream would wait for every running launch of the process, serving launches
// included.
ctlproj_status copyOut(ctlproj_projector* p, const float* d, float* host, size_t n)
{
    hipError_t e = hipMemcpyAsync(host, d, n * sizeof(float), hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, "result copy");
}

} // namespace

ctlproj_status ctlproj_project_spectral(ctlproj_projector* p, const ctlproj_volume* const* densities,
                                        uint32_t n_materials, const float* mu, const double* intensity,
                                        const float* response, uint32_t n_bins, const double* total_intensity,
                                        float* out, size_t out_count)
{
    if (!p || !p->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    if (!p->configured) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "project() called before configure()");
    if (n_materials == 0 || n_materials > ctlk::kMaxMaterials)
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "1 to 8 materials are supported");
    if (!densities) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volumes is NULL");
    const size_t per = size_t(p->V) * p->M * p->R * p->C;
    if (out_count != per) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output size does not match the configured geometry");
    if (per == 0) return CTLPROJ_OK;
    if (!out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output is NULL");
    ctlproj_status st = begin(p->ctx);
    if (st != CTLPROJ_OK) return st;
    DeviceBuffer d_proj, d_out;
    if ((st = deviceAlloc(d_proj, n_materials * per * sizeof(float), "material projections")) != CTLPROJ_OK ||
        (st = deviceAlloc(d_out, per * sizeof(float), "spectral output")) != CTLPROJ_OK)
        return st;
    if ((st = projectAll(p, densities, n_materials, d_proj.as<float>(), per)) != CTLPROJ_OK) return st;
    const size_t pixels = per / p->V;
    if ((st = ctlproj_spectral_extinction(p->ctx, d_proj.as<float>(), n_materials, p->V, pixels, mu, intensity,
                                          response, n_bins, total_intensity, d_out.as<float>())) != CTLPROJ_OK)
        return st;
    return copyOut(p, d_out.as<float>(), out, per);
}

ctlproj_status ctlproj_project_linear_dynamic(ctlproj_projector* p, const ctlproj_volume* slope,
                                              const ctlproj_volume* offset, const double* time_stamps, float* out,
                                              size_t out_count)
{
    if (!p || !p->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "projector is NULL");
    if (!p->configured) return fail(CTLPROJ_ERR_NOT_CONFIGURED, "project() called before configure()");
    if (!slope || !offset) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume is NULL");
    for (int k = 0; k < 3; ++k)
        if (slope->n[k] != offset->n[k])
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "offset and slope volumes have different sizes");
    const size_t per = size_t(p->V) * p->M * p->R * p->C;
    if (out_count != per) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output size does not match the configured geometry");
    if (per == 0) return CTLPROJ_OK;
    if (!out || !time_stamps) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    ctlproj_status st = begin(p->ctx);
    if (st != CTLPROJ_OK) return st;
    DeviceBuffer d_proj;
    if ((st = deviceAlloc(d_proj, 2 * per * sizeof(float), "dynamic projections")) != CTLPROJ_OK) return st;
    const ctlproj_volume* vols[2] = { slope, offset };
    if ((st = projectAll(p, vols, 2, d_proj.as<float>(), per)) != CTLPROJ_OK) return st;
    // time stamps enter as float, as LinearDynamicVolume::updateVolume does (float(time()))
    std::vector<float> ca(p->V), cb(p->V, 1.0f);
    for (uint32_t v = 0; v < p->V; ++v) ca[v] = float(time_stamps[v]);
    float* a = d_proj.as<float>();
    if ((st = ctlproj_combine_views(p->ctx, a, ca.data(), a + per, cb.data(), p->V, per / p->V, a)) != CTLPROJ_OK)
        return st;
    return copyOut(p, a, out, per);
}
