// ctlproj.hpp -- header-only C++ mirror of CTL's projector interface over the C ABI (ctlproj.h).
//
// Mirrors CTL::OCL::RayCasterProjector (modules/src/projectors/raycasterprojector.h:20-58) and
// AbstractProjector (modules/src/projectors/abstractprojector.h:65-112) without Qt, and the OCL
// processing classes VolumeSlicer, VolumeResampler and RadonTransform3D on resident volumes:
//   RayCasterProjector p;  p.settings().raySampling = 0.2f;
//   p.configure(pmats, viewDim);            // FullGeometry as [views][modules][3][4] doubles
//   ProjectionData d = p.project(volume);   // VoxelVolume<float>, x fastest
// Errors follow the reference: an invalid device throws std::runtime_error (as the reference's
// constructor and prepareOpenCLDeviceList do); a volume without data returns ProjectionData with 0
// views; other failures throw std::runtime_error carrying ctlproj_last_error().
#ifndef CTLPROJ_HPP
#define CTLPROJ_HPP

#include "ctlproj.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <functional>
#include <memory>
#include <stdexcept>
#include <map>
#include <string>
#include <vector>

namespace ctl_amd {

inline void check(ctlproj_status st)
{
    if (st != CTLPROJ_OK)
        throw std::runtime_error(std::string("ctlproj error ") + std::to_string(int(st)) + ": " +
                                 ctlproj_last_error());
}

// VoxelVolume<float> (img/voxelvolume.h): dims X, Y, Z; data x fastest; voxel size and offset in mm.
struct VoxelVolume {
    std::array<uint32_t, 3> dims{ { 0, 0, 0 } };
    std::array<float, 3> voxelSize{ { 1.0f, 1.0f, 1.0f } };
    std::array<float, 3> offset{ { 0.0f, 0.0f, 0.0f } };
    std::vector<float> data;
    bool hasData() const
    {
        return !data.empty() && data.size() == size_t(dims[0]) * dims[1] * dims[2];
    }
};

// SingleViewData::Dimensions (img/singleviewdata.h).
struct ViewDimensions {
    uint32_t nbChannels = 0, nbRows = 0, nbModules = 0;
};

// ProjectionData (img/projectiondata.h): views -> modules -> rows -> channels, channel fastest.
struct ProjectionData {
    ViewDimensions viewDim;
    uint32_t nbViews = 0;
    std::vector<float> data;
    float& operator()(uint32_t view, uint32_t module, uint32_t x, uint32_t y)
    {
        return data[((size_t(view) * viewDim.nbModules + module) * viewDim.nbRows + y) *
                        viewDim.nbChannels + x];
    }
};

// One context per device (the replacement of OpenCLConfig::instance()).
class Context {
public:
    explicit Context(int device = 0) { check(ctlproj_context_create(device, &h_)); }
    ~Context() { ctlproj_context_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ctlproj_context* get() const { return h_; }

private:
    ctlproj_context* h_ = nullptr;
};

// A volume kept resident in HBM across projections.
class ResidentVolume {
public:
    ResidentVolume(Context& ctx, const VoxelVolume& v, ctlproj_dtype storage = CTLPROJ_F32)
    {
        check(ctlproj_volume_create(ctx.get(), v.data.data(), v.dims.data(), v.voxelSize.data(),
                                    v.offset.data(), storage, &h_));
    }
    ~ResidentVolume() { ctlproj_volume_destroy(h_); }
    ResidentVolume(const ResidentVolume&) = delete;
    ResidentVolume& operator=(const ResidentVolume&) = delete;
    const ctlproj_volume* get() const { return h_; }

private:
    ctlproj_volume* h_ = nullptr;
};

class RayCasterProjector {
public:
    // RayCasterProjector::Settings (raycasterprojector.h:27-37)
    struct Settings {
        uint32_t raysPerPixel[2] = { 1, 1 };
        float raySampling = 0.3f;
        uint32_t volumeUpSampling = 1;
        bool interpolate = true;

        // Settings::optimizedFor (raycasterprojector.cpp:482-521)
        static Settings optimizedFor(const VoxelVolume& volume, double pixelWidthMm, double pixelHeightMm)
        {
            ctlproj_settings c;
            check(ctlproj_settings_optimized_for(volume.voxelSize.data(), pixelWidthMm, pixelHeightMm, &c));
            Settings s;
            s.raysPerPixel[0] = c.rays_per_pixel[0];
            s.raysPerPixel[1] = c.rays_per_pixel[1];
            s.raySampling = c.ray_sampling;
            s.volumeUpSampling = c.volume_upsampling;
            s.interpolate = c.interpolate != 0;
            return s;
        }
    };

    // parameter() / setParameter() (raycasterprojector.cpp:69-91): the reference's keys
    std::map<std::string, double> parameter() const
    {
        return { { "Rays per pixel X", settings_.raysPerPixel[0] },
                 { "Rays per pixel Y", settings_.raysPerPixel[1] },
                 { "Ray sampling step length", settings_.raySampling },
                 { "Volume upsampling factor", settings_.volumeUpSampling },
                 { "Interpolate", settings_.interpolate ? 1.0 : 0.0 } };
    }
    void setParameter(const std::map<std::string, double>& p)
    {
        auto get = [&](const char* k, double d) { auto it = p.find(k); return it == p.end() ? d : it->second; };
        settings_.raysPerPixel[0] = uint32_t(get("Rays per pixel X", 1));
        settings_.raysPerPixel[1] = uint32_t(get("Rays per pixel Y", 1));
        settings_.raySampling = float(get("Ray sampling step length", 0.3));
        settings_.volumeUpSampling = uint32_t(get("Volume upsampling factor", 1));
        settings_.interpolate = get("Interpolate", 1.0) != 0.0;
    }

    explicit RayCasterProjector(std::shared_ptr<Context> ctx = std::make_shared<Context>(0))
        : ctx_(std::move(ctx))
    {
        ctlproj_settings s;
        ctlproj_settings_default(&s);
        check(ctlproj_projector_create(ctx_->get(), &s, &h_));
    }
    ~RayCasterProjector() { ctlproj_projector_destroy(h_); }
    RayCasterProjector(const RayCasterProjector&) = delete;
    RayCasterProjector& operator=(const RayCasterProjector&) = delete;

    Settings& settings() { return settings_; }
    bool isLinear() const { return true; }

    // ProjectorNotifier::projectionFinished(int) (abstractprojector.h:57-63)
    std::function<void(int)> projectionFinished;

    // configure (raycasterprojector.cpp:97-107) from an encoded FullGeometry
    void configure(const std::vector<double>& pmats, const ViewDimensions& viewDim)
    {
        const size_t per_view = size_t(viewDim.nbModules) * 12;
        if (per_view == 0 || pmats.size() % per_view != 0)
            throw std::domain_error("FullGeometry size does not match the view dimensions");
        viewDim_ = viewDim;
        nbViews_ = uint32_t(pmats.size() / per_view);
        check(ctlproj_projector_configure(h_, pmats.data(), nbViews_, viewDim.nbModules,
                                          viewDim.nbChannels, viewDim.nbRows));
    }

    // project (raycasterprojector.cpp:139-375): uploads the volume, as the reference does
    ProjectionData project(const VoxelVolume& volume)
    {
        ProjectionData ret;
        ret.viewDim = viewDim_;
        if (!volume.hasData()) return ret; // the reference returns 0 views (.cpp:156-160)
        pushSettings();
        ret.nbViews = nbViews_;
        ret.data.resize(size_t(nbViews_) * viewDim_.nbModules * viewDim_.nbRows * viewDim_.nbChannels);
        check(ctlproj_project_host_volume(h_, volume.data.data(), volume.dims.data(),
                                          volume.voxelSize.data(), volume.offset.data(),
                                          ret.data.data(), ret.data.size()));
        if (projectionFinished)
            for (uint32_t v = 0; v < nbViews_; ++v) projectionFinished(int(v));
        return ret;
    }

    // project from a volume that stays resident on the device
    ProjectionData project(const ResidentVolume& volume)
    {
        pushSettings();
        ProjectionData ret;
        ret.viewDim = viewDim_;
        ret.nbViews = nbViews_;
        ret.data.resize(size_t(nbViews_) * viewDim_.nbModules * viewDim_.nbRows * viewDim_.nbChannels);
        check(ctlproj_project(h_, volume.get(), ret.data.data(), ret.data.size(), &progressThunk, this));
        return ret;
    }

    // projectComposite (abstractprojector.h:99-112): the sum of the sub-volume projections
    ProjectionData projectComposite(const std::vector<VoxelVolume>& volumes)
    {
        ProjectionData sum;
        for (const auto& v : volumes) {
            ProjectionData p = project(v);
            if (sum.data.empty()) sum = std::move(p);
            else
                for (size_t i = 0; i < sum.data.size() && i < p.data.size(); ++i) sum.data[i] += p.data[i];
        }
        return sum;
    }

    // the C handle with the current settings applied, for the extensions below
    ctlproj_projector* prepared()
    {
        pushSettings();
        return h_;
    }
    ViewDimensions viewDimensions() const { return viewDim_; }
    uint32_t nbViews() const { return nbViews_; }
    Context& context() const { return *ctx_; }

private:
    static void progressThunk(int view, void* self)
    {
        auto* p = static_cast<RayCasterProjector*>(self);
        if (p->projectionFinished) p->projectionFinished(view);
    }

    void pushSettings()
    {
        ctlproj_settings s;
        ctlproj_settings_default(&s);
        s.rays_per_pixel[0] = settings_.raysPerPixel[0];
        s.rays_per_pixel[1] = settings_.raysPerPixel[1];
        s.ray_sampling = settings_.raySampling;
        s.volume_upsampling = settings_.volumeUpSampling;
        s.interpolate = settings_.interpolate ? 1 : 0;
        check(ctlproj_projector_set_settings(h_, &s));
    }

    std::shared_ptr<Context> ctx_;
    ctlproj_projector* h_ = nullptr;
    Settings settings_;
    ViewDimensions viewDim_;
    uint32_t nbViews_ = 0;
};

// SpectralEffectsExtension (projectors/spectraleffectsextension.cpp:232-352) with the ray caster as
// nested projector: one projection per material density and the sum over energy bins on the
// device. The spectrum (per-bin, per-view intensities; I0 per view) and each material's mass
// attenuation coefficient per bin (cm^2/g) are inputs; the detector response is optional.
class SpectralEffectsExtension {
public:
    explicit SpectralEffectsExtension(RayCasterProjector& nested) : nested_(nested) {}
    // intensities: [bin][view]; totalIntensity: one value per view (empty: the sum over bins)
    void setSpectralInformation(uint32_t nbBins, std::vector<double> intensities, std::vector<double> totalIntensity = {})
    {
        bins_ = nbBins;
        inten_ = std::move(intensities);
        total_ = std::move(totalIntensity);
    }
    void setDetectorResponse(std::vector<float> response) { resp_ = std::move(response); }
    bool isLinear() const { return false; }
    // volumes[k] is a density volume, massAttenuation[k] its coefficients, one per bin
    ProjectionData projectComposite(const std::vector<const ResidentVolume*>& volumes,
                                    const std::vector<std::vector<float>>& massAttenuation)
    {
        const uint32_t nm = uint32_t(volumes.size());
        const uint32_t V = nested_.nbViews();
        if (massAttenuation.size() != nm) throw std::domain_error("one coefficient list per material");
        if (inten_.size() != size_t(bins_) * V) throw std::domain_error("spectral information does not match the views");
        std::vector<float> mu(size_t(bins_) * nm);
        for (uint32_t k = 0; k < nm; ++k) {
            if (massAttenuation[k].size() != bins_) throw std::domain_error("one coefficient per energy bin");
            for (uint32_t b = 0; b < bins_; ++b) mu[size_t(b) * nm + k] = massAttenuation[k][b] * 0.1f; // 1/cm -> 1/mm
        }
        std::vector<double> total = total_;
        if (total.empty()) {
            total.assign(V, 0.0);
            for (uint32_t b = 0; b < bins_; ++b)
                for (uint32_t v = 0; v < V; ++v) total[v] += inten_[size_t(b) * V + v];
        }
        std::vector<const ctlproj_volume*> h(nm);
        for (uint32_t k = 0; k < nm; ++k) h[k] = volumes[k]->get();
        ProjectionData ret;
        ret.viewDim = nested_.viewDimensions();
        ret.nbViews = V;
        ret.data.resize(size_t(V) * ret.viewDim.nbModules * ret.viewDim.nbRows * ret.viewDim.nbChannels);
        check(ctlproj_project_spectral(nested_.prepared(), h.data(), nm, mu.data(), inten_.data(),
                                       resp_.empty() ? nullptr : resp_.data(), bins_, total.data(), ret.data.data(),
                                       ret.data.size()));
        return ret;
    }

private:
    RayCasterProjector& nested_;
    uint32_t bins_ = 0;
    std::vector<double> inten_, total_;
    std::vector<float> resp_;
};

// DynamicProjectorExtension (projectors/dynamicprojectorextension.cpp:39-66) for a
// LinearDynamicVolume (img/lineardynamicvolume.cpp:194-198: value = slope * t + offset): view v
// sees the volume at timeStamps[v].
class DynamicProjectorExtension {
public:
    explicit DynamicProjectorExtension(RayCasterProjector& nested) : nested_(nested) {}
    ProjectionData project(const ResidentVolume& slope, const ResidentVolume& offset,
                           const std::vector<double>& timeStamps)
    {
        const uint32_t V = nested_.nbViews();
        if (timeStamps.size() != V) throw std::domain_error("one time stamp per view");
        ProjectionData ret;
        ret.viewDim = nested_.viewDimensions();
        ret.nbViews = V;
        ret.data.resize(size_t(V) * ret.viewDim.nbModules * ret.viewDim.nbRows * ret.viewDim.nbChannels);
        check(ctlproj_project_linear_dynamic(nested_.prepared(), slope.get(), offset.get(), timeStamps.data(),
                                             ret.data.data(), ret.data.size()));
        return ret;
    }

private:
    RayCasterProjector& nested_;
};

// Chunk2D<float> (img/chunk2d.h): width x height, x fastest.
struct Chunk2D {
    uint32_t width = 0, height = 0;
    std::vector<float> data;
    float operator()(uint32_t x, uint32_t y) const { return data[size_t(y) * width + x]; }
};

// Unit normal from polar coordinates, as volumeslicer.cpp:172-176 and radontransform3d.cpp:226-230.
inline std::array<double, 3> planeNormal(double azimuth, double polar)
{
    return { { std::sin(polar) * std::cos(azimuth), std::sin(polar) * std::sin(azimuth), std::cos(polar) } };
}

// OCL::VolumeSlicer (processing/volumeslicer.h:37-60) on a resident volume.
class VolumeSlicer {
public:
    VolumeSlicer(Context& ctx, const VoxelVolume& volume, ctlproj_dtype storage = CTLPROJ_F32)
        : vol_(ctx, volume, storage)
    {
        uint32_t side = 0;
        check(ctlproj_slicer_default_side(volume.dims.data(), &side));
        width_ = height_ = side;
        reso_ = std::min(volume.voxelSize[0], std::min(volume.voxelSize[1], volume.voxelSize[2]));
    }
    void setSliceDimensions(uint32_t width, uint32_t height) { width_ = width; height_ = height; }
    void setSliceResolution(float pixelResolution) { reso_ = pixelResolution; }
    float sliceResolution() const { return reso_; }
    Chunk2D slice(const std::array<double, 3>& planeUnitNormal, double planeDistanceFromOrigin) const
    {
        Chunk2D ret;
        ret.width = width_;
        ret.height = height_;
        ret.data.resize(size_t(width_) * height_);
        check(ctlproj_slice(vol_.get(), planeUnitNormal.data(), &planeDistanceFromOrigin, 1, width_, height_, reso_,
                            ret.data.data(), ret.data.size()));
        return ret;
    }
    Chunk2D slice(double azimuth, double polar, double distance) const
    {
        return slice(planeNormal(azimuth, polar), distance);
    }

private:
    ResidentVolume vol_;
    uint32_t width_ = 0, height_ = 0;
    float reso_ = 1.0f;
};

// OCL::VolumeResampler (processing/volumeresampler.h:11-56) on a resident volume. Ranges are
// {lo1, hi1, lo2, hi2, lo3, hi3} (SamplingRange per axis).
class VolumeResampler {
public:
    VolumeResampler(Context& ctx, const VoxelVolume& volume, ctlproj_dtype storage = CTLPROJ_F32)
        : vol_(ctx, volume, storage)
    {
        check(ctlproj_resampler_default_ranges(vol_.get(), ranges_.data()));
    }
    void setSamplingRanges(const std::array<float, 6>& ranges) { ranges_ = ranges; }
    const std::array<float, 6>& ranges() const { return ranges_; }
    // resample: out[i + j*n1 + k*n1*n2] at (p1[i], p2[j], p3[k])
    std::vector<float> resample(const std::vector<float>& p1, const std::vector<float>& p2,
                                const std::vector<float>& p3) const
    {
        std::vector<float> out(p1.size() * p2.size() * p3.size());
        check(ctlproj_resample(vol_.get(), ranges_.data(), p1.data(), uint32_t(p1.size()), p2.data(),
                               uint32_t(p2.size()), p3.data(), uint32_t(p3.size()), out.data(), out.size()));
        return out;
    }
    // sample: xyz holds 3 floats per point (Generic3DCoord)
    std::vector<float> sample(const std::vector<float>& xyz) const
    {
        std::vector<float> out(xyz.size() / 3);
        check(ctlproj_sample(vol_.get(), ranges_.data(), xyz.data(), out.size(), out.data()));
        return out;
    }

private:
    ResidentVolume vol_;
    std::array<float, 6> ranges_{};
};

// OCL::RadonTransform3D (processing/radontransform3d.h:92-173) on a resident volume.
class RadonTransform3D {
public:
    RadonTransform3D(Context& ctx, const VoxelVolume& volume, ctlproj_dtype storage = CTLPROJ_F32)
        : vol_(ctx, volume, storage)
    {
        check(ctlproj_radon3d_default_slice(volume.dims.data(), volume.voxelSize.data(), &side_, &reso_));
    }
    void setSliceResolution(float pixelResolution)
    {
        check(ctlproj_radon3d_side_for_resolution(side_, reso_, pixelResolution, &side_));
        reso_ = pixelResolution;
    }
    uint32_t sliceSide() const { return side_; }
    float sliceResolution() const { return reso_; }
    float planeIntegral(const std::array<double, 3>& planeUnitNormal, double planeDistanceFromOrigin) const
    {
        const float d = float(planeDistanceFromOrigin);
        float r = 0.0f;
        check(ctlproj_radon3d_plane_integrals(vol_.get(), planeUnitNormal.data(), 1, &d, 1, side_, reso_, &r, 1));
        return r;
    }
    float planeIntegral(double azimuth, double polar, double distance) const
    {
        return planeIntegral(planeNormal(azimuth, polar), distance);
    }
    // sampleTransform(azimuths, polars, distances): out[a + nAzi*(p + nPol*d)]
    std::vector<float> sampleTransform(const std::vector<float>& azimuth, const std::vector<float>& polar,
                                       const std::vector<float>& distance) const
    {
        std::vector<float> out(azimuth.size() * polar.size() * distance.size());
        check(ctlproj_radon3d_sample_transform(vol_.get(), azimuth.data(), uint32_t(azimuth.size()), polar.data(),
                                               uint32_t(polar.size()), distance.data(), uint32_t(distance.size()),
                                               side_, reso_, out.data(), out.size()));
        return out;
    }

private:
    ResidentVolume vol_;
    uint32_t side_ = 0;
    float reso_ = 1.0f;
};

} // namespace ctl_amd

#endif // CTLPROJ_HPP
