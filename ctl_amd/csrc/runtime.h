// Internal objects behind the C ABI handles.
#pragma once

#include "ctlproj.h"
#include "errors.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

struct ctlproj_context {
    int device = 0;
    hipStream_t stream = nullptr; // uploads and volume conversions
    std::atomic<int> inject{ 0 }; // failure injection (tests of error paths)
    std::mutex mu;
};

struct ctlproj_volume {
    ctlproj_context* ctx = nullptr;
    void* d_data = nullptr;
    size_t bytes = 0;
    uint32_t* d_occ = nullptr; // occupancy table of view.occ (NULL: none)
    ctlk::VolumeView view{};
    uint32_t n[3] = { 0, 0, 0 };
    float voxel_mm[3] = { 0, 0, 0 };
    float offset_mm[3] = { 0, 0, 0 };
    // FP32 brick copy for the plane integrals and its kind, made on first use (guarded by ctx->mu)
    mutable float* d_brick = nullptr;
    mutable ctlk::BrickKind brick_kind = ctlk::BrickKind::Auto;
    // largest |voxel| of the brick copy (negative: not measured yet; NaN: not finite)
    mutable float brick_maxabs = -1.0f;
};

struct ctlproj_projector {
    ctlproj_context* ctx = nullptr;
    ctlproj_settings s{};
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;
    bool configured = false;
    uint32_t V = 0, M = 0, C = 0, R = 0;
    std::vector<double> pmats;
    double* d_pmats = nullptr;
    size_t d_pmats_cap = 0;
    ctlk::RayGeom* d_geom = nullptr;
    size_t geom_cap = 0;
    float* d_out = nullptr; // device staging of the host-output path
    size_t out_cap = 0;
    float* h_stage[2] = { nullptr, nullptr }; // pinned host staging
    size_t stage_cap = 0;
    unsigned long long* d_count = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timed = false;
};

namespace ctlrt {

// thread-local error message helpers (fail, clearError, lastError: errors.h)
ctlproj_status hipFail(hipError_t e, const char* what);

// checks the failure-injection counter of the context
bool injected(ctlproj_context* ctx);

// Effective (possibly upsampled) volume for a projection: either the resident volume itself or
// a temporary device copy owned by `tmp`.
struct EffectiveVolume {
    ctlk::VolumeView view{};
    ctlk::VolumeGeom geom{};
    void* tmp = nullptr;
    ~EffectiveVolume();
};

ctlproj_status effectiveVolume(const ctlproj_volume* vol, const ctlproj_settings& s,
                               hipStream_t stream, EffectiveVolume& ev);
// True for a volume of 4 GiB or more: only the projections (ctlproj_project*) read those.
bool largeVolume(const ctlproj_volume* vol);

// Device allocation released on scope exit (callers synchronise their stream before that).
struct DeviceBuffer {
    void* p = nullptr;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer()
    {
        if (p) (void)hipFree(p);
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

inline ctlproj_status deviceAlloc(DeviceBuffer& b, size_t bytes, const char* what)
{
    hipError_t e = hipMalloc(&b.p, bytes ? bytes : 1);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, what);
}

// Allocates b and copies `bytes` from host memory on stream s.
inline ctlproj_status deviceUpload(DeviceBuffer& b, const void* host, size_t bytes, hipStream_t s, const char* what)
{
    ctlproj_status st = deviceAlloc(b, bytes, what);
    if (st != CTLPROJ_OK || bytes == 0) return st;
    hipError_t e = hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, s);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, what);
}

ctlproj_status validateSettings(const ctlproj_settings& s);
// True if projections with these settings and this volume storage use the LDS-staged kernel.
bool stagedWorkload(const ctlproj_settings& s, ctlk::Storage st, ctlk::Layout ly);
ctlk::LaunchShape shapeFor(const ctlproj_settings& s, ctlk::Storage st, ctlk::Layout ly,
                           uint32_t channels, uint32_t rows);

} // namespace ctlrt
