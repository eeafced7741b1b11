// Internal objects behind the C ABI handles.
#pragma once

#include "ctlproj.h"
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
    ctlk::VolumeView view{};
    uint32_t n[3] = { 0, 0, 0 };
    float voxel_mm[3] = { 0, 0, 0 };
    float offset_mm[3] = { 0, 0, 0 };
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

// thread-local error message helpers
ctlproj_status fail(ctlproj_status st, const std::string& msg);
ctlproj_status hipFail(hipError_t e, const char* what);
void clearError();
const char* lastError();

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

ctlproj_status validateSettings(const ctlproj_settings& s);
// True if projections with these settings and this volume storage use the LDS-staged kernel.
bool stagedWorkload(const ctlproj_settings& s, ctlk::Storage st, ctlk::Layout ly);
ctlk::LaunchShape shapeFor(const ctlproj_settings& s, ctlk::Storage st, ctlk::Layout ly,
                           uint32_t channels, uint32_t rows);

} // namespace ctlrt
