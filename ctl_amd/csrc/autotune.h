// Autotuner for the ray-casting launch shape (wavefront pixel tile and ray-step unroll).
#pragma once

#include "kernels.h"

#include <functional>
#include <string>
#include <vector>

namespace ctltune {

struct Choice {
    ctlk::LaunchShape shape;
    float ms = 0.0f;
};

// Candidate shapes. Gathers from HBM: wavefront tiles 8x8, 16x4, 4x16 (pixels along channels x
// rows), unroll 2, 4 and a cap of 0 (none) or 4 resident workgroups per CU. LDS-staged workloads
// (`staged`): tiles 8x8, 16x4 with 48- or 64-sample chunks in LDS boxes of 2560, 3072 or 6400
// texels (unroll 4), and the default unstaged shape.
// With `channels` > 0 (single projections), each candidate that narrowModuleShape changes for
// modules of that width is followed by its stacked form.
std::vector<ctlk::LaunchShape> candidates(bool staged, uint32_t channels = 0);

// Times every candidate with `run` (returns milliseconds, or a negative value on failure), keeps
// the fastest and records it in the process-wide table under `key`.
bool tune(const std::string& key, const std::function<float(const ctlk::LaunchShape&)>& run,
          Choice* best, bool staged = false, uint32_t channels = 0);

// Shape recorded for `key` by an earlier tune(); false if none.
bool lookup(const std::string& key, ctlk::LaunchShape* shape);

// Key of a workload class: storage, layout, interpolation, detector size class and whether the
// LDS-staged kernel applies.
std::string key(ctlk::Storage st, ctlk::Layout ly, int interpolate, uint32_t channels, uint32_t rows,
                bool staged);

// The whole table as JSON text.
std::string exportTable();

} // namespace ctltune
