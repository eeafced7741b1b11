// Seeded random-ellipsoid phantom generator (host side).
#pragma once

#include <cstdint>

namespace ctlphantom {

struct PhantomSpec {
    uint32_t dims[3];
    float voxel_mm[3];
    uint32_t n_ellipsoids;
    uint64_t seed;
    float semi_axis_min_mm, semi_axis_max_mm;
    float mu_min, mu_max;
    float background_mu;
};

// Writes dims[0]*dims[1]*dims[2] floats, x fastest.
void ellipsoids(const PhantomSpec& spec, float* out);

} // namespace ctlphantom
