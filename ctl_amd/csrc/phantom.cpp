// Seeded random-ellipsoid phantoms for tests and benchmarks (host side, multithreaded).
//
// Generator (fixed so that every machine draws the same phantom from the same seed):
//   RNG: std::mt19937_64(seed); uniform U[0,1) = (x >> 11) * 2^-53; Gaussians by Box-Muller.
//   Per ellipsoid: centre uniform in the inner 70% of each axis extent, semi-axes U[a_min, a_max]
//   mm, rotation from a normalised Gaussian quaternion, additive attenuation U[mu_min, mu_max].
//   Background: an axis-aligned ellipsoid with semi-axes 0.45 x each extent and attenuation
//   `background_mu`. Voxel value = sum of the attenuations of the ellipsoids that contain the
//   voxel centre. The volume is centred on the origin (zero offset).
#include "phantom.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <thread>
#include <vector>

namespace ctlphantom {

namespace {
struct Ellipsoid {
    double c[3];
    double R[9]; // rows: ellipsoid axes in world coordinates
    double inv_a2[3];
    double mu;
    double bbox_lo[3], bbox_hi[3];
};

struct Rng {
    std::mt19937_64 g;
    explicit Rng(uint64_t seed) : g(seed) {}
    double uniform() { return double(g() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double gauss()
    {
        double u1 = uniform();
        while (u1 <= 0.0) u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
};

void quatToRot(double w, double x, double y, double z, double* R)
{
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z);     R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z);     R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y);     R[7] = 2 * (y * z + w * x);     R[8] = 1 - 2 * (x * x + y * y);
}

void finishBox(Ellipsoid& e, const double a[3])
{
    // axis-aligned bounding box of a rotated ellipsoid: half-width_k = sqrt(sum_j (R_jk a_j)^2)
    for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += (e.R[3 * j + k] * a[j]) * (e.R[3 * j + k] * a[j]);
        const double h = std::sqrt(s);
        e.bbox_lo[k] = e.c[k] - h;
        e.bbox_hi[k] = e.c[k] + h;
        e.inv_a2[k] = 1.0 / (a[k] * a[k]);
    }
}
} // namespace

void ellipsoids(const PhantomSpec& s, float* out)
{
    const uint32_t n[3] = { s.dims[0], s.dims[1], s.dims[2] };
    double ext[3];
    for (int k = 0; k < 3; ++k) ext[k] = double(n[k]) * s.voxel_mm[k];

    std::vector<Ellipsoid> el;
    Rng rng(s.seed);
    if (s.background_mu != 0.0f) {
        Ellipsoid b{};
        for (int k = 0; k < 3; ++k) b.c[k] = 0.0;
        quatToRot(1, 0, 0, 0, b.R);
        const double a[3] = { 0.45 * ext[0], 0.45 * ext[1], 0.45 * ext[2] };
        b.mu = s.background_mu;
        finishBox(b, a);
        el.push_back(b);
    }
    for (uint32_t i = 0; i < s.n_ellipsoids; ++i) {
        Ellipsoid e{};
        for (int k = 0; k < 3; ++k) e.c[k] = rng.uniform(-0.35 * ext[k], 0.35 * ext[k]);
        double a[3];
        for (int k = 0; k < 3; ++k) a[k] = rng.uniform(s.semi_axis_min_mm, s.semi_axis_max_mm);
        double q[4];
        double qn = 0.0;
        for (double& x : q) { x = rng.gauss(); qn += x * x; }
        qn = std::sqrt(qn);
        quatToRot(q[0] / qn, q[1] / qn, q[2] / qn, q[3] / qn, e.R);
        e.mu = rng.uniform(s.mu_min, s.mu_max);
        finishBox(e, a);
        el.push_back(e);
    }

    const size_t slice = size_t(n[0]) * n[1];
    auto work = [&](uint32_t z0, uint32_t z1) {
        std::vector<float> row(n[0]);
        for (uint32_t z = z0; z < z1; ++z) {
            const double pz = (z + 0.5) * s.voxel_mm[2] - 0.5 * ext[2];
            for (uint32_t y = 0; y < n[1]; ++y) {
                const double py = (y + 0.5) * s.voxel_mm[1] - 0.5 * ext[1];
                std::fill(row.begin(), row.end(), 0.0f);
                for (const Ellipsoid& e : el) {
                    if (pz < e.bbox_lo[2] || pz > e.bbox_hi[2] || py < e.bbox_lo[1] || py > e.bbox_hi[1])
                        continue;
                    const double xlo = (e.bbox_lo[0] + 0.5 * ext[0]) / s.voxel_mm[0] - 0.5;
                    const double xhi = (e.bbox_hi[0] + 0.5 * ext[0]) / s.voxel_mm[0] - 0.5;
                    const int32_t x0 = std::max<int32_t>(0, int32_t(std::floor(xlo)));
                    const int32_t x1 = std::min<int32_t>(int32_t(n[0]) - 1, int32_t(std::ceil(xhi)));
                    const double dy = py - e.c[1], dz = pz - e.c[2];
                    for (int32_t x = x0; x <= x1; ++x) {
                        const double dx = (x + 0.5) * s.voxel_mm[0] - 0.5 * ext[0] - e.c[0];
                        double r = 0.0;
                        for (int j = 0; j < 3; ++j) {
                            const double t = e.R[3 * j] * dx + e.R[3 * j + 1] * dy + e.R[3 * j + 2] * dz;
                            r += t * t * e.inv_a2[j];
                        }
                        if (r <= 1.0) row[x] += float(e.mu);
                    }
                }
                std::copy(row.begin(), row.end(), out + z * slice + size_t(y) * n[0]);
            }
        }
    };
    const uint32_t nt = std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), 32u));
    std::vector<std::thread> th;
    const uint32_t chunk = (n[2] + nt - 1) / nt;
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t z0 = t * chunk, z1 = std::min(n[2], z0 + chunk);
        if (z0 < z1) th.emplace_back(work, z0, z1);
    }
    for (auto& t : th) t.join();
}

} // namespace ctlphantom
