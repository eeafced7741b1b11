// Volume processing on resident volumes behind the C ABI: VolumeSlicer, VolumeResampler and
// RadonTransform3D (modules/src/processing/volumeslicer.cpp, volumeresampler.cpp,
// radontransform3d.cpp). The plane geometry is computed in FP64 on the host and cast to FP32,
// as the reference's host code does; the sampling runs in processing.hip.
#include "ctlproj.h"
#include "runtime.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace ctlrt;

namespace {

constexpr uint32_t kPatch = 16; // radontransform3d.cpp:6

ctlproj_status alloc(DeviceBuffer& b, size_t bytes, const char* what) { return deviceAlloc(b, bytes, what); }

ctlproj_status upload(DeviceBuffer& b, const void* host, size_t bytes, hipStream_t s, const char* what)
{
    return deviceUpload(b, host, bytes, s, what);
}

ctlproj_status begin(const ctlproj_volume* vol)
{
    if (!vol || !vol->ctx) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume is NULL");
    hipError_t e = hipSetDevice(vol->ctx->device);
    if (e != hipSuccess) return hipFail(e, "hipSetDevice");
    if (injected(vol->ctx)) return CTLPROJ_ERR_INJECTED;
    if (largeVolume(vol))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "volume processing reads volumes below 4 GiB only");
    return CTLPROJ_OK;
}

ctlproj_status finish(hipError_t e, hipStream_t s, const char* what)
{
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? CTLPROJ_OK : hipFail(e, what);
}

// Unit plane normals (the reference asserts qFuzzyCompare(norm, 1)).
ctlproj_status checkNormals(const double* n, uint32_t count)
{
    if (count && !n) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "normals is NULL");
    for (uint32_t i = 0; i < count; ++i) {
        const double* v = n + 3 * size_t(i);
        if (!(std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2])))
            return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "plane normal is not finite");
        const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (std::fabs(len - 1.0) > 1e-6) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "plane normal is not a unit vector");
    }
    return CTLPROJ_OK;
}

bool validResolution(float r) { return std::isfinite(r) && r > 0.0f; }

// Rotation taking the xy-plane to the plane with unit normal n (volumeslicer.cpp:207-217,
// radontransform3d.cpp:425-438): third column n; second column n x e_a for the axis a of n's
// smallest component (ties go to the later axis), normalised; first column (second x third).
void planeRotation(const double n[3], double R[3][3])
{
    int a = std::fabs(n[0]) < std::fabs(n[1]) ? 0 : 1;
    if (!(std::fabs(n[a]) < std::fabs(n[2]))) a = 2;
    double c2[3] = { 0.0, 0.0, 0.0 };
    // n x e_a
    const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
    c2[a1] = n[a2];
    c2[a2] = -n[a1];
    const double len = std::sqrt(c2[0] * c2[0] + c2[1] * c2[1] + c2[2] * c2[2]);
    for (double& v : c2) v /= len;
    const double c1[3] = { c2[1] * n[2] - c2[2] * n[1], c2[2] * n[0] - c2[0] * n[2], c2[0] * n[1] - c2[1] * n[0] };
    for (int i = 0; i < 3; ++i) {
        R[i][0] = c1[i];
        R[i][1] = c2[i];
        R[i][2] = n[i];
    }
}

// ceil(1.4142 * largest dimension) in float, as volumeslicer.cpp:226-227 and
// radontransform3d.cpp:281-283 compute it.
uint32_t sqrt2Side(const uint32_t d[3])
{
    const uint32_t m = std::max(d[0], std::max(d[1], d[2]));
    return uint32_t(std::ceil(1.4142f * float(m)));
}

// RadonTransform3D::nextMultipleOfN (radontransform3d.cpp:293-301).
uint32_t nextMultiple(uint32_t v, uint32_t n) { return v % n ? uint32_t(std::floor(double(v) / double(n)) + 1.0) * n : v; }

// Slicer plane maps (volumeslicer.cpp:118-130).
std::vector<ctlk::PlaneMap> slicerMaps(const ctlproj_volume* vol, const double* normals, const double* distances,
                                       uint32_t n, float resolution)
{
    std::vector<ctlk::PlaneMap> maps(n);
    float corner[3];
    for (int k = 0; k < 3; ++k) corner[k] = -0.5f * float(vol->n[k] - 1u) + vol->offset_mm[k] / vol->voxel_mm[k];
    for (uint32_t i = 0; i < n; ++i) {
        double R[3][3];
        planeRotation(normals + 3 * size_t(i), R);
        ctlk::PlaneMap& m = maps[i];
        for (int r = 0; r < 3; ++r) {
            const double inv = 1.0 / double(vol->voxel_mm[r]);
            for (int c = 0; c < 3; ++c) m.h[4 * r + c] = float(inv * (double(resolution) * R[r][c]));
            m.h[4 * r + 3] = float(inv * (R[r][2] * distances[i]));
            m.c[r] = corner[r];
        }
        m.pad = 0.0f;
    }
    return maps;
}

// Plane-integral maps (radontransform3d.cpp:306-311, 355-367, 440-450).
std::vector<ctlk::PlaneMap> radonMaps(const ctlproj_volume* vol, const double* normals, uint32_t n, uint32_t side,
                                      float resolution)
{
    std::vector<ctlk::PlaneMap> maps(n);
    double corner[3];
    for (int k = 0; k < 3; ++k) corner[k] = double(vol->offset_mm[k]) - 0.5 * double(vol->n[k]) * double(vol->voxel_mm[k]);
    const double start = -double(resolution) * 0.5 * double(side - 1u); // template plane corner, x and y
    for (uint32_t i = 0; i < n; ++i) {
        const double* nv = normals + 3 * size_t(i);
        double R[3][3];
        planeRotation(nv, R);
        ctlk::PlaneMap& m = maps[i];
        for (int r = 0; r < 3; ++r) {
            const double inv = 1.0 / double(vol->voxel_mm[r]);
            const double t = R[r][0] * start + R[r][1] * start - corner[r];
            for (int c = 0; c < 3; ++c) m.h[4 * r + c] = float(inv * (double(resolution) * R[r][c]));
            m.h[4 * r + 3] = float(inv * t);
            m.c[r] = float(nv[r]) / vol->voxel_mm[r];
        }
        m.pad = 0.0f;
    }
    return maps;
}

// Plane integrals of `maps` into d_out ([d][o], row stride n_orient) on the context stream.
ctlproj_status planeIntegrals(const ctlproj_volume* vol, const std::vector<ctlk::PlaneMap>& maps,
                              const float* distances, uint32_t n_dist, uint32_t side, float resolution, float* d_out)
{
    const uint32_t n_orient = uint32_t(maps.size());
    hipStream_t s = vol->ctx->stream;
    if (n_orient == 0 || n_dist == 0) return CTLPROJ_OK;
    if (side == 0) {
        hipError_t e = ctlk::launchZero(d_out, size_t(n_orient) * n_dist, s);
        return finish(e, s, "zero fill");
    }
    // the volume's brick copy (made once, kept with the volume); without it the kernel reads the
    // resident texels. CTLPROJ_PLANE_READS=texels keeps the texel reads; =pairs asks for the
    // Pair16 copy whatever the volume's size (a copy of the other kind is replaced).
    const float* brick = nullptr;
    const char* env = std::getenv("CTLPROJ_PLANE_READS");
    const ctlk::BrickKind want = (env && std::strcmp(env, "pairs") == 0) ? ctlk::BrickKind::Pair16 : ctlk::BrickKind::Auto;
    uint32_t nb[3];
    size_t brick_bytes = 0;
    ctlk::BrickKind kind = ctlk::BrickKind::Auto;
    if (!(env && std::strcmp(env, "texels") == 0) && ctlk::brickShape(vol->n, want, &kind, nb, &brick_bytes)) {
        if (vol->d_brick && vol->brick_kind != kind) {
            (void)hipFree(vol->d_brick);
            vol->d_brick = nullptr;
            vol->brick_maxabs = -1.0f;
        }
        if (!vol->d_brick) {
            void* p = nullptr;
            if (hipMalloc(&p, brick_bytes) == hipSuccess) {
                hipError_t e = ctlk::launchBrick(vol->view, kind, static_cast<float*>(p), s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) {
                    (void)hipFree(p);
                    return hipFail(e, "brick kernel");
                }
                vol->d_brick = static_cast<float*>(p);
                vol->brick_kind = kind;
            } else {
                (void)hipGetLastError(); // no room for the copy: read the resident texels
            }
        }
        brick = vol->d_brick;
    }
    DeviceBuffer d_maps, d_dist, d_part;
    ctlproj_status st = upload(d_maps, maps.data(), maps.size() * sizeof(ctlk::PlaneMap), s, "plane maps");
    if (st == CTLPROJ_OK) st = upload(d_dist, distances, size_t(n_dist) * sizeof(float), s, "distances");
    if (st != CTLPROJ_OK) return st;

    // By default the integrals are summed brick by brick from LDS (needs the brick copy and finite
    // voxels); CTLPROJ_PLANE_KERNEL=tiles keeps the tile kernel that gathers through the caches.
    const char* kenv = std::getenv("CTLPROJ_PLANE_KERNEL");
    if (brick && !(kenv && std::strcmp(kenv, "tiles") == 0)) {
        if (vol->brick_maxabs < 0.0f) {
            DeviceBuffer d_max;
            if ((st = alloc(d_max, sizeof(uint32_t), "brick maximum")) != CTLPROJ_OK) return st;
            uint32_t bits = 0;
            hipError_t e = ctlk::launchZero(d_max.as<float>(), 1, s); // zero bits
            if (e == hipSuccess) e = ctlk::launchBrickMaxAbs(brick, brick_bytes, d_max.as<uint32_t>(), s);
            if (e == hipSuccess) e = hipMemcpyAsync(&bits, d_max.p, sizeof(bits), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            if (e != hipSuccess) return hipFail(e, "brick maximum");
            float m;
            std::memcpy(&m, &bits, sizeof(m));
            vol->brick_maxabs = bits >= 0x7f800000u ? std::nanf("") : m;
        }
        // fixed-point units: every total (at most side^2 samples of at most the largest |voxel|)
        // stays below 2^62
        const double bound = double(vol->brick_maxabs) * double(side) * double(side) * 1.01 + 1.0;
        const int shift = std::isfinite(bound) ? std::min(62 - int(std::ceil(std::log2(bound))), 60) : -1;
        if (shift >= 16) {
            std::vector<ctlk::PlaneInv> inv(n_orient);
            for (uint32_t o = 0; o < n_orient; ++o) {
                const ctlk::PlaneMap& m = maps[o];
                double a[3][3]; // rows: axes; columns: x, y, distance
                for (int k = 0; k < 3; ++k) {
                    a[k][0] = m.h[4 * k];
                    a[k][1] = m.h[4 * k + 1];
                    a[k][2] = m.c[k];
                }
                const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                                   a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                                   a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
                ctlk::PlaneInv& r = inv[o];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        // inverse = adjugate / det: entry (i, j) is the cofactor of (j, i)
                        const int r0 = (j + 1) % 3, r1 = (j + 2) % 3, c0 = (i + 1) % 3, c1 = (i + 2) % 3;
                        r.m[i][j] = float((a[r0][c0] * a[r1][c1] - a[r0][c1] * a[r1][c0]) / det);
                    }
                r.pad[0] = r.pad[1] = r.pad[2] = 0.0f;
            }
            DeviceBuffer d_inv, d_acc;
            if ((st = upload(d_inv, inv.data(), inv.size() * sizeof(ctlk::PlaneInv), s, "plane inverses")) != CTLPROJ_OK)
                return st;
            const uint32_t per_launch = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(n_orient, (64ull << 20) / n_dist)));
            if ((st = alloc(d_acc, size_t(per_launch) * n_dist * 8, "plane totals")) != CTLPROJ_OK) return st;
            // orientations per workgroup: enough workgroups for the bricks to fill the device
            // orientations per workgroup: 16-1024 measured within 15% of each other on 256^3 x 1024
            // orientations, 64-128 best
            const uint32_t o_chunk = 128;
            for (uint32_t o0 = 0; o0 < n_orient; o0 += per_launch) {
                const uint32_t oc = std::min(per_launch, n_orient - o0);
                hipError_t e = ctlk::launchPlaneBricks(vol->view, brick, vol->brick_kind,
                                                       d_maps.as<ctlk::PlaneMap>() + o0, d_inv.as<ctlk::PlaneInv>() + o0,
                                                       d_dist.as<float>(), oc, n_dist, side, side, o_chunk, shift,
                                                       d_acc.as<unsigned long long>(), std::pow(resolution, 2.0f),
                                                       d_out + o0, n_orient, s);
                if (e != hipSuccess) return hipFail(e, "plane brick kernel");
            }
            return finish(hipSuccess, s, "plane brick kernel");
        }
    }

    ctlk::PlaneJob job{};
    job.n_dist = n_dist;
    job.width = job.height = side;
    job.tiles_x = job.tiles_y = (side + 7) / 8;
    job.n_tiles = ((job.tiles_x + 3) / 4) * ((job.tiles_y + 3) / 4) * 16;
    const uint64_t chunks = (uint64_t(n_dist) + ctlk::kPlaneDists - 1) / ctlk::kPlaneDists;
    // Orientations per launch: enough workgroups to fill the device several times over, each
    // split of the template plane at least 16 tiles, the partial sums within 64 Mi floats.
    const uint64_t target = 4096, max_splits = std::max<uint64_t>(1, job.n_tiles / 16);
    const uint64_t max_partial = 64ull << 20;
    const uint32_t per_launch = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(
        n_orient, std::min<uint64_t>(1u << 16, max_partial / n_dist))));
    size_t part_floats = 0;
    for (uint32_t o0 = 0; o0 < n_orient; o0 += per_launch) {
        const uint32_t oc = std::min(per_launch, n_orient - o0);
        const uint64_t splits = std::max<uint64_t>(1, std::min<uint64_t>(
            {max_splits, (target + oc * chunks - 1) / (oc * chunks), max_partial / (uint64_t(oc) * n_dist)}));
        job.n_orient = oc;
        job.splits = uint32_t(splits);
        const size_t need = size_t(oc) * n_dist * splits;
        if (need > part_floats) {
            if (d_part.p) (void)hipFree(d_part.p);
            d_part.p = nullptr;
            if ((st = alloc(d_part, need * sizeof(float), "plane partial sums")) != CTLPROJ_OK) return st;
            part_floats = need;
        }
        hipError_t e = ctlk::launchPlaneIntegrals(vol->view, brick, vol->brick_kind, d_maps.as<ctlk::PlaneMap>() + o0,
                                                  d_dist.as<float>(), job,
                                                  d_part.as<float>(), std::pow(resolution, 2.0f), d_out + o0,
                                                  n_orient, s);
        if (e != hipSuccess) return hipFail(e, "plane integral kernel");
    }
    return finish(hipSuccess, s, "plane integral kernel");
}

ctlproj_status checkRanges(const float* r)
{
    if (!r) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "ranges is NULL");
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(r[k])) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "sampling range is not finite");
    return CTLPROJ_OK;
}

} // namespace

ctlproj_status ctlproj_slicer_default_side(const uint32_t dims[3], uint32_t* side)
{
    if (!dims || !side) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *side = sqrt2Side(dims);
    return CTLPROJ_OK;
}

static ctlproj_status sliceImpl(const ctlproj_volume* vol, const double* normals, const double* distances,
                                uint32_t n_planes, uint32_t width, uint32_t height, float resolution, float* out,
                                size_t out_count, bool device)
{
    ctlproj_status st = begin(vol);
    if (st != CTLPROJ_OK) return st;
    if ((st = checkNormals(normals, n_planes)) != CTLPROJ_OK) return st;
    if (n_planes && !distances) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "distances is NULL");
    for (uint32_t i = 0; i < n_planes; ++i)
        if (!std::isfinite(distances[i])) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "plane distance is not finite");
    if (!validResolution(resolution)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "slice resolution must be positive");
    const size_t count = size_t(n_planes) * width * height;
    if (out_count != count) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output size does not match planes x height x width");
    if (count == 0) return CTLPROJ_OK;
    if (!out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output is NULL");
    std::lock_guard<std::mutex> lk(vol->ctx->mu);
    hipStream_t s = vol->ctx->stream;
    const std::vector<ctlk::PlaneMap> maps = slicerMaps(vol, normals, distances, n_planes, resolution);
    DeviceBuffer d_maps, d_out;
    if ((st = upload(d_maps, maps.data(), maps.size() * sizeof(ctlk::PlaneMap), s, "slice maps")) != CTLPROJ_OK)
        return st;
    float* dst = out;
    if (!device) {
        if ((st = alloc(d_out, count * sizeof(float), "slice output")) != CTLPROJ_OK) return st;
        dst = d_out.as<float>();
    }
    hipError_t e = ctlk::launchSlice(vol->view, d_maps.as<ctlk::PlaneMap>(), n_planes, width, height, dst, s);
    if (e == hipSuccess && !device) e = hipMemcpyAsync(out, dst, count * sizeof(float), hipMemcpyDeviceToHost, s);
    return finish(e, s, "slice kernel");
}

ctlproj_status ctlproj_slice(const ctlproj_volume* vol, const double* normals, const double* distances,
                             uint32_t n_planes, uint32_t width, uint32_t height, float resolution, float* out,
                             size_t out_count)
{
    return sliceImpl(vol, normals, distances, n_planes, width, height, resolution, out, out_count, false);
}

ctlproj_status ctlproj_slice_device(const ctlproj_volume* vol, const double* normals, const double* distances,
                                    uint32_t n_planes, uint32_t width, uint32_t height, float resolution,
                                    float* device_out, size_t out_count)
{
    return sliceImpl(vol, normals, distances, n_planes, width, height, resolution, device_out, out_count, true);
}

ctlproj_status ctlproj_resampler_default_ranges(const ctlproj_volume* vol, float ranges[6])
{
    if (!vol || !ranges) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int k = 0; k < 3; ++k) { // volumeresampler.cpp:74-83
        const float half = 0.5f * vol->voxel_mm[k] * float(vol->n[k] - 1u);
        ranges[2 * k] = vol->offset_mm[k] - half;
        ranges[2 * k + 1] = vol->offset_mm[k] + half;
    }
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_resample(const ctlproj_volume* vol, const float ranges[6], const float* p1, uint32_t n1,
                                const float* p2, uint32_t n2, const float* p3, uint32_t n3, float* out,
                                size_t out_count)
{
    ctlproj_status st = begin(vol);
    if (st != CTLPROJ_OK) return st;
    if ((st = checkRanges(ranges)) != CTLPROJ_OK) return st;
    const size_t count = size_t(n1) * n2 * n3;
    if (out_count != count) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output size does not match n1 x n2 x n3");
    if (count == 0) return CTLPROJ_OK;
    if (!p1 || !p2 || !p3 || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(vol->ctx->mu);
    hipStream_t s = vol->ctx->stream;
    DeviceBuffer d1, d2, d3, d_out;
    if ((st = upload(d1, p1, n1 * sizeof(float), s, "sampling points")) != CTLPROJ_OK ||
        (st = upload(d2, p2, n2 * sizeof(float), s, "sampling points")) != CTLPROJ_OK ||
        (st = upload(d3, p3, n3 * sizeof(float), s, "sampling points")) != CTLPROJ_OK ||
        (st = alloc(d_out, count * sizeof(float), "resample output")) != CTLPROJ_OK)
        return st;
    hipError_t e = ctlk::launchResample(vol->view, ranges, d1.as<float>(), n1, d2.as<float>(), n2, d3.as<float>(), n3,
                                        d_out.as<float>(), s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out.p, count * sizeof(float), hipMemcpyDeviceToHost, s);
    return finish(e, s, "resample kernel");
}

ctlproj_status ctlproj_sample(const ctlproj_volume* vol, const float ranges[6], const float* xyz, size_t n,
                              float* out)
{
    ctlproj_status st = begin(vol);
    if (st != CTLPROJ_OK) return st;
    if ((st = checkRanges(ranges)) != CTLPROJ_OK) return st;
    if (n == 0) return CTLPROJ_OK;
    if (!xyz || !out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(vol->ctx->mu);
    hipStream_t s = vol->ctx->stream;
    DeviceBuffer d_xyz, d_out;
    if ((st = upload(d_xyz, xyz, 3 * n * sizeof(float), s, "sampling points")) != CTLPROJ_OK ||
        (st = alloc(d_out, n * sizeof(float), "sample output")) != CTLPROJ_OK)
        return st;
    hipError_t e = ctlk::launchSamplePoints(vol->view, ranges, d_xyz.as<float>(), n, d_out.as<float>(), s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out.p, n * sizeof(float), hipMemcpyDeviceToHost, s);
    return finish(e, s, "sample kernel");
}

ctlproj_status ctlproj_sample_device(const ctlproj_volume* vol, const float ranges[6], const float* device_xyz,
                                     size_t n, float* device_out)
{
    ctlproj_status st = begin(vol);
    if (st != CTLPROJ_OK) return st;
    if ((st = checkRanges(ranges)) != CTLPROJ_OK) return st;
    if (n == 0) return CTLPROJ_OK;
    if (!device_xyz || !device_out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(vol->ctx->mu);
    hipStream_t s = vol->ctx->stream;
    // the caller's arrays may be written by work on other streams: order after the device
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = ctlk::launchSamplePoints(vol->view, ranges, device_xyz, n, device_out, s);
    return finish(e, s, "sample kernel");
}

ctlproj_status ctlproj_radon3d_default_slice(const uint32_t dims[3], const float voxel_mm[3], uint32_t* side,
                                             float* resolution)
{
    if (!dims || !voxel_mm || !side || !resolution) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    *side = nextMultiple(sqrt2Side(dims), kPatch);
    *resolution = std::min(voxel_mm[0], std::min(voxel_mm[1], voxel_mm[2]));
    return CTLPROJ_OK;
}

ctlproj_status ctlproj_radon3d_side_for_resolution(uint32_t side, float resolution, float new_resolution,
                                                   uint32_t* new_side)
{
    if (!new_side) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!validResolution(resolution) || !validResolution(new_resolution))
        return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "slice resolution must be positive");
    const float factor = resolution / new_resolution; // radontransform3d.cpp:239-251
    const float px = float(side) * factor;
    if (!(px < 4294967040.0f)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "slice too large");
    *new_side = nextMultiple(uint32_t(px), kPatch);
    return CTLPROJ_OK;
}

static ctlproj_status radonImpl(const ctlproj_volume* vol, const double* normals, uint32_t n_orient,
                                const float* distances, uint32_t n_dist, uint32_t side, float resolution, float* out,
                                size_t out_count, bool device)
{
    ctlproj_status st = begin(vol);
    if (st != CTLPROJ_OK) return st;
    if ((st = checkNormals(normals, n_orient)) != CTLPROJ_OK) return st;
    if (n_dist && !distances) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "distances is NULL");
    if (!validResolution(resolution)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "slice resolution must be positive");
    if (side > 65536) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "slice side must be at most 65536 pixels");
    const size_t count = size_t(n_orient) * n_dist;
    if (out_count != count) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output size does not match distances x orientations");
    if (count == 0) return CTLPROJ_OK;
    if (count > 0xffffff00ull) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "too many plane integrals for one call");
    if (!out) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "output is NULL");
    std::lock_guard<std::mutex> lk(vol->ctx->mu);
    hipStream_t s = vol->ctx->stream;
    const std::vector<ctlk::PlaneMap> maps = radonMaps(vol, normals, n_orient, side, resolution);
    DeviceBuffer d_out;
    float* dst = out;
    if (!device) {
        if ((st = alloc(d_out, count * sizeof(float), "plane integral output")) != CTLPROJ_OK) return st;
        dst = d_out.as<float>();
    }
    if ((st = planeIntegrals(vol, maps, distances, n_dist, side, resolution, dst)) != CTLPROJ_OK) return st;
    if (device) return CTLPROJ_OK;
    hipError_t e = hipMemcpyAsync(out, dst, count * sizeof(float), hipMemcpyDeviceToHost, s);
    return finish(e, s, "plane integral copy");
}

ctlproj_status ctlproj_radon3d_plane_integrals(const ctlproj_volume* vol, const double* normals, uint32_t n_orient,
                                               const float* distances, uint32_t n_dist, uint32_t side,
                                               float resolution, float* out, size_t out_count)
{
    return radonImpl(vol, normals, n_orient, distances, n_dist, side, resolution, out, out_count, false);
}

ctlproj_status ctlproj_radon3d_plane_integrals_device(const ctlproj_volume* vol, const double* normals,
                                                      uint32_t n_orient, const float* distances, uint32_t n_dist,
                                                      uint32_t side, float resolution, float* device_out,
                                                      size_t out_count)
{
    return radonImpl(vol, normals, n_orient, distances, n_dist, side, resolution, device_out, out_count, true);
}

ctlproj_status ctlproj_radon3d_sample_transform(const ctlproj_volume* vol, const float* azimuth, uint32_t n_azi,
                                                const float* polar, uint32_t n_pol, const float* distances,
                                                uint32_t n_dist, uint32_t side, float resolution, float* out,
                                                size_t out_count)
{
    if ((n_azi && !azimuth) || (n_pol && !polar)) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "NULL angle array");
    const size_t n_orient = size_t(n_azi) * n_pol;
    if (n_orient > 0xffffffffull) return fail(CTLPROJ_ERR_INVALID_ARGUMENT, "too many orientations");
    // radontransform3d.cpp:117-141, 391-399: polar outer, azimuth inner; angles in double
    std::vector<double> normals(3 * n_orient);
    for (uint32_t p = 0; p < n_pol; ++p)
        for (uint32_t a = 0; a < n_azi; ++a) {
            const double pol = double(polar[p]), azi = double(azimuth[a]);
            double* v = normals.data() + 3 * (size_t(p) * n_azi + a);
            v[0] = std::sin(pol) * std::cos(azi);
            v[1] = std::sin(pol) * std::sin(azi);
            v[2] = std::cos(pol);
        }
    return radonImpl(vol, normals.data(), uint32_t(n_orient), distances, n_dist, side, resolution, out, out_count,
                     false);
}
