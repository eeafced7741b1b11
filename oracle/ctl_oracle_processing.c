/*
 * ctl_oracle_processing.c -- CPU oracle for CTL's volume processing on OpenCL images:
 * VolumeSlicer, VolumeResampler and RadonTransform3D's plane integrals.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE: loaded only by tests/ as the checker.
 *
 * Every function has a ref32 form (fp64 = 0: the host's FP64 geometry cast to float as the
 * reference's host code does, the kernel's arithmetic in float, OpenCL's linear sampler in float)
 * and a ref64 form (fp64 = 1: the same sample positions with every quantity in FP64). Volumes are
 * float, x fastest (index = z*X*Y + y*X + x).
 */
#include "ctl_oracle.h"

#include <math.h>
#include <stddef.h>

/* Columns r1, r2, r3 of the rotation that takes the xy-plane to the plane with unit normal n:
 * r3 = n; r2 = n x e_a, normalised, for the axis a along which n has its smallest component
 * (ties resolved towards the later axis); r1 = r2 x r3. Follows
 * modules/src/processing/volumeslicer.cpp:207-217 (identical in radontransform3d.cpp:425-438).
 * R is row-major: R[3*i + j] = component i of column j. */
static void plane_rotation(const double n[3], double R[9])
{
    int axis = fabs(n[0]) < fabs(n[1]) ? 0 : 1;
    axis = fabs(n[axis]) < fabs(n[2]) ? axis : 2;
    double e[3] = { 0.0, 0.0, 0.0 };
    e[axis] = 1.0;
    double r2[3] = { n[1] * e[2] - n[2] * e[1], n[2] * e[0] - n[0] * e[2], n[0] * e[1] - n[1] * e[0] };
    const double len = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
    for (int k = 0; k < 3; ++k) r2[k] /= len;
    const double r1[3] = { r2[1] * n[2] - r2[2] * n[1], r2[2] * n[0] - r2[0] * n[2],
                           r2[0] * n[1] - r2[1] * n[0] };
    for (int i = 0; i < 3; ++i) {
        R[3 * i + 0] = r1[i];
        R[3 * i + 1] = r2[i];
        R[3 * i + 2] = n[i];
    }
}

int oracle_slice(const float* vol, const uint32_t dims[3], const float voxel_mm[3],
                 const float offset_mm[3], const double normal[3], double distance, uint32_t width,
                 uint32_t height, float resolution, int fp64, float* out)
{
    if (!vol || !dims || !voxel_mm || !offset_mm || !normal || !out) return 1;
    /* homography (volumeslicer.cpp:204-222): diag(1/voxel) * [reso * R | R * (0, 0, d)] */
    double R[9], H[12];
    plane_rotation(normal, R);
    for (int i = 0; i < 3; ++i) {
        const double inv = 1.0 / (double)voxel_mm[i];
        for (int j = 0; j < 3; ++j) H[4 * i + j] = inv * ((double)resolution * R[3 * i + j]);
        H[4 * i + 3] = inv * (R[3 * i + 2] * distance);
    }
    if (!fp64) {
        /* host side in float (volumeslicer.cpp:122-130), kernel in float (volumeSlicer.cl) */
        float h[12], corner[3];
        for (int k = 0; k < 12; ++k) h[k] = (float)H[k];
        for (int k = 0; k < 3; ++k)
            corner[k] = -0.5f * (float)(dims[k] - 1u) + offset_mm[k] / voxel_mm[k];
        const float cx = 0.5f * (float)(width - 1u), cy = 0.5f * (float)(height - 1u);
        for (uint32_t y = 0; y < height; ++y)
            for (uint32_t x = 0; x < width; ++x) {
                const float vx = (float)x - cx, vy = (float)y - cy;
                float u[3];
                for (int k = 0; k < 3; ++k)
                    u[k] = h[4 * k] * vx + h[4 * k + 1] * vy + h[4 * k + 2] * 0.0f + h[4 * k + 3] -
                           corner[k] + 0.5f;
                out[(size_t)y * width + x] = oracle_sample_linear32(vol, dims, u[0], u[1], u[2]);
            }
    } else {
        double corner[3];
        for (int k = 0; k < 3; ++k)
            corner[k] = -0.5 * ((double)dims[k] - 1.0) + (double)offset_mm[k] / (double)voxel_mm[k];
        const double cx = 0.5 * ((double)width - 1.0), cy = 0.5 * ((double)height - 1.0);
        for (uint32_t y = 0; y < height; ++y)
            for (uint32_t x = 0; x < width; ++x) {
                const double vx = (double)x - cx, vy = (double)y - cy;
                double u[3];
                for (int k = 0; k < 3; ++k)
                    u[k] = H[4 * k] * vx + H[4 * k + 1] * vy + H[4 * k + 3] - corner[k] + 0.5;
                out[(size_t)y * width + x] = (float)oracle_sample_linear64(vol, dims, u[0], u[1], u[2]);
            }
    }
    return 0;
}

/* Sampler coordinate of a point along axis k (volumeResampler.cl:20-32): the range [lo, hi]
 * spans the voxel centres 0 .. N-1. */
static float resample_coord32(const uint32_t* dims, const float* ranges, int k, float p)
{
    const float scale = ((float)dims[k] - 1.0f) / (ranges[2 * k + 1] - ranges[2 * k]);
    return scale * (p - ranges[2 * k]) + 0.5f;
}

static double resample_coord64(const uint32_t* dims, const float* ranges, int k, float p)
{
    const double scale = ((double)dims[k] - 1.0) / ((double)ranges[2 * k + 1] - (double)ranges[2 * k]);
    return scale * ((double)p - (double)ranges[2 * k]) + 0.5;
}

int oracle_resample(const float* vol, const uint32_t dims[3], const float ranges[6],
                    const float* p1, uint32_t n1, const float* p2, uint32_t n2, const float* p3,
                    uint32_t n3, int fp64, float* out)
{
    if (!vol || !dims || !ranges || !out) return 1;
    for (uint32_t k = 0; k < n3; ++k)
        for (uint32_t j = 0; j < n2; ++j)
            for (uint32_t i = 0; i < n1; ++i) {
                const size_t o = (size_t)i + (size_t)j * n1 + (size_t)k * n1 * n2;
                if (!fp64)
                    out[o] = oracle_sample_linear32(vol, dims, resample_coord32(dims, ranges, 0, p1[i]),
                                                    resample_coord32(dims, ranges, 1, p2[j]),
                                                    resample_coord32(dims, ranges, 2, p3[k]));
                else
                    out[o] = (float)oracle_sample_linear64(vol, dims, resample_coord64(dims, ranges, 0, p1[i]),
                                                           resample_coord64(dims, ranges, 1, p2[j]),
                                                           resample_coord64(dims, ranges, 2, p3[k]));
            }
    return 0;
}

int oracle_sample_points(const float* vol, const uint32_t dims[3], const float ranges[6],
                         const float* xyz, size_t n, int fp64, float* out)
{
    if (!vol || !dims || !ranges || (n && (!xyz || !out))) return 1;
    for (size_t i = 0; i < n; ++i) {
        const float* p = xyz + 3 * i;
        if (!fp64)
            out[i] = oracle_sample_linear32(vol, dims, resample_coord32(dims, ranges, 0, p[0]),
                                            resample_coord32(dims, ranges, 1, p[1]),
                                            resample_coord32(dims, ranges, 2, p[2]));
        else
            out[i] = (float)oracle_sample_linear64(vol, dims, resample_coord64(dims, ranges, 0, p[0]),
                                                   resample_coord64(dims, ranges, 1, p[1]),
                                                   resample_coord64(dims, ranges, 2, p[2]));
    }
    return 0;
}

/* Plane integrals (radontransform3d.cpp:55-156, 188-216, 303-450; planeIntegral.cl):
 * the template plane is a width x height pixel grid of `resolution` mm centred on the origin,
 * rotated onto the normal and shifted by the distance along it; every pixel reads the volume with
 * the linear sampler, 16x16 patches are summed on the device (row sums, then the total), the
 * host adds the patch sums in order and scales by resolution^2. out is [n_dist][n_orient]. */
int oracle_plane_integrals(const float* vol, const uint32_t dims[3], const float voxel_mm[3],
                           const float offset_mm[3], const double* normals, uint32_t n_orient,
                           const float* distances, uint32_t n_dist, uint32_t width, uint32_t height,
                           float resolution, int fp64, float* out)
{
    if (!vol || !dims || !voxel_mm || !offset_mm || (n_orient && !normals) || (n_dist && !distances) ||
        !out)
        return 1;
    const uint32_t P = 16; /* PATCH_SIZE */
    double corner[3];
    for (int k = 0; k < 3; ++k) /* radontransform3d.cpp:306-308 */
        corner[k] = (double)offset_mm[k] - 0.5 * (double)dims[k] * (double)voxel_mm[k];
    const double start[3] = { -(double)resolution * 0.5 * (double)(width - 1u),
                              -(double)resolution * 0.5 * (double)(height - 1u), 0.0 };
    for (uint32_t o = 0; o < n_orient; ++o) {
        const double* n = normals + 3 * (size_t)o;
        double R[9], H[12];
        plane_rotation(n, R);
        for (int i = 0; i < 3; ++i) { /* transformXYPlaneToCentralPlane, radontransform3d.cpp:440-450 */
            const double inv = 1.0 / (double)voxel_mm[i];
            const double t = R[3 * i] * start[0] + R[3 * i + 1] * start[1] + R[3 * i + 2] * start[2] - corner[i];
            for (int j = 0; j < 3; ++j) H[4 * i + j] = inv * ((double)resolution * R[3 * i + j]);
            H[4 * i + 3] = inv * t;
        }
        for (uint32_t d = 0; d < n_dist; ++d) {
            if (!fp64) {
                float h[12], shift[3];
                for (int k = 0; k < 12; ++k) h[k] = (float)H[k];
                for (int k = 0; k < 3; ++k) shift[k] = (float)n[k] / voxel_mm[k]; /* :359-361 */
                for (int k = 0; k < 3; ++k) h[4 * k + 3] += distances[d] * shift[k];
                float total = 0.0f; /* host loop, radontransform3d.cpp:86-94 */
                for (uint32_t gy = 0; gy * P < height; ++gy)
                    for (uint32_t gx = 0; gx * P < width; ++gx) {
                        float patch = 0.0f;
                        for (uint32_t py = 0; py < P && gy * P + py < height; ++py) {
                            float row = 0.0f;
                            for (uint32_t px = 0; px < P && gx * P + px < width; ++px) {
                                const float x = (float)(gx * P + px), y = (float)(gy * P + py);
                                float u[3];
                                for (int k = 0; k < 3; ++k)
                                    u[k] = h[4 * k] * x + h[4 * k + 1] * y + h[4 * k + 2] * 0.0f + h[4 * k + 3];
                                row += oracle_sample_linear32(vol, dims, u[0], u[1], u[2]);
                            }
                            patch += row;
                        }
                        total += patch;
                    }
                out[(size_t)d * n_orient + o] = total * powf(resolution, 2.0f);
            } else {
                double shift[3], h3[3];
                for (int k = 0; k < 3; ++k) {
                    shift[k] = n[k] / (double)voxel_mm[k];
                    h3[k] = H[4 * k + 3] + (double)distances[d] * shift[k];
                }
                double total = 0.0;
                for (uint32_t y = 0; y < height; ++y)
                    for (uint32_t x = 0; x < width; ++x) {
                        double u[3];
                        for (int k = 0; k < 3; ++k) u[k] = H[4 * k] * x + H[4 * k + 1] * y + h3[k];
                        total += oracle_sample_linear64(vol, dims, u[0], u[1], u[2]);
                    }
                out[(size_t)d * n_orient + o] = (float)(total * (double)resolution * (double)resolution);
            }
        }
    }
    return 0;
}
