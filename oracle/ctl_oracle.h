/*
 * ctl_oracle.h -- CPU oracle for CTL's ray-caster forward projector.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker (or as the timed
 * CPU baseline). The HIP product path never calls it.
 *
 * Two restatements of the reference algorithm live here:
 *   ref32 -- the OpenCL kernel's numerics as a CPU OpenCL runtime executes them
 *            (modules/src/ocl/cl_src/projectors/raycasterprojector_interp.cl:18-148 and
 *             raycasterprojector_no_interp.cl:16-154, host side
 *             modules/src/projectors/raycasterprojector.cpp:139-375, 527-672);
 *   ref64 -- the same sample positions and loop bounds with every quantity in FP64
 *            (the "FP64 CPU ray-caster" that error statements are made against);
 * and the reference's own CPU projector:
 *   refcpu -- RayCasterProjectorCPU (modules/src/projectors/raycasterprojectorcpu.cpp:132-407):
 *            FP64 direction, positions (std::fma), clip and sum; FP32 weights and values; its
 *            unmasked read at a position exactly n + 1/2 (n for nearest) on an axis.
 * Plus the FP64 geometry helpers the host side uses (QR of M, source position).
 */
#ifndef CTL_ORACLE_H
#define CTL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_settings {
    uint32_t rays_per_pixel[2]; /* raycasterprojector.h:31 */
    float ray_sampling;         /* raycasterprojector.h:32 */
    uint32_t volume_upsampling; /* raycasterprojector.h:33 */
    int32_t interpolate;        /* raycasterprojector.h:34 */
} oracle_settings;

/* Householder QR of a 3x3 matrix (row-major), restating mat::QRdecomposition
 * (modules/src/mat/matrix_algorithm.cpp:14-37) and qrdcmp (:125-171). Returns 1 if singular. */
int oracle_qr(const double A[9], double Q[9], double R[9]);

/* decomposeM (raycasterprojector.cpp:527-541): QR of M = P[:, 0:3] with R negated when
 * R00*R11*R22 < 0, packed as [Q row-major (9), R00, R01, R02, R11, R12, R22, 0]. */
void oracle_decompose_m(const double P[12], double qr16[16]);

/* ProjectionMatrix::sourcePosition (modules/src/mat/projectionmatrix.cpp:264-274). */
void oracle_source_position(const double P[12], double c[3]);

/* Ray direction for detector pixel (x, y) from the packed QR, in FP64, before the FP32 cast
 * (raycasterprojector_interp.cl:74-86 without the final normalize). */
void oracle_direction64(const double qr16[16], double x, double y, double d[3]);

/*
 * Forward projection of a float volume (x fastest: index = z*X*Y + y*X + x,
 * voxelvolume.tpp:310-332) for all configured views, or for a subset of them.
 *
 *   pmats    [n_views][n_modules][3][4] row-major FP64 projection matrices
 *   out      [n_out][n_modules][rows][channels] float, n_out = n_subset or n_views
 *   out64    optional FP64 copy of the same values (ref64 only; may be NULL)
 *   fp64     0 -> ref32, 1 -> ref64, 2 -> refcpu (rays per pixel and interpolation apply, the
 *            upsampling factor does not: the CPU projector has none)
 *   n_samples  optional: total number of sampling-loop iterations (reference loop count)
 * Returns 0 on success, nonzero on invalid arguments.
 */
int oracle_project(const float* volume, const uint32_t dims[3], const float voxel_mm[3],
                   const float offset_mm[3], const double* pmats, uint32_t n_views,
                   uint32_t n_modules, uint32_t channels, uint32_t rows,
                   const oracle_settings* settings, int fp64, int n_threads,
                   const uint32_t* view_subset, uint32_t n_subset, float* out, double* out64,
                   uint64_t* n_samples);

/* ref32 single-sample helpers, exposed so tests can pin them against brute force. */
float oracle_sample_linear32(const float* vol, const uint32_t dims[3], float px, float py, float pz);
double oracle_sample_linear64(const float* vol, const uint32_t dims[3], double px, double py,
                              double pz);
/* refcpu single-sample reads (interpolatedRead / nonInterpolatedRead). */
float oracle_cpu_interpolated_read(const float* vol, const uint32_t dims[3], double px, double py, double pz);
float oracle_cpu_nearest_read(const float* vol, const uint32_t dims[3], double px, double py, double pz);
/* ref32 slab clip (raycasterprojector_interp.cl:88-148); writes (lambda_min, lambda_max). */
void oracle_clip32(const float source[3], const float direction[3], const float vol_size[3],
                   const float vol_corner[3], int interpolate, float minmax[2]);

/* ---- volume processing (ctl_oracle_processing.c); fp64: 0 -> ref32, 1 -> ref64 ---------- */

/* VolumeSlicer::slice (modules/src/processing/volumeslicer.cpp:107-161, 204-222 and
 * modules/src/ocl/cl_src/processing/volumeSlicer.cl): the plane {p : normal . p = distance},
 * width x height pixels of `resolution` mm centred on distance * normal; out[y][x]. */
int oracle_slice(const float* vol, const uint32_t dims[3], const float voxel_mm[3],
                 const float offset_mm[3], const double normal[3], double distance, uint32_t width,
                 uint32_t height, float resolution, int fp64, float* out);
/* VolumeResampler::resample / sample (volumeresampler.cpp:117-213, volumeResampler.cl):
 * ranges = {lo1, hi1, lo2, hi2, lo3, hi3} map onto the voxel centres 0 .. N-1 of each axis.
 * resample: out[k][j][i] at (p1[i], p2[j], p3[k]); sample: out[i] at xyz[3i .. 3i+2]. */
int oracle_resample(const float* vol, const uint32_t dims[3], const float ranges[6],
                    const float* p1, uint32_t n1, const float* p2, uint32_t n2, const float* p3,
                    uint32_t n3, int fp64, float* out);
int oracle_sample_points(const float* vol, const uint32_t dims[3], const float ranges[6],
                         const float* xyz, size_t n, int fp64, float* out);
/* RadonTransform3D plane integrals (radontransform3d.cpp:55-156, 188-216, 303-450,
 * planeIntegral.cl): out[d][o] = integral over the plane {p : normals[o] . p = distances[d]}
 * sampled on a width x height grid of `resolution` mm. */
int oracle_plane_integrals(const float* vol, const uint32_t dims[3], const float voxel_mm[3],
                           const float offset_mm[3], const double* normals, uint32_t n_orient,
                           const float* distances, uint32_t n_dist, uint32_t width, uint32_t height,
                           float resolution, int fp64, float* out);

#ifdef __cplusplus
}
#endif

#endif
