/*
 * ctl_oracle.c -- CPU oracle for CTL's ray-caster forward projector.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, as the checker or as the timed CPU baseline.
 *
 * ref32 restates the reference's OpenCL kernel as a CPU OpenCL runtime executes it:
 *   - FP64 direction from the packed QR, cast to float3 and normalised in FP32
 *     (raycasterprojector_interp.cl:74-86);
 *   - FP32 slab clip with the half-voxel margin and 0.01-voxel relaxation (:88-148);
 *   - positions mad((float)i, step, source - corner) in FP32, taken here as a fused multiply-add;
 *   - the OpenCL 1.2 linear filter (CLK_FILTER_LINEAR, CLK_ADDRESS_CLAMP, unnormalised
 *     coordinates): texel centres at k + 1/2, i0 = floor(u - 1/2), a = frac(u - 1/2), border 0,
 *     full-precision FP32 weights (a CPU runtime does not quantise them);
 *   - FP64 accumulation, output increment_mm * (float)sum / (float)(rx * ry) (:61-71).
 * ref64 keeps the same loop structure with every quantity in FP64.
 *
 * Implementation-defined points of the OpenCL source and the choice made here:
 *   mad(a, b, c)  -> fmaf(a, b, c) (x86 CPU runtimes emit FMA);
 *   normalize(v)  -> v / sqrtf(dot(v, v)) with a left-to-right FP32 dot product;
 *   (uint)lambda for a miss (FLT_MAX) -> the ray contributes nothing;
 *   contraction of a*b + c elsewhere -> none (the file is built with -ffp-contract=off).
 */
#include "ctl_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------------- */
/* FP64 3x3 helpers                                                                             */
/* ------------------------------------------------------------------------------------------- */

/* mat::det (modules/src/mat/matrix_algorithm.cpp, det()) -- cofactor expansion along row 0. */
static double det3(const double m[3][3])
{
    return m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
         - m[0][1] * (m[1][0] * m[2][2] - m[2][0] * m[1][2])
         + m[0][2] * (m[1][0] * m[2][1] - m[2][0] * m[1][1]);
}

/* qFuzzyIsNull(double): |d| <= 1e-12 (Qt's definition). */
static int fuzzy_null(double d) { return fabs(d) <= 0.000000000001; }

/* Householder QR in place, restating qrdcmp (matrix_algorithm.cpp:125-171). */
static int qrdcmp3(double a[3][3], double d[3])
{
    const int n = 3;
    double c[2];
    int sing = 0;
    for (int k = 0; k < n - 1; k++) {
        double scale = 0.0;
        for (int i = k; i < n; i++)
            if (scale < fabs(a[i][k])) scale = fabs(a[i][k]);
        if (fuzzy_null(scale)) {
            sing = 1;
            c[k] = 0.0;
            d[k] = 0.0;
        } else {
            double sum = 0.0;
            for (int i = k; i < n; i++) a[i][k] /= scale;
            for (int i = k; i < n; i++) sum += a[i][k] * a[i][k];
            const double sigma = copysign(sqrt(sum), a[k][k]);
            a[k][k] += sigma;
            c[k] = sigma * a[k][k];
            d[k] = -scale * sigma;
            for (int j = k + 1; j < n; j++) {
                double s = 0.0;
                for (int i = k; i < n; i++) s += a[i][k] * a[i][j];
                const double tau = s / c[k];
                for (int i = k; i < n; i++) a[i][j] -= tau * a[i][k];
            }
        }
    }
    d[n - 1] = a[n - 1][n - 1];
    if (fuzzy_null(d[n - 1])) sing = 1;
    return sing;
}

/* mat::QRdecomposition (matrix_algorithm.cpp:14-37): Q = Q1*Q2 from the two Householder
 * vectors left in the lower part of the factored matrix, R = upper part with diagonal d. */
int oracle_qr(const double A[9], double Q[9], double R[9])
{
    double a[3][3];
    double d[3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = A[3 * r + c];
    const int sing = qrdcmp3(a, d);

    /* R = [d0, a01, a02; 0, d1, a12; 0, 0, d2] (B = a^T, so B(1,0) = a[0][1] etc.) */
    R[0] = d[0]; R[1] = a[0][1]; R[2] = a[0][2];
    R[3] = 0.0;  R[4] = d[1];    R[5] = a[1][2];
    R[6] = 0.0;  R[7] = 0.0;     R[8] = d[2];

    double u1[3] = { a[0][0], a[1][0], a[2][0] };
    double u2[3] = { 0.0, a[1][1], a[2][1] };
    const double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    const double n2 = sqrt(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
    for (int i = 0; i < 3; ++i) { u1[i] /= n1; u2[i] /= n2; }

    double Q1[3][3], Q2[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            Q1[r][c] = (r == c ? 1.0 : 0.0) - 2.0 * u1[r] * u1[c];
            Q2[r][c] = (r == c ? 1.0 : 0.0) - 2.0 * u2[r] * u2[c];
        }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Q1[r][k] * Q2[k][c];
            Q[3 * r + c] = s;
        }
    return sing;
}

/* decomposeM (raycasterprojector.cpp:527-541). */
void oracle_decompose_m(const double P[12], double qr16[16])
{
    double M[9], Q[9], R[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[3 * r + c] = P[4 * r + c];
    oracle_qr(M, Q, R);
    if (signbit(R[0] * R[4] * R[8]))
        for (int i = 0; i < 9; ++i) R[i] = -R[i];
    for (int i = 0; i < 9; ++i) qr16[i] = Q[i];
    qr16[9] = R[0]; qr16[10] = R[1]; qr16[11] = R[2];
    qr16[12] = R[4]; qr16[13] = R[5];
    qr16[14] = R[8];
    qr16[15] = 0.0;
}

/* ProjectionMatrix::sourcePosition (projectionmatrix.cpp:264-274): Cramer's rule for the null
 * space of P, c = [det(p2,p3,p4), -det(p1,p3,p4), det(p1,p2,p4)] / (-det(p1,p2,p3)). */
void oracle_source_position(const double P[12], double c[3])
{
    double m[3][3];
    int cols[4][3] = { { 0, 1, 2 }, { 1, 2, 3 }, { 0, 2, 3 }, { 0, 1, 3 } };
    double dets[4];
    for (int k = 0; k < 4; ++k) {
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) m[r][j] = P[4 * r + cols[k][j]];
        dets[k] = det3(m);
    }
    const double hom2cart = -dets[0];
    c[0] = dets[1] / hom2cart;
    c[1] = -dets[2] / hom2cart;
    c[2] = dets[3] / hom2cart;
}

/* calculateDirection (raycasterprojector_interp.cl:74-86) up to the FP32 cast:
 * d = R^-1 Q^T [x, y, 1]^T by back substitution. */
void oracle_direction64(const double q[16], double x, double y, double d[3])
{
    const double qx = q[0] * x + q[3] * y + q[6];
    const double qy = q[1] * x + q[4] * y + q[7];
    const double qz = q[2] * x + q[5] * y + q[8];
    const double dz = qz / q[14];
    const double dy = (qy - dz * q[13]) / q[12];
    const double dx = (qx - dy * q[10] - dz * q[11]) / q[9];
    d[0] = dx; d[1] = dy; d[2] = dz;
}

/* ------------------------------------------------------------------------------------------- */
/* ref32 pieces                                                                                 */
/* ------------------------------------------------------------------------------------------- */

/* checkFace (raycasterprojector_interp.cl:136-148). NaN comparisons are false, which is what
 * rejects the faces of axis-parallel rays. */
static void check_face32(float h0, float h1, float c10, float c11, float c20, float c21,
                         float lambda, float mm[2])
{
    const int intersects = (h0 >= c10) && (h1 >= c11) && (h0 <= c20) && (h1 <= c21);
    if (intersects && lambda < mm[0]) mm[0] = lambda;
    if (intersects && lambda > mm[1]) mm[1] = lambda;
}

/* calculateIntersections (raycasterprojector_interp.cl:88-134; no_interp.cl:95-140 uses the box
 * without the half-voxel margin). */
void oracle_clip32(const float s[3], const float d[3], const float vol_size[3],
                   const float vol_corner[3], int interpolate, float mm[2])
{
    const float margin = interpolate ? 0.5f : 0.0f;
    float c1[3], c2[3], l1[3], l2[3];
    for (int k = 0; k < 3; ++k) {
        c1[k] = vol_corner[k] - margin;
        c2[k] = vol_corner[k] + vol_size[k] + margin;
        l1[k] = (c1[k] - s[k]) / d[k];
        l2[k] = (c2[k] - s[k]) / d[k];
    }
    for (int k = 0; k < 3; ++k) {
        c1[k] -= 0.01f;
        c2[k] += 0.01f;
    }
    mm[0] = FLT_MAX;
    mm[1] = 0.0f;
    const float* ls[2] = { l1, l2 };
    for (int side = 0; side < 2; ++side) {
        const float* l = ls[side];
        /* yz-face */
        check_face32(s[1] + l[0] * d[1], s[2] + l[0] * d[2], c1[1], c1[2], c2[1], c2[2], l[0], mm);
        /* xz-face */
        check_face32(s[0] + l[1] * d[0], s[2] + l[1] * d[2], c1[0], c1[2], c2[0], c2[2], l[1], mm);
        /* xy-face */
        check_face32(s[0] + l[2] * d[0], s[1] + l[2] * d[1], c1[0], c1[1], c2[0], c2[1], l[2], mm);
    }
    mm[0] = fmaxf(mm[0], 0.0f);
    mm[1] = fmaxf(mm[1], 0.0f);
}

static float texel32(const float* v, const uint32_t* n, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0 || i >= (int)n[0] || j >= (int)n[1] || k >= (int)n[2]) return 0.0f;
    return v[((size_t)k * n[1] + (size_t)j) * n[0] + (size_t)i];
}

/* OpenCL 1.2 linear filter for a 3D image with unnormalised coordinates and CLK_ADDRESS_CLAMP
 * (border colour 0): T = sum over the 8 texels of (1-a|a)(1-b|b)(1-c|c) T_ijk. */
float oracle_sample_linear32(const float* v, const uint32_t n[3], float px, float py, float pz)
{
    /* Every texel is border beyond one voxel outside the grid: return early so the integer
     * conversions below stay in range for far-away positions. */
    if (!(px > -1.0f && py > -1.0f && pz > -1.0f && px < (float)n[0] + 1.0f &&
          py < (float)n[1] + 1.0f && pz < (float)n[2] + 1.0f))
        return 0.0f;
    const float u = px - 0.5f, w = py - 0.5f, t = pz - 0.5f;
    const float fu = floorf(u), fw = floorf(w), ft = floorf(t);
    const int i0 = (int)fu, j0 = (int)fw, k0 = (int)ft;
    const float a = u - fu, b = w - fw, c = t - ft;
    const float ma = 1.0f - a, mb = 1.0f - b, mc = 1.0f - c;
    return ma * mb * mc * texel32(v, n, i0, j0, k0)
         + a * mb * mc * texel32(v, n, i0 + 1, j0, k0)
         + ma * b * mc * texel32(v, n, i0, j0 + 1, k0)
         + a * b * mc * texel32(v, n, i0 + 1, j0 + 1, k0)
         + ma * mb * c * texel32(v, n, i0, j0, k0 + 1)
         + a * mb * c * texel32(v, n, i0 + 1, j0, k0 + 1)
         + ma * b * c * texel32(v, n, i0, j0 + 1, k0 + 1)
         + a * b * c * texel32(v, n, i0 + 1, j0 + 1, k0 + 1);
}

double oracle_sample_linear64(const float* v, const uint32_t n[3], double px, double py, double pz)
{
    if (!(px > -1.0 && py > -1.0 && pz > -1.0 && px < (double)n[0] + 1.0 &&
          py < (double)n[1] + 1.0 && pz < (double)n[2] + 1.0))
        return 0.0;
    const double u = px - 0.5, w = py - 0.5, t = pz - 0.5;
    const double fu = floor(u), fw = floor(w), ft = floor(t);
    const int i0 = (int)fu, j0 = (int)fw, k0 = (int)ft;
    const double a = u - fu, b = w - fw, c = t - ft;
    const double ma = 1.0 - a, mb = 1.0 - b, mc = 1.0 - c;
    return ma * mb * mc * texel32(v, n, i0, j0, k0)
         + a * mb * mc * texel32(v, n, i0 + 1, j0, k0)
         + ma * b * mc * texel32(v, n, i0, j0 + 1, k0)
         + a * b * mc * texel32(v, n, i0 + 1, j0 + 1, k0)
         + ma * mb * c * texel32(v, n, i0, j0, k0 + 1)
         + a * mb * c * texel32(v, n, i0 + 1, j0, k0 + 1)
         + ma * b * c * texel32(v, n, i0, j0 + 1, k0 + 1)
         + a * b * c * texel32(v, n, i0 + 1, j0 + 1, k0 + 1);
}

/* Nearest lookup (raycasterprojector_no_interp.cl:63-73): floor (convert_int3_rtn) and a
 * bounds check. */
static float sample_nearest32(const float* v, const uint32_t* n, float px, float py, float pz)
{
    if (!(px >= 0.0f && py >= 0.0f && pz >= 0.0f && px < (float)n[0] && py < (float)n[1] &&
          pz < (float)n[2]))
        return 0.0f;
    return texel32(v, n, (int)floorf(px), (int)floorf(py), (int)floorf(pz));
}

static double sample_nearest64(const float* v, const uint32_t* n, double px, double py, double pz)
{
    if (!(px >= 0.0 && py >= 0.0 && pz >= 0.0 && px < (double)n[0] && py < (double)n[1] &&
          pz < (double)n[2]))
        return 0.0;
    return texel32(v, n, (int)floor(px), (int)floor(py), (int)floor(pz));
}

/* ------------------------------------------------------------------------------------------- */
/* ref64 clip                                                                                   */
/* ------------------------------------------------------------------------------------------- */

static void check_face64(double h0, double h1, double c10, double c11, double c20, double c21,
                         double lambda, double mm[2])
{
    const int intersects = (h0 >= c10) && (h1 >= c11) && (h0 <= c20) && (h1 <= c21);
    if (intersects && lambda < mm[0]) mm[0] = lambda;
    if (intersects && lambda > mm[1]) mm[1] = lambda;
}

static void clip64(const double s[3], const double d[3], const double vol_size[3],
                   const double vol_corner[3], int interpolate, double mm[2])
{
    const double margin = interpolate ? 0.5 : 0.0;
    double c1[3], c2[3], l1[3], l2[3];
    for (int k = 0; k < 3; ++k) {
        c1[k] = vol_corner[k] - margin;
        c2[k] = vol_corner[k] + vol_size[k] + margin;
        l1[k] = (c1[k] - s[k]) / d[k];
        l2[k] = (c2[k] - s[k]) / d[k];
    }
    for (int k = 0; k < 3; ++k) {
        c1[k] -= 0.01;
        c2[k] += 0.01;
    }
    mm[0] = DBL_MAX;
    mm[1] = 0.0;
    const double* ls[2] = { l1, l2 };
    for (int side = 0; side < 2; ++side) {
        const double* l = ls[side];
        check_face64(s[1] + l[0] * d[1], s[2] + l[0] * d[2], c1[1], c1[2], c2[1], c2[2], l[0], mm);
        check_face64(s[0] + l[1] * d[0], s[2] + l[1] * d[2], c1[0], c1[2], c2[0], c2[2], l[1], mm);
        check_face64(s[0] + l[2] * d[0], s[1] + l[2] * d[1], c1[0], c1[1], c2[0], c2[1], l[2], mm);
    }
    mm[0] = fmax(mm[0], 0.0);
    mm[1] = fmax(mm[1], 0.0);
}

/* ------------------------------------------------------------------------------------------- */
/* refcpu: RayCasterProjectorCPU (modules/src/projectors/raycasterprojectorcpu.cpp)              */
/* ------------------------------------------------------------------------------------------- */

/* VoxelVolume<float>::operator() (voxelvolume.tpp:324-332): the linear index z*X*Y + y*X + x
 * without a per-axis bounds check. The reference asserts lup < size in debug builds and reads
 * past the array in release builds; an index beyond the array reads 0 here. */
static float voxel_linear(const float* v, const uint32_t* n, int x, int y, int z)
{
    const size_t lup = (size_t)(uint32_t)z * n[0] * n[1] + (size_t)(uint32_t)y * n[0] + (size_t)(uint32_t)x;
    return lup < (size_t)n[0] * n[1] * n[2] ? v[lup] : 0.0f;
}

/* interpolatedRead (raycasterprojectorcpu.cpp:239-285): FP64 position, FP32 weights and values,
 * taps at or beyond a low border read 0, taps beyond the high border read 0 except the v0xx taps
 * of a position exactly at n + 1/2 on that axis, where floor(p - 1/2) = n is not masked and the
 * read lands on the linear index (the reference's out-of-bounds read). */
float oracle_cpu_interpolated_read(const float* v, const uint32_t n[3], double px, double py, double pz)
{
    if (px < -0.5 || py < -0.5 || pz < -0.5 || px > n[0] + 0.5 || py > n[1] + 0.5 || pz > n[2] + 0.5)
        return 0.0f;
    const int vx = (int)floor(px - 0.5), vy = (int)floor(py - 0.5), vz = (int)floor(pz - 0.5);
    const int lo0 = vx < 0, lo1 = vy < 0, lo2 = vz < 0;
    const int hi0 = vx >= (int)n[0] - 1, hi1 = vy >= (int)n[1] - 1, hi2 = vz >= (int)n[2] - 1;
    const float w0 = (float)(px - (vx + 0.5)), w1 = (float)(py - (vy + 0.5)), w2 = (float)(pz - (vz + 0.5));
    const float v000 = (lo0 || lo1 || lo2) ? 0.0f : voxel_linear(v, n, vx, vy, vz);
    const float v001 = (lo0 || lo1 || hi2) ? 0.0f : voxel_linear(v, n, vx, vy, vz + 1);
    const float v010 = (lo0 || hi1 || lo2) ? 0.0f : voxel_linear(v, n, vx, vy + 1, vz);
    const float v011 = (lo0 || hi1 || hi2) ? 0.0f : voxel_linear(v, n, vx, vy + 1, vz + 1);
    const float v100 = (hi0 || lo1 || lo2) ? 0.0f : voxel_linear(v, n, vx + 1, vy, vz);
    const float v101 = (hi0 || lo1 || hi2) ? 0.0f : voxel_linear(v, n, vx + 1, vy, vz + 1);
    const float v110 = (hi0 || hi1 || lo2) ? 0.0f : voxel_linear(v, n, vx + 1, vy + 1, vz);
    const float v111 = (hi0 || hi1 || hi2) ? 0.0f : voxel_linear(v, n, vx + 1, vy + 1, vz + 1);
    const float w0o = 1.0f - w0;
    const float c00 = w0o * v000 + w0 * v100;
    const float c01 = w0o * v001 + w0 * v101;
    const float c10 = w0o * v010 + w0 * v110;
    const float c11 = w0o * v011 + w0 * v111;
    const float w1o = 1.0f - w1;
    const float c0 = c00 * w1o + c10 * w1;
    const float c1 = c01 * w1o + c11 * w1;
    return c0 * (1.0f - w2) + c1 * w2;
}

/* nonInterpolatedRead (raycasterprojectorcpu.cpp:287-299): floor of the FP64 position; a
 * position exactly at n on an axis passes the range test and reads the linear index. */
float oracle_cpu_nearest_read(const float* v, const uint32_t n[3], double px, double py, double pz)
{
    if (px < 0.0 || py < 0.0 || pz < 0.0 || px > n[0] || py > n[1] || pz > n[2]) return 0.0f;
    return voxel_linear(v, n, (int)floor(px), (int)floor(py), (int)floor(pz));
}

/* ------------------------------------------------------------------------------------------- */
/* projection driver                                                                            */
/* ------------------------------------------------------------------------------------------- */

typedef struct job {
    /* volume after VolumeSpecs::upSampleVolume (raycasterprojector.cpp:576-672) */
    const float* vol;
    uint32_t n[3];
    float vs[3];     /* voxel size (FP32, as the host passes it) */
    float corner[3]; /* volumeCorner (raycasterprojector.cpp:552-558), FP32 */
    double corner64[3];
    float increment; /* min(voxel size) * raySampling (raycasterprojector.cpp:178-179) */
    double increment64;
    const double* pmats;
    uint32_t n_modules, channels, rows;
    const uint32_t* views;
    uint32_t n_out;
    oracle_settings s;
    int fp64;
    int cpu;  /* 1: RayCasterProjectorCPU restatement (refcpu) */
    float* out;
    double* out64;
    /* work distribution: one task = one (output view, module, row) */
    uint64_t n_tasks;
    uint64_t next;
    pthread_mutex_t lock;
    uint64_t samples;
} job;

/* One detector row of RayCasterProjectorCPU::computeView (raycasterprojectorcpu.cpp:132-235):
 * FP64 direction from the packed QR normalised in FP64, scaled per axis by the FP32 increment in
 * voxels (increment_mm / voxel size, both FP32); FP64 source, corner (FP32 volumeCorner / voxel
 * size) and slab clip; positions std::fma(i, direction, corner-to-source) in FP64; FP32
 * interpolation; FP64 sum; out = increment_mm * (float)sum / (float)(rx * ry). */
static void run_row_cpu(const job* J, const double qr[16], const double src64[3], size_t out_base,
                        uint32_t y, uint64_t* samples)
{
    const uint32_t rx = J->s.rays_per_pixel[0], ry = J->s.rays_per_pixel[1];
    const int interp = J->s.interpolate;
    double source[3], corner[3], c2s[3], vsize[3], inc_vox[3];
    for (int k = 0; k < 3; ++k) {
        source[k] = src64[k] / (double)J->vs[k];
        corner[k] = (double)J->corner[k] / (double)J->vs[k];
        inc_vox[k] = (double)(J->increment / J->vs[k]);
        c2s[k] = source[k] - corner[k];
        vsize[k] = (double)J->n[k];
    }
    const double ipx = 1.0 / (double)rx, ipy = 1.0 / (double)ry;
    for (uint32_t x = 0; x < J->channels; ++x) {
        const double pcx = ((double)x - 0.5) + 0.5 * ipx;
        const double pcy = ((double)y - 0.5) + 0.5 * ipy;
        double acc = 0.0;
        for (uint32_t a = 0; a < rx; ++a)
            for (uint32_t b = 0; b < ry; ++b) {
                double d[3];
                oracle_direction64(qr, pcx + (double)a * ipx, pcy + (double)b * ipy, d);
                const double len = sqrt(0.0 + d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                for (int k = 0; k < 3; ++k) d[k] = d[k] / len * inc_vox[k];
                double mm[2];
                clip64(source, d, vsize, corner, interp, mm);
                /* a miss (DBL_MAX) contributes nothing, as in ref32: the reference's (uint) cast
                 * of DBL_MAX is undefined and, on x86, samples two positions next to the source */
                if (mm[0] == DBL_MAX) continue;
                const uint32_t i0 = (uint32_t)mm[0];
                const uint32_t i1 = (uint32_t)mm[1] + 1u;
                for (uint32_t i = i0; i <= i1; ++i) {
                    const double fi = (double)i;
                    const double p0 = fma(fi, d[0], c2s[0]);
                    const double p1 = fma(fi, d[1], c2s[1]);
                    const double p2 = fma(fi, d[2], c2s[2]);
                    acc += (double)(interp ? oracle_cpu_interpolated_read(J->vol, J->n, p0, p1, p2)
                                           : oracle_cpu_nearest_read(J->vol, J->n, p0, p1, p2));
                }
                if (i1 >= i0) *samples += (uint64_t)(i1 - i0) + 1u;
            }
        J->out[out_base + x] = J->increment * (float)acc / (float)(rx * ry);
    }
}

static void run_row(job* J, uint32_t ov, uint32_t m, uint32_t y, uint64_t* samples)
{
    const uint32_t view = J->views ? J->views[ov] : ov;
    const double* P = J->pmats + ((size_t)view * J->n_modules + m) * 12;
    const double* P0 = J->pmats + (size_t)view * J->n_modules * 12; /* module 0: source */
    double qr[16], src64[3];
    oracle_decompose_m(P, qr);
    oracle_source_position(P0, src64);
    const uint32_t rx = J->s.rays_per_pixel[0], ry = J->s.rays_per_pixel[1];
    const size_t out_base = (((size_t)ov * J->n_modules + m) * J->rows + y) * J->channels;
    const int interp = J->s.interpolate;

    if (J->cpu) {
        run_row_cpu(J, qr, src64, out_base, y, samples);
        return;
    }
    if (!J->fp64) {
        /* kernel prologue, raycasterprojector_interp.cl:35-40 */
        const float src_mm[3] = { (float)src64[0], (float)src64[1], (float)src64[2] };
        float source[3], corner[3], c2s[3], vsize[3];
        for (int k = 0; k < 3; ++k) {
            source[k] = src_mm[k] / J->vs[k];
            corner[k] = J->corner[k] / J->vs[k];
            c2s[k] = source[k] - corner[k];
            vsize[k] = (float)J->n[k];
        }
        const float ipx = 1.0f / (float)rx, ipy = 1.0f / (float)ry;
        for (uint32_t x = 0; x < J->channels; ++x) {
            const float pcx = (float)x - 0.5f + 0.5f * ipx;
            const float pcy = (float)y - 0.5f + 0.5f * ipy;
            double acc = 0.0;
            for (uint32_t a = 0; a < rx; ++a)
                for (uint32_t b = 0; b < ry; ++b) {
                    const float px = pcx + (float)a * ipx;
                    const float py = pcy + (float)b * ipy;
                    double d64[3];
                    oracle_direction64(qr, (double)px, (double)py, d64);
                    float d[3] = { (float)d64[0], (float)d64[1], (float)d64[2] };
                    const float dot = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    const float len = sqrtf(dot);
                    for (int k = 0; k < 3; ++k) d[k] = d[k] / len;
                    for (int k = 0; k < 3; ++k) d[k] = J->increment * d[k];
                    for (int k = 0; k < 3; ++k) d[k] /= J->vs[k];
                    float mm[2];
                    oracle_clip32(source, d, vsize, corner, interp, mm);
                    if (mm[0] == FLT_MAX) continue; /* miss */
                    const uint32_t i0 = (uint32_t)mm[0];
                    const uint32_t i1 = (uint32_t)mm[1] + 1u;
                    for (uint32_t i = i0; i <= i1; ++i) {
                        const float fi = (float)i;
                        const float p0 = fmaf(fi, d[0], c2s[0]);
                        const float p1 = fmaf(fi, d[1], c2s[1]);
                        const float p2 = fmaf(fi, d[2], c2s[2]);
                        const float v = interp ? oracle_sample_linear32(J->vol, J->n, p0, p1, p2)
                                               : sample_nearest32(J->vol, J->n, p0, p1, p2);
                        acc += (double)v;
                    }
                    if (i1 >= i0) *samples += (uint64_t)(i1 - i0) + 1u;
                }
            J->out[out_base + x] = J->increment * (float)acc / (float)(rx * ry);
        }
    } else {
        double source[3], corner[3], c2s[3], vsize[3];
        for (int k = 0; k < 3; ++k) {
            source[k] = src64[k] / (double)J->vs[k];
            corner[k] = J->corner64[k] / (double)J->vs[k];
            c2s[k] = source[k] - corner[k];
            vsize[k] = (double)J->n[k];
        }
        const double ipx = 1.0 / (double)rx, ipy = 1.0 / (double)ry;
        for (uint32_t x = 0; x < J->channels; ++x) {
            const double pcx = (double)x - 0.5 + 0.5 * ipx;
            const double pcy = (double)y - 0.5 + 0.5 * ipy;
            double acc = 0.0;
            for (uint32_t a = 0; a < rx; ++a)
                for (uint32_t b = 0; b < ry; ++b) {
                    double d[3];
                    oracle_direction64(qr, pcx + (double)a * ipx, pcy + (double)b * ipy, d);
                    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    for (int k = 0; k < 3; ++k) d[k] = J->increment64 * (d[k] / len) / (double)J->vs[k];
                    double mm[2];
                    clip64(source, d, vsize, corner, interp, mm);
                    if (mm[0] == DBL_MAX) continue;
                    const uint32_t i0 = (uint32_t)mm[0];
                    const uint32_t i1 = (uint32_t)mm[1] + 1u;
                    for (uint32_t i = i0; i <= i1; ++i) {
                        const double fi = (double)i;
                        const double p0 = fma(fi, d[0], c2s[0]);
                        const double p1 = fma(fi, d[1], c2s[1]);
                        const double p2 = fma(fi, d[2], c2s[2]);
                        acc += interp ? oracle_sample_linear64(J->vol, J->n, p0, p1, p2)
                                      : sample_nearest64(J->vol, J->n, p0, p1, p2);
                    }
                    if (i1 >= i0) *samples += (uint64_t)(i1 - i0) + 1u;
                }
            const double val = J->increment64 * acc / (double)(rx * ry);
            J->out[out_base + x] = (float)val;
            if (J->out64) J->out64[out_base + x] = val;
        }
    }
}

static void* worker(void* arg)
{
    job* J = (job*)arg;
    uint64_t samples = 0;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        const uint64_t t = J->next++;
        pthread_mutex_unlock(&J->lock);
        if (t >= J->n_tasks) break;
        const uint32_t y = (uint32_t)(t % J->rows);
        const uint64_t vm = t / J->rows;
        run_row(J, (uint32_t)(vm / J->n_modules), (uint32_t)(vm % J->n_modules), y, &samples);
    }
    pthread_mutex_lock(&J->lock);
    J->samples += samples;
    pthread_mutex_unlock(&J->lock);
    return NULL;
}

int oracle_project(const float* volume, const uint32_t dims[3], const float voxel_mm[3],
                   const float offset_mm[3], const double* pmats, uint32_t n_views,
                   uint32_t n_modules, uint32_t channels, uint32_t rows,
                   const oracle_settings* settings, int fp64, int n_threads,
                   const uint32_t* view_subset, uint32_t n_subset, float* out, double* out64,
                   uint64_t* n_samples)
{
    if (!volume || !dims || !voxel_mm || !offset_mm || !pmats || !settings || !out) return 1;
    if (n_modules == 0 || channels == 0 || rows == 0) return 1;
    if (settings->rays_per_pixel[0] == 0 || settings->rays_per_pixel[1] == 0) return 1;
    if (dims[0] == 0 || dims[1] == 0 || dims[2] == 0) return 1;
    if (view_subset)
        for (uint32_t i = 0; i < n_subset; ++i)
            if (view_subset[i] >= n_views) return 1;

    job J;
    memset(&J, 0, sizeof J);
    J.s = *settings;
    /* non-interpolating mode forces the upsampling factor to 1 (raycasterprojector.cpp:148-151);
     * RayCasterProjectorCPU has no upsampling setting */
    const uint32_t f = (J.s.interpolate && fp64 != 2) ? J.s.volume_upsampling : 1u;

    /* VolumeSpecs::upSampleVolume (raycasterprojector.cpp:576-672) */
    float* up = NULL;
    float mean = 0.0f;
    if (f == 0) {
        for (size_t i = 0, n = (size_t)dims[0] * dims[1] * dims[2]; i < n; ++i) mean += volume[i];
        mean /= (float)((size_t)dims[0] * dims[1] * dims[2]);
        J.vol = &mean;
        J.n[0] = J.n[1] = J.n[2] = 1;
        for (int k = 0; k < 3; ++k) J.vs[k] = voxel_mm[k] * (float)dims[k];
    } else if (f == 1) {
        J.vol = volume;
        for (int k = 0; k < 3; ++k) { J.n[k] = dims[k]; J.vs[k] = voxel_mm[k]; }
    } else {
        const size_t X = dims[0], Y = dims[1], Z = dims[2];
        const size_t nX = X * f, nY = Y * f, nZ = Z * f;
        up = (float*)malloc(nX * nY * nZ * sizeof(float));
        if (!up) return 2;
        for (size_t z = 0; z < Z; ++z)
            for (size_t y = 0; y < Y; ++y)
                for (size_t x = 0; x < X; ++x) {
                    const float val = volume[(z * Y + y) * X + x];
                    for (size_t iz = 0; iz < f; ++iz)
                        for (size_t iy = 0; iy < f; ++iy)
                            for (size_t ix = 0; ix < f; ++ix)
                                up[x * f + ix + (y * f + iy) * nX + (z * f + iz) * nX * nY] = val;
                }
        J.vol = up;
        J.n[0] = (uint32_t)nX; J.n[1] = (uint32_t)nY; J.n[2] = (uint32_t)nZ;
        for (int k = 0; k < 3; ++k) J.vs[k] = voxel_mm[k] / (float)f;
    }
    float smallest = J.vs[0];
    if (J.vs[1] < smallest) smallest = J.vs[1];
    if (J.vs[2] < smallest) smallest = J.vs[2];
    J.increment = smallest * J.s.ray_sampling;
    J.increment64 = (double)smallest * (double)J.s.ray_sampling;
    for (int k = 0; k < 3; ++k) {
        J.corner[k] = offset_mm[k] - 0.5f * (float)J.n[k] * J.vs[k];
        J.corner64[k] = (double)offset_mm[k] - 0.5 * (double)J.n[k] * (double)J.vs[k];
    }
    J.pmats = pmats;
    J.n_modules = n_modules;
    J.channels = channels;
    J.rows = rows;
    J.views = view_subset;
    J.n_out = view_subset ? n_subset : n_views;
    J.fp64 = fp64 == 1;
    J.cpu = fp64 == 2;
    J.out = out;
    J.out64 = out64;
    J.n_tasks = (uint64_t)J.n_out * n_modules * rows;
    pthread_mutex_init(&J.lock, NULL);

    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 1; t < n_threads; ++t)
        if (pthread_create(&th[started], NULL, worker, &J) == 0) ++started;
    worker(&J);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&J.lock);
    if (n_samples) *n_samples = J.samples;
    free(up);
    return 0;
}
