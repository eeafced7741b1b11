// HIP kernels of the forward projector, written for gfx950 (CDNA4, wave64).
//
//   setup_kernel    FP64, one thread per (view, module): source position (Cramer's rule) and
//                   sign-corrected M^-1, folded into an FP32 per-pixel direction basis.
//   raycast_kernel  one thread per detector pixel, wavefront = 8x8 (tunable) pixel tile; FP32
//                   slab clip, fixed-step trilinear samples gathered from the padded HBM volume
//                   (FP32 / FP16 / BF16 storage), FP32 interpolation and accumulation.
//   count_kernel    the reference's sampling-loop count of every ray (for Gsamples/s).
//   pack / upsample / sum / fill kernels for resident volumes.
//
// Gathers use raw buffer loads: an offset past the end of the volume returns 0 instead of
// touching memory, so no index can fault even if a position were computed wrongly.
#include "kernels.h"
#include "sampler.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cstdint>
#include <mutex>

namespace ctlk {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNumXcd = 8; // MI355X: 8 XCDs, workgroups dealt round-robin

// Workgroup b of nb is dispatched to XCD b % 8; give each XCD a contiguous range of tiles so that
// neighbouring tiles (which gather the same voxels) share that XCD's L2.
__device__ __forceinline__ uint32_t xcdRemap(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    const uint32_t x = b % kNumXcd, k = b / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------------------------------------
// FP64 setup
//
// One thread per (view, module) restates the reference's host-side preparation in the same
// operations and order, so that every ray starts from the same FP64 and FP32 values the
// reference computes: decomposeM (raycasterprojector.cpp:527-541, Householder QR of M as
// mat::QRdecomposition / qrdcmp, matrix_algorithm.cpp:14-37, 125-171, with R negated when
// R00 R11 R22 < 0), the source position by Cramer's rule (projectionmatrix.cpp:264-274, module 0
// of the view, raycasterprojector.cpp:323-325) cast to FP32, and the kernel prologue's FP32
// source, corner and corner-to-source vector (raycasterprojector_interp.cl:35-37). No
// multiply-add is fused here (fp contract off); FP64 division and square root are IEEE.

__device__ double det3(const double* m)
{
#pragma clang fp contract(off)
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) +
           m[2] * (m[3] * m[7] - m[6] * m[4]);
}

// qrdcmp of the 3x3 matrix a (row-major), in place; d receives R's diagonal.
__device__ void householder3(double* a, double* d)
{
#pragma clang fp contract(off)
    double c[2];
    for (int k = 0; k < 2; ++k) {
        double scale = 0.0;
        for (int i = k; i < 3; ++i)
            if (scale < fabs(a[3 * i + k])) scale = fabs(a[3 * i + k]);
        if (fabs(scale) <= 1e-12) { // qFuzzyIsNull: a singular column
            c[k] = 0.0;
            d[k] = 0.0;
        } else {
            for (int i = k; i < 3; ++i) a[3 * i + k] /= scale;
            double sum = 0.0;
            for (int i = k; i < 3; ++i) sum += a[3 * i + k] * a[3 * i + k];
            const double sigma = copysign(sqrt(sum), a[3 * k + k]);
            a[3 * k + k] += sigma;
            c[k] = sigma * a[3 * k + k];
            d[k] = -scale * sigma;
            for (int j = k + 1; j < 3; ++j) {
                double t = 0.0;
                for (int i = k; i < 3; ++i) t += a[3 * i + k] * a[3 * i + j];
                const double tau = t / c[k];
                for (int i = k; i < 3; ++i) a[3 * i + j] -= tau * a[3 * i + k];
            }
        }
    }
    d[2] = a[8];
}

__global__ void setup_kernel(const double* __restrict__ pm, uint32_t views, uint32_t modules,
                             VolumeGeom vg, float shift, RayGeom* __restrict__ out)
{
#pragma clang fp contract(off)
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= views * modules) return;
    const uint32_t v = t / modules;
    const double* P = pm + size_t(t) * 12;
    const double* P0 = pm + size_t(v) * modules * 12;

    // source = null space of P0 by Cramer's rule
    double m[9];
    const int cols[4][3] = { { 0, 1, 2 }, { 1, 2, 3 }, { 0, 2, 3 }, { 0, 1, 3 } };
    double dets[4];
    for (int k = 0; k < 4; ++k) {
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) m[3 * r + j] = P0[4 * r + cols[k][j]];
        dets[k] = det3(m);
    }
    const double h = -dets[0];
    const double src[3] = { dets[1] / h, -dets[2] / h, dets[3] / h };

    // decomposeM: Q = Q1 Q2 from the Householder vectors, R upper with diagonal d
    double a[9], d[3];
    for (int r = 0; r < 3; ++r)
        for (int j = 0; j < 3; ++j) a[3 * r + j] = P[4 * r + j];
    householder3(a, d);
    double R[6] = { d[0], a[1], a[2], d[1], a[5], d[2] }; // R00 R01 R02 R11 R12 R22
    double u1[3] = { a[0], a[3], a[6] };
    double u2[3] = { 0.0, a[4], a[7] };
    const double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    const double n2 = sqrt(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
    for (int i = 0; i < 3; ++i) {
        u1[i] /= n1;
        u2[i] /= n2;
    }
    double Q1[9], Q2[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            Q1[3 * r + c] = (r == c ? 1.0 : 0.0) - 2.0 * u1[r] * u1[c];
            Q2[3 * r + c] = (r == c ? 1.0 : 0.0) - 2.0 * u2[r] * u2[c];
        }
    RayGeom g;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            double q = 0.0;
            for (int k = 0; k < 3; ++k) q += Q1[3 * r + k] * Q2[3 * k + c];
            g.qr[3 * r + c] = q;
        }
    if (signbit(R[0] * R[3] * R[5]))
        for (int i = 0; i < 6; ++i) R[i] = -R[i];
    for (int i = 0; i < 6; ++i) g.qr[9 + i] = R[i];

    // kernel prologue (FP32): source and corner in voxel units and their difference; volumeCorner
    // (raycasterprojector.cpp:552-558) as the host computes it in FP32
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
        const float corner_mm = vg.offset_mm[k] - 0.5f * float(vg.n[k]) * vg.voxel_mm[k];
        g.src[k] = float(src[k]) / vg.voxel_mm[k];
        g.corner[k] = corner_mm / vg.voxel_mm[k];
        // sample positions of the gathers: the reference's position plus `shift` (the sampler's
        // -1/2 and the padding's +1)
        g.c[k] = (g.src[k] - g.corner[k]) + shift;
        ok = ok && isfinite(g.c[k]);
    }
    for (int i = 0; i < 15; ++i) ok = ok && isfinite(g.qr[i]);
    g.valid = ok ? 1u : 0u;
    out[t] = g;
}

// ---------------------------------------------------------------------------------------------
// per-ray setup (FP32)

struct Ray {
    float s[3]; // step per sample, texel units
    int i0, i1; // inclusive sample index range (reference loop bounds)
    bool hit;
};

// checkFace (raycasterprojector_interp.cl:136-148)
__device__ __forceinline__ void checkFace(float h0, float h1, float c10, float c11, float c20,
                                          float c21, float lambda, float& lo, float& hi)
{
#pragma clang fp contract(off)
    const bool in = (h0 >= c10) & (h1 >= c11) & (h0 <= c20) & (h1 <= c21);
    lo = (in & (lambda < lo)) ? lambda : lo;
    hi = (in & (lambda > hi)) ? lambda : hi;
}

// The sub-ray through detector coordinate (px, py), as the reference's kernel computes it:
// calculateDirection in FP64 from the packed QR (raycasterprojector_interp.cl:74-86), cast to
// FP32 and normalised (v / sqrt(dot(v, v))), scaled by the increment and divided by the voxel size
// (:53-57); then calculateIntersections (:88-134) in the reference's own frame (source and volume
// corner in voxel units, the box [corner - margin, corner + n + margin], 0.01-voxel relaxation)
// and the loop bounds (uint)lambda_min .. (uint)lambda_max + 1. The operations and their order are
// the reference's, none fused (fp contract off), so the step, the bounds and (with the gathers'
// position frame c = source - corner + shift, computed once per view) the sample positions are
// the reference's bit for bit. nhi = n + 1 per axis; blo = 0 selects the trilinear box (margin
// 1/2), blo = 1 the nearest-mode box (no margin, raycasterprojector_no_interp.cl:95-140).
template <class G> // RayGeom, or RayGeom in the constant address space
__device__ __forceinline__ Ray setupRay(const G& g, float px, float py, float inc,
                                        const float vs[3], const float nhi[3], float blo)
{
#pragma clang fp contract(off)
    Ray r;
    const double x = double(px), y = double(py);
    const auto* q = g.qr;
    const double qx = q[0] * x + q[3] * y + q[6];
    const double qy = q[1] * x + q[4] * y + q[7];
    const double qz = q[2] * x + q[5] * y + q[8];
    const double dz = qz / q[14];
    const double dy = (qy - dz * q[13]) / q[12];
    const double dx = (qx - dy * q[10] - dz * q[11]) / q[9];
    float d[3] = { float(dx), float(dy), float(dz) };
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k) r.s[k] = (inc * (d[k] / len)) / vs[k];

    const float margin = blo == 0.0f ? 0.5f : 0.0f;
    float c1[3], c2[3], l1[3], l2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c1[k] = g.corner[k] - margin;
        c2[k] = g.corner[k] + (nhi[k] - 1.0f) + margin;
        l1[k] = (c1[k] - g.src[k]) / r.s[k];
        l2[k] = (c2[k] - g.src[k]) / r.s[k];
        c1[k] -= 0.01f;
        c2[k] += 0.01f;
    }
    float lo = FLT_MAX, hi = 0.0f;
    const auto* c = g.src;
    const float* s = r.s;
    checkFace(c[1] + l1[0] * s[1], c[2] + l1[0] * s[2], c1[1], c1[2], c2[1], c2[2], l1[0], lo, hi);
    checkFace(c[0] + l1[1] * s[0], c[2] + l1[1] * s[2], c1[0], c1[2], c2[0], c2[2], l1[1], lo, hi);
    checkFace(c[0] + l1[2] * s[0], c[1] + l1[2] * s[1], c1[0], c1[1], c2[0], c2[1], l1[2], lo, hi);
    checkFace(c[1] + l2[0] * s[1], c[2] + l2[0] * s[2], c1[1], c1[2], c2[1], c2[2], l2[0], lo, hi);
    checkFace(c[0] + l2[1] * s[0], c[2] + l2[1] * s[2], c1[0], c1[2], c2[0], c2[2], l2[1], lo, hi);
    checkFace(c[0] + l2[2] * s[0], c[1] + l2[2] * s[1], c1[0], c1[1], c2[0], c2[1], l2[2], lo, hi);
    r.hit = lo != FLT_MAX;
    lo = fmaxf(lo, 0.0f);
    hi = fmaxf(hi, 0.0f);
    // (uint) conversions of the loop bounds; lambda stays far below 2^30 for any real geometry,
    // the clamp only keeps the integer arithmetic defined.
    r.i0 = int(fminf(lo, 1073741824.0f));
    r.i1 = int(fminf(hi, 1073741824.0f)) + 1;
    return r;
}

// Detector coordinate of sub-ray k of pixel x with ip = 1 / rays per pixel
// (raycasterprojector_interp.cl:39-40, 52), unfused as the reference computes it.
__device__ __forceinline__ float subRayCoord(uint32_t x, float ip, uint32_t k)
{
#pragma clang fp contract(off)
    const float corner = float(x) - 0.5f + 0.5f * ip;
    return corner + float(k) * ip;
}

// Samples whose position lies outside [blo, n+1) on any axis read only border texels and add
// exactly 0; drop them from both ends so the hot loop needs no bounds test. Positions are
// monotone in i on every axis, so the remaining range is contiguous.
__device__ __forceinline__ bool insideQ(int i, const float s[3], const float c[3], float blo,
                                        const float nhi[3])
{
#pragma clang fp contract(off)
    const float fi = float(i);
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float q = fmaf(fi, s[k], c[k]);
        in = in & (q >= blo) & (q < nhi[k]);
    }
    return in;
}

// The walks run a step or two per end; kept as plain loops (the compiler's four-way unrolled
// forms spilled registers to scratch in the staged caster and cost it 2%).
__device__ __forceinline__ void tighten(Ray& r, const float c[3], float blo, const float nhi[3])
{
    if (!r.hit) { r.i0 = 1; r.i1 = 0; return; }
#pragma clang loop unroll(disable) vectorize(disable) interleave(disable)
    while (r.i0 <= r.i1 && !insideQ(r.i0, r.s, c, blo, nhi)) ++r.i0;
#pragma clang loop unroll(disable) vectorize(disable) interleave(disable)
    while (r.i1 >= r.i0 && !insideQ(r.i1, r.s, c, blo, nhi)) --r.i1;
}

// Wave-wide min / max: DPP reductions within each 16-lane row, then the four row results. All
// 64 lanes must be active.
template <bool MAX>
__device__ __forceinline__ int waveReduce(int v)
{
    auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xF, 0xF, false)); // row_ror:4
    v = op(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false)); // row_ror:8
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}

__device__ __forceinline__ int waveMin(int v) { return waveReduce<false>(v); }
__device__ __forceinline__ int waveMax(int v) { return waveReduce<true>(v); }

// Wave-wide minimum (maximum) of a float: DPP steps within each row of 16 lanes, then the four
// rows' values. All 64 lanes must be active.
template <bool MAX>
__device__ __forceinline__ float waveReduceF(float v)
{
    auto op = [](float a, float b) { return MAX ? fmaxf(a, b) : fminf(a, b); };
    v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0xB1, 0xF, 0xF, false)));
    v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x4E, 0xF, 0xF, false)));
    v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x124, 0xF, 0xF, false)));
    v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x128, 0xF, 0xF, false)));
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return op(op(r0, r1), op(r2, r3));
}

__device__ __forceinline__ float waveMinF(float v) { return waveReduceF<false>(v); }
__device__ __forceinline__ float waveMaxF(float v) { return waveReduceF<true>(v); }

// Adds this wave's lanes' counts to *tally (lane 0's atomic); all lanes active.
__device__ __forceinline__ void tallyWave(unsigned long long* tally, uint32_t n)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) n += __shfl_xor(n, m, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(tally, (unsigned long long)n);
}

// Samples per segment of the unstaged casters' walk when a volume has an occupancy table.
constexpr int kSegment = 64;

// True if no voxel in the padded-grid box [lo, hi] (inclusive, clamped to the table) is non-zero:
// eight corner reads of the summed-area table (uniform indices, so scalar loads).
__device__ __forceinline__ bool bricksEmpty(const VolumeView& vol, const uint32_t vlo[3], const uint32_t vhi[3])
{
    uint32_t lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const uint32_t l = vlo[k] >> kOccShift, h = vhi[k] >> kOccShift, last = vol.occ_n[k] - 1;
        lo[k] = __builtin_amdgcn_readfirstlane(l < last ? l : last);
        hi[k] = __builtin_amdgcn_readfirstlane((h < last ? h : last) + 1);
    }
    const uint32_t nx = vol.occ_n[0] + 1, nxy = nx * (vol.occ_n[1] + 1);
    // read as constant memory (the table is not written while kernels run): with uniform indices
    // these are scalar loads, which leave the texture path to the gathers
    typedef __attribute__((address_space(4))) const uint32_t occ_word;
    const occ_word* t = (const occ_word*)(vol.occ);
    auto at = [&](uint32_t i, uint32_t j, uint32_t k) { return t[k * nxy + j * nx + i]; };
    const uint32_t n = at(hi[0], hi[1], hi[2]) - at(lo[0], hi[1], hi[2]) - at(hi[0], lo[1], hi[2]) -
                       at(hi[0], hi[1], lo[2]) + at(lo[0], lo[1], hi[2]) + at(lo[0], hi[1], lo[2]) +
                       at(hi[0], lo[1], lo[2]) - at(lo[0], lo[1], lo[2]);
    return n == 0;
}

// True if the wave's samples with indices in [s0, s1] (0 <= s0 <= s1) read only zero voxels. All
// rays of a (view, module) start from the same point c, so along each axis the positions
// fma(i, s, c) of the active rays lie between fma(i, smin, c) and fma(i, smax, c) (the FMA is
// monotonic in s for i >= 0, and in i), extremes at i = s0 or s1; a sample's taps span
// floor(q) .. floor(q) + 1, and samples outside [0, n + 1) are never taken.
__device__ __forceinline__ bool rangeEmpty(const VolumeView& vol, const float nhi[3], int s0, int s1,
                                           const float smin[3], const float smax[3], const float c[3])
{
    uint32_t lo[3], hi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a0 = fmaf(float(s0), smin[k], c[k]), a1 = fmaf(float(s1), smin[k], c[k]);
        const float b0 = fmaf(float(s0), smax[k], c[k]), b1 = fmaf(float(s1), smax[k], c[k]);
        const float ql = fmaxf(fminf(a0, a1), 0.0f), qh = fminf(fmaxf(b0, b1), nhi[k]);
        if (!(ql == ql) || !(qh == qh)) return false; // not a number: take the samples
        if (qh < ql) return true;                     // the segment's positions miss the support
        lo[k] = uint32_t(ql);
        hi[k] = uint32_t(qh) + 1u;
    }
    return bricksEmpty(vol, lo, hi);
}

// LARGE: a volume of 4 GiB or more, read at 64-bit addresses from `base` (interleaved layout
// only); otherwise through the buffer `rs`.
template <Storage ST, Layout LY, bool INTERP, bool LARGE = false>
__device__ __forceinline__ float sampleAt(__amdgpu_buffer_rsrc_t rs, const Pitch& p, float fi,
                                          const float s[3], const float c[3], const void* base = nullptr)
{
    const float qx = fmaf(fi, s[0], c[0]);
    const float qy = fmaf(fi, s[1], c[1]);
    const float qz = fmaf(fi, s[2], c[2]);
    if constexpr (LARGE)
        return INTERP ? trilinearWide<ST>(base, p, qx, qy, qz) : nearestWide<ST>(base, p, qx, qy, qz);
    else
        return INTERP ? trilinear<ST, LY>(rs, p, qx, qy, qz) : nearest<ST, LY>(rs, p, qx, qy, qz);
}

// Processes logical tile b of one projection: one thread per detector pixel. All rays of a
// wavefront step through the same sample indices i (the same distance from the source), so the
// 64 gathers of one load instruction fall into a small patch of the volume and share cache lines.
// Each ray's own range [i0, i1] (the reference's loop bounds, trimmed to samples that can touch
// the volume) is honoured with the exec mask in the head and tail of the wave's range; in between
// every active ray samples, unmasked and unrolled.
// Pixel (lx, ly) of `lane` in its wavefront tile. Each group of 16 lanes covers 8 x 2 pixels
// (4 x 4 or 2 x 8 in narrower tiles) as four 2 x 2-pixel quads: the lanes that the LDS and the
// texture path serve together then sample neighbouring positions. On the 256^3 / 1024 x 768
// workload this is 5% (FP32, staged) to 8% (FP16) faster than row-major lanes, with the same
// output. Tiles one pixel wide or tall stay row-major. Either way lx and ly each gather a subset
// of the lane's six bits, in order: up to three runs of bits, each a shift and a mask (9 bits of
// the code, from laneCode). Branch-free on purpose: a branch here made the compiler serialise
// the gather loop's loads.
__device__ __forceinline__ uint32_t gatherBits(uint32_t lane, uint32_t e)
{
    return ((lane >> (e & 7u)) & ((e >> 3) & 63u)) | ((lane >> ((e >> 9) & 7u)) & ((e >> 12) & 63u)) |
           ((lane >> ((e >> 18) & 7u)) & ((e >> 21) & 63u));
}

template <Storage ST, Layout LY, bool INTERP, int UNROLL, bool LARGE = false>
__device__ __forceinline__ void castTile(const VolumeView& vol, const RayGeom* __restrict__ geom,
                                         const RayParams& rp, const CastConst& kc,
                                         float* __restrict__ out, uint32_t b, uint32_t wave)
{
    const uint32_t vm = b / kc.tiles_per_vm;
    const uint32_t t = b - vm * kc.tiles_per_vm;
    const uint32_t ty = t / kc.tiles_x, tx = t - ty * kc.tiles_x;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wy = wave / kc.bwx, wx = wave - wy * kc.bwx;
    const uint32_t lx = gatherBits(lane, kc.lane_x), ly = gatherBits(lane, kc.lane_y);
    const uint32_t x = (tx * kc.bwx + wx) * kc.wave_w + lx;
    const uint32_t y = (ty * kc.bwy + wy) * kc.wave_h + ly;
    const bool pixel = x < rp.channels && y < rp.rows;

    const RayGeom* gp = geom + vm;
    const __amdgpu_buffer_rsrc_t rs = makeRsrc(vol.data, vol.bytes); // unused by LARGE gathers

    // kernel prologue of raycasterprojector_interp.cl:39-40
    const float ipx = 1.0f / float(rp.rpp_x), ipy = 1.0f / float(rp.rpp_y);

    float acc = 0.0f;
    if (gp->valid) {
        for (uint32_t ra = 0; ra < rp.rpp_x; ++ra)
            for (uint32_t rb = 0; rb < rp.rpp_y; ++rb) {
                // the 160-byte record is read again for each sub-ray (the empty asm hides the
                // pointer's value): held in registers across the sample loop, its FP64 QR pushed the
                // persistent kernels to 124 VGPRs, where the compiler issued the unrolled gathers one
                // at a time (served FP16 scans 1.6x slower)
                const RayGeom* gr = gp;
                asm volatile("" : "+s"(gr));
                // read as constant memory (nothing writes it while kernels run): scalar loads that
                // leave the texture path to the gathers; the sample positions' origin goes to
                // registers once
                typedef __attribute__((address_space(4))) const RayGeom geom_const;
                const geom_const& g = *(const geom_const*)gr;
                const float gc[3] = { g.c[0], g.c[1], g.c[2] };
                Ray ray = setupRay(g, subRayCoord(x, ipx, ra), subRayCoord(y, ipy, rb), kc.inc, kc.vs,
                                   kc.nhi, kc.blo);
                if (!pixel) ray.hit = false;
                tighten(ray, gc, kc.blo, kc.nhi);
                const bool act = pixel && ray.i0 <= ray.i1;
                const int i0 = act ? ray.i0 : INT_MAX, i1 = act ? ray.i1 : INT_MIN;
                const int lo = waveMin(i0), hi = waveMax(i1);
                if (lo > hi) continue; // no ray of this wave touches the volume
                int a = waveMax(act ? i0 : INT_MIN), bb = waveMin(act ? i1 : INT_MAX);
                if (a > bb) { a = hi + 1; bb = hi; } // no common range: everything masked
                const uint32_t span = uint32_t(i1 - i0);
                // With an occupancy table the range is walked in segments of kSegment samples,
                // anchored at the common start `a` (kSegment is a multiple of UNROLL, so every sample
                // lands in the partial sum it lands in without segments), and a segment whose
                // samples read only zero voxels is skipped: its samples are exactly 0.
                float smin[3] = { 0.0f, 0.0f, 0.0f }, smax[3] = { 0.0f, 0.0f, 0.0f };
                if (vol.occ) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        smin[k] = waveMinF(act ? ray.s[k] : FLT_MAX);
                        smax[k] = waveMaxF(act ? ray.s[k] : -FLT_MAX);
                    }
                }
                const int anchor = a <= hi ? a : lo;
                const int seg = vol.occ ? kSegment : (hi - lo + 1);

                float part[UNROLL];
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) part[u] = 0.0f;
                for (int c0 = anchor - (anchor - lo + seg - 1) / seg * seg; c0 <= hi; c0 += seg) {
                    const int s0 = max(c0, lo), s1 = min(c0 + seg - 1, hi);
                    if (vol.occ && rangeEmpty(vol, kc.nhi, s0, s1, smin, smax, gc)) continue; // uniform
                    if (kc.tally) {
                        const int ta = max(s0, i0), te = min(s1, i1);
                        tallyWave(kc.tally, act && ta <= te ? uint32_t(te - ta + 1) : 0u);
                    }
                    // head [lo, a) and, when there is no common range, everything
                    for (int i = s0; i <= min(s1, a - 1); ++i)
                        if (act && uint32_t(i - i0) <= span)
                            part[0] += sampleAt<ST, LY, INTERP, LARGE>(rs, kc.pitch, float(i), ray.s, gc, vol.data);
                    // body [a, bb]: every active ray samples
                    if (act) {
                        const int be = min(s1, bb);
                        int i = max(s0, a);
                        float fi = float(i);
                        for (; i + UNROLL - 1 <= be; i += UNROLL, fi += float(UNROLL)) {
#pragma unroll
                            for (int u = 0; u < UNROLL; ++u)
                                part[u] += sampleAt<ST, LY, INTERP, LARGE>(rs, kc.pitch, fi + float(u), ray.s, gc, vol.data);
                        }
                        for (; i <= be; ++i, fi += 1.0f)
                            part[0] += sampleAt<ST, LY, INTERP, LARGE>(rs, kc.pitch, fi, ray.s, gc, vol.data);
                    }
                    // tail (bb, hi]
                    for (int i = max(s0, bb + 1); i <= s1; ++i)
                        if (act && uint32_t(i - i0) <= span)
                            part[0] += sampleAt<ST, LY, INTERP, LARGE>(rs, kc.pitch, float(i), ray.s, gc, vol.data);
                }
                float sum = 0.0f;
#pragma unroll
                for (int u = 0; u < UNROLL; ++u) sum += part[u];
                acc += sum;
            }
    }
    if (pixel) { // raycasterprojector_interp.cl:70-71; accumulating launches add to the output
        const float r = kc.inc * acc / float(rp.rpp_x * rp.rpp_y);
        float& o = out[(size_t(vm) * rp.rows + y) * rp.channels + x];
        o = kc.accumulate ? o + r : r;
    }
}

template <Storage ST, Layout LY, bool INTERP, int UNROLL, bool LARGE = false>
__global__ __launch_bounds__(kBlock) void raycast_kernel(VolumeView vol, const RayGeom* __restrict__ geom,
                                                         RayParams rp, CastConst kc,
                                                         float* __restrict__ out)
{
    castTile<ST, LY, INTERP, UNROLL, LARGE>(vol, geom, rp, kc, out, xcdRemap(blockIdx.x, gridDim.x),
                                            threadIdx.x >> 6);
}

// ---------------------------------------------------------------------------------------------
// LDS-staged ray casting (interleaved volume, trilinear, one ray per pixel)
//
// The 256 rays of a workgroup walk their sample indices together in chunks of kc.chunk samples.
// For each chunk the workgroup takes the exact box of texels that its rays' taps touch (sample
// positions are monotone in i, so the chunk's end samples bound it), copies that box from HBM into
// LDS with direct buffer-to-LDS loads, and samples from LDS. A gather from LDS costs a fraction
// of a gather through the vector L1, whose scattered cache lines bound the unstaged kernel. A
// chunk whose box does not fit the LDS budget samples from HBM as the unstaged kernel does.
// Samples, their order within a ray and the FP32 arithmetic of each sample are the same as in
// castTile. The box holds the volume's own 8-byte texels: FP32 (v, dz) (two 16-byte LDS reads per
// sample, rows y and y+1) or FP16/BF16 (y,z)-quads (one 16-byte read).

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Minimum of each 16-bit half (v_pk_min_u16).
__device__ __forceinline__ uint32_t pkMin(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}

// Wave-wide pkMin, as waveReduce. All 64 lanes must be active.
__device__ __forceinline__ uint32_t wavePkMin(uint32_t v)
{
    v = pkMin(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0xB1, 0xF, 0xF, false)));
    v = pkMin(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x4E, 0xF, 0xF, false)));
    v = pkMin(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x124, 0xF, 0xF, false)));
    v = pkMin(v, uint32_t(__builtin_amdgcn_update_dpp(int(v), int(v), 0x128, 0xF, 0xF, false)));
    const uint32_t r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const uint32_t r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return __builtin_amdgcn_readfirstlane(pkMin(pkMin(r0, r1), pkMin(r2, r3)));
}

typedef __attribute__((address_space(3))) void lds_void;

typedef float lds_f2 __attribute__((ext_vector_type(2)));
typedef uint32_t lds_u2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) const lds_f2 lds_float2;
typedef __attribute__((address_space(3))) const lds_u2 lds_uint2;

// LDS byte address of a pointer into shared memory
__device__ __forceinline__ uint32_t ldsAddr(const void* p)
{
    return uint32_t(uintptr_t((__attribute__((address_space(3))) const char*)p));
}

// Trilinear sample from the LDS box ((v, dz) texels) at LDS byte address `at`, with rows px8 bytes
// apart; the same arithmetic in the same order as Sampler<F32, Interleaved>.
__device__ __forceinline__ float triLds(uint32_t at, uint32_t px8, float wx, float wy, float wz)
{
    const lds_float2* p = (const lds_float2*)uintptr_t(at);
    const lds_float2* q = (const lds_float2*)uintptr_t(at + px8);
    const lds_f2 a0 = p[0], a1 = p[1];
    const lds_f2 b0 = q[0], b1 = q[1];
    return triDeltaZ(make_float2(a0.x, a0.y), make_float2(a1.x, a1.y), make_float2(b0.x, b0.y),
                     make_float2(b1.x, b1.y), wx, wy, wz);
}

// a * b + c on the low 24 bits of a and b as one v_mad_u32_u24 (b uniform); written out because
// the compiler otherwise splits two of them into two multiplies and a three-way add
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}

// FP16/BF16 quads: texels x and x+1 at LDS byte address `at`; the same lerps as QuadHalf::tri.
template <Storage ST>
__device__ __forceinline__ float quadLds(uint32_t at, float wx, float wy, float wz)
{
    const lds_uint2* p = (const lds_uint2*)uintptr_t(at);
    const lds_u2 a = p[0], b = p[1];
    return quadTri<ST>(a.x, a.y, b.x, b.y, wx, wy, wz);
}

// Samples [wlo, whi] of a chunk for a lane whose own range is [a, e] (`live` if non-empty), with
// [wa, we] the range every live lane of the wave shares: head and tail under the exec mask, the
// shared body unmasked and unrolled, UNROLL samples a step. With UNROLL = 3 or 4 the step's last
// sample comes from `g` instead of `f` (the same value, read another way); `f` everywhere else.
template <int UNROLL, typename F, typename G>
__device__ __forceinline__ void sampleRange(int wlo, int wa, int we, int whi, int a, int e, bool live,
                                            float* part, F&& f, G&& g)
{
    const uint32_t span = uint32_t(e - a);
    for (int i = wlo; i < wa; ++i)
        if (live && uint32_t(i - a) <= span) part[0] += f(float(i));
    if (live) {
        int i = wa;
        float fi = float(wa);
        for (; i + UNROLL - 1 <= we; i += UNROLL, fi += float(UNROLL)) {
#pragma unroll
            for (int u = 0; u < UNROLL; ++u)
                part[u] += (UNROLL >= 3 && u == UNROLL - 1) ? g(fi + float(u)) : f(fi + float(u));
        }
        for (; i <= we; ++i, fi += 1.0f) part[0] += f(fi);
    }
    for (int i = we + 1; i <= whi; ++i)
        if (live && uint32_t(i - a) <= span) part[0] += f(float(i));
}

// True if every voxel that a tap of a chunk with texel box [o, o + nb) can read is exactly 0: the
// box's voxels with their z + 1 neighbours (held in the texels) and y + 1 (the quads' second row).
// Every sample of such a chunk is exactly 0, so skipping it leaves each ray's sum unchanged.
__device__ __forceinline__ bool boxEmpty(const VolumeView& vol, const int o[3], const int nb[3])
{
    if (!vol.occ) return false;
    const uint32_t lo[3] = { uint32_t(o[0]), uint32_t(o[1]), uint32_t(o[2]) };
    const uint32_t hi[3] = { uint32_t(o[0] + nb[0] - 1), uint32_t(o[1] + nb[1]), uint32_t(o[2] + nb[2]) };
    return bricksEmpty(vol, lo, hi);
}

// Texel box [o, o + nb) per axis of the taps of samples s0..s1 (0 <= s0 <= s1) of rays whose
// steps lie in [smin, smax] per axis. All rays of a (view, module) start from the same point c, so
// along each axis the positions fma(i, s, c) lie between fma(i, smin, c) and fma(i, smax, c) (the
// FMA is monotonic in s for i >= 0, and in i), extremes at i = s0 or s1; positions outside
// [blo, nhi) are never taken. A tap of position q reads texels floor(q) .. floor(q) + 1 along x,
// and along y for FP32 (the texels carry z + 1, the FP16/BF16 quads y + 1 too). False if the
// positions miss [blo, nhi) on some axis.
template <Storage ST>
__device__ __forceinline__ bool boxOf(int s0, int s1, const float smin[3], const float smax[3], const float c[3],
                                      float blo, const float nhi[3], int o[3], int nb[3])
{
    bool any = true;
    const float f0 = float(s0), f1 = float(s1);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float a0 = fmaf(f0, smin[k], c[k]), a1 = fmaf(f1, smin[k], c[k]);
        const float b0 = fmaf(f0, smax[k], c[k]), b1 = fmaf(f1, smax[k], c[k]);
        const float ql = fmaxf(fminf(a0, a1), blo), qh = fminf(fmaxf(b0, b1), nhi[k]);
        const int extra = (k == 0 || (k == 1 && ST == Storage::F32)) ? 1 : 0;
        any = any && ql <= qh;
        o[k] = int(ql);
        nb[k] = min(int(qh), int(nhi[k]) - 1) + extra - o[k] + 1;
    }
    return any;
}

// LDS row pitch of a staged box nb[0] texels wide: even rows (a 16-byte load never straddles two
// rows), not a multiple of 16 texels (rows would share LDS banks).
__device__ __forceinline__ int boxPitch(int nbx)
{
    int px = (nbx + 1) & ~1;
    if ((px & 15) == 0) px += 2;
    return px;
}

// A box the staging indexes can address (a large volume stages through a window of the box's
// slices, addressed with 32 bits).
template <bool LARGE>
__device__ __forceinline__ bool boxSmall(const int nb[3], const CastConst& kc)
{
    return nb[0] >= 1 && nb[0] <= 256 && nb[1] >= 1 && nb[1] <= 256 && nb[2] >= 1 && nb[2] <= 256 &&
           (!LARGE || uint64_t(nb[2] + 1) * kc.pitch.tz * 8u < 0xffff0000ull);
}

// A box that is staged: addressable and within the LDS box.
template <bool LARGE>
__device__ __forceinline__ bool boxFits(const int nb[3], const CastConst& kc)
{
    if (!boxSmall<LARGE>(nb, kc)) return false;
    const int n_ins = (boxPitch(nb[0]) * nb[1] * nb[2] + 127) >> 7;
    return n_ins * 128 <= int(kc.lds_texels);
}

// ints per wave in the staged caster's exchange array
constexpr int kRed = 10;

template <Storage ST, int UNROLL, bool LARGE = false>
__device__ __forceinline__ void castTileStaged(const VolumeView& vol, const RayGeom* __restrict__ geom,
                                               const RayParams& rp, const CastConst& kc,
                                               float* __restrict__ out, uint32_t b, uint32_t wave,
                                               float2* __restrict__ box, int (*red)[4][kRed])
{
    const uint32_t vm = b / kc.tiles_per_vm;
    const uint32_t t = b - vm * kc.tiles_per_vm;
    const uint32_t ty = t / kc.tiles_x, tx = t - ty * kc.tiles_x;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wy = wave / kc.bwx, wx = wave - wy * kc.bwx;
    const uint32_t lx = gatherBits(lane, kc.lane_x), ly = gatherBits(lane, kc.lane_y);
    const uint32_t x = (tx * kc.bwx + wx) * kc.wave_w + lx;
    const uint32_t y = (ty * kc.bwy + wy) * kc.wave_h + ly;
    const bool pixel = x < rp.channels && y < rp.rows;

    // the geometry record read as constant memory (scalar loads; in the batch kernel the record's
    // address comes from the item table, where the compiler would otherwise use vector loads)
    typedef __attribute__((address_space(4))) const RayGeom geom_const;
    const geom_const& g = *(const geom_const*)(geom + vm);
    const float gc[3] = { g.c[0], g.c[1], g.c[2] };
    const __amdgpu_buffer_rsrc_t rs = makeRsrc(vol.data, vol.bytes); // unused by LARGE gathers
    float acc = 0.0f;
    if (g.valid) { // uniform over the workgroup: one (view, module) per workgroup
        // one ray per pixel: the sub-ray coordinate x - 1/2 + 1/2 is x exactly
        Ray ray = setupRay(g, float(x), float(y), kc.inc, kc.vs, kc.nhi, kc.blo);
        if (!pixel) ray.hit = false;
        tighten(ray, gc, kc.blo, kc.nhi);
        const bool act = pixel && ray.i0 <= ray.i1;
        // sample ranges, once per tile: the union [r_lo, r_hi] and the intersection [r_alo, r_ahi]
        // of the wave's rays, and the workgroup's union. Clipped to a chunk, the union covers
        // every live lane's samples and every live lane samples all of the intersection.
        const int r_lo = waveReduce<false>(act ? ray.i0 : INT_MAX);
        const int r_hi = waveReduce<true>(act ? ray.i1 : INT_MIN);
        const int r_alo = waveReduce<true>(act ? ray.i0 : INT_MIN);
        const int r_ahi = waveReduce<false>(act ? ray.i1 : INT_MAX);
        // the extremes of the active rays' steps per axis, over the wave and the workgroup (the
        // texel boxes of the chunks follow from them, boxOf)
        float smin[3], smax[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            smin[k] = waveMinF(act ? ray.s[k] : FLT_MAX);
            smax[k] = waveMaxF(act ? ray.s[k] : -FLT_MAX);
        }
        if (lane == 0) {
            red[2][wave][0] = r_lo;
            red[2][wave][1] = r_hi;
            red[2][wave][2] = r_alo;
            red[2][wave][3] = r_ahi;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                red[2][wave][4 + k] = __float_as_int(smin[k]);
                red[2][wave][7 + k] = __float_as_int(smax[k]);
            }
        }
        __syncthreads();
        const int blo = min(min(red[2][0][0], red[2][1][0]), min(red[2][2][0], red[2][3][0]));
        const int bhi = max(max(red[2][0][1], red[2][1][1]), max(red[2][2][1], red[2][3][1]));
        // the samples every active ray of the workgroup takes: in chunks inside this range the
        // boxOf boxes are exactly the boxes of the lanes' taps (no lane's range starts or ends
        // there); elsewhere the lanes' own boxes are reduced
        const int g_alo = max(max(red[2][0][2], red[2][1][2]), max(red[2][2][2], red[2][3][2]));
        const int g_ahi = min(min(red[2][0][3], red[2][1][3]), min(red[2][2][3], red[2][3][3]));
        float gmin[3], gmax[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gmin[k] = fminf(fminf(__int_as_float(red[2][0][4 + k]), __int_as_float(red[2][1][4 + k])),
                            fminf(__int_as_float(red[2][2][4 + k]), __int_as_float(red[2][3][4 + k])));
            gmax[k] = fmaxf(fmaxf(__int_as_float(red[2][0][7 + k]), __int_as_float(red[2][1][7 + k])),
                            fmaxf(__int_as_float(red[2][2][7 + k]), __int_as_float(red[2][3][7 + k])));
        }
        const int K = int(kc.chunk);
        float part[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) part[u] = 0.0f;
        int parity = 0; // half of red[0..1] the next box exchange writes
        for (int c0 = blo; c0 <= bhi; c0 += K) {
            const int c1 = min(c0 + K - 1, bhi);
            const int a = act ? max(c0, ray.i0) : INT_MAX;
            const int e = act ? min(c1, ray.i1) : INT_MIN;
            const bool live = a <= e;
            // the wave's sample range [wlo, whi] in this chunk and the part [wa, we] all its live
            // lanes share (empty: all masked)
            int wlo = max(c0, r_lo), whi = min(c1, r_hi), wa = max(c0, r_alo), we = min(c1, r_ahi);
            if (wlo > whi) { wlo = INT_MAX; whi = INT_MIN; wa = INT_MIN + 1; we = INT_MIN; }
            else if (wa > we) { wa = whi + 1; we = whi; }
            // the workgroup's texel box for the chunk (uniform): a chunk whose samples miss the
            // support, or whose box holds only zero voxels, adds exactly 0 to every ray
            int o[3] = { 0, 0, 0 }, nb[3] = { 0, 0, 0 };
            uint32_t pk[3] = { ~0u, ~0u, ~0u };
            bool synced = false; // a workgroup barrier since the previous chunk's sampling
            // the boxOf box: exact where no lane's range starts or ends in the chunk, else a
            // superset, which serves unless it overflows the LDS box (uniform). Where rays enter
            // or leave the volume at very different samples (oblique rays at the silhouette,
            // coarse detector pixels) the superset spans the positions the late rays would have
            // outside the volume along the other axes; the lanes' own boxes are then reduced.
            // (64 projections of 64^3-256^3 volumes on 257^2-766^2 panels: 628 -> 585 ms; the
            // 256^3 / 1024 x 768 scan 40.1 -> 40.6 ms)
            bool exact = false;
            if (kc.lds_texels) {
                if (!boxOf<ST>(c0, c1, gmin, gmax, gc, kc.blo, kc.nhi, o, nb)) continue;
                exact = !(c0 >= g_alo && c1 <= g_ahi) && !boxFits<LARGE>(nb, kc);
            }
            if (kc.lds_texels && !exact) {
                if (boxEmpty(vol, o, nb)) continue;
            } else if (kc.lds_texels) {
                // Texel box of this lane's taps in the chunk, [lo, hi] per axis (texels carry z+1,
                // and y+1 for the FP16/BF16 quads, so only x and for FP32 y need one more).
                // Reduced over the wave as packed 16-bit minima (a maximum m as the minimum of
                // 0xffff - m): padded sizes stay below 65535 when a box is staged. 0xffff marks
                // "no live lane".
                {
                    const int as = live ? a : c0, es = live ? e : c0; // in range for every lane
                    uint32_t lo[3], hi[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float qa = fmaf(float(as), ray.s[k], gc[k]), qe = fmaf(float(es), ray.s[k], gc[k]);
                        const uint32_t extra = (k == 0 || (k == 1 && ST == Storage::F32)) ? 1u : 0u;
                        lo[k] = uint32_t(int(fminf(qa, qe))) & 0xffffu;
                        hi[k] = 0xffffu - ((uint32_t(int(fmaxf(qa, qe))) + extra) & 0xffffu);
                    }
                    pk[0] = live ? lo[0] | lo[1] << 16 : ~0u;
                    pk[1] = live ? hi[0] | hi[1] << 16 : ~0u;
                    pk[2] = live ? lo[2] | hi[2] << 16 : ~0u;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) pk[k] = wavePkMin(pk[k]);
                if (lane == 0)
#pragma unroll
                    for (int k = 0; k < 3; ++k) red[parity][wave][k] = int(pk[k]);
                // the box of the previous chunk is no longer read; red[parity] is complete
                __syncthreads();
                synced = true;
                uint32_t bpk[3];
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    bpk[k] = pkMin(pkMin(uint32_t(red[parity][0][k]), uint32_t(red[parity][1][k])),
                                   pkMin(uint32_t(red[parity][2][k]), uint32_t(red[parity][3][k])));
                parity ^= 1; // the next exchange writes the other half (its readers passed this barrier)
                o[0] = int(bpk[0] & 0xffffu);
                o[1] = int(bpk[0] >> 16);
                o[2] = int(bpk[2] & 0xffffu);
                nb[0] = int(0xffffu - (bpk[1] & 0xffffu)) - o[0] + 1;
                nb[1] = int(0xffffu - (bpk[1] >> 16)) - o[1] + 1;
                nb[2] = int(0xffffu - (bpk[2] >> 16)) - o[2] + 1;
                // 0xffff: no lane of the workgroup samples in this chunk
                if (o[0] == 0xffff || boxEmpty(vol, o, nb)) continue;
            }
            // so does this wave's part of a chunk whose own box is empty: the wave then only helps
            // to stage the box for the others (uniform over the wave)
            bool taken = live;
            if (kc.lds_texels && vol.occ && wlo <= whi) {
                int wo[3], wn[3];
                bool empty;
                if (!exact) {
                    empty = !boxOf<ST>(wlo, whi, smin, smax, gc, kc.blo, kc.nhi, wo, wn) || boxEmpty(vol, wo, wn);
                } else {
                    wo[0] = int(pk[0] & 0xffffu);
                    wo[1] = int(pk[0] >> 16);
                    wo[2] = int(pk[2] & 0xffffu);
                    wn[0] = int(0xffffu - (pk[1] & 0xffffu)) - wo[0] + 1;
                    wn[1] = int(0xffffu - (pk[1] >> 16)) - wo[1] + 1;
                    wn[2] = int(0xffffu - (pk[2] >> 16)) - wo[2] + 1;
                    empty = wo[0] != 0xffff && boxEmpty(vol, wo, wn);
                }
                if (empty) {
                    wlo = INT_MAX; whi = INT_MIN; wa = INT_MIN + 1; we = INT_MIN;
                    taken = false;
                }
            }
            if (kc.tally) tallyWave(kc.tally, taken ? uint32_t(e - a + 1) : 0u);
            const bool small = kc.lds_texels && boxSmall<LARGE>(nb, kc);
            const int px = boxPitch(nb[0]);
            const int pz = small ? px * nb[1] : 0;
            const int n_ins = small ? (pz * nb[2] + 127) >> 7 : 0; // 16-byte loads of 64 lanes = 128 texels
            if (small && n_ins * 128 <= int(kc.lds_texels)) {
                if (!synced) __syncthreads(); // every wave is done with the previous chunk's box
                // texel pairs t = 128 k + 2 lane of the box for k = wave, wave + 4, ...: position
                // (cx, cy, cz) of the first by division, then steps of 512 texels with carries
                int cx, cy, cz, dx, dy, dz;
                {
                    // approximate reciprocals (1 ulp): every quotient below is < 512, so the
                    // truncated products are off by at most one, which the carries correct
                    const float rpx = __builtin_amdgcn_rcpf(float(px));
                    const float rby = __builtin_amdgcn_rcpf(float(nb[1]));
                    const int t0 = int(wave) * 128 + 2 * int(lane);
                    int row = int(float(t0) * rpx);
                    cx = t0 - row * px;
                    if (cx < 0) { --row; cx += px; } else if (cx >= px) { ++row; cx -= px; }
                    cz = int(float(row) * rby);
                    cy = row - cz * nb[1];
                    if (cy < 0) { --cz; cy += nb[1]; } else if (cy >= nb[1]) { ++cz; cy -= nb[1]; }
                    int drow = int(512.0f * rpx);
                    dx = 512 - drow * px;
                    if (dx < 0) { --drow; dx += px; } else if (dx >= px) { ++drow; dx -= px; }
                    dz = int(float(drow) * rby);
                    dy = drow - dz * nb[1];
                    if (dy < 0) { --dz; dy += nb[1]; } else if (dy >= nb[1]) { ++dz; dy -= nb[1]; }
                }
                // the staging loads' buffer: the volume, or for a large volume the window that starts at
                // the box's first slice (its slices then count from 0)
                __amdgpu_buffer_rsrc_t win = rs;
                uint32_t z0 = uint32_t(o[2]);
                if constexpr (LARGE) {
                    const uint64_t slice = uint64_t(kc.pitch.tz) * 8u;
                    const uint64_t total = uint64_t(vol.rows) * vol.row_bytes;
                    // o is the workgroup's box origin: uniform, so the window is a scalar resource
                    const uint64_t first = uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(o[2]))) * slice;
                    const uint64_t rest = total > first ? total - first : 0;
                    win = makeRsrc(static_cast<const char*>(vol.data) + first,
                                   uint32_t(rest < 0xffff0000ull ? rest : 0xffff0000ull));
                    z0 = 0;
                }
                for (int k = int(wave); k < n_ins; k += 4) {
                    // texels past the box land in padding of the LDS box; past the volume they read 0
                    const uint32_t src = (uint32_t(o[0] + cx) + __umul24(uint32_t(o[1] + cy), kc.pitch.ty) +
                                          (z0 + uint32_t(cz)) * kc.pitch.tz) << 3;
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(win, (lds_void*)(box + k * 128), 16, src, 0, 0, 0);
                    cx += dx;
                    cy += dy;
                    cz += dz;
                    if (cx >= px) { cx -= px; ++cy; }
                    if (cy >= nb[1]) { cy -= nb[1]; ++cz; }
                }
                __syncthreads(); // box loaded (the compiler waits for the loads before the barrier)
                // byte offset of texel (ix, iy, iz) in the box: pitches in bytes, so that the address
                // is one shift-add and two 24-bit multiply-adds (mod 2^32; the result is in range)
                const uint32_t px8 = uint32_t(px) * 8u, pz8 = uint32_t(pz) * 8u;
                const uint32_t base8 = (uint32_t(o[0]) + uint32_t(o[1]) * uint32_t(px) + uint32_t(o[2]) * uint32_t(pz)) * 8u;
                // LDS address of texel (ix, iy, iz): ix * 8 + iy * px8 + iz * pz8 + bias (mod 2^32)
                const uint32_t bias = ldsAddr(box) - base8;
                sampleRange<UNROLL>(wlo, wa, we, whi, a, e, live, part, [&](float fi) {
                    const float qx = fmaf(fi, ray.s[0], gc[0]);
                    const float qy = fmaf(fi, ray.s[1], gc[1]);
                    const float qz = fmaf(fi, ray.s[2], gc[2]);
                    const uint32_t at = mad24(uint32_t(int(qz)), pz8,
                                              mad24(uint32_t(int(qy)), px8, (uint32_t(int(qx)) << 3) + bias));
                    const float wx = __builtin_amdgcn_fractf(qx), wy = __builtin_amdgcn_fractf(qy);
                    const float wz = __builtin_amdgcn_fractf(qz);
                    if constexpr (ST == Storage::F32) return triLds(at, px8, wx, wy, wz);
                    else return quadLds<ST>(at, wx, wy, wz);
                }, [&](float fi) {
                    // with unroll 3 (4) a third (a quarter) of the shared body gathers through the
                    // vector L1 instead (the box was just read from there): the LDS pipe and the
                    // texture path then share the taps. 256^3 / 1024 x 768, 90 views: a third
                    // 19.0 ms, a quarter 19.5, half 21.1, none 22.7
                    return sampleAt<ST, Layout::Interleaved, true, LARGE>(rs, kc.pitch, fi, ray.s, gc, vol.data);
                });
            } else {
                auto gather = [&](float fi) {
                    return sampleAt<ST, Layout::Interleaved, true, LARGE>(rs, kc.pitch, fi, ray.s, gc, vol.data);
                };
                sampleRange<UNROLL>(wlo, wa, we, whi, a, e, live, part, gather, gather);
            }
            // two-level summation: the chunk's partial sums (at most kc.chunk samples) are added
            // to the ray's total once per chunk, which keeps the FP32 rounding of long rays small
            float chunk_sum = 0.0f;
#pragma unroll
            for (int u = 0; u < UNROLL; ++u) {
                chunk_sum += part[u];
                part[u] = 0.0f;
            }
            acc += chunk_sum;
        }
    }
    if (pixel) { // raycasterprojector_interp.cl:70-71; accumulating launches add to the output
        const float r = kc.inc * acc / float(rp.rpp_x * rp.rpp_y);
        float& o = out[(size_t(vm) * rp.rows + y) * rp.channels + x];
        o = kc.accumulate ? o + r : r;
    }
}

template <Storage ST, int UNROLL, bool LARGE = false>
__global__ __launch_bounds__(kBlock) void staged_kernel(VolumeView vol, const RayGeom* __restrict__ geom,
                                                        RayParams rp, CastConst kc, float* __restrict__ out)
{
    extern __shared__ __attribute__((aligned(16))) float2 s_box[];
    __shared__ int s_red[3][4][kRed];
    static_assert(sizeof(s_red) <= kStaticLds, "static LDS beyond the reserve");
    castTileStaged<ST, UNROLL, LARGE>(vol, geom, rp, kc, out, xcdRemap(blockIdx.x, gridDim.x), threadIdx.x >> 6,
                                      s_box, s_red);
}

// Persistent batch kernel: a fixed grid of workgroups drains a queue of tiles that spans many
// projections (volumes, geometries and detector sizes of their own). A tile is a workgroup's
// 2x2 wavefronts of neighbouring pixels, so the four waves on a CU gather from the same patch
// of the volume. The tile space is cut into one contiguous range per XCD; a workgroup takes tiles
// from its own XCD's range first (neighbouring tiles of one view then share that XCD's L2) and
// then from the others. `stride` > 1 visits every stride-th run of kSampleRun consecutive tiles
// (used by the autotuner to time a sample of the work; runs keep neighbouring tiles together, as a
// full launch has them, so the sample's cache behaviour is the launch's).
__device__ __forceinline__ unsigned long long sampledTile(unsigned long long t, uint32_t stride)
{
    return (t / kSampleRun) * stride * kSampleRun + t % kSampleRun;
}

// Takes the next tile for workgroup-home XCD `home` (own XCD's range first, then the others):
// sets *t to the tile (~0 once every range is drained) and *k to its item. Called by all threads.
__device__ __forceinline__ void grabTile(const BatchItem* __restrict__ items, uint32_t n_items,
                                         BatchQueue* __restrict__ q, uint32_t stride, uint32_t home,
                                         unsigned long long* s_tile, uint32_t* s_item,
                                         unsigned long long* t, uint32_t* k)
{
    if (threadIdx.x == 0) {
        unsigned long long got = ~0ull;
        for (uint32_t j = 0; j < kNumXcd && got == ~0ull; ++j) {
            const uint32_t x = (home + j) % kNumXcd;
            if (__hip_atomic_load(&q->next[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= q->end[x]) continue;
            const unsigned long long tt = atomicAdd(&q->next[x], 1ull);
            if (tt < q->end[x]) got = tt;
        }
        uint32_t lo = 0;
        if (got != ~0ull) {
            const unsigned long long tt = sampledTile(got, stride);
            uint32_t hi = n_items; // last item with tile_begin <= tt
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (items[mid].tile_begin <= tt) lo = mid;
                else hi = mid;
            }
        }
        *s_tile = got;
        *s_item = lo;
    }
    __syncthreads();
    // wave-uniform copies: the item's fields are then scalar loads
    const uint32_t tlo = __builtin_amdgcn_readfirstlane(uint32_t(*s_tile));
    const uint32_t thi = __builtin_amdgcn_readfirstlane(uint32_t(*s_tile >> 32));
    *k = __builtin_amdgcn_readfirstlane(*s_item);
    __syncthreads();
    *t = (unsigned long long)thi << 32 | tlo;
}

template <Storage ST, Layout LY, bool INTERP, int UNROLL>
__global__ __launch_bounds__(kBlock) void batch_kernel(const BatchItem* __restrict__ items, uint32_t n_items,
                                                       BatchQueue* __restrict__ q, uint32_t stride)
{
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_item;
    const uint32_t home = blockIdx.x % kNumXcd;
    for (;;) {
        unsigned long long t;
        uint32_t k;
        grabTile(items, n_items, q, stride, home, &s_tile, &s_item, &t, &k);
        if (t == ~0ull) break; // every workgroup exits once all ranges are drained
        const BatchItem& it = items[k];
        castTile<ST, LY, INTERP, UNROLL>(it.vol, it.geom, it.rp, it.kc, it.out,
                                         uint32_t(sampledTile(t, stride) - it.tile_begin), threadIdx.x >> 6);
    }
}

// The same queue over interleaved trilinear items that all have kc.chunk > 0, cast with the
// LDS-staged caster (kc.lds_texels must not exceed the launch's dynamic LDS).
template <Storage ST, int UNROLL>
__global__ __launch_bounds__(kBlock) void batch_staged_kernel(const BatchItem* __restrict__ items,
                                                              uint32_t n_items, BatchQueue* __restrict__ q,
                                                              uint32_t stride)
{
    extern __shared__ __attribute__((aligned(16))) float2 s_box[];
    __shared__ int s_red[3][4][kRed];
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_item;
    static_assert(sizeof(s_red) + 16 <= kStaticLds, "static LDS beyond the reserve");
    const uint32_t home = blockIdx.x % kNumXcd;
    for (;;) {
        unsigned long long t;
        uint32_t k;
        grabTile(items, n_items, q, stride, home, &s_tile, &s_item, &t, &k);
        if (t == ~0ull) break;
        const BatchItem& it = items[k];
        castTileStaged<ST, UNROLL>(it.vol, it.geom, it.rp, it.kc, it.out, uint32_t(sampledTile(t, stride) - it.tile_begin),
                                   threadIdx.x >> 6, s_box, s_red);
    }
}

__global__ void queue_init_kernel(BatchQueue* q, unsigned long long total)
{
    const uint32_t x = threadIdx.x;
    if (x < kNumXcd) {
        const unsigned long long per = total / kNumXcd, rem = total % kNumXcd;
        const unsigned long long begin = x * per + (x < rem ? x : rem);
        q->next[x] = begin;
        q->end[x] = begin + per + (x < rem ? 1 : 0);
    }
}

__global__ __launch_bounds__(kBlock) void count_kernel(const RayGeom* __restrict__ geom, RayParams rp,
                                                       CastConst kc, unsigned long long* count)
{
    __shared__ unsigned long long block_sum;
    if (threadIdx.x == 0) block_sum = 0;
    __syncthreads();
    const size_t pix = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t per_vm = size_t(rp.channels) * rp.rows;
    unsigned long long mine = 0;
    if (pix < per_vm * rp.views * rp.modules) {
        const uint32_t vm = uint32_t(pix / per_vm);
        const uint32_t rem = uint32_t(pix - size_t(vm) * per_vm);
        const uint32_t y = rem / rp.channels, x = rem - y * rp.channels;
        const RayGeom g = geom[vm];
        const float ipx = 1.0f / float(rp.rpp_x), ipy = 1.0f / float(rp.rpp_y);
        if (g.valid)
            for (uint32_t ra = 0; ra < rp.rpp_x; ++ra)
                for (uint32_t rb = 0; rb < rp.rpp_y; ++rb) {
                    const Ray ray = setupRay(g, subRayCoord(x, ipx, ra), subRayCoord(y, ipy, rb), kc.inc,
                                             kc.vs, kc.nhi, kc.blo);
                    if (ray.hit && ray.i1 >= ray.i0) mine += (unsigned long long)(ray.i1 - ray.i0 + 1);
                }
    }
    if (mine) atomicAdd(&block_sum, mine);
    __syncthreads();
    if (threadIdx.x == 0 && block_sum) atomicAdd(count, block_sum);
}

// ---------------------------------------------------------------------------------------------
// Mean of k projections of n values each (contiguous), in the extinction domain or, for
// intensity-domain averaging, as -ln(mean(exp(-e))) (ArealFocalSpotExtension,
// arealfocalspotextension.cpp:82-166, with initial intensity 1).

__global__ __launch_bounds__(kBlock) void average_kernel(const float* __restrict__ in, uint32_t k, size_t n,
                                                         int intensity, float* __restrict__ out)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        float s = 0.0f;
        for (uint32_t j = 0; j < k; ++j) {
            const float e = in[size_t(j) * n + i];
            s += intensity ? expf(-e) : e;
        }
        s /= float(k);
        out[i] = intensity ? -logf(s) : s;
    }
}

// Spectral extinction from material projections (SpectralEffectsExtension::projectLinear,
// spectraleffectsextension.cpp:232-300, singleBinIntensityLinear :330-352, and the FP32 intensity
// transforms of singleviewdata.cpp:135-156): for a pixel of view v with material projections p_m,
//   e_b = sum_m p_m mu[b][m],  s = sum_b (i[b][v] exp(-e_b)) r[b],  out = ln(I0[v] / s),
// the sums in material and bin order. Workgroup b covers pixels [256 (b % pb), ...) of view b / pb.
__global__ __launch_bounds__(kBlock) void spectral_kernel(const float* __restrict__ proj, uint32_t n_mat,
                                                          uint32_t views, size_t n, uint32_t pb,
                                                          const float* __restrict__ mu,
                                                          const float* __restrict__ inten,
                                                          const float* __restrict__ resp, uint32_t n_bins,
                                                          const float* __restrict__ total, float* __restrict__ out)
{
    const uint32_t v = blockIdx.x / pb;
    const size_t i = size_t(blockIdx.x % pb) * kBlock + threadIdx.x;
    if (i >= n) return;
    float p[kMaxMaterials];
#pragma unroll
    for (uint32_t m = 0; m < kMaxMaterials; ++m)
        p[m] = m < n_mat ? proj[(size_t(m) * views + v) * n + i] : 0.0f;
    float s = 0.0f;
    for (uint32_t b = 0; b < n_bins; ++b) {
        float e = 0.0f;
#pragma unroll
        for (uint32_t m = 0; m < kMaxMaterials; ++m)
            if (m < n_mat) e += p[m] * mu[b * n_mat + m];
        s += inten[size_t(b) * views + v] * expf(-e) * resp[b];
    }
    out[size_t(v) * n + i] = logf(total[v] / s);
}

// Per-view linear combination out[v] = ca[v] a[v] + cb[v] b[v] (b may be NULL) of projections of
// n values per view (DynamicProjectorExtension on volumes linear in time).
__global__ __launch_bounds__(kBlock) void combine_views_kernel(const float* __restrict__ a, const float* __restrict__ ca,
                                                               const float* __restrict__ b, const float* __restrict__ cb,
                                                               size_t n, uint32_t pb, float* __restrict__ out)
{
    const uint32_t v = blockIdx.x / pb;
    const size_t i = size_t(blockIdx.x % pb) * kBlock + threadIdx.x;
    if (i >= n) return;
    const size_t k = size_t(v) * n + i;
    float r = ca[v] * a[k];
    if (b) r += cb[v] * b[k];
    out[k] = r;
}

// ---------------------------------------------------------------------------------------------
// Gather-ceiling probes. Diagnostics that measure the ceiling of the gather path; no projection
// runs through them.
//
//   probe_gather_kernel  the ray caster's loads without its interpolation: the same tiles, rays,
//                        sample ranges, wave-aligned loop and 16-byte texel loads over an
//                        interleaved volume, with the raw taps folded into an integer sink.
//   probe_stream_kernel  coalesced 16-byte loads from a 16 KiB window that stays in the vector
//                        L1: the cache-to-register data path without any gather.

typedef uint32_t probe_u4 __attribute__((ext_vector_type(4)));

template <Storage ST>
__device__ __forceinline__ uint32_t rawTaps(__amdgpu_buffer_rsrc_t r, const Pitch& p, float fi,
                                            const float s[3], const float c[3])
{
    const float qx = fmaf(fi, s[0], c[0]);
    const float qy = fmaf(fi, s[1], c[1]);
    const float qz = fmaf(fi, s[2], c[2]);
    const uint32_t idx = uint32_t(int(qx)) + __umul24(uint32_t(int(qy)), p.ty) +
                         __umul24(uint32_t(int(qz)), p.tz);
    const probe_u4 a = __builtin_bit_cast(probe_u4, __builtin_amdgcn_raw_buffer_load_b128(r, idx << 3, 0, 0));
    uint32_t h = a.x ^ a.y ^ a.z ^ a.w;
    if constexpr (ST == Storage::F32) { // second row of taps (y+1)
        const probe_u4 b = __builtin_bit_cast(probe_u4, __builtin_amdgcn_raw_buffer_load_b128(r, idx << 3, p.y, 0));
        h ^= b.x ^ b.y ^ b.z ^ b.w;
    }
    return h;
}

template <Storage ST>
__global__ __launch_bounds__(kBlock) void probe_gather_kernel(VolumeView vol, const RayGeom* __restrict__ geom,
                                                              RayParams rp, CastConst kc,
                                                              float* __restrict__ out)
{
    constexpr int U = 4;
    const uint32_t b = xcdRemap(blockIdx.x, gridDim.x), wave = threadIdx.x >> 6;
    const uint32_t vm = b / kc.tiles_per_vm;
    const uint32_t t = b - vm * kc.tiles_per_vm;
    const uint32_t ty = t / kc.tiles_x, tx = t - ty * kc.tiles_x;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wy = wave / kc.bwx, wx = wave - wy * kc.bwx;
    const uint32_t lx = gatherBits(lane, kc.lane_x), ly = gatherBits(lane, kc.lane_y);
    const uint32_t x = (tx * kc.bwx + wx) * kc.wave_w + lx;
    const uint32_t y = (ty * kc.bwy + wy) * kc.wave_h + ly;
    const bool pixel = x < rp.channels && y < rp.rows;
    const RayGeom g = geom[vm];
    const __amdgpu_buffer_rsrc_t rs = makeRsrc(vol.data, vol.bytes);

    uint32_t acc = 0;
    if (g.valid) {
        Ray ray = setupRay(g, float(x), float(y), kc.inc, kc.vs, kc.nhi, kc.blo); // one ray per pixel
        if (!pixel) ray.hit = false;
        tighten(ray, g.c, kc.blo, kc.nhi);
        const bool act = pixel && ray.i0 <= ray.i1;
        const int i0 = act ? ray.i0 : INT_MAX, i1 = act ? ray.i1 : INT_MIN;
        const int lo = waveMin(i0), hi = waveMax(i1);
        if (lo <= hi) {
            int a = waveMax(act ? i0 : INT_MIN), bb = waveMin(act ? i1 : INT_MAX);
            if (a > bb) { a = hi + 1; bb = hi; }
            const uint32_t span = uint32_t(i1 - i0);
            for (int i = lo; i < a; ++i)
                if (act && uint32_t(i - i0) <= span) acc += rawTaps<ST>(rs, kc.pitch, float(i), ray.s, g.c);
            if (act) {
                int i = a;
                float fi = float(a);
                for (; i + U - 1 <= bb; i += U, fi += float(U)) {
#pragma unroll
                    for (int u = 0; u < U; ++u) acc += rawTaps<ST>(rs, kc.pitch, fi + float(u), ray.s, g.c);
                }
                for (; i <= bb; ++i, fi += 1.0f) acc += rawTaps<ST>(rs, kc.pitch, fi, ray.s, g.c);
            }
            for (int i = bb + 1; i <= hi; ++i)
                if (act && uint32_t(i - i0) <= span) acc += rawTaps<ST>(rs, kc.pitch, float(i), ray.s, g.c);
        }
    }
    if (pixel) out[(size_t(vm) * rp.rows + y) * rp.channels + x] = float(acc & 0xffffu);
}

// `step` and `mask` are runtime values (4096 and 16383) so that no load address is loop-invariant
// at compile time and every load is issued.
__global__ __launch_bounds__(kBlock) void probe_stream_kernel(const void* __restrict__ src, uint32_t iters,
                                                              uint32_t step, uint32_t mask,
                                                              float* __restrict__ out)
{
    const __amdgpu_buffer_rsrc_t r = makeRsrc(src, mask + 1);
    uint32_t ofs = threadIdx.x * 16; // a wavefront reads 1 KiB contiguous per load
    uint32_t acc = 0;
    for (uint32_t k = 0; k < iters; k += 4) {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const probe_u4 a = __builtin_bit_cast(probe_u4, __builtin_amdgcn_raw_buffer_load_b128(r, ofs, 0, 0));
            acc += a.x ^ a.y ^ a.z ^ a.w;
            ofs = (ofs + step) & mask;
        }
    }
    out[size_t(blockIdx.x) * kBlock + threadIdx.x] = float(acc & 0xffffu);
}

// LDS read ceiling: each workgroup fills a 16 KiB LDS window, then every lane reads 16-byte
// elements, a wavefront 1 KiB contiguous per ds_read_b128 (no bank conflict), 16 reads per step at
// immediate offsets from a base that moves by a runtime `step` (0 here, so the compiler cannot
// hoist the reads out of the loop). Two three-operand adds per read keep the values live.
// PAIRS = true reads the same 16 bytes per lane as two adjacent 8-byte elements at an odd 8-byte
// index (ds_read2_b64 ... offset1:1), the way the staged ray caster reads a texel pair from its box.
template <bool PAIRS>
__global__ __launch_bounds__(kBlock) void probe_lds_kernel(uint32_t iters, uint32_t step, float* __restrict__ out)
{
    __shared__ __attribute__((aligned(16))) probe_u4 win[1024]; // 16 KiB
    for (uint32_t i = threadIdx.x; i < 1024; i += kBlock) win[i] = probe_u4{ i, i * 3u, i * 5u, i * 7u };
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    uint32_t acc0 = 0, acc1 = 0, base = 0;
    for (uint32_t k = 0; k < iters; k += 16) {
        if constexpr (PAIRS) {
            typedef uint32_t u2 __attribute__((ext_vector_type(2)));
            const u2* w = reinterpret_cast<const u2*>(win) + ((base + 2 * lane + 1) & 2047u);
#pragma unroll
            for (uint32_t u = 0; u < 15; ++u) { // 15 x 1 KiB: element 2047 is the last one read
                const u2 a = w[u * 128], b = w[u * 128 + 1];
                acc0 += a.x + a.y;
                acc1 += b.x + b.y;
            }
        } else {
            const probe_u4* w = win + ((base + lane) & 1023u);
#pragma unroll
            for (uint32_t u = 0; u < 16; ++u) {
                const probe_u4 a = w[u * 64];
                acc0 += a.x + a.y;
                acc1 += a.z + a.w;
            }
        }
        base += step;
    }
    out[size_t(blockIdx.x) * kBlock + threadIdx.x] = float((acc0 + acc1) & 0xffffu);
}

// ---------------------------------------------------------------------------------------------
// volume conversion: every padded texel (x, y, z) is written from a value function v(x, y, z)
// of the padded frame that returns 0 outside the interior.

__device__ __forceinline__ uint16_t toHalfBits(float f)
{
    return __builtin_bit_cast(uint16_t, _Float16(f)); // round to nearest even
}

__device__ __forceinline__ uint16_t toBf16Bits(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
    return uint16_t(u >> 16);
}

__device__ __forceinline__ uint16_t toBits(float v, Storage st)
{
    return st == Storage::F16 ? toHalfBits(v) : toBf16Bits(v);
}

__device__ __forceinline__ float fromBits(uint16_t h, Storage st)
{
    return st == Storage::F16 ? float(__builtin_bit_cast(_Float16, h)) : __uint_as_float(uint32_t(h) << 16);
}

// Value of voxel (x, y, z) of the padded frame stored in a view (first component of its texel).
__device__ __forceinline__ float viewValue(const VolumeView& s, size_t x, size_t y, size_t z)
{
    const size_t i = z * s.pitch_z + y * s.pitch_y + x;
    if (s.storage == Storage::F32)
        return static_cast<const float*>(s.data)[s.layout == Layout::Interleaved ? 2 * i : i];
    const uint16_t* h = static_cast<const uint16_t*>(s.data);
    return fromBits(h[s.layout == Layout::Interleaved ? 4 * i : i], s.storage);
}

template <class F>
__device__ __forceinline__ void writeTexel(void* dst, size_t i, size_t x, size_t y, size_t z,
                                           Storage st, Layout ly, const F& v)
{
    if (ly == Layout::Plain) {
        if (st == Storage::F32) static_cast<float*>(dst)[i] = v(x, y, z);
        else static_cast<uint16_t*>(dst)[i] = toBits(v(x, y, z), st);
    } else if (st == Storage::F32) {
        // (v(z), v(z+1) - v(z)): the difference a z-lerp would form, rounded once here
        float2 t;
        t.x = v(x, y, z);
        t.y = v(x, y, z + 1) - t.x;
        static_cast<float2*>(dst)[i] = t;
    } else {
        ushort4 t;
        t.x = toBits(v(x, y, z), st);
        t.y = toBits(v(x, y + 1, z), st);
        t.z = toBits(v(x, y, z + 1), st);
        t.w = toBits(v(x, y + 1, z + 1), st);
        static_cast<ushort4*>(dst)[i] = t;
    }
}

struct FromFloat { // source float array (nx, ny, nz) replicated f times per axis
    const float* src;
    uint32_t nx, ny, nz, f;
    __device__ float operator()(size_t x, size_t y, size_t z) const
    {
        if (x < 1 || y < 1 || z < 1 || x > size_t(nx) * f || y > size_t(ny) * f || z > size_t(nz) * f)
            return 0.0f;
        return src[((z - 1) / f * ny + (y - 1) / f) * nx + (x - 1) / f];
    }
};

struct FromView { // another padded volume replicated f times per axis
    VolumeView s;
    uint32_t f;
    __device__ float operator()(size_t x, size_t y, size_t z) const
    {
        if (x < 1 || y < 1 || z < 1 || x > size_t(s.n[0]) * f || y > size_t(s.n[1]) * f ||
            z > size_t(s.n[2]) * f)
            return 0.0f;
        return viewValue(s, (x - 1) / f + 1, (y - 1) / f + 1, (z - 1) / f + 1);
    }
};

struct FromMean { // one interior voxel holding sum / count
    const double* sum;
    double count;
    __device__ float operator()(size_t x, size_t y, size_t z) const
    {
        // the reference sums in float (raycasterprojector.cpp:603-606); the mean is rounded once
        return (x == 1 && y == 1 && z == 1) ? float(*sum / count) : 0.0f;
    }
};

template <class F>
__global__ void build_kernel(F v, uint32_t px, uint32_t py, uint32_t pz, Storage st, Layout ly,
                             void* __restrict__ dst)
{
    const size_t total = size_t(px) * py * pz;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += size_t(gridDim.x) * blockDim.x) {
        const size_t x = i % px, yz = i / px, y = yz % py, z = yz / py;
        writeTexel(dst, i, x, y, z, st, ly, v);
    }
}

__global__ void sum_kernel(VolumeView s, double* sum)
{
    __shared__ double part[kBlock];
    double acc = 0.0;
    const size_t nx = s.n[0], ny = s.n[1], total = size_t(s.n[0]) * s.n[1] * s.n[2];
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += size_t(gridDim.x) * blockDim.x) {
        const size_t x = i % nx, yz = i / nx, y = yz % ny, z = yz / ny;
        acc += viewValue(s, x + 1, y + 1, z + 1);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(sum, part[0]);
}

// Occupancy table: one thread per brick writes 1 if any voxel of its kOccBrick^3 cube of the padded
// grid is not exactly 0 (NaN included), at table entry (i + 1, j + 1, k + 1); the table's first
// row, column and slice are 0. Prefix sums along x, then y, then z make it a summed-area table.
__global__ void occ_flag_kernel(VolumeView v, uint32_t* __restrict__ occ, uint32_t ox, uint32_t oy, uint32_t oz)
{
    const size_t total = size_t(ox + 1) * (oy + 1) * (oz + 1);
    const size_t px = size_t(v.n[0]) + 2, py = size_t(v.n[1]) + 2, pz = size_t(v.n[2]) + 2;
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += size_t(gridDim.x) * blockDim.x) {
        const size_t i = t % (ox + 1), jk = t / (ox + 1), j = jk % (oy + 1), k = jk / (oy + 1);
        uint32_t flag = 0;
        if (i && j && k) {
            const size_t x0 = (i - 1) << kOccShift, y0 = (j - 1) << kOccShift, z0 = (k - 1) << kOccShift;
            for (size_t z = z0; z < z0 + kOccBrick && z < pz && !flag; ++z)
                for (size_t y = y0; y < y0 + kOccBrick && y < py && !flag; ++y)
                    for (size_t x = x0; x < x0 + kOccBrick && x < px; ++x)
                        if (!(viewValue(v, x, y, z) == 0.0f)) { flag = 1; break; }
        }
        occ[t] = flag;
    }
}

// Inclusive prefix sums along one axis of the (ox + 1) x (oy + 1) x (oz + 1) table: one thread per
// line; `stride` is the distance of neighbours on that axis, `len` the line's length, and the
// line's start is (line % inner) + (line / inner) * outer.
__global__ void occ_scan_kernel(uint32_t* __restrict__ occ, size_t lines, size_t inner, size_t outer, size_t stride,
                                uint32_t len)
{
    for (size_t l = size_t(blockIdx.x) * blockDim.x + threadIdx.x; l < lines; l += size_t(gridDim.x) * blockDim.x) {
        uint32_t* p = occ + (l % inner) + (l / inner) * outer;
        uint32_t acc = 0;
        for (uint32_t k = 0; k < len; ++k) {
            acc += p[k * stride];
            p[k * stride] = acc;
        }
    }
}

__global__ void zero_kernel(float* d, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x)
        d[i] = 0.0f;
}

uint32_t gridFor(size_t n)
{
    const size_t blocks = (n + kBlock - 1) / kBlock;
    return uint32_t(blocks < 65536 ? (blocks ? blocks : 1) : 65536);
}

} // namespace

// gatherBits code that moves the lane bits set in `mask` (in order) to the low bits
static uint32_t laneCode(uint32_t mask)
{
    uint32_t e = 0, run = 0, p = 0;
    for (uint32_t k = 0; k < 6;) {
        if (!((mask >> k) & 1u)) {
            ++k;
            continue;
        }
        const uint32_t lo = k;
        while (k < 6 && ((mask >> k) & 1u)) ++k;
        const uint32_t n = k - lo;
        e |= ((lo - p) | ((((1u << n) - 1u) << p) << 3)) << (9 * run);
        p += n;
        ++run; // at most three runs: the x bits of the quad layout are 0, 2.., 4..
    }
    return e;
}

// Lane bits that give a lane's x in a ww x wh wavefront tile (powers of two, ww * wh = 64): bit 0
// (within a quad), the low bits of the quad index within a 16-lane group, the low bits of the
// group index; all of the low bits for row-major tiles one pixel wide or tall.
static uint32_t laneXMask(uint32_t ww, uint32_t wh)
{
    auto lg = [](uint32_t v) { uint32_t k = 0; while ((2u << k) <= v) ++k; return k; };
    if (ww < 2 || wh < 2) return ww - 1u;
    const uint32_t g = ww < 8 ? ww : 8;
    return 1u | (((1u << lg(g / 2)) - 1u) << 2) | (((1u << lg(ww / g)) - 1u) << 4);
}

CastConst makeCastConst(const VolumeView* vol, const VolumeGeom& vg, const RayParams& rp,
                    const LaunchShape& sh)
{
    CastConst kc{};
    kc.wave_w = sh.wave_w;
    kc.wave_h = sh.wave_h;
    const uint32_t xm = laneXMask(sh.wave_w, sh.wave_h);
    kc.lane_x = laneCode(xm);
    kc.lane_y = laneCode(63u & ~xm);
    kc.bwx = sh.block_waves_x;
    kc.bwy = sh.block_waves_y;
    const uint32_t btw = sh.wave_w * sh.block_waves_x, bth = sh.wave_h * sh.block_waves_y;
    kc.tiles_x = (rp.channels + btw - 1) / btw;
    kc.tiles_y = (rp.rows + bth - 1) / bth;
    kc.tiles_per_vm = kc.tiles_x * kc.tiles_y;
    kc.inc = vg.increment_mm;
    for (int k = 0; k < 3; ++k) {
        kc.vs[k] = vg.voxel_mm[k];
        kc.nhi[k] = float(vg.n[k]) + 1.0f;
    }
    kc.blo = rp.interpolate ? 0.0f : 1.0f;
    kc.chunk = (vol && stageable(vol->storage, vol->layout, rp)) ? sh.chunk : 0u;
    // the staged kernel packs texel coordinates into 16 bits: larger volumes chunk their samples
    // but never stage a box
    const bool fits16 = vol && vol->n[0] + 2 < 0xffffu && vol->n[1] + 2 < 0xffffu && vol->n[2] + 2 < 0xffffu;
    kc.lds_texels = fits16 ? sh.lds_texels : 0u;
    if (vol) {
        const uint32_t eb = uint32_t(texelBytes(vol->storage, vol->layout));
        kc.pitch.ty = vol->pitch_y;
        kc.pitch.tz = vol->pitch_z;
        kc.pitch.py = vol->n[1] + 2;
        kc.pitch.y = vol->pitch_y * eb;
        kc.pitch.z = vol->pitch_z * eb;
        kc.pitch.yz = (vol->pitch_y + vol->pitch_z) * eb;
    }
    return kc;
}

namespace {

template <Storage ST, Layout LY, bool INTERP>
hipError_t launchCastU(const VolumeView& vol, const RayGeom* g, const RayParams& rp,
                       const CastConst& kc, uint32_t unroll, uint32_t blocks, uint32_t lds, float* out,
                       hipStream_t s)
{
    if constexpr (LY == Layout::Interleaved) {
        if (vol.large) {
            switch (unroll) {
            case 1: raycast_kernel<ST, LY, INTERP, 1, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
            case 3:
            case 4: raycast_kernel<ST, LY, INTERP, 4, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
            default: raycast_kernel<ST, LY, INTERP, 2, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
            }
            return hipGetLastError();
        }
    } else if (vol.large) {
        return hipErrorInvalidValue; // large volumes are interleaved
    }
    switch (unroll) {
    case 1: raycast_kernel<ST, LY, INTERP, 1><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    case 3: // the unstaged kernels run 3 as 4
    case 4: raycast_kernel<ST, LY, INTERP, 4><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    default: raycast_kernel<ST, LY, INTERP, 2><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    }
    return hipGetLastError();
}

template <Storage ST>
hipError_t launchStaged(const VolumeView& vol, const RayGeom* g, const RayParams& rp, const CastConst& kc,
                        uint32_t unroll, uint32_t blocks, uint32_t lds, float* out, hipStream_t s)
{
    if (vol.large) {
        switch (unroll) {
        case 1: staged_kernel<ST, 1, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
        case 3: staged_kernel<ST, 3, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
        case 4: staged_kernel<ST, 4, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
        default: staged_kernel<ST, 2, true><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
        }
        return hipGetLastError();
    }
    switch (unroll) {
    case 1: staged_kernel<ST, 1><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    case 3: staged_kernel<ST, 3><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    case 4: staged_kernel<ST, 4><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    default: staged_kernel<ST, 2><<<blocks, kBlock, lds, s>>>(vol, g, rp, kc, out); break;
    }
    return hipGetLastError();
}

template <Storage ST>
hipError_t launchCastI(const VolumeView& vol, const RayGeom* g, const RayParams& rp,
                       const CastConst& kc, uint32_t unroll, uint32_t blocks, uint32_t lds, float* out,
                       hipStream_t s)
{
    if (vol.layout == Layout::Interleaved)
        return rp.interpolate
                   ? launchCastU<ST, Layout::Interleaved, true>(vol, g, rp, kc, unroll, blocks, lds, out, s)
                   : launchCastU<ST, Layout::Interleaved, false>(vol, g, rp, kc, unroll, blocks, lds, out, s);
    return rp.interpolate ? launchCastU<ST, Layout::Plain, true>(vol, g, rp, kc, unroll, blocks, lds, out, s)
                          : launchCastU<ST, Layout::Plain, false>(vol, g, rp, kc, unroll, blocks, lds, out, s);
}

} // namespace

hipError_t launchProbeGather(const VolumeView& vol, const VolumeGeom& vg, const RayGeom* d_geom,
                             const RayParams& rp, float* d_out, hipStream_t s)
{
    if (vol.layout != Layout::Interleaved || !rp.interpolate || rp.rpp_x != 1 || rp.rpp_y != 1)
        return hipErrorInvalidValue;
    const CastConst kc = makeCastConst(&vol, vg, rp, LaunchShape{});
    const uint64_t blocks = uint64_t(kc.tiles_per_vm) * rp.views * rp.modules;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0xffffffffull) return hipErrorInvalidConfiguration;
    if (vol.storage == Storage::F32)
        probe_gather_kernel<Storage::F32><<<uint32_t(blocks), kBlock, 0, s>>>(vol, d_geom, rp, kc, d_out);
    else
        probe_gather_kernel<Storage::F16><<<uint32_t(blocks), kBlock, 0, s>>>(vol, d_geom, rp, kc, d_out);
    return hipGetLastError();
}

hipError_t launchAverage(const float* d_in, uint32_t k, size_t n, int intensity, float* d_out, hipStream_t s)
{
    if (k == 0 || n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 65536);
    average_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(d_in, k, n, intensity, d_out);
    return hipGetLastError();
}

hipError_t launchSpectral(const float* d_proj, uint32_t n_mat, uint32_t views, size_t n, const float* d_mu,
                          const float* d_inten, const float* d_resp, uint32_t n_bins, const float* d_total,
                          float* d_out, hipStream_t s)
{
    if (n_mat == 0 || n_mat > kMaxMaterials) return hipErrorInvalidValue;
    if (views == 0 || n == 0) return hipSuccess;
    const uint64_t pb = (uint64_t(n) + kBlock - 1) / kBlock, blocks = pb * views;
    if (blocks > 0x7fffffffull) return hipErrorInvalidConfiguration;
    spectral_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(d_proj, n_mat, views, n, uint32_t(pb), d_mu, d_inten,
                                                        d_resp, n_bins, d_total, d_out);
    return hipGetLastError();
}

hipError_t launchCombineViews(const float* d_a, const float* d_ca, const float* d_b, const float* d_cb,
                              uint32_t views, size_t n, float* d_out, hipStream_t s)
{
    if (views == 0 || n == 0) return hipSuccess;
    const uint64_t pb = (uint64_t(n) + kBlock - 1) / kBlock, blocks = pb * views;
    if (blocks > 0x7fffffffull) return hipErrorInvalidConfiguration;
    combine_views_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(d_a, d_ca, d_b, d_cb, n, uint32_t(pb), d_out);
    return hipGetLastError();
}

hipError_t launchProbeLds(uint32_t grid, uint32_t iters, bool pairs, float* d_out, hipStream_t s)
{
    if (grid == 0 || iters % 16) return hipErrorInvalidValue;
    if (pairs) probe_lds_kernel<true><<<grid, kBlock, 0, s>>>(iters, 0u, d_out);
    else probe_lds_kernel<false><<<grid, kBlock, 0, s>>>(iters, 0u, d_out);
    return hipGetLastError();
}

hipError_t launchProbeStream(const void* d_src, uint32_t grid, uint32_t iters, float* d_out, hipStream_t s)
{
    if (grid == 0 || iters % 4) return hipErrorInvalidValue;
    probe_stream_kernel<<<grid, kBlock, 0, s>>>(d_src, iters, 4096u, 16383u, d_out); // 16 KiB window
    return hipGetLastError();
}

bool validShape(const LaunchShape& s)
{
    return s.wave_w * s.wave_h == 64 && s.block_waves_x * s.block_waves_y == kBlock / 64 &&
           (s.unroll >= 1 && s.unroll <= 4) && s.blocks_per_cu <= 8;
}

hipError_t launchSetup(const double* d_pmats, uint32_t views, uint32_t modules, uint32_t channels,
                       uint32_t rows, const VolumeGeom& vg, int interpolate, RayGeom* d_geom,
                       hipStream_t s)
{
    (void)channels;
    (void)rows;
    // position frame of the gathers: trilinear q = p + 1/2 (the sampler's -1/2, padding +1),
    // nearest q = p + 1 (padding +1), p = the reference's position relative to the volume corner
    const float shift = interpolate ? 0.5f : 1.0f;
    const uint32_t n = views * modules;
    if (n == 0) return hipSuccess;
    setup_kernel<<<(n + 63) / 64, 64, 0, s>>>(d_pmats, views, modules, vg, shift, d_geom);
    return hipGetLastError();
}

hipError_t launchRaycast(const VolumeView& vol, const VolumeGeom& vg, const RayGeom* d_geom,
                         const RayParams& rp, const LaunchShape& shape, float* d_out,
                         hipStream_t s, int accumulate, unsigned long long* tally)
{
    if (!validShape(shape)) return hipErrorInvalidValue;
    CastConst kc = makeCastConst(&vol, vg, rp, shape);
    kc.accumulate = accumulate ? 1u : 0u;
    kc.tally = tally;
    const uint64_t blocks = uint64_t(kc.tiles_per_vm) * rp.views * rp.modules;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0xffffffffull) return hipErrorInvalidConfiguration;
    const uint32_t lds = ldsForBlocks(shape.blocks_per_cu);
    if (kc.chunk) {
        if (kc.chunk < 16 || kc.chunk > 1024 || kc.lds_texels > kMaxLdsTexels || kc.lds_texels % 128)
            return hipErrorInvalidValue;
        const uint32_t box = kc.lds_texels * 8;
        const uint32_t dyn = box > lds ? box : lds;
        switch (vol.storage) {
        case Storage::F16: return launchStaged<Storage::F16>(vol, d_geom, rp, kc, shape.unroll, uint32_t(blocks), dyn, d_out, s);
        case Storage::BF16: return launchStaged<Storage::BF16>(vol, d_geom, rp, kc, shape.unroll, uint32_t(blocks), dyn, d_out, s);
        default: return launchStaged<Storage::F32>(vol, d_geom, rp, kc, shape.unroll, uint32_t(blocks), dyn, d_out, s);
        }
    }
    switch (vol.storage) {
    case Storage::F16: return launchCastI<Storage::F16>(vol, d_geom, rp, kc, shape.unroll, uint32_t(blocks), lds, d_out, s);
    case Storage::BF16: return launchCastI<Storage::BF16>(vol, d_geom, rp, kc, shape.unroll, uint32_t(blocks), lds, d_out, s);
    default: return launchCastI<Storage::F32>(vol, d_geom, rp, kc, shape.unroll, uint32_t(blocks), lds, d_out, s);
    }
}

hipError_t launchCount(const VolumeGeom& vg, const RayGeom* d_geom, const RayParams& rp,
                       unsigned long long* d_count, hipStream_t s)
{
    const CastConst kc = makeCastConst(nullptr, vg, rp, LaunchShape{});
    const uint64_t rays = uint64_t(rp.channels) * rp.rows * rp.modules * rp.views;
    const uint64_t blocks = (rays + kBlock - 1) / kBlock;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0xffffffffull) return hipErrorInvalidConfiguration;
    count_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(d_geom, rp, kc, d_count);
    return hipGetLastError();
}

hipError_t launchPack(const float* d_src, const uint32_t n[3], uint32_t f, Storage st, Layout ly,
                      void* d_dst, hipStream_t s)
{
    const uint32_t px = n[0] * f + 2, py = n[1] * f + 2, pz = n[2] * f + 2;
    const FromFloat v{ d_src, n[0], n[1], n[2], f };
    build_kernel<<<gridFor(size_t(px) * py * pz), kBlock, 0, s>>>(v, px, py, pz, st, ly, d_dst);
    return hipGetLastError();
}

hipError_t launchOccupancy(const VolumeView& view, uint32_t* d_occ, hipStream_t s)
{
    const size_t ox = occBricks(view.n[0]), oy = occBricks(view.n[1]), oz = occBricks(view.n[2]);
    const size_t nx = ox + 1, ny = oy + 1, nz = oz + 1;
    occ_flag_kernel<<<gridFor(nx * ny * nz), kBlock, 0, s>>>(view, d_occ, uint32_t(ox), uint32_t(oy), uint32_t(oz));
    // x lines: start (j, k) = l * nx; y lines: start i + k * nx * ny; z lines: start i + j * nx
    occ_scan_kernel<<<gridFor(ny * nz), kBlock, 0, s>>>(d_occ, ny * nz, 1, nx, 1, uint32_t(nx));
    occ_scan_kernel<<<gridFor(nx * nz), kBlock, 0, s>>>(d_occ, nx * nz, nx, nx * ny, nx, uint32_t(ny));
    occ_scan_kernel<<<gridFor(nx * ny), kBlock, 0, s>>>(d_occ, nx * ny, nx * ny, 0, nx * ny, uint32_t(nz));
    return hipGetLastError();
}

hipError_t launchUpsample(const VolumeView& src, uint32_t f, void* d_dst, hipStream_t s)
{
    const uint32_t px = src.n[0] * f + 2, py = src.n[1] * f + 2, pz = src.n[2] * f + 2;
    const FromView v{ src, f };
    build_kernel<<<gridFor(size_t(px) * py * pz), kBlock, 0, s>>>(v, px, py, pz, src.storage,
                                                                  src.layout, d_dst);
    return hipGetLastError();
}

hipError_t launchSum(const VolumeView& src, double* d_sum, hipStream_t s)
{
    const size_t total = size_t(src.n[0]) * src.n[1] * src.n[2];
    sum_kernel<<<gridFor(total) < 4096 ? gridFor(total) : 4096, kBlock, 0, s>>>(src, d_sum);
    return hipGetLastError();
}

hipError_t launchFillMean(const double* d_sum, double count, Storage st, Layout ly, void* d_dst,
                          hipStream_t s)
{
    const FromMean v{ d_sum, count };
    build_kernel<<<1, 64, 0, s>>>(v, 3u, 3u, 3u, st, ly, d_dst);
    return hipGetLastError();
}

hipError_t launchZero(float* d, size_t n, hipStream_t s)
{
    zero_kernel<<<gridFor(n), kBlock, 0, s>>>(d, n);
    return hipGetLastError();
}

hipError_t launchQueueInit(BatchQueue* d_q, unsigned long long total, hipStream_t s)
{
    queue_init_kernel<<<1, 64, 0, s>>>(d_q, total);
    return hipGetLastError();
}

namespace {
template <Storage ST, Layout LY, bool INTERP>
const void* batchKernelU(uint32_t unroll)
{
    switch (unroll) {
    case 1: return reinterpret_cast<const void*>(&batch_kernel<ST, LY, INTERP, 1>);
    case 3: // the unstaged kernels run 3 as 4
    case 4: return reinterpret_cast<const void*>(&batch_kernel<ST, LY, INTERP, 4>);
    default: return reinterpret_cast<const void*>(&batch_kernel<ST, LY, INTERP, 2>);
    }
}

template <Storage ST>
const void* batchKernelS(Layout ly, int interp, uint32_t unroll)
{
    if (ly == Layout::Interleaved)
        return interp ? batchKernelU<ST, Layout::Interleaved, true>(unroll)
                      : batchKernelU<ST, Layout::Interleaved, false>(unroll);
    return interp ? batchKernelU<ST, Layout::Plain, true>(unroll) : batchKernelU<ST, Layout::Plain, false>(unroll);
}

template <Storage ST>
const void* batchStagedKernelU(uint32_t unroll)
{
    switch (unroll) {
    case 1: return reinterpret_cast<const void*>(&batch_staged_kernel<ST, 1>);
    case 3: return reinterpret_cast<const void*>(&batch_staged_kernel<ST, 3>);
    case 4: return reinterpret_cast<const void*>(&batch_staged_kernel<ST, 4>);
    default: return reinterpret_cast<const void*>(&batch_staged_kernel<ST, 2>);
    }
}

const void* batchStagedKernel(Storage st, uint32_t unroll)
{
    switch (st) {
    case Storage::F16: return batchStagedKernelU<Storage::F16>(unroll);
    case Storage::BF16: return batchStagedKernelU<Storage::BF16>(unroll);
    default: return batchStagedKernelU<Storage::F32>(unroll);
    }
}

const void* batchKernel(Storage st, Layout ly, int interp, uint32_t unroll)
{
    switch (st) {
    case Storage::F16: return batchKernelS<Storage::F16>(ly, interp, unroll);
    case Storage::BF16: return batchKernelS<Storage::BF16>(ly, interp, unroll);
    default: return batchKernelS<Storage::F32>(ly, interp, unroll);
    }
}

// Dynamic LDS of a batch launch: the staged kernel's box, or the occupancy-capping reservation.
uint32_t batchLds(uint32_t lds_texels, uint32_t blocks_per_cu)
{
    const uint32_t cap = ldsForBlocks(blocks_per_cu), box = lds_texels * 8u;
    return box > cap ? box : cap;
}

bool batchStaged(Storage, Layout ly, int interp, uint32_t lds_texels)
{
    return lds_texels && ly == Layout::Interleaved && interp;
}

} // namespace

hipError_t batchGrid(Storage st, Layout ly, int interpolate, uint32_t unroll, uint32_t blocks_per_cu,
                     uint32_t lds_texels, uint32_t* grid)
{
    if (lds_texels > kMaxLdsTexels || lds_texels % 128) return hipErrorInvalidValue;
    const bool staged = batchStaged(st, ly, interpolate, lds_texels);
    int dev = 0, cus = 0, per_cu = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, staged ? batchStagedKernel(st, unroll) : batchKernel(st, ly, interpolate, unroll), kBlock,
            staged ? batchLds(lds_texels, blocks_per_cu) : ldsForBlocks(blocks_per_cu));
    if (e != hipSuccess) return e;
    *grid = uint32_t(cus * (per_cu > 0 ? per_cu : 1));
    return hipSuccess;
}

hipError_t launchBatch(const BatchItem* d_items, uint32_t n_items, Storage st, Layout ly,
                       int interpolate, uint32_t unroll, uint32_t blocks_per_cu, uint32_t lds_texels,
                       uint32_t grid, BatchQueue* d_q, uint32_t stride, hipStream_t s)
{
    if (n_items == 0 || grid == 0) return hipSuccess;
    if (lds_texels > kMaxLdsTexels || lds_texels % 128) return hipErrorInvalidValue;
    const bool staged = batchStaged(st, ly, interpolate, lds_texels);
    const void* fn = staged ? batchStagedKernel(st, unroll) : batchKernel(st, ly, interpolate, unroll);
    void* args[] = { const_cast<BatchItem**>(&d_items), &n_items, &d_q, &stride };
    return hipLaunchKernel(fn, dim3(grid), dim3(kBlock), args,
                           staged ? batchLds(lds_texels, blocks_per_cu) : ldsForBlocks(blocks_per_cu), s);
}

// ---------------------------------------------------------------------------------------------
// Static device scratch of the allocation-free projection: the ray geometry records of one
// projection. Each device holds its own copy for the life of the process. Host code never uses
// the array's name as an address (in host code it is a stand-in, not device memory): initScratch
// asks the runtime for the current device's address once, and every launch takes that pointer.

namespace {
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};
} // namespace

hipError_t initScratch(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice; // the symbol's address is the current device's
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void* scratchPtr(int device)
{
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

} // namespace ctlk
