// HIP kernels of the volume processing on resident volumes, written for gfx950 (CDNA4, wave64).
//
//   slice_kernel     one thread per slice pixel of any number of planes (VolumeSlicer).
//   resample_kernel  one thread per grid point, sample_kernel one per listed point
//                    (VolumeResampler).
//   plane_brick_kernel  plane integrals (RadonTransform3D), the default: a workgroup stages one
//                    brick of the volume in LDS and sums, for many orientations, the samples of
//                    every plane through it (details at the kernel); plane_fix_kernel converts its
//                    fixed-point totals.
//   plane_kernel     plane integrals by tiles (volumes without a brick copy or with non-finite
//                    voxels): a workgroup takes one plane orientation,
//                    up to kPlaneDists distances and a range of 8x8 pixel tiles of the template
//                    plane; every wavefront walks its tiles, reading all distances of a pixel
//                    back to back (the same small patch of the volume, so the gathers share L1
//                    lines), and keeps one FP32 sum per distance in registers. The workgroup's
//                    sums go to a partial buffer that plane_reduce_kernel adds up in a fixed
//                    order, so results do not depend on scheduling.
//
// Every read goes through sampleImage: OpenCL's linear sampler with border colour 0 on the padded
// resident volume, FP32 interpolation for every storage type.
#include "kernels.h"
#include "sampler.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace ctlk {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNumXcd = 8;

// Contiguous logical ranges per XCD (workgroup b runs on XCD b % 8), so that neighbouring
// logical workgroups -- neighbouring orientations in the plane kernel -- share an L2.
__device__ __forceinline__ uint32_t xcdOrder(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    const uint32_t x = b % kNumXcd, k = b / kNumXcd;
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ Pitch pitchOf(const VolumeView& v)
{
    const uint32_t eb = v.layout == Layout::Interleaved ? 8u : (v.storage == Storage::F32 ? 4u : 2u);
    Pitch p;
    p.ty = v.pitch_y;
    p.tz = v.pitch_z;
    p.y = v.pitch_y * eb;
    p.z = v.pitch_z * eb;
    p.yz = (v.pitch_y + v.pitch_z) * eb;
    return p;
}

__device__ __forceinline__ float waveSum(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// volumeSlicer.cl: u = H (x - cx, y - cy, 0, 1)^T - corner + 1/2 per plane; out[plane][y][x].
template <Storage ST, Layout LY>
__global__ __launch_bounds__(kBlock) void slice_kernel(VolumeView vol, const PlaneMap* __restrict__ maps,
                                                       uint32_t width, uint32_t height, uint32_t tiles_x,
                                                       uint32_t tiles_per_plane, float cx, float cy,
                                                       float* __restrict__ out)
{
    const uint32_t plane = blockIdx.x / tiles_per_plane, t = blockIdx.x % tiles_per_plane;
    const uint32_t x = (t % tiles_x) * 16 + (threadIdx.x & 15), y = (t / tiles_x) * 16 + (threadIdx.x >> 4);
    if (x >= width || y >= height) return;
    const PlaneMap& m = maps[plane];
    const float vx = float(x) - cx, vy = float(y) - cy;
    float u[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) u[k] = fmaf(m.h[4 * k], vx, fmaf(m.h[4 * k + 1], vy, m.h[4 * k + 3])) - m.c[k] + 0.5f;
    const float nhi[3] = { float(vol.n[0]) + 1.0f, float(vol.n[1]) + 1.0f, float(vol.n[2]) + 1.0f };
    out[(size_t(plane) * height + y) * width + x] =
        sampleImage<ST, LY>(makeRsrc(vol.data, vol.bytes), pitchOf(vol), nhi, u[0], u[1], u[2]);
}

// volumeResampler.cl: u = (N - 1) / (hi - lo) * (p - lo) + 1/2 per axis.
struct Ranges {
    float lo[3], scale[3];
};

template <Storage ST, Layout LY>
__global__ __launch_bounds__(kBlock) void resample_kernel(VolumeView vol, Ranges r, const float* __restrict__ p1,
                                                          const float* __restrict__ p2, const float* __restrict__ p3,
                                                          uint32_t n1, uint32_t n2, uint32_t count,
                                                          float* __restrict__ out)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    const uint32_t s1 = i % n1, s23 = i / n1, s2 = s23 % n2, s3 = s23 / n2;
    const float nhi[3] = { float(vol.n[0]) + 1.0f, float(vol.n[1]) + 1.0f, float(vol.n[2]) + 1.0f };
    out[i] = sampleImage<ST, LY>(makeRsrc(vol.data, vol.bytes), pitchOf(vol), nhi,
                                 r.scale[0] * (p1[s1] - r.lo[0]) + 0.5f, r.scale[1] * (p2[s2] - r.lo[1]) + 0.5f,
                                 r.scale[2] * (p3[s3] - r.lo[2]) + 0.5f);
}

template <Storage ST, Layout LY>
__global__ __launch_bounds__(kBlock) void sample_kernel(VolumeView vol, Ranges r, const float* __restrict__ xyz,
                                                        size_t count, float* __restrict__ out)
{
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= count) return;
    const float* p = xyz + 3 * i;
    const float nhi[3] = { float(vol.n[0]) + 1.0f, float(vol.n[1]) + 1.0f, float(vol.n[2]) + 1.0f };
    out[i] = sampleImage<ST, LY>(makeRsrc(vol.data, vol.bytes), pitchOf(vol), nhi,
                                 r.scale[0] * (p[0] - r.lo[0]) + 0.5f, r.scale[1] * (p[1] - r.lo[1]) + 0.5f,
                                 r.scale[2] * (p[2] - r.lo[2]) + 0.5f);
}

// A volume's occupancy table (VolumeView::occ, occ_n) for the plane kernel; t NULL: none.
struct OccTable {
    const uint32_t* t;
    uint32_t n[3];
};

// Reads of the plane kernel: the resident volume's own texels, or its FP32 brick copy.
template <Storage ST, Layout LY> struct TexelFetch {
    VolumeView vol;
    __device__ __forceinline__ float operator()(const float nhi[3], float qx, float qy, float qz) const
    {
        return sampleTexel<ST, LY>(makeRsrc(vol.data, vol.bytes), pitchOf(vol), nhi, qx, qy, qz);
    }
};

// Brick copies for the plane kernel, FP32 whatever the volume's storage, in 128-byte bricks (one
// L1/L2 line each), bricks x fastest. Both hold the interleaved FP32 texels {v, v(z+1) - v} and
// interpolate exactly as the FP32 interleaved sampler does (same operands, same order):
//   Texel8: 8-byte texels in bricks of 4 x 2 x 2; a sample reads four texels (x and x + 1 of rows
//           y and y + 1) with four 8-byte loads.
//   Pair16: 16-byte elements (the texels of x and x + 1) in bricks of 2 x 2 x 2; a sample reads
//           two elements with two 16-byte loads, touching fewer lines for twice the memory.
// On 1024 x 128 planes: Texel8 44.8 ms at 256^3 (its 138 MB copy stays in the last-level cache)
// and 277 ms at 512^3, Pair16 55.0 / 249 ms, against 58.8 / 270 ms for eight dword gathers of
// 4 x 4 x 2-voxel bricks of plain voxels. brickShape picks Texel8 while its copy is at most
// kTexel8CacheBytes, Pair16 beyond.
struct Texel8Fetch {
    const float* data;
    uint32_t bytes;
    uint32_t row, slab; // bytes per row of bricks (nbx * 128) and per slab (nbx * nby * 128)
    __device__ __forceinline__ float operator()(const float nhi[3], float qx, float qy, float qz) const
    {
        // branch-free like sampleTexel: outside positions read texel 0 and are discarded
        const bool in = qx >= 0.0f && qy >= 0.0f && qz >= 0.0f && qx < nhi[0] && qy < nhi[1] && qz < nhi[2];
        qx = in ? qx : 0.0f;
        qy = in ? qy : 0.0f;
        qz = in ? qz : 0.0f;
        const __amdgpu_buffer_rsrc_t r = makeRsrc(data, bytes);
        const uint32_t ix = uint32_t(int(qx)), iy = uint32_t(int(qy)), iz = uint32_t(int(qz));
        const float wx = __builtin_amdgcn_fractf(qx), wy = __builtin_amdgcn_fractf(qy);
        const float wz = __builtin_amdgcn_fractf(qz);
        const uint32_t z0 = (iz >> 1) * slab + ((iz & 1u) << 6);
        const uint32_t x0 = ((ix >> 2) << 7) + ((ix & 3u) << 3), x1 = (((ix + 1) >> 2) << 7) + (((ix + 1) & 3u) << 3);
        const uint32_t y0 = __umul24(iy >> 1, row) + ((iy & 1u) << 5) + z0;
        const uint32_t y1 = __umul24((iy + 1) >> 1, row) + (((iy + 1) & 1u) << 5) + z0;
        auto at = [&](uint32_t o) { return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, o, 0, 0)); };
        const float v = triDeltaZ(at(x0 + y0), at(x1 + y0), at(x0 + y1), at(x1 + y1), wx, wy, wz);
        return in ? v : 0.0f;
    }
};

struct Pair16Fetch {
    const float* data;
    uint32_t bytes;
    uint32_t row, slab; // bytes per row of bricks (nbx * 128) and per slab (nbx * nby * 128)
    __device__ __forceinline__ float operator()(const float nhi[3], float qx, float qy, float qz) const
    {
        const bool in = qx >= 0.0f && qy >= 0.0f && qz >= 0.0f && qx < nhi[0] && qy < nhi[1] && qz < nhi[2];
        qx = in ? qx : 0.0f;
        qy = in ? qy : 0.0f;
        qz = in ? qz : 0.0f;
        const __amdgpu_buffer_rsrc_t r = makeRsrc(data, bytes);
        const uint32_t ix = uint32_t(int(qx)), iy = uint32_t(int(qy)), iz = uint32_t(int(qz));
        const float wx = __builtin_amdgcn_fractf(qx), wy = __builtin_amdgcn_fractf(qy);
        const float wz = __builtin_amdgcn_fractf(qz);
        const uint32_t xz = ((ix >> 1) << 7) + ((ix & 1u) << 4) + (iz >> 1) * slab + ((iz & 1u) << 6);
        const uint32_t y0 = __umul24(iy >> 1, row) + ((iy & 1u) << 5);
        const uint32_t y1 = __umul24((iy + 1) >> 1, row) + (((iy + 1) & 1u) << 5);
        const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, xz + y0, 0, 0));
        const float4 b = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, xz + y1, 0, 0));
        const float v = triDeltaZ(make_float2(a.x, a.y), make_float2(a.z, a.w), make_float2(b.x, b.y),
                                  make_float2(b.z, b.w), wx, wy, wz);
        return in ? v : 0.0f;
    }
};

// Value of voxel (x, y, z) of the padded frame (0 beyond it), as FP32.
__device__ __forceinline__ float padValue(const VolumeView& src, uint32_t x, uint32_t y, uint32_t z)
{
    if (x >= src.n[0] + 2 || y >= src.n[1] + 2 || z >= src.n[2] + 2) return 0.0f;
    const size_t j = size_t(z) * src.pitch_z + size_t(y) * src.pitch_y + x;
    if (src.storage == Storage::F32)
        return static_cast<const float*>(src.data)[src.layout == Layout::Interleaved ? 2 * j : j];
    const uint16_t h = static_cast<const uint16_t*>(src.data)[src.layout == Layout::Interleaved ? 4 * j : j];
    return src.storage == Storage::F16 ? float(__builtin_bit_cast(_Float16, h)) : __uint_as_float(uint32_t(h) << 16);
}

// Texel8 copy: texel i = 16 * brick + (x & 3) + 4 (y & 1) + 8 (z & 1)
__global__ __launch_bounds__(kBlock) void texel8_kernel(VolumeView src, uint32_t nbx, uint32_t nby, size_t total,
                                                        float2* __restrict__ dst)
{
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += size_t(gridDim.x) * kBlock) {
        const size_t b = i >> 4;
        const uint32_t e = uint32_t(i & 15);
        const uint32_t bx = uint32_t(b % nbx), by = uint32_t((b / nbx) % nby), bz = uint32_t(b / (size_t(nbx) * nby));
        const uint32_t x = bx * 4 + (e & 3u), y = by * 2 + ((e >> 2) & 1u), z = bz * 2 + (e >> 3);
        float2 t;
        t.x = padValue(src, x, y, z);
        t.y = padValue(src, x, y, z + 1) - t.x;
        dst[i] = t;
    }
}

// Pair16 copy: element i = 8 * brick + (x & 1) + 2 (y & 1) + 4 (z & 1)
__global__ __launch_bounds__(kBlock) void pair16_kernel(VolumeView src, uint32_t nbx, uint32_t nby, size_t total,
                                                        float4* __restrict__ dst)
{
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < total; i += size_t(gridDim.x) * kBlock) {
        const size_t b = i >> 3;
        const uint32_t e = uint32_t(i & 7);
        const uint32_t bx = uint32_t(b % nbx), by = uint32_t((b / nbx) % nby), bz = uint32_t(b / (size_t(nbx) * nby));
        const uint32_t x = bx * 2 + (e & 1u), y = by * 2 + ((e >> 1) & 1u), z = bz * 2 + (e >> 2);
        float4 t;
        t.x = padValue(src, x, y, z);
        t.y = padValue(src, x, y, z + 1) - t.x;
        t.z = padValue(src, x + 1, y, z);
        t.w = padValue(src, x + 1, y, z + 1) - t.z;
        dst[i] = t;
    }
}

// planeIntegral.cl: u = H (x, y, 0, 1)^T + d * shift for pixel (x, y) and distance d, read at the
// texel position q = u + 1/2 (the 1/2 is folded into H's last column).
// Logical workgroup b = (distance chunk, orientation, split), orientation fastest after the split.
// Every lane reads the chunk's distances of its pixel back to back (independent gathers in
// flight). The four wavefronts take the tiles of a row of a 4x4 tile group (32x32 pixels)
// together and the group's rows one after the other, so tiles sharing texel rows are read close
// together in time. A wavefront skips a tile when the box spanned by its corner pixels and the
// chunk's distance range misses the volume's support on some axis (with half a texel to spare):
// every sample of such a tile reads the zero border. With the volume's occupancy table, a wavefront
// also skips the distances of a tile whose samples read only bricks that hold no non-zero voxel:
// lane 8 j + c reads corner c of the summed-area table for distance j (one gather for all eight
// distances), and the eight lanes of a distance add their signed corners. Skipped samples are
// exactly 0, so every sum is the one without skipping, bit for bit.
template <class FETCH>
__global__ __launch_bounds__(kBlock) void plane_kernel(FETCH fetch, uint32_t n0, uint32_t n1, uint32_t n2,
                                                       const PlaneMap* __restrict__ maps,
                                                       const float* __restrict__ dist, PlaneJob job,
                                                       OccTable occ, float* __restrict__ partial)
{
    const uint32_t b = xcdOrder(blockIdx.x, gridDim.x);
    const uint32_t split = b % job.splits, rest = b / job.splits;
    const uint32_t o = rest % job.n_orient, d0 = (rest / job.n_orient) * kPlaneDists;
    const uint32_t nd = min(uint32_t(kPlaneDists), job.n_dist - d0);
    const uint32_t v0 = uint32_t(uint64_t(job.n_tiles) * split / job.splits);
    const uint32_t v1 = uint32_t(uint64_t(job.n_tiles) * (split + 1) / job.splits);
    const uint32_t groups_x = (job.tiles_x + 3) / 4;
    const PlaneMap& m = maps[o];
    float h[3][3], s[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        h[k][0] = m.h[4 * k];
        h[k][1] = m.h[4 * k + 1];
        h[k][2] = m.h[4 * k + 3] + 0.5f;
        s[k] = m.c[k];
    }
    float dd[kPlaneDists], acc[kPlaneDists];
    float dlo = dist[d0], dhi = dlo;
#pragma unroll
    for (int j = 0; j < kPlaneDists; ++j) {
        dd[j] = uint32_t(j) < nd ? dist[d0 + j] : dlo;
        dlo = fminf(dlo, dd[j]);
        dhi = fmaxf(dhi, dd[j]);
        acc[j] = 0.0f;
    }
    const float nhi[3] = { float(n0) + 1.0f, float(n1) + 1.0f, float(n2) + 1.0f };
    // constant part and distance term's range of each texel coordinate
    float elo[3], ehi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        elo[k] = h[k][2] + fminf(dlo * s[k], dhi * s[k]);
        ehi[k] = h[k][2] + fmaxf(dlo * s[k], dhi * s[k]);
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the occupancy test's distance (lane / 8) and table corner (lane % 8: bit k picks the upper
    // bound on axis k), and that distance's constant part of each texel coordinate
    const uint32_t cj = lane >> 3, cc = lane & 7;
    const float cd = dist[d0 + (cj < nd ? cj : 0u)];
    float ce[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) ce[k] = fmaf(cd, s[k], h[k][2]);
    const uint32_t all = nd >= 32 ? ~0u : (1u << nd) - 1u;
    for (uint32_t v = v0 + wave; v < v1; v += kBlock / 64) {
        const uint32_t g = v >> 4, w = v & 15;
        const uint32_t tx = (g % groups_x) * 4 + (w & 3), ty = (g / groups_x) * 4 + (w >> 2);
        if (tx >= job.tiles_x || ty >= job.tiles_y) continue;
        const float x0 = float(tx * 8), x1 = float(tx * 8 + 7), y0 = float(ty * 8), y1 = float(ty * 8 + 7);
        bool miss = false;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float ax = h[k][0] * x0, bx = h[k][0] * x1, ay = h[k][1] * y0, by = h[k][1] * y1;
            const float lo = elo[k] + fminf(ax, bx) + fminf(ay, by);
            const float hi = ehi[k] + fmaxf(ax, bx) + fmaxf(ay, by);
            miss = miss || hi < -0.5f || lo >= nhi[k] + 0.5f;
        }
        if (miss) continue; // uniform: the tile's values are the same for every lane
        uint32_t live = all; // distances of the chunk whose samples may read a non-zero voxel
        if (occ.t) {
            bool empty = cj >= nd;
            uint32_t idx[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                // positions of this distance's samples in the tile, with half a texel to spare; a
                // sample at q reads voxels floor(q) and floor(q) + 1, and only samples in [0, n + 1)
                // read any
                const float ax = h[k][0] * x0, bx = h[k][0] * x1, ay = h[k][1] * y0, by = h[k][1] * y1;
                const float lo = fmaxf(ce[k] + fminf(ax, bx) + fminf(ay, by) - 0.5f, 0.0f);
                const float hi = fminf(ce[k] + fmaxf(ax, bx) + fmaxf(ay, by) + 0.5f, nhi[k]);
                empty = empty || !(lo <= hi);
                const uint32_t last = occ.n[k] - 1u;
                const uint32_t bl = min(uint32_t(lo) >> kOccShift, last);
                const uint32_t bh = min((uint32_t(hi) + 1u) >> kOccShift, last) + 1u;
                idx[k] = ((cc >> k) & 1u) ? bh : bl;
            }
            const uint32_t nx = occ.n[0] + 1u, nxy = nx * (occ.n[1] + 1u);
            int cnt = int(occ.t[idx[2] * nxy + idx[1] * nx + idx[0]]);
            cnt = (__popc(~cc & 7u) & 1u) ? -cnt : cnt; // summed-area corner sign
            cnt += __shfl_xor(cnt, 1, 64);
            cnt += __shfl_xor(cnt, 2, 64);
            cnt += __shfl_xor(cnt, 4, 64);
            const uint64_t bal = __ballot(cc == 0 && !empty && cnt != 0);
            live = 0;
#pragma unroll
            for (int j = 0; j < kPlaneDists; ++j) live |= uint32_t((bal >> (8 * j)) & 1u) << j;
            if (!live) continue; // uniform
        }
        const uint32_t x = tx * 8 + (lane & 7), y = ty * 8 + (lane >> 3);
        if (x >= job.width || y >= job.height) continue;
        const float fx = float(x), fy = float(y);
        const float b0 = fmaf(h[0][0], fx, fmaf(h[0][1], fy, h[0][2]));
        const float b1 = fmaf(h[1][0], fx, fmaf(h[1][1], fy, h[1][2]));
        const float b2 = fmaf(h[2][0], fx, fmaf(h[2][1], fy, h[2][2]));
        // every lane reads all kPlaneDists positions (a short chunk repeats its first distance and
        // drops those sums), with no branch between the samples
#pragma unroll
        for (int j = 0; j < kPlaneDists; ++j)
            if ((live >> j) & 1u) // uniform
                acc[j] += fetch(nhi, fmaf(dd[j], s[0], b0), fmaf(dd[j], s[1], b1), fmaf(dd[j], s[2], b2));
    }
    __shared__ float red[kBlock / 64][kPlaneDists];
#pragma unroll
    for (int j = 0; j < kPlaneDists; ++j) {
        const float r = waveSum(acc[j]);
        if (lane == 0) red[wave][j] = r;
    }
    __syncthreads();
    if (threadIdx.x < nd) {
        const uint32_t j = threadIdx.x;
        const float r = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
        partial[(size_t(o) * job.n_dist + d0 + j) * job.splits + split] = r;
    }
}

// out[d * out_stride + o] = scale * (sum of the splits of (o, d), in split order).
__global__ __launch_bounds__(kBlock) void plane_reduce_kernel(const float* __restrict__ partial, uint32_t n_orient,
                                                              uint32_t n_dist, uint32_t splits, float scale,
                                                              float* __restrict__ out, uint32_t out_stride)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_orient * n_dist) return;
    const uint32_t o = i % n_orient, d = i / n_orient;
    const float* src = partial + (size_t(o) * n_dist + d) * splits;
    float s = 0.0f;
    for (uint32_t k = 0; k < splits; ++k) s += src[k];
    out[size_t(d) * out_stride + o] = s * scale;
}

// ---------------------------------------------------------------------------------------------
// Plane integrals by volume bricks (the default when the volume has an FP32 brick copy).
//
// A workgroup takes one brick of kPlaneBrick^3 cells of the padded texel frame (a sample at q
// belongs to the cell floor(q)) and a range of orientations. It copies the brick's texels (the
// cells' taps: x and y one further, z carried by the (v, dz) texels) from the brick copy into
// LDS once and then, for every orientation and every distance whose plane cuts the brick, sums
// the samples of that plane whose cell lies in the brick: the template pixels of the plane's
// intersection polygon with the brick, walked in 8 x 8-pixel tiles of its bounding rectangle.
// Every template sample inside the support belongs to exactly one brick (its position is computed
// with the same FP32 operations as in plane_kernel, so the same cell), and a brick without a
// non-zero voxel (occupancy table) is skipped. One workgroup thus serves many orientations from one
// staging of its voxels instead of every orientation streaming the volume through the caches.
//
// Each (brick, orientation, distance) sum is added to its plane's total as a 64-bit fixed-point
// integer (2^fix_shift units per value), so the totals do not depend on the order in which
// workgroups finish; plane_fix_kernel converts them back. The host picks fix_shift from the
// volume's largest |value| so that no total can overflow.
constexpr int kPlaneBrick = 14;
constexpr int kPBX = kPlaneBrick + 1, kPBY = kPlaneBrick + 1, kPBZ = kPlaneBrick;
constexpr int kPBTexels = kPBX * kPBY * kPBZ; // 3150 texels, 25,200 bytes of LDS

// Texel (x, y, z) of the padded frame, (v, dz), from a brick copy.
struct Texel8Src {
    const float* data;
    uint32_t bytes, row, slab;
    __device__ __forceinline__ float2 operator()(uint32_t x, uint32_t y, uint32_t z) const
    {
        const uint32_t o = (z >> 1) * slab + (y >> 1) * row + ((x >> 2) << 7) + ((z & 1u) << 6) + ((y & 1u) << 5) +
                           ((x & 3u) << 3);
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(makeRsrc(data, bytes), o, 0, 0));
    }
};
struct Pair16Src {
    const float* data;
    uint32_t bytes, row, slab;
    __device__ __forceinline__ float2 operator()(uint32_t x, uint32_t y, uint32_t z) const
    {
        const uint32_t o = (z >> 1) * slab + (y >> 1) * row + ((x >> 1) << 7) + ((z & 1u) << 6) + ((y & 1u) << 5) +
                           ((x & 1u) << 4);
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(makeRsrc(data, bytes), o, 0, 0));
    }
};

// Sum over the wave's 64 lanes (all active), the same value returned to every lane: DPP within
// each 16-lane row, then the four rows.
__device__ __forceinline__ float waveSumDpp(float v)
{
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xF, 0xF, false)); // row_ror:4
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false)); // row_ror:8
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return (r0 + r1) + (r2 + r3);
}

template <class SRC>
__global__ __launch_bounds__(kBlock) void plane_brick_kernel(SRC src, uint32_t n0, uint32_t n1, uint32_t n2,
                                                             OccTable occ, const PlaneMap* __restrict__ maps,
                                                             const PlaneInv* __restrict__ inv,
                                                             const float* __restrict__ dist, uint32_t n_orient,
                                                             uint32_t n_dist, uint32_t width, uint32_t height,
                                                             uint32_t o_chunk, double fix_scale,
                                                             unsigned long long* __restrict__ acc)
{
    __shared__ __attribute__((aligned(16))) float2 box[kPBTexels];
    __shared__ uint32_t s_j[kBlock / 64][64]; // each wave's list of distances to cut the brick
    __shared__ float s_d[kBlock / 64][64];
    const uint32_t nb[3] = { (n0 + kPlaneBrick) / kPlaneBrick, (n1 + kPlaneBrick) / kPlaneBrick,
                             (n2 + kPlaneBrick) / kPlaneBrick }; // bricks over cells 0 .. n
    const uint32_t nbricks = nb[0] * nb[1] * nb[2];
    const uint32_t brick = blockIdx.x % nbricks, ob = (blockIdx.x / nbricks) * o_chunk;
    const uint32_t oe = min(ob + o_chunk, n_orient);
    const uint32_t bx = brick % nb[0], by = (brick / nb[0]) % nb[1], bz = brick / (nb[0] * nb[1]);
    const uint32_t nn[3] = { n0, n1, n2 };
    const uint32_t c0[3] = { bx * kPlaneBrick, by * kPlaneBrick, bz * kPlaneBrick };
    uint32_t nc[3]; // cells of this brick per axis
#pragma unroll
    for (int k = 0; k < 3; ++k) nc[k] = min(uint32_t(kPlaneBrick), nn[k] + 1u - c0[k]);
    if (occ.t) { // the brick's samples read voxels c0 .. c0 + nc
        uint32_t lo[3], hi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t last = occ.n[k] - 1u;
            lo[k] = min(c0[k] >> kOccShift, last);
            hi[k] = min((c0[k] + nc[k]) >> kOccShift, last) + 1u;
        }
        const uint32_t nx = occ.n[0] + 1u, nxy = nx * (occ.n[1] + 1u);
        typedef __attribute__((address_space(4))) const uint32_t occ_word;
        const occ_word* t = (const occ_word*)(occ.t);
        auto at = [&](uint32_t i, uint32_t j, uint32_t k) { return t[k * nxy + j * nx + i]; };
        const uint32_t cnt = at(hi[0], hi[1], hi[2]) - at(lo[0], hi[1], hi[2]) - at(hi[0], lo[1], hi[2]) -
                             at(hi[0], hi[1], lo[2]) + at(lo[0], lo[1], hi[2]) + at(lo[0], hi[1], lo[2]) +
                             at(hi[0], lo[1], lo[2]) - at(lo[0], lo[1], lo[2]);
        if (cnt == 0) return; // uniform: every sample of the brick is exactly 0
    }
    // stage the brick's texels (texels past the brick's cells read as 0 and are never sampled)
    for (uint32_t t = threadIdx.x; t < uint32_t(kPBTexels); t += kBlock) {
        const uint32_t lx = t % kPBX, ly = (t / kPBX) % kPBY, lz = t / (kPBX * kPBY);
        const bool in = lx <= nc[0] && ly <= nc[1] && lz < nc[2];
        box[t] = in ? src(c0[0] + lx, c0[1] + ly, c0[2] + lz) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float qlo[3] = { float(c0[0]), float(c0[1]), float(c0[2]) };
    const float qhi[3] = { float(c0[0] + nc[0]), float(c0[1] + nc[1]), float(c0[2] + nc[2]) };
    const uint32_t base8 = (c0[2] * uint32_t(kPBX * kPBY) + c0[1] * uint32_t(kPBX) + c0[0]) * 8u;
    for (uint32_t o = ob + wave; o < oe; o += kBlock / 64) {
        typedef __attribute__((address_space(4))) const PlaneMap map_const;
        typedef __attribute__((address_space(4))) const PlaneInv inv_const;
        const map_const& m = *(const map_const*)(maps + o);
        const inv_const& mi = *(const inv_const*)(inv + o);
        float h[3][3], c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            h[k][0] = m.h[4 * k];
            h[k][1] = m.h[4 * k + 1];
            h[k][2] = m.h[4 * k + 3] + 0.5f;
            c[k] = m.c[k];
        }
        // (X, Y, D) = template pixel and distance of a position, affine in q: at the brick's
        // lowest corner and per unit of each axis' extent
        float v0[3], vk[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            v0[r] = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v0[r] = fmaf(mi.m[r][k], qlo[k] - h[k][2], v0[r]);
                vk[r][k] = mi.m[r][k] * float(nc[k]);
            }
        }
        float dmin = v0[2], dmax = v0[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dmin += fminf(vk[2][k], 0.0f);
            dmax += fmaxf(vk[2][k], 0.0f);
        }
        const float dm = 1e-3f + 1e-5f * (fabsf(dmin) + fabsf(dmax));
        // the distances whose plane can cut the brick, gathered into this wave's list (64 at a
        // time) so that one pass of the wave's lanes computes their polygons
        uint32_t listed = 0;
        auto flush = [&]() {
            __builtin_amdgcn_wave_barrier();
            const uint32_t j = lane < listed ? s_j[wave][lane] : 0u;
            const float d = s_d[wave][lane];
            bool ok = lane < listed;
            // bounding rectangle of the plane's polygon in the brick: the plane cuts the brick's
            // edges where D - d changes sign
            float xl = FLT_MAX, xh = -FLT_MAX, yl = FLT_MAX, yh = -FLT_MAX;
            if (ok) {
                float cx[8], cy[8], cf[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    float X = v0[0], Y = v0[1], D = v0[2];
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        if ((q >> k) & 1) {
                            X += vk[0][k];
                            Y += vk[1][k];
                            D += vk[2][k];
                        }
                    cx[q] = X;
                    cy[q] = Y;
                    cf[q] = D - d;
                }
#pragma unroll
                for (int q = 0; q < 8; ++q)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        if ((q >> k) & 1) continue;
                        const int r = q | (1 << k);
                        const float fa = cf[q], fb = cf[r];
                        if ((fa <= 0.0f) != (fb <= 0.0f) || fa == 0.0f) {
                            const float den = fa - fb;
                            const float t = den != 0.0f ? fminf(fmaxf(fa / den, 0.0f), 1.0f) : 0.0f;
                            const float X = fmaf(t, cx[r] - cx[q], cx[q]), Y = fmaf(t, cy[r] - cy[q], cy[q]);
                            xl = fminf(xl, X);
                            xh = fmaxf(xh, X);
                            yl = fminf(yl, Y);
                            yh = fmaxf(yh, Y);
                        }
                    }
            }
            const float xm = 0.05f + 1e-5f * fmaxf(fabsf(xl), fabsf(xh)), ym = 0.05f + 1e-5f * fmaxf(fabsf(yl), fabsf(yh));
            const int ixl = int(fmaxf(ceilf(xl - xm), 0.0f));
            const int ixh = int(fminf(floorf(xh + xm), float(width - 1u)));
            const int iyl = int(fmaxf(ceilf(yl - ym), 0.0f));
            const int iyh = int(fminf(floorf(yh + ym), float(height - 1u)));
            ok = ok && xl <= xh && ixl <= ixh && iyl <= iyh;
            uint64_t todo = __ballot(ok);
            while (todo) {
                const int L = __builtin_ctzll(todo);
                todo &= todo - 1;
                const int rxl = __builtin_amdgcn_readlane(ixl, L), rxh = __builtin_amdgcn_readlane(ixh, L);
                const int ryl = __builtin_amdgcn_readlane(iyl, L), ryh = __builtin_amdgcn_readlane(iyh, L);
                const float dj = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(d), L));
                float sum = 0.0f;
                // the rectangle's pixels in row-major order, 64 a pass: lane pixel (px, py) relative
                // to the rectangle, advanced by 64 pixels with a carry
                const int w = rxh - rxl + 1, hgt = ryh - ryl + 1;
                const float rw = 1.0f / float(w);
                int py = int(float(lane) * rw), px = int(lane) - py * w;
                if (px < 0) { --py; px += w; } else if (px >= w) { ++py; px -= w; }
                int dy = int(64.0f * rw), dx = 64 - dy * w;
                if (dx < 0) { --dy; dx += w; } else if (dx >= w) { ++dy; dx -= w; }
                for (int base = 0; base < w * hgt; base += 64) {
                    const int x = rxl + px, y = ryl + py;
                    const float fx = float(x), fy = float(y);
                    // the sample position as plane_kernel computes it
                    const float qx = fmaf(dj, c[0], fmaf(h[0][0], fx, fmaf(h[0][1], fy, h[0][2])));
                    const float qy = fmaf(dj, c[1], fmaf(h[1][0], fx, fmaf(h[1][1], fy, h[1][2])));
                    const float qz = fmaf(dj, c[2], fmaf(h[2][0], fx, fmaf(h[2][1], fy, h[2][2])));
                    const bool mine = py < hgt && qx >= qlo[0] && qx < qhi[0] && qy >= qlo[1] && qy < qhi[1] &&
                                      qz >= qlo[2] && qz < qhi[2];
                    if (mine) {
                        const uint32_t ofs = __umul24(uint32_t(qz), uint32_t(kPBX * kPBY * 8)) +
                                             (__umul24(uint32_t(qy), uint32_t(kPBX * 8)) + ((uint32_t(qx) << 3) - base8));
                        const float2* p = reinterpret_cast<const float2*>(reinterpret_cast<const char*>(box) + ofs);
                        const float2* q = p + kPBX;
                        sum += triDeltaZ(p[0], p[1], q[0], q[1], __builtin_amdgcn_fractf(qx),
                                         __builtin_amdgcn_fractf(qy), __builtin_amdgcn_fractf(qz));
                    }
                    px += dx;
                    py += dy;
                    if (px >= w) { px -= w; ++py; }
                }
                sum = waveSumDpp(sum);
                if (lane == 0 && sum != 0.0f)
                    atomicAdd(acc + size_t(o) * n_dist + uint32_t(__builtin_amdgcn_readlane(int(j), L)),
                              (unsigned long long)__double2ll_rn(double(sum) * fix_scale));
            }
            __builtin_amdgcn_wave_barrier();
            listed = 0;
        };
        for (uint32_t jb = 0; jb < n_dist; jb += 64) {
            const uint32_t j = jb + lane;
            const float d = j < n_dist ? dist[j] : 0.0f;
            const bool in = j < n_dist && d >= dmin - dm && d <= dmax + dm;
            const uint64_t m = __ballot(in);
            const uint32_t cnt = uint32_t(__popcll(m));
            if (cnt == 0) continue; // uniform
            if (listed + cnt > 64) flush();
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
            if (in) {
                s_j[wave][listed + rank] = j;
                s_d[wave][listed + rank] = d;
            }
            listed += cnt;
        }
        if (listed) flush();
    }
}

// out[d * out_stride + o] = scale * total(o, d) of the fixed-point totals.
__global__ __launch_bounds__(kBlock) void plane_fix_kernel(const unsigned long long* __restrict__ acc, uint32_t n_orient,
                                                           uint32_t n_dist, double inv_fix, float scale,
                                                           float* __restrict__ out, uint32_t out_stride)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_orient * n_dist) return;
    const uint32_t o = i % n_orient, d = i / n_orient;
    const float total = float(double((long long)acc[size_t(o) * n_dist + d]) * inv_fix);
    out[size_t(d) * out_stride + o] = total * scale;
}

// Largest |value| of a brick copy's voxels (the v of each 8-byte texel of every 16-byte element
// half), as the bits of a non-negative float into *d_max (atomicMax on the bits orders them as the
// values); NaN and infinity give bits at or above 0x7f800000.
__global__ __launch_bounds__(kBlock) void brick_maxabs_kernel(const float2* __restrict__ t, size_t n, uint32_t* d_max)
{
    uint32_t m = 0;
    for (size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += size_t(gridDim.x) * kBlock)
        m = max(m, __float_as_uint(t[i].x) & 0x7fffffffu);
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) m = max(m, uint32_t(__shfl_xor(int(m), s, 64)));
    if ((threadIdx.x & 63) == 0) atomicMax(d_max, m);
}

#define CTLK_DISPATCH(KERNEL, vol, grid, ...)                                                       \
    do {                                                                                           \
        const bool il_ = (vol).layout == Layout::Interleaved;                                      \
        switch ((vol).storage) {                                                                   \
        case Storage::F16:                                                                         \
            if (il_) KERNEL<Storage::F16, Layout::Interleaved><<<grid, kBlock, 0, s>>>(__VA_ARGS__); \
            else KERNEL<Storage::F16, Layout::Plain><<<grid, kBlock, 0, s>>>(__VA_ARGS__);         \
            break;                                                                                 \
        case Storage::BF16:                                                                        \
            if (il_) KERNEL<Storage::BF16, Layout::Interleaved><<<grid, kBlock, 0, s>>>(__VA_ARGS__); \
            else KERNEL<Storage::BF16, Layout::Plain><<<grid, kBlock, 0, s>>>(__VA_ARGS__);        \
            break;                                                                                 \
        default:                                                                                   \
            if (il_) KERNEL<Storage::F32, Layout::Interleaved><<<grid, kBlock, 0, s>>>(__VA_ARGS__); \
            else KERNEL<Storage::F32, Layout::Plain><<<grid, kBlock, 0, s>>>(__VA_ARGS__);         \
            break;                                                                                 \
        }                                                                                          \
    } while (0)

Ranges makeRanges(const VolumeView& vol, const float ranges[6])
{
    Ranges r;
    for (int k = 0; k < 3; ++k) {
        r.lo[k] = ranges[2 * k];
        r.scale[k] = (float(vol.n[k]) - 1.0f) / (ranges[2 * k + 1] - ranges[2 * k]);
    }
    return r;
}

} // namespace

hipError_t launchSlice(const VolumeView& vol, const PlaneMap* d_maps, uint32_t n_planes, uint32_t width,
                       uint32_t height, float* d_out, hipStream_t s)
{
    const uint32_t tiles_x = (width + 15) / 16, tiles_y = (height + 15) / 16;
    const uint64_t per_plane = uint64_t(tiles_x) * tiles_y, blocks = per_plane * n_planes;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffull) return hipErrorInvalidConfiguration;
    const float cx = 0.5f * float(width - 1u), cy = 0.5f * float(height - 1u);
    CTLK_DISPATCH(slice_kernel, vol, uint32_t(blocks), vol, d_maps, width, height, tiles_x, uint32_t(per_plane),
                  cx, cy, d_out);
    return hipGetLastError();
}

hipError_t launchResample(const VolumeView& vol, const float ranges[6], const float* d_p1, uint32_t n1,
                          const float* d_p2, uint32_t n2, const float* d_p3, uint32_t n3, float* d_out,
                          hipStream_t s)
{
    const uint64_t count = uint64_t(n1) * n2 * n3;
    if (count == 0) return hipSuccess;
    if (count > 0xffffffffull - kBlock) return hipErrorInvalidConfiguration;
    const Ranges r = makeRanges(vol, ranges);
    const uint32_t blocks = uint32_t((count + kBlock - 1) / kBlock);
    CTLK_DISPATCH(resample_kernel, vol, blocks, vol, r, d_p1, d_p2, d_p3, n1, n2, uint32_t(count), d_out);
    return hipGetLastError();
}

hipError_t launchSamplePoints(const VolumeView& vol, const float ranges[6], const float* d_xyz, size_t n,
                              float* d_out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (uint64_t(n) + kBlock - 1) / kBlock;
    if (blocks > 0x7fffffffull) return hipErrorInvalidConfiguration;
    const Ranges r = makeRanges(vol, ranges);
    CTLK_DISPATCH(sample_kernel, vol, uint32_t(blocks), vol, r, d_xyz, n, d_out);
    return hipGetLastError();
}

bool brickShape(const uint32_t n[3], BrickKind want, BrickKind* kind, uint32_t nb[3], size_t* bytes)
{
    const uint64_t p[3] = { uint64_t(n[0]) + 2, uint64_t(n[1]) + 2, uint64_t(n[2]) + 2 };
    const uint64_t t8 = ((p[0] + 3) / 4) * ((p[1] + 1) / 2) * ((p[2] + 1) / 2) * 128;
    BrickKind k = want;
    if (k == BrickKind::Auto) k = t8 <= kTexel8CacheBytes ? BrickKind::Texel8 : BrickKind::Pair16;
    const uint32_t bx = k == BrickKind::Texel8 ? 4 : 2;
    nb[0] = uint32_t((p[0] + bx - 1) / bx);
    nb[1] = uint32_t((p[1] + 1) / 2);
    nb[2] = uint32_t((p[2] + 1) / 2);
    const uint64_t b = uint64_t(nb[0]) * nb[1] * nb[2] * 128;
    if (b >= (1ull << 32) - 256 || uint64_t(nb[0]) * 128 >= (1ull << 24)) return false;
    *kind = k;
    *bytes = size_t(b);
    return true;
}

hipError_t launchBrick(const VolumeView& vol, BrickKind kind, float* d_dst, hipStream_t s)
{
    uint32_t nb[3];
    size_t bytes = 0;
    BrickKind k;
    if (!brickShape(vol.n, kind, &k, nb, &bytes) || k != kind) return hipErrorInvalidValue;
    if (kind == BrickKind::Texel8) {
        const size_t total = bytes / 8;
        const size_t blocks = std::min<size_t>((total + kBlock - 1) / kBlock, 1u << 20);
        texel8_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(vol, nb[0], nb[1], total, reinterpret_cast<float2*>(d_dst));
    } else {
        const size_t total = bytes / 16;
        const size_t blocks = std::min<size_t>((total + kBlock - 1) / kBlock, 1u << 20);
        pair16_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(vol, nb[0], nb[1], total, reinterpret_cast<float4*>(d_dst));
    }
    return hipGetLastError();
}

hipError_t launchPlaneIntegrals(const VolumeView& vol, const float* d_brick, BrickKind brick_kind, const PlaneMap* d_maps,
                                const float* d_dist,
                                const PlaneJob& job, float* d_partial, float scale, float* d_out,
                                uint32_t out_stride, hipStream_t s)
{
    if (job.n_orient == 0 || job.n_dist == 0) return hipSuccess;
    if (job.splits == 0 || job.tiles_x == 0 || job.tiles_y == 0 || job.n_tiles == 0) return hipErrorInvalidValue;
    const uint64_t chunks = (uint64_t(job.n_dist) + kPlaneDists - 1) / kPlaneDists;
    const uint64_t blocks = chunks * job.n_orient * job.splits;
    if (blocks > 0x7fffffffull || uint64_t(job.n_orient) * job.n_dist > 0xffffffffull - kBlock)
        return hipErrorInvalidConfiguration;
    const uint32_t n0 = vol.n[0], n1 = vol.n[1], n2 = vol.n[2];
    const OccTable oc{ vol.occ, { vol.occ_n[0], vol.occ_n[1], vol.occ_n[2] } };
    if (d_brick) {
        uint32_t nb[3];
        size_t bytes = 0;
        BrickKind k;
        if (!brickShape(vol.n, brick_kind, &k, nb, &bytes) || k != brick_kind) return hipErrorInvalidValue;
        if (k == BrickKind::Texel8) {
            const Texel8Fetch f{ d_brick, uint32_t(bytes), nb[0] * 128u, nb[0] * nb[1] * 128u };
            plane_kernel<Texel8Fetch><<<uint32_t(blocks), kBlock, 0, s>>>(f, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
        } else {
            const Pair16Fetch f{ d_brick, uint32_t(bytes), nb[0] * 128u, nb[0] * nb[1] * 128u };
            plane_kernel<Pair16Fetch><<<uint32_t(blocks), kBlock, 0, s>>>(f, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
        }
    } else {
        const bool il = vol.layout == Layout::Interleaved;
        switch (vol.storage) {
        case Storage::F16:
            if (il) plane_kernel<TexelFetch<Storage::F16, Layout::Interleaved>><<<uint32_t(blocks), kBlock, 0, s>>>({ vol }, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
            else plane_kernel<TexelFetch<Storage::F16, Layout::Plain>><<<uint32_t(blocks), kBlock, 0, s>>>({ vol }, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
            break;
        case Storage::BF16:
            if (il) plane_kernel<TexelFetch<Storage::BF16, Layout::Interleaved>><<<uint32_t(blocks), kBlock, 0, s>>>({ vol }, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
            else plane_kernel<TexelFetch<Storage::BF16, Layout::Plain>><<<uint32_t(blocks), kBlock, 0, s>>>({ vol }, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
            break;
        default:
            if (il) plane_kernel<TexelFetch<Storage::F32, Layout::Interleaved>><<<uint32_t(blocks), kBlock, 0, s>>>({ vol }, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
            else plane_kernel<TexelFetch<Storage::F32, Layout::Plain>><<<uint32_t(blocks), kBlock, 0, s>>>({ vol }, n0, n1, n2, d_maps, d_dist, job, oc, d_partial);
            break;
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const uint32_t rb = uint32_t((uint64_t(job.n_orient) * job.n_dist + kBlock - 1) / kBlock);
    plane_reduce_kernel<<<rb, kBlock, 0, s>>>(d_partial, job.n_orient, job.n_dist, job.splits, scale, d_out,
                                              out_stride);
    return hipGetLastError();
}

uint32_t planeBricks(const uint32_t n[3])
{
    return ((n[0] + kPlaneBrick) / kPlaneBrick) * ((n[1] + kPlaneBrick) / kPlaneBrick) *
           ((n[2] + kPlaneBrick) / kPlaneBrick);
}

hipError_t launchBrickMaxAbs(const float* d_brick, size_t bytes, uint32_t* d_max, hipStream_t s)
{
    const size_t n = bytes / 8;
    const size_t blocks = std::min<size_t>((n + kBlock - 1) / kBlock, 4096);
    if (n == 0) return hipSuccess;
    brick_maxabs_kernel<<<uint32_t(blocks), kBlock, 0, s>>>(reinterpret_cast<const float2*>(d_brick), n, d_max);
    return hipGetLastError();
}

hipError_t launchPlaneBricks(const VolumeView& vol, const float* d_brick, BrickKind brick_kind, const PlaneMap* d_maps,
                             const PlaneInv* d_inv, const float* d_dist, uint32_t n_orient, uint32_t n_dist,
                             uint32_t width, uint32_t height, uint32_t o_chunk, int fix_shift,
                             unsigned long long* d_acc, float scale, float* d_out, uint32_t out_stride,
                             hipStream_t s)
{
    if (n_orient == 0 || n_dist == 0) return hipSuccess;
    if (!d_brick || o_chunk == 0 || width == 0 || height == 0 || fix_shift < 0 || fix_shift > 62)
        return hipErrorInvalidValue;
    uint32_t nb[3];
    size_t bytes = 0;
    BrickKind k;
    if (!brickShape(vol.n, brick_kind, &k, nb, &bytes) || k != brick_kind) return hipErrorInvalidValue;
    const uint64_t count = uint64_t(n_orient) * n_dist;
    const uint64_t blocks = uint64_t(planeBricks(vol.n)) * ((n_orient + o_chunk - 1) / o_chunk);
    if (blocks > 0x7fffffffull || count > 0xffffffffull - kBlock) return hipErrorInvalidConfiguration;
    hipError_t e = launchZero(reinterpret_cast<float*>(d_acc), size_t(count) * 2, s);
    if (e != hipSuccess) return e;
    const OccTable oc{ vol.occ, { vol.occ_n[0], vol.occ_n[1], vol.occ_n[2] } };
    const double fix = std::ldexp(1.0, fix_shift);
    if (k == BrickKind::Texel8) {
        const Texel8Src f{ d_brick, uint32_t(bytes), nb[0] * 128u, nb[0] * nb[1] * 128u };
        plane_brick_kernel<Texel8Src><<<uint32_t(blocks), kBlock, 0, s>>>(
            f, vol.n[0], vol.n[1], vol.n[2], oc, d_maps, d_inv, d_dist, n_orient, n_dist, width, height, o_chunk, fix,
            d_acc);
    } else {
        const Pair16Src f{ d_brick, uint32_t(bytes), nb[0] * 128u, nb[0] * nb[1] * 128u };
        plane_brick_kernel<Pair16Src><<<uint32_t(blocks), kBlock, 0, s>>>(
            f, vol.n[0], vol.n[1], vol.n[2], oc, d_maps, d_inv, d_dist, n_orient, n_dist, width, height, o_chunk, fix,
            d_acc);
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t rb = uint32_t((count + kBlock - 1) / kBlock);
    plane_fix_kernel<<<rb, kBlock, 0, s>>>(d_acc, n_orient, n_dist, 1.0 / fix, scale, d_out, out_stride);
    return hipGetLastError();
}

} // namespace ctlk
