// Texel samplers of resident volumes (device code shared by the projector and the volume
// processing kernels).
//
// A resident volume is padded by one zero texel on every side (VolumeView). The OpenCL image
// sampler the reference reads volumes with (CLK_FILTER_LINEAR, CLK_ADDRESS_CLAMP, border 0,
// unnormalised coordinate u) interpolates the voxels around u - 1/2; in the padded frame that is
// the texel position q = u + 1/2, whose taps are texels floor(q) and floor(q) + 1.
#pragma once

#include "kernels.h"

#include <cstdint>

namespace ctlk {

namespace {

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// buffer resource word 3 for gfx9-family raw buffers
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t makeRsrc(const void* p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, kRsrcWord3);
}

// ---------------------------------------------------------------------------------------------
// trilinear / nearest sample at texel-frame position q (padded frame, blo <= q < n+1)

__device__ __forceinline__ float lerp(float a, float b, float w) { return fmaf(w, b - a, a); }

__device__ __forceinline__ float halfLo(uint32_t u) { return float(__builtin_bit_cast(half2_t, u).x); }
__device__ __forceinline__ float halfHi(uint32_t u) { return float(__builtin_bit_cast(half2_t, u).y); }
__device__ __forceinline__ float bfLo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bfHi(uint32_t u) { return __uint_as_float(u & 0xffff0000u); }

// x-lerp of a packed pair (a = element at x, b = element at x+1) in FP32 arithmetic; for FP16 the
// two FMAs read the halves directly (v_fma_mix).
template <Storage ST>
__device__ __forceinline__ float lerpPair(uint32_t lo_word, bool lo_hi, uint32_t hi_word, bool hi_hi, float w)
{
    if constexpr (ST == Storage::F16) {
        const float a = lo_hi ? halfHi(lo_word) : halfLo(lo_word);
        const float b = hi_hi ? halfHi(hi_word) : halfLo(hi_word);
        return fmaf(w, b, fmaf(-w, a, a));
    } else {
        const float a = lo_hi ? bfHi(lo_word) : bfLo(lo_word);
        const float b = hi_hi ? bfHi(hi_word) : bfLo(hi_word);
        return lerp(a, b, w);
    }
}

// Trilinear value from the four (v(z), v(z+1) - v(z)) texels around a sample: a0, a1 at x, x+1 of
// row y, b0, b1 of row y+1. The z-lerps take the stored difference (one FMA each, the same
// rounding as lerp), then x, then y.
__device__ __forceinline__ float triDeltaZ(float2 a0, float2 a1, float2 b0, float2 b1, float wx, float wy,
                                           float wz)
{
    const float c00 = fmaf(wz, a0.y, a0.x), c10 = fmaf(wz, a1.y, a1.x); // y0: x0, x1
    const float c01 = fmaf(wz, b0.y, b0.x), c11 = fmaf(wz, b1.y, b1.x); // y1: x0, x1
    return lerp(lerp(c00, c10, wx), lerp(c01, c11, wx), wy);
}

template <Storage ST, Layout LY> struct Sampler;

// FP32, one voxel per texel: four 8-byte loads (x-pairs at y/z corners). The z-differences are
// formed here, so the result is the interleaved sampler's bit for bit.
template <> struct Sampler<Storage::F32, Layout::Plain> {
    static constexpr uint32_t kShift = 2;
    static __device__ __forceinline__ float2 load(__amdgpu_buffer_rsrc_t r, uint32_t ofs, uint32_t so)
    {
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, ofs, so, 0));
    }
    static __device__ __forceinline__ float tri(__amdgpu_buffer_rsrc_t r, const Pitch& p, uint32_t ofs,
                                                float wx, float wy, float wz)
    {
        const float2 t00 = load(r, ofs, 0), t10 = load(r, ofs, p.y); // z0: rows y, y+1
        const float2 t01 = load(r, ofs, p.z), t11 = load(r, ofs, p.yz); // z1
        return triDeltaZ(make_float2(t00.x, t01.x - t00.x), make_float2(t00.y, t01.y - t00.y),
                         make_float2(t10.x, t11.x - t10.x), make_float2(t10.y, t11.y - t10.y), wx, wy, wz);
    }
};

// FP16/BF16, one voxel per texel: four 4-byte loads.
template <Storage ST> struct PlainHalf {
    static constexpr uint32_t kShift = 1;
    static __device__ __forceinline__ float tri(__amdgpu_buffer_rsrc_t r, const Pitch& p, uint32_t ofs,
                                                float wx, float wy, float wz)
    {
        const uint32_t t00 = __builtin_amdgcn_raw_buffer_load_b32(r, ofs, 0, 0);
        const uint32_t t10 = __builtin_amdgcn_raw_buffer_load_b32(r, ofs, p.y, 0);
        const uint32_t t01 = __builtin_amdgcn_raw_buffer_load_b32(r, ofs, p.z, 0);
        const uint32_t t11 = __builtin_amdgcn_raw_buffer_load_b32(r, ofs, p.yz, 0);
        // z, then x, then y, as the FP32 samplers
        const float c00 = lerpPair<ST>(t00, false, t01, false, wz), c10 = lerpPair<ST>(t00, true, t01, true, wz);
        const float c01 = lerpPair<ST>(t10, false, t11, false, wz), c11 = lerpPair<ST>(t10, true, t11, true, wz);
        return lerp(lerp(c00, c10, wx), lerp(c01, c11, wx), wy);
    }
};
template <> struct Sampler<Storage::F16, Layout::Plain> : PlainHalf<Storage::F16> {};
template <> struct Sampler<Storage::BF16, Layout::Plain> : PlainHalf<Storage::BF16> {};

// FP32 (v, dz) texels: two 16-byte loads, rows y and y+1, each the texels at x and x+1.
template <> struct Sampler<Storage::F32, Layout::Interleaved> {
    static constexpr uint32_t kShift = 3;
    static __device__ __forceinline__ float tri(__amdgpu_buffer_rsrc_t r, const Pitch& p, uint32_t ofs,
                                                float wx, float wy, float wz)
    {
        const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, ofs, 0, 0));
        const float4 b = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, ofs, p.y, 0));
        return triDeltaZ(make_float2(a.x, a.y), make_float2(a.z, a.w), make_float2(b.x, b.y),
                         make_float2(b.z, b.w), wx, wy, wz);
    }
};

// Trilinear value from the (y,z)-quads at x (words p0: z, p1: z+1) and x+1 (q0, q1), each word
// [v(y) v(y+1)]: z, then x, then y, as the FP32 samplers.
template <Storage ST>
__device__ __forceinline__ float quadTri(uint32_t p0, uint32_t p1, uint32_t q0, uint32_t q1, float wx, float wy,
                                         float wz)
{
    const float c00 = lerpPair<ST>(p0, false, p1, false, wz), c10 = lerpPair<ST>(q0, false, q1, false, wz); // y0
    const float c01 = lerpPair<ST>(p0, true, p1, true, wz), c11 = lerpPair<ST>(q0, true, q1, true, wz);     // y1
    return lerp(lerp(c00, c10, wx), lerp(c01, c11, wx), wy);
}

// FP16/BF16 (y,z)-quad texels: one 16-byte load = the quads at x and x+1, each
// [v(y,z) v(y+1,z) | v(y,z+1) v(y+1,z+1)] as two 32-bit words.
template <Storage ST> struct QuadHalf {
    static constexpr uint32_t kShift = 3;
    static __device__ __forceinline__ float tri(__amdgpu_buffer_rsrc_t r, const Pitch&, uint32_t ofs,
                                                float wx, float wy, float wz)
    {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const u4 t = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(r, ofs, 0, 0));
        return quadTri<ST>(t.x, t.y, t.z, t.w, wx, wy, wz);
    }
};
template <> struct Sampler<Storage::F16, Layout::Interleaved> : QuadHalf<Storage::F16> {};
template <> struct Sampler<Storage::BF16, Layout::Interleaved> : QuadHalf<Storage::BF16> {};

template <Storage ST, Layout LY>
__device__ __forceinline__ float trilinear(__amdgpu_buffer_rsrc_t r, const Pitch& p, float qx,
                                           float qy, float qz)
{
    const int ix = int(qx), iy = int(qy), iz = int(qz); // q >= 0: truncation is floor
    const float wx = __builtin_amdgcn_fractf(qx);
    const float wy = __builtin_amdgcn_fractf(qy);
    const float wz = __builtin_amdgcn_fractf(qz);
    const uint32_t idx = uint32_t(ix) + __umul24(uint32_t(iy), p.ty) + __umul24(uint32_t(iz), p.tz);
    return Sampler<ST, LY>::tri(r, p, idx << Sampler<ST, LY>::kShift, wx, wy, wz);
}

template <Storage ST, Layout LY>
__device__ __forceinline__ float nearest(__amdgpu_buffer_rsrc_t r, const Pitch& p, float qx,
                                         float qy, float qz)
{
    const uint32_t idx = uint32_t(int(qx)) + __umul24(uint32_t(int(qy)), p.ty) +
                         __umul24(uint32_t(int(qz)), p.tz);
    constexpr uint32_t shift = LY == Layout::Interleaved ? 3 : (ST == Storage::F32 ? 2 : 1);
    if constexpr (ST == Storage::F32) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, idx << shift, 0, 0));
    } else {
        const uint32_t raw = __builtin_amdgcn_raw_buffer_load_b16(r, idx << shift, 0, 0);
        if constexpr (ST == Storage::F16) return float(__builtin_bit_cast(_Float16, uint16_t(raw)));
        else return __uint_as_float(raw << 16);
    }
}

// ---------------------------------------------------------------------------------------------
// Volumes of 4 GiB or more (VolumeView::large): a buffer's offsets (index * stride included) are
// 32-bit, so these read with 64-bit addresses, row (y + z*(Y+2)) times the row bytes plus the
// byte offset of x. Every tap of a sample in the support (blo <= q < n+1) lies inside the array.
// Only the interleaved layouts are read this way.

typedef uint32_t wideu4 __attribute__((ext_vector_type(4), aligned(8))); // 8-byte aligned texel pairs
typedef uint32_t wideu2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ const char* wideAddr(const void* base, const Pitch& p, uint32_t ix, uint32_t iy,
                                                uint32_t iz)
{
    const uint32_t row = iy + __umul24(iz, p.py);
    return static_cast<const char*>(base) + (uint64_t(row) * p.y + (ix << 3));
}

template <Storage ST>
__device__ __forceinline__ float trilinearWide(const void* base, const Pitch& p, float qx, float qy, float qz)
{
    const uint32_t ix = uint32_t(int(qx)), iy = uint32_t(int(qy)), iz = uint32_t(int(qz)); // q >= 0
    const float wx = __builtin_amdgcn_fractf(qx);
    const float wy = __builtin_amdgcn_fractf(qy);
    const float wz = __builtin_amdgcn_fractf(qz);
    const char* t = wideAddr(base, p, ix, iy, iz);
    if constexpr (ST == Storage::F32) { // (v, dz) texels at x and x+1 of rows y and y+1
        const float4 a = __builtin_bit_cast(float4, *reinterpret_cast<const wideu4*>(t));
        const float4 b = __builtin_bit_cast(float4, *reinterpret_cast<const wideu4*>(t + p.y));
        return triDeltaZ(make_float2(a.x, a.y), make_float2(a.z, a.w), make_float2(b.x, b.y),
                         make_float2(b.z, b.w), wx, wy, wz);
    } else { // (y,z)-quads at x and x+1
        const wideu4 q = *reinterpret_cast<const wideu4*>(t);
        return quadTri<ST>(q.x, q.y, q.z, q.w, wx, wy, wz);
    }
}

template <Storage ST>
__device__ __forceinline__ float nearestWide(const void* base, const Pitch& p, float qx, float qy, float qz)
{
    const wideu2 t = *reinterpret_cast<const wideu2*>(
        wideAddr(base, p, uint32_t(int(qx)), uint32_t(int(qy)), uint32_t(int(qz)))); // the texel at x
    if constexpr (ST == Storage::F32) return __uint_as_float(t.x);
    else if constexpr (ST == Storage::F16) return float(__builtin_bit_cast(_Float16, uint16_t(t.x & 0xffffu)));
    else return __uint_as_float(t.x << 16);
}

// OpenCL linear image read at unnormalised coordinate u (the reference's read_imagef with
// CLK_FILTER_LINEAR | CLK_ADDRESS_CLAMP) = sampleTexel at q = u + 1/2: zero wherever all eight
// taps are border texels, that is unless 0 <= q < n + 1 on every axis (nhi = n + 1). NaN
// coordinates read zero.
// Branch-free, so that the loads of several samples can be in flight together: a position
// outside the support reads texel 0 instead and its value is discarded.
template <Storage ST, Layout LY>
__device__ __forceinline__ float sampleTexel(__amdgpu_buffer_rsrc_t rs, const Pitch& p, const float nhi[3],
                                             float qx, float qy, float qz)
{
    const bool in = qx >= 0.0f && qy >= 0.0f && qz >= 0.0f && qx < nhi[0] && qy < nhi[1] && qz < nhi[2];
    const float v = trilinear<ST, LY>(rs, p, in ? qx : 0.0f, in ? qy : 0.0f, in ? qz : 0.0f);
    return in ? v : 0.0f;
}

template <Storage ST, Layout LY>
__device__ __forceinline__ float sampleImage(__amdgpu_buffer_rsrc_t rs, const Pitch& p, const float nhi[3],
                                             float ux, float uy, float uz)
{
    return sampleTexel<ST, LY>(rs, p, nhi, ux + 0.5f, uy + 0.5f, uz + 0.5f);
}

} // namespace
} // namespace ctlk
