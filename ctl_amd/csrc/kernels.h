// Device data structures and kernel launchers shared by the host runtime and kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace ctlk {

enum class Storage : int { F32 = 0, F16 = 1, BF16 = 2 };

inline size_t storageBytes(Storage s) { return s == Storage::F32 ? 4 : 2; }

// Texel layouts of a resident volume.
//   Plain:       one voxel per texel (4 B FP32, 2 B FP16/BF16).
//   Interleaved: 8-byte texels that carry the neighbours a trilinear sample needs, so one 16-byte
//                load fetches the taps of two x-adjacent texels:
//                  FP32       texel(x,y,z) = [v(x,y,z), v(x,y,z+1) - v(x,y,z)] (2 loads/sample;
//                             the stored difference makes each z-lerp one FMA)
//                  FP16/BF16  texel(x,y,z) = [v(x,y,z), v(x,y+1,z), v(x,y,z+1), v(x,y+1,z+1)]
//                                                                              (1 load/sample)
enum class Layout : int { Plain = 0, Interleaved = 1 };

inline size_t texelBytes(Storage s, Layout l) { return l == Layout::Interleaved ? 8 : storageBytes(s); }

// A resident volume in HBM: (X+2) x (Y+2) x (Z+2) texels, x fastest, with a one-texel zero
// border on every side, so that every tap of a sample inside the support is an in-bounds read.
// A volume of 4 GiB or more (`large`, interleaved layout only) is beyond the 32-bit offsets of a
// buffer: its gathers use 64-bit addresses (texel row y + z*(Y+2) times row_bytes, plus x), and
// its LDS staging loads a buffer window that starts at the staged box's first slice.
struct VolumeView {
    const void* data;
    uint32_t n[3];      // interior voxel counts
    uint32_t pitch_y;   // X + 2        (texels per row)
    uint32_t pitch_z;   // (X+2)*(Y+2)  (texels per slice)
    uint32_t bytes;     // total bytes of the padded array (buffer range for the gather loads); for
                        // a large volume the largest 32-bit range, used by no gather
    Storage storage;
    Layout layout;
    uint32_t large;     // 1: 64-bit addressing
    uint32_t row_bytes; // bytes per texel row, (X+2) * texel bytes
    uint32_t rows;      // texel rows, (Y+2) * (Z+2)
    // Occupancy of the padded voxel grid in bricks of kOccBrick^3 voxels, as a summed-area table:
    // occ[((k * (oy + 1)) + j) * (ox + 1) + i] = number of bricks holding a non-zero voxel (NaN
    // counts) among bricks [0, i) x [0, j) x [0, k), for (ox, oy, oz) = occ_n. NULL: no table.
    const uint32_t* occ;
    uint32_t occ_n[3];
};

// Edge of the occupancy table's bricks, in voxels of the padded grid.
constexpr uint32_t kOccShift = 2;
constexpr uint32_t kOccBrick = 1u << kOccShift;
// Bricks per axis of a padded grid of n + 2 voxels, and the table's size in bytes.
inline uint32_t occBricks(uint32_t n) { return (n + 2 + kOccBrick - 1) >> kOccShift; }
inline size_t occBytes(const uint32_t n[3])
{
    return size_t(occBricks(n[0]) + 1) * (occBricks(n[1]) + 1) * (occBricks(n[2]) + 1) * sizeof(uint32_t);
}

// Volume geometry after upsampling, as the host passes it (FP32, raycasterprojector.cpp:169-179).
struct VolumeGeom {
    float voxel_mm[3];
    float offset_mm[3];
    uint32_t n[3];
    float increment_mm; // min(voxel size) * raySampling
};

// Per (view, module) ray geometry written by the FP64 setup kernel: the reference's packed QR of
// M (decomposeM, raycasterprojector.cpp:527-541: Q row-major, then R00 R01 R02 R11 R12 R22), the
// FP32 source and volume corner in voxel units (raycasterprojector_interp.cl:35-36), and `c`, the
// reference's corner-to-source vector plus the gathers' frame shift (sample i is at
// fma(i, step, c) in the padded texel frame).
struct alignas(16) RayGeom {
    double qr[15];
    float src[3];
    float corner[3];
    float c[3];
    uint32_t valid; // 0 if any of these values is not finite (singular M)
};
static_assert(sizeof(RayGeom) == 160, "RayGeom is one 160-byte record");

struct LaunchShape {
    uint32_t wave_w = 8, wave_h = 8; // pixel tile per wavefront (wave_w * wave_h = 64)
    uint32_t block_waves_x = 2, block_waves_y = 2; // waves per workgroup
    uint32_t unroll = 4;
    uint32_t blocks_per_cu = 0; // 0: as many as registers allow; else capped with reserved LDS
    // LDS staging (interleaved volumes, trilinear, one ray per pixel): samples per chunk
    // (0 = unstaged) and LDS box capacity in texels (a multiple of 128)
    uint32_t chunk = 64;
    uint32_t lds_texels = 3072; // 24 KiB: six workgroups per CU
};

// Default launch shape: 8x8 wavefront tiles unrolled 4x for the gathers from HBM; 16x4 tiles
// unrolled 3x (a third of the samples gathered through L1) with 64-sample chunks in a 3072-texel
// LDS box for the staged kernel (the fastest shapes measured on the 256^3 / 1024x768 workload:
// 76.3 ms at 360 views against 78.6 with unroll 4).
inline LaunchShape defaultShape(bool staged)
{
    LaunchShape s;
    if (staged) {
        s.wave_w = 16;
        s.wave_h = 4;
        s.unroll = 3;
    } else {
        s.chunk = 0;
    }
    return s;
}

// The shape for modules no wider than one wavefront tile (the 16-channel modules of cylindrical
// detectors): the workgroup's wavefronts stacked vertically instead of a column of them idle,
// and for the staged kernel, whose ray bundle is then taller, a box of 6400 texels (three
// workgroups per CU; 32.6 ms for 500 views of 40 modules of 16 x 64 pixels, against 36.2 ms with
// 48-sample chunks in 3840 texels and 38.2 ms with 7680 texels). Unchanged unless the shape has
// more than one wavefront across and channels <= its tile width. The staged kernel keeps unroll
// 4 there (a quarter of the samples through L1: 132.0 ms for the 2000-view helical scan against
// 141.6 ms with unroll 3). The unstaged kernel keeps at most five workgroups per CU wherever one
// workgroup spans a module's width: the workgroups of neighbouring modules share few cache lines,
// and more resident waves only evict each other's (FP16, 500 views of the helical scan: 25.3 ms
// against 27.8 uncapped, 25.3 with four; on 1024-channel panels the cap costs 5-12%).
inline LaunchShape narrowModuleShape(LaunchShape s, uint32_t channels, bool staged)
{
    if (channels <= s.wave_w && s.block_waves_x >= 2) {
        s.block_waves_y *= s.block_waves_x;
        s.block_waves_x = 1;
        if (staged) {
            s.lds_texels = 6400;
            s.unroll = 4;
        }
    }
    if (!staged && channels <= s.wave_w * s.block_waves_x) s.blocks_per_cu = 5;
    return s;
}

// LDS of a CU and of one workgroup on gfx950; kernels keep at most kStaticLds bytes of static LDS
// beside the dynamic allocation.
constexpr uint32_t kLdsBytes = 160u * 1024u;
constexpr uint32_t kStaticLds = 1024u;
// Largest LDS box of the staged kernel, in 8-byte texels (a multiple of 128).
constexpr uint32_t kMaxLdsTexels = ((kLdsBytes - kStaticLds) / 8u) & ~127u;

// Dynamic LDS that caps the resident workgroups per CU at `blocks`.
inline uint32_t ldsForBlocks(uint32_t blocks)
{
    return blocks ? (((kLdsBytes - kStaticLds) / blocks) / 256u) * 256u : 0u;
}

struct RayParams {
    uint32_t channels, rows, modules, views;
    uint32_t rpp_x, rpp_y;
    int interpolate;
};

// True if a projection with this storage, layout and mode can use the LDS-staged kernel.
inline bool stageable(Storage, Layout ly, const RayParams& rp)
{
    return ly == Layout::Interleaved && rp.interpolate && rp.rpp_x == 1 && rp.rpp_y == 1;
}

struct Pitch {
    uint32_t y, z, yz; // byte offsets of the (y+1), (z+1), (y+1,z+1) texel rows
    uint32_t ty, tz;   // texels per row / slice
    uint32_t py;       // texel rows per slice (64-bit addressing of large volumes)
};

// Launch constants of one projection (tile decomposition, step, box, texel pitches).
struct CastConst {
    uint32_t tiles_x, tiles_y, tiles_per_vm;
    uint32_t wave_w, wave_h, bwx, bwy;
    float inc;
    float vs[3];
    float nhi[3]; // n + 1 per axis
    float blo;    // 0 (trilinear) or 1 (nearest)
    Pitch pitch;
    uint32_t chunk;      // LDS-staged kernel: samples per ray and chunk (0: unstaged kernel)
    uint32_t lds_texels; // LDS-staged kernel: capacity of the LDS box in 8-byte texels
    uint32_t accumulate; // 1: add to the output instead of overwriting it
    uint32_t lane_x, lane_y; // lane -> pixel in a wavefront tile: bit-gather codes (makeCastConst)
    // diagnostics: when not NULL, the casters add the number of samples they take (those of chunks
    // and segments skipped as empty excluded) here
    unsigned long long* tally;
};

CastConst makeCastConst(const VolumeView* vol, const VolumeGeom& vg, const RayParams& rp,
                        const LaunchShape& shape);

// One projection of a batch: volume, per-view geometry, output and its first tile in the batch.
struct BatchItem {
    VolumeView vol;
    const RayGeom* geom;
    float* out;
    RayParams rp;
    CastConst kc;
    unsigned long long tile_begin;
};

// Tiles per run of the batch autotuner's sample (every stride-th run of consecutive tiles).
constexpr uint32_t kSampleRun = 64;

// Work queue of the persistent batch kernel: one tile range per XCD.
struct BatchQueue {
    unsigned long long next[8];
    unsigned long long end[8];
};

// Resets the queue to [0, total) split over the XCDs (a kernel, so it can sit in a graph).
hipError_t launchQueueInit(BatchQueue* d_q, unsigned long long total, hipStream_t s);
// Persistent kernel over `n_items` projections; all items share storage, layout and mode.
// lds_texels > 0 with interleaved trilinear items selects the LDS-staged kernel with that box
// (every item must then have kc.chunk > 0); 0 selects the unstaged kernel.
hipError_t launchBatch(const BatchItem* d_items, uint32_t n_items, Storage st, Layout ly,
                       int interpolate, uint32_t unroll, uint32_t blocks_per_cu, uint32_t lds_texels,
                       uint32_t grid, BatchQueue* d_q, uint32_t stride, hipStream_t s);
// Workgroups of the batch kernel (as selected by launchBatch) that fit on the device at once.
hipError_t batchGrid(Storage st, Layout ly, int interpolate, uint32_t unroll, uint32_t blocks_per_cu,
                     uint32_t lds_texels, uint32_t* grid);

// FP64 setup: one thread per (view, module).
hipError_t launchSetup(const double* d_pmats, uint32_t views, uint32_t modules, uint32_t channels,
                       uint32_t rows, const VolumeGeom& vg, int interpolate, RayGeom* d_geom,
                       hipStream_t s);

// Ray casting of all views in one launch; out is [views][modules][rows][channels]. accumulate != 0
// adds the projection to what d_out holds.
hipError_t launchRaycast(const VolumeView& vol, const VolumeGeom& vg, const RayGeom* d_geom,
                         const RayParams& rp, const LaunchShape& shape, float* d_out,
                         hipStream_t s, int accumulate = 0,
                         unsigned long long* tally = nullptr);

// Reference loop-count of all rays, added into *d_count.
hipError_t launchCount(const VolumeGeom& vg, const RayGeom* d_geom, const RayParams& rp,
                       unsigned long long* d_count, hipStream_t s);

// float (X,Y,Z) -> padded storage in `layout`, replicating each voxel f^3 times (f >= 1).
hipError_t launchPack(const float* d_src, const uint32_t n[3], uint32_t f, Storage st, Layout layout,
                      void* d_dst, hipStream_t s);
// padded storage -> padded storage upsampled by f (f >= 1), same storage and layout.
hipError_t launchUpsample(const VolumeView& src, uint32_t f, void* d_dst, hipStream_t s);
// Builds view.occ's table (occBytes(view.n) bytes at d_occ, which view.occ need not point to yet)
// from the view's voxels.
hipError_t launchOccupancy(const VolumeView& view, uint32_t* d_occ, hipStream_t s);
// Mean of the interior of a padded volume, into *d_sum (double, accumulated).
hipError_t launchSum(const VolumeView& src, double* d_sum, hipStream_t s);
// Fill a one-texel padded volume (1x1x1 interior) with the value *d_sum / count.
hipError_t launchFillMean(const double* d_sum, double count, Storage st, Layout layout, void* d_dst,
                          hipStream_t s);
// Zero-fill of a float array with a plain kernel (graph-capture safe).
hipError_t launchZero(float* d, size_t n, hipStream_t s);

// Gather-ceiling probes (diagnostics). launchProbeGather issues the ray caster's texel loads for
// every ray of the projection (interleaved layout, trilinear, one ray per pixel) without the
// interpolation; FP16 and BF16 share one load pattern. launchProbeStream runs `grid` workgroups
// of coalesced 16-byte loads, `iters` (a multiple of 4) per thread, from a 16 KiB window of d_src;
// d_out receives grid * 256 floats.
hipError_t launchProbeGather(const VolumeView& vol, const VolumeGeom& vg, const RayGeom* d_geom,
                             const RayParams& rp, float* d_out, hipStream_t s);
hipError_t launchProbeStream(const void* d_src, uint32_t grid, uint32_t iters, float* d_out, hipStream_t s);
// LDS read ceiling: `grid` workgroups of conflict-free 16-byte LDS reads (pairs: as two adjacent
// 8-byte elements, 15 of every 16 reads), `iters` (a multiple of 16) per thread, from a 16 KiB LDS
// window; d_out receives grid * 256 floats.
hipError_t launchProbeLds(uint32_t grid, uint32_t iters, bool pairs, float* d_out, hipStream_t s);

// Mean of k contiguous projections of n values (intensity != 0: -ln of the mean of exp(-value)).
hipError_t launchAverage(const float* d_in, uint32_t k, size_t n, int intensity, float* d_out, hipStream_t s);

// Spectral extinction of views x n pixels from n_mat (<= kMaxMaterials) material projections
// stored one after the other ([m][v][pixel]); mu [bins][n_mat], inten [bins][views], resp [bins],
// total [views].
constexpr uint32_t kMaxMaterials = 8;
hipError_t launchSpectral(const float* d_proj, uint32_t n_mat, uint32_t views, size_t n, const float* d_mu,
                          const float* d_inten, const float* d_resp, uint32_t n_bins, const float* d_total,
                          float* d_out, hipStream_t s);
// out[v] = ca[v] a[v] + cb[v] b[v] per view of n values (d_b and d_cb may be NULL).
hipError_t launchCombineViews(const float* d_a, const float* d_ca, const float* d_b, const float* d_cb,
                              uint32_t views, size_t n, float* d_out, hipStream_t s);

bool validShape(const LaunchShape& s);

// Static device scratch of the allocation-free projection (one copy per device, held until the
// process ends): room for kScratchRayGeoms ray geometry records. initScratch(device), with that
// device current, looks up the copy's address once; scratchPtr(device) returns it afterwards
// (NULL before). Every user of a device's scratch shares it: calls that overlap on one device race.
constexpr size_t kScratchRayGeoms = 131072;
constexpr size_t kScratchBytes = kScratchRayGeoms * sizeof(RayGeom); // 20 MiB
hipError_t initScratch(int device);
void* scratchPtr(int device);

// ---- volume processing (processing.hip) -----------------------------------------------------

// Affine map of a sampling plane, FP32 as the host casts it: sampler coordinate
//   slices:          u_k = h[4k] (x - cx) + h[4k+1] (y - cy) + h[4k+3] - c[k] + 1/2
//   plane integrals: u_k = h[4k] x + h[4k+1] y + h[4k+3] + d c[k]   (c = normal / voxel size)
struct alignas(16) PlaneMap {
    float h[12];
    float c[3];
    float pad;
};
static_assert(sizeof(PlaneMap) == 64, "PlaneMap is one 64-byte record");

// Inverse of a plane map's linear part: rows give the template pixel x, y and the distance d of a
// texel position q as m[r] . (q - h2), h2 the map's constant column plus 1/2.
struct alignas(16) PlaneInv {
    float m[3][3];
    float pad[3];
};
static_assert(sizeof(PlaneInv) == 48, "PlaneInv is one 48-byte record");

// Distances per workgroup of the plane-integral kernel.
constexpr int kPlaneDists = 8;

// One launch of plane integrals: n_orient orientations x n_dist distances over a width x height
// template plane cut into tiles_x x tiles_y tiles of 8x8 pixels, visited in groups of 4x4 tiles
// (n_tiles = 16 * groups, counting the unused places of partial groups); each (orientation,
// distance chunk) is split into `splits` ranges of that order.
struct PlaneJob {
    uint32_t n_orient, n_dist;
    uint32_t width, height;
    uint32_t tiles_x, tiles_y, n_tiles;
    uint32_t splits;
};

// n_planes slices of width x height pixels into d_out ([plane][y][x]).
hipError_t launchSlice(const VolumeView& vol, const PlaneMap* d_maps, uint32_t n_planes, uint32_t width,
                       uint32_t height, float* d_out, hipStream_t s);
// ranges = {lo1, hi1, lo2, hi2, lo3, hi3}; d_out[k][j][i] at (p1[i], p2[j], p3[k]).
hipError_t launchResample(const VolumeView& vol, const float ranges[6], const float* d_p1, uint32_t n1,
                          const float* d_p2, uint32_t n2, const float* d_p3, uint32_t n3, float* d_out,
                          hipStream_t s);
// d_out[i] at the point d_xyz[3i .. 3i+2].
hipError_t launchSamplePoints(const VolumeView& vol, const float ranges[6], const float* d_xyz, size_t n,
                              float* d_out, hipStream_t s);
// FP32 brick copies of a padded volume for the plane kernel (128-byte bricks): Texel8, the
// interleaved texels in bricks of 4x2x2; Pair16, 16-byte elements (the texels of x and x + 1) in
// bricks of 2x2x2. Auto takes Texel8 while that copy is at most kTexel8CacheBytes (it then stays
// in the last-level cache), Pair16 beyond.
enum class BrickKind : int { Auto = 0, Texel8 = 1, Pair16 = 2 };
constexpr uint64_t kTexel8CacheBytes = 192ull << 20;
// The kind brickShape resolves `want` to, nb[] bricks per axis and the copy's size. False if the
// copy would not fit one 32-bit-addressed buffer.
bool brickShape(const uint32_t n[3], BrickKind want, BrickKind* kind, uint32_t nb[3], size_t* bytes);
hipError_t launchBrick(const VolumeView& vol, BrickKind kind, float* d_dst, hipStream_t s);
// Plane integrals: d_partial holds n_orient * n_dist * splits floats; d_out[d * out_stride + o]
// receives scale * (the sum over the template plane). d_brick (the volume's brick copy of kind
// brick_kind, or NULL to read the resident texels) selects the reads.
hipError_t launchPlaneIntegrals(const VolumeView& vol, const float* d_brick, BrickKind brick_kind,
                                const PlaneMap* d_maps, const float* d_dist, const PlaneJob& job, float* d_partial,
                                float scale, float* d_out, uint32_t out_stride, hipStream_t s);

// Plane integrals by volume bricks staged in LDS (needs the volume's brick copy): n_orient x n_dist
// integrals into d_out[d * out_stride + o], accumulated in d_acc (n_orient * n_dist 64-bit
// fixed-point totals, 2^fix_shift units per value; every |total| must stay below 2^62) by
// workgroups of one brick and o_chunk orientations. planeBricks() is the number of bricks.
uint32_t planeBricks(const uint32_t n[3]);
hipError_t launchPlaneBricks(const VolumeView& vol, const float* d_brick, BrickKind brick_kind, const PlaneMap* d_maps,
                             const PlaneInv* d_inv, const float* d_dist, uint32_t n_orient, uint32_t n_dist,
                             uint32_t width, uint32_t height, uint32_t o_chunk, int fix_shift,
                             unsigned long long* d_acc, float scale, float* d_out, uint32_t out_stride,
                             hipStream_t s);
// Largest |v| of a brick copy of `bytes` bytes, as float bits, atomically maxed into *d_max (which
// the caller clears); NaN and infinity give bits >= 0x7f800000.
hipError_t launchBrickMaxAbs(const float* d_brick, size_t bytes, uint32_t* d_max, hipStream_t s);

} // namespace ctlk
