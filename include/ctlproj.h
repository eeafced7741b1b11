/*
 * ctlproj.h -- C ABI of the MI355X forward projector (drop-in for CTL::OCL::RayCasterProjector).
 *
 * Every entry point returns a ctlproj_status; no exception crosses the ABI. The message of the
 * last failure on the calling thread is available from ctlproj_last_error().
 *
 * Reference interfaces replaced (paths relative to the CTL source tree):
 *   OpenCLConfig / ClFileLoader / PinnedBuf*      modules/src/ocl/openclconfig.h:38-131,
 *                                                 clfileloader.h:47-77, pinnedmem.h:143-173
 *   RayCasterProjector::Settings                  modules/src/projectors/raycasterprojector.h:27-37
 *   RayCasterProjector::configure                 modules/src/projectors/raycasterprojector.cpp:97-107
 *   RayCasterProjector::project                   modules/src/projectors/raycasterprojector.cpp:139-375
 *   GeometryEncoder::encodeFullGeometry           modules/src/acquisition/geometryencoder.cpp:83-135
 *
 * Data layouts:
 *   volume      float, x fastest: index = (z*Y + y)*X + x        (voxelvolume.tpp:310-332)
 *   P-matrices  double [views][modules][3][4], row-major          (viewgeometry.h:21-102)
 *   projections float [views][modules][rows][channels]           (projectiondata.cpp:134-228)
 */
#ifndef CTLPROJ_H
#define CTLPROJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ctlproj_status {
    CTLPROJ_OK = 0,
    CTLPROJ_ERR_INVALID_ARGUMENT = 1, /* bad pointer, size or setting */
    CTLPROJ_ERR_NO_DEVICE = 2,        /* no HIP device, or an invalid device ordinal */
    CTLPROJ_ERR_DEVICE = 3,           /* a HIP runtime call failed */
    CTLPROJ_ERR_OUT_OF_MEMORY = 4,    /* device or host allocation failed */
    CTLPROJ_ERR_NOT_CONFIGURED = 5,   /* project() before configure() */
    CTLPROJ_ERR_NO_DATA = 6,          /* volume without data (reference returns 0 views) */
    CTLPROJ_ERR_OVER_BUDGET = 7,      /* a request larger than the server's HBM budget */
    CTLPROJ_ERR_INJECTED = 8,         /* failure forced by ctlproj_debug_inject_failure */
    CTLPROJ_ERR_TIMEOUT = 9,          /* a wait ended before the request finished */
    CTLPROJ_ERR_IO = 10               /* a file could not be read or written, or is malformed */
} ctlproj_status;

typedef enum ctlproj_dtype {
    CTLPROJ_F32 = 0,
    CTLPROJ_F16 = 1,
    CTLPROJ_BF16 = 2
} ctlproj_dtype;

/* RayCasterProjector::Settings (raycasterprojector.h:27-37) plus tuning fields that the reference
 * does not have. ctlproj_settings_default() fills the reference defaults: rays per pixel {1,1},
 * ray sampling 0.3, upsampling 1, interpolate on; tuning fields 0 = automatic. */
typedef struct ctlproj_settings {
    uint32_t rays_per_pixel[2];
    float ray_sampling;
    uint32_t volume_upsampling;
    int32_t interpolate;
    uint32_t tile_w;   /* pixels per wavefront tile along channels (0 = auto) */
    uint32_t tile_h;   /* pixels per wavefront tile along rows (0 = auto) */
    uint32_t unroll;   /* ray-step unroll of the sampling loop, 1-4 (0 = auto); with 3 or 4 the
                          LDS-staged kernel takes each step's last sample through L1; the unstaged
                          kernels run 3 as 4 */
    uint32_t blocks_per_cu; /* cap on resident workgroups per CU, 1..8 (0 = auto) */
    int32_t lds_chunk;      /* LDS staging of trilinear projections with one ray per pixel:
                               samples per chunk, 16..1024 (0 = auto: on for FP32 storage, off
                               for FP16/BF16; -1 = off: gather from HBM through the L1) */
    uint32_t lds_texels;    /* LDS box capacity in 8-byte texels, 1024..20352 (0 = auto) */
} ctlproj_settings;

typedef struct ctlproj_context ctlproj_context;
typedef struct ctlproj_volume ctlproj_volume;
typedef struct ctlproj_projector ctlproj_projector;

/* Per-view progress, the analogue of ProjectorNotifier::projectionFinished(int)
 * (abstractprojector.h:57-63). Called on the calling thread once a view's data is in `out`. */
typedef void (*ctlproj_progress_fn)(int view, void* user);

/* ---- environment (replaces OpenCLConfig, openclconfig.cpp:15-171) ----------------------- */

void ctlproj_settings_default(ctlproj_settings* s);

/* RayCasterProjector::Settings::optimizedFor (raycasterprojector.cpp:482-521): rays per pixel from
 * the pixel aspect ratio (presets 1:2, 1:3, 1:4, 2:3, 3:4), a volume upsampling factor for voxels
 * larger than a sub-pixel, and more rays for voxels smaller than a sub-pixel; other fields take
 * their defaults. voxel_mm: the volume's voxel size (its smallest component counts); pixel sizes
 * of the detector in mm. A zero pixel size returns CTLPROJ_ERR_INVALID_ARGUMENT (the reference
 * throws). Needs no device. */
ctlproj_status ctlproj_settings_optimized_for(const float voxel_mm[3], double pixel_width_mm,
                                              double pixel_height_mm, ctlproj_settings* out);
const char* ctlproj_last_error(void);
ctlproj_status ctlproj_device_count(int* count);
ctlproj_status ctlproj_context_create(int device, ctlproj_context** out);
ctlproj_status ctlproj_context_destroy(ctlproj_context* ctx);

/* ---- resident volumes (replaces the per-call Image3D upload, raycasterprojector.cpp:236-251) */

/* Uploads `host` (float, X*Y*Z) and keeps it resident in HBM in `storage` precision.
 * voxel_mm: voxel size, offset_mm: volume offset (VoxelVolume::offset). Volumes of 4 GiB or more
 * (padded 8-byte texels) are projected but refused by batches, the server and volume processing.
 * The volume also keeps an occupancy table of its 4x4x4-voxel bricks (a summed-area table, about
 * 1/64 of the voxel count in 4-byte entries), with which the ray casters skip chunks of samples
 * that read only zero voxels; the environment variable CTLPROJ_EMPTY_SKIP=0 at creation leaves it
 * out. */
ctlproj_status ctlproj_volume_create(ctlproj_context* ctx, const float* host, const uint32_t dims[3],
                                     const float voxel_mm[3], const float offset_mm[3],
                                     ctlproj_dtype storage, ctlproj_volume** out);
/* Same, from a float array already in device memory on the context's device. */
ctlproj_status ctlproj_volume_create_from_device(ctlproj_context* ctx, const float* device_ptr,
                                                 const uint32_t dims[3], const float voxel_mm[3],
                                                 const float offset_mm[3], ctlproj_dtype storage,
                                                 ctlproj_volume** out);
ctlproj_status ctlproj_volume_destroy(ctlproj_volume* vol);
/* Device bytes held by the volume. */
ctlproj_status ctlproj_volume_bytes(const ctlproj_volume* vol, size_t* bytes);

/* ---- projector (RayCasterProjector, raycasterprojector.h:20-58) -------------------------- */

ctlproj_status ctlproj_projector_create(ctlproj_context* ctx, const ctlproj_settings* settings,
                                        ctlproj_projector** out);
ctlproj_status ctlproj_projector_destroy(ctlproj_projector* p);
ctlproj_status ctlproj_projector_set_settings(ctlproj_projector* p, const ctlproj_settings* s);
ctlproj_status ctlproj_projector_get_settings(const ctlproj_projector* p, ctlproj_settings* s);

/* configure (raycasterprojector.cpp:97-107): `pmats` holds views*modules FP64 3x4 matrices;
 * channels and rows are the pixels per module. The matrices are copied. */
ctlproj_status ctlproj_projector_configure(ctlproj_projector* p, const double* pmats, uint32_t views,
                                           uint32_t modules, uint32_t channels, uint32_t rows);

/* project (raycasterprojector.cpp:139-375) into host memory: out holds
 * views*modules*rows*channels floats. Blocks until done. `progress` may be NULL. */
ctlproj_status ctlproj_project(ctlproj_projector* p, const ctlproj_volume* vol, float* out,
                               size_t out_count, ctlproj_progress_fn progress, void* user);

/* Reference-style one-shot project(VolumeData): uploads the host volume (FP32 storage),
 * projects into host memory and frees the device copy. */
ctlproj_status ctlproj_project_host_volume(ctlproj_projector* p, const float* host_volume,
                                           const uint32_t dims[3], const float voxel_mm[3],
                                           const float offset_mm[3], float* out, size_t out_count);

/* Asynchronous projection into device memory on the projector's stream (no host copy).
 * Call ctlproj_projector_synchronize before reading `device_out`. */
ctlproj_status ctlproj_project_device(ctlproj_projector* p, const ctlproj_volume* vol,
                                      float* device_out, size_t out_count);
/* Allocation-free form of ctlproj_project_device. No counterpart in the reference: it is for
 * callers that cannot allocate device memory during a projection (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc/hipFree: the
 * per-(view, module) ray geometry goes into a static device scratch of 20 MiB per device, held
 * until the process ends, which limits a projection to CTLPROJ_NOALLOC_MAX_VIEW_MODULES
 * (view, module) pairs (larger ones return CTLPROJ_ERR_INVALID_ARGUMENT, as does a volume
 * upsampling factor other than 1, which would need a temporary volume). Same kernels and results
 * as ctlproj_project_device; asynchronous on the projector's stream. Every call on one device
 * shares that device's scratch: calls that overlap on one device (two projectors on two streams
 * or threads) race and must be serialised by the caller. */
#define CTLPROJ_NOALLOC_MAX_VIEW_MODULES 131072u
ctlproj_status ctlproj_project_device_noalloc(ctlproj_projector* p, const ctlproj_volume* vol,
                                              float* device_out, size_t out_count);

/* AbstractProjector::projectComposite (abstractprojector.h:99-112): the sum of the projections of
 * n resident volumes (any voxel sizes and offsets), accumulated on the device in volume order on
 * the projector's stream. The host form copies the sum to `out` ([V][M][R][C]). */
ctlproj_status ctlproj_project_composite(ctlproj_projector* p, const ctlproj_volume* const* vols,
                                         uint32_t n, float* out, size_t out_count);
ctlproj_status ctlproj_project_composite_device(ctlproj_projector* p, const ctlproj_volume* const* vols,
                                                uint32_t n, float* device_out, size_t out_count);

/* Mean of k projections of n values stored one after the other in device memory. intensity_domain:
 * -ln(mean(exp(-value))) (averaging in the intensity domain with initial intensity 1, as
 * ArealFocalSpotExtension does, arealfocalspotextension.cpp:82-166); otherwise the plain mean
 * (its low-extinction approximation). Synchronous on the context's stream. */
ctlproj_status ctlproj_average_projections(ctlproj_context* ctx, const float* device_in, uint32_t k,
                                           size_t n, int intensity_domain, float* device_out);

/* SpectralEffectsExtension::projectLinear (spectraleffectsextension.cpp:232-300, 330-352): the
 * extinction seen through an X-ray spectrum, from the projections of n_materials (1..8) material
 * densities stored one after the other in device memory (n_materials blocks of views *
 * pixels_per_view floats). For each pixel of view v:
 *   out = ln( I0[v] / sum_b response[b] * intensity[b][v] * exp(-sum_m mu[b][m] * proj_m) )
 * mu [n_bins][n_materials]: mass attenuation coefficients of each bin in the units of the density
 * projections (the reference multiplies cm^2/g by 0.1 for g/cm^3 densities in mm);
 * intensity [n_bins][views]: the bins' initial intensities per view; response [n_bins]: detector
 * response per bin or NULL (no response model); total_intensity [views]: I0. Bins whose
 * intensities sum to zero are skipped, as in the reference. Synchronous on the context's stream. */
ctlproj_status ctlproj_spectral_extinction(ctlproj_context* ctx, const float* device_material_projs,
                                           uint32_t n_materials, uint32_t views, size_t pixels_per_view,
                                           const float* mu, const double* intensity, const float* response,
                                           uint32_t n_bins, const double* total_intensity, float* device_out);
/* Per-view linear combination of two projections in device memory:
 * out[v] = coeff_a[v] * a[v] + coeff_b[v] * b[v] (device_b NULL: the second term is absent).
 * DynamicProjectorExtension (dynamicprojectorextension.cpp:39-66) on volumes that are linear in
 * time (LinearDynamicVolume, lineardynamicvolume.cpp:194-198: slope * t + offset) is
 * t_v * P_v(slope) + P_v(offset). Synchronous on the context's stream. */
ctlproj_status ctlproj_combine_views(ctlproj_context* ctx, const float* device_a, const float* coeff_a,
                                     const float* device_b, const float* coeff_b, uint32_t views,
                                     size_t pixels_per_view, float* device_out);

/* Host-level forms of the two above for callers without device buffers, on the configured
 * projector: the material densities are projected into internal device memory, combined there and
 * the extinction ([V][M][R][C]) copied to `out`.
 * ctlproj_project_spectral: SpectralEffectsExtension::projectLinear with the spectrum and
 * attenuation given as for ctlproj_spectral_extinction (n_materials densities).
 * ctlproj_project_linear_dynamic: DynamicProjectorExtension::project of a LinearDynamicVolume
 * (value = slope * t + offset at view v's time stamp t = time_stamps[v], taken as float). */
ctlproj_status ctlproj_project_spectral(ctlproj_projector* p, const ctlproj_volume* const* densities,
                                        uint32_t n_materials, const float* mu, const double* intensity,
                                        const float* response, uint32_t n_bins, const double* total_intensity,
                                        float* out, size_t out_count);
ctlproj_status ctlproj_project_linear_dynamic(ctlproj_projector* p, const ctlproj_volume* slope,
                                              const ctlproj_volume* offset, const double* time_stamps, float* out,
                                              size_t out_count);

ctlproj_status ctlproj_projector_synchronize(ctlproj_projector* p);

/* Duration of the last ray-casting kernel on the projector's stream (HIP events), in ms. */
ctlproj_status ctlproj_projector_last_kernel_ms(ctlproj_projector* p, float* ms);

/* Number of sampling-loop iterations the reference performs for this volume and geometry
 * (sum over rays of floor(lambda_max) + 2 - floor(lambda_min)), counted on the device. */
ctlproj_status ctlproj_projector_count_samples(ctlproj_projector* p, const ctlproj_volume* vol,
                                               uint64_t* n_samples);

/* Diagnostics (no reference counterpart): projects all views into device_out, as
 * ctlproj_project_device does, and returns the number of samples the kernels actually took. The
 * casters skip chunks of samples that read only zero voxels (exactly 0, so every output is the
 * same as with all samples taken); n_taken excludes them. */
ctlproj_status ctlproj_projector_count_taken(ctlproj_projector* p, const ctlproj_volume* vol, float* device_out,
                                            size_t out_count, uint64_t* n_taken);

/* Measures the candidate launch shapes (wavefront pixel tile, ray-step unroll, LDS chunk and box;
 * for narrow modules also with the wavefronts stacked) on up to 24 of the configured views and
 * remembers the fastest for this workload class (storage, layout, mode, detector size); later
 * projections with automatic tuning fields use it.
 * No counterpart in the reference (it has no launch-shape choice). */
ctlproj_status ctlproj_projector_autotune(ctlproj_projector* p, const ctlproj_volume* vol,
                                         float* best_ms);
/* The autotuner's table as JSON text (needed = length + 1). */
ctlproj_status ctlproj_autotune_table(char* buf, size_t cap, size_t* needed);

/* ---- batched projection (no counterpart in the reference) ---------------------------------
 * Many projections of their own volume, geometry and detector size run in one persistent kernel
 * captured in a hipGraph: ctlproj_batch_add each entry (device output of views*modules*rows*
 * channels floats, owned by the caller), ctlproj_batch_finalize once (uploads geometry, optional
 * autotuning on a sample of the batch's tiles, graph capture), then ctlproj_batch_launch as often
 * as needed. One batch holds one storage type and one interpolation mode. The callers that
 * multiply project() calls in the reference (projectComposite, abstractprojector.h:99-112;
 * SpectralEffectsExtension, spectraleffectsextension.cpp:385-395; ArealFocalSpotExtension,
 * arealfocalspotextension.cpp:104-150) map onto it. */
typedef struct ctlproj_batch ctlproj_batch;
ctlproj_status ctlproj_batch_create(ctlproj_context* ctx, ctlproj_batch** out);
ctlproj_status ctlproj_batch_destroy(ctlproj_batch* b);
ctlproj_status ctlproj_batch_add(ctlproj_batch* b, const ctlproj_volume* vol, const double* pmats,
                                 uint32_t views, uint32_t modules, uint32_t channels, uint32_t rows,
                                 const ctlproj_settings* settings, float* device_out,
                                 size_t out_count, uint32_t* index);
ctlproj_status ctlproj_batch_finalize(ctlproj_batch* b, int autotune);
ctlproj_status ctlproj_batch_launch(ctlproj_batch* b);
ctlproj_status ctlproj_batch_synchronize(ctlproj_batch* b);
ctlproj_status ctlproj_batch_last_ms(ctlproj_batch* b, float* ms);
ctlproj_status ctlproj_batch_info(const ctlproj_batch* b, uint32_t* tile_w, uint32_t* tile_h,
                                  uint32_t* unroll, uint32_t* grid, uint64_t* tiles);
ctlproj_status ctlproj_batch_count_samples(ctlproj_batch* b, uint64_t* n_samples);
/* The finalized batch's launch shape as the tuning fields of *settings (tile, unroll, workgroups
 * per CU, LDS chunk (-1: unstaged) and box); the other fields are left as they are. A single
 * projection with these fields runs the same per-pixel code as the batch and gives the same
 * bits. */
ctlproj_status ctlproj_batch_shape(const ctlproj_batch* b, ctlproj_settings* settings);

/* ---- serving (no counterpart in the reference) ---------------------------------------------
 * A request loop over resident volumes: requests are admitted when their device footprint
 * (output + geometry) fits into the server's HBM arena, split into chunks of views, and
 * batched across volumes and detector sizes into persistent-kernel launches on three streams of
 * the server's own, so requests that arrive while others run join the next launch; single-view
 * requests run beside a scan launch. The streams order themselves against the null stream.
 * Each result is copied to the caller's host array (pinned memory makes the copy asynchronous);
 * latency is measured from submit to the result being in host memory. Volumes must outlive the
 * requests that use them. Calls may come from any thread. */
typedef struct ctlproj_server ctlproj_server;

typedef struct ctlproj_server_stats {
    uint64_t submitted, completed, failed, rejected;
    uint64_t views;      /* views of completed requests */
    uint64_t launches;   /* persistent-kernel launches */
    double p50_ms, p99_ms, mean_ms, max_ms;          /* all completed requests */
    double p50_single_ms, p99_single_ms;             /* one-view requests */
    double p50_multi_ms, p99_multi_ms;               /* requests of more than one view */
    double span_s;       /* first submit to last completion */
    /* launch slots: how many slot streams carry a CU mask of the whole device (each then has a
     * hardware queue of its own; 0 if the runtime refused the mask and plain streams were made),
     * the most scan launches (more than one view) in flight at once (at most slots - 1), and the
     * single-view launches that completed while a scan launch issued before them was still running */
    uint32_t slots, cu_masked_slots, max_scan_launches_in_flight, reserved;
    uint64_t drr_launches_beside_scan;
} ctlproj_server_stats;

/* hbm_budget: bytes of the arena (outputs + geometry of admitted requests).
 * max_launch_pixels: pixels (views x modules x rows x channels) per launch, 0 = 16 Mi (a few ms of
 * work: single-view requests wait for at most that, and run in launches of their own). */
ctlproj_status ctlproj_server_create(ctlproj_context* ctx, size_t hbm_budget,
                                     uint64_t max_launch_pixels, ctlproj_server** out);
/* Waits for every submitted request, then stops the server. */
ctlproj_status ctlproj_server_destroy(ctlproj_server* s);
/* Queues a projection; returns at once. A request larger than the whole arena is rejected with
 * CTLPROJ_ERR_OVER_BUDGET. Storage type and interpolation mode may differ between requests. */
ctlproj_status ctlproj_server_submit(ctlproj_server* s, const ctlproj_volume* vol,
                                     const double* pmats, uint32_t views, uint32_t modules,
                                     uint32_t channels, uint32_t rows,
                                     const ctlproj_settings* settings, float* host_out,
                                     size_t out_count, uint64_t* request_id);
/* Waits for one request (timeout_s < 0: no limit); returns its status (CTLPROJ_ERR_TIMEOUT if
 * it is still running). latency_ms may be NULL. */
ctlproj_status ctlproj_server_wait(ctlproj_server* s, uint64_t request_id, double timeout_s,
                                   double* latency_ms);
/* Waits for every request submitted so far. */
ctlproj_status ctlproj_server_drain(ctlproj_server* s);
ctlproj_status ctlproj_server_get_stats(ctlproj_server* s, ctlproj_server_stats* out);
ctlproj_status ctlproj_server_reset_stats(ctlproj_server* s);

/* ---- geometry front end (GeometryEncoder, geometryencoder.cpp:83-135) -------------------- */

typedef struct ctlproj_detector {
    uint32_t pixels[2];       /* channels, rows per module */
    double pixel_mm[2];       /* pixel width, height */
    uint32_t modules;         /* 1 and cylindrical = 0: FlatPanelDetector */
    int32_t cylindrical;      /* CylindricalDetector (cylindricaldetector.cpp:29-41) */
    double angulation;        /* rad between adjacent modules */
    double module_spacing;    /* mm */
    double skew;
} ctlproj_detector;

typedef enum ctlproj_gantry_type { CTLPROJ_CARM = 0, CTLPROJ_TUBULAR = 1 } ctlproj_gantry_type;

typedef struct ctlproj_gantry {
    int32_t type;             /* ctlproj_gantry_type */
    double carm_span;         /* CarmGantry(cArmSpan), carmgantry.cpp:12-16 */
    double sdd, sid;          /* TubularGantry(SDD, SID, rotation, pitch, tilt), */
    double rotation, pitch, tilt; /*   tubulargantry.cpp:20-33 */
} ctlproj_gantry;

typedef enum ctlproj_trajectory_type {
    CTLPROJ_SHORT_SCAN = 0,       /* p = {source_to_iso, start_angle, angle_span (<0: 180deg+fan)} */
    CTLPROJ_WOBBLE = 1,           /* p = {angle_span, source_to_iso, start_angle, wobble_angle, freq} */
    CTLPROJ_CIRCLE_PLUS_LINE = 2, /* p = {angle_span, source_to_iso, line_length, fraction, start} */
    CTLPROJ_HELICAL = 3,          /* p = {angle_increment, pitch_increment, start_pitch, start_angle} */
    CTLPROJ_AXIAL = 4             /* p = {start_angle} */
} ctlproj_trajectory_type;

typedef struct ctlproj_trajectory {
    int32_t type;
    double p[6];
} ctlproj_trajectory;

/* encodeFullGeometry for a system and a trajectory protocol (trajectories.cpp). out holds
 * views*modules*12 doubles. */
ctlproj_status ctlproj_encode_trajectory(const ctlproj_gantry* gantry, const ctlproj_detector* det,
                                         const ctlproj_trajectory* traj, uint32_t views,
                                         double* out, size_t out_count);

/* C-arm views at explicit source locations (prepare::CarmGantryParam::setLocation,
 * preparesteps.cpp:71-84): positions [views][3] mm, rotations [views][3][3] row-major. */
ctlproj_status ctlproj_encode_carm_poses(double carm_span, const ctlproj_detector* det,
                                         const double* positions, const double* rotations,
                                         uint32_t views, double* out, size_t out_count);

/* The same with the source displaced by source_shift (mm, in the source's own frame: the
 * gantry's source displacement, abstractgantry.h:159-163, as prepare::GantryDisplacementParam
 * sets it; ArealFocalSpotExtension shifts it over the focal spot). */
ctlproj_status ctlproj_encode_trajectory_shifted(const ctlproj_gantry* gantry, const ctlproj_detector* det,
                                                 const ctlproj_trajectory* traj, uint32_t views,
                                                 const double source_shift[3], double* out,
                                                 size_t out_count);
ctlproj_status ctlproj_encode_carm_poses_shifted(double carm_span, const ctlproj_detector* det,
                                                 const double* positions, const double* rotations,
                                                 uint32_t views, const double source_shift[3],
                                                 double* out, size_t out_count);

/* GeometryDecoder::decodeFullGeometry (geometrydecoder.cpp): the physical geometry behind
 * P-matrices [V][M][3][4] of modules with channels x rows pixels. reference: CTLPROJ_DECODE_PIXEL_SIZE
 * (a, b = pixel width, height in mm) or the PhysicalDimension whose length in mm is `a`
 * (pixel width, pixel height, source-detector distance). Outputs per view: source position (3),
 * source rotation (9, row-major; computeSourceRotation's columns r1, r2, unit source-to-detector
 * direction), pixel size (2), skew (1); per view and module: centre position (3) and rotation
 * (9, row-major, world -> module). */
typedef enum ctlproj_decode_reference {
    CTLPROJ_DECODE_PIXEL_SIZE = 0,
    CTLPROJ_DECODE_PIXEL_WIDTH = 1,
    CTLPROJ_DECODE_PIXEL_HEIGHT = 2,
    CTLPROJ_DECODE_SOURCE_DETECTOR_DISTANCE = 3
} ctlproj_decode_reference;
ctlproj_status ctlproj_decode_geometry(const double* pmats, uint32_t views, uint32_t modules,
                                       uint32_t channels, uint32_t rows, int reference, double a,
                                       double b, double* source_pos, double* source_rot,
                                       double* pixel_size, double* skew, double* module_pos,
                                       double* module_rot);

/* GeometryEncoder for generic components: P-matrices [V][M][3][4] (normalised) from source
 * positions (V x 3), module centres (V x M x 3), module rotations (V x M x 9, world -> module),
 * pixel sizes (V x 2) and skew coefficients (V, nullable = 0). The inverse of
 * ctlproj_decode_geometry. */
ctlproj_status ctlproj_encode_generic(const double* source_pos, const double* module_pos,
                                      const double* module_rot, const double* pixel_size,
                                      const double* skew, uint32_t views, uint32_t modules,
                                      uint32_t channels, uint32_t rows, double* out, size_t out_count);

/* ---- synthetic phantoms ------------------------------------------------------------------ */

/* Seeded random-ellipsoid phantom (host memory, x fastest), see the generator description in
 * ctl_amd/csrc/phantom.cpp. */
ctlproj_status ctlproj_phantom_ellipsoids(const uint32_t dims[3], const float voxel_mm[3],
                                          uint32_t n_ellipsoids, uint64_t seed,
                                          float semi_axis_min_mm, float semi_axis_max_mm,
                                          float mu_min, float mu_max, float background_mu,
                                          float* out);

/* ---- DEN and NRRD files (io/den/denfileio.h, io/nrrd/nrrdfileio.h) ------------------------
 * DEN: 6-byte header of three uint16 rows, cols, count (den_header.tpp), then count matrices of
 * rows x cols elements, cols fastest; the element size (1, 2, 4 or 8 bytes: uchar, ushort, float,
 * double) follows from the file size. Volumes: cols = X, rows = Y, count = Z; projections:
 * cols = channels, rows = detector rows, count = modules x views; geometry: 3 x 4 per (view,
 * module). NRRD: "NRRD000<v>" header with type/dimension/sizes/encoding raw/endian little,
 * optional spacings and space origin, key:=value pairs and comments, an empty line, raw data
 * (nrrdfileio.tpp). All data little-endian; sizes of at most 4 dimensions, first fastest. */
typedef enum ctlproj_elem {
    CTLPROJ_ELEM_U8 = 1,
    CTLPROJ_ELEM_U16 = 2,
    CTLPROJ_ELEM_F32 = 4,
    CTLPROJ_ELEM_F64 = 8
} ctlproj_elem;

typedef struct ctlproj_nrrd_header {
    uint32_t version;       /* the digit of "NRRD000<v>" */
    uint32_t ndim;          /* "dimension:" */
    uint32_t nsizes;        /* number of "sizes:" entries (== ndim) */
    int32_t elem;           /* ctlproj_elem of "type:" */
    uint32_t elem_bytes;
    uint64_t sizes[4];
    uint64_t count;         /* product of the sizes */
    uint64_t header_bytes;  /* offset of the raw data */
    int32_t has_spacing;
    double spacing[3];      /* "spacings:" (voxel size, mm) */
    int32_t has_origin;
    double origin[3];       /* "space origin:" (volume offset, mm) */
    int32_t data_complete;  /* the file holds count x elem_bytes bytes of data */
} ctlproj_nrrd_header;

ctlproj_status ctlproj_den_info(const char* path, uint32_t* rows, uint32_t* cols, uint32_t* count,
                                uint32_t* elem_bytes);
/* Reads the data (out_bytes must equal rows*cols*count*elem_bytes). */
ctlproj_status ctlproj_den_read(const char* path, void* out, size_t out_bytes);
/* Sizes 1..65535 each (den_header.h isOutOfBounds). */
ctlproj_status ctlproj_den_write(const char* path, const void* data, uint32_t elem_bytes,
                                 uint32_t rows, uint32_t cols, uint32_t count);
/* Header fields; keyvals (nullable) receives the key:=value lines, NUL-terminated. */
ctlproj_status ctlproj_nrrd_read_header(const char* path, ctlproj_nrrd_header* info, char* keyvals,
                                 size_t keyvals_cap);
/* Reads the raw data (out_bytes must equal count x elem_bytes). */
ctlproj_status ctlproj_nrrd_read(const char* path, void* out, size_t out_bytes);
/* labels: the "labels:" value (nullable, e.g. "\"num vox x\" \"num vox y\" \"num vox z\"");
 * spacing and origin: 3 values each or NULL; keyvals: "key:=value" lines or NULL. */
ctlproj_status ctlproj_nrrd_write(const char* path, const void* data, int elem, uint32_t ndim,
                                  const uint64_t* sizes, const char* labels, const double* spacing,
                                  const double* origin, const char* keyvals);

/* ---- volume processing on resident volumes ----------------------------------------------
 * VolumeSlicer (processing/volumeslicer.h:15-60), VolumeResampler (processing/volumeresampler.h:
 * 13-55) and RadonTransform3D (processing/radontransform3d.h:16-90) upload a volume into an
 * OpenCL image and read it with the linear sampler (CLK_FILTER_LINEAR, CLK_ADDRESS_CLAMP,
 * border colour 0). These entry points read a resident volume of any storage type the same way,
 * with FP32 interpolation. Plane geometry is computed in FP64 and cast to FP32 as the reference's
 * host code does. Every call is synchronous on the volume's context stream; device arrays
 * belong to the volume's device. */

/* VolumeSlicer::sliceDim (volumeslicer.cpp:224-229): default slice side ceil(1.4142 * max(dims))
 * pixels; the default resolution is the smallest voxel size. */
ctlproj_status ctlproj_slicer_default_side(const uint32_t dims[3], uint32_t* side);
/* VolumeSlicer::slice (volumeslicer.cpp:107-178) for n_planes planes at once: plane k is
 * {p : normals[3k..3k+2] . p = distances[k]} (unit normal, mm), sampled on width x height pixels of
 * `resolution` mm centred on distances[k] * normal. out / device_out: n_planes*height*width floats,
 * [plane][y][x]. Non-unit normals and resolution <= 0 return CTLPROJ_ERR_INVALID_ARGUMENT. */
ctlproj_status ctlproj_slice(const ctlproj_volume* vol, const double* normals, const double* distances,
                             uint32_t n_planes, uint32_t width, uint32_t height, float resolution, float* out,
                             size_t out_count);
ctlproj_status ctlproj_slice_device(const ctlproj_volume* vol, const double* normals, const double* distances,
                                    uint32_t n_planes, uint32_t width, uint32_t height, float resolution,
                                    float* device_out, size_t out_count);

/* VolumeResampler's default sampling ranges (volumeresampler.cpp:74-83): per axis
 * {offset - voxel*(N-1)/2, offset + voxel*(N-1)/2}, i.e. the first and last voxel centres.
 * ranges = {lo1, hi1, lo2, hi2, lo3, hi3}. */
ctlproj_status ctlproj_resampler_default_ranges(const ctlproj_volume* vol, float ranges[6]);
/* VolumeResampler::resample (volumeresampler.cpp:117-171): the ranges map onto the voxel centres
 * 0 .. N-1 of each axis; out holds n1*n2*n3 floats, out[i + j*n1 + k*n1*n2] at (p1[i], p2[j], p3[k]). */
ctlproj_status ctlproj_resample(const ctlproj_volume* vol, const float ranges[6], const float* p1, uint32_t n1,
                                const float* p2, uint32_t n2, const float* p3, uint32_t n3, float* out,
                                size_t out_count);
/* VolumeResampler::sample (volumeresampler.cpp:173-213; the cl::Buffer overload :215-255 is the
 * device form): out[i] at the point xyz[3i .. 3i+2]. */
ctlproj_status ctlproj_sample(const ctlproj_volume* vol, const float ranges[6], const float* xyz, size_t n,
                              float* out);
ctlproj_status ctlproj_sample_device(const ctlproj_volume* vol, const float ranges[6], const float* device_xyz,
                                     size_t n, float* device_out);

/* RadonTransform3D's default template plane (radontransform3d.cpp:18-24, 278-301): side =
 * ceil(1.4142 * max(dims)) rounded up to a multiple of 16 pixels, resolution = smallest voxel. */
ctlproj_status ctlproj_radon3d_default_slice(const uint32_t dims[3], const float voxel_mm[3], uint32_t* side,
                                             float* resolution);
/* RadonTransform3D::setSliceResolution (radontransform3d.cpp:239-251): the side that keeps the
 * template plane's extent at a new resolution (a multiple of 16). */
ctlproj_status ctlproj_radon3d_side_for_resolution(uint32_t side, float resolution, float new_resolution,
                                                   uint32_t* new_side);
/* Plane integrals (RadonTransform3D::planeIntegral, radontransform3d.cpp:188-216, and the device
 * loop of sampleTransform, :55-156): out[d*n_orient + o] = integral of the volume over the plane
 * {p : normals[o] . p = distances[d]}, sampled on a side x side template plane of `resolution` mm
 * (sum of the samples times resolution^2). */
ctlproj_status ctlproj_radon3d_plane_integrals(const ctlproj_volume* vol, const double* normals,
                                               uint32_t n_orient, const float* distances, uint32_t n_dist,
                                               uint32_t side, float resolution, float* out, size_t out_count);
ctlproj_status ctlproj_radon3d_plane_integrals_device(const ctlproj_volume* vol, const double* normals,
                                                      uint32_t n_orient, const float* distances, uint32_t n_dist,
                                                      uint32_t side, float resolution, float* device_out,
                                                      size_t out_count);
/* RadonTransform3D::sampleTransform (radontransform3d.cpp:55-156): the grid of azimuth x polar x
 * distance samples, normal (sin(polar) cos(azimuth), sin(polar) sin(azimuth), cos(polar));
 * out[a + n_azi*(p + n_pol*d)] (the returned VoxelVolume's x-fastest order). */
ctlproj_status ctlproj_radon3d_sample_transform(const ctlproj_volume* vol, const float* azimuth, uint32_t n_azi,
                                                const float* polar, uint32_t n_pol, const float* distances,
                                                uint32_t n_dist, uint32_t side, float resolution, float* out,
                                                size_t out_count);

/* ---- gather-ceiling probes (diagnostics) -------------------------------------------------
 * No counterpart in the reference: they measure the ceiling against which the ray caster's gather
 * rate is judged, and compute no projection.
 *   CTLPROJ_PROBE_GATHER  the configured projection's texel loads (same rays, sample ranges,
 *                         wave-aligned loop and 16-byte loads) without the interpolation; needs
 *                         trilinear, rays_per_pixel 1x1, upsampling 1 and an interleaved volume;
 *                         device_out as for ctlproj_project_device (its values are meaningless).
 *                         *bytes = samples x 8 taps x element size.
 *   CTLPROJ_PROBE_STREAM  coalesced 16-byte loads from a 16 KiB window of the volume that stays in
 *                         the vector L1; device_out needs 2048 floats per CU. *bytes = bytes loaded.
 *   CTLPROJ_PROBE_LDS     conflict-free 16-byte LDS reads (ds_read_b128, a wavefront 1 KiB
 *                         contiguous) from a 16 KiB LDS window per workgroup; the volume is not
 *                         read; device_out needs 2048 floats per CU. *bytes = bytes read from LDS.
 *   CTLPROJ_PROBE_LDS_PAIRS  the same 16 bytes per lane read as two adjacent 8-byte elements
 *                         (ds_read2_b64), the instruction the staged ray caster reads texel pairs with.
 * *ms is the probe kernel's duration (HIP events on the projector's stream). */
typedef enum ctlproj_probe_mode {
    CTLPROJ_PROBE_GATHER = 0,
    CTLPROJ_PROBE_STREAM = 1,
    CTLPROJ_PROBE_LDS = 2,
    CTLPROJ_PROBE_LDS_PAIRS = 3
} ctlproj_probe_mode;
ctlproj_status ctlproj_diag_gather_probe(ctlproj_projector* p, const ctlproj_volume* vol, int mode,
                                         float* device_out, size_t out_count, float* ms, double* bytes);

/* ---- failure injection (tests of the error paths) --------------------------------------- */

/* The next `count` device operations of the given context fail with CTLPROJ_ERR_INJECTED. */
ctlproj_status ctlproj_debug_inject_failure(ctlproj_context* ctx, int count);

#ifdef __cplusplus
}
#endif

#endif /* CTLPROJ_H */
