/*
 * sf_seepage.h -- C-ABI of the MI355X (gfx950) coupled SPH-DEM seepage step.
 *
 * One handle = one scene on one GPU: particle state (SoA, sorted by grid cell
 * every step), boundary particles, uniform-grid searcher, solver passes and the
 * integrator.  Every entry point returns 0 on success and a negative SF_E* code
 * on failure; sf_last_error() then describes the failure (thread-local).
 * Nothing here throws, allocates torch objects or needs a Python interpreter.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   sf_params            <- CudaSeepageflowParams     seepage_flows_cuda/include/kiri_pbs_cuda/data/cuda_seepageflow_params.h:25-69
 *   sf_boundary_params   <- CudaBoundaryParams        .../data/cuda_boundary_params.h:19-27
 *   sf_solver_options    <- CudaWCSphSFSolver ctor    .../solver/seepageflow/cuda_wcsph_sf_solver.cuh:22-32
 *   sf_create            <- CudaSFParticles ctor + CudaGNSearcher ctor + CudaWCSphSFSolver ctor + CudaSFSystem ctor
 *                           (cuda_sf_particles.cuh:201-235, cuda_neighbor_searcher.cu:19-34, cuda_sf_system.cu:17-64)
 *   sf_upload_particles  <- CudaSFParticles(maxNum, pos, col, label, mass, radius, cda0asat, amcamcp) cuda_sf_particles.cuh:201-235
 *   sf_upload_boundary   <- CudaBoundaryParticles(pos, label) cuda_boundary_particles.cuh:21-28
 *                           + CudaBaseSystem ctor boundary sort (cuda_base_system.cu:26)
 *                           + ComputeBoundaryVolume (cuda_sf_system.cu:121-136)
 *   sf_build_neighbors   <- CudaGNBaseSearcher::BuildGNSearcher  cuda_neighbor_searcher.cu:36-51
 *   sf_update_system     <- CudaBaseSystem::UpdateSystem  cuda_base_system.cu:29-56 (-> CudaSFSystem::OnUpdateSolver cuda_sf_system.cu:66-119)
 *   sf_step              <- the same per-step pass sequence with a caller-fixed dt (cuda_wcsph_sf_solver.cpp:44-177)
 *   sf_num_sub_steps     <- CudaBaseSystem::GetNumOfSubTimeSteps  cuda_base_system.cuh:38
 *   sf_update_frame      <- the frame loop of Update()  seepage_flows/src/seepageflow/main.cpp:235-259
 *   sf_download / sf_device_buffer <- CudaSFParticles::Get*Ptr()  cuda_sf_particles.cuh:170-190,
 *                           CudaGNBaseSearcher::GetCellStart()  cuda_neighbor_searcher.cuh:44-50
 *   sf_particle_view     <- the same getters in the reference's array layout: CudaSFParticles::Get*Ptr()
 *                           (cuda_sf_particles.cuh:170-190, CudaParticles::GetPosPtr cuda_particles.cuh:57)
 *                           and GetCellStart() (cuda_neighbor_searcher.cuh:44-50)
 *   sf_update_solver     <- CudaWCSphSFSolver::UpdateSolver  cuda_wcsph_sf_solver.cuh:36-43,
 *                           cuda_wcsph_sf_solver.cpp:34-179 (and CudaSphSFSolver::UpdateSolver,
 *                           cuda_sph_sf_solver.cpp:14-152) on the current neighbour build
 *   sf_export_bgeo       <- ExportBgeoFileCUDA  seepage_flows/include/kiri_utils.h:364-448
 *   sf_export_bgeo_device <- ExportBgeoFileCUDA(folder, name, pos, col, radius, label, n)
 *                           seepage_flows/include/kiri_utils.h:364-371 (device arrays in)
 *   sf_write_bgeo        <- Partio::writeBGEO (uncompressed V5)  extlib/include/partio/io/BGEO.cpp:311-476
 *   sf_build_world_boundary <- CudaBoundaryEmitter::BuildWorldBoundary  cuda_boundary_emitter.cpp:12-80
 *   sf_read_bgeo         <- ReadBgeoFileForGPU  seepage_flows/include/kiri_utils.h:252-308
 *                           over Partio::readBGEO  extlib/include/partio/io/BGEO.cpp:196-309
 *   sf_build_seepage_shape_volume / sf_build_seepage_shape_multi_volume
 *                        <- CudaVolumeEmitter::BuildSeepageflowShape(Multi)Volume  cuda_volume_emitter.cpp:109-175
 *   sf_bgeo_to_vtk       <- Partio2VTK  seepage_flows/include/kiri_utils.h:54-250
 *   sf_grid_size         <- make_int3((highest - lowest) / kernel_radius)  main.cpp:104, cuda_neighbor_searcher.cu:26
 *   sf_grid_build        <- CudaGNBaseSearcher::BuildGNSearcher  cuda_neighbor_searcher.cu:36-51 on caller arrays
 *   sf_set_emitter / sf_get_emitter <- CudaEmitter + CudaSphEmitterParams (see their declarations)
 *   sf_grid_build_noalloc, sf_batch_*, sf_step_graph, sf_*_launch_config, sf_autotune, sf_set_field, sf_diff_*,
 *   sf_set_sort_strategy, sf_set_prefix_scan_threshold, sf_set_interface_lanes, sf_step_timed(_every):
 *                           no counterpart (see their declarations)
 *
 * Threading: a handle is single-threaded; all its GPU work is ordered on the
 * handle's own HIP stream (sf_set_stream can replace it).
 */
#ifndef SF_SEEPAGE_H
#define SF_SEEPAGE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_ABI_VERSION 2

/* ---- error codes ------------------------------------------------------- */
#define SF_OK 0
#define SF_E_INVALID (-1)   /* bad argument / parameter                     */
#define SF_E_CAPACITY (-2)  /* particle count above the handle's capacity   */
#define SF_E_HIP (-3)       /* a HIP runtime call failed                    */
#define SF_E_STATE (-4)     /* call out of order (e.g. step before upload)  */
#define SF_E_IO (-5)        /* file could not be written                    */
#define SF_E_NOTFOUND (-6)  /* unknown field id                             */

/* ---- plain vector types (layout of CUDA float3 / int3) ----------------- */
typedef struct sf_float3 { float x, y, z; } sf_float3;
typedef struct sf_int3 { int32_t x, y, z; } sf_int3;

/* CudaSphSolverType (cuda_sph_params.h:19-24) */
#define SF_SPH_SOLVER 0
#define SF_WCSPH_SOLVER 1
/* CudaSeepageflowType (cuda_seepageflow_params.h:19-23) */
#define SF_TYPE_SF 0
#define SF_TYPE_MULTI_SF 1

/* Mirrors CudaSeepageflowParams field for field (cuda_seepageflow_params.h:25-69). */
typedef struct sf_params {
    /* water (SPH) */
    float sph_mass;
    float sph_density;
    float sph_particle_radius;
    float sph_kernel_radius;
    float sph_stiff;
    float sph_visc;
    float sph_nu;
    float sph_bnu;
    /* sand (DEM) */
    float dem_mass;
    float dem_density;
    float dem_particle_radius;
    float dem_kernel_radius;
    float dem_young;
    float dem_poisson;
    float dem_tan_friction_angle;
    float dem_damping;
    /* seepage flow */
    float sf_c0;
    float sf_csat;
    float sf_cmc;
    float sf_cmc_p;
    float sf_cd;
    float sf_a0;
    float sf_asat;
    float sf_amc;
    float sf_amc_p;
    sf_float3 sf_dry_sand_color;
    sf_float3 sf_wet_sand_color;
    sf_float3 gravity;
    float dt;
    int32_t solver_type; /* SF_SPH_SOLVER | SF_WCSPH_SOLVER */
    int32_t sf_type;     /* SF_TYPE_SF | SF_TYPE_MULTI_SF  */
} sf_params;

/* Mirrors CudaBoundaryParams (cuda_boundary_params.h:19-27). grid_size must equal
 * sf_grid_size(lowest_point, highest_point, kernel_radius). */
typedef struct sf_boundary_params {
    float kernel_radius;
    sf_float3 lowest_point;
    sf_float3 highest_point;
    sf_float3 world_size;
    sf_float3 world_center;
    sf_int3 grid_size;
} sf_boundary_params;

/* Mirrors the CudaWCSphSFSolver constructor (cuda_wcsph_sf_solver.cuh:22-32):
 * negativeScale = 0, timeStepLimitScale = 3, speedOfSound = 100. */
typedef struct sf_solver_options {
    float negative_scale;
    float time_step_limit_scale;
    float speed_of_sound;
} sf_solver_options;

/* ---- field ids for sf_download / sf_device_buffer ----------------------
 * float3 fields download as packed float[3*n]; labels as uint64 (the
 * reference's size_t); keys, permutation and cell starts as uint32.       */
enum sf_field {
    SF_FIELD_POS = 0,            /* float3  particle position (current order)      */
    SF_FIELD_VEL = 1,            /* float3                                          */
    SF_FIELD_ACC = 2,            /* float3  acceleration of the last step           */
    SF_FIELD_COL = 3,            /* float3  colour                                  */
    SF_FIELD_LABEL = 4,          /* uint64  0 = water, 1 = sand                      */
    SF_FIELD_MASS = 5,           /* float                                           */
    SF_FIELD_RADIUS = 6,         /* float                                           */
    SF_FIELD_DENSITY = 7,        /* float                                           */
    SF_FIELD_PRESSURE = 8,       /* float                                           */
    SF_FIELD_VOIDAGE = 9,        /* float                                           */
    SF_FIELD_SATURATION = 10,    /* float                                           */
    SF_FIELD_MAX_SATURATION = 11,/* float                                           */
    SF_FIELD_AVG_FLOW_VEL = 12,  /* float3                                          */
    SF_FIELD_AVG_DRAG_FORCE = 13,/* float3  sand entries; water entries are 0       */
    SF_FIELD_ADHESION_FORCE = 14,/* float3  water entries; sand entries are 0       */
    SF_FIELD_AVG_ADHESION_FORCE = 15, /* float3 water entries; sand entries are 0   */
    SF_FIELD_CDA0ASAT = 16,      /* float3  per-particle (cd, a0, asat)             */
    SF_FIELD_AMCAMCP = 17,       /* float2  per-particle (amc, amcp)                */
    SF_FIELD_PARTICLE2CELL = 18, /* uint32  cell key of each particle before the last sort */
    SF_FIELD_SORT_PERM = 19,     /* uint32  perm[new] = old index of the last sort  */
    SF_FIELD_SORTED_KEYS = 20,   /* uint32  keys after the last sort                */
    SF_FIELD_CELL_START = 21,    /* uint32  numCells + 1 entries; [k+1] is cellEnd[k] */
    SF_FIELD_NBR_COUNT = 22,     /* uint32  per particle: 27-cell candidates with r^2 < h^2 (self included) */
    SF_FIELD_NBR_CANDIDATES = 23,/* uint32  per particle: 27-cell candidate count   */
    SF_FIELD_B_POS = 24,         /* float3  boundary particles, sorted by cell      */
    SF_FIELD_B_LABEL = 25,       /* uint64                                          */
    SF_FIELD_B_VOLUME = 26,      /* float                                           */
    SF_FIELD_B_CELL_START = 27,  /* uint32  numCells + 1 entries                    */
    SF_FIELD_COUNT_ = 28
};

typedef struct sf_system* sf_handle;

/* ---- host-only helpers (no GPU needed) --------------------------------- */
int sf_abi_version(void);
const char* sf_last_error(void);
/* Validates parameter structs exactly as sf_create does, without a GPU. */
int sf_validate_params(const sf_params* p, const sf_boundary_params* b);
/* make_int3((hi - lo) / h), FP32, truncation (main.cpp:104). */
int sf_grid_size(sf_float3 lowest, sf_float3 highest, float kernel_radius, sf_int3* out);
/* Six wall planes of boundary particles at spacing 2r, label 0, in the
 * reference's plane order. Call with pos3 == NULL to get the count in *n_out;
 * otherwise `capacity` is the number of particles pos3/label can hold. */
int sf_build_world_boundary(sf_float3 lowest, sf_float3 highest, float particle_radius,
                            size_t capacity, float* pos3, uint64_t* label, size_t* n_out);
/* Writes an uncompressed Houdini bgeo V5 file with position, Cd, pscale, label. */
int sf_write_bgeo(const char* path, size_t n, const float* pos3, const float* col3,
                  const float* radius, const uint64_t* label);

/* ---- scene input (host only) --------------------------------------------- */
/* Every point of a classic Houdini bgeo V5 file (uncompressed or gzip) as
 * (x, y, z, pscale) float4 rows; pscale is 0.01 where the file has no
 * "pscale" point attribute, and flip_yz stores (x, z, y).  shape4 == NULL
 * only counts (*n_out).  Attribute types float / int / vector / indexed
 * string are read past; a string attribute, another version or the newer
 * JSON bgeo are SF_E_IO. */
int sf_read_bgeo(const char* path, int flip_yz, float* shape4, size_t capacity, size_t* n_out);
/* Converts a bgeo V5 file to a binary legacy VTK 4.1 unstructured grid: the
 * positions as POINTS, one vertex cell per particle (its point is the "id"
 * attribute when the file has one, else the particle index), and every other
 * float / int / vector point attribute as POINT_DATA FIELD arrays with all
 * their components.  An empty file writes nothing. */
int sf_bgeo_to_vtk(const char* bgeo_path, const char* vtk_path);

/* SeepageflowVolumeData / SeepageflowMultiVolumeData (cuda_volume_emitter.cuh:43-63)
 * as caller-owned arrays of `capacity` entries, `size` of them in use. */
typedef struct sf_volume_data {
    float* pos3;
    float* col3;
    uint64_t* label;
    float* mass;
    float* radius;
    float* cda0asat3;       /* multi-material builder only */
    float* amcamcp2;        /* multi-material builder only */
    size_t size;            /* in/out */
    size_t capacity;
    float sand_min_radius;  /* in/out */
} sf_volume_data;
/* Appends the n shape points (sf_read_bgeo rows) as sand, label 1, with
 * mass = sand_density * ((4/3) pi r^3).  sand_min_radius restarts from
 * float(SIZE_MAX) on every call (the multi builder: only while size == 0)
 * and takes the smallest r.  offset_y != 0 then lowers every particle held
 * (earlier shapes too) by min_y - (world_lowest_y + 5 sand_min_radius), with
 * min_y this shape's lowest y.  SF_E_CAPACITY if size + n > capacity. */
int sf_build_seepage_shape_volume(sf_volume_data* d, const float* shape4, size_t n, sf_float3 color,
                                  float sand_density, int offset_y, float world_lowest_y);
/* The same, also writing (cd, a0, asat) and (amc, amcp) for the shape's
 * particles (the MULTI_SF per-particle coefficients). */
int sf_build_seepage_shape_multi_volume(sf_volume_data* d, const float* shape4, size_t n, sf_float3 color,
                                        float sand_density, sf_float3 cda0asat, float amc, float amcp,
                                        int offset_y, float world_lowest_y);

/* ---- scene lifetime ----------------------------------------------------- */
/* opts may be NULL (reference defaults). device < 0 keeps the current device. */
int sf_create(const sf_params* p, const sf_boundary_params* b, const sf_solver_options* opts,
              size_t max_num, int device, sf_handle* out);
int sf_destroy(sf_handle h);
/* Replace the parameter structs used by the next steps (the reference passes
 * them by value to every UpdateSolver call). grid/box must not change. */
int sf_set_params(sf_handle h, const sf_params* p);
/* hipStream_t as void*; NULL restores the handle's own stream.  Work the
 * handle queued on the previous stream stays ahead of its work on the new
 * one (the new stream waits for an event recorded on the old). */
int sf_set_stream(sf_handle h, void* hip_stream);
int sf_synchronize(sf_handle h);

/* Host arrays; vel3, col3, cda0asat3, amcamcp2 may be NULL (zero / default).
 * Labels must be 0 (water) or 1 (sand); masses must be > 0. */
int sf_upload_particles(sf_handle h, size_t n, const float* pos3, const float* vel3,
                        const float* col3, const uint64_t* label, const float* mass,
                        const float* radius, const float* cda0asat3, const float* amcamcp2);
/* Sorts boundary particles by cell and computes their volumes. */
int sf_upload_boundary(sf_handle h, size_t nb, const float* pos3, const uint64_t* label);

/* ---- stepping ----------------------------------------------------------- */
int sf_build_neighbors(sf_handle h);
int sf_step(sf_handle h, float dt);
int sf_step_n(sf_handle h, float dt, int n_steps);
int sf_update_system(sf_handle h, float render_interval, float* ms_out);
/* The solver part of sf_update_system on the neighbour structures the last
 * sf_build_neighbors built: the passes, the sub-step count (CFL for the WCSPH
 * solver) and the advect, without the build and without emission (the
 * reference emits in CudaSFSystem::OnUpdateSolver, after UpdateSolver).
 * SF_E_STATE unless sf_build_neighbors ran since the last step. */
int sf_update_solver(sf_handle h, float render_interval, float* ms_out);
int sf_num_sub_steps(sf_handle h, size_t* out);
/* One frame of the reference driver's loop, with the results of calling
 * sf_update_system from the host the way main.cpp:235-259 does:
 *   adaptive != 0:  remaining = render_interval;
 *                   while (remaining > 1e-6f) { sf_update_system(h, remaining);
 *                                               remaining -= remaining / nsub; }
 *   adaptive == 0:  k = nsub (as left by the previous call);
 *                   k times sf_update_system(h, render_interval);
 * The remaining-time update and the emitter clock run on the device, and the
 * host reads them back only after as many sub-steps as the loop certainly
 * still runs (the CFL dt has a fixed upper bound); the particle state, nsub
 * and the emitter state afterwards are those of the host loop, bit for bit.
 * *sub_steps (may be NULL) = sub-steps run; *ms_out (may be NULL) = GPU time. */
int sf_update_frame(sf_handle h, float render_interval, int adaptive, size_t* sub_steps, float* ms_out);
/* Passes of one step, in launch order, for sf_step_profiled. */
enum sf_pass {
    SF_PASS_HASH = 0, SF_PASS_SORT = 1, SF_PASS_GATHER = 2, SF_PASS_CELL_START = 3,
    SF_PASS_DENSITY = 4, SF_PASS_AVG_FLOW = 5, SF_PASS_VOIDAGE_ADHESION = 6,
    SF_PASS_SAND_MOMENTUM = 7, SF_PASS_WATER_MOMENTUM = 8, SF_PASS_ADVECT = 9,
    SF_PASS_COUNT = 10
};
/* n_steps fixed-dt steps with HIP events recorded on the handle's stream
 * around every pass (no host synchronisation inside the region).  pass_ms
 * (SF_PASS_COUNT floats, may be NULL) receives the mean duration of each pass
 * per step; total_ms the elapsed time of the whole region.  No counterpart in
 * the reference: it replaces the cudaEvent timing of UpdateSystem
 * (cuda_base_system.cu:31-55) with per-pass resolution. */
int sf_step_profiled(sf_handle h, float dt, int n_steps, float* pass_ms, float* total_ms);
/* The same with events around one pass only (pass in SF_PASS_HASH ..
 * SF_PASS_ADVECT): every event recorded between two kernels costs the stream
 * a few microseconds, so this times a whole run with the least disturbance
 * while still timing that pass inside it.  *pass_ms (may be NULL) receives
 * the pass's mean duration per step, *total_ms the whole region.  No
 * counterpart in the reference. */
int sf_step_timed(sf_handle h, float dt, int n_steps, int pass, float* pass_ms, float* total_ms);
/* sf_step_timed with the pass's events on steps 0, every, 2 every, ... only:
 * *pass_ms is the mean over those steps, and the other steps run without
 * events.  every = 1 is sf_step_timed.  No counterpart in the reference. */
int sf_step_timed_every(sf_handle h, float dt, int n_steps, int pass, int every, float* pass_ms, float* total_ms);
int sf_size(sf_handle h, size_t* n, size_t* nb, size_t* num_cells);

/* ---- data access -------------------------------------------------------- */
int sf_download(sf_handle h, int field, void* dst, size_t dst_bytes);
/* Overwrites a persistent per-particle field in the current slot order:
 * SF_FIELD_VEL / SF_FIELD_COL / SF_FIELD_CDA0ASAT (float3), SF_FIELD_MAX_SATURATION
 * (float), SF_FIELD_AMCAMCP (float2).  No counterpart in the reference, whose
 * particle constructor zeroes maxSaturation (cuda_sf_particles.cuh:201-235);
 * it gives scenes an initial saturation history. */
int sf_set_field(sf_handle h, int field, const void* src, size_t src_bytes);
/* Raw device pointer and element stride in bytes of a field's storage
 * (float3 fields are stored as float4, stride 16). Valid until the next step.
 * Once the position buffer (SF_FIELD_POS / MASS / LABEL) has been handed out,
 * every later neighbour build of the handle hashes the positions again, so
 * writes through it between steps are seen (otherwise the step's advect
 * already computed the next build's cell keys), and the water momentum pass
 * reads the slot arrays instead of its two-load records (the masses may
 * change: 1-4 % slower per step at 1M particles, same results). */
int sf_device_buffer(sf_handle h, int field, void** dptr, size_t* stride_bytes);
/* A device array of the field in the reference's layout, as
 * CudaSFParticles::Get*Ptr() returns it: packed float3 (12-byte elements) for
 * SF_FIELD_POS, VEL, ACC, COL, AVG_FLOW_VEL, AVG_DRAG_FORCE, ADHESION_FORCE,
 * AVG_ADHESION_FORCE, CDA0ASAT; float for MASS, RADIUS, DENSITY, PRESSURE,
 * VOIDAGE, SATURATION, MAX_SATURATION; uint64 (size_t) for LABEL; float2 for
 * AMCAMCP; uint64 for SORTED_KEYS (GetGridIdxArray(), the keys after the
 * last sort), CELL_START and B_CELL_START (numCells + 1 entries, as
 * GetCellStart()).  The array has room for max_num particles, belongs to the
 * handle and keeps its address until sf_destroy.  The handle keeps every view
 * it handed out current: each call that changes the state (upload, set_field,
 * build, step, update, frame, differentiable window) refreshes it before
 * returning, in the current slot order (the particles are re-sorted by cell
 * every step, as the reference's arrays are).  Values written into a view of
 * a persistent field (POS, VEL, COL, LABEL, MASS, RADIUS, MAX_SATURATION,
 * CDA0ASAT, AMCAMCP) between calls are taken into the state at the start of
 * the next build or step; a label other than 0 reads as sand, masses must
 * stay > 0.  Writes into the other views are overwritten by the next step,
 * which recomputes those fields (in the reference too).  sf_download and
 * sf_export_bgeo read the state, not the views.  Handing out the position,
 * mass or label view makes every later build hash the positions again and
 * turns the water momentum pass's two-load records off, as sf_device_buffer
 * does. */
int sf_particle_view(sf_handle h, int field, void** dptr);
/* Writes folder/name.bgeo from the current state (creates folder). */
int sf_export_bgeo(sf_handle h, const char* folder, const char* name);
/* Writes folder/name.bgeo (creates folder) from n particles in device memory
 * of the current device, in the reference's layout: packed float3 positions
 * and colours, float radii, uint64 labels (sf_particle_view arrays). */
int sf_export_bgeo_device(const char* folder, const char* name, const float* pos3, const float* col3,
                          const float* radius, const uint64_t* label, size_t n);

/* ---- the searcher on caller arrays ----------------------------------------
 * CudaGNBaseSearcher::BuildGNSearcher (cuda_neighbor_searcher.cu:36-51) for
 * any particle set the caller holds on the current device: pos3 = n packed
 * float3 positions (device memory).  Writes sorted_keys[n] (the cell keys in
 * sorted order, GetGridIdxArray()), perm[n] (the stable sort's permutation,
 * perm[new] = old; the reference permutes the particle arrays by it) and
 * cell_start[numCells + 1] (GetCellStart(); [k + 1] is cellEnd[k]), with
 * numCells = grid_size.x * y * z, keys and cells exactly as the handle's
 * searcher computes them.  The positions are not moved.  Work is ordered on
 * `stream` (a hipStream_t; NULL = the null stream) and finished on return;
 * the call allocates and frees its own work arrays (about 8 bytes per
 * particle).  n < 2^27, numCells < 2^31. */
int sf_grid_build(const float* pos3, size_t n, sf_float3 lowest, float cell_size, sf_int3 grid_size,
                  uint32_t* sorted_keys, uint32_t* perm, uint32_t* cell_start, void* stream);
/* No counterpart in the reference.  sf_grid_build for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same outputs, bit for bit, with the work arrays in
 * a library-owned device array of 8.25 MiB per GPU (held until the process
 * ends) instead of an allocation.  At most 1048576 particles (SF_E_CAPACITY
 * above).  The call allocates and frees nothing, enqueues its kernels on
 * `stream` and returns without waiting for them.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other threads,
 * other streams without ordering) race and must not be issued. */
int sf_grid_build_noalloc(const float* pos3, size_t n, sf_float3 lowest, float cell_size, sf_int3 grid_size,
                          uint32_t* sorted_keys, uint32_t* perm, uint32_t* cell_start, void* stream);

/* ---- inflow emitter ------------------------------------------------------
 * sf_emitter_params mirrors CudaSphEmitterParams (cuda_sph_params.h:49-70);
 * sf_set_emitter replaces the CudaEmitter ctor + Build*Emitter calls of the
 * CudaSFSystem ctor (cuda_emitter.cuh:24-55, cuda_emitter.cpp:12-97,
 * cuda_sf_system.cu:36-58): the sample points are built with the water
 * particle radius exactly as the reference's float loops build them.  Each
 * sf_update_system then emits after its step as CudaSFSystem::OnUpdateSolver
 * does (cuda_sf_system.cu:86-115): WCSPH, when |v| * elapsed / (2 r_w) reaches
 * the emission counter (elapsed += render_interval / sub-steps); SPH, every
 * int(2 r_w / (|v| dt))-th call.  Emitted particles (AddSphParticles,
 * cuda_sf_particles.cu:50-68) are water of mass sph_mass and radius
 * sph_particle_radius with the emitter's velocity and colour, appended after
 * the current particles; when they would not fit below max_num the emitter
 * switches itself off.  sf_step / sf_step_n / sf_step_graph do not emit.
 * The emission counter and elapsed time are the system's (zero at sf_create)
 * and survive later sf_set_emitter calls, as the reference's CudaSFSystem
 * members survive changes to its emitter and to CUDA_SPH_EMITTER_PARAMS.run. */
#define SF_EMITTER_SQUARE 0
#define SF_EMITTER_CIRCLE 1
#define SF_EMITTER_RECTANGLE 2
typedef struct sf_emitter_params {
    int enable;           /* CudaSphEmitterParams::enable: build the samples, emitter on */
    int run;              /* CudaSphEmitterParams::run: emission switched on */
    sf_float3 emit_pos;
    sf_float3 emit_vel;   /* also the emitted particles' velocity; must not be zero */
    sf_float3 emit_col;
    int emit_type;        /* SF_EMITTER_SQUARE / CIRCLE / RECTANGLE */
    float emit_radius;    /* square and circle */
    float emit_width;     /* rectangle */
    float emit_height;    /* rectangle */
} sf_emitter_params;
int sf_set_emitter(sf_handle h, const sf_emitter_params* e);
/* enabled: the emitter's status (it turns itself off at capacity);
 * num_points: samples per emission; emitted: particles emitted so far. */
int sf_get_emitter(sf_handle h, int* enabled, size_t* num_points, size_t* emitted);

/* ---- batched scenes ------------------------------------------------------
 * No counterpart in the reference (one CudaSFSystem per process, main.cpp:106-221):
 * many independent scenes of uneven particle counts in one handle, every pass
 * one launch over all of them.  Scenes share the box, grid, boundary particles
 * and every sf_params field except these, whose names follow sf_params. */
typedef struct sf_scene_params {
    float dem_tan_friction_angle;
    float sf_c0, sf_cmc, sf_cmc_p, sf_csat; /* capillary Bezier controls */
    float sf_cd;                           /* drag coefficient (SF mode)  */
} sf_scene_params;
/* counts[s] particles for scene s; the handle's capacity is their sum.  Upload
 * with sf_upload_particles, scene-major (scene 0's particles first).  Scene s
 * stays in the slot range [off[s], off[s+1]) of every field for the handle's
 * lifetime (off from sf_batch_scene_offsets); within it the particles are in
 * cell order, exactly as a single-scene handle of that scene keeps them.
 * Batched handles step with a fixed dt (sf_step*, sf_step_graph). */
int sf_batch_create(const sf_params* p, const sf_boundary_params* b, const sf_solver_options* opts,
                    size_t n_scenes, const size_t* counts, const sf_scene_params* scenes, int device,
                    sf_handle* out);
int sf_batch_set_scene_params(sf_handle h, const sf_scene_params* scenes);
/* offsets: n_scenes + 1 entries; *n_scenes may be NULL. */
int sf_batch_scene_offsets(sf_handle h, size_t* n_scenes, size_t* offsets, size_t capacity);

/* ---- graph replay and launch tuning -------------------------------------
 * No counterpart in the reference (which synchronises after every pass,
 * kiri_pbs_pch.cuh:35-50).  sf_step_graph captures two fixed-dt steps into a
 * HIP graph once (again after any change of dt, parameters, uploads or launch
 * shapes) and replays it; results are identical to sf_step_n. */
int sf_step_graph(sf_handle h, float dt, int n_steps);
/* Sort of the neighbour build.  SF_SORT_INCREMENTAL (default) rebuilds the
 * stable sort from the previous step's keys and cellStart when the particles
 * are still in that step's order (most keep their cell); the first build
 * after an upload always uses the LSD radix sort.  Both give the stable
 * sort's permutation element for element. */
#define SF_SORT_RADIX 0
#define SF_SORT_INCREMENTAL 1
int sf_set_sort_strategy(sf_handle h, int strategy);
/* Largest number of 2048-slot tiles up to which the slot prefixes of the
 * incremental sort and of the water prefix sum the earlier tiles' counts
 * inside each prefix block; above it a one-block scan launch runs in between
 * (default 4096 tiles, about 8.4M slots).  Both branches give the same
 * prefixes; the switch lets a test run both on one input.  Process-wide (no
 * handle); graphs captured before keep the branch they were captured with.
 * tiles < 0 restores the default.  Returns the previous threshold.  No
 * counterpart in the reference. */
int sf_set_prefix_scan_threshold(int tiles);
/* Launch shape of a neighbour pass (SF_PASS_DENSITY .. SF_PASS_WATER_MOMENTUM):
 * workgroup size nb in {64, 128, 256}; LDS candidate-list depth cap in {8, 16}
 * (sand / water momentum only).  Results do not depend on the shape. */
int sf_set_launch_config(sf_handle h, int pass, int nb, int cap);
int sf_get_launch_config(sf_handle h, int pass, int* nb, int* cap);
/* Lanes per particle (1, 2 or 4; default 4) of the two interface passes
 * (SF_PASS_AVG_FLOW, SF_PASS_VOIDAGE_ADHESION), whose neighbour sums run only
 * for the particles with neighbours of the other phase.  More lanes shorten
 * the walks where the phases mix; they regroup those sums, so results change
 * at rounding level, and handles that must agree bit for bit use the same
 * value.  The autotuner keeps it. */
int sf_set_interface_lanes(sf_handle h, int lanes);
/* Times every shape of every neighbour pass on the handle's current particles
 * (reps launches each, after one neighbour build) and keeps the fastest.
 * best_ms / default_ms (5 floats each, may be NULL) receive per pass the
 * chosen and the default (256, 16) shape's mean time. */
int sf_autotune(sf_handle h, int reps, float* best_ms, float* default_ms);

/* ---- differentiable window ----------------------------------------------
 * No counterpart in the reference.  Gradients of a loss on the state after
 * n_steps fixed-dt steps with respect to the initial positions and velocities
 * and to four parameters: sf_cd (drag / permeability), sf_c0 and sf_cmc
 * (capillary curve), dem_tan_friction_angle.  Hand-written reverse passes of
 * every force pass and the integrator replay each step's recorded sort,
 * neighbour ranges and branch decisions; the forward keeps a state checkpoint
 * every checkpoint_every steps and the backward recomputes one segment at a
 * time.  Single-scene SF-mode handles without label-1 collider particles.
 * Every per-particle array here is in id order: the slot order of the state
 * the window started from.  on_device: pointers are device memory on the
 * handle's GPU (else host memory). */
int sf_diff_forward(sf_handle h, float dt, int n_steps, int checkpoint_every);
/* Back to the window's initial state; vel3 (may be NULL) replaces its velocities. */
int sf_diff_reset(sf_handle h, const float* vel3, int on_device);
/* SF_FIELD_POS / SF_FIELD_VEL after the window, SF_FIELD_SATURATION of its last step. */
int sf_diff_get(sf_handle h, int field, float* dst, int on_device);
/* Seeds (any may be NULL = 0): d loss / d final position, velocity, last-step
 * saturation.  Results (any may be NULL): param_bar4 (host) = d loss / d
 * (sf_cd, sf_c0, sf_cmc, dem_tan_friction_angle); d loss / d initial position
 * and velocity.  The handle is left at the end of the window. */
int sf_diff_backward(sf_handle h, const float* pos_bar3, const float* vel_bar3, const float* sat_bar,
                     float* param_bar4, float* pos0_bar3, float* vel0_bar3, int on_device);

#ifdef __cplusplus
}
#endif
#endif /* SF_SEEPAGE_H */
