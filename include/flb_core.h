/*
 * flb_core.h -- C ABI of the MI355X-native FlowLB lattice-Boltzmann core.
 *
 * The FlowLB C++ host runs its solver through one templated call,
 * FLB::h_launchCudaCalculations2D<T> (src/cuda/runCuda.cuh:26), which does the
 * set-up, the per-step loop and the readback internally.  This library splits
 * that call into a handle with explicit set-up, advance and readback entry
 * points, with plain pointers and sizes only, integer status codes and no
 * process exit on error.  Every entry point below names the reference code it
 * replaces.
 *
 * Conventions (same as the reference):
 *   - cell index n = x + y*nx, y = 0 at the bed;
 *   - vector fields are stored [x components | y components] (2*N floats);
 *   - flags are one byte per cell with the bit values of FLB::TypesNodes
 *     (src/utils.h:91-108), see FLB_* below;
 *   - the D2Q9 velocity order is 0:(0,0) 1:(+1,0) 2:(-1,0) 3:(0,+1) 4:(0,-1)
 *     5:(+1,+1) 6:(-1,-1) 7:(+1,-1) 8:(-1,+1) (src/cuda/cudaInitData.cu:76-78).
 *
 * Threading: one handle per host thread; calls on one handle are serialised by
 * the caller.  flb_step only enqueues work on the handle's HIP stream (a
 * free-surface handle keeps at most 8 steps queued: the host sizes each
 * stream-collide launch from a bounding box the device published a few steps
 * earlier, so flb_step may wait for the step 8 back); the readback calls
 * (flb_read*, flb_mass) and flb_sync wait for it.
 */
#ifndef FLB_CORE_H
#define FLB_CORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLB_ABI_VERSION 1

/* Status codes (the reference calls exit(EXIT_FAILURE), src/cuda/cudaUtils.cuh:17-26). */
enum {
  FLB_OK = 0,
  FLB_ERR_ARGUMENT = 1,    /* invalid argument or option */
  FLB_ERR_DEVICE = 2,      /* HIP runtime error (no device, launch failure, ...) */
  FLB_ERR_MEMORY = 3,      /* device or host allocation failed */
  FLB_ERR_STATE = 4,       /* call not valid in the handle's state */
  FLB_ERR_UNSUPPORTED = 5  /* option combination not implemented */
};

/* Cell flags, FLB::TypesNodes (src/utils.h:91-108). */
enum {
  FLB_FLUID = 0x01, FLB_GAS = 0x02, FLB_INTERFACE = 0x04, FLB_WALL = 0x08,
  FLB_MOVING_WALL = 0x10, FLB_INLET = 0x20, FLB_OUTLET = 0x40, FLB_GAS_INTERFACE = 0x80
};

/* Population storage in device memory.  All formats store f_i - w_i
   (DDF-shifted; FP16 scaled by 2^15).  FP32, FP16S and BF16S compute in FP32
   with FP32 cell fields.  FP64 is the reference's PRECISION = 64
   (h_launchCudaCalculations2D<double>, src/ui/app.cpp:300-310): populations,
   fields and arithmetic in FP64, with the float-typed constants and float
   temporaries of the reference kept (1/tau, F, Smagorinsky sqrtf, TRT
   temporaries, neighbour averages).  Free-surface and monofluid runs; no
   adjoint or batch. */
enum { FLB_STORAGE_FP32 = 0, FLB_STORAGE_FP16S = 1, FLB_STORAGE_BF16S = 2, FLB_STORAGE_FP64 = 3 };
/* OptionsCalculation::typeProblem (src/io/reader.h:104). */
enum { FLB_MONOFLUID = 0, FLB_FREE_SURFACE = 1 };
/* OptionsCalculation::collisionOperator: 0 = SRT, 1 = TRT (src/cuda/lbmKernel2D.cu:444-511).
   As written the TRT step does not collide when a volumetric force is applied
   (:485-488).  FLB_TRT_GUO (no reference counterpart) defines that branch:
   TRT plus the Guo terms split into their pair-symmetric part, weighted
   1 - w+/2, and antisymmetric part, weighted 1 - w-/2; unforced it is TRT. */
enum { FLB_SRT = 0, FLB_TRT = 1, FLB_TRT_GUO = 2 };
/* Excess-mass accounting of the free-surface passes.
   FLB_MASS_AS_WRITTEN: the reference's rules (receivers counted on the
   pre-pass flags, fresh own fill level in the exchange weight, fluid cells
   exchange over every link).  FLB_MASS_CONSERVATIVE: mass crosses only
   fluid/interface links with symmetric weights and excess goes only to cells
   that pull it, so total mass is conserved to round-off. */
enum { FLB_MASS_AS_WRITTEN = 0, FLB_MASS_CONSERVATIVE = 1 };

/* Solver options: the lattice-unit content of OptionsCalculation
   (src/io/reader.h:98-127) after FLB::h_initConstantDataCuda2D
   (src/cuda/cudaInitData.cu:57-108) converted it. */
typedef struct flb_options {
  float tau;                   /* relaxation time; 1/tau is formed in float */
  float fx, fy;                /* body force per cell (Guo forcing), lattice units */
  int32_t use_force;           /* USE_VOLUMETRIC_FORCE */
  int32_t collision;           /* FLB_SRT / FLB_TRT / FLB_TRT_GUO */
  int32_t smagorinsky;         /* SUBGRID_TURBULENCE_MODEL */
  int32_t problem;             /* FLB_MONOFLUID / FLB_FREE_SURFACE */
  int32_t storage;             /* FLB_STORAGE_* (PRECISION 32 -> FP32, 16 -> FP16S) */
  int32_t gas_check_corrected; /* 0: gas-reconstruction neighbour test as written
                                  (src/cuda/cellCalculations.cuh:250-259); 1: test
                                  the neighbour the population comes from */
  int32_t mass_mode;           /* FLB_MASS_* */
  int32_t block_x;             /* threads per block along x (64..1024, multiple of 64); 0 = 256 */
  int32_t device;              /* HIP device ordinal */
  int32_t row_pitch;           /* cells per row of the device arrays (>= nx); 0 = nx
                                  rounded up to 64 (see flb_get_layout) */
  int32_t whole_grid;          /* free surface: 1 launches stream-collide over the whole
                                  grid every step instead of over the window that can hold
                                  work (same results; a check and a fallback); 0 = window.
                                  No reference counterpart. */
} flb_options;

typedef struct flb_handle flb_handle;

/* Fills *opt with the defaults (SRT, FP32 storage, monofluid, tau = 0.6). */
void flb_default_options(flb_options* opt);

/* Set-up.  Replaces the set-up half of h_launchCudaCalculations2D
   (src/cuda/runCuda.cu:31-85): initFields (src/initData.h:125-165),
   h_initConstantDataCuda2D (src/cuda/cudaInitData.cu:57-108) and
   h_initDataCuda2D (src/cuda/cudaInitData.cuh:205-240).  flags: N bytes.
   rho0 (N floats) may be NULL for rho = 1; u0 (2N floats, [ux | uy]) may be
   NULL for u = 0 (inlet cells then need u0 for their prescribed velocity).
   Populations start at feq(rho0, u0) stored with the odd-step pattern. */
int flb_create(const flb_options* opt, const uint8_t* flags, uint32_t nx, uint32_t ny,
               const float* rho0, const float* u0, flb_handle** out);

/* Free-surface fields after initFreeSurfaceFields (src/initData.h:175-206).
   excess may be NULL (zero).  The populations are not touched (nor, with
   16-bit storage, the rounding residual stream-collide keeps for them), so
   writing back the fields of flb_read / flb_read_fs leaves a run unchanged.
   The _f64 form takes doubles: exact for FP64 storage (restoring an FP64
   checkpoint), rounded to float otherwise. */
int flb_set_fs_fields(flb_handle* h, const float* mass, const float* phi, const float* excess);
int flb_set_fs_fields_f64(flb_handle* h, const double* mass, const double* phi,
                          const double* excess);

/* Advance nsteps lattice steps.  Replaces the step loops
   h_runCudaMonoFluidNoGraphics2D / h_runCudaFreeSurfaceNoGraphics2D
   (src/cuda/runCuda.cu:473-490, 506-527): one d_StreamCollide2D per step, or
   d_FreeSurface2D_1 -> d_StreamCollide2D -> d_FreeSurface2D_2 -> _3 -> _4.
   The handle owns the time counter t (parity of the esoteric pull).  The
   density/velocity readback fields hold the values of the last step. */
int flb_step(flb_handle* h, uint64_t nsteps);

/* Waits for all work enqueued on the handle. */
int flb_sync(flb_handle* h);

/* Readback into caller-owned host buffers; every pointer may be NULL.
   Replaces CudaUtils::copyDataFromDevice (src/cuda/cudaUtils.cuh:42-53) with
   the fields the run loops save (src/cuda/runCuda.cu:483-487, 518-523):
   u (2N), rho (N), flags (N), phi (N, free surface only). */
int flb_read(flb_handle* h, float* u, float* rho, uint8_t* flags, float* phi);
/* The same in double: exact for FP64 storage, widened from float otherwise. */
int flb_read_f64(flb_handle* h, double* u, double* rho, uint8_t* flags, double* phi);

/* Free-surface mass and excess-mass fields (N values each, NULL allowed). */
int flb_read_fs(flb_handle* h, float* mass, float* excess);
int flb_read_fs_f64(flb_handle* h, double* mass, double* excess);

/* Raw population array f[q*N + n] in the esoteric-pull slot layout, as
   unshifted values (9N floats / doubles). */
int flb_read_populations(flb_handle* h, float* f);
int flb_read_populations_f64(flb_handle* h, double* f);
/* Writes the raw population array (same layout).  With 16-bit storage the
   stored values are rounded; a free-surface handle's rounding residual of the
   replaced populations is cleared.  The _f64 form is exact for FP64 storage. */
int flb_write_populations(flb_handle* h, const float* f);
int flb_write_populations_f64(flb_handle* h, const double* f);

/* Device pointers of the fields, for a HIP-GL interop consumer
   (the reference maps u onto a GL buffer, src/cuda/runCuda.cu:137-147).
   Any output pointer may be NULL.  u, rho and phi hold doubles for FP64
   storage, floats otherwise.  Device fields are pitched: cell (x, y) is
   element x + y * pitch of a plane of `plane` elements, and u holds the ux
   plane followed by the uy plane (flb_get_layout).  They stay valid until
   flb_destroy; the flag buffer of a free-surface handle alternates every
   step, so query it again after stepping. */
int flb_device_fields(flb_handle* h, void** d_u, void** d_rho, void** d_flags, void** d_phi);

/* Total fluid mass in FP64: stored mass plus the excess-mass shares that the
   next step pulls in (free surface), or the sum of all populations
   (monofluid).  No reference counterpart (the reference has no mass
   accounting). */
int flb_mass(flb_handle* h, double* total);

/* Lattice time t (number of steps done). */
uint64_t flb_time(const flb_handle* h);

/* Grid size and problem type (FLB_MONOFLUID / FLB_FREE_SURFACE) of a handle;
   any output pointer may be NULL. */
int flb_get_shape(const flb_handle* h, uint32_t* nx, uint32_t* ny, int32_t* problem);

/* Viewer output (the GUI's device kernels).  Both enqueue on the handle's
   stream and return; call flb_sync before the destination is read or its
   GL resource unmapped (the reference synchronises the device once per frame,
   src/cuda/runCuda.cu:258).  Destinations are device memory the caller owns,
   e.g. a GL pixel buffer mapped with hipGraphicsMapResources /
   hipGraphicsResourceGetMappedPointer and copied into the texture.

   flb_export_velocity replaces CudaUtils::save2DDataToOpenGL
   (src/cuda/cudaUtils.cu:107-118), launched per frame at
   src/cuda/runCuda.cu:201-218: texel (x, y) of a float2 image with rows of
   dst_pitch_bytes (>= 8 * nx, a multiple of 8, 8-byte aligned) receives
   (u_x, u_y) of cell (x, y) -- no flip, as in the reference -- converted to
   float for FP64 storage. */
int flb_export_velocity(flb_handle* h, void* d_dst, size_t dst_pitch_bytes);

/* flb_isosurface replaces d_CreateIsoSurfaceVector2D
   (src/cuda/cudaIsosurface.cuh:33-198, with getInterpolatedPoint, drawLine and
   marchingSquares, src/cuda/cudaIsosurface.cu:16-86), launched at
   src/cuda/runCuda.cu:246-251 (velocity) and :418-426 (fill level).  Marching
   squares over the squares of the camera window [x_min, x_max) x
   [y_min, y_max), counted in display rows from the top (display row y spans
   lattice rows ny-1-y and ny-2-y); each isoline segment is rasterised into
   an 8-bit image of width x height pixels (rows of tex_pitch_bytes) by
   Bresenham's line as 255, y flipped to GL's lower-left origin.  Vertex
   values are |u| (FLB_ISO_VELOCITY) or the fill level phi (FLB_ISO_PHI,
   free surface only; the reference passes iso_value 0.5 for it).  The
   reference runs the velocity kernel on phi and so reads a second plane past
   the end of phi (runCuda.cu:424); here phi is read as the scalar it is.
   clear != 0 zeroes the image first (the reference leaves clearing to the
   GL side).  Only the squares of the window are launched. */
enum { FLB_ISO_VELOCITY = 0, FLB_ISO_PHI = 1 };
typedef struct flb_iso_view {
  uint32_t width, height;   /* texture size in pixels */
  uint32_t x_min, x_max;    /* camera bounds in lattice nodes (getCameraDomainBounds) */
  uint32_t y_min, y_max;
  float iso_value;
  int32_t field;            /* FLB_ISO_VELOCITY or FLB_ISO_PHI */
  int32_t clear;
} flb_iso_view;
int flb_isosurface(flb_handle* h, const flb_iso_view* view, void* d_tex, size_t tex_pitch_bytes);

/* Attainable-bandwidth reference for stream-collide (no reference
   counterpart; the reference reports no bandwidth).  On a buffer of its own
   (nine planes of an nx x ny grid, pitched and strided as a handle's) on
   `device`, times `launches` launches after four warm-ups and returns the mean
   device time per launch; bytes_per_launch (may be NULL) receives the bytes
   the launch moves: 18 values per cell, 72 bytes in FP32, 36 in 16-bit.
   FLB_PROBE_ESOTERIC: stream-collide's esoteric-pull loads and stores in place
   in FP32 with no arithmetic; FLB_PROBE_ESOTERIC16: the same with 16-bit
   values; FLB_PROBE_COPY: nine FP32 planes copied to nine others;
   FLB_PROBE_INPLACE: nine FP32 planes incremented in place. */
enum { FLB_PROBE_ESOTERIC = 0, FLB_PROBE_COPY = 1, FLB_PROBE_INPLACE = 2, FLB_PROBE_ESOTERIC16 = 3 };
int flb_bandwidth_probe(int32_t device, uint32_t nx, uint32_t ny, int32_t kind, uint32_t launches,
                        float* ms_per_launch, double* bytes_per_launch);

/* Layout of the device fields of flb_device_fields: row stride `pitch`
   (cells, >= nx) and plane size `plane` (= pitch * ny elements).  No
   reference counterpart (the reference's device fields are unpadded). */
int flb_get_layout(const flb_handle* h, uint32_t* pitch, uint64_t* plane);

/* Enqueue nsteps steps between two HIP events on the handle's stream and
   return the elapsed device time in ms.  With per_kernel_ms non-NULL (array of
   FLB_NUM_KERNELS floats) each kernel launch is bracketed by its own events and
   the summed device time per kernel is returned instead.  FLB_K_FS_UPDATE is
   the one launch that runs free-surface passes 2, 3 and 4
   (d_FreeSurface2D_2/_3/_4, src/cuda/lbmKernel2D.cu:123-303). */
enum { FLB_K_STREAM_COLLIDE = 0, FLB_K_FS1 = 1, FLB_K_FS_UPDATE = 2, FLB_NUM_KERNELS = 3 };
int flb_step_timed(flb_handle* h, uint64_t nsteps, float* total_ms, float* per_kernel_ms);

/* Histogram of flags (256 counts), computed on the device. */
int flb_flag_histogram(flb_handle* h, uint64_t* counts);

/* flb_flag_histogram without allocating: no counterpart in the reference.  For
   callers that cannot allocate device memory (code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  The 256 device
   counters live in a file-scope array of the library (2 KiB per GPU, kept
   until the process ends), resolved per device by flb_create; the histogram
   always has 256 bins, so no grid exceeds it.  Every call on one GPU shares
   that array: calls that overlap on one GPU (from several threads, or on
   handles whose streams run concurrently) race and must be serialised by the
   caller.  The counters are cleared by a kernel at the start of each call. */
int flb_flag_histogram_noalloc(flb_handle* h, uint64_t* counts);

/* Last error message of the handle ("" if none), and a name for a status. */
const char* flb_last_error(const flb_handle* h);
const char* flb_status_string(int status);

/* Frees every device buffer of the handle (src/cuda/runCuda.cu:111-127). */
void flb_destroy(flb_handle* h);

/* ------------------------------------------------------------ adjoint */
/* Gradients of a time-averaged discharge.  No reference counterpart (FlowLB
   has no adjoint); these entry points extend the solver for optimisation of a
   culvert inlet.  Monofluid, SRT, FP32 or FP64 storage, no subgrid model.
   With FP64 storage the forward sweep, its trace and the adjoint are FP64
   throughout (the gradient of the FP64 discrete model); with FP32 storage the
   forward and its trace are FP32 and the adjoint FP64. */

/* Brinkman drag: per-cell solid fraction alpha (N floats) adds the force
   -alpha*rho*u to the Guo forcing of non-inlet/outlet cells, with the half-force
   velocity shift solved for u.  NULL removes it. */
int flb_set_brinkman(flb_handle* h, const float* alpha);

/* Objective J = (1/(t1-t0+1)) * sum_{t=t0..t1} Q(t), with Q(t) the sum over the
   non-wall cells of `column` of rho*u_x at step t, over a run of nsteps steps
   from the handle's initial state (t = 0).  Checkpoints every ckpt steps. */
typedef struct flb_adjoint_spec {
  uint64_t nsteps;
  uint64_t t0, t1;
  uint32_t column;
  uint32_t ckpt;
  /* Device bytes for per-step traces kept from the forward sweep (0: half the
     free device memory).  Checkpoint segments whose trace does not fit are
     recomputed from their checkpoint in the reverse sweep. */
  uint64_t trace_budget_bytes;
} flb_adjoint_spec;

/* Forward objective only; the handle is left at t = 0. */
int flb_objective(flb_handle* h, const flb_adjoint_spec* spec, double* J);

/* Forward sweep keeping the per-step trace of the last checkpoint segments
   that fit in the trace budget and a checkpoint every ckpt steps for the
   others; reverse sweep that recomputes only the segments without a resident
   trace.  The adjoint state and dJ/dalpha are FP64; dJ/dalpha is returned
   rounded to float.  Outputs J, dJ/dalpha (N floats, NULL allowed; zeros
   without Brinkman drag), dJ/du_inlet (x component of the velocity held by
   the inlet cells).  phase_ms (NULL allowed) receives the device time of the
   forward sweep, the recompute passes and the reverse sweep.  The handle is
   left at t = 0. */
int flb_adjoint(flb_handle* h, const flb_adjoint_spec* spec, double* J, float* dj_dalpha,
                double* dj_duin, float* phase_ms);

/* ------------------------------------------------------------ batch */
/* Many independent monofluid SRT domains of uneven sizes advanced together, one
   kernel launch per step for all of them, runs of steps replayed from HIP
   graphs.  No reference counterpart (FlowLB runs one domain per process, each
   step being launch + device synchronisation, src/cuda/runCuda.cu:473-490). */
typedef struct flb_instance {
  uint32_t nx, ny;
  float tau, fx, fy;
  int32_t use_force;
  const uint8_t* flags;  /* N bytes */
  const float* rho0;     /* N floats or NULL */
  const float* u0;       /* 2N floats or NULL (inlet velocity at inlet cells) */
} flb_instance;

typedef struct flb_batch flb_batch;

int flb_batch_create(const flb_instance* instances, uint32_t count, int32_t storage,
                     int32_t device, flb_batch** out);
/* Launch modes of flb_batch_step. */
enum { FLB_BATCH_LAUNCH = 0, FLB_BATCH_PERSISTENT = 1 };
/* Order of the packed tile list: instances interleaved row by row (default),
   instance after instance row-major, or instance after instance in bands of
   32 rows walked column by column. */
enum { FLB_BATCH_ORDER_ROWS = 0, FLB_BATCH_ORDER_INSTANCE = 1, FLB_BATCH_ORDER_BANDS = 2 };
int flb_batch_set_order(flb_batch* b, int32_t order);
/* FLB_BATCH_LAUNCH: threads per block (multiple of 64), blocks of the
   grid-stride launch (0 = one block per tile), steps per captured graph (even;
   0 = plain launches). */
int flb_batch_configure(flb_batch* b, int32_t block_x, int32_t grid, int32_t graph_chunk);
/* FLB_BATCH_PERSISTENT: one cooperative launch runs up to `chunk` steps, its
   blocks_per_cu resident blocks per CU walking the tile list and meeting at a
   grid barrier between steps; the last step of each flb_batch_step call runs
   as a plain launch and writes the readback fields.  Fails with
   FLB_ERR_ARGUMENT when that many blocks cannot be resident together. */
int flb_batch_configure_persistent(flb_batch* b, int32_t block_x, int32_t blocks_per_cu,
                                   int32_t chunk);
/* Times tile orders x block sizes x grid sizes of both modes over `steps`
   steps each (the state advances) and keeps the fastest (flb_batch_order
   reports the order).  best_grid: blocks of the launch
   (0 = one per tile), or -(blocks per CU) when the persistent mode won. */
int flb_batch_autotune(flb_batch* b, uint32_t steps, int32_t* best_block, int32_t* best_grid,
                       float* best_ms_per_step);
int flb_batch_step(flb_batch* b, uint64_t nsteps);
int flb_batch_sync(flb_batch* b);
int flb_batch_step_timed(flb_batch* b, uint64_t nsteps, float* ms);
/* Readback of one instance (fields of the last step; raw FP32 populations). */
int flb_batch_read(flb_batch* b, uint32_t instance, float* u, float* rho);
int flb_batch_read_populations(flb_batch* b, uint32_t instance, float* f);
uint64_t flb_batch_time(const flb_batch* b);
/* Current configuration: block size, grid (blocks of the launch, 0 = one per
   tile; -(blocks per CU) in the persistent mode), steps per graph (per
   cooperative launch in the persistent mode), tiles of the packed list. */
int flb_batch_info(const flb_batch* b, int32_t* block_x, int32_t* grid, int32_t* graph_chunk,
                   uint32_t* ntiles);
int32_t flb_batch_order(const flb_batch* b);
const char* flb_batch_last_error(const flb_batch* b);
void flb_batch_destroy(flb_batch* b);

/* ------------------------------------------------------------ output */
/* VTK ImageData writer, VTIWriter (src/io/writer.cpp:15-252).  ASCII output
   reproduces the reference's text layout; binary = 1 writes the same arrays
   as raw appended data (UInt64 block headers), for large grids. */
enum { FLB_VTI_UINT8 = 0, FLB_VTI_FLOAT32 = 1 };
typedef struct flb_vti_field {
  const char* name;
  int32_t type;          /* FLB_VTI_UINT8 (1 component) or FLB_VTI_FLOAT32 */
  uint32_t components;   /* 1, or 2 for [x | y] vector fields */
  const void* data;      /* N (or 2N) values */
  double scale;          /* multiplies every value (lattice -> SI) */
  int32_t flip_y;        /* negate the y component */
} flb_vti_field;
int flb_write_vti(const char* path, uint32_t nx, uint32_t ny, double origin_x, double origin_y,
                  double spacing, const flb_vti_field* fields, uint32_t nfields, int32_t binary);
/* loaderFiles.vti.series (Writer::createTimeSeriesFile, src/io/writer.cpp:159-181). */
int flb_write_vti_series(const char* path, uint32_t count, double time_interval);
/* Readback + write of Flags, U, rho (and phi for free surface) of a handle
   (the save step of h_runCuda*NoGraphics2D, src/cuda/runCuda.cu:483-487). */
int flb_save_vti(flb_handle* h, const char* path, double origin_x, double origin_y,
                 double spacing, double velocity_scale, double density_scale, int32_t binary);

/* ------------------------------------------------------------ host set-up */

/* Rectangle in SI units: centre (x0, y0), height (y), width (x)
   (RectangleShape, src/geometry/shapes.cpp:58-126). */
typedef struct flb_rect { double x0, y0, height, width; } flb_rect;

/* Flat-bed channel whose cross-drainage work spans the whole domain. */
typedef struct flb_channel {
  uint32_t nx, ny;
  double dx;                 /* SIZE */
  double x_min, y_min;
  int32_t free_surface;      /* default node GAS instead of FLUID */
  uint8_t bc_left, bc_right, bc_up, bc_down;  /* FLB_INLET/OUTLET/WALL/FLUID */
  uint32_t n_obstacles, n_initial;
  const flb_rect* obstacles; /* OBSTACLE(k) -> WALL */
  const flb_rect* initial;   /* INITIAL_CONDITION_FLAG(k) -> FLUID */
} flb_channel;

/* Flags of a channel, initData (src/initData.h:220-333): default node,
   boundary rows/columns (down, up, left, right priority), obstacles, initial
   fluid; for free surface followed by flb_seed_free_surface. */
int flb_build_channel_flags(const flb_channel* g, uint8_t* flags);

/* initFreeSurfaceFlags (src/initData.cpp:13-37) in index order. */
int flb_seed_free_surface(uint8_t* flags, uint32_t nx, uint32_t ny);

/* initFreeSurfaceFields (src/initData.h:175-206). excess may be NULL. */
int flb_init_free_surface_fields(const uint8_t* flags, size_t n, float* mass, float* phi,
                                 float* excess);

/* Synthetic culvert in lattice units: walls/inlet/outlet on the sides, an
   embankment of WALL cells over columns [emb_x0, emb_x1] above a culvert
   opening of `opening` cells at the bed, optional initial water block over
   columns [1, water_x1] and rows [1, water_depth].  rho0 = 1 + U(-noise,
   noise) drawn with std::mt19937_64(seed) in index order (rho0 may be NULL);
   u0 = (u_inlet, 0) at inlet cells, 0 elsewhere (u0 may be NULL). */
typedef struct flb_culvert {
  uint32_t nx, ny;
  int32_t free_surface;
  uint8_t bc_left, bc_right, bc_up, bc_down;
  uint32_t emb_x0, emb_x1, opening;
  uint32_t water_x1, water_depth;  /* water_x1 = 0: no water block */
  float u_inlet;
  double rho_noise;
  uint64_t seed;
} flb_culvert;
int flb_build_culvert(const flb_culvert* c, uint8_t* flags, float* rho0, float* u0);

/* Unit conversion, FLB::Units (src/utils.cpp:7-81) as used by
   h_initConstantDataCuda2D (src/cuda/cudaInitData.cu:57-84): lattice tau and
   body force from SI inputs. */
typedef struct flb_units_in {
  double length_si;          /* x extent of the domain (m) */
  uint32_t nx;               /* nodes along x */
  double ref_velocity_si, ref_velocity_lb, density_si;
  double relaxation_time;    /* used when != 0 */
  double kinematic_viscosity;
  int32_t use_force;
  double acceleration_x, acceleration_y;       /* precedence when != 0 */
  double volumetric_force_x, volumetric_force_y;
  double velocity_si_x, velocity_si_y;
} flb_units_in;
typedef struct flb_units_out {
  float tau, inv_tau, fx, fy;
  double velocity_lb_x, velocity_lb_y;
  double length_scale, time_scale, mass_scale;
} flb_units_out;
int flb_units(const flb_units_in* in, flb_units_out* out);

#ifdef __cplusplus
}
#endif
#endif /* FLB_CORE_H */
