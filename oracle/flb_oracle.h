/*
 * flb_oracle.h -- C API of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP core in
 * flowlb_amd/: it restates the reference FlowLB D2Q9 kernels
 * (src/cuda/lbmKernel2D.cu, src/cuda/cellCalculations.cuh) and the reference
 * host set-up (src/initData.h, src/initData.cpp) as serial CPU loops.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load
 * it.  The product never links, loads or calls it.
 *
 * Every function is available in two arithmetic types: the `orc_*` entry
 * points compute in FP64 (the parity reference), the `orcf_*` entry points in
 * FP32 (used for the CPU baseline timing with OpenMP threads).
 */
#ifndef FLB_ORACLE_H
#define FLB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cell flags, same bit values as FLB::TypesNodes (src/utils.h:91-108). */
enum {
  ORC_FLUID = 0x01, ORC_GAS = 0x02, ORC_INTERFACE = 0x04, ORC_WALL = 0x08,
  ORC_MWALL = 0x10, ORC_INLET = 0x20, ORC_OUTLET = 0x40, ORC_GAS_INTERFACE = 0x80
};

typedef struct orc_params {
  uint32_t nx, ny;
  float inv_tau;          /* 1/tau kept as float (src/cuda/cudaInitData.cu:82) */
  float fx, fy;           /* body force, float (src/cuda/cudaInitData.cu:66-72) */
  int32_t use_force;      /* d_useVolumetricForce */
  int32_t collision;      /* 0 = SRT, 1 = TRT, 2 = TRT with Guo forcing */
  int32_t smagorinsky;    /* d_useSubgridModel */
  int32_t free_surface;   /* TYPE_PROBLEM = FREE_SURFACE */
  int32_t gas_check_corrected; /* 0: neighbour test of storeReconstructedGasf as written */
  int32_t mass_mode;      /* 0: excess-mass receivers as written, 1: conservative */
  int32_t threads;        /* OpenMP threads for the data-parallel loops (<=0: 1) */
} orc_params;

/* Rectangle in SI coordinates (RectangleShape, src/geometry/shapes.cpp:58-126). */
typedef struct orc_rect { double x0, y0, height, width; } orc_rect;

typedef struct orc_channel {
  uint32_t nx, ny;
  double dx;              /* SIZE: spacing in metres */
  double x_min, y_min;
  int32_t free_surface;   /* default node GAS instead of FLUID */
  uint8_t bc_left, bc_right, bc_up, bc_down;  /* INLET / OUTLET / WALL / FLUID */
  uint32_t n_obstacles, n_initial;
  const orc_rect* obstacles;  /* -> WALL */
  const orc_rect* initial;    /* -> FLUID */
} orc_channel;

/* Set-up (src/initData.h:220-333, src/initData.cpp:13-151). */
int orc_build_channel_flags(const orc_channel* g, uint8_t* flags);
void orc_seed_free_surface(uint8_t* flags, uint32_t nx, uint32_t ny);
void orc_init_free_surface_fields(const uint8_t* flags, size_t n, double* mass, double* phi,
                                  double* excess);

/* Solver state.  rho0 and u0 (u0 = [ux | uy], 2N) may be NULL (rho = 1, u = 0
   except inlet cells, which take u_inlet).  alpha (Brinkman solid fraction) may
   be NULL. */
void* orc_create(const orc_params* p, const uint8_t* flags, const double* rho0, const double* u0,
                 const double* alpha);
void orc_set_fs_fields(void* s, const double* mass, const double* phi, const double* excess);
void orc_step(void* s, uint64_t nsteps);
uint64_t orc_time(void* s);
/* Any pointer may be NULL.  f is the raw slot array f[q*N + n]. */
void orc_get(void* s, double* f, double* rho, double* u, uint8_t* flags, double* mass,
             double* excess, double* phi);
void orc_set_f(void* s, const double* f);
double orc_total_mass(void* s);
void orc_destroy(void* s);

/* Free-surface discharge through a column: sum of phi * rho * u_x over the
   cells with a fluid or interface bit. */
double orc_fs_discharge(void* s, uint32_t column);

/* Single passes on a hand-built state, for the tests that pin each kernel.
   orc_set_cells overwrites flags / rho / u (each pointer may be NULL);
   orc_set_time sets the lattice time that selects the slot parity.
   orc_run_pass runs one kernel without advancing time: 1 d_FreeSurface2D_1,
   2 d_StreamCollide2D, 3 d_FreeSurface2D_2, 4 d_FreeSurface2D_3,
   5 d_FreeSurface2D_4; it returns nonzero for an unknown pass.
   orc_get_reconstructed reads the populations the last pass 1 reconstructed
   (rec[k*N + n]) and, per cell, the bit mask of the directions k it replaced. */
void orc_set_cells(void* s, const uint8_t* flags, const double* rho, const double* u);
void orc_set_time(void* s, uint64_t t);
int orc_run_pass(void* s, int pass);
void orc_get_reconstructed(void* s, double* rec, uint16_t* mask);

/* Discharge through a column: sum over non-wall cells of rho*ux using the
   moments of the last step. */
double orc_discharge(void* s, uint32_t column);

/* Single-phase adjoint.  Objective J = mean over t in [t0, t1] of the
   discharge through `column` computed from the in-step moments.  Forward run
   of nsteps from the state at creation, reverse sweep with checkpoints every
   `ckpt` steps.  Outputs dJ/dalpha (N values, NULL allowed), dJ/du_inlet (x
   component of the inlet velocity) and J. */
int orc_adjoint(void* s, uint64_t nsteps, uint32_t column, uint64_t t0, uint64_t t1,
                uint32_t ckpt, double* dj_dalpha, double* dj_duin, double* j_out);
/* Forward objective only (for finite differences). */
double orc_objective(void* s, uint64_t nsteps, uint32_t column, uint64_t t0, uint64_t t1);

/* FP32 variants (CPU baseline). */
void* orcf_create(const orc_params* p, const uint8_t* flags, const double* rho0, const double* u0,
                  const double* alpha);
void orcf_set_fs_fields(void* s, const double* mass, const double* phi, const double* excess);
void orcf_step(void* s, uint64_t nsteps);
void orcf_get(void* s, double* f, double* rho, double* u, uint8_t* flags, double* mass,
              double* excess, double* phi);
double orcf_total_mass(void* s);
void orcf_destroy(void* s);

#ifdef __cplusplus
}
#endif
#endif
