/* flb_deck.h -- FlowLB case decks (Geometry.txt, OptionsCalc.txt and the CSV
   outlines they name) read on the host and turned into the cell flags of the
   domain, as FlowLB's "Create Domain" and "Calculate" actions do
   (src/ui/app.cpp:203-239, 241-300).  Plain C ABI; link libflbcore.so. */
#ifndef FLB_DECK_H
#define FLB_DECK_H

#include <stddef.h>
#include <stdint.h>

#include "flb_core.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Which options of OptionsCalc.txt are read.
   FLB_DECK_DOMAIN: TYPE_PROBLEM and the four boundaries only, unknown keys
     ignored (CalculationReader::readSomeOptionsCalculation,
     src/io/reader.cpp:738-768, as App::createDomain calls it).
   FLB_DECK_CALCULATION: every option, unknown keys rejected, values checked
     (CalculationReader::readOptionsCalculation, reader.cpp:710-736, and
     checkOptionsCalculation, reader.cpp:138-265). */
enum { FLB_DECK_DOMAIN = 0, FLB_DECK_CALCULATION = 1 };

typedef struct flb_deck flb_deck;

typedef struct flb_deck_info {
  /* mesh (Mesh::calculateNumberPointsMesh, src/geometry/mesh.cpp:45-52) */
  uint32_t nx, ny;
  double x_min, x_max, y_min, y_max, size;
  uint32_t n_points, n_cdws, n_obstacles, n_initial;
  /* options (OptionsCalculation, src/io/reader.h:98-127) */
  int32_t problem;                    /* FLB_MONOFLUID / FLB_FREE_SURFACE */
  uint8_t bc_left, bc_right, bc_up, bc_down; /* FLB_INLET/OUTLET/WALL/FLUID(periodic) */
  int32_t collision;                  /* 0 SRT, 1 TRT */
  int32_t smagorinsky, surface_tension, use_force, precision;
  int32_t block_x, block_y, block_z;  /* CUDA_BLOCK_SIZE */
  int32_t graphics;                   /* 0 NONE, 1 OPENGL */
  double ref_velocity_lb, ref_velocity_si, flow, relaxation_time;
  double kinematic_viscosity, density;
  double velocity_si[3], acceleration[3], volumetric_force[3];
  float time_simulation, time_save;
} flb_deck_info;

/* Reads <dir>/Geometry.txt (GeometryReader::readGeometryOptions,
   src/io/reader.cpp:1180-1217, with getMeshElements/getCDWData/
   getObstaclesData/modifyGeometryData, :785-1178), builds the mesh
   (Mesh::init, src/geometry/mesh.cpp:11-18) and reads <dir>/OptionsCalc.txt in
   the given mode.  CSV outlines are read from dir.  On failure *out is NULL,
   the status is returned and err (if non-NULL) receives the message, with the
   line number where the reference's reader reports one.  Only 2D decks. */
int flb_deck_load(const char* dir, int mode, flb_deck** out, char* err, size_t err_len);

int flb_deck_get_info(const flb_deck* d, flb_deck_info* info);

/* Cell flags, nx*ny bytes in index order x + y*nx: the flag loop of initData
   (src/initData.h:220-333) followed, for free-surface decks, by
   initFreeSurfaceFlags (src/initData.cpp:13-37). */
int flb_deck_build_flags(const flb_deck* d, uint8_t* flags);

/* The unit-conversion inputs the deck holds (h_initConstantDataCuda2D,
   src/cuda/cudaInitData.cu:57-82): length = x_max - x_min over nx - 1 cells. */
int flb_deck_units_in(const flb_deck* d, flb_units_in* in);

void flb_deck_free(flb_deck* d);

#ifdef __cplusplus
}
#endif

#endif /* FLB_DECK_H */
