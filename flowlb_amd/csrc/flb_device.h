// flb_device.h -- device-side building blocks shared by the FlowLB HIP kernels:
// D2Q9 lattice constants, the population storage codecs and the kernel
// parameter block.  gfx950 only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flb {

// D2Q9 velocities in the reference order (src/cuda/cudaInitData.cu:76-78);
// opposite pairs (1,2) (3,4) (5,6) (7,8).
__host__ __device__ constexpr int cx(int i) {
  return i == 1 || i == 5 || i == 7 ? 1 : (i == 2 || i == 6 || i == 8 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int i) {
  return i == 3 || i == 5 || i == 8 ? 1 : (i == 4 || i == 6 || i == 7 ? -1 : 0);
}
// Weights in float as the host builds them (src/ui/app.cpp:288).
__host__ __device__ constexpr float wgt(int i) {
  return i == 0 ? 4.0f / 9.0f : (i < 5 ? 1.0f / 9.0f : 1.0f / 36.0f);
}

enum : uint8_t {
  kFluid = 0x01, kGas = 0x02, kInterface = 0x04, kWall = 0x08, kMovingWall = 0x10,
  kInlet = 0x20, kOutlet = 0x40, kGasInterface = 0x80,
  kFGI = 0x07, kIF = 0x05, kIG = 0x06, kInletFluid = 0x21
};

// Weights in the arithmetic type R: float as above, double for the FP64 path
// (the host builds them in double there, src/ui/app.cpp:303).
template <class R>
__host__ __device__ constexpr R wgt_r(int i) {
  return i == 0 ? R(4) / R(9) : (i < 5 ? R(1) / R(9) : R(1) / R(36));
}

// ---------------------------------------------------------------- storage
// All formats hold the DDF-shifted population f_i - w_i.  T is the stored
// type, R the arithmetic type: FP32 for the 32- and 16-bit formats, FP64 for
// StoreF64 (the reference's PRECISION = 64, whose fields -- density,
// velocity, mass, excess, fill level -- are double as well).
struct StoreF32 {
  using T = float;
  using R = float;
  static constexpr int kBytes = 4;
  __device__ static float load(const T* p, size_t i) { return p[i]; }
  __device__ static void store(T* p, size_t i, float v) { p[i] = v; }
  __device__ static float round(float v) { return v; }  // value a store keeps
};

struct StoreF64 {
  using T = double;
  using R = double;
  static constexpr int kBytes = 8;
  __device__ static double load(const T* p, size_t i) { return p[i]; }
  __device__ static void store(T* p, size_t i, double v) { p[i] = v; }
  __device__ static double round(double v) { return v; }
};

// IEEE binary16, scaled by 2^15 so that the shifted populations (|f - w| < 1)
// sit in the normal range.
struct StoreF16S {
  using T = _Float16;
  using R = float;
  static constexpr int kBytes = 2;
  __device__ static float load(const T* p, size_t i) { return float(p[i]) * 3.0517578125e-05f; }
  __device__ static void store(T* p, size_t i, float v) { p[i] = _Float16(v * 32768.0f); }
  __device__ static float round(float v) {
    return float(_Float16(v * 32768.0f)) * 3.0517578125e-05f;
  }
};

// bfloat16 (upper half of an FP32), round to nearest even on store.
struct StoreBF16S {
  using T = uint16_t;
  using R = float;
  static constexpr int kBytes = 2;
  __device__ static float load(const T* p, size_t i) {
    return __uint_as_float(uint32_t(p[i]) << 16);
  }
  __device__ static void store(T* p, size_t i, float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    p[i] = uint16_t(u >> 16);
  }
  __device__ static float round(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
  }
};

// Tile of the fused free-surface passes 2-4: one block of kFsTileW x kFsTileH
// threads per tile, one wavefront per tile row.
constexpr int kFsTileW = 64, kFsTileH = 8;

// ---------------------------------------------------------------- params
struct Params {
  uint32_t nx, ny;
  uint32_t tiles_x;      // blocks per row
  uint32_t nblocks;      // total blocks of a full-grid launch
  uint64_t N;            // cells
  uint64_t Np;           // plane stride of the population array (elements)
  float inv_tau, fx, fy;
  int32_t use_force, collision, smagorinsky, gas_check_corrected, mass_mode;
  void* f;               // populations, 9 planes of Np elements of the storage type
  uint8_t* flags;        // current flags
  uint8_t* flags_next;   // flags written by the free-surface update (passes 2-4)
  // Cell fields.  They hold elements of the storage's arithmetic type R
  // (double for StoreF64); kernels access them through fld<R>().
  float* rho;            // readback density (N)
  float* u;              // readback velocity (2N)
  float* pub_rho[2];     // densities published for inlet cells, by parity
  float* pub_u[2];       // velocities published for outlet cells, by parity
  float* mass;
  float* excess;
  float* phi;
  float* rec;            // reconstructed gas populations, 9 planes of N
  const float* alpha;    // Brinkman solid fraction or nullptr
  float* trace;          // adjoint trace of one step: 3 planes of N (or nullptr)
};

// A cell field viewed in the arithmetic type R.
template <class R>
__device__ __forceinline__ R* fld(float* p) {
  return reinterpret_cast<R*>(p);
}

// Block -> tile mapping.  Workgroups are dispatched round-robin over the 8
// XCDs (block b to XCD b mod 8), each with its own L2.  Tiles are numbered row
// by row and taken in that order: with a row of a multiple of 8 tiles the tiles
// above and below a tile run on the same XCD, so the row-shifted loads meet in
// its L2, and every XCD gets an equal share of every band of rows.  (Giving
// each XCD a contiguous band of rows instead left the XCDs holding the gas
// rows of a free-surface domain idle while the others worked on the water.)
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t) { return b; }

}  // namespace flb
