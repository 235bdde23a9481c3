// flb_cell.h -- per-cell device code of the D2Q9 step, shared by the
// single-domain kernels (flb_kernels.hip) and the batched kernel
// (flb_batch.hip).  gfx950 only.
#pragma once

#include "flb_device.h"

namespace flb {

// Equilibrium minus weights (feq_i - w_i), calculateEquilibriumDDF2Q9
// (src/cuda/cellCalculations.cuh:32-52) rearranged for shifted storage:
// feq_i - w_i = w_i * ((rho - 1) + rho * A_i).
__device__ __forceinline__ void equilibrium_shifted(float rho, float rho1, float ux, float uy,
                                                    float* fe) {
  const float u2 = 3.0f * (ux * ux + uy * uy);
  const float ax = 3.0f * ux, ay = 3.0f * uy;
  const float a0 = ax - ay, a1 = ax + ay;
  const float h = 0.5f * u2;
  const float w0 = wgt(0), w1 = wgt(1), w5 = wgt(5);
  fe[0] = w0 * (rho1 - rho * h);
  fe[1] = w1 * (rho1 + rho * (ax - h + 0.5f * ax * ax));
  fe[2] = w1 * (rho1 + rho * (-ax - h + 0.5f * ax * ax));
  fe[3] = w1 * (rho1 + rho * (ay - h + 0.5f * ay * ay));
  fe[4] = w1 * (rho1 + rho * (-ay - h + 0.5f * ay * ay));
  fe[5] = w5 * (rho1 + rho * (a1 - h + 0.5f * a1 * a1));
  fe[6] = w5 * (rho1 + rho * (-a1 - h + 0.5f * a1 * a1));
  fe[7] = w5 * (rho1 + rho * (a0 - h + 0.5f * a0 * a0));
  fe[8] = w5 * (rho1 + rho * (-a0 - h + 0.5f * a0 * a0));
}

// Guo forcing terms, calculateForcingTerms2D (src/cuda/cellCalculations.cuh:64-73),
// with the reference's literal 0.33333334f.
__device__ __forceinline__ void guo_terms(float ux, float uy, float gx, float gy, float* fi) {
  const float uF = -0.33333334f * (ux * gx + uy * gy);
  fi[0] = 9.0f * wgt(0) * uF;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    const float cF = float(cx(i)) * gx + float(cy(i)) * gy;
    fi[i] = 9.0f * wgt(i) * (cF * (float(cx(i)) * ux + float(cy(i)) * uy + 0.33333334f) + uF);
  }
}

__device__ __forceinline__ float fill_level(float m, float rho) {  // calculatePhi
  return rho > 0.0f ? __saturatef(m / rho) : 0.5f;
}

struct Cell {
  uint32_t x, y, n;
  bool valid;
};

__device__ __forceinline__ Cell locate(const Params& p) {
  const uint32_t L = xcd_tile(blockIdx.x, p.nblocks);
  const uint32_t y = L / p.tiles_x;
  const uint32_t x = (L - y * p.tiles_x) * blockDim.x + threadIdx.x;
  return Cell{x, y, x + y * p.nx, x < p.nx};
}

// Periodic neighbour indices (src/cuda/cellCalculations.cuh:82-109) without
// integer division: the wrap is a compare at the domain edge.
__device__ __forceinline__ void neighbours(const Params& p, uint32_t x, uint32_t y, uint32_t* nb) {
  const uint32_t nx = p.nx;
  const uint32_t xr = x + 1 == nx ? 0u : x + 1;
  const uint32_t xl = x == 0 ? nx - 1 : x - 1;
  const uint32_t yc = y * nx;
  const uint32_t yu = (y + 1 == p.ny ? 0u : y + 1) * nx;
  const uint32_t yd = (y == 0 ? p.ny - 1 : y - 1) * nx;
  nb[0] = x + yc;
  nb[1] = xr + yc;
  nb[2] = xl + yc;
  nb[3] = x + yu;
  nb[4] = x + yd;
  nb[5] = xr + yu;
  nb[6] = xl + yd;
  nb[7] = xr + yd;
  nb[8] = xl + yu;
}

// Fluid/gas/interface bits of the eight neighbours, issued together so that
// they cost one memory round trip.
__device__ __forceinline__ void neighbour_flags(const Params& p, const uint32_t* nb, uint8_t* nf) {
#pragma unroll
  for (int i = 1; i < 9; i++) nf[i] = p.flags[nb[i]] & kFGI;
}

// A non-inlet, non-outlet fluid cell whose eight neighbours are fluid or
// interface (also while interface neighbours turn interface->fluid or ->gas
// inside stream-collide).  For such a cell the mass exchange of free-surface
// pass 1 over all links, sum_i (in_opp(i) - out_i), equals rho_in(t) -
// rho*(t-1): the density of the incoming populations minus the sum of the
// populations the cell emitted, and its mass before the pass is the density it
// had, so the updated mass is rho_in(t) plus the excess shares it pulls (up to
// the collision's rounding and Guo residual, ~1e-8 per step).
//
// In the conservative mode a fluid cell carries its mass between pass 1 and
// pass 4 as the deviation mass - rho: stream-collide writes the pulled shares
// of an interior cell, pass 1 the deviation of any other fluid cell, and pass 4
// turns the deviation into the excess.  Shares far below the rounding step of
// rho (~6e-8) are never rounded onto rho's grid, where they would be lost.
__device__ __forceinline__ bool interior_fluid(uint8_t fl, const uint8_t* nf) {
  if ((fl & (kFGI | kInlet | kOutlet)) != kFluid) return false;
  bool all = true;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    const uint8_t e = nf[i];
    all = all && (e == kFluid || e == kInterface || e == kIF || e == kIG);
  }
  return all;
}

// loadf (src/cuda/cellCalculations.cuh:185-196) with ODD = t % 2.
template <class S, int ODD>
__device__ __forceinline__ void load_incoming(const Params& p, const uint32_t* nb, float* g) {
  const typename S::T* f = static_cast<const typename S::T*>(p.f);
  const uint64_t Np = p.Np;
  g[0] = S::load(f, nb[0]);
#pragma unroll
  for (int i = 1; i < 9; i += 2) {
    g[i] = S::load(f + (ODD ? i : i + 1) * Np, nb[0]);
    g[i + 1] = S::load(f + (ODD ? i + 1 : i) * Np, nb[i]);
  }
}

// loadOutgoingf (src/cuda/cellCalculations.cuh:207-217).
template <class S, int ODD>
__device__ __forceinline__ void load_outgoing(const Params& p, const uint32_t* nb, float* o) {
  const typename S::T* f = static_cast<const typename S::T*>(p.f);
  const uint64_t Np = p.Np;
#pragma unroll
  for (int i = 1; i < 9; i += 2) {
    o[i] = S::load(f + (ODD ? i : i + 1) * Np, nb[i]);
    o[i + 1] = S::load(f + (ODD ? i + 1 : i) * Np, nb[0]);
  }
}

// storef (src/cuda/cellCalculations.cuh:228-238).
template <class S, int ODD>
__device__ __forceinline__ void store_outgoing(const Params& p, const uint32_t* nb, const float* g) {
  typename S::T* f = static_cast<typename S::T*>(p.f);
  const uint64_t Np = p.Np;
  S::store(f, nb[0], g[0]);
#pragma unroll
  for (int i = 1; i < 9; i += 2) {
    S::store(f + (ODD ? i + 1 : i) * Np, nb[i], g[i]);
    S::store(f + (ODD ? i : i + 1) * Np, nb[0], g[i + 1]);
  }
}

// ------------------------------------------------------------- stream-collide
// d_StreamCollide2D (src/cuda/lbmKernel2D.cu:305-519).  MOM: write the
// density/velocity readback fields (always done in free-surface runs, where
// the free-surface passes read them).  TRACE: record (rho, m_x, m_y) of every
// cell -- or, for inlet/outlet cells, the density and velocity they read -- in
// p.trace for the adjoint sweep.
template <class S, int ODD, bool FS, bool MOM, bool TRACE>
__device__ __forceinline__ void stream_collide_cell(const Params& p, const Cell c) {
  const uint32_t n = c.n;
  const uint64_t N = p.N;
  // Density of cell n is read next step by an inlet at n-1, velocity by an
  // outlet at n+1 (lbmKernel2D.cu:336,344-345); both roles are static.  The
  // three flags are loaded unconditionally (clamped index, same cache line) so
  // that they share one memory round trip instead of three branch-guarded ones.
  const uint8_t fl = p.flags[n];
  const uint8_t fl_prev = p.flags[n > 0 ? n - 1 : n];
  const uint8_t fl_next = p.flags[n + 1 < N ? n + 1 : n];
  const bool pub_r = n > 0 && (fl_prev & kInlet);
  const bool pub_v = n + 1 < N && (fl_next & kOutlet) && !(fl & kInlet);
  if (fl & (kGas | kWall)) {
    if (FS) {  // a source cell that is gas this step keeps its last value
      if (pub_r) p.pub_rho[ODD][n] = p.pub_rho[ODD ^ 1][n];
      if (pub_v) {
        p.pub_u[ODD][n] = p.pub_u[ODD ^ 1][n];
        p.pub_u[ODD][N + n] = p.pub_u[ODD ^ 1][N + n];
      }
    }
    return;
  }
  uint32_t nb[9];
  neighbours(p, c.x, c.y, nb);
  float g[9];
  load_incoming<S, ODD>(p, nb, g);
  // Free surface: every load that depends only on the own flag is issued here,
  // with the populations, so that the cell waits for memory once.
  uint8_t nf[9];
  float ex[9], rc[9], m_if = 0.0f;
  const bool cons = FS && p.mass_mode != 0;
  const bool fluid_dev = cons && (fl & (kFGI | kInlet | kOutlet)) == kFluid;
  const bool iface = FS && (fl & kFGI) == kInterface;
  if (FS) neighbour_flags(p, nb, nf);
  if (fluid_dev) {
#pragma unroll
    for (int i = 1; i < 9; i++) ex[i] = p.excess[nb[i]];
  }
  if (iface) {
    m_if = p.mass[n];
#pragma unroll
    for (int k = 1; k < 9; k++) rc[k] = p.rec[k * N + n];
    // Reconstructed gas populations of free-surface pass 1 replace incoming
    // ones (storeReconstructedGasf, cellCalculations.cuh:250-259).
#pragma unroll
    for (int k = 1; k < 9; k++) {
      const int opp = (k & 1) ? k + 1 : k - 1;
      const uint8_t gate = p.gas_check_corrected ? nf[opp] : nf[k];
      if (gate == kGas) g[k] = rc[k];
    }
  }
  const uint8_t io = fl & (kInlet | kOutlet);
  float rho, rho1, ux, uy;
  if (io & kInlet) {
    rho = p.pub_rho[ODD ^ 1][n + 1];
    rho1 = rho - 1.0f;
    ux = p.u[n];
    uy = p.u[N + n];
    if (TRACE) {
      p.trace[n] = rho;
      p.trace[N + n] = ux;
      p.trace[2 * N + n] = uy;
    }
  } else if (fl & kOutlet) {
    rho = p.rho[n];
    rho1 = rho - 1.0f;
    ux = p.pub_u[ODD ^ 1][n - 1];
    uy = p.pub_u[ODD ^ 1][N + n - 1];
    if (TRACE) {
      p.trace[n] = rho;
      p.trace[N + n] = ux;
      p.trace[2 * N + n] = uy;
    }
  } else {
    rho1 = g[0];
#pragma unroll
    for (int i = 1; i < 9; i++) rho1 += g[i];
    rho = 1.0f + rho1;
    const float mx = g[1] - g[2] + g[5] - g[6] + g[7] - g[8];
    const float my = g[3] - g[4] + g[5] - g[6] + g[8] - g[7];
    ux = mx / rho;
    uy = my / rho;
    if (TRACE) {
      p.trace[n] = rho;
      p.trace[N + n] = mx;
      p.trace[2 * N + n] = my;
    }
  }
  if (fluid_dev && interior_fluid(fl, nf)) {  // deviation = pulled shares
    double m = 0.0;
#pragma unroll
    for (int i = 1; i < 9; i++) m += double(ex[i]);
    p.mass[n] = float(m);
  }
  if (iface) {  // interface -> fluid / gas test (:353-367)
    bool not_fluid = true, not_gas = true;
#pragma unroll
    for (int i = 1; i < 9; i++) {
      not_fluid = not_fluid && (nf[i] != kFluid);
      not_gas = not_gas && (nf[i] != kGas);
    }
    if ((m_if > 1.001f * rho) || not_gas) p.flags[n] = (fl & uint8_t(~kFGI)) | kIF;
    else if ((m_if < -0.001f) || not_fluid) p.flags[n] = (fl & uint8_t(~kFGI)) | kIG;
  }
  float fi[9];
  bool forced = false;
  if (p.alpha && !io) {
    // Brinkman drag G = F - alpha rho u with the Guo half-force shift solved
    // for u (no reference counterpart).
    const float a = p.alpha[n];
    const float gx0 = p.use_force ? p.fx : 0.0f, gy0 = p.use_force ? p.fy : 0.0f;
    const float rhol2 = 0.5f / rho;
    const float s = 1.0f / (1.0f + 0.5f * a);
    ux = (ux + rhol2 * gx0) * s;
    uy = (uy + rhol2 * gy0) * s;
    guo_terms(ux, uy, gx0 - a * rho * ux, gy0 - a * rho * uy, fi);
    forced = true;
  } else if (p.use_force) {
    const float rhol2 = 0.5f / rho;
    ux = ux + rhol2 * p.fx;
    uy = uy + rhol2 * p.fy;
    guo_terms(ux, uy, p.fx, p.fy, fi);
    forced = true;
  }
  if (MOM || FS) {
    if (!io) {
      p.rho[n] = rho;
      p.u[n] = ux;
      p.u[N + n] = uy;
    } else if (fl & kOutlet) {
      p.u[n] = ux;
      p.u[N + n] = uy;
    }
    p.rho[n] = rho;
  }
  if (pub_r) p.pub_rho[ODD][n] = rho;
  if (pub_v) {
    p.pub_u[ODD][n] = ux;
    p.pub_u[ODD][N + n] = uy;
  }
  float fe[9];
  equilibrium_shifted(rho, rho1, ux, uy, fe);
  float inv_tau = p.inv_tau;
  if (p.smagorinsky) {  // (:417-442)
    float hxx = 0.0f, hyy = 0.0f, hxy = 0.0f;
#pragma unroll
    for (int i = 1; i < 9; i++) {
      const float d = g[i] - fe[i];
      hxx += float(cx(i) * cx(i)) * d;
      hyy += float(cy(i) * cy(i)) * d;
      hxy += float(cx(i) * cy(i)) * d;
    }
    const float Q = hxx * hxx + hyy * hyy + 2.0f * hxy * hxy;
    const float tau = 1.0f / inv_tau;
    inv_tau = 2.0f / (tau + sqrtf(tau * tau + 0.54037965f * sqrtf(Q) / rho));
  }
  if (io) {
#pragma unroll
    for (int i = 0; i < 9; i++) g[i] = fe[i];
  } else if (p.collision == 0) {  // SRT
    const float omc = 1.0f - inv_tau;
    if (forced) {
      const float tl = 1.0f - 0.5f * inv_tau;
#pragma unroll
      for (int i = 0; i < 9; i++) g[i] = omc * g[i] + inv_tau * fe[i] + tl * fi[i];
    } else {
#pragma unroll
      for (int i = 0; i < 9; i++) g[i] = omc * g[i] + inv_tau * fe[i];
    }
  } else if (!forced) {  // TRT, magic parameter 1/4; the forced variant only streams
    const float ip = inv_tau;
    const float in = 1.0f / (0.25f / (1.0f / inv_tau - 0.5f) + 0.5f);
    g[0] = g[0] - ip * (g[0] - fe[0]);
#pragma unroll
    for (int i = 1; i < 9; i += 2) {
      const float sym = g[i] - fe[i] + g[i + 1] - fe[i + 1];
      const float fa = g[i] - g[i + 1];
      const float fea = fe[i] - fe[i + 1];
      g[i] = g[i] - 0.5f * ip * sym - 0.5f * in * (fa - fea);
      g[i + 1] = g[i + 1] - 0.5f * ip * sym - 0.5f * in * (fea - fa);
    }
  }
  store_outgoing<S, ODD>(p, nb, g);
}

}  // namespace flb
