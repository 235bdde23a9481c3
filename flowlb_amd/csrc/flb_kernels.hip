// flb_kernels.hip -- hand-written gfx950 kernels of the FlowLB D2Q9 hot path.
//
// One thread per cell, 64-lane wavefronts along x (a block is one row segment of
// blockDim.x cells), populations in structure-of-arrays planes so that every
// wavefront load/store of a population plane is one contiguous run.  The
// esoteric-pull slot algebra (src/cuda/cellCalculations.cuh:185-259) makes the
// update in place: at each parity every cell reads and writes the same nine
// slots and no two cells share a slot.
//
// Cross-cell reads inside a kernel only touch data no thread of that kernel
// writes (or data whose tested bits no thread changes), so results do not
// depend on the order in which wavefronts run:
//   * inlet/outlet cells read the neighbour density/velocity of the previous
//     step from a parity-indexed side buffer (pub_rho/pub_u);
//   * the free-surface pass 1 writes reconstructed gas populations to a side
//     array that stream-collide merges when it loads the cell;
//   * the gas->interface and interface->fluid neighbour rules are evaluated by
//     the cell that changes (gather form), never written into a neighbour;
//   * pass 4 reads the flags of the previous pass and writes a second buffer.
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
// the free-surface passes read them).
template <class S, int ODD, bool FS, bool MOM>
__global__ void __launch_bounds__(1024) k_stream_collide(const Params p) {
  const Cell c = locate(p);
  if (!c.valid) return;
  const uint32_t n = c.n;
  const uint64_t N = p.N;
  const uint8_t fl = p.flags[n];
  // Density of cell n is read next step by an inlet at n-1, velocity by an
  // outlet at n+1 (lbmKernel2D.cu:336,344-345); both roles are static.
  const bool pub_r = n > 0 && (p.flags[n - 1] & kInlet);
  const bool pub_v = n + 1 < N && (p.flags[n + 1] & kOutlet) && !(fl & kInlet);
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
  if (FS && (fl & kFGI) == kInterface) {
    // Reconstructed gas populations of free-surface pass 1 replace incoming
    // ones (storeReconstructedGasf, cellCalculations.cuh:250-259).
#pragma unroll
    for (int k = 1; k < 9; k++) {
      const int opp = (k & 1) ? k + 1 : k - 1;
      const uint32_t gate = p.gas_check_corrected ? nb[opp] : nb[k];
      if ((p.flags[gate] & kFGI) == kGas) g[k] = p.rec[k * N + n];
    }
  }
  const uint8_t io = fl & (kInlet | kOutlet);
  float rho, rho1, ux, uy;
  if (io & kInlet) {
    rho = p.pub_rho[ODD ^ 1][n + 1];
    rho1 = rho - 1.0f;
    ux = p.u[n];
    uy = p.u[N + n];
  } else if (fl & kOutlet) {
    rho = p.rho[n];
    rho1 = rho - 1.0f;
    ux = p.pub_u[ODD ^ 1][n - 1];
    uy = p.pub_u[ODD ^ 1][N + n - 1];
  } else {
    rho1 = g[0];
#pragma unroll
    for (int i = 1; i < 9; i++) rho1 += g[i];
    rho = 1.0f + rho1;
    ux = (g[1] - g[2] + g[5] - g[6] + g[7] - g[8]) / rho;
    uy = (g[3] - g[4] + g[5] - g[6] + g[8] - g[7]) / rho;
  }
  if (FS && (fl & kInterface) == kInterface) {  // interface -> fluid / gas test (:353-367)
    const float m = p.mass[n];
    bool not_fluid = true, not_gas = true;
#pragma unroll
    for (int i = 1; i < 9; i++) {
      const uint8_t e = p.flags[nb[i]] & kFGI;
      not_fluid = not_fluid && (e != kFluid);
      not_gas = not_gas && (e != kGas);
    }
    if ((m > 1.001f * rho) || not_gas) p.flags[n] = (fl & uint8_t(~kFGI)) | kIF;
    else if ((m < -0.001f) || not_fluid) p.flags[n] = (fl & uint8_t(~kFGI)) | kIG;
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

// ------------------------------------------------------------- free surface
// d_FreeSurface2D_1 (src/cuda/lbmKernel2D.cu:13-121): mass exchange and gas
// reconstruction for interface cells.
template <class S, int ODD>
__global__ void __launch_bounds__(1024) k_fs1(const Params p) {
  const Cell c = locate(p);
  if (!c.valid) return;
  const uint32_t n = c.n;
  const uint64_t N = p.N;
  const uint8_t fl = p.flags[n];
  if (fl & (kGas | kWall | kInlet)) return;
  uint32_t nb[9];
  neighbours(p, c.x, c.y, nb);
  float in[9], out[9];
  load_incoming<S, ODD>(p, nb, in);
  out[0] = in[0];
  load_outgoing<S, ODD>(p, nb, out);
  float m = p.mass[n];
#pragma unroll
  for (int i = 1; i < 9; i++) m += p.excess[nb[i]];
  const uint8_t e = fl & kFGI;
  if (e == kFluid) {
    if (p.mass_mode == 0) {
#pragma unroll
      for (int i = 1; i < 9; i += 2) {
        m += in[i + 1] - out[i];
        m += in[i] - out[i + 1];
      }
    } else {
#pragma unroll
      for (int i = 1; i < 9; i += 2) {
        const uint8_t a = p.flags[nb[i]] & kFGI, b = p.flags[nb[i + 1]] & kFGI;
        if (a == kFluid || a == kInterface) m += in[i + 1] - out[i];
        if (b == kFluid || b == kInterface) m += in[i] - out[i + 1];
      }
    }
  } else if (e == kInterface) {
    float rho, ux, uy;
    if (fl & (kInlet | kOutlet)) {
      rho = p.rho[n];
      ux = p.u[n];
      uy = p.u[N + n];
    } else {
      float s = out[0];
#pragma unroll
      for (int i = 1; i < 9; i++) s += out[i];
      rho = 1.0f + s;
      ux = (out[1] - out[2] + out[5] - out[6] + out[7] - out[8]) / rho;
      uy = (out[3] - out[4] + out[5] - out[6] + out[8] - out[7]) / rho;
    }
    const float phi0 = p.mass_mode == 0 ? fill_level(m, rho) : p.phi[n];
    const float rhol2 = 0.5f / rho;
    ux = ux + rhol2 * p.fx;
    uy = uy + rhol2 * p.fy;
    float fe[9];
    equilibrium_shifted(1.0f, 0.0f, ux, uy, fe);  // atmospheric equilibrium, rho = 1
    uint8_t nf[9];
#pragma unroll
    for (int i = 1; i < 9; i++) nf[i] = p.flags[nb[i]] & kFGI;
#pragma unroll
    for (int i = 1; i < 9; i += 2) {
      if (nf[i] == kFluid) m += in[i + 1] - out[i];
      else if (nf[i] == kInterface) m += (in[i + 1] - out[i]) * 0.5f * (p.phi[nb[i]] + phi0);
      if (nf[i + 1] == kFluid) m += in[i] - out[i + 1];
      else if (nf[i + 1] == kInterface) m += (in[i] - out[i + 1]) * 0.5f * (p.phi[nb[i + 1]] + phi0);
    }
#pragma unroll
    for (int k = 1; k < 9; k++) {
      const int opp = (k & 1) ? k + 1 : k - 1;
      const uint8_t gate = p.gas_check_corrected ? nf[opp] : nf[k];
      if (gate == kGas) p.rec[k * N + n] = fe[k] - out[opp] + fe[opp];
    }
  }
  p.mass[n] = m;
}

// d_FreeSurface2D_2 (src/cuda/lbmKernel2D.cu:123-148) in gather form: a gas or
// interface->gas cell next to an interface->fluid cell changes itself.
__global__ void __launch_bounds__(1024) k_fs2(const Params p) {
  const Cell c = locate(p);
  if (!c.valid) return;
  const uint8_t fl = p.flags[c.n];
  const uint8_t e = fl & kIG;
  if (e != kIG && e != kGas) return;
  uint32_t nb[9];
  neighbours(p, c.x, c.y, nb);
  bool near_if = false;
#pragma unroll
  for (int i = 1; i < 9; i++) near_if = near_if || ((p.flags[nb[i]] & kIF) == kIF);
  if (!near_if) return;
  p.flags[c.n] = e == kIG ? uint8_t(fl & ~kGas) : uint8_t((fl & ~kGas) | kGasInterface);
}

// d_FreeSurface2D_3 (src/cuda/lbmKernel2D.cu:150-190): gas->interface cells
// take feq of the neighbour averages; fluid cells next to an interface->gas
// cell become interface (gather form).
template <class S, int ODD>
__global__ void __launch_bounds__(1024) k_fs3(const Params p) {
  const Cell c = locate(p);
  if (!c.valid) return;
  const uint32_t n = c.n;
  const uint64_t N = p.N;
  const uint8_t fl = p.flags[n];
  const bool gi = (fl & (kGasInterface | kIG)) == kGasInterface;
  const uint8_t fi = fl & (kInletFluid | kInterface);
  const bool cand = fi == kFluid || fi == kIF;
  if (!gi && !cand) return;
  uint32_t nb[9];
  neighbours(p, c.x, c.y, nb);
  if (gi) {
    // calculateAverageValuesNeighbors2D (cellCalculations.cuh:124-145)
    float cnt = 0.0f, sr = 0.0f, sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int i = 1; i < 9; i++) {
      if (p.flags[nb[i]] & kIF) {
        cnt += 1.0f;
        sr += p.rho[nb[i]];
        sx += p.u[nb[i]];
        sy += p.u[N + nb[i]];
      }
    }
    float rho = cnt > 0.0f ? sr / cnt : 0.0f;
    float ux = cnt > 0.0f ? sx / cnt : 0.0f;
    float uy = cnt > 0.0f ? sy / cnt : 0.0f;
    const float rhol2 = 0.5f / rho;
    ux = ux + rhol2 * p.fx;
    uy = uy + rhol2 * p.fy;
    float fe[9];
    equilibrium_shifted(rho, rho - 1.0f, ux, uy, fe);
    store_outgoing<S, ODD>(p, nb, fe);
  } else {
    bool near_ig = false;
#pragma unroll
    for (int i = 1; i < 9; i++)
      near_ig = near_ig || ((p.flags[nb[i]] & (kGasInterface | kIG)) == kIG);
    if (near_ig) p.flags[n] = uint8_t((fl & ~kIF) | kInterface);
  }
}

// Flag a cell carries after its own pass-4 transition.
__device__ __forceinline__ uint8_t settle(uint8_t fl) {
  if (fl & (kWall | kMovingWall | kInlet)) return fl;
  const uint8_t e = fl & (kFGI | kGasInterface);
  if (e == kIF || e == kIG) return fl & uint8_t(~kInterface);
  if (e == kGasInterface) return uint8_t((fl & ~kGasInterface) | kInterface);
  return fl;
}

// d_FreeSurface2D_4 (src/cuda/lbmKernel2D.cu:192-303): flag transitions,
// mass clamping, excess mass split over the receiving neighbours.
__global__ void __launch_bounds__(1024) k_fs4(const Params p) {
  const Cell c = locate(p);
  if (!c.valid) return;
  const uint32_t n = c.n;
  const uint8_t fl = p.flags[n];
  if (fl & (kWall | kMovingWall | kInlet)) {
    p.flags_next[n] = fl;
    return;
  }
  const uint8_t e = fl & (kFGI | kGasInterface);
  float m = p.mass[n];
  if (e == kGas && m == 0.0f) {
    // Empty gas cell: only a share left from the previous step needs clearing.
    p.flags_next[n] = fl;
    if (p.excess[n] != 0.0f) p.excess[n] = 0.0f;
    return;
  }
  const float rho = p.rho[n];
  float ex = 0.0f, ph = 0.0f;
  uint8_t nfl = fl;
  switch (e) {
    case kFluid:
      ex = m - rho;
      m = rho;
      ph = 1.0f;
      break;
    case kGas:
      ex = m;
      m = 0.0f;
      break;
    case kIF:
      nfl = fl & uint8_t(~kInterface);
      ex = m - rho;
      m = rho;
      ph = 1.0f;
      break;
    case kIG:
      nfl = fl & uint8_t(~kInterface);
      ex = m;
      m = 0.0f;
      break;
    case kInterface:
    case kGasInterface:
      if (e == kGasInterface) nfl = uint8_t((fl & ~kGasInterface) | kInterface);
      if (m > rho) {
        ex = m - rho;
        m = rho;
      } else if (m < 0.0f) {
        ex = m;
        m = 0.0f;
      }
      ph = fill_level(m, rho);
      break;
    default:
      break;
  }
  uint32_t nb[9];
  neighbours(p, c.x, c.y, nb);
  int cont = 0;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    const uint8_t q = p.flags[nb[i]];
    if (p.mass_mode == 0) {
      const uint8_t r = q & (kIF | kGasInterface);
      cont += int(r == kFluid || r == kInterface || r == kIF || r == kGasInterface);
    } else {
      cont += int((settle(q) & (kGas | kWall | kInlet)) == 0);
    }
  }
  if (cont > 0) ex = ex / float(cont);
  else {
    m += ex;
    ex = 0.0f;
  }
  p.mass[n] = m;
  p.excess[n] = ex;
  p.phi[n] = ph;
  p.flags_next[n] = nfl;
}

// ------------------------------------------------------------- utilities
__device__ __forceinline__ double block_sum(double v, double* lds) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < int(blockDim.x >> 6); w++) s += lds[w];
  return s;
}

// Free-surface mass: stored mass + excess shares pulled next step, FP64.
__global__ void __launch_bounds__(256) k_fs_mass(const Params p, double* partial) {
  __shared__ double lds[16];
  const Cell c = locate(p);
  double v = 0.0;
  if (c.valid) {
    v = p.mass[c.n];
    if (!(p.flags[c.n] & (kGas | kWall | kInlet))) {
      uint32_t nb[9];
      neighbours(p, c.x, c.y, nb);
      for (int i = 1; i < 9; i++) v += double(p.excess[nb[i]]);
    }
  }
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Monofluid mass: sum of all populations (shifted values summed, + N).
template <class S>
__global__ void __launch_bounds__(256) k_pop_mass(const Params p, double* partial) {
  __shared__ double lds[16];
  const Cell c = locate(p);
  double v = 0.0;
  if (c.valid) {
    const typename S::T* f = static_cast<const typename S::T*>(p.f);
    for (int q = 0; q < 9; q++) v += double(S::load(f + q * p.Np, c.n));
  }
  const double s = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <class S>
__global__ void k_pop_to_f32(const Params p, float* out) {
  const Cell c = locate(p);
  if (!c.valid) return;
  const typename S::T* f = static_cast<const typename S::T*>(p.f);
  for (int q = 0; q < 9; q++) out[q * p.N + c.n] = S::load(f + q * p.Np, c.n) + wgt(q);
}

// in: shifted populations (f - w), FP32, plane stride N.
template <class S>
__global__ void k_f32_to_pop(const Params p, const float* in) {
  const Cell c = locate(p);
  if (!c.valid) return;
  typename S::T* f = static_cast<typename S::T*>(p.f);
  for (int q = 0; q < 9; q++) S::store(f + q * p.Np, c.n, in[q * p.N + c.n]);
}

__global__ void k_flag_hist(const Params p, unsigned long long* counts) {
  __shared__ unsigned int h[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const Cell c = locate(p);
  if (c.valid) atomicAdd(&h[p.flags[c.n]], 1u);
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
}

// ------------------------------------------------------------- launch table
using KernelFn = void (*)(const Params);

template <class S>
KernelFn stream_collide_fn(int odd, bool fs, bool mom) {
  if (fs) return odd ? k_stream_collide<S, 1, true, false> : k_stream_collide<S, 0, true, false>;
  if (mom) return odd ? k_stream_collide<S, 1, false, true> : k_stream_collide<S, 0, false, true>;
  return odd ? k_stream_collide<S, 1, false, false> : k_stream_collide<S, 0, false, false>;
}
template <class S>
KernelFn fs1_fn(int odd) { return odd ? k_fs1<S, 1> : k_fs1<S, 0>; }
template <class S>
KernelFn fs3_fn(int odd) { return odd ? k_fs3<S, 1> : k_fs3<S, 0>; }

}  // namespace flb

// Host-side accessors used by flb_core.hip (same translation unit group).
namespace flb {
KernelFn get_stream_collide(int storage, int odd, bool fs, bool mom) {
  switch (storage) {
    case 1: return stream_collide_fn<StoreF16S>(odd, fs, mom);
    case 2: return stream_collide_fn<StoreBF16S>(odd, fs, mom);
    default: return stream_collide_fn<StoreF32>(odd, fs, mom);
  }
}
KernelFn get_fs1(int storage, int odd) {
  switch (storage) {
    case 1: return fs1_fn<StoreF16S>(odd);
    case 2: return fs1_fn<StoreBF16S>(odd);
    default: return fs1_fn<StoreF32>(odd);
  }
}
KernelFn get_fs3(int storage, int odd) {
  switch (storage) {
    case 1: return fs3_fn<StoreF16S>(odd);
    case 2: return fs3_fn<StoreBF16S>(odd);
    default: return fs3_fn<StoreF32>(odd);
  }
}
KernelFn get_fs2() { return k_fs2; }
KernelFn get_fs4() { return k_fs4; }

using MassFn = void (*)(const Params, double*);
MassFn get_pop_mass(int storage) {
  switch (storage) {
    case 1: return k_pop_mass<StoreF16S>;
    case 2: return k_pop_mass<StoreBF16S>;
    default: return k_pop_mass<StoreF32>;
  }
}
MassFn get_fs_mass() { return k_fs_mass; }
using ConvOutFn = void (*)(const Params, float*);
using ConvInFn = void (*)(const Params, const float*);
ConvOutFn get_pop_to_f32(int storage) {
  switch (storage) {
    case 1: return k_pop_to_f32<StoreF16S>;
    case 2: return k_pop_to_f32<StoreBF16S>;
    default: return k_pop_to_f32<StoreF32>;
  }
}
ConvInFn get_f32_to_pop(int storage) {
  switch (storage) {
    case 1: return k_f32_to_pop<StoreF16S>;
    case 2: return k_f32_to_pop<StoreBF16S>;
    default: return k_f32_to_pop<StoreF32>;
  }
}
using HistFn = void (*)(const Params, unsigned long long*);
HistFn get_flag_hist() { return k_flag_hist; }
}  // namespace flb
