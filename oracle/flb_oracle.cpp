// flb_oracle.cpp -- CPU oracle of the FlowLB D2Q9 hot path.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product).  A serial
// restatement of the reference kernels, written from their semantics:
//   * src/cuda/lbmKernel2D.cu          d_StreamCollide2D, d_FreeSurface2D_1..4
//   * src/cuda/cellCalculations.cuh    helpers, esoteric-pull slot algebra
//   * src/initData.h, src/initData.cpp host set-up of flags and populations
// Arithmetic follows the reference's double instantiation (PRECISION = double)
// including its float-typed constants (1/tau, F, 0.33333334f, 1.001f, the float
// neighbour averages and TRT temporaries).  Cross-cell reads inside one kernel
// see the pre-kernel state (the reference's in-kernel races resolved that way).
//
// Pinning of each part (closed forms, reference fixtures, brute force) lives in
// tests/test_oracle_*.py.  The Brinkman term and the adjoint are extensions with
// no reference counterpart; they are pinned by finite differences.
#include "flb_oracle.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int CX[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};   // cudaInitData.cu:76
constexpr int CY[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};   // cudaInitData.cu:78

enum : uint8_t {
  FLUID = 0x01, GAS = 0x02, INTERFACE = 0x04, WALL = 0x08, MWALL = 0x10, INLET = 0x20,
  OUTLET = 0x40, GAS_INTERFACE = 0x80,
  FLUID_GAS_INTERFACE = 0x07, INTERFACE_FLUID = 0x05, INTERFACE_GAS = 0x06, INLET_FLUID = 0x21
};

// Lattice weights as the host builds them (src/ui/app.cpp:288,300).
template <class T> struct Weights;
template <> struct Weights<float> {
  static constexpr float w[9] = {4.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f, 1.0f / 9.0f,
                                 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
};
template <> struct Weights<double> {
  static constexpr double w[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                  1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

// __saturatef: clamp to [0, 1], NaN -> 0.  It takes a float, so a double
// argument is rounded to float first (the PRECISION = 64 kernels pass doubles).
inline float saturate(float v) { return v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f; }

// Equilibrium, calculateEquilibriumDDF2Q9 (cellCalculations.cuh:32-52).
template <class T>
inline void equilibrium(T rho, T ux, T uy, T* feq) {
  const T* w = Weights<T>::w;
  const T u2 = 3.0f * (ux * ux + uy * uy);
  const T rhow14 = w[1] * rho;
  const T rhow58 = w[5] * rho;
  ux *= 3.0f;
  uy *= 3.0f;
  const T u0 = ux - uy;
  const T u1 = ux + uy;
  feq[0] = w[0] * rho * (1.0f - 0.5f * u2);
  feq[1] = rhow14 * (1.0f + ux - 0.5f * u2 + 0.5f * ux * ux);
  feq[2] = rhow14 * (1.0f - ux - 0.5f * u2 + 0.5f * ux * ux);
  feq[3] = rhow14 * (1.0f + uy - 0.5f * u2 + 0.5f * uy * uy);
  feq[4] = rhow14 * (1.0f - uy - 0.5f * u2 + 0.5f * uy * uy);
  feq[5] = rhow58 * (1.0f + u1 - 0.5f * u2 + 0.5f * u1 * u1);
  feq[6] = rhow58 * (1.0f - u1 - 0.5f * u2 + 0.5f * u1 * u1);
  feq[7] = rhow58 * (1.0f + u0 - 0.5f * u2 + 0.5f * u0 * u0);
  feq[8] = rhow58 * (1.0f - u0 - 0.5f * u2 + 0.5f * u0 * u0);
}

// Guo forcing terms, calculateForcingTerms2D (cellCalculations.cuh:64-73).  The
// force components are floats in the reference (d_Fx, d_Fy); c_i . F is formed
// in float.
template <class T>
inline void forcing(T ux, T uy, float fx, float fy, T* fi) {
  const T* w = Weights<T>::w;
  const T uF = -0.33333334f * (ux * fx + uy * fy);
  fi[0] = 9.0f * w[0] * uF;
  for (int i = 1; i < 9; i++) {
    const float cF = float(CX[i]) * fx + float(CY[i]) * fy;
    fi[i] = 9.0f * w[i] * (cF * (CX[i] * ux + CY[i] * uy + 0.33333334f) + uF);
  }
}

// Same, with a per-cell force in T (Brinkman extension).
template <class T>
inline void forcing_t(T ux, T uy, T gx, T gy, T* fi) {
  const T* w = Weights<T>::w;
  const T uF = -0.33333334f * (ux * gx + uy * gy);
  fi[0] = 9.0f * w[0] * uF;
  for (int i = 1; i < 9; i++) {
    const T cG = CX[i] * gx + CY[i] * gy;
    fi[i] = 9.0f * w[i] * (cG * (CX[i] * ux + CY[i] * uy + 0.33333334f) + uF);
  }
}

template <class T>
inline T phi_of(T m, T rho) {  // calculatePhi (cellCalculations.cuh:153-158)
  if (rho > 0.0f) return T(saturate(float(m / rho)));
  return 0.5f;
}

// Periodic neighbour indices, calculateNeighborsIdxD2Q9 (cellCalculations.cuh:82-109).
inline void neighbours(uint32_t x, uint32_t y, uint32_t nx, uint32_t ny, uint32_t* nb) {
  const uint32_t yc = y * nx;
  const uint32_t yu = ((y + 1u) % ny) * nx;
  const uint32_t yd = ((y + ny - 1u) % ny) * nx;
  const uint32_t xr = (x + 1u) % nx;
  const uint32_t xl = (x + nx - 1u) % nx;
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

// Esoteric-pull slot algebra (cellCalculations.cuh:185-238).
template <class T>
inline void loadf(const T* f, size_t N, uint32_t n, const uint32_t* nb, uint64_t t, T* lf) {
  const bool odd = t & 1u;
  lf[0] = f[n];
  for (int i = 1; i < 9; i += 2) {
    lf[i] = f[n + N * (odd ? i : i + 1)];
    lf[i + 1] = f[nb[i] + N * (odd ? i + 1 : i)];
  }
}
template <class T>
inline void load_outgoing(const T* f, size_t N, uint32_t n, const uint32_t* nb, uint64_t t, T* of) {
  const bool odd = t & 1u;
  for (int i = 1; i < 9; i += 2) {
    of[i] = f[nb[i] + N * (odd ? i : i + 1)];
    of[i + 1] = f[n + N * (odd ? i + 1 : i)];
  }
}
template <class T>
inline void storef(T* f, size_t N, uint32_t n, const uint32_t* nb, uint64_t t, const T* lf) {
  const bool odd = t & 1u;
  f[n] = lf[0];
  for (int i = 1; i < 9; i += 2) {
    f[nb[i] + N * (odd ? i + 1 : i)] = lf[i];
    f[n + N * (odd ? i : i + 1)] = lf[i + 1];
  }
}

struct Params {
  uint32_t nx, ny;
  float inv_tau, fx, fy;
  bool use_force, smagorinsky, free_surface, gas_check_corrected;
  int collision, mass_mode, threads;
};

template <class T>
struct Sim {
  Params p;
  uint32_t nx, ny;
  size_t N;
  uint64_t t = 0;
  std::vector<T> f, rho, u, mass, excess, phi, alpha;
  std::vector<uint8_t> flags;
  // scratch
  std::vector<T> rho_snap, u_snap, rec;
  std::vector<uint16_t> rec_mask;
  std::vector<uint8_t> flags_snap;

  int threads() const { return p.threads > 0 ? p.threads : 1; }

  // -------------------------------------------------------------- set-up
  // initFields (src/initData.h:125-165): u = inlet velocity at inlets, 0 elsewhere
  // (here taken from u0), rho = 1 (here rho0), f = feq stored with the odd
  // pattern (storef with t = 1).  The equilibrium is evaluated in double as the
  // host does (initEquilibriumDDF2Q9, src/initData.h:75-100).
  void init(const uint8_t* fl, const double* rho0, const double* u0, const double* al) {
    nx = p.nx;
    ny = p.ny;
    N = size_t(nx) * ny;
    f.assign(9 * N, T(0));
    rho.assign(N, T(1));
    u.assign(2 * N, T(0));
    flags.assign(fl, fl + N);
    if (p.free_surface) {
      mass.assign(N, T(0));
      excess.assign(N, T(0));
      phi.assign(N, T(0));
    }
    if (al) alpha.assign(al, al + N);
    for (size_t n = 0; n < N; n++) {
      if (rho0) rho[n] = T(rho0[n]);
      if (u0) {
        u[n] = T(u0[n]);
        u[n + N] = T(u0[n + N]);
      }
    }
    for (uint32_t y = 0; y < ny; y++)
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        T ux = u[n], uy = u[n + N];
        const double r = rho0 ? rho0[n] : 1.0;
        const double u2 = 3.0 * (ux * ux + uy * uy);
        const double rw14 = double(Weights<T>::w[1]) * r, rw58 = double(Weights<T>::w[5]) * r;
        ux *= 3.0;
        uy *= 3.0;
        const double a0 = ux - uy, a1 = ux + uy;
        T lf[9];
        lf[0] = T(double(Weights<T>::w[0]) * r * (1.0 - 0.5 * u2));
        lf[1] = T(rw14 * (1.0 + ux - 0.5 * u2 + 0.5 * ux * ux));
        lf[2] = T(rw14 * (1.0 - ux - 0.5 * u2 + 0.5 * ux * ux));
        lf[3] = T(rw14 * (1.0 + uy - 0.5 * u2 + 0.5 * uy * uy));
        lf[4] = T(rw14 * (1.0 - uy - 0.5 * u2 + 0.5 * uy * uy));
        lf[5] = T(rw58 * (1.0 + a1 - 0.5 * u2 + 0.5 * a1 * a1));
        lf[6] = T(rw58 * (1.0 - a1 - 0.5 * u2 + 0.5 * a1 * a1));
        lf[7] = T(rw58 * (1.0 + a0 - 0.5 * u2 + 0.5 * a0 * a0));
        lf[8] = T(rw58 * (1.0 - a0 - 0.5 * u2 + 0.5 * a0 * a0));
        storef(f.data(), N, n, nb, 1, lf);
      }
    rho_snap.resize(N);
    u_snap.resize(2 * N);
    if (p.free_surface) {
      rec.assign(9 * N, T(0));
      rec_mask.assign(N, 0);
      flags_snap.resize(N);
    }
  }

  // -------------------------------------------------------------- kernels
  // d_StreamCollide2D (lbmKernel2D.cu:305-519).
  void stream_collide() {
    const uint64_t tt = t;
    std::memcpy(rho_snap.data(), rho.data(), N * sizeof(T));
    std::memcpy(u_snap.data(), u.data(), 2 * N * sizeof(T));
    if (p.free_surface) std::memcpy(flags_snap.data(), flags.data(), N);
    const uint8_t* fsnap = p.free_surface ? flags_snap.data() : flags.data();
#pragma omp parallel for schedule(static) num_threads(threads())
    for (int64_t yy = 0; yy < int64_t(ny); yy++) {
      const uint32_t y = uint32_t(yy);
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        const uint8_t flag = fsnap[n];
        if (flag & (GAS | WALL)) continue;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        T lf[9];
        loadf(f.data(), N, n, nb, tt, lf);
        if (p.free_surface && rec_mask[n]) {
          // Reconstructed populations written by the free-surface pass into
          // this cell's incoming slots (storeReconstructedGasf).
          for (int i = 1; i < 9; i++)
            if (rec_mask[n] & (1u << i)) lf[i] = rec[i * N + n];
        }
        T r, ux, uy;
        const uint8_t io = flag & (INLET | OUTLET);
        if (io & INLET) {
          r = rho_snap[n + 1 < N ? n + 1 : n];
          ux = u_snap[n];
          uy = u_snap[n + N];
        } else if (flag & OUTLET) {
          const size_t s = n > 0 ? n - 1 : n;
          r = rho_snap[n];
          ux = u_snap[s];
          uy = u_snap[s + N];
        } else {
          r = lf[0];
          for (int i = 1; i < 9; i++) r += lf[i];
          ux = (lf[1] - lf[2] + lf[5] - lf[6] + lf[7] - lf[8]) / r;
          uy = (lf[3] - lf[4] + lf[5] - lf[6] + lf[8] - lf[7]) / r;
        }
        if ((flag & INTERFACE) == INTERFACE) {  // interface transition test (:353-367)
          const T m = mass[n];
          bool not_fluid = true, not_gas = true;
          for (int i = 1; i < 9; i++) {
            const uint8_t e = fsnap[nb[i]] & FLUID_GAS_INTERFACE;
            not_fluid = not_fluid && (e != FLUID);
            not_gas = not_gas && (e != GAS);
          }
          if ((m > 1.001f * r) || not_gas)
            flags[n] = (flag & ~FLUID_GAS_INTERFACE) | INTERFACE_FLUID;
          else if ((m < -0.001f) || not_fluid)
            flags[n] = (flag & ~FLUID_GAS_INTERFACE) | INTERFACE_GAS;
        }
        T fi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const bool brink = !alpha.empty() && !io;
        if (brink) {
          // Brinkman drag G = F - alpha*rho*u with the Guo half-force shift
          // solved for u: u = (u_raw + F/(2 rho)) / (1 + alpha/2).
          const T a = alpha[n];
          const T fx = p.use_force ? T(p.fx) : T(0), fy = p.use_force ? T(p.fy) : T(0);
          const T rhol2 = 0.5f / r;
          ux = (ux + rhol2 * fx) / (T(1) + T(0.5) * a);
          uy = (uy + rhol2 * fy) / (T(1) + T(0.5) * a);
          forcing_t(ux, uy, fx - a * r * ux, fy - a * r * uy, fi);
        } else if (p.use_force) {
          const T rhol2 = 0.5f / r;
          ux = ux + rhol2 * p.fx;
          uy = uy + rhol2 * p.fy;
          forcing(ux, uy, p.fx, p.fy, fi);
        }
        if (!io) {
          rho[n] = r;
          u[n] = ux;
          u[n + N] = uy;
        } else if (flag & OUTLET) {
          u[n] = ux;
          u[n + N] = uy;
        }
        rho[n] = r;
        T feq[9];
        equilibrium(r, ux, uy, feq);
        float inv_tau = p.inv_tau;
        if (p.smagorinsky) {  // (:417-442)
          T hxx = 0.0f, hyy = 0.0f, hxy = 0.0f;
          for (int i = 1; i < 9; i++) {
            const T fne = lf[i] - feq[i];
            const T cx = T(CX[i]), cy = T(CY[i]);
            hxx += cx * cx * fne;
            hyy += cy * cy * fne;
            hxy += cx * cy * fne;
          }
          const T Q = hxx * hxx + hyy * hyy + 2.0f * hxy * hxy;
          const float tau = 1.0f / inv_tau;
          inv_tau = 2.0f / (tau + sqrtf(float(tau * tau + 0.54037965f * sqrtf(float(Q)) / r)));
        }
        const bool forced = p.use_force || brink;
        if (p.collision == 0) {  // SRT (:446-473)
          if (!io) {
            if (forced) {
              const float taul = 1.0f - 0.5f * inv_tau;
              for (int i = 0; i < 9; i++)
                lf[i] = (1.0f - inv_tau) * lf[i] + inv_tau * feq[i] + taul * fi[i];
            } else {
              for (int i = 0; i < 9; i++) lf[i] = (1.0f - inv_tau) * lf[i] + inv_tau * feq[i];
            }
          } else {
            for (int i = 0; i < 9; i++) lf[i] = feq[i];
          }
        } else {  // TRT (:475-510); with forcing the reference performs no collision
          // collision 2 (no reference counterpart) defines the forced branch:
          // the Guo terms split into the pair-symmetric part, weighted
          // 1 - w+/2, and the antisymmetric part, weighted 1 - w-/2.
          const float ip = inv_tau;
          const float in = 1.0f / (0.25f / (1.0f / inv_tau - 0.5f) + 0.5f);
          if (!io) {
            if (!forced || p.collision == 2) {
              lf[0] = lf[0] - ip * (lf[0] - feq[0]);
              for (int i = 1; i < 9; i += 2) {
                const float fmf = float(lf[i] - feq[i] + lf[i + 1] - feq[i + 1]);
                const float fa = float(lf[i] - lf[i + 1]);
                const float fea = float(feq[i] - feq[i + 1]);
                lf[i] = lf[i] - 0.5f * ip * fmf - 0.5f * in * (fa - fea);
                lf[i + 1] = lf[i + 1] - 0.5f * ip * fmf - 0.5f * in * (fea - fa);
              }
              if (forced) {
                const T ap = T(1.0f - 0.5f * ip), an = T(1.0f - 0.5f * in);
                lf[0] = lf[0] + ap * fi[0];
                for (int i = 1; i < 9; i += 2) {
                  const T fs = T(0.5f) * (fi[i] + fi[i + 1]), fd = T(0.5f) * (fi[i] - fi[i + 1]);
                  lf[i] = lf[i] + (ap * fs + an * fd);
                  lf[i + 1] = lf[i + 1] + (ap * fs - an * fd);
                }
              }
            }
          } else {
            for (int i = 0; i < 9; i++) lf[i] = feq[i];
          }
        }
        storef(f.data(), N, n, nb, tt, lf);
      }
    }
  }

  // d_FreeSurface2D_1 (lbmKernel2D.cu:13-121).  The reconstructed gas
  // populations are collected and applied when the next kernel loads them, so
  // every outgoing-population read sees the pre-kernel array.
  void fs1() {
    const uint64_t tt = t;
    std::fill(rec_mask.begin(), rec_mask.end(), uint16_t(0));
#pragma omp parallel for schedule(static) num_threads(threads())
    for (int64_t yy = 0; yy < int64_t(ny); yy++) {
      const uint32_t y = uint32_t(yy);
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        const uint8_t flag = flags[n];
        if (flag & (GAS | WALL | INLET)) continue;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        T in[9], out[9];
        loadf(f.data(), N, n, nb, tt, in);
        out[0] = in[0];
        load_outgoing(f.data(), N, n, nb, tt, out);
        // As written the increments are added to the mass one by one; the
        // conservative mode sums them first and adds the total once, so that
        // excess shares far below the rounding step of the mass are not lost.
        T m0 = mass[n];
        T dm = 0.0f;
        T& m = p.mass_mode == 0 ? m0 : dm;
        for (int i = 1; i < 9; i++) m += excess[nb[i]];
        const uint8_t e = flag & FLUID_GAS_INTERFACE;
        if (e == FLUID && p.mass_mode == 0) {
          for (int i = 1; i < 9; i += 2) {
            m += in[i + 1] - out[i];
            m += in[i] - out[i + 1];
          }
        } else if (e == FLUID) {
          // Conservative mode: mass crosses only fluid/interface links, so both
          // cells of a link book the same amount with opposite signs.
          for (int i = 1; i < 9; i += 2) {
            const uint8_t a = flags[nb[i]] & FLUID_GAS_INTERFACE;
            const uint8_t b = flags[nb[i + 1]] & FLUID_GAS_INTERFACE;
            if (a == FLUID || a == INTERFACE) m += in[i + 1] - out[i];
            if (b == FLUID || b == INTERFACE) m += in[i] - out[i + 1];
          }
        } else if (e == INTERFACE) {
          T r, ux, uy;
          if (flag & (INLET | OUTLET)) {
            r = rho[n];
            ux = u[n];
            uy = u[n + N];
          } else {
            r = out[0];
            for (int i = 1; i < 9; i++) r += out[i];
            ux = (out[1] - out[2] + out[5] - out[6] + out[7] - out[8]) / r;
            uy = (out[3] - out[4] + out[5] - out[6] + out[8] - out[7]) / r;
          }
          T nphi[9];
          for (int i = 1; i < 9; i++) nphi[i] = phi[nb[i]];
          // As written the own fill level is the fresh one; the conservative
          // mode uses the stored one so that both cells of a link agree.
          nphi[0] = p.mass_mode == 0 ? phi_of(m0, r) : phi[n];
          const T rhol2 = 0.5f / r;
          ux = ux + rhol2 * p.fx;
          uy = uy + rhol2 * p.fy;
          T feq[9];
          equilibrium(T(1.0f - 0.0f), ux, uy, feq);
          uint8_t nfl[9];
          for (int i = 1; i < 9; i++) nfl[i] = flags[nb[i]] & FLUID_GAS_INTERFACE;
          for (int i = 1; i < 9; i += 2) {
            uint8_t ne = nfl[i];
            if (ne == FLUID) m += in[i + 1] - out[i];
            else if (ne == INTERFACE) m += (in[i + 1] - out[i]) * 0.5f * (nphi[i] + nphi[0]);
            ne = nfl[i + 1];
            if (ne == FLUID) m += in[i] - out[i + 1];
            else if (ne == INTERFACE) m += (in[i] - out[i + 1]) * 0.5f * (nphi[i + 1] + nphi[0]);
          }
          T rc[9];
          for (int i = 1; i < 9; i += 2) {
            rc[i] = feq[i] - out[i + 1] + feq[i + 1];
            rc[i + 1] = feq[i + 1] - out[i] + feq[i];
          }
          // storeReconstructedGasf (cellCalculations.cuh:250-259): rc[k] replaces
          // the incoming population k.  As written the gate is the neighbour in
          // direction +c_k; corrected it is the source neighbour at -c_k.
          uint16_t mask = 0;
          for (int k = 1; k < 9; k++) {
            const int opp = (k & 1) ? k + 1 : k - 1;
            const uint8_t gate = p.gas_check_corrected ? nfl[opp] : nfl[k];
            if (gate == GAS) {
              mask |= uint16_t(1u << k);
              rec[k * N + n] = rc[k];
            }
          }
          rec_mask[n] = mask;
        }
        if (p.mass_mode != 0) m0 = m0 + dm;
        mass[n] = m0;
      }
    }
  }

  // d_FreeSurface2D_2 (lbmKernel2D.cu:123-148), scatter form as written.
  void fs2() {
    for (uint32_t y = 0; y < ny; y++)
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        if ((flags[n] & INTERFACE_FLUID) != INTERFACE_FLUID) continue;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        for (int i = 1; i < 9; i++) {
          const uint8_t nf = flags[nb[i]];
          const uint8_t e = nf & INTERFACE_GAS;
          if (e == INTERFACE_GAS) flags[nb[i]] = nf & ~GAS;
          else if (e == GAS) flags[nb[i]] = (nf & ~GAS) | GAS_INTERFACE;
        }
      }
  }

  // d_FreeSurface2D_3 (lbmKernel2D.cu:150-190).
  void fs3() {
    const uint64_t tt = t;
    std::memcpy(flags_snap.data(), flags.data(), N);
    // Gas -> interface cells: populations = feq of the neighbour averages
    // (calculateAverageValuesNeighbors2D accumulates in float).
    for (uint32_t y = 0; y < ny; y++)
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        if ((flags_snap[n] & (GAS_INTERFACE | INTERFACE_GAS)) != GAS_INTERFACE) continue;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        float cnt = 0.0f, sr = 0.0f, sx = 0.0f, sy = 0.0f;
        for (int i = 1; i < 9; i++) {
          if (flags_snap[nb[i]] & INTERFACE_FLUID) {
            cnt += 1.0f;
            sr += rho[nb[i]];
            sx += u[nb[i]];
            sy += u[nb[i] + N];
          }
        }
        T r = cnt > 0.0f ? sr / cnt : 0.0f;
        T ux = cnt > 0.0f ? sx / cnt : 0.0f;
        T uy = cnt > 0.0f ? sy / cnt : 0.0f;
        const T rhol2 = 0.5f / r;
        ux = ux + rhol2 * p.fx;
        uy = uy + rhol2 * p.fy;
        T feq[9];
        equilibrium(r, ux, uy, feq);
        storef(f.data(), N, n, nb, tt, feq);
      }
    // Interface -> gas cells: fluid neighbours become interface (scatter form).
    for (uint32_t y = 0; y < ny; y++)
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        if ((flags[n] & (GAS_INTERFACE | INTERFACE_GAS)) != INTERFACE_GAS) continue;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        for (int i = 1; i < 9; i++) {
          const uint8_t nf = flags[nb[i]];
          const uint8_t e = nf & (INLET_FLUID | INTERFACE);
          if (e != INLET_FLUID && (e == FLUID || e == INTERFACE_FLUID))
            flags[nb[i]] = (nf & ~INTERFACE_FLUID) | INTERFACE;
        }
      }
  }

  // Flag a cell will carry after d_FreeSurface2D_4's own transition.
  static uint8_t transition(uint8_t fl) {
    if (fl & (WALL | MWALL | INLET)) return fl;
    switch (fl & (FLUID_GAS_INTERFACE | GAS_INTERFACE)) {
      case INTERFACE_FLUID: return fl & ~INTERFACE;
      case INTERFACE_GAS: return fl & ~INTERFACE;
      case GAS_INTERFACE: return (fl & ~GAS_INTERFACE) | INTERFACE;
      default: return fl;
    }
  }

  // d_FreeSurface2D_4 (lbmKernel2D.cu:192-303).  Receivers are counted on the
  // pre-kernel flags (mass_mode 0), or as the cells that pull excess mass in
  // the next d_FreeSurface2D_1: not GAS, WALL or INLET after the transition
  // (mass_mode 1).
  void fs4() {
    std::memcpy(flags_snap.data(), flags.data(), N);
#pragma omp parallel for schedule(static) num_threads(threads())
    for (int64_t yy = 0; yy < int64_t(ny); yy++) {
      const uint32_t y = uint32_t(yy);
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        const uint8_t flag = flags_snap[n];
        if (flag & (WALL | MWALL | INLET)) continue;
        const T r = rho[n];
        T m = mass[n];
        T ex = 0.0f;
        T ph = 0.0f;
        switch (flag & (FLUID_GAS_INTERFACE | GAS_INTERFACE)) {
          case FLUID:
            ex = m - r;
            m = r;
            ph = 1.0f;
            break;
          case GAS:
            ex = m;
            m = 0.0f;
            ph = 0.0f;
            break;
          case INTERFACE:
            if (m > r) {
              ex = m - r;
              m = r;
            } else if (m < 0.0f) {
              ex = m;
              m = 0.0f;
            } else
              ex = 0.0f;
            ph = phi_of(m, r);
            break;
          case INTERFACE_FLUID:
            flags[n] = flag & ~INTERFACE;
            ex = m - r;
            m = r;
            ph = 1.0f;
            break;
          case INTERFACE_GAS:
            flags[n] = flag & ~INTERFACE;
            ex = m;
            m = 0.0f;
            ph = 0.0f;
            break;
          case GAS_INTERFACE:
            flags[n] = (flag & ~GAS_INTERFACE) | INTERFACE;
            if (m > r) {
              ex = m - r;
              m = r;
            } else if (m < 0.0f) {
              ex = m;
              m = 0.0f;
            } else
              ex = 0.0f;
            ph = phi_of(m, r);
            break;
          default:
            break;
        }
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        int cont = 0;
        for (int i = 1; i < 9; i++) {
          const uint8_t nf = flags_snap[nb[i]];
          if (p.mass_mode == 0) {
            const uint8_t e = nf & (INTERFACE_FLUID | GAS_INTERFACE);
            cont += int(e == FLUID || e == INTERFACE || e == INTERFACE_FLUID || e == GAS_INTERFACE);
          } else {
            cont += int((transition(nf) & (GAS | WALL | INLET)) == 0);
          }
        }
        m += cont > 0 ? T(0.0f) : ex;
        ex = cont > 0 ? ex / float(cont) : T(0.0f);
        mass[n] = m;
        excess[n] = ex;
        phi[n] = ph;
      }
    }
  }

  void step() {
    if (p.free_surface) {
      fs1();
      stream_collide();
      fs2();
      fs3();
      fs4();
    } else {
      stream_collide();
    }
    t++;
  }

  // Fluid mass: stored mass plus the excess shares the next d_FreeSurface2D_1
  // pulls in (cells that are not GAS, WALL or INLET collect the shares of all
  // eight neighbours).
  double total_mass() const {
    double s = 0.0;
    for (uint32_t y = 0; y < ny; y++)
      for (uint32_t x = 0; x < nx; x++) {
        const uint32_t n = x + y * nx;
        s += double(mass[n]);
        if (flags[n] & (GAS | WALL | INLET)) continue;
        uint32_t nb[9];
        neighbours(x, y, nx, ny, nb);
        for (int i = 1; i < 9; i++) s += double(excess[nb[i]]);
      }
    return s;
  }
};

Params to_params(const orc_params* q) {
  Params p;
  p.nx = q->nx;
  p.ny = q->ny;
  p.inv_tau = q->inv_tau;
  p.fx = q->use_force ? q->fx : 0.0f;
  p.fy = q->use_force ? q->fy : 0.0f;
  p.use_force = q->use_force != 0;
  p.smagorinsky = q->smagorinsky != 0;
  p.free_surface = q->free_surface != 0;
  p.gas_check_corrected = q->gas_check_corrected != 0;
  p.collision = q->collision;
  p.mass_mode = q->mass_mode;
  p.threads = q->threads;
  return p;
}

template <class T>
void* create(const orc_params* q, const uint8_t* flags, const double* rho0, const double* u0,
             const double* alpha) {
  auto* s = new Sim<T>();
  s->p = to_params(q);
  s->init(flags, rho0, u0, alpha);
  return s;
}

template <class T>
void get(Sim<T>* s, double* f, double* rho, double* u, uint8_t* flags, double* mass, double* excess,
         double* phi) {
  const size_t N = s->N;
  if (f)
    for (size_t i = 0; i < 9 * N; i++) f[i] = double(s->f[i]);
  if (rho)
    for (size_t i = 0; i < N; i++) rho[i] = double(s->rho[i]);
  if (u)
    for (size_t i = 0; i < 2 * N; i++) u[i] = double(s->u[i]);
  if (flags) std::memcpy(flags, s->flags.data(), N);
  if (s->p.free_surface) {
    if (mass)
      for (size_t i = 0; i < N; i++) mass[i] = double(s->mass[i]);
    if (excess)
      for (size_t i = 0; i < N; i++) excess[i] = double(s->excess[i]);
    if (phi)
      for (size_t i = 0; i < N; i++) phi[i] = double(s->phi[i]);
  }
}

template <class T>
void set_fs(Sim<T>* s, const double* mass, const double* phi, const double* excess) {
  for (size_t i = 0; i < s->N; i++) {
    if (mass) s->mass[i] = T(mass[i]);
    if (phi) s->phi[i] = T(phi[i]);
    s->excess[i] = excess ? T(excess[i]) : T(0);
  }
}

}  // namespace

// ---------------------------------------------------------------- set-up API

namespace {
// RectangleShape::isNodeInside without rotation (shapes.cpp:58-126), epsilon 1e-12.
bool rect_inside(const orc_rect& r, double xs, double ys) {
  const double eps = 1e-12;
  const double xmax = r.x0 + r.width / 2.0 + eps, xmin = r.x0 - r.width / 2.0 - eps;
  const double ymax = r.y0 + r.height / 2.0 + eps, ymin = r.y0 - r.height / 2.0 - eps;
  return !(xs > xmax || xs < xmin || ys > ymax || ys < ymin);
}
}  // namespace

// initData (src/initData.h:220-333) for a flat bed whose cross-drainage work
// covers the whole domain: every node starts as the default node, the boundary
// rows and columns take their boundary type (down, up, left, right in that
// priority, :296-299), then obstacles become WALL and initial-condition shapes
// become FLUID.
int orc_build_channel_flags(const orc_channel* g, uint8_t* flags) {
  if (!g || !flags || g->nx < 2 || g->ny < 2) return 1;
  const double eps = 1e-12;
  const uint8_t def = g->free_surface ? GAS : FLUID;
  const double ymax = g->y_min + double(g->ny - 1) * g->dx;
  for (uint32_t y = 0; y < g->ny; y++)
    for (uint32_t x = 0; x < g->nx; x++) {
      const size_t n = x + size_t(y) * g->nx;
      const double xs = g->x_min + x * g->dx;
      const double ys = g->y_min + y * g->dx;
      uint8_t fl = def;
      const double base = g->y_min;
      if (ys <= base + eps && fl == def) fl = g->bc_down;
      else if (ys >= ymax - eps && fl == def) fl = g->bc_up;
      else if (x == 0 && fl == def) fl = g->bc_left;
      else if (x == g->nx - 1 && fl == def) fl = g->bc_right;
      for (uint32_t k = 0; k < g->n_obstacles; k++)
        if (rect_inside(g->obstacles[k], xs, ys)) fl = WALL;
      for (uint32_t k = 0; k < g->n_initial; k++)
        if (rect_inside(g->initial[k], xs, ys)) fl = FLUID;
      flags[n] = fl;
    }
  return 0;
}

// initFreeSurfaceFlags (src/initData.cpp:13-37) with the bounded neighbour
// stencil calculateNeighborsIndexD2Q9BBox (:39-124), visited in index order.
void orc_seed_free_surface(uint8_t* flags, uint32_t nx, uint32_t ny) {
  for (uint32_t y = 0; y < ny; y++)
    for (uint32_t x = 0; x < nx; x++) {
      const size_t n = x + size_t(y) * nx;
      uint8_t& fl = flags[n];
      if (fl & ~(FLUID | INLET | OUTLET)) continue;
      if (fl & INLET) fl |= FLUID;
      else if (fl & OUTLET) fl |= GAS;
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          if (!dx && !dy) continue;
          const int64_t xx = int64_t(x) + dx, yy = int64_t(y) + dy;
          if (xx < 0 || yy < 0 || xx >= int64_t(nx) || yy >= int64_t(ny)) continue;
          uint8_t& nf = flags[size_t(xx) + size_t(yy) * nx];
          if (nf == GAS && (fl & FLUID)) nf = INTERFACE;
        }
    }
}

// initFreeSurfaceFields (src/initData.h:175-206).
void orc_init_free_surface_fields(const uint8_t* flags, size_t n, double* mass, double* phi,
                                  double* excess) {
  for (size_t i = 0; i < n; i++) {
    double m = 0.0, ph = 0.0;
    switch (flags[i]) {
      case FLUID:
      case FLUID | INLET:
      case FLUID | OUTLET:
        m = 1.0;
        ph = 1.0;
        break;
      case INTERFACE:
        m = 0.5;
        ph = 0.5;
        break;
      default:
        break;
    }
    mass[i] = m;
    phi[i] = ph;
    excess[i] = 0.0;
  }
}

// ---------------------------------------------------------------- FP64 API

void* orc_create(const orc_params* p, const uint8_t* flags, const double* rho0, const double* u0,
                 const double* alpha) {
  return create<double>(p, flags, rho0, u0, alpha);
}
void orc_set_fs_fields(void* s, const double* mass, const double* phi, const double* excess) {
  set_fs(static_cast<Sim<double>*>(s), mass, phi, excess);
}
void orc_step(void* s, uint64_t nsteps) {
  auto* q = static_cast<Sim<double>*>(s);
  for (uint64_t k = 0; k < nsteps; k++) q->step();
}
uint64_t orc_time(void* s) { return static_cast<Sim<double>*>(s)->t; }
void orc_get(void* s, double* f, double* rho, double* u, uint8_t* flags, double* mass,
             double* excess, double* phi) {
  get(static_cast<Sim<double>*>(s), f, rho, u, flags, mass, excess, phi);
}
void orc_set_f(void* s, const double* f) {
  auto* q = static_cast<Sim<double>*>(s);
  std::memcpy(q->f.data(), f, 9 * q->N * sizeof(double));
}
double orc_total_mass(void* s) { return static_cast<Sim<double>*>(s)->total_mass(); }
void orc_destroy(void* s) { delete static_cast<Sim<double>*>(s); }

double orc_fs_discharge(void* s, uint32_t column) {
  auto* q = static_cast<Sim<double>*>(s);
  if (!q->p.free_surface) return 0.0;
  double acc = 0.0;
  for (uint32_t y = 0; y < q->ny; y++) {
    const size_t n = column + size_t(y) * q->nx;
    if (q->flags[n] & (FLUID | INTERFACE)) acc += q->phi[n] * q->rho[n] * q->u[n];
  }
  return acc;
}

void orc_set_cells(void* s, const uint8_t* flags, const double* rho, const double* u) {
  auto* q = static_cast<Sim<double>*>(s);
  if (flags) std::memcpy(q->flags.data(), flags, q->N);
  if (rho) std::memcpy(q->rho.data(), rho, q->N * sizeof(double));
  if (u) std::memcpy(q->u.data(), u, 2 * q->N * sizeof(double));
}
void orc_set_time(void* s, uint64_t t) { static_cast<Sim<double>*>(s)->t = t; }

int orc_run_pass(void* s, int pass) {
  auto* q = static_cast<Sim<double>*>(s);
  if (!q->p.free_surface && pass != 2) return 1;
  switch (pass) {
    case 1: q->fs1(); return 0;
    case 2: q->stream_collide(); return 0;
    case 3: q->fs2(); return 0;
    case 4: q->fs3(); return 0;
    case 5: q->fs4(); return 0;
    default: return 1;
  }
}

void orc_get_reconstructed(void* s, double* rec, uint16_t* mask) {
  auto* q = static_cast<Sim<double>*>(s);
  if (!q->p.free_surface) return;
  if (rec) std::memcpy(rec, q->rec.data(), 9 * q->N * sizeof(double));
  if (mask) std::memcpy(mask, q->rec_mask.data(), q->N * sizeof(uint16_t));
}

double orc_discharge(void* s, uint32_t column) {
  auto* q = static_cast<Sim<double>*>(s);
  double acc = 0.0;
  for (uint32_t y = 0; y < q->ny; y++) {
    const size_t n = column + size_t(y) * q->nx;
    if (q->flags[n] & (WALL | GAS)) continue;
    acc += q->rho[n] * q->u[n];
  }
  return acc;
}

// ---------------------------------------------------------------- adjoint
// Discrete adjoint of the single-phase step (SRT, Guo forcing, Brinkman drag,
// inlet/outlet, implicit bounce-back), FP64.  Extension without reference
// counterpart; pinned by central finite differences in the tests.
//
// Each cell reads and writes one slot set per parity, so the reverse step is
// per cell and in place: the adjoints of the populations a cell stored are
// read from its store slots, the adjoints of the populations it loaded are
// written to its load slots.  Inlet and outlet cells depend on the previous
// step's published density / velocity of a neighbour; those adjoints travel
// backwards through rbar / ubar (two parity buffers, like the forward).

namespace {

struct Snapshot {
  std::vector<double> f, rho, u;
};

// Store slot of population i of cell n at parity odd (storef).
inline size_t store_slot(size_t N, uint32_t n, const uint32_t* nb, bool odd, int i) {
  if (i == 0) return n;
  if (i & 1) return nb[i] + N * (odd ? i + 1 : i);
  return n + N * (odd ? i - 1 : i);
}
// Load slot of population i of cell n at parity odd (loadf).
inline size_t load_slot(size_t N, uint32_t n, const uint32_t* nb, bool odd, int i) {
  if (i == 0) return n;
  if (i & 1) return n + N * (odd ? i : i + 1);
  return nb[i - 1] + N * (odd ? i : i - 1);
}

struct Adjoint {
  Sim<double>& s;
  uint32_t col;
  uint64_t t0, t1;
  double qbar;
  std::vector<double> abar, rbar[2], ubar[2], galpha;
  double duin = 0.0;

  Adjoint(Sim<double>& sim, uint32_t c, uint64_t a, uint64_t b)
      : s(sim), col(c), t0(a), t1(b), qbar(1.0 / double(b - a + 1)) {
    abar.assign(9 * s.N, 0.0);
    for (int k = 0; k < 2; k++) {
      rbar[k].assign(s.N, 0.0);
      ubar[k].assign(2 * s.N, 0.0);
    }
    galpha.assign(s.N, 0.0);
  }

  static void feq_and_grad(double rho, double ux, double uy, double* fe, double* dfx, double* dfy) {
    const double* w = Weights<double>::w;
    const double usq = ux * ux + uy * uy;
    for (int i = 0; i < 9; i++) {
      const double cu = CX[i] * ux + CY[i] * uy;
      fe[i] = w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
      dfx[i] = w[i] * rho * (3.0 * CX[i] + 9.0 * cu * CX[i] - 3.0 * ux);
      dfy[i] = w[i] * rho * (3.0 * CY[i] + 9.0 * cu * CY[i] - 3.0 * uy);
    }
  }

  // Reverse of step t.  st: state before step t (populations and fields).
  void backward(uint64_t t, const Snapshot& st) {
    const size_t N = s.N;
    const bool odd = t & 1u;
    const int cur = int(t & 1u), prev = cur ^ 1;
    const bool in_window = t >= t0 && t <= t1;
    const double k3 = double(0.33333334f);
    const double om = double(s.p.inv_tau);
    const double tl = double(1.0f - 0.5f * s.p.inv_tau);
    const double Fx = s.p.use_force ? double(s.p.fx) : 0.0, Fy = s.p.use_force ? double(s.p.fy) : 0.0;
    const double* w = Weights<double>::w;
    for (uint32_t y = 0; y < s.ny; y++)
      for (uint32_t x = 0; x < s.nx; x++) {
        const uint32_t n = x + y * s.nx;
        const uint8_t fl = s.flags[n];
        if (fl & (GAS | WALL)) continue;
        uint32_t nb[9];
        neighbours(x, y, s.nx, s.ny, nb);
        double v[9];
        for (int i = 0; i < 9; i++) v[i] = abar[store_slot(N, n, nb, odd, i)];
        const bool pub_r = n > 0 && (s.flags[n - 1] & INLET);
        const bool pub_v = n + 1 < N && (s.flags[n + 1] & OUTLET) && !(fl & INLET);
        double rb_pub = 0.0, ubx_pub = 0.0, uby_pub = 0.0;
        if (pub_r) {
          rb_pub = rbar[cur][n];
          rbar[cur][n] = 0.0;
        }
        if (pub_v) {
          ubx_pub = ubar[cur][n];
          uby_pub = ubar[cur][N + n];
          ubar[cur][n] = ubar[cur][N + n] = 0.0;
        }
        const uint8_t io = fl & (INLET | OUTLET);
        double gb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (io) {
          double rho, ux, uy;
          if (io & INLET) {
            rho = st.rho[n + 1 < N ? n + 1 : n];
            ux = st.u[n];
            uy = st.u[N + n];
          } else {
            rho = st.rho[n];
            ux = st.u[n - 1];
            uy = st.u[N + n - 1];
          }
          const double uxt = ux + 0.5 / rho * Fx, uyt = uy + 0.5 / rho * Fy;
          double fe[9], dfx[9], dfy[9];
          feq_and_grad(rho, uxt, uyt, fe, dfx, dfy);
          double rbt = rb_pub, ubx = ubx_pub, uby = uby_pub;
          for (int i = 0; i < 9; i++) {
            rbt += v[i] * fe[i] / rho;
            ubx += v[i] * dfx[i];
            uby += v[i] * dfy[i];
          }
          // u = u_read + F/(2 rho)
          rbt += -(ubx * Fx + uby * Fy) * 0.5 / (rho * rho);
          if (io & INLET) {
            if (n + 1 < N) rbar[prev][n + 1] += rbt;
            duin += ubx;
          } else {
            ubar[prev][n - 1] += ubx;
            ubar[prev][N + n - 1] += uby;
          }
        } else {
          double g[9];
          loadf(st.f.data(), N, n, nb, t, g);
          double rho = 0.0, mx = 0.0, my = 0.0;
          for (int i = 0; i < 9; i++) {
            rho += g[i];
            mx += CX[i] * g[i];
            my += CY[i] * g[i];
          }
          const double a = s.alpha.empty() ? 0.0 : s.alpha[n];
          const double sa = 1.0 / (1.0 + 0.5 * a);
          const double ux = (mx + 0.5 * Fx) * sa / rho, uy = (my + 0.5 * Fy) * sa / rho;
          const double Gx = Fx - a * rho * ux, Gy = Fy - a * rho * uy;
          const bool forced = s.p.use_force || !s.alpha.empty();
          double fe[9], dfx[9], dfy[9];
          feq_and_grad(rho, ux, uy, fe, dfx, dfy);
          double rb = rb_pub, ubx = ubx_pub, uby = uby_pub, Gbx = 0.0, Gby = 0.0;
          for (int i = 0; i < 9; i++) {
            gb[i] = (1.0 - om) * v[i];
            const double feb = om * v[i];
            rb += feb * fe[i] / rho;
            ubx += feb * dfx[i];
            uby += feb * dfy[i];
            if (forced) {
              const double Fb = tl * v[i];
              const double cG = CX[i] * Gx + CY[i] * Gy;
              const double cu = CX[i] * ux + CY[i] * uy;
              ubx += Fb * 9.0 * w[i] * (cG * CX[i] - k3 * Gx);
              uby += Fb * 9.0 * w[i] * (cG * CY[i] - k3 * Gy);
              Gbx += Fb * 9.0 * w[i] * (CX[i] * (cu + k3) - k3 * ux);
              Gby += Fb * 9.0 * w[i] * (CY[i] * (cu + k3) - k3 * uy);
            }
          }
          // G = F - a rho u
          ubx += -a * rho * Gbx;
          uby += -a * rho * Gby;
          rb += -a * (ux * Gbx + uy * Gby);
          double ab = -rho * (ux * Gbx + uy * Gby);
          if (in_window && x == col) {  // objective: qbar * rho * ux
            rb += qbar * ux;
            ubx += qbar * rho;
          }
          // u = (m + F/2) sa / rho
          const double mbx = ubx * sa / rho, mby = uby * sa / rho;
          const double ubu = ubx * ux + uby * uy;
          rb += -ubu / rho;
          ab += -0.5 * sa * ubu;
          for (int i = 0; i < 9; i++) gb[i] += rb + CX[i] * mbx + CY[i] * mby;
          if (!s.alpha.empty()) galpha[n] += ab;
        }
        for (int i = 0; i < 9; i++) abar[load_slot(N, n, nb, odd, i)] = gb[i];
      }
  }
};

Snapshot snap(const Sim<double>& s) { return Snapshot{s.f, s.rho, s.u}; }
void restore(Sim<double>& s, const Snapshot& z, uint64_t t) {
  s.f = z.f;
  s.rho = z.rho;
  s.u = z.u;
  s.t = t;
}

}  // namespace

double orc_objective(void* sp, uint64_t nsteps, uint32_t column, uint64_t t0, uint64_t t1) {
  auto& s = *static_cast<Sim<double>*>(sp);
  const Snapshot z0 = snap(s);
  const uint64_t tstart = s.t;
  double J = 0.0;
  for (uint64_t k = 0; k < nsteps; k++) {
    s.step();
    if (k >= t0 && k <= t1) J += orc_discharge(sp, column);
  }
  restore(s, z0, tstart);
  return J / double(t1 - t0 + 1);
}

int orc_adjoint(void* sp, uint64_t nsteps, uint32_t column, uint64_t t0, uint64_t t1,
                uint32_t ckpt, double* dj_dalpha, double* dj_duin, double* j_out) {
  auto& s = *static_cast<Sim<double>*>(sp);
  if (s.p.free_surface || s.p.collision != 0 || s.p.smagorinsky) return 1;
  if (t1 < t0 || t1 >= nsteps || ckpt == 0 || column >= s.nx) return 2;
  const uint64_t tstart = s.t;
  if (tstart != 0) return 3;
  const Snapshot z0 = snap(s);
  // forward with checkpoints every ckpt steps
  std::vector<Snapshot> cps;
  double J = 0.0;
  for (uint64_t k = 0; k < nsteps; k++) {
    if (k % ckpt == 0) cps.push_back(snap(s));
    s.step();
    if (k >= t0 && k <= t1) J += orc_discharge(sp, column);
  }
  J /= double(t1 - t0 + 1);
  Adjoint adj(s, column, t0, t1);
  std::vector<Snapshot> seg;
  for (int64_t c = int64_t(cps.size()) - 1; c >= 0; c--) {
    const uint64_t ts = uint64_t(c) * ckpt;
    const uint64_t te = std::min<uint64_t>(ts + ckpt, nsteps);
    restore(s, cps[size_t(c)], ts);
    seg.clear();
    for (uint64_t k = ts; k < te; k++) {
      seg.push_back(snap(s));
      s.step();
    }
    for (uint64_t k = te; k-- > ts;) adj.backward(k, seg[size_t(k - ts)]);
  }
  // The initial populations feq(rho0, u0) of inlet cells depend on u_inlet.
  const size_t N = s.N;
  for (uint32_t y = 0; y < s.ny; y++)
    for (uint32_t x = 0; x < s.nx; x++) {
      const uint32_t n = x + y * s.nx;
      if (!(s.flags[n] & INLET) || (s.flags[n] & (GAS | WALL))) continue;
      uint32_t nb[9];
      neighbours(x, y, s.nx, s.ny, nb);
      double fe[9], dfx[9], dfy[9];
      Adjoint::feq_and_grad(z0.rho[n], z0.u[n], z0.u[N + n], fe, dfx, dfy);
      for (int i = 0; i < 9; i++) adj.duin += adj.abar[store_slot(N, n, nb, true, i)] * dfx[i];
    }
  restore(s, z0, tstart);
  if (dj_dalpha)
    for (size_t n = 0; n < N; n++) dj_dalpha[n] = adj.galpha[n];
  if (dj_duin) *dj_duin = adj.duin;
  if (j_out) *j_out = J;
  return 0;
}

// ---------------------------------------------------------------- FP32 API

void* orcf_create(const orc_params* p, const uint8_t* flags, const double* rho0, const double* u0,
                  const double* alpha) {
  return create<float>(p, flags, rho0, u0, alpha);
}
void orcf_set_fs_fields(void* s, const double* mass, const double* phi, const double* excess) {
  set_fs(static_cast<Sim<float>*>(s), mass, phi, excess);
}
void orcf_step(void* s, uint64_t nsteps) {
  auto* q = static_cast<Sim<float>*>(s);
  for (uint64_t k = 0; k < nsteps; k++) q->step();
}
void orcf_get(void* s, double* f, double* rho, double* u, uint8_t* flags, double* mass,
              double* excess, double* phi) {
  get(static_cast<Sim<float>*>(s), f, rho, u, flags, mass, excess, phi);
}
double orcf_total_mass(void* s) { return static_cast<Sim<float>*>(s)->total_mass(); }
void orcf_destroy(void* s) { delete static_cast<Sim<float>*>(s); }
