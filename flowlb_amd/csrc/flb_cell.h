// flb_cell.h -- per-cell device code of the D2Q9 step, shared by the
// single-domain kernels (flb_kernels.hip) and the batched kernel
// (flb_batch.hip).  gfx950 only.
#pragma once

#include "flb_device.h"

namespace flb {

// Quotients a / d for one divisor d (the density of a cell: velocity, Guo's
// half-force shift).  In FP32 each quotient is the IEEE-rounded a / d: the
// refined reciprocal and the two correction steps are the sequence hipcc emits
// for a float division, without its pre-scaling of operands near the ends of
// the exponent range (none here: d is a density near 1, a is zero or far above
// 2^-100) and with its final fix-up of zeros and specials; the reciprocal is
// formed once per cell instead of once per division (22 -> 15 instructions
// for the velocity, 33 -> 21 with the force).  FP64 divides.
template <class R>
struct Quotient {
  R d;
  __device__ explicit Quotient(R den) : d(den) {}
  __device__ R operator()(R a) const { return a / d; }
};
template <>
struct Quotient<float> {
  float d, r;
  __device__ explicit Quotient(float den) : d(den) {
    const float r0 = __builtin_amdgcn_rcpf(den);
    r = __builtin_fmaf(__builtin_fmaf(-den, r0, 1.0f), r0, r0);
  }
  __device__ float operator()(float a) const {
    const float q0 = a * r;
    const float q1 = __builtin_fmaf(__builtin_fmaf(-d, q0, a), r, q0);
    const float q2 = __builtin_fmaf(__builtin_fmaf(-d, q1, a), r, q1);
    return __builtin_amdgcn_div_fixupf(q2, d, a);
  }
};

// Equilibrium minus weights (feq_i - w_i), calculateEquilibriumDDF2Q9
// (src/cuda/cellCalculations.cuh:32-52) rearranged for shifted storage:
// feq_i - w_i = w_i * ((rho - 1) + rho * A_i).  R is the arithmetic type.
template <class R>
__device__ __forceinline__ void equilibrium_shifted(R rho, R rho1, R ux, R uy, R* fe) {
  const R u2 = R(3) * (ux * ux + uy * uy);
  const R ax = R(3) * ux, ay = R(3) * uy;
  const R a0 = ax - ay, a1 = ax + ay;
  const R h = R(0.5f) * u2;
  const R w0 = wgt_r<R>(0), w1 = wgt_r<R>(1), w5 = wgt_r<R>(5);
  const R q = R(0.5f);
  fe[0] = w0 * (rho1 - rho * h);
  fe[1] = w1 * (rho1 + rho * (ax - h + q * ax * ax));
  fe[2] = w1 * (rho1 + rho * (-ax - h + q * ax * ax));
  fe[3] = w1 * (rho1 + rho * (ay - h + q * ay * ay));
  fe[4] = w1 * (rho1 + rho * (-ay - h + q * ay * ay));
  fe[5] = w5 * (rho1 + rho * (a1 - h + q * a1 * a1));
  fe[6] = w5 * (rho1 + rho * (-a1 - h + q * a1 * a1));
  fe[7] = w5 * (rho1 + rho * (a0 - h + q * a0 * a0));
  fe[8] = w5 * (rho1 + rho * (-a0 - h + q * a0 * a0));
}

// Guo forcing terms, calculateForcingTerms2D (src/cuda/cellCalculations.cuh:64-73),
// with the reference's literal 0.33333334f.  The force is a float constant
// there, so c_i . F is formed in float also in the double path.
template <class R>
__device__ __forceinline__ void guo_terms(R ux, R uy, float gx, float gy, R* fi) {
  const R uF = R(-0.33333334f) * (ux * R(gx) + uy * R(gy));
  fi[0] = R(9.0f) * wgt_r<R>(0) * uF;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    const float cF = float(cx(i)) * gx + float(cy(i)) * gy;
    fi[i] = R(9.0f) * wgt_r<R>(i) *
            (R(cF) * (R(cx(i)) * ux + R(cy(i)) * uy + R(0.33333334f)) + uF);
  }
}

// calculatePhi (cellCalculations.cuh:153-158): __saturatef takes a float, so
// the double path rounds the fill level to float as well.
template <class R>
__device__ __forceinline__ R fill_level(R m, R rho) {
  return rho > R(0) ? R(__saturatef(float(m / rho))) : R(0.5f);
}

struct Cell {
  uint32_t x, y, n;
  bool valid;
};

// 16-bit storage residual of a fluid cell (free surface, conservative mode):
// rho - 1 minus the sum of the populations as stored, kept in plane 0 of rec
// as binary16 of residual * 2^10 (2 bytes read and written per fluid cell
// instead of 4).  A residual is at most 9 half-ulps of the stored values,
// |r| < 0.04 for either 16-bit format (typically 1e-5 for FP16S, 1e-4 for
// BF16S), so the scaled value is a normal binary16 and keeps it to 2^-11
// relative: ~4e-9 absolute per cell and step, unbiased, against the ~1e-5 the
// residual restores.  The first Nd halves of the plane are used.
constexpr float kResidScale = 1024.0f;
__device__ __forceinline__ float resid_load(const Params& p, uint32_t n) {
  return float(at(reinterpret_cast<const _Float16*>(p.rec), n)) * (1.0f / kResidScale);
}
__device__ __forceinline__ void resid_store(const Params& p, uint32_t n, float r) {
  at(reinterpret_cast<_Float16*>(p.rec), n) = _Float16(r * kResidScale);
}

// Block index in the order a step walks its tiles: the odd steps (ODD = t % 2)
// take the bands of the launch last to first, so that a step starts on the
// rows the previous step wrote last, which are still in the 256 MB Infinity
// Cache.  The launch is the window of Params (the whole grid: ncols = tiles_x,
// win_magic = band_magic).
template <class S, int ODD>
__device__ __forceinline__ uint32_t walk_block(const Params& p) {
  if (!ODD) return blockIdx.x;
  return band_reversed(blockIdx.x, p.walk_last, p.walk_bands, kBandRows * p.ncols, p.win_magic);
}

// Cell of this thread: block b (in walk order) = one row segment of
// blockDim.x cells.
__device__ __forceinline__ Cell locate(const Params& p, uint32_t b = blockIdx.x) {
  uint32_t tx, y;
  block_tile(b, p.tiles_x, p.ny, p.band_magic, p.last_magic, tx, y);
  const uint32_t x = tx * blockDim.x + threadIdx.x;
  return Cell{x, y, x + y * p.pitch, x < p.nx};
}

// Cell of this thread in a free-surface stream-collide launch over the window
// of Params (block_tile_window).
__device__ __forceinline__ Cell locate_window(const Params& p, uint32_t b = blockIdx.x) {
  uint32_t tx, y;
  block_tile_window(p, b, tx, y);
  const uint32_t x = tx * blockDim.x + threadIdx.x;
  return Cell{x, y, x + y * p.pitch, x < p.nx};
}

// Free surface: does the 64-cell row chunk of this wavefront have work for the
// pass whose bits start at `base` (kWorkSC, kWorkFS1)?  In conservative mode
// pass 1 has none in chunks of gas and interior fluid, stream-collide none in
// chunks of gas and walls.  The chunk is uniform over the wavefront (a block
// is one row segment that starts at a multiple of 64), so its tile word is a
// scalar load.
// Returns the work bits of the chunk (bit kWorkSC: stream-collide, bit
// kWorkFS1: pass 1, kChunkBody: the tile is a body of fluid, Params::tile_body),
// 0 for a wavefront past the row's end.
constexpr uint32_t kChunkBody = 1u << 30;
__device__ __forceinline__ uint32_t chunk_work(const Params& p, const Cell& c, bool with_pub,
                                               bool with_body) {
  const uint32_t tx = uint32_t(__builtin_amdgcn_readfirstlane(c.x)) / kFsTileW;
  const uint32_t y = uint32_t(__builtin_amdgcn_readfirstlane(c.y));
  if (tx >= p.work_x) return 0u;
  const uint32_t i = (y / kFsTileH) * p.work_x + tx;
  const int r = int(y % kFsTileH);
  const uint32_t w = p.tile_work[i];
  const uint32_t sc = (w | (with_pub ? p.tile_pub[i] : 0u)) >> (kWorkSC + r);
  const uint32_t body = with_body && p.tile_body[i] ? kChunkBody : 0u;
  return (sc & 1u) << kWorkSC | ((w >> (kWorkFS1 + r)) & 1u) << kWorkFS1 | body;
}

// Periodic neighbour indices (src/cuda/cellCalculations.cuh:82-109) without
// integer division: the wrap is a compare at the domain edge.
__device__ __forceinline__ void neighbours(const Params& p, uint32_t x, uint32_t y, uint32_t* nb) {
  const uint32_t nx = p.nx, pitch = p.pitch;
  const uint32_t xr = x + 1 == nx ? 0u : x + 1;
  const uint32_t xl = x == 0 ? nx - 1 : x - 1;
  const uint32_t yc = y * pitch;
  const uint32_t yu = (y + 1 == p.ny ? 0u : y + 1) * pitch;
  const uint32_t yd = (y == 0 ? p.ny - 1 : y - 1) * pitch;
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

// The cells before and after (x, y) in the reference's row-major numbering
// n = x + y * nx, as device indices: the inlet reads rho[n + 1] and the outlet
// u[n - 1] (lbmKernel2D.cu:336,344), and the last cell of a row is followed by
// the first cell of the next row.  Where a neighbour does not exist the index
// is the cell itself and has_* is false.
struct Linear {
  uint32_t prev, next;
  bool has_prev, has_next;
};
__device__ __forceinline__ Linear linear_neighbours(const Params& p, uint32_t x, uint32_t y,
                                                    uint32_t n) {
  Linear l;
  l.has_prev = x > 0 || y > 0;
  l.has_next = x + 1 < p.nx || y + 1 < p.ny;
  l.prev = x > 0 ? n - 1 : (y > 0 ? n - p.pitch + p.nx - 1 : n);
  l.next = x + 1 < p.nx ? n + 1 : (y + 1 < p.ny ? n + p.pitch - p.nx + 1 : n);
  return l;
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

// Buffer resource of population plane q.
template <class S>
__device__ __forceinline__ Rsrc pop_plane(const Params& p, int q) {
  using T = typename S::T;
  return plane_rsrc(static_cast<const T*>(p.f) + uint64_t(q) * p.Np, p.Np * sizeof(T));
}

// Population access through the plane's buffer resource (S::kBuffer) or a
// 32-bit offset from the plane's base.
template <class S>
__device__ __forceinline__ typename S::R pop_load(const Params& p, int q, uint32_t cell) {
  using T = typename S::T;
  if (S::kBuffer) return S::bload(pop_plane<S>(p, q), cell);
  return S::load(static_cast<const T*>(p.f) + uint64_t(q) * p.Np, cell);
}
template <class S, bool BUF = S::kBuffer>
__device__ __forceinline__ void pop_store(const Params& p, int q, uint32_t cell,
                                          typename S::R v) {
  using T = typename S::T;
  if (BUF) S::bstore(pop_plane<S>(p, q), cell, v);
  else S::store(static_cast<T*>(p.f) + uint64_t(q) * p.Np, cell, v);
}

// loadf (src/cuda/cellCalculations.cuh:185-196) with ODD = t % 2.
template <class S, int ODD>
__device__ __forceinline__ void load_incoming(const Params& p, const uint32_t* nb,
                                              typename S::R* g) {
  g[0] = pop_load<S>(p, 0, nb[0]);
#pragma unroll
  for (int i = 1; i < 9; i += 2) {
    g[i] = pop_load<S>(p, ODD ? i : i + 1, nb[0]);
    g[i + 1] = pop_load<S>(p, ODD ? i + 1 : i, nb[i]);
  }
}

// loadOutgoingf (src/cuda/cellCalculations.cuh:207-217).
template <class S, int ODD>
__device__ __forceinline__ void load_outgoing(const Params& p, const uint32_t* nb,
                                              typename S::R* o) {
#pragma unroll
  for (int i = 1; i < 9; i += 2) {
    o[i] = pop_load<S>(p, ODD ? i : i + 1, nb[i]);
    o[i + 1] = pop_load<S>(p, ODD ? i + 1 : i, nb[0]);
  }
}

// storef (src/cuda/cellCalculations.cuh:228-238).  BUF false: plain stores
// from each plane's base (two scalar registers a plane instead of a
// four-register buffer descriptor), for kernels short of scalar registers.
template <class S, int ODD, bool BUF = S::kBuffer>
__device__ __forceinline__ void store_outgoing(const Params& p, const uint32_t* nb,
                                               const typename S::R* g) {
  pop_store<S, BUF>(p, 0, nb[0], g[0]);
#pragma unroll
  for (int i = 1; i < 9; i += 2) {
    pop_store<S, BUF>(p, ODD ? i + 1 : i, nb[i], g[i]);
    pop_store<S, BUF>(p, ODD ? i : i + 1, nb[0], g[i + 1]);
  }
}

// ------------------------------------------------------------- stream-collide
// d_StreamCollide2D (src/cuda/lbmKernel2D.cu:305-519).  MOM: write the
// density/velocity readback fields (always done in free-surface runs, where
// the free-surface passes read them).  TRACE: record (rho, m_x, m_y) of every
// cell -- or, for inlet/outlet cells, the density and velocity they read -- in
// p.trace for the adjoint sweep.
// DENSE (free surface, conservative mode): the cell lies in a row chunk
// without pass-1 work, so it is gas, wall or an interior fluid cell -- no
// interface, inlet or outlet cell -- and the neighbour flags need not be
// loaded.  Compiling that case separately keeps the kernel that runs the
// bodies of fluid to fewer registers.
// BODY (with DENSE): the cell's tile and its halo are plain fluid
// (Params::tile_body) and the update kernel will not visit it, so the cell
// also does the update's body-of-fluid step: mass = rho, its excess (the
// deviation) shared by eight receivers into excess_out, flag kept.  Its flag
// is known (plain fluid, no inlet/outlet neighbour role: such tiles are never
// bodies), so the wavefront goes straight to the population loads.
template <class S, int ODD, bool FS, bool MOM, bool TRACE, bool DENSE = false, bool BODY = false>
__device__ __forceinline__ void stream_collide_cell(const Params& p, const Cell c) {
  constexpr bool body = DENSE && BODY;
  using R = typename S::R;
  R* const rho_f = fld<R>(p.rho);
  R* const u_f = fld<R>(p.u);
  R* const prho0 = fld<R>(p.pub_rho[ODD]);
  R* const prho1 = fld<R>(p.pub_rho[ODD ^ 1]);
  R* const pu0 = fld<R>(p.pub_u[ODD]);
  R* const pu1 = fld<R>(p.pub_u[ODD ^ 1]);
  const uint32_t n = c.n;
  const uint64_t N = p.N;
  // Density of cell n is read next step by an inlet at n-1, velocity by an
  // outlet at n+1 (lbmKernel2D.cu:336,344-345); both roles are static.  The
  // three flags are loaded unconditionally (a missing neighbour reads the cell
  // itself) so that they share one memory round trip instead of three
  // branch-guarded ones.
  const Linear lin = linear_neighbours(p, c.x, c.y, n);
  // Single phase with 16-bit storage: the populations are loaded together
  // with the flags (also for wall cells, whose slots are in bounds and simply
  // not used), so that a wavefront waits for memory once before its collision
  // instead of twice.  Measured on the 8192 x 2048 culvert: FP16S / BF16S
  // +5-7 %, FP32 -5 % (its wall cells' extra 36 bytes cost more than the
  // round trip), so FP32 and FP64 keep the flag test first.
  constexpr bool kEarly = !FS && S::kBytes == 2;
  uint32_t nb[9];
  R g[9];
  if (kEarly) {
    neighbours(p, c.x, c.y, nb);
    load_incoming<S, ODD>(p, nb, g);
  }
  const uint8_t fl = body ? uint8_t(kFluid) : p.flags[n];
  const uint8_t fl_prev = body ? uint8_t(kFluid) : p.flags[lin.prev];
  const uint8_t fl_next = body ? uint8_t(kFluid) : p.flags[lin.next];
  const bool pub_r = !body && lin.has_prev && (fl_prev & kInlet);
  const bool pub_v = !body && lin.has_next && (fl_next & kOutlet) && !(fl & kInlet);
  if (fl & (kGas | kWall)) {
    if (FS) {  // a source cell that is gas this step keeps its last value
      if (pub_r) at(prho0, n) = at(prho1, n);
      if (pub_v) {
        at(pu0, n) = at(pu1, n);
        at(pu0 + N, n) = at(pu1 + N, n);
      }
    }
    return;
  }
  if (!kEarly) {
    neighbours(p, c.x, c.y, nb);
    load_incoming<S, ODD>(p, nb, g);
  }
  // Free surface: every load that depends only on the own flag is issued here,
  // with the populations, so that the cell waits for memory once.
  uint8_t nf[9];
  R ex[9], rc[9], m_if = R(0), resid = R(0);
  const bool cons = FS && (DENSE || p.mass_mode != 0);
  const bool fluid_dev = cons && (DENSE || (fl & (kFGI | kInlet | kOutlet)) == kFluid);
  const bool iface = FS && !DENSE && (fl & kFGI) == kInterface;
  // 16-bit storage rounds every stored population; that rounding changes the
  // population mass, which the interior-fluid shortcut below would take over.
  constexpr bool kRounded = S::kBytes < 4;
  if (FS && !DENSE) neighbour_flags(p, nb, nf);
  if (fluid_dev) {
#pragma unroll
    for (int i = 1; i < 9; i++) ex[i] = at(fld<R>(p.excess), nb[i]);
    if (kRounded) resid = resid_load(p, n);  // plane 0: last step's residual
  }
  if (iface) {
    m_if = at(fld<R>(p.mass), n);
#pragma unroll
    for (int k = 1; k < 9; k++) rc[k] = at(fld<R>(p.rec) + k * N, n);
    // Reconstructed gas populations of free-surface pass 1 replace incoming
    // ones (storeReconstructedGasf, cellCalculations.cuh:250-259).
#pragma unroll
    for (int k = 1; k < 9; k++) {
      const int opp = (k & 1) ? k + 1 : k - 1;
      const uint8_t gate = p.gas_check_corrected ? nf[opp] : nf[k];
      if (gate == kGas) g[k] = rc[k];
    }
  }
  const uint8_t io = DENSE ? 0 : fl & (kInlet | kOutlet);
  R rho, rho1, ux, uy, mx = R(0), my = R(0);
  if (io & kInlet) {
    rho = at(prho1, lin.next);
    rho1 = rho - R(1);
    ux = at(u_f, n);
    uy = at(u_f + N, n);
    if (TRACE) {
      fld<R>(p.trace)[n] = rho;
      fld<R>(p.trace)[N + n] = ux;
      fld<R>(p.trace)[2 * N + n] = uy;
    }
  } else if (io & kOutlet) {
    rho = at(rho_f, n);
    rho1 = rho - R(1);
    ux = at(pu1, lin.prev);
    uy = at(pu1 + N, lin.prev);
    if (TRACE) {
      fld<R>(p.trace)[n] = rho;
      fld<R>(p.trace)[N + n] = ux;
      fld<R>(p.trace)[2 * N + n] = uy;
    }
  } else {
    rho1 = g[0];
#pragma unroll
    for (int i = 1; i < 9; i++) rho1 += g[i];
    rho = R(1) + rho1;
    mx = g[1] - g[2] + g[5] - g[6] + g[7] - g[8];
    my = g[3] - g[4] + g[5] - g[6] + g[8] - g[7];
    if (TRACE) {
      fld<R>(p.trace)[n] = rho;
      fld<R>(p.trace)[N + n] = mx;
      fld<R>(p.trace)[2 * N + n] = my;
    }
  }
  const Quotient<R> over_rho(rho);
  if (!io) {
    ux = over_rho(mx);
    uy = over_rho(my);
  }
  // A fluid cell whose eight neighbours are fluid or interface (general path).
  const bool inner = (FS && !DENSE) ? interior_fluid(fl, nf) : true;
  if (fluid_dev && (DENSE || inner)) {
    // Deviation m - rho that link-by-link accounting would give: the pulled
    // shares, plus the rounding of rho itself, plus (16-bit storage) the mass
    // the previous step's stores rounded away.
    R dev;
    if (kRounded) {
      // 16-bit storage, in float: the FP64 sums below were a third of the
      // vector instructions of a body-of-fluid cell, and the 16-bit kernel is
      // bound by instruction issue.  rho - 1 is exact for rho in [0.5, 2], so
      // rho1 - (rho - 1) is exactly the rounding of 1 + rho1 (Fast2Sum) -- the
      // term the FP64 sum is there for.  Left out are the roundings of the float
      // sums of rho1 and of the shares (~1e-10), far below the binary16
      // residual's own (~4e-9).
      R sx = ex[1];
#pragma unroll
      for (int i = 2; i < 9; i++) sx += ex[i];
      dev = sx + (rho1 - (rho - R(1))) + resid;
    } else {
      double m = 0.0;
#pragma unroll
      for (int i = 1; i < 9; i++) m += double(ex[i]);
      double sg = 0.0;
#pragma unroll
      for (int i = 0; i < 9; i++) sg += double(g[i]);
      m += (1.0 + sg) - double(rho);
      dev = R(m);
    }
    if (body) {
      at(fld<R>(p.mass), n) = rho;
      at(fld<R>(p.excess_out), n) = dev / R(8.0f);
      p.flags_next[n] = fl;
    } else {
      at(fld<R>(p.mass), n) = dev;
    }
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
  R fi[9];
  bool forced = false;
  if (!FS && p.alpha && !io) {  // Brinkman drag is offered for monofluid handles only
    // Brinkman drag G = F - alpha rho u with the Guo half-force shift solved
    // for u (no reference counterpart; FP32 storage only).
    const R a = R(p.alpha[n]);
    const R gx0 = R(p.use_force ? p.fx : 0.0f), gy0 = R(p.use_force ? p.fy : 0.0f);
    const R rhol2 = over_rho(R(0.5f));
    const R s = R(1) / (R(1) + R(0.5f) * a);
    ux = (ux + rhol2 * gx0) * s;
    uy = (uy + rhol2 * gy0) * s;
    guo_terms<R>(ux, uy, float(gx0 - a * rho * ux), float(gy0 - a * rho * uy), fi);
    forced = true;
  } else if (p.use_force) {
    const R rhol2 = over_rho(R(0.5f));
    ux = ux + rhol2 * R(p.fx);
    uy = uy + rhol2 * R(p.fy);
    guo_terms<R>(ux, uy, p.fx, p.fy, fi);
    forced = true;
  }
  if (MOM || FS) {
    // Free surface: every cell's density is read by pass 4, but the velocity
    // of a fluid cell only at readback -- pass 1 reads the velocity of
    // inlet/outlet interface cells, pass 3 averages the fluid and interface
    // neighbours of a gas->interface cell, which are interface cells as long
    // as no fluid cell borders gas.  The initial flags can hold such a fluid
    // cell where water touches an open periodic edge (the seeding does not
    // wrap), so every fluid cell that is not interior (general path only)
    // stores its velocity too.  The density of a body-of-fluid cell is read
    // only at readback: the cell sets its mass itself, and no gas cell
    // borders it.
    const bool want_u = MOM || (!DENSE && ((fl & kFGI) != kFluid || !inner));
    if (!io) {
      if (want_u) {
        at(u_f, n) = ux;
        at(u_f + N, n) = uy;
      }
    } else if (io & kOutlet) {
      at(u_f, n) = ux;
      at(u_f + N, n) = uy;
    }
    if (MOM || !body) at(rho_f, n) = rho;
  }
  if (pub_r) at(prho0, n) = rho;
  if (pub_v) {
    at(pu0, n) = ux;
    at(pu0 + N, n) = uy;
  }
  R fe[9];
  equilibrium_shifted<R>(rho, rho1, ux, uy, fe);
  if (FS && (DENSE || p.mass_mode != 0)) {
    // The float weights sum to 1 + 7.45e-9, so every collision adds
    // omega * 7.45e-9 * (rho - 1) of population mass; interior fluid cells take
    // their mass from the populations, which turned this into a drift of
    // 3.7e-7 per 10^4 steps on a 512 x 128 dam break.  The rest population
    // absorbs the weight error instead: sum_i fe_i = rho - 1.
    fe[0] = rho1 - (((fe[1] + fe[2]) + (fe[3] + fe[4])) + ((fe[5] + fe[6]) + (fe[7] + fe[8])));
  }
  // 1/tau is a float constant in the reference also for PRECISION = 64, and
  // the Smagorinsky correction goes through float sqrtf (:414-442).
  float inv_tau = p.inv_tau;
  if (p.smagorinsky) {
    R hxx = R(0), hyy = R(0), hxy = R(0);
#pragma unroll
    for (int i = 1; i < 9; i++) {
      const R d = g[i] - fe[i];
      hxx += R(cx(i) * cx(i)) * d;
      hyy += R(cy(i) * cy(i)) * d;
      hxy += R(cx(i) * cy(i)) * d;
    }
    const R Q = hxx * hxx + hyy * hyy + R(2.0f) * hxy * hxy;
    const float tau = 1.0f / inv_tau;
    inv_tau = 2.0f / (tau + sqrtf(float(R(tau * tau) + R(0.54037965f * sqrtf(float(Q))) / rho)));
  }
  if (io) {
#pragma unroll
    for (int i = 0; i < 9; i++) g[i] = fe[i];
  } else if (p.collision == 0) {  // SRT
    const float omc = 1.0f - inv_tau;
    if (forced) {
      const float tl = 1.0f - 0.5f * inv_tau;
#pragma unroll
      for (int i = 0; i < 9; i++) g[i] = R(omc) * g[i] + R(inv_tau) * fe[i] + R(tl) * fi[i];
    } else {
#pragma unroll
      for (int i = 0; i < 9; i++) g[i] = R(omc) * g[i] + R(inv_tau) * fe[i];
    }
  } else if (!forced || p.collision == 2) {
    // TRT, magic parameter 1/4 (:475-510).  As written a forced TRT step only
    // streams; collision 2 (no reference counterpart) defines the forced
    // branch: Guo terms split into the pair-symmetric part, weighted 1 - w+/2,
    // and the antisymmetric part, weighted 1 - w-/2 (SRT + Guo when w+ = w-).
    // The reference keeps these temporaries in float also for PRECISION = 64.
    const float ip = inv_tau;
    const float in = 1.0f / (0.25f / (1.0f / inv_tau - 0.5f) + 0.5f);
    g[0] = g[0] - R(ip) * (g[0] - fe[0]);
#pragma unroll
    for (int i = 1; i < 9; i += 2) {
      const float sym = float(g[i] - fe[i] + g[i + 1] - fe[i + 1]);
      const float fa = float(g[i] - g[i + 1]);
      const float fea = float(fe[i] - fe[i + 1]);
      g[i] = g[i] - R(0.5f * ip * sym) - R(0.5f * in * (fa - fea));
      g[i + 1] = g[i + 1] - R(0.5f * ip * sym) - R(0.5f * in * (fea - fa));
    }
    if (forced) {
      const R ap = R(1.0f - 0.5f * ip), an = R(1.0f - 0.5f * in);
      g[0] = g[0] + ap * fi[0];
#pragma unroll
      for (int i = 1; i < 9; i += 2) {
        const R fs = R(0.5f) * (fi[i] + fi[i + 1]), fd = R(0.5f) * (fi[i] - fi[i + 1]);
        g[i] = g[i] + (ap * fs + an * fd);
        g[i + 1] = g[i + 1] + (ap * fs - an * fd);
      }
    }
  }
  if (!kRounded && S::kBytes == 4 && cons && !(fl & kInlet)) {
    // FP32 storage: the collision's float rounding would make the stored
    // populations carry slightly more or less than rho, which the next step's
    // interior-fluid shortcut (it books rho as emitted) would lose -- a
    // systematic drift of ~4e-7 per 10^4 steps on the 4096 x 1024 dam break.
    // The rest population takes that rounding instead: the nine stored values
    // sum to rho - 1 up to one rounding of g[0] (~1e-9, unbiased).
    double so = 0.0;
#pragma unroll
    for (int i = 1; i < 9; i++) so += double(g[i]);
    g[0] = R((double(rho) - 1.0) - so);
  }
  if (kRounded && cons && !(fl & kInlet)) {
    // Residual of this step's collision and storage rounding, rho - 1 - sum of
    // the stored populations, for the next step's interior-fluid deviation
    // (float sum, as the deviation above: its rounding, ~1e-10, is far below
    // the binary16 store's).
    R so = S::round(g[0]);
#pragma unroll
    for (int i = 1; i < 9; i++) so += S::round(g[i]);
    resid_store(p, n, (rho - R(1)) - so);
  }
  store_outgoing<S, ODD>(p, nb, g);
}

}  // namespace flb
