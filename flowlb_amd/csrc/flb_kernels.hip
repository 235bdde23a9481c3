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
//   * passes 2-4 run as one kernel on an LDS tile of the flags and write the
//     new flags to a second buffer.
#include "flb_cell.h"

namespace flb {

// One block per row segment of blockDim.x cells.  (A grid of resident blocks
// looping over the tiles was measured 30 % slower on the 8192 x 2048 dam
// break: the hardware overlaps the short all-gas wavefronts of independent
// blocks better than a loop does.)  Free surface: the launch covers the window
// of Params (locate_window), and a wavefront whose row chunk has no
// stream-collide work leaves after reading its tile word.  DENSE (conservative
// mode, FP64 storage; the other formats use k_stream_collide_all): the kernel
// runs only the chunks without pass-1 work -- gas, walls and interior fluid --
// and k_stream_collide_surface the others.
template <class S, int ODD, bool FS, bool MOM, bool TRACE = false, bool DENSE = false>
__global__ void __launch_bounds__(1024) k_stream_collide(const Params p) {
  const uint32_t b = walk_block<S, ODD>(p);
  const Cell c = FS ? locate_window(p, b) : locate(p, b);
  uint32_t work = 0;
  if (FS) {
    work = chunk_work(p, c, true, DENSE);
    if (!(work & (1u << kWorkSC))) return;
    if (DENSE && (work & (1u << kWorkFS1))) return;
  }
  if (!c.valid) return;
  if (DENSE && (work & kChunkBody))
    stream_collide_cell<S, ODD, FS, MOM, TRACE, DENSE, true>(p, c);
  else
    stream_collide_cell<S, ODD, FS, MOM, TRACE, DENSE>(p, c);
}

// Conservative free-surface stream-collide in one launch over the grid: chunks
// with pass-1 work (the surface) take the general path, the others the
// low-register paths for interior fluid and bodies of fluid.  The register
// budget is capped at 72 VGPRs (7 wavefronts per SIMD): the general path
// spills a little, which costs less than a second launch for the surface
// chunks (8192 x 2048 dam break: 118 -> 114.6 us per step in FP32, 107.4 ->
// 103.1 us in FP16S; a cap of 64 VGPRs spills more and gained less).  The
// FP64 storage needs more registers and keeps the two launches.
template <class S, int ODD, bool MOM>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(7)))
k_stream_collide_all(const Params p) {
  const Cell c = locate_window(p, walk_block<S, ODD>(p));
  const uint32_t work = chunk_work(p, c, true, true);
  if (!(work & (1u << kWorkSC)) || !c.valid) return;
  if (work & (1u << kWorkFS1))
    stream_collide_cell<S, ODD, true, MOM, false, false>(p, c);
  else if (work & kChunkBody)
    stream_collide_cell<S, ODD, true, MOM, false, true, true>(p, c);
  else
    stream_collide_cell<S, ODD, true, MOM, false, true>(p, c);
}

// ------------------------------------------------------------- free surface
// d_FreeSurface2D_1 (src/cuda/lbmKernel2D.cu:13-121): mass exchange and gas
// reconstruction for interface cells.
template <class S, int ODD>
__device__ __forceinline__ void fs1_cell(const Params& p, const Cell c) {
  using R = typename S::R;
  R* const mass = fld<R>(p.mass);
  R* const phi = fld<R>(p.phi);
  const uint32_t n = c.n;
  const uint64_t N = p.N;
  // Every value a fluid or interface cell may need (flags, the 17 population
  // slots, the neighbours' excess shares and fill levels, the own mass) is
  // loaded together with the own flag, so that a cell waits for memory once;
  // gas, wall and inlet cells then leave, and their loads (all in bounds) go
  // unused.
  uint32_t nb[9];
  neighbours(p, c.x, c.y, nb);
  const uint8_t fl = p.flags[n];
  uint8_t nf[9];
  neighbour_flags(p, nb, nf);
  R in[9], out[9], ex[9], ph[9];
  load_incoming<S, ODD>(p, nb, in);
  out[0] = in[0];
  load_outgoing<S, ODD>(p, nb, out);
#pragma unroll
  for (int i = 1; i < 9; i++) ex[i] = at(fld<R>(p.excess), nb[i]);
#pragma unroll
  for (int i = 1; i < 9; i++) ph[i] = at(phi, nb[i]);
  const R phi_n = at(phi, n);
  R m = at(mass, n);
  const bool cons = p.mass_mode != 0;
  // Conservative mode: the mass of a fluid cell whose eight neighbours are all
  // fluid or interface is updated by stream-collide instead (interior_fluid).
  // (One test after all loads: a separate test of the own flag first would let
  // the compiler move every load behind it, one more wait for memory.)
  const bool idle = ((fl & (kGas | kWall | kInlet)) != 0) | (cons & interior_fluid(fl, nf));
  if (idle) return;
  // As written the increments are added to the mass one by one in FP32
  // (lbmKernel2D.cu:38-52, 94-106).  The conservative mode sums them in FP64
  // and rounds once, so excess shares far below the rounding step of the
  // mass are not lost (FP32 one-by-one drifts ~1e-5 relative per 10^4 steps).
  double dm = 0.0;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    if (cons) dm += double(ex[i]);
    else m += ex[i];
  }
  const uint8_t e = fl & kFGI;
  if (e == kFluid) {
    if (!cons) {
#pragma unroll
      for (int i = 1; i < 9; i += 2) {
        m += in[i + 1] - out[i];
        m += in[i] - out[i + 1];
      }
    } else {
#pragma unroll
      for (int i = 1; i < 9; i += 2) {
        const uint8_t a = nf[i], b = nf[i + 1];
        if (a == kFluid || a == kInterface) dm += double(in[i + 1] - out[i]);
        if (b == kFluid || b == kInterface) dm += double(in[i] - out[i + 1]);
      }
      // Deviation from the density stream-collide computes from the same
      // incoming populations in the same order (an outlet keeps its density).
      R rho;
      if (fl & kOutlet) {
        rho = at(fld<R>(p.rho), n);
      } else {
        R s = in[0];
#pragma unroll
        for (int i = 1; i < 9; i++) s += in[i];
        rho = R(1) + s;
      }
      at(mass, n) = R(dm + (double(m) - double(rho)));
      return;
    }
  } else if (e == kInterface) {
    R rho, ux, uy, mx = R(0), my = R(0);
    const bool io = fl & (kInlet | kOutlet);
    if (io) {
      rho = at(fld<R>(p.rho), n);
      ux = at(fld<R>(p.u), n);
      uy = at(fld<R>(p.u) + N, n);
    } else {
      R s = out[0];
#pragma unroll
      for (int i = 1; i < 9; i++) s += out[i];
      rho = R(1) + s;
      mx = out[1] - out[2] + out[5] - out[6] + out[7] - out[8];
      my = out[3] - out[4] + out[5] - out[6] + out[8] - out[7];
    }
    const Quotient<R> over_rho(rho);
    if (!io) {
      ux = over_rho(mx);
      uy = over_rho(my);
    }
    const R phi0 = cons ? phi_n : fill_level<R>(m, rho);
    const R rhol2 = over_rho(R(0.5f));
    ux = ux + rhol2 * R(p.fx);
    uy = uy + rhol2 * R(p.fy);
    R fe[9];
    equilibrium_shifted<R>(R(1), R(0), ux, uy, fe);  // atmospheric equilibrium, rho = 1
#pragma unroll
    for (int i = 1; i < 9; i += 2) {
      R d = R(0);
      bool any = true;
      if (nf[i] == kFluid) d = in[i + 1] - out[i];
      else if (nf[i] == kInterface) d = (in[i + 1] - out[i]) * R(0.5f) * (ph[i] + phi0);
      else any = false;
      if (any) {
        if (cons) dm += double(d);
        else m += d;
      }
      any = true;
      if (nf[i + 1] == kFluid) d = in[i] - out[i + 1];
      else if (nf[i + 1] == kInterface) d = (in[i] - out[i + 1]) * R(0.5f) * (ph[i + 1] + phi0);
      else any = false;
      if (any) {
        if (cons) dm += double(d);
        else m += d;
      }
    }
#pragma unroll
    for (int k = 1; k < 9; k++) {
      const int opp = (k & 1) ? k + 1 : k - 1;
      const uint8_t gate = p.gas_check_corrected ? nf[opp] : nf[k];
      if (gate == kGas) at(fld<R>(p.rec) + k * N, n) = fe[k] - out[opp] + fe[opp];
    }
  }
  if (cons) m = R(double(m) + dm);
  at(mass, n) = m;
}


// First entry a list walker takes, loaded together with the list length (the
// index is kept inside the list's capacity, so the load is in bounds whatever
// the length; the value is used only if i < length): the walker then waits for
// memory once before its first item instead of twice.
__device__ __forceinline__ uint32_t first_entry(const uint32_t* list, uint32_t i, uint32_t cap) {
  return at(list, i < cap ? i : cap - 1u);
}

// Flag transition a cell applies to itself in pass 4.
__device__ __forceinline__ uint8_t settle(uint8_t fl) {
  if (fl & (kWall | kMovingWall | kInlet)) return fl;
  const uint8_t e = fl & (kFGI | kGasInterface);
  if (e == kIF || e == kIG) return fl & uint8_t(~kInterface);
  if (e == kGasInterface) return uint8_t((fl & ~kGasInterface) | kInterface);
  return fl;
}

// ---------------------------------------------------------- passes 2, 3, 4
// d_FreeSurface2D_2, _3 and _4 (src/cuda/lbmKernel2D.cu:123-303) fused into one
// launch.  The flags a cell sees in pass 4 depend on the flags within three
// cells of it after stream-collide (pass 2 looks one cell further than pass 3,
// pass 3 one further than pass 4), so each block loads its tile of flags with
// a three-cell halo into LDS once and evaluates
//   pass 2 on the tile + 2 rings:  gas / interface->gas cell next to an
//          interface->fluid cell -> gas->interface / interface (gather form);
//   pass 3 on the tile + 1 ring:   fluid / interface->fluid cell next to an
//          interface->gas cell -> interface (gather form);
//   pass 4 on the tile:            transitions, mass clamp, excess split;
// plus the gas->interface initialisation of pass 3 (populations from the
// neighbour averages) for its own cells.  Every cell is evaluated from the
// same input flags by every block that holds it, so blocks agree without
// communicating, and the result equals running the three passes as separate
// kernels.  The new flags go to the second flag buffer.
constexpr int kTileW = kFsTileW, kTileH = kFsTileH, kHalo = 3;
constexpr int kLdsW = kTileW + 2 * kHalo, kLdsH = kTileH + 2 * kHalo;
constexpr int kUpdateThreads = kTileW * kTileH;

// Periodic index for the halo, v in [-kHalo, n + kTileW + kHalo): a select
// when n covers that range, the remainder for domains narrower or lower than
// a tile plus halo (SMALL, chosen on the host so that the common kernel has no
// branch between its loads).
template <bool SMALL>
__device__ __forceinline__ int wrap(int v, int n) {
  if (SMALL) {
    const int w = v % n;
    return w < 0 ? w + n : w;
  }
  return v < 0 ? v + n : (v >= n ? v - n : v);
}

// LDS offsets of the eight neighbours, in the order of neighbours().
__device__ __forceinline__ int lds_offset(int i) {
  switch (i) {
    case 1: return 1;
    case 2: return -1;
    case 3: return kLdsW;
    case 4: return -kLdsW;
    case 5: return kLdsW + 1;
    case 6: return -kLdsW - 1;
    case 7: return -kLdsW + 1;
    default: return kLdsW - 1;
  }
}

// Pass 3's gas->interface initialisation and pass 4 for one cell of the tile:
// f0 its flag after stream-collide, f2s[j] its flag after pass 3 (neighbours
// at lds_offset), m its mass.
// Receiver bits of a cell for its neighbours' excess shares, from its flag
// after passes 2 and 3 (computed once per cell of the tile and its ring in
// pass 3; each cell of the tile then adds up its eight neighbours' bits):
// kRcvWritten -- a receiver as written (fluid, interface, interface->fluid or
// gas->interface before the transitions, lbmKernel2D.cu:285-292);
// kRcvCons -- a receiver in the conservative mode (not gas, wall or inlet
// after its own transition); kRcvFI -- fluid or interface after it.
constexpr uint8_t kRcvCons = 1, kRcvFI = 2, kRcvWritten = 4;
__device__ __forceinline__ uint8_t receiver_bits(uint8_t q) {
  const uint8_t r = q & (kIF | kGasInterface);
  const bool written = r == kFluid || r == kInterface || r == kIF || r == kGasInterface;
  const uint8_t sq = settle(q);
  const bool cons = (sq & (kGas | kWall | kInlet)) == 0;
  const bool fi = (sq & kFGI) == kFluid || (sq & kFGI) == kInterface;
  return uint8_t((written ? kRcvWritten : 0) | (cons ? kRcvCons : 0) | (fi ? kRcvFI : 0));
}

template <class S, int ODD>
__device__ __forceinline__ uint8_t update_cell(const Params& p, const uint8_t* f2s,
                                               const uint8_t* rcv, int j,
                                               uint8_t f0, typename S::R m, typename S::R rho_pre,
                                               uint32_t n, int x, int y, bool& interior) {
  using R = typename S::R;
  R* const mass = fld<R>(p.mass);
  const uint8_t fl = f2s[j];
  interior = false;
  if (fl & (kWall | kMovingWall | kInlet)) {
    p.flags_next[n] = fl;
    return fl;
  }
  const bool cons = p.mass_mode != 0;
  const uint8_t e = fl & (kFGI | kGasInterface);
  if (e == kGas && m == 0.0f) {
    // Empty gas cell.  A cell that turned gas in the previous step left an
    // excess share for its neighbours and marked its mass -0.0; clear both.
    p.flags_next[n] = fl;
    if (signbit(m)) {  // both share buffers: this step's readers are done
      at(mass, n) = R(0);
      at(fld<R>(p.excess), n) = R(0);
      at(fld<R>(p.excess_out), n) = R(0);
    }
    return fl;
  }
  const R rho = rho_pre;  // loaded before the flag passes (nothing here writes rho)
  if (e == kGasInterface) {
    // Pass 3: calculateAverageValuesNeighbors2D (cellCalculations.cuh:124-145)
    // over the fluid/interface neighbours, accumulated in float as the
    // reference does for both precisions, then feq with the Guo shift.
    uint32_t nb[9];
    neighbours(p, uint32_t(x), uint32_t(y), nb);
    const uint64_t N = p.N;
    // The 24 loads are issued before any neighbour's flag is tested (a load
    // under the test would be waited for before the next neighbour's).
    float nr[9], nu[9], nv[9];
#pragma unroll
    for (int k = 1; k < 9; k++) {
      nr[k] = float(at(fld<R>(p.rho), nb[k]));
      nu[k] = float(at(fld<R>(p.u), nb[k]));
      nv[k] = float(at(fld<R>(p.u) + N, nb[k]));
    }
    float cnt = 0.0f, sr = 0.0f, sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int k = 1; k < 9; k++) {
      if (f2s[j + lds_offset(k)] & kIF) {
        cnt += 1.0f;
        sr += nr[k];
        sx += nu[k];
        sy += nv[k];
      }
    }
    const R ar = R(cnt > 0.0f ? sr / cnt : 0.0f);
    R ux = R(cnt > 0.0f ? sx / cnt : 0.0f);
    R uy = R(cnt > 0.0f ? sy / cnt : 0.0f);
    const R rhol2 = R(0.5f) / ar;
    ux = ux + rhol2 * R(p.fx);
    uy = uy + rhol2 * R(p.fy);
    R fe[9];
    equilibrium_shifted<R>(ar, ar - R(1), ux, uy, fe);
    store_outgoing<S, ODD, false>(p, nb, fe);  // the update is short of scalar registers
  }
  // A fluid cell carries its mass as the deviation from rho in the
  // conservative mode (interior_fluid); one that pass 3 made interface takes
  // its absolute mass back.
  const bool dev = cons && (f0 & (kFGI | kInlet)) == kFluid;
  if (dev && e == kInterface) m = m + rho;
  R ex = R(0), ph = R(0);
  bool write_phi = true;
  uint8_t nfl = fl;
  switch (e) {
    case kFluid:
      ex = dev ? m : m - rho;
      m = rho;
      ph = R(1);
      write_phi = false;  // a cell that is fluid here was fluid after the last step
      break;
    case kGas:
      ex = m;
      m = -R(0);
      break;
    case kIF:
      nfl = fl & uint8_t(~kInterface);
      ex = m - rho;
      m = rho;
      ph = R(1);
      break;
    case kIG:
      nfl = fl & uint8_t(~kInterface);
      ex = m;
      m = -R(0);
      break;
    case kInterface:
    case kGasInterface:
      if (e == kGasInterface) nfl = uint8_t((fl & ~kGasInterface) | kInterface);
      if (m > rho) {
        ex = m - rho;
        m = rho;
      } else if (m < R(0)) {
        ex = m;
        m = R(0);
      }
      ph = fill_level<R>(m, rho);
      break;
    default:
      break;
  }
  // Receivers of the excess (receiver_bits).  As written: neighbours fluid,
  // interface, interface->fluid or gas->interface before the transitions
  // (lbmKernel2D.cu:285-292).  Conservative: the cells that pull it in the
  // next pass 1, i.e. not gas, wall or inlet after their own transition.
  // The same settled flags tell whether all eight neighbours are fluid or
  // interface next step (interior_fluid of the next pass 1).
  int cont = 0;
  uint8_t all = 0xff;
#pragma unroll
  for (int k = 1; k < 9; k++) {
    const uint8_t q = rcv[j + lds_offset(k)];
    cont += (q >> (cons ? 0 : 2)) & 1;
    all &= q;
  }
  const bool all_fi = (all & kRcvFI) != 0;
  interior = cons && all_fi;
  if (cont > 0) ex = ex / R(float(cont));
  else {
    m += ex;
    ex = R(0);
  }
  at(mass, n) = m;
  at(fld<R>(p.excess_out), n) = ex;
  if (write_phi) at(fld<R>(p.phi), n) = ph;
  p.flags_next[n] = nfl;
  return nfl;
}

// End of a tile in the update kernel.  One wavefront per tile row computed
// whether its row chunk has work next step for stream-collide (a cell that is
// not gas or wall) and pass 1 (such a cell that is not, in the conservative
// mode, an interior fluid cell; inlets are included although pass 1 passes
// them by, so that a chunk without pass-1 work holds only interior fluid
// cells besides gas and walls, which stream-collide relies on).  Thread 0 stores the tile's
// work word (Params::tile_work) if it changed and appends the chunks with
// pass-1 work to the next step's pass-1 list.  Block-wide: every thread of
// the block calls it.
__device__ __forceinline__ void finish_tile(const Params& p, int tid, uint32_t tx, uint32_t ty,
                                            uint32_t old, bool sc, bool fs1, bool nonfluid,
                                            uint32_t* bits, uint32_t quiet) {
  const bool any_sc = __any(sc), any_fs1 = __any(fs1), any_nf = __any(nonfluid);
  const int r = tid / 64;
  if ((tid & 63) == 0 && (any_sc || any_fs1 || any_nf))
    atomicOr(bits, (any_sc ? 1u << (kWorkSC + r) : 0u) | (any_fs1 ? 1u << (kWorkFS1 + r) : 0u) |
                       (any_nf ? kWorkFluid : 0u));
  __syncthreads();
  if (tid == 0) {
    const uint32_t w = (*bits ^ kWorkFluid) | quiet;  // the fluid bit: no wavefront saw another cell
    if (w != old) at(p.tile_work, ty * p.work_x + tx) = w;
    const uint32_t rows_fs1 = (w >> kWorkFS1) & 0xffu;
    if (rows_fs1) {
      const uint32_t chunk0 = ty * kFsTileH * p.work_x + tx;
      uint32_t k = atomicAdd(&p.counts_out[kListFS1], uint32_t(__popc(rows_fs1)));
      for (uint32_t m = rows_fs1; m; m &= m - 1)
        at(p.fs1_list_out, k++) = chunk0 + uint32_t(__ffs(m) - 1) * p.work_x;
    }
  }
}

// Passes 2-4 of one tile (one kFsTileW x kFsTileH block).
template <class S, int ODD, bool SMALL>
__device__ __forceinline__ void update_tile(const Params& p, uint32_t tx, uint32_t ty,
                                            uint8_t* f0s, uint8_t* f1s, uint8_t* f2s, uint8_t* rcv,
                                            uint32_t* work_bits, uint32_t* busy_waves) {
  using R = typename S::R;
  R* const mass = fld<R>(p.mass);
  const int nx = int(p.nx), ny = int(p.ny), pitch = int(p.pitch);
  const int x0 = int(tx) * kTileW, y0 = int(ty) * kTileH;
  const int tid = int(threadIdx.x);
  const int x = x0 + (tid & (kTileW - 1)), y = y0 + tid / kTileW;
  const bool valid = x < nx && y < ny;
  const uint32_t n = valid ? uint32_t(x + y * pitch) : 0u;
  const uint32_t old_work = at(p.tile_work, ty * p.work_x + tx);
  if (tid == 0) *work_bits = 0;  // ordered before its use by the next barrier
  // All global loads of the tile are issued before any is used: the two halo
  // bytes of this thread, its own flag, mass and density (read for every
  // cell, though only non-gas cells and gas cells turning interface use it).
  // Loads are unconditional with selected indices (lanes without a use read
  // cell 0, one cached line): a load under a divergent branch makes the
  // compiler wait for it where the branch merges.
  static_assert(kLdsH * kLdsW <= 2 * kUpdateThreads, "two halo bytes per thread");
  const int i0 = tid, i1 = tid + kUpdateThreads;
  const bool has1 = i1 < kLdsH * kLdsW;
  const int i1c = has1 ? i1 : i0;
  const int r0 = i0 / kLdsW, c0 = i0 - r0 * kLdsW;
  const int r1 = i1c / kLdsW, c1 = i1c - r1 * kLdsW;
  const uint8_t own = p.flags[n];
  R m = at(mass, n);
  const uint8_t h0 =
      p.flags[wrap<SMALL>(x0 + c0 - kHalo, nx) + wrap<SMALL>(y0 + r0 - kHalo, ny) * pitch];
  const uint8_t h1 =
      p.flags[wrap<SMALL>(x0 + c1 - kHalo, nx) + wrap<SMALL>(y0 + r1 - kHalo, ny) * pitch];
  const R rho_pre = at(fld<R>(p.rho), n);
  // Keeps the loads together: the scheduler would otherwise place the test of
  // the mass between them and wait for it before issuing the halo loads (the
  // empty statement takes the mass as input, after every load of the tile).
  asm volatile("" : "+v"(m) : : "memory");
  f0s[i0] = h0;
  f0s[i1] = h1;
  // A flag changes in passes 2-4 only within three cells of a fluid or
  // interface cell, and only a gas cell holding mass needs more than its flag
  // copied.  Tiles without either skip the passes.
  constexpr uint8_t kWet = kFluid | kInterface | kGasInterface;
  const bool own_plain = !valid || (own & (kWall | kMovingWall | kInlet)) ||
                         ((own & (kFGI | kGasInterface)) == kGas && m == R(0));
  const bool busy = (h0 & kWet) || (has1 && (h1 & kWet)) || !own_plain;
  // Inside a body of fluid (every flag of the tile and its halo plain fluid)
  // passes 2 and 3 change nothing and every cell has eight receivers.
  const bool thin = h0 != kFluid || (has1 && h1 != kFluid);
  const uint32_t wave_bits = (__any(busy) ? 1u : 0u) | (__any(thin) ? 2u : 0u);
  if ((tid & 63) == 0) reinterpret_cast<uint8_t*>(busy_waves)[tid >> 6] = uint8_t(wave_bits);
  __syncthreads();
  const uint32_t any_bits = busy_waves[0] | busy_waves[1];
  const bool quiet = (any_bits & 0x01010101u) == 0;
  const bool body = (any_bits & 0x02020202u) == 0;
  if (body) {
    // The fluid branch of update_cell with cont = 8: the excess (the
    // deviation in the conservative mode) is shared by the eight neighbours.
    const bool cons = p.mass_mode != 0;
    if (valid) {
      const R ex = cons ? m : m - rho_pre;
      at(mass, n) = rho_pre;
      at(fld<R>(p.excess_out), n) = ex / R(8.0f);
      p.flags_next[n] = own;
    }
    finish_tile(p, tid, tx, ty, old_work, valid, valid && !cons, false, work_bits, 0u);
    return;
  }
  if (quiet) {
    if (valid) {
      p.flags_next[n] = own;
      if (!(own & (kWall | kMovingWall | kInlet)) && (own & (kFGI | kGasInterface)) == kGas &&
          signbit(m)) {  // both share buffers: this step's readers are done
        at(mass, n) = R(0);
        at(fld<R>(p.excess), n) = R(0);
        at(fld<R>(p.excess_out), n) = R(0);
      }
    }
    const bool sc = valid && !(own & (kGas | kWall));
    finish_tile(p, tid, tx, ty, old_work, sc, sc, valid && own != kFluid, work_bits, kWorkQuiet);
    return;
  }
  for (int i = tid; i < (kLdsH - 2) * (kLdsW - 2); i += kUpdateThreads) {  // pass 2
    const int r = 1 + i / (kLdsW - 2), j = r * kLdsW + 1 + i % (kLdsW - 2);
    const uint8_t fl = f0s[j];
    const uint8_t e = fl & kIG;
    uint8_t o = fl;
    if (e == kIG || e == kGas) {
      bool near_if = false;
#pragma unroll
      for (int k = 1; k < 9; k++) near_if = near_if || ((f0s[j + lds_offset(k)] & kIF) == kIF);
      if (near_if) o = e == kIG ? uint8_t(fl & ~kGas) : uint8_t((fl & ~kGas) | kGasInterface);
    }
    f1s[j] = o;
  }
  __syncthreads();
  for (int i = tid; i < (kLdsH - 4) * (kLdsW - 4); i += kUpdateThreads) {  // pass 3 flags
    const int r = 2 + i / (kLdsW - 4), j = r * kLdsW + 2 + i % (kLdsW - 4);
    const uint8_t fl = f1s[j];
    const uint8_t fi = fl & (kInletFluid | kInterface);
    uint8_t o = fl;
    if (fi == kFluid || fi == kIF) {
      bool near_ig = false;
#pragma unroll
      for (int k = 1; k < 9; k++)
        near_ig = near_ig || ((f1s[j + lds_offset(k)] & (kGasInterface | kIG)) == kIG);
      if (near_ig) o = uint8_t((fl & ~kIF) | kInterface);
    }
    f2s[j] = o;
    rcv[j] = receiver_bits(o);
  }
  __syncthreads();
  bool sc = false, fs1 = false, nonfluid = false;
  if (valid) {
    const int j = (tid / kTileW + kHalo) * kLdsW + (tid & (kTileW - 1)) + kHalo;
    bool interior;
    const uint8_t fin = update_cell<S, ODD>(p, f2s, rcv, j, f0s[j], m, rho_pre, n, x, y, interior);
    sc = !(fin & (kGas | kWall));
    fs1 = sc && !(interior && (fin & (kFGI | kInlet | kOutlet)) == kFluid);
    nonfluid = fin != kFluid;
  }
  finish_tile(p, tid, tx, ty, old_work, sc, fs1, nonfluid, work_bits, 0u);
}

// One block of the update reduces the per-wavefront boxes pass 1 wrote this
// step and sends the result to the host (mapped memory): first the box, then
// -- after a system-scope fence, so that a host that reads the step and then
// the box gets a box at least as new -- the step.  Every thread of the block
// calls it.
__device__ __forceinline__ void publish_box(const Params& p, uint32_t* red) {
  const uint32_t parts = (p.work_x * p.tile_rows + 63u) / 64u;
  uint32_t ylo = ~0u, yhi = 0u, xlo = ~0u, xhi = 0u;
  for (uint32_t i = threadIdx.x; i < parts; i += blockDim.x) {
    const uint32_t* e = p.box + 4u * i;
    ylo = min(ylo, e[0]);
    yhi = max(yhi, e[1]);
    xlo = min(xlo, e[2]);
    xhi = max(xhi, e[3]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    ylo = min(ylo, uint32_t(__shfl_xor(int(ylo), o, 64)));
    yhi = max(yhi, uint32_t(__shfl_xor(int(yhi), o, 64)));
    xlo = min(xlo, uint32_t(__shfl_xor(int(xlo), o, 64)));
    xhi = max(xhi, uint32_t(__shfl_xor(int(xhi), o, 64)));
  }
  const uint32_t wave = threadIdx.x / 64u, waves = blockDim.x / 64u;
  if ((threadIdx.x & 63u) == 0) {
    red[4 * wave] = ylo;
    red[4 * wave + 1] = yhi;
    red[4 * wave + 2] = xlo;
    red[4 * wave + 3] = xhi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < waves; w++) {
      ylo = min(ylo, red[4 * w]);
      yhi = max(yhi, red[4 * w + 1]);
      xlo = min(xlo, red[4 * w + 2]);
      xhi = max(xhi, red[4 * w + 3]);
    }
    const unsigned long long packed =
        (unsigned long long)(ylo & 0xffffu) | (unsigned long long)(yhi & 0xffffu) << 16 |
        (unsigned long long)(xlo & 0xffffu) << 32 | (unsigned long long)(xhi & 0xffffu) << 48;
    p.box_host[0] = packed;
    __threadfence_system();
    p.box_host[1] = (unsigned long long)p.step;
  }
}

// The update kernel walks the tile list pass 1 built for this step.  A tile
// left off the list was quiet in its last update and its neighbour tiles held
// only gas and walls after the last step, so stream-collide and pass 1 left
// them alone, its flags and halo are unchanged, both flag buffers hold them
// and no cell of it turned gas then: it has nothing to do.  The barrier at
// the end of finish_tile orders a tile's LDS reads before the next tile's
// writes; the next tile's index is loaded while the current one is evaluated.
template <class S, int ODD, bool SMALL>
__device__ __forceinline__ void fs_update(const Params& p) {
  // f0s is padded to two bytes per thread so that every lane writes both of
  // its halo bytes unconditionally (lanes past the tile write the padding).
  __shared__ uint8_t f0s[2 * kUpdateThreads], f1s[kLdsH * kLdsW], f2s[kLdsH * kLdsW];
  __shared__ uint8_t rcv[kLdsH * kLdsW];  // receiver_bits of f2s
  __shared__ uint32_t work_bits;
  __shared__ uint32_t busy_waves[2];  // one byte per wavefront
  static_assert(kUpdateThreads / 64 <= 8, "one byte per wavefront");
  // The last block publishes the box: with fewer listed tiles than blocks it
  // has no tile, so the publication does not lengthen a tile's chain.
  if (blockIdx.x + 1 == gridDim.x) {
    __shared__ uint32_t red[4 * (kUpdateThreads / 64)];
    publish_box(p, red);
  }
  uint32_t i = blockIdx.x;
  uint32_t tile = first_entry(p.tile_list, i, p.work_x * p.tile_rows);
  const uint32_t count = *p.tile_count;
  // Keeps the first entry's load beside the length's: the compiler would
  // otherwise move it behind the length test, a second wait for memory.
  asm volatile("" ::: "memory");
  while (i < count) {
    const uint32_t j = i + gridDim.x;
    const uint32_t next = j < count ? at(p.tile_list, j) : 0u;
    const uint32_t ty = tile / p.work_x, tx = tile - ty * p.work_x;
    update_tile<S, ODD, SMALL>(p, tx, ty, f0s, f1s, f2s, rcv, &work_bits, busy_waves);
    i = j;
    tile = next;
  }
}

// The update kernel at 8 waves per SIMD (64 VGPRs; the scalar registers over
// the budget spill to vector-register lanes): four 512-thread blocks per CU
// instead of three.  On the 8192 x 2048 dam break the update went 14.3 ->
// 12.7 us in FP32 and 13.7 -> 12.1 us in FP16S.  FP64 storage would spill
// vector registers to scratch under that cap and keeps the compiler's choice.
template <class S, int ODD, bool SMALL>
__global__ void __launch_bounds__(kUpdateThreads) __attribute__((amdgpu_waves_per_eu(8)))
k_fs_update(const Params p) {
  fs_update<S, ODD, SMALL>(p);
}
template <int ODD, bool SMALL>
__global__ void __launch_bounds__(kUpdateThreads) k_fs_update_f64(const Params p) {
  fs_update<StoreF64, ODD, SMALL>(p);
}

// Pass 1 over the chunk list of this step (one wavefront per 64-cell row
// chunk), by a grid of resident blocks.  It also empties the pass-1 list the
// update kernel of this step fills and the counter of the next step's tile
// list (their last readers, the kernels of the previous step, have finished),
// and lists the tiles the update of this step must evaluate: all but those
// quiet in their last update (kWorkQuiet) with only gas and walls in the 3 x 3
// tiles around them (periodic, as the lattice) after the last step.
template <class S, int ODD>
__global__ void __launch_bounds__(kListThreads) k_fs1(const Params p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) p.counts_out[kListFS1] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 1) *p.tile_count_next = 0;
  const uint32_t lane = threadIdx.x & 63u;
  // Every 8th block classifies tiles, the others walk the chunk list (both
  // kinds spread over the dispatch order); a grid under 8 blocks does both.
  // Blocks that did both made the kernel as long as the two latency chains
  // together (8192 x 2048 dam break: pass 1 13.2 -> 11.8 us).
  const bool split = gridDim.x >= 8u;
  const uint32_t ncls = split ? gridDim.x / 8u : gridDim.x;
  const bool cls = !split || (blockIdx.x % 8u) == 7u;
  const bool chk = !split || !cls;
  const uint32_t cb = split ? blockIdx.x - (blockIdx.x + 1u) / 8u : blockIdx.x;
  const uint32_t nchk = split ? gridDim.x - ncls : gridDim.x;
  const uint32_t nw = nchk * (blockDim.x / 64u);
  uint32_t i = cb * (blockDim.x / 64u) + threadIdx.x / 64u;
  uint32_t c = first_entry(p.fs1_list, i, p.work_x * p.ny);
  const uint32_t count = chk ? p.counts_in[kListFS1] : 0u;
  for (; i < count; i += nw) {
    const uint32_t j = i + nw;
    const uint32_t next = j < count ? at(p.fs1_list, j) : 0u;
    const uint32_t y = c / p.work_x, x = (c - y * p.work_x) * kFsTileW + lane;
    const Cell cell{x, y, x + y * p.pitch, x < p.nx};
    if (cell.valid) fs1_cell<S, ODD>(p, cell);
    c = next;
  }
  if (!cls) return;
  const uint32_t tiles = p.work_x * p.tile_rows;
  const uint32_t nt = ncls * blockDim.x;
  const uint32_t kb = split ? blockIdx.x / 8u : blockIdx.x;
  // A tile's three-cell halo lies in the 3 x 3 tiles around it, taken
  // periodically as the lattice is, unless a last tile column (row) of fewer
  // than three cells lets it spill one tile further: then the first tile
  // column (row) and the one before the last are never skipped.
  const bool thin_x = p.nx - (p.work_x - 1u) * kTileW < uint32_t(kHalo) && p.work_x >= 4u;
  const bool thin_y = p.ny - (p.tile_rows - 1u) * kTileH < uint32_t(kHalo) && p.tile_rows >= 4u;
  // One tile per lane; the loop is uniform over the wavefront (ballots).  Each
  // lane loads the publisher word and the 3 x 3 work words at once, so the
  // classification waits for memory once.
  for (uint32_t t0 = kb * blockDim.x + (threadIdx.x & ~63u); t0 < tiles; t0 += nt) {
    const uint32_t t = t0 + lane;
    bool need = false, has_sc = false;
    uint32_t ty = 0, tx = 0;
    if (t < tiles) {
      ty = t / p.work_x;
      tx = t - ty * p.work_x;
      const uint32_t xm = tx > 0 ? tx - 1 : p.work_x - 1, xp = tx + 1 < p.work_x ? tx + 1 : 0;
      const uint32_t ym = ty > 0 ? ty - 1 : p.tile_rows - 1;
      const uint32_t yp = ty + 1 < p.tile_rows ? ty + 1 : 0;
      const uint32_t pub = at(p.tile_pub, t);
      uint32_t w[9];
      const uint32_t rows[3] = {ym * p.work_x, ty * p.work_x, yp * p.work_x};
#pragma unroll
      for (int k = 0; k < 3; k++) {
        w[3 * k] = at(p.tile_work, rows[k] + xm);
        w[3 * k + 1] = at(p.tile_work, rows[k] + tx);
        w[3 * k + 2] = at(p.tile_work, rows[k] + xp);
      }
      const bool spill = (thin_x && (tx == 0 || tx + 2 == p.work_x)) ||
                         (thin_y && (ty == 0 || ty + 2 == p.tile_rows));
      has_sc = ((w[4] | pub) & kWorkSCMask) != 0;
      bool skip = false, body = false;
      if (!pub && !spill) {
        uint32_t any = 0, all = ~0u;
#pragma unroll
        for (int k = 0; k < 9; k++) {
          any |= w[k];
          all &= w[k];
        }
        const uint32_t own = w[4];
        skip = (own & kWorkQuiet) && !(any & kWorkSCMask);
        // Conservative mode: the tile and its halo are plain fluid and stay
        // so through stream-collide, which then hands out the cells' excess
        // and sets their mass, as the update's body-of-fluid path would.
        body = p.mass_mode != 0 && (all & kWorkFluid) && !(own & (0xffu << kWorkFS1));
      }
      at(p.tile_body, t) = body;
      need = !skip && !body;
    }
    // Bounding box of this wavefront's tiles with stream-collide work (empty:
    // lo > hi), one entry of Params::box per 64 tiles; the update reduces them.
    {
      uint32_t ylo = has_sc ? ty : ~0u, yhi = has_sc ? ty : 0u;
      uint32_t xlo = has_sc ? tx : ~0u, xhi = has_sc ? tx : 0u;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        ylo = min(ylo, uint32_t(__shfl_xor(int(ylo), o, 64)));
        yhi = max(yhi, uint32_t(__shfl_xor(int(yhi), o, 64)));
        xlo = min(xlo, uint32_t(__shfl_xor(int(xlo), o, 64)));
        xhi = max(xhi, uint32_t(__shfl_xor(int(xhi), o, 64)));
      }
      if (lane == 0) {
        uint32_t* e = p.box + 4u * (t0 / 64u);
        e[0] = ylo;
        e[1] = yhi;
        e[2] = xlo;
        e[3] = xhi;
      }
    }
    // One atomic per wavefront appends its listed tiles.
    const uint64_t m = __ballot(need);
    uint32_t base = 0;
    if (lane == 0 && m) base = atomicAdd(p.tile_count, uint32_t(__popcll(m)));
    base = __shfl(base, 0, 64);
    if (need) at(p.tile_list, base + uint32_t(__popcll(m & ((1ull << lane) - 1ull)))) = t;
  }
}

// Free-surface stream-collide of the chunks with pass-1 work (the surface,
// inlets and outlets), from the pass-1 list, by a grid of resident blocks;
// k_stream_collide<..., DENSE> runs the rest.
template <class S, int ODD, bool MOM>
__global__ void __launch_bounds__(kListThreads) k_stream_collide_surface(const Params p) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t nw = gridDim.x * (blockDim.x / 64u);
  uint32_t i = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u;
  uint32_t c = first_entry(p.fs1_list, i, p.work_x * p.ny);
  const uint32_t count = p.counts_in[kListFS1];
  for (; i < count; i += nw) {
    const uint32_t j = i + nw;
    const uint32_t next = j < count ? p.fs1_list[j] : 0u;
    const uint32_t y = c / p.work_x, x = (c - y * p.work_x) * kFsTileW + lane;
    const Cell cell{x, y, x + y * p.pitch, x < p.nx};
    if (cell.valid) stream_collide_cell<S, ODD, true, MOM, false>(p, cell);
    c = next;
  }
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
template <class R>
__global__ void __launch_bounds__(256) k_fs_mass(const Params p, double* partial) {
  __shared__ double lds[16];
  const Cell c = locate(p);
  double v = 0.0;
  if (c.valid) {
    v = double(fld<R>(p.mass)[c.n]);
    if (!(p.flags[c.n] & (kGas | kWall | kInlet))) {
      uint32_t nb[9];
      neighbours(p, c.x, c.y, nb);
      for (int i = 1; i < 9; i++) v += double(fld<R>(p.excess)[nb[i]]);
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

// Full populations f = (f - w) + w, in the output type O (float or double),
// to a staging array of 9 planes of nx * ny cells in row-major order.
template <class S, class O>
__global__ void k_pop_out(const Params p, O* out) {
  using R = typename S::R;
  const Cell c = locate(p);
  if (!c.valid) return;
  const typename S::T* f = static_cast<const typename S::T*>(p.f);
  const uint64_t m = c.x + uint64_t(c.y) * p.nx;
  for (int q = 0; q < 9; q++) out[q * p.NL + m] = O(S::load(f + q * p.Np, c.n) + wgt_r<R>(q));
}

// in: shifted populations (f - w) of type I, 9 planes of nx * ny cells.
template <class S, class I>
__global__ void k_pop_in(const Params p, const I* in) {
  using R = typename S::R;
  const Cell c = locate(p);
  if (!c.valid) return;
  typename S::T* f = static_cast<typename S::T*>(p.f);
  const uint64_t m = c.x + uint64_t(c.y) * p.nx;
  for (int q = 0; q < 9; q++) S::store(f + q * p.Np, c.n, R(in[q * p.NL + m]));
}

// Flag histogram, grid-stride over all cells.  The lanes of a wavefront that
// hold the same flag add their count once (one LDS atomic per distinct value
// per wavefront), and each block adds its non-zero bins to the global counts
// once: a per-cell atomic on a few hot bins (gas, fluid) serialises.
__global__ void __launch_bounds__(256) k_flag_hist(const Params p, unsigned long long* counts) {
  __shared__ unsigned int h[256];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t m = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; m < p.NL; m += stride) {
    const uint64_t y = m / p.nx;
    const int v = p.flags[m + y * (p.pitch - p.nx)];  // skip the row padding
    uint64_t todo = __ballot(1);
    while (todo) {
      const int leader = __ffsll((unsigned long long)todo) - 1;
      const int lv = __shfl(v, leader, 64);
      const uint64_t same = __ballot(v == lv);
      if (lane == leader) atomicAdd(&h[lv], unsigned(__popcll(same)));
      todo &= ~same;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += blockDim.x)
    if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
}

// ------------------------------------------------------------- launch table
using KernelFn = void (*)(const Params);
using MassFn = void (*)(const Params, double*);
using ConvOutFn = void (*)(const Params, float*);
using ConvInFn = void (*)(const Params, const float*);
using ConvOut64Fn = void (*)(const Params, double*);
using ConvIn64Fn = void (*)(const Params, const double*);
using HistFn = void (*)(const Params, unsigned long long*);

template <class S>
KernelFn stream_collide_fn(int odd, bool fs, bool mom) {
  if (fs && mom) return odd ? k_stream_collide<S, 1, true, true> : k_stream_collide<S, 0, true, true>;
  if (fs) return odd ? k_stream_collide<S, 1, true, false> : k_stream_collide<S, 0, true, false>;
  if (mom) return odd ? k_stream_collide<S, 1, false, true> : k_stream_collide<S, 0, false, true>;
  return odd ? k_stream_collide<S, 1, false, false> : k_stream_collide<S, 0, false, false>;
}
template <class S>
KernelFn dense_fn(int odd, bool mom) {
  if (mom)
    return odd ? k_stream_collide<S, 1, true, true, false, true>
               : k_stream_collide<S, 0, true, true, false, true>;
  return odd ? k_stream_collide<S, 1, true, false, false, true>
             : k_stream_collide<S, 0, true, false, false, true>;
}
template <class S>
KernelFn all_fn(int odd, bool mom) {
  if (mom) return odd ? k_stream_collide_all<S, 1, true> : k_stream_collide_all<S, 0, true>;
  return odd ? k_stream_collide_all<S, 1, false> : k_stream_collide_all<S, 0, false>;
}
template <class S>
KernelFn surface_fn(int odd, bool mom) {
  if (mom) return odd ? k_stream_collide_surface<S, 1, true> : k_stream_collide_surface<S, 0, true>;
  return odd ? k_stream_collide_surface<S, 1, false> : k_stream_collide_surface<S, 0, false>;
}
template <class S>
KernelFn fs1_fn(int odd) { return odd ? k_fs1<S, 1> : k_fs1<S, 0>; }
template <class S>
KernelFn fs_update_fn(int odd, bool small) {
  if (small) return odd ? k_fs_update<S, 1, true> : k_fs_update<S, 0, true>;
  return odd ? k_fs_update<S, 1, false> : k_fs_update<S, 0, false>;
}
template <>
KernelFn fs_update_fn<StoreF64>(int odd, bool small) {
  if (small) return odd ? k_fs_update_f64<1, true> : k_fs_update_f64<0, true>;
  return odd ? k_fs_update_f64<1, false> : k_fs_update_f64<0, false>;
}

}  // namespace flb

// Host-side accessors used by flb_core.hip (same translation unit group).
// Storage codes: 0 FP32, 1 FP16S, 2 BF16S, 3 FP64.
namespace flb {
template <template <class> class Pick>
auto by_storage(int storage) {
  switch (storage) {
    case 1: return Pick<StoreF16S>::get();
    case 2: return Pick<StoreBF16S>::get();
    case 3: return Pick<StoreF64>::get();
    default: return Pick<StoreF32>::get();
  }
}
KernelFn get_stream_collide(int storage, int odd, bool fs, bool mom) {
  switch (storage) {
    case 1: return stream_collide_fn<StoreF16S>(odd, fs, mom);
    case 2: return stream_collide_fn<StoreBF16S>(odd, fs, mom);
    case 3: return stream_collide_fn<StoreF64>(odd, fs, mom);
    default: return stream_collide_fn<StoreF32>(odd, fs, mom);
  }
}
KernelFn get_stream_collide_dense(int storage, int odd, bool mom) {
  switch (storage) {
    case 1: return dense_fn<StoreF16S>(odd, mom);
    case 2: return dense_fn<StoreBF16S>(odd, mom);
    case 3: return dense_fn<StoreF64>(odd, mom);
    default: return dense_fn<StoreF32>(odd, mom);
  }
}
KernelFn get_stream_collide_all(int storage, int odd, bool mom) {
  switch (storage) {
    case 1: return all_fn<StoreF16S>(odd, mom);
    case 2: return all_fn<StoreBF16S>(odd, mom);
    case 3: return nullptr;  // FP64: dense + surface launches
    default: return all_fn<StoreF32>(odd, mom);
  }
}
KernelFn get_stream_collide_surface(int storage, int odd, bool mom) {
  switch (storage) {
    case 1: return surface_fn<StoreF16S>(odd, mom);
    case 2: return surface_fn<StoreBF16S>(odd, mom);
    case 3: return surface_fn<StoreF64>(odd, mom);
    default: return surface_fn<StoreF32>(odd, mom);
  }
}
KernelFn get_fs1(int storage, int odd) {
  switch (storage) {
    case 1: return fs1_fn<StoreF16S>(odd);
    case 2: return fs1_fn<StoreBF16S>(odd);
    case 3: return fs1_fn<StoreF64>(odd);
    default: return fs1_fn<StoreF32>(odd);
  }
}
KernelFn get_fs_update(int storage, int odd, uint32_t nx, uint32_t ny) {
  const bool small = nx < uint32_t(kTileW + 2 * kHalo) || ny < uint32_t(kTileH + 2 * kHalo);
  switch (storage) {
    case 1: return fs_update_fn<StoreF16S>(odd, small);
    case 2: return fs_update_fn<StoreBF16S>(odd, small);
    case 3: return fs_update_fn<StoreF64>(odd, small);
    default: return fs_update_fn<StoreF32>(odd, small);
  }
}
KernelFn get_stream_collide_trace(int storage, int odd, bool mom) {
  if (storage == 3) {  // FP64
    if (mom) return odd ? k_stream_collide<StoreF64, 1, false, true, true>
                        : k_stream_collide<StoreF64, 0, false, true, true>;
    return odd ? k_stream_collide<StoreF64, 1, false, false, true>
               : k_stream_collide<StoreF64, 0, false, false, true>;
  }
  if (mom) return odd ? k_stream_collide<StoreF32, 1, false, true, true>
                      : k_stream_collide<StoreF32, 0, false, true, true>;
  return odd ? k_stream_collide<StoreF32, 1, false, false, true>
             : k_stream_collide<StoreF32, 0, false, false, true>;
}

template <class S>
struct PopMass {
  static MassFn get() { return k_pop_mass<S>; }
};
template <class S>
struct FsMass {
  static MassFn get() { return k_fs_mass<typename S::R>; }
};
template <class S>
struct PopOut32 {
  static ConvOutFn get() { return k_pop_out<S, float>; }
};
template <class S>
struct PopOut64 {
  static ConvOut64Fn get() { return k_pop_out<S, double>; }
};
template <class S>
struct PopIn32 {
  static ConvInFn get() { return k_pop_in<S, float>; }
};
template <class S>
struct PopIn64 {
  static ConvIn64Fn get() { return k_pop_in<S, double>; }
};
MassFn get_pop_mass(int storage) { return by_storage<PopMass>(storage); }
MassFn get_fs_mass(int storage) { return by_storage<FsMass>(storage); }
ConvOutFn get_pop_to_f32(int storage) { return by_storage<PopOut32>(storage); }
ConvOut64Fn get_pop_to_f64(int storage) { return by_storage<PopOut64>(storage); }
ConvInFn get_f32_to_pop(int storage) { return by_storage<PopIn32>(storage); }
ConvIn64Fn get_f64_to_pop(int storage) { return by_storage<PopIn64>(storage); }
HistFn get_flag_hist() { return k_flag_hist; }
}  // namespace flb
