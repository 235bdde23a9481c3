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
#include "flb_cell.h"

namespace flb {

template <class S, int ODD, bool FS, bool MOM, bool TRACE = false>
__global__ void __launch_bounds__(1024) k_stream_collide(const Params p) {
  const Cell c = locate(p);
  if (!c.valid) return;
  stream_collide_cell<S, ODD, FS, MOM, TRACE>(p, c);
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
  // Conservative mode: the mass of a fluid cell whose eight neighbours are all
  // fluid or interface is updated by stream-collide instead (interior_fluid).
  if (p.mass_mode != 0 && interior_fluid(p, fl, nb)) return;
  float in[9], out[9];
  load_incoming<S, ODD>(p, nb, in);
  out[0] = in[0];
  load_outgoing<S, ODD>(p, nb, out);
  // As written the increments are added to the mass one by one in FP32
  // (lbmKernel2D.cu:38-52, 94-106).  The conservative mode sums them in FP64
  // and rounds once, so excess shares far below the rounding step of the
  // mass are not lost (FP32 one-by-one drifts ~1e-5 relative per 10^4 steps).
  const bool cons = p.mass_mode != 0;
  float m = p.mass[n];
  double dm = 0.0;
#pragma unroll
  for (int i = 1; i < 9; i++) {
    if (cons) dm += double(p.excess[nb[i]]);
    else m += p.excess[nb[i]];
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
        const uint8_t a = p.flags[nb[i]] & kFGI, b = p.flags[nb[i + 1]] & kFGI;
        if (a == kFluid || a == kInterface) dm += double(in[i + 1] - out[i]);
        if (b == kFluid || b == kInterface) dm += double(in[i] - out[i + 1]);
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
    const float phi0 = cons ? p.phi[n] : fill_level(m, rho);
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
      float d = 0.0f;
      bool any = true;
      if (nf[i] == kFluid) d = in[i + 1] - out[i];
      else if (nf[i] == kInterface) d = (in[i + 1] - out[i]) * 0.5f * (p.phi[nb[i]] + phi0);
      else any = false;
      if (any) {
        if (cons) dm += double(d);
        else m += d;
      }
      any = true;
      if (nf[i + 1] == kFluid) d = in[i] - out[i + 1];
      else if (nf[i + 1] == kInterface) d = (in[i] - out[i + 1]) * 0.5f * (p.phi[nb[i + 1]] + phi0);
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
      if (gate == kGas) p.rec[k * N + n] = fe[k] - out[opp] + fe[opp];
    }
  }
  if (cons) m = float(double(m) + dm);
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
  // As written the receivers are counted on the flags before this pass, so the
  // new flags go to the second buffer.  In the conservative mode the count
  // uses the settled flag of each neighbour, which is the same before and
  // after its own transition: flags are updated in place.
  const bool cons = p.mass_mode != 0;
  const uint8_t fl = p.flags[n];
  if (fl & (kWall | kMovingWall | kInlet)) {
    if (!cons) p.flags_next[n] = fl;
    return;
  }
  const uint8_t e = fl & (kFGI | kGasInterface);
  float m = p.mass[n];
  if (e == kGas && m == 0.0f) {
    // Empty gas cell.  A cell that turned gas in the previous step left an
    // excess share for its neighbours and marked its mass -0.0; clear both.
    if (!cons) p.flags_next[n] = fl;
    if (signbit(m)) {
      p.mass[n] = 0.0f;
      p.excess[n] = 0.0f;
    }
    return;
  }
  const float rho = p.rho[n];
  float ex = 0.0f, ph = 0.0f;
  bool write_phi = true;
  uint8_t nfl = fl;
  switch (e) {
    case kFluid:
      ex = m - rho;
      m = rho;
      ph = 1.0f;
      write_phi = false;  // a cell that is fluid here was fluid after the last pass
      break;
    case kGas:
      ex = m;
      m = -0.0f;
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
      m = -0.0f;
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
    if (!cons) {
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
  if (write_phi) p.phi[n] = ph;
  if (!cons) p.flags_next[n] = nfl;
  else if (nfl != fl) p.flags[n] = nfl;
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
KernelFn get_stream_collide_trace(int odd, bool mom) {
  if (mom) return odd ? k_stream_collide<StoreF32, 1, false, true, true>
                      : k_stream_collide<StoreF32, 0, false, true, true>;
  return odd ? k_stream_collide<StoreF32, 1, false, false, true>
             : k_stream_collide<StoreF32, 0, false, false, true>;
}
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
