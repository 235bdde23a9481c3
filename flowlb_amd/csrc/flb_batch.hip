// flb_batch.hip -- many independent single-phase culverts of uneven sizes in
// one kernel launch per step.  No reference counterpart: FlowLB runs one domain
// per process.
//
// Every instance keeps the single-domain layout (nine population planes, flags,
// fields, publication buffers; rows padded to a pitch of 64 cells so that each
// wavefront row segment starts on a 256-byte boundary) at its own offset of
// shared allocations.  The work of a step is a packed list of tiles --
// (instance, row, x-tile) -- over all instances, interleaved row by row.  Two
// launch modes:
//   * one launch per step, a block per tile or a grid-stride walk, runs of
//     steps replayed from hipGraphs captured for an even number of steps;
//   * persistent: one cooperative launch runs a chunk of steps, its resident
//     blocks walking the tile list and meeting at a grid barrier between
//     steps.
// The odd steps take the tile list's bands last to first (a second copy of
// the list), so that a step starts on what the Infinity Cache still holds of
// the previous one.  An autotuner times block sizes, grid sizes and both modes
// on the actual instance set and keeps the fastest.
#include <algorithm>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "flb_cell.h"
#include "flb_internal.h"

namespace flb {

struct Tile {
  uint32_t inst;
  uint32_t row;
  uint32_t tx;
  uint32_t pad;
};

// One step over `tiles`, the list in this step's order (the odd steps get the
// second copy).
template <class S, int ODD, bool MOM>
__global__ void __launch_bounds__(1024) k_batch_step(const Tile* __restrict__ tiles, uint32_t ntiles,
                                                      const Params* __restrict__ inst) {
  uint32_t k = blockIdx.x;
  if (k >= ntiles) return;
  Tile t = tiles[k];
  // A grid-stride block loads its next tile's descriptor while it evaluates
  // this one (with a block per tile there is no next: the same line again).
  for (;;) {
    const uint32_t kn = k + gridDim.x;
    const bool more = kn < ntiles;
    const Tile tn = tiles[more ? kn : k];
    const Params& p = inst[t.inst];
    const uint32_t x = t.tx * blockDim.x + threadIdx.x;
    if (x < p.nx) {
      const Cell c{x, t.row, x + t.row * p.pitch, true};
      stream_collide_cell<S, ODD, false, MOM, false>(p, c);
    }
    if (!more) break;
    k = kn;
    t = tn;
  }
}

// Persistent mode: `nsteps` steps from parity odd0 in one cooperative launch
// (every block resident, so the grid barrier between steps completes); the
// readback fields are not written (the last step of a call runs as a plain
// launch).  Every block reaches every barrier.
template <class S>
__global__ void __launch_bounds__(1024) k_batch_persistent(const Tile* __restrict__ tiles,
                                                            uint32_t ntiles,
                                                            const Params* __restrict__ inst,
                                                            uint32_t nsteps, uint32_t odd0) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  for (uint32_t s = 0; s < nsteps; ++s) {
    const bool odd = ((odd0 + s) & 1u) != 0;
    const Tile* const list = tiles + (odd ? ntiles : 0u);  // the odd steps' order
    for (uint32_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
      const Tile t = list[k];
      const Params& p = inst[t.inst];
      const uint32_t x = t.tx * blockDim.x + threadIdx.x;
      if (x < p.nx) {
        const Cell c{x, t.row, x + t.row * p.pitch, true};
        if (odd) stream_collide_cell<S, 1, false, false, false>(p, c);
        else stream_collide_cell<S, 0, false, false, false>(p, c);
      }
    }
    if (s + 1 < nsteps) grid.sync();
  }
}

// Populations of instance i to / from nx * ny host-ordered float planes
// (unshifted), through the pitched device planes.
template <class S>
__global__ void k_batch_to_f32(const Params* inst, uint32_t i, float* out) {
  const Params& p = inst[i];
  for (uint64_t n = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; n < p.NL;
       n += uint64_t(gridDim.x) * blockDim.x) {
    const typename S::T* f = static_cast<const typename S::T*>(p.f);
    const uint64_t y = n / p.nx, d = n + y * (p.pitch - p.nx);
    for (int q = 0; q < 9; q++) out[q * p.NL + n] = S::load(f + q * p.Np, d) + wgt(q);
  }
}

template <class S>
__global__ void k_batch_from_f32(const Params* inst, uint32_t i, const float* in) {
  const Params& p = inst[i];
  for (uint64_t n = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; n < p.NL;
       n += uint64_t(gridDim.x) * blockDim.x) {
    typename S::T* f = static_cast<typename S::T*>(p.f);
    const uint64_t y = n / p.nx, d = n + y * (p.pitch - p.nx);
    for (int q = 0; q < 9; q++) S::store(f + q * p.Np, d, in[q * p.NL + n]);
  }
}

using BatchFn = void (*)(const Tile*, uint32_t, const Params*);
template <class S>
BatchFn batch_fn(int odd, bool mom) {
  if (mom) return odd ? k_batch_step<S, 1, true> : k_batch_step<S, 0, true>;
  return odd ? k_batch_step<S, 1, false> : k_batch_step<S, 0, false>;
}
BatchFn get_batch(int storage, int odd, bool mom) {
  switch (storage) {
    case 1: return batch_fn<StoreF16S>(odd, mom);
    case 2: return batch_fn<StoreBF16S>(odd, mom);
    default: return batch_fn<StoreF32>(odd, mom);
  }
}
using PersistFn = void (*)(const Tile*, uint32_t, const Params*, uint32_t, uint32_t);
PersistFn get_persistent(int storage) {
  switch (storage) {
    case 1: return k_batch_persistent<StoreF16S>;
    case 2: return k_batch_persistent<StoreBF16S>;
    default: return k_batch_persistent<StoreF32>;
  }
}

}  // namespace flb

struct flb_batch {
  int storage = 0;
  int block_x = 256;
  int grid = 0;            // blocks of the grid-stride launch
  int graph_chunk = 100;   // steps per captured graph (even), 0 = no graphs
  int mode = FLB_BATCH_LAUNCH;
  int order = FLB_BATCH_ORDER_ROWS;
  int persist_per_cu = 0;  // persistent mode: resident blocks per CU
  int persist_chunk = 100; // persistent mode: steps per cooperative launch
  int ncu = 256;
  hipStream_t stream = nullptr;
  std::vector<flb::Params> host_inst;
  flb::Params* inst = nullptr;
  flb::Tile* tiles = nullptr;
  uint32_t ntiles = 0;
  std::vector<void*> owned;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  uint64_t t = 0;
  std::string err;
};

namespace {

#define BAT_HIP(b, call)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
      return FLB_ERR_DEVICE;                                                               \
    }                                                                                      \
  } while (0)

template <class T>
int balloc(flb_batch* b, T** p, size_t count) {
  void* v = nullptr;
  if (hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    b->err = "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed";
    return FLB_ERR_MEMORY;
  }
  b->owned.push_back(v);
  *p = static_cast<T*>(v);
  return FLB_OK;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Tile list for block size bx in the batch's tile order:
//   FLB_BATCH_ORDER_ROWS      instances interleaved row by row (consecutive
//                             tiles come from different instances);
//   FLB_BATCH_ORDER_INSTANCE  instance after instance, each row-major;
//   FLB_BATCH_ORDER_BANDS     instance after instance, each in bands of 32
//                             rows walked column by column (the single-domain
//                             kernels' block order).
int build_tiles(flb_batch* b, int bx) {
  std::vector<flb::Tile> tl;
  const uint32_t n = uint32_t(b->host_inst.size());
  if (b->order == FLB_BATCH_ORDER_ROWS) {
    uint32_t max_rows = 0;
    for (const auto& p : b->host_inst) max_rows = std::max(max_rows, p.ny);
    for (uint32_t r = 0; r < max_rows; ++r)
      for (uint32_t i = 0; i < n; ++i) {
        const auto& p = b->host_inst[i];
        if (r >= p.ny) continue;
        const uint32_t tx = (p.nx + bx - 1) / bx;
        for (uint32_t k = 0; k < tx; ++k) tl.push_back(flb::Tile{i, r, k, 0});
      }
  } else {
    constexpr uint32_t kBand = 32;
    for (uint32_t i = 0; i < n; ++i) {
      const auto& p = b->host_inst[i];
      const uint32_t tx = (p.nx + bx - 1) / bx;
      if (b->order == FLB_BATCH_ORDER_INSTANCE) {
        for (uint32_t r = 0; r < p.ny; ++r)
          for (uint32_t k = 0; k < tx; ++k) tl.push_back(flb::Tile{i, r, k, 0});
      } else {
        // In a full band the tile at list position q runs on XCD q mod 8 (one
        // block per tile, or a grid-stride walk over a multiple of 8 blocks) and
        // takes row (q mod 8) * 4 + j / 8 of its column (j its place in the
        // column): each XCD gets four adjacent rows, as in the single-domain
        // band order, so that row-neighbour reads stay in one L2.
        for (uint32_t r0 = 0; r0 < p.ny; r0 += kBand) {
          const uint32_t h = std::min(p.ny - r0, kBand);
          for (uint32_t k = 0; k < tx; ++k) {
            const uint32_t base = uint32_t(tl.size());
            for (uint32_t j = 0; j < h; ++j) {
              const uint32_t r = h == kBand ? r0 + ((base + j) % 8u) * 4u + j / 8u : r0 + j;
              tl.push_back(flb::Tile{i, r, k, 0});
            }
          }
        }
      }
    }
  }
  if (b->tiles) {
    (void)hipFree(b->tiles);
    b->owned.erase(std::remove(b->owned.begin(), b->owned.end(), static_cast<void*>(b->tiles)),
                   b->owned.end());
    b->tiles = nullptr;
  }
  // The odd steps take the list's bands (runs of tiles of one instance and one
  // 32-row band) last to first, each band's tiles in their order, so that a
  // step starts on the tiles the previous one wrote last, which are still in
  // the Infinity Cache.  Both orders go into one allocation, the odd one second.
  std::vector<flb::Tile> both(tl);
  both.reserve(2 * tl.size());
  std::vector<std::pair<size_t, size_t>> runs;  // [begin, end) of each band run
  for (size_t a = 0; a < tl.size();) {
    size_t e = a + 1;
    while (e < tl.size() && tl[e].inst == tl[a].inst && tl[e].row / 32 == tl[a].row / 32) ++e;
    runs.emplace_back(a, e);
    a = e;
  }
  for (size_t r = runs.size(); r-- > 0;)
    both.insert(both.end(), tl.begin() + ptrdiff_t(runs[r].first), tl.begin() + ptrdiff_t(runs[r].second));
  int rc = balloc(b, &b->tiles, both.size());
  if (rc) return rc;
  BAT_HIP(b, hipMemcpy(b->tiles, both.data(), both.size() * sizeof(flb::Tile), hipMemcpyHostToDevice));
  b->ntiles = uint32_t(tl.size());
  b->block_x = bx;
  return FLB_OK;
}

void drop_graph(flb_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = 0;
}

int launch_step(flb_batch* b, int odd, bool mom) {
  const uint32_t grid = b->grid > 0 ? uint32_t(b->grid) : b->ntiles;  // 0: a block per tile
  hipLaunchKernelGGL(flb::get_batch(b->storage, odd, mom), dim3(grid), dim3(b->block_x), 0,
                     b->stream, b->tiles + (odd ? b->ntiles : 0u), b->ntiles, b->inst);
  BAT_HIP(b, hipGetLastError());
  return FLB_OK;
}

// Resident blocks per CU the persistent kernel can have at block size bx.
int persistent_capacity(flb_batch* b, int bx, int* per_cu) {
  int n = 0;
  BAT_HIP(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                 &n, reinterpret_cast<const void*>(flb::get_persistent(b->storage)), bx, 0));
  *per_cu = n;
  return FLB_OK;
}

// One cooperative launch of nsteps steps starting at parity odd0.  The runtime
// rejects a grid that cannot be resident at once, so the barrier always
// completes.
int launch_persistent(flb_batch* b, uint32_t nsteps, uint32_t odd0) {
  const uint32_t grid = uint32_t(b->persist_per_cu * b->ncu);
  const flb::Tile* tiles = b->tiles;
  uint32_t ntiles = b->ntiles;
  const flb::Params* inst = b->inst;
  void* args[] = {&tiles, &ntiles, &inst, &nsteps, &odd0};
  BAT_HIP(b, hipLaunchCooperativeKernel(reinterpret_cast<const void*>(flb::get_persistent(b->storage)),
                                        dim3(grid), dim3(b->block_x), args, 0, b->stream));
  return FLB_OK;
}

// Graph of graph_chunk steps starting at an even t (parities alternate).
int ensure_graph(flb_batch* b) {
  if (b->graph || b->graph_chunk <= 0) return FLB_OK;
  hipGraph_t g = nullptr;
  BAT_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = FLB_OK;
  for (int s = 0; s < b->graph_chunk && !rc; ++s) rc = launch_step(b, s & 1, false);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  BAT_HIP(b, e);
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BAT_HIP(b, e);
  b->graph_steps = b->graph_chunk;
  return FLB_OK;
}

}  // namespace

extern "C" {

const char* flb_batch_last_error(const flb_batch* b) { return b ? b->err.c_str() : ""; }

void flb_batch_destroy(flb_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (void* v : b->owned) (void)hipFree(v);
  delete b;
}

int flb_batch_create(const flb_instance* in, uint32_t count, int32_t storage, int32_t device,
                     flb_batch** out) {
  if (!out) return FLB_ERR_ARGUMENT;
  *out = nullptr;
  if (!in || count == 0 || storage < 0 || storage > 2) return FLB_ERR_ARGUMENT;
  for (uint32_t i = 0; i < count; ++i) {
    const flb_instance& s = in[i];
    if (!s.flags || s.nx < 2 || s.ny < 2 || !(s.tau > 0.5f)) return FLB_ERR_ARGUMENT;
    const size_t N = size_t(s.nx) * s.ny;
    if (s.flags[N - 1] & FLB_INLET) return FLB_ERR_ARGUMENT;
    // Planes addressed through 32-bit offsets and buffer descriptors.
    if (flb::plane_stride(align_up(s.nx, 64) * s.ny, 4) * 4 >= (1ull << 32)) return FLB_ERR_ARGUMENT;
    if ((s.flags[0] & FLB_OUTLET) && !(s.flags[0] & FLB_INLET)) return FLB_ERR_ARGUMENT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return FLB_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return FLB_ERR_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return FLB_ERR_DEVICE;
  flb_batch* b = new flb_batch();
  b->storage = storage;
  auto bail = [&](int rc) {
    flb_batch_destroy(b);
    return rc;
  };
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) b->ncu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(FLB_ERR_DEVICE);
  const size_t eb = storage == 0 ? 4 : 2;
  // Offsets of every instance in the shared allocations (elements).
  // Rows of every instance padded to a pitch of 64 cells.
  std::vector<size_t> off_f(count), off_c(count);
  size_t tot_f = 0, tot_c = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const size_t Nd = align_up(in[i].nx, 64) * in[i].ny;
    off_f[i] = tot_f;
    tot_f += 9 * flb::plane_stride(Nd, uint32_t(eb));
    off_c[i] = tot_c;
    tot_c += align_up(Nd, 64);
  }
  uint8_t* f = nullptr;
  uint8_t* flags = nullptr;
  float *rho = nullptr, *u = nullptr, *pr0 = nullptr, *pr1 = nullptr, *pu0 = nullptr, *pu1 = nullptr;
  int rc;
  if ((rc = balloc(b, &f, tot_f * eb)) || (rc = balloc(b, &flags, tot_c)) ||
      (rc = balloc(b, &rho, tot_c)) || (rc = balloc(b, &u, 2 * tot_c)) ||
      (rc = balloc(b, &pr0, tot_c)) || (rc = balloc(b, &pr1, tot_c)) ||
      (rc = balloc(b, &pu0, 2 * tot_c)) || (rc = balloc(b, &pu1, 2 * tot_c)) ||
      (rc = balloc(b, &b->inst, count)))
    return bail(rc);
  float* staging = nullptr;
  size_t max_n = 0;
  for (uint32_t i = 0; i < count; ++i) max_n = std::max(max_n, size_t(in[i].nx) * in[i].ny);
  if (hipMalloc(&staging, 9 * max_n * sizeof(float)) != hipSuccess) return bail(FLB_ERR_MEMORY);
  b->host_inst.resize(count);
  std::vector<float> fs, hr, hu;
  for (uint32_t i = 0; i < count && !rc; ++i) {
    const flb_instance& s = in[i];
    const size_t N = size_t(s.nx) * s.ny;
    flb::Params& p = b->host_inst[i];
    p = flb::Params{};
    p.nx = s.nx;
    p.ny = s.ny;
    p.pitch = uint32_t(align_up(s.nx, 64));
    p.N = size_t(p.pitch) * s.ny;  // device plane of a cell field
    p.NL = N;                      // host plane
    p.Np = flb::plane_stride(p.N, uint32_t(eb));
    p.inv_tau = 1.0f / s.tau;
    p.use_force = s.use_force;
    p.fx = s.use_force ? s.fx : 0.0f;
    p.fy = s.use_force ? s.fy : 0.0f;
    p.f = f + off_f[i] * eb;
    p.flags = flags + off_c[i];
    p.rho = rho + off_c[i];
    p.u = u + 2 * off_c[i];
    p.pub_rho[0] = pr0 + off_c[i];
    p.pub_rho[1] = pr1 + off_c[i];
    p.pub_u[0] = pu0 + 2 * off_c[i];
    p.pub_u[1] = pu1 + 2 * off_c[i];
    hr.assign(N, 1.0f);
    hu.assign(2 * N, 0.0f);
    if (s.rho0) std::memcpy(hr.data(), s.rho0, N * sizeof(float));
    if (s.u0) std::memcpy(hu.data(), s.u0, 2 * N * sizeof(float));
    flb::host_initial_populations(s.nx, s.ny, s.rho0, s.u0, fs);
    // Host planes of nx * ny cells into pitched device planes (a device plane
    // is exactly ny rows of pitch cells, so `planes` planes are planes * ny rows).
    auto in2d = [&](void* dst, const void* src, size_t elem, int planes) {
      return hipMemcpy2D(dst, p.pitch * elem, src, s.nx * elem, s.nx * elem, size_t(s.ny) * planes,
                         hipMemcpyHostToDevice);
    };
    hipError_t e = in2d(p.flags, s.flags, 1, 1);
    if (e == hipSuccess) e = in2d(p.rho, hr.data(), 4, 1);
    if (e == hipSuccess) e = in2d(p.u, hu.data(), 4, 2);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
      e = in2d(p.pub_rho[k], hr.data(), 4, 1);
      if (e == hipSuccess) e = in2d(p.pub_u[k], hu.data(), 4, 2);
    }
    if (e == hipSuccess) e = hipMemcpy(staging, fs.data(), 9 * N * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->inst + i, &p, sizeof(p), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      void (*conv)(const flb::Params*, uint32_t, const float*) =
          storage == 1 ? flb::k_batch_from_f32<flb::StoreF16S>
                       : (storage == 2 ? flb::k_batch_from_f32<flb::StoreBF16S>
                                       : flb::k_batch_from_f32<flb::StoreF32>);
      hipLaunchKernelGGL(conv, dim3(512), dim3(256), 0, b->stream, b->inst, i, staging);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    }
    if (e != hipSuccess) {
      b->err = std::string("batch upload: ") + hipGetErrorString(e);
      rc = FLB_ERR_DEVICE;
    }
  }
  (void)hipFree(staging);
  if (rc) return bail(rc);
  b->grid = 8 * b->ncu;
  if ((rc = build_tiles(b, 256))) return bail(rc);
  *out = b;
  return FLB_OK;
}

int flb_batch_configure(flb_batch* b, int32_t block_x, int32_t grid, int32_t graph_chunk) {
  if (!b) return FLB_ERR_ARGUMENT;
  if (block_x % 64 || block_x < 64 || block_x > 1024 || grid < 0 || graph_chunk < 0 ||
      graph_chunk % 2)
    return FLB_ERR_ARGUMENT;
  drop_graph(b);
  b->mode = FLB_BATCH_LAUNCH;
  b->grid = grid;
  b->graph_chunk = graph_chunk;
  if (block_x != b->block_x) return build_tiles(b, block_x);
  return FLB_OK;
}

int flb_batch_set_order(flb_batch* b, int32_t order) {
  if (!b || order < FLB_BATCH_ORDER_ROWS || order > FLB_BATCH_ORDER_BANDS) return FLB_ERR_ARGUMENT;
  if (order == b->order) return FLB_OK;
  drop_graph(b);
  b->order = order;
  return build_tiles(b, b->block_x);
}

int flb_batch_configure_persistent(flb_batch* b, int32_t block_x, int32_t blocks_per_cu,
                                   int32_t chunk) {
  if (!b) return FLB_ERR_ARGUMENT;
  if (block_x % 64 || block_x < 64 || block_x > 1024 || blocks_per_cu < 1 || chunk < 1)
    return FLB_ERR_ARGUMENT;
  int cap = 0, rc;
  if ((rc = persistent_capacity(b, block_x, &cap))) return rc;
  if (blocks_per_cu > cap) {
    b->err = "persistent batch: " + std::to_string(blocks_per_cu) + " blocks of " +
             std::to_string(block_x) + " threads per CU cannot be resident (at most " +
             std::to_string(cap) + ")";
    return FLB_ERR_ARGUMENT;
  }
  drop_graph(b);
  if (block_x != b->block_x && (rc = build_tiles(b, block_x))) return rc;
  b->mode = FLB_BATCH_PERSISTENT;
  b->persist_per_cu = blocks_per_cu;
  b->persist_chunk = chunk;
  return FLB_OK;
}

int flb_batch_step(flb_batch* b, uint64_t nsteps) {
  if (!b) return FLB_ERR_ARGUMENT;
  int rc = FLB_OK;
  uint64_t left = nsteps;
  while (left > 0 && !rc) {
    if (b->mode == FLB_BATCH_PERSISTENT && left > 1) {
      // The last step of the call runs as a plain launch and writes the fields.
      const uint64_t k = std::min<uint64_t>(left - 1, uint64_t(b->persist_chunk));
      rc = launch_persistent(b, uint32_t(k), uint32_t(b->t & 1));
      b->t += k;
      left -= k;
      continue;
    }
    // Graph replays need an even start; the last step writes the readback fields.
    if (b->graph_chunk > 0 && (b->t & 1) == 0 && left > uint64_t(b->graph_chunk)) {
      if ((rc = ensure_graph(b))) break;
      BAT_HIP(b, hipGraphLaunch(b->graph, b->stream));
      b->t += uint64_t(b->graph_steps);
      left -= uint64_t(b->graph_steps);
    } else {
      rc = launch_step(b, int(b->t & 1), left == 1);
      b->t += 1;
      left -= 1;
    }
  }
  return rc;
}

int flb_batch_sync(flb_batch* b) {
  if (!b) return FLB_ERR_ARGUMENT;
  BAT_HIP(b, hipStreamSynchronize(b->stream));
  return FLB_OK;
}

int flb_batch_step_timed(flb_batch* b, uint64_t nsteps, float* ms) {
  if (!b) return FLB_ERR_ARGUMENT;
  hipEvent_t e0, e1;
  BAT_HIP(b, hipEventCreate(&e0));
  BAT_HIP(b, hipEventCreate(&e1));
  BAT_HIP(b, hipEventRecord(e0, b->stream));
  int rc = flb_batch_step(b, nsteps);
  BAT_HIP(b, hipEventRecord(e1, b->stream));
  BAT_HIP(b, hipEventSynchronize(e1));
  float v = 0.0f;
  BAT_HIP(b, hipEventElapsedTime(&v, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = v;
  return rc;
}

// For each tile order, times block sizes {64, 128, 256, 512} x {4, 16, 64
// blocks per CU, one block per tile} with one launch per step, and the
// persistent kernel at block sizes {256, 512} with 1, 2 or 4 resident blocks
// per CU (as far as they fit), over `steps` steps each (the state advances),
// and keeps the fastest.
// best_grid reports blocks of the launch (0: one per tile); a persistent
// winner reports -(blocks per CU).
int flb_batch_autotune(flb_batch* b, uint32_t steps, int32_t* best_block, int32_t* best_grid,
                       float* best_ms) {
  if (!b || steps == 0) return FLB_ERR_ARGUMENT;
  const int blocks[4] = {64, 128, 256, 512};
  const int per_cu[4] = {4, 16, 64, 0};
  const int chunk = b->graph_chunk;
  float best = 1e30f;
  int bb = b->block_x, bg = b->grid, bo = b->order;
  bool bp = false;
  int rc = FLB_OK;
  auto time_it = [&](float* ms) -> int {
    int r = flb_batch_step(b, 2);  // warm
    if (!r) r = flb_batch_step_timed(b, steps, ms);
    return r;
  };
  for (int order : {FLB_BATCH_ORDER_ROWS, FLB_BATCH_ORDER_INSTANCE, FLB_BATCH_ORDER_BANDS}) {
    if ((rc = flb_batch_set_order(b, order))) return rc;
    for (int bx : blocks)
      for (int pc : per_cu) {
        if ((rc = flb_batch_configure(b, bx, pc * b->ncu, 0))) return rc;
        float ms = 0.0f;
        if ((rc = time_it(&ms))) return rc;
        if (ms < best) {
          best = ms;
          bb = bx;
          bg = pc * b->ncu;
          bo = order;
          bp = false;
        }
      }
    for (int bx : {256, 512})
      for (int pc : {1, 2, 4}) {
        int cap = 0;
        if ((rc = persistent_capacity(b, bx, &cap))) return rc;
        if (pc > cap) continue;
        if ((rc = flb_batch_configure_persistent(b, bx, pc, int(steps)))) return rc;
        float ms = 0.0f;
        if ((rc = time_it(&ms))) return rc;
        if (ms < best) {
          best = ms;
          bb = bx;
          bg = pc;
          bo = order;
          bp = true;
        }
      }
  }
  if ((rc = flb_batch_set_order(b, bo))) return rc;
  if (bp) {
    if ((rc = flb_batch_configure_persistent(b, bb, bg, chunk > 0 ? chunk : 100))) return rc;
  } else if ((rc = flb_batch_configure(b, bb, bg, chunk))) {
    return rc;
  }
  if (best_block) *best_block = bb;
  if (best_grid) *best_grid = bp ? -bg : bg;
  if (best_ms) *best_ms = best / float(steps);
  return FLB_OK;
}

int flb_batch_read(flb_batch* b, uint32_t i, float* u, float* rho) {
  if (!b || i >= b->host_inst.size()) return FLB_ERR_ARGUMENT;
  const flb::Params& p = b->host_inst[i];
  BAT_HIP(b, hipStreamSynchronize(b->stream));
  const size_t row = p.nx * sizeof(float), dp = p.pitch * sizeof(float);
  if (u) BAT_HIP(b, hipMemcpy2D(u, row, p.u, dp, row, 2 * size_t(p.ny), hipMemcpyDeviceToHost));
  if (rho) BAT_HIP(b, hipMemcpy2D(rho, row, p.rho, dp, row, p.ny, hipMemcpyDeviceToHost));
  return FLB_OK;
}

int flb_batch_read_populations(flb_batch* b, uint32_t i, float* f) {
  if (!b || !f || i >= b->host_inst.size()) return FLB_ERR_ARGUMENT;
  const flb::Params& p = b->host_inst[i];
  float* tmp = nullptr;
  BAT_HIP(b, hipMalloc(&tmp, 9 * p.NL * sizeof(float)));
  void (*conv)(const flb::Params*, uint32_t, float*) =
      b->storage == 1 ? flb::k_batch_to_f32<flb::StoreF16S>
                      : (b->storage == 2 ? flb::k_batch_to_f32<flb::StoreBF16S>
                                         : flb::k_batch_to_f32<flb::StoreF32>);
  hipLaunchKernelGGL(conv, dim3(512), dim3(256), 0, b->stream, b->inst, i, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e == hipSuccess) e = hipMemcpy(f, tmp, 9 * p.NL * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  BAT_HIP(b, e);
  return FLB_OK;
}

uint64_t flb_batch_time(const flb_batch* b) { return b ? b->t : 0; }

int32_t flb_batch_order(const flb_batch* b) { return b ? b->order : -1; }

int flb_batch_info(const flb_batch* b, int32_t* block_x, int32_t* grid, int32_t* graph_chunk,
                   uint32_t* ntiles) {
  if (!b) return FLB_ERR_ARGUMENT;
  const bool pers = b->mode == FLB_BATCH_PERSISTENT;
  if (block_x) *block_x = b->block_x;
  if (grid) *grid = pers ? -b->persist_per_cu : b->grid;
  if (graph_chunk) *graph_chunk = pers ? b->persist_chunk : b->graph_chunk;
  if (ntiles) *ntiles = b->ntiles;
  return FLB_OK;
}

}  // extern "C"
