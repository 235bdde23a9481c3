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

// Element i of an array through a 32-bit byte offset from its base.  With a
// wavefront-uniform base the compiler emits the SGPR-base form of the memory
// instruction (global_load/store ... saddr: 64-bit base in SGPRs, one VGPR of
// offset) instead of a per-lane 64-bit address -- two VALU instructions and a
// VGPR pair saved per access.  Every device plane is under 4 GiB (flb_create
// and flb_batch_create check), so an in-plane offset cannot wrap.
template <class T>
__device__ __forceinline__ T& at(T* base, uint32_t i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * uint32_t(sizeof(T)));
}
template <class T>
__device__ __forceinline__ const T& at(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * uint32_t(sizeof(T)));
}

// Buffer resource (V#) of one device plane of `bytes` bytes (< 4 GiB): the
// memory instructions then take the plane's base and size in SGPRs and one
// 32-bit VGPR offset per lane (buffer_load/store ... offen), and an offset
// past the plane reads zero / writes nothing instead of faulting.  The third
// configuration word is the gfx9-family value composable_kernel uses for
// gfx950.
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc plane_rsrc(const void* base, uint64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(uint32_t(bytes)),
                                           0x00020000);
}
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------- storage
// All formats hold the DDF-shifted population f_i - w_i.  T is the stored
// type, R the arithmetic type: FP32 for the 32- and 16-bit formats, FP64 for
// StoreF64 (the reference's PRECISION = 64, whose fields -- density,
// velocity, mass, excess, fill level -- are double as well).
struct StoreF32 {
  using T = float;
  using R = float;
  static constexpr int kBytes = 4;
  static constexpr bool kBuffer = true;
  __device__ static float load(const T* p, uint32_t i) { return at(p, i); }
  __device__ static void store(T* p, uint32_t i, float v) { at(p, i) = v; }
  __device__ static float bload(Rsrc r, uint32_t i) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(i * 4u), 0, 0));
  }
  __device__ static void bstore(Rsrc r, uint32_t i, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(i * 4u), 0, 0);
  }
  __device__ static float round(float v) { return v; }  // value a store keeps
};

struct StoreF64 {
  using T = double;
  using R = double;
  static constexpr int kBytes = 8;
  static constexpr bool kBuffer = true;
  __device__ static double load(const T* p, uint32_t i) { return at(p, i); }
  __device__ static void store(T* p, uint32_t i, double v) { at(p, i) = v; }
  __device__ static double bload(Rsrc r, uint32_t i) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(i * 8u), 0, 0));
  }
  __device__ static void bstore(Rsrc r, uint32_t i, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, int(i * 8u), 0, 0);
  }
  __device__ static double round(double v) { return v; }
};

// IEEE binary16, scaled by 2^15 so that the shifted populations (|f - w| < 1)
// sit in the normal range.
struct StoreF16S {
  using T = _Float16;
  using R = float;
  static constexpr int kBytes = 2;
  static constexpr bool kBuffer = true;
  __device__ static float load(const T* p, uint32_t i) { return float(at(p, i)) * 3.0517578125e-05f; }
  __device__ static void store(T* p, uint32_t i, float v) { at(p, i) = _Float16(v * 32768.0f); }
  __device__ static float bload(Rsrc r, uint32_t i) {
    const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, int(i * 2u), 0, 0);
    return float(__builtin_bit_cast(_Float16, h)) * 3.0517578125e-05f;
  }
  __device__ static void bstore(Rsrc r, uint32_t i, float v) {
    const _Float16 h = _Float16(v * 32768.0f);
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), r, int(i * 2u), 0, 0);
  }
  __device__ static float round(float v) {
    return float(_Float16(v * 32768.0f)) * 3.0517578125e-05f;
  }
};

// bfloat16 (upper half of an FP32), round to nearest even on store: gfx950's
// FP32 -> BF16 conversion (one instruction; for finite values the same bits as
// adding 0x7fff plus the lowest kept bit and truncating).
struct StoreBF16S {
  using T = uint16_t;
  using R = float;
  static constexpr int kBytes = 2;
  static constexpr bool kBuffer = true;
  __device__ static unsigned short bits(float v) { return __builtin_bit_cast(unsigned short, __bf16(v)); }
  __device__ static float load(const T* p, uint32_t i) {
    return __uint_as_float(uint32_t(at(p, i)) << 16);
  }
  __device__ static void store(T* p, uint32_t i, float v) { at(p, i) = bits(v); }
  __device__ static float bload(Rsrc r, uint32_t i) {
    const unsigned short h = __builtin_amdgcn_raw_buffer_load_b16(r, int(i * 2u), 0, 0);
    return __uint_as_float(uint32_t(h) << 16);
  }
  __device__ static void bstore(Rsrc r, uint32_t i, float v) {
    __builtin_amdgcn_raw_buffer_store_b16(bits(v), r, int(i * 2u), 0, 0);
  }
  __device__ static float round(float v) {
    return __uint_as_float(uint32_t(bits(v)) << 16);
  }
};

// Tile of the fused free-surface passes 2-4: one block of kFsTileW x kFsTileH
// threads per tile, one wavefront per tile row.
constexpr int kFsTileW = 64, kFsTileH = 8;
static_assert(kFsTileH <= 8, "one byte of row bits per pass in a tile_work word");

// ---------------------------------------------------------------- params
// Device cell fields and population planes store a row of nx cells every
// `pitch` cells (pitch >= nx): cell (x, y) is element x + y * pitch of a plane.
// The default pitch rounds nx up to 64, so every wavefront row segment starts
// 256-byte aligned (FP32).
struct Params {
  uint32_t nx, ny;
  uint32_t pitch;        // row stride of the device planes (cells)
  uint32_t tiles_x;      // blocks per row
  uint32_t nblocks;      // total blocks of a full-grid launch
  uint64_t N;            // plane size of the device cell fields (pitch * ny)
  uint64_t NL;           // cells of the domain (nx * ny): host arrays, staging
  uint64_t Np;           // plane stride of the population array (elements)
  float inv_tau, fx, fy;
  int32_t use_force, collision, smagorinsky, gas_check_corrected, mass_mode;
  void* f;               // populations, 9 planes of Np elements of the storage type
  uint8_t* flags;        // current flags
  uint8_t* flags_next;   // flags written by the free-surface update (passes 2-4)
  // Cell fields.  They hold elements of the storage's arithmetic type R
  // (double for StoreF64); kernels access them through fld<R>().
  float* rho;            // readback density (N)
  float* u;              // readback velocity (2 planes of N)
  float* pub_rho[2];     // densities published for inlet cells, by parity
  float* pub_u[2];       // velocities published for outlet cells, by parity
  float* mass;
  // Excess shares, double-buffered by step: excess holds the shares pulled this
  // step (by pass 1 and stream-collide), excess_out receives the shares the
  // update -- and stream-collide for bodies of fluid -- hand out for the next.
  float* excess;
  float* excess_out;
  float* phi;
  float* rec;            // reconstructed gas populations, 9 planes of N
  const float* alpha;    // Brinkman solid fraction or nullptr
  float* trace;          // adjoint trace of one step: 3 planes of N elements of R (or nullptr)
  // Free surface: one word per kFsTileW x kFsTileH tile (work_x per tile row)
  // telling which of its 64-cell row chunks have work next step: bit r for
  // stream-collide in row r of the tile, bit 8 + r for pass 1.  Written by the
  // update kernel; tile_pub holds the stream-collide bits of the chunks with a
  // cell that publishes for an inlet/outlet neighbour (static).
  uint32_t* tile_work;
  const uint32_t* tile_pub;
  uint32_t work_x;     // tiles (and 64-cell chunks) per row
  uint32_t tile_rows;  // tile rows
  // Pass-1 work list of this step (read) and of the next step (written by the
  // update kernel): row chunks, index = y * work_x + chunk; counts_* hold the
  // list length at kListFS1.
  const uint32_t* fs1_list;
  const uint32_t* counts_in;
  uint32_t* fs1_list_out;
  uint32_t* counts_out;
  // Tiles the update of this step evaluates (written by pass 1), their number,
  // and the counter of the next step's list (emptied by pass 1).
  uint32_t* tile_list;
  uint32_t* tile_count;
  uint32_t* tile_count_next;
  // Per tile, written by pass 1: the tile and its three-cell halo are plain
  // fluid, so stream-collide also does the update's work for its cells.
  uint32_t* tile_body;
  // Free surface: bounding box, in tiles (row lo, row hi, column lo, column
  // hi), of the chunks with stream-collide work.  Pass 1 writes one box per
  // 64 tiles; the update of step t reduces them and publishes the result to
  // box_host (mapped host memory: the packed box, then t).
  uint32_t* box;
  unsigned long long* box_host;
  uint64_t step;  // t
  // Window of the free-surface stream-collide launch (locate_window): bands of
  // kBandRows rows from band0, block columns col0 .. col0 + ncols - 1.  The
  // whole grid is 0, 0, tiles_x.
  uint32_t band0, col0, ncols;
  // Blocks in the last band of the launch and bands of the launch (the odd
  // steps take the bands last to first: walk_block).
  uint32_t walk_last, walk_bands;
  // div_magic(kBandRows * tiles_x), div_magic(kBandRows * ncols) and
  // div_magic(height of the last band): block_tile and block_tile_window
  // divide on the scalar unit (0 = divide instead).
  uint32_t band_magic, win_magic, last_magic;
};

enum : int { kListFS1 = 0 };
constexpr int kListThreads = 256;  // block size of the list-driven pass 1
// Grid of pass 1 and of the surface stream-collide launch (FP64 storage),
// which walk the pass-1 chunk list with one wavefront per chunk.  On the
// 8192 x 2048 dam break (about 2500 chunks) the step took 132 us with 1
// block per CU, 126 us with 2 and 121 us with 4 or 8.
constexpr int kListBlocksPerCU = 4;
// Grid of the update's walk over its tile list.  On the 8192 x 2048 dam break
// the update took 47 us with 4 blocks per CU (all resident), 39 us with 16
// and 40 us with 32; one block per tile, the idle ones leaving at once, 55 us
// (all-gas domain: 7.8 us with 16 per CU, 31 us with one block per tile).
constexpr int kUpdateBlocksPerCU = 16;

constexpr int kWorkSC = 0, kWorkFS1 = 8;  // bit offsets in a tile_work word
constexpr uint32_t kWorkQuiet = 1u << 16;  // the update found the tile quiet
constexpr uint32_t kWorkFluid = 1u << 17;  // every cell of the tile is plain fluid
constexpr uint32_t kWorkReset = 0xffffu;   // every chunk has work, nothing known
constexpr uint32_t kWorkSCMask = 0xffu << kWorkSC;

// A cell field viewed in the arithmetic type R.
template <class R>
__device__ __forceinline__ R* fld(float* p) {
  return reinterpret_cast<R*>(p);
}

// Block -> tile mapping.  Workgroups are dispatched round-robin: block b goes
// to XCD b mod 8, and within an XCD consecutive blocks go to its shader
// engines in turn, so block b lands on one of 32 fixed dispatch slots by
// b mod 32.  In row-major order with a multiple of 32 tiles per row every slot
// owns fixed column strips: on the 8192 x 2048 dam break (rows wet left of
// x = 3600, 32 tiles per row) half of the shader engines never saw a wet tile
// and stream-collide took 177 us.  Blocks therefore walk bands of kBandRows
// rows column by column (block b -> row band * kBandRows + b mod h of its
// band of h rows): the slot follows the row, every slot gets every column,
// and stream-collide takes 118 us (single phase at 8192 x 2048: 286 -> 254
// us).  Rotating each row by y tiles was as fast at 32 tiles per row but not
// at 33.  A whole band (32 rows) is in flight at once, as 32 rows were in
// row-major order.  Within a column of a full band the 32 consecutive blocks
// take the rows in the order 0, 4, 8, .., 28, 1, 5, .., so that XCD k (blocks
// b = k mod 8) runs rows 4k .. 4k + 3: a cell reads its neighbours' excess
// shares and slots in the rows above and below, and with consecutive rows on
// consecutive XCDs each XCD's L2 fetched those lines again (8192 x 2048 dam
// break: FP16S stream-collide 46.6 -> 42.3 us, FP32 69.3 -> 67.1 us; single
// phase FP32 +2 %).  8 rows per XCD in bands of 64 measured the same.
constexpr uint32_t kTileOrder = 2;  // 0 row-major, 1 rows rotated by y, 2 bands
constexpr uint32_t kBandRows = 32;

// Division of a wave-uniform index by a launch constant d >= 2 on the scalar
// unit.  With m = ceil(2^32 / d), q = (n m) >> 32 is floor(n / d) or one more
// for any n < 2^31 (m exceeds 2^32 / d by less than 1), and the compare
// corrects it.  The compiler's own unsigned division goes through the vector
// unit's float reciprocal: ~10 of the ~185 vector instructions of a 16-bit
// stream-collide wavefront.  m = 0 divides instead.
__host__ __device__ inline uint32_t div_magic(uint64_t d) {
  return d >= 2 ? uint32_t(((1ull << 32) + d - 1) / d) : 0u;
}
__host__ __device__ __forceinline__ uint32_t div_by(uint32_t n, uint32_t d, uint32_t m) {
  if (m == 0) return n / d;
  const uint32_t q = uint32_t((uint64_t(n) * m) >> 32);
  return q * d > n ? q - 1 : q;
}
// Row and block column of block b within a band of h rows: full bands (h =
// kBandRows) divide by a power of two and take their rows in XCD groups of
// four (above), the last band (h = last, its magic mlast) by magic
// multiplication, rows in order; both are formed and one is selected, so the
// mapping stays branch-free scalar code.
__host__ __device__ __forceinline__ void band_cell(uint32_t r, uint32_t h, uint32_t last,
                                                   uint32_t mlast, uint32_t& c, uint32_t& row) {
  const uint32_t cl = div_by(r, last, mlast);
  const bool full = h == kBandRows;
  const uint32_t q = r % kBandRows;
  c = full ? r / kBandRows : cl;
  row = full ? (q & 7u) * 4u + (q >> 3) : r - cl * last;
}
// Height of the last band of a grid of `rows` rows and its magic.
__host__ __device__ inline uint32_t last_band_rows(uint32_t rows) {
  return rows - (rows - 1) / kBandRows * kBandRows;
}

__host__ __device__ __forceinline__ void block_tile(uint32_t b, uint32_t tiles_x, uint32_t rows,
                                                    uint32_t magic, uint32_t last_magic,
                                                    uint32_t& tx, uint32_t& y) {
  if (kTileOrder == 2) {
    const uint32_t per_band = kBandRows * tiles_x;
    const uint32_t band = div_by(b, per_band, magic);
    const uint32_t r = b - band * per_band;
    const uint32_t y0 = band * kBandRows;
    const uint32_t h = rows - y0 < kBandRows ? rows - y0 : kBandRows;
    uint32_t row;
    band_cell(r, h, last_band_rows(rows), last_magic, tx, row);
    y = y0 + row;
    return;
  }
  y = b / tiles_x;
  tx = b - y * tiles_x;
  if (kTileOrder == 1) tx = (tx + y % tiles_x) % tiles_x;
}

// Forward-order index of block b when a launch takes its bands last to first,
// each band's blocks in their forward order: `last` blocks in the last band,
// `nbands` bands, `per` blocks in a full band (div_magic m).  The row
// grouping of a band is kept: every band but the last is a whole number of
// 32-block columns, so rows congruent mod 8 stay on one XCD.
__host__ __device__ __forceinline__ uint32_t band_reversed(uint32_t b, uint32_t last,
                                                           uint32_t nbands, uint32_t per,
                                                           uint32_t m) {
  if (b < last) return (nbands - 1u) * per + b;
  const uint32_t b2 = b - last;
  const uint32_t k = div_by(b2, per, m);
  return (nbands - 2u - k) * per + (b2 - k * per);
}

// Params::walk_last / walk_bands of a launch over `rows` rows of `cols` block
// columns.
inline void set_walk(Params& p, uint32_t rows, uint32_t cols) {
  p.walk_bands = (rows + kBandRows - 1) / kBandRows;
  p.walk_last = (rows - (p.walk_bands - 1) * kBandRows) * cols;
}

// block_tile (band order) restricted to the window band0 / col0 / ncols of
// Params: the same mapping as block_tile over the whole grid when the window
// is (0, 0, tiles_x).
__host__ __device__ __forceinline__ void block_tile_window(const Params& p, uint32_t b,
                                                           uint32_t& tx, uint32_t& y) {
  const uint32_t per_band = kBandRows * p.ncols;
  const uint32_t band = div_by(b, per_band, p.win_magic);
  const uint32_t r = b - band * per_band;
  const uint32_t y0 = (p.band0 + band) * kBandRows;
  const uint32_t h = p.ny - y0 < kBandRows ? p.ny - y0 : kBandRows;
  uint32_t c, row;
  band_cell(r, h, last_band_rows(p.ny), p.last_magic, c, row);
  tx = p.col0 + c;
  y = y0 + row;
}

}  // namespace flb
