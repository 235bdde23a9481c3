// flb_core.hip -- handle, device memory layout and the C ABI of the FlowLB core.
//
// Device memory of one handle.  Every plane stores the nx x ny cells in rows
// of `pitch` cells (Nd = pitch * ny elements per plane of a cell field):
//   f          9 planes x Np elements of the storage type (Np = Nd rounded up
//              to 8 KiB plus 4 KiB: flb::plane_stride; every plane starts
//              256-byte aligned)
//   flags[2]   Nd bytes each; free-surface pass 4 writes the second buffer
//   rho, u     readback fields, 1 and 2 planes
//   pub_rho[2], pub_u[2]   inlet/outlet neighbour values by step parity
//   mass, excess, phi      free surface, 1 plane each
//   rec        free surface, reconstructed gas populations, 9 planes
// Host arrays of the ABI hold nx * ny cells per plane, row-major.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/flb_core.h"
#include "flb_device.h"

#include "flb_internal.h"

// Device scratch of flb_flag_histogram_noalloc: its 256 flag counters, one
// copy per GPU, held for the life of the process (2 KiB).  In host code this
// name is not a device address; the per-device address comes from
// hipGetSymbolAddress (scratch_address below).
constexpr size_t kScratchCounters = 256;
static __device__ __align__(256) unsigned long long scratch[kScratchCounters];

namespace {

constexpr int kUtilBlock = 256;

// Address of `scratch` on each device, looked up once per device with that
// device current (flb_create).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned long long* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_address(int device, unsigned long long** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

__global__ void k_clear_u64(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

int fail(flb_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

int hip_fail(flb_handle* h, hipError_t e, const char* what) {
  return fail(h, FLB_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define FLB_HIP(h, call)                                   \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail((h), e_, #call); \
  } while (0)

uint64_t round_up(uint64_t v, uint64_t m) { return (v + m - 1) / m * m; }

int block_x(const flb_handle* h) { return h->opt.block_x > 0 ? h->opt.block_x : 256; }

flb::Params params(const flb_handle* h, int bx) {
  flb::Params p{};
  p.nx = h->nx;
  p.ny = h->ny;
  p.pitch = h->pitch;
  p.tiles_x = uint32_t((h->nx + bx - 1) / bx);
  p.nblocks = p.tiles_x * h->ny;
  p.N = h->Nd;
  p.NL = h->N;
  p.Np = h->Np;
  p.inv_tau = h->inv_tau;
  const bool force = h->opt.use_force != 0;
  p.fx = force ? h->opt.fx : 0.0f;
  p.fy = force ? h->opt.fy : 0.0f;
  p.use_force = h->opt.use_force;
  p.collision = h->opt.collision;
  p.smagorinsky = h->opt.smagorinsky;
  p.gas_check_corrected = h->opt.gas_check_corrected;
  p.mass_mode = h->opt.mass_mode;
  p.f = h->f;
  p.flags = h->flags[h->cur];
  p.flags_next = h->flags[h->cur ^ 1];
  p.rho = h->rho;
  p.u = h->u;
  p.pub_rho[0] = h->pub_rho[0];
  p.pub_rho[1] = h->pub_rho[1];
  p.pub_u[0] = h->pub_u[0];
  p.pub_u[1] = h->pub_u[1];
  p.mass = h->mass;
  p.excess = h->excess[h->cur];
  p.excess_out = h->excess[h->cur ^ 1];
  p.phi = h->phi;
  p.rec = h->rec;
  p.alpha = h->alpha;
  p.trace = nullptr;
  p.tile_work = h->tile_work;
  p.tile_pub = h->tile_pub;
  p.work_x = h->work_x;
  p.tile_rows = h->tile_rows;
  p.tile_list = h->tile_list;
  p.tile_body = h->tile_body;
  p.box = h->box;
  p.box_host = h->box_host_dev;
  p.step = h->t;
  p.band0 = 0;
  p.col0 = 0;
  p.ncols = p.tiles_x;
  flb::set_walk(p, h->ny, p.tiles_x);
  p.band_magic = flb::div_magic(uint64_t(flb::kBandRows) * p.tiles_x);
  p.win_magic = p.band_magic;
  p.last_magic = flb::div_magic(flb::last_band_rows(h->ny));
  return p;
}

template <class T>
int dev_alloc(flb_handle* h, T** ptr, size_t count) {
  void* v = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = hipMalloc(&v, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, FLB_ERR_MEMORY, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                       " bytes failed: " + hipGetErrorString(e));
  }
  h->owned.push_back(v);
  *ptr = static_cast<T*>(v);
  // Zeroed, so the row padding of pitched planes holds defined values.
  FLB_HIP(h, hipMemsetAsync(v, 0, bytes, h->stream));
  return FLB_OK;
}

int upload(flb_handle* h, void* dst, const void* src, size_t bytes) {
  FLB_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  FLB_HIP(h, hipStreamSynchronize(h->stream));
  return FLB_OK;
}

int download(flb_handle* h, void* dst, const void* src, size_t bytes) {
  FLB_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
  FLB_HIP(h, hipStreamSynchronize(h->stream));
  return FLB_OK;
}

// Host planes of nx * ny cells <-> device planes of Nd cells in rows of
// `pitch`: one strided copy over ny * planes rows (a device plane is exactly
// ny rows of pitch cells).
int cells_in(flb_handle* h, void* dev, const void* host, size_t elem, int planes) {
  FLB_HIP(h, hipMemcpy2DAsync(dev, h->pitch * elem, host, h->nx * elem, h->nx * elem,
                              size_t(h->ny) * planes, hipMemcpyHostToDevice, h->stream));
  FLB_HIP(h, hipStreamSynchronize(h->stream));
  return FLB_OK;
}

int cells_out(flb_handle* h, void* host, const void* dev, size_t elem, int planes) {
  FLB_HIP(h, hipMemcpy2DAsync(host, h->nx * elem, dev, h->pitch * elem, h->nx * elem,
                              size_t(h->ny) * planes, hipMemcpyDeviceToHost, h->stream));
  FLB_HIP(h, hipStreamSynchronize(h->stream));
  return FLB_OK;
}

// Cell fields hold float, or double for FP64 storage (h->field_bytes); the
// host side of the ABI passes float arrays and these helpers convert.
// count: device elements.
int field_alloc(flb_handle* h, float** ptr, size_t count) {
  uint8_t* b = nullptr;
  int rc = dev_alloc(h, &b, count * size_t(h->field_bytes));
  *ptr = reinterpret_cast<float*>(b);
  return rc;
}

template <class In>
int upload_field(flb_handle* h, float* dst, const In* src, int planes) {
  const size_t count = h->N * planes;
  if (h->field_bytes == 8) {
    std::vector<double> t(src, src + count);
    return cells_in(h, dst, t.data(), sizeof(double), planes);
  }
  if (std::is_same<In, float>::value) return cells_in(h, dst, src, sizeof(float), planes);
  std::vector<float> t(src, src + count);
  return cells_in(h, dst, t.data(), sizeof(float), planes);
}

template <class Out>
int download_field(flb_handle* h, Out* dst, const float* src, int planes) {
  const size_t count = h->N * planes;
  if (h->field_bytes == 8) {
    if (std::is_same<Out, double>::value) return cells_out(h, dst, src, sizeof(double), planes);
    std::vector<double> t(count);
    int rc = cells_out(h, t.data(), src, sizeof(double), planes);
    for (size_t i = 0; i < count && !rc; ++i) dst[i] = Out(t[i]);
    return rc;
  }
  if (std::is_same<Out, float>::value) return cells_out(h, dst, src, sizeof(float), planes);
  std::vector<float> t(count);
  int rc = cells_out(h, t.data(), src, sizeof(float), planes);
  for (size_t i = 0; i < count && !rc; ++i) dst[i] = Out(t[i]);
  return rc;
}

// Free surface: the next step evaluates everything -- every tile in the
// update, every chunk in pass 1 and stream-collide -- after which the update
// kernel narrows the work map again.  Used at creation and whenever the host
// replaces cell fields.
int reset_work(flb_handle* h) {
  const size_t tiles = size_t(h->work_x) * h->tile_rows;
  const size_t chunks = size_t(h->work_x) * h->ny;
  const int rd = int(h->t & 1u);
  std::vector<uint32_t> ids(chunks), all(tiles, flb::kWorkReset);
  for (size_t i = 0; i < ids.size(); ++i) ids[i] = uint32_t(i);
  uint32_t counts[8] = {};
  counts[4 * rd + flb::kListFS1] = uint32_t(chunks);
  const uint32_t zero2[2] = {0u, 0u};
  int rc;
  if ((rc = upload(h, h->tile_work, all.data(), tiles * sizeof(uint32_t)))) return rc;
  if ((rc = upload(h, h->fs1_list[rd], ids.data(), chunks * sizeof(uint32_t)))) return rc;
  if ((rc = upload(h, h->tile_counts, zero2, sizeof(zero2)))) return rc;
  if ((rc = upload(h, h->counts, counts, sizeof(counts)))) return rc;
  // The published stream-collide box is marked unknown once the uploads
  // (which wait for the stream) are done, so that no update still in flight
  // can overwrite the mark.
  *static_cast<volatile unsigned long long*>(&h->box_host[1]) = ~0ull;
  return FLB_OK;
}

// count: device elements.
int zero_field(flb_handle* h, float* dst, size_t count) {
  FLB_HIP(h, hipMemsetAsync(dst, 0, count * size_t(h->field_bytes), h->stream));
  return FLB_OK;
}

int launch(flb_handle* h, flb::KernelFn fn, const flb::Params& p, int bx) {
  hipLaunchKernelGGL(fn, dim3(p.nblocks), dim3(bx), 0, h->stream, p);
  FLB_HIP(h, hipGetLastError());
  return FLB_OK;
}

// Host initial populations, initFields (src/initData.h:125-165): feq(rho0, u0)
// evaluated in double as initEquilibriumDDF2Q9 does (:75-100), stored with
// the odd-step pattern (storef with t = 1), DDF-shifted (f - w).
template <class T>
void initial_populations(uint32_t nx, uint32_t ny, const float* rho0, const float* u0,
                         std::vector<T>& fs) {
  const size_t N = size_t(nx) * ny;
  fs.assign(9 * N, T(0));
  const double w[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                       1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
  const int cxs[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
  const int cys[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
  for (uint32_t y = 0; y < ny; ++y)
    for (uint32_t x = 0; x < nx; ++x) {
      const size_t n = size_t(y) * nx + x;
      const double r = rho0 ? double(rho0[n]) : 1.0;
      const double ux = u0 ? double(u0[n]) : 0.0, uy = u0 ? double(u0[N + n]) : 0.0;
      const double usq = 1.5 * (ux * ux + uy * uy);
      double feq[9];
      for (int i = 0; i < 9; ++i) {
        const double cu = 3.0 * (cxs[i] * ux + cys[i] * uy);
        // w (r - 1) + w r (cu + cu^2/2 - 3u^2/2): the shifted equilibrium
        feq[i] = w[i] * (r - 1.0) + w[i] * r * (cu + 0.5 * cu * cu - usq);
      }
      const uint32_t xr = x + 1 == nx ? 0 : x + 1, xl = x == 0 ? nx - 1 : x - 1;
      const uint32_t yu = y + 1 == ny ? 0 : y + 1, yd = y == 0 ? ny - 1 : y - 1;
      const size_t nb[9] = {n,
                            size_t(y) * nx + xr,
                            size_t(y) * nx + xl,
                            size_t(yu) * nx + x,
                            size_t(yd) * nx + x,
                            size_t(yu) * nx + xr,
                            size_t(yd) * nx + xl,
                            size_t(yd) * nx + xr,
                            size_t(yu) * nx + xl};
      // storef at odd parity: f_i -> slot (i+1, n + c_i), f_{i+1} -> slot (i, n).
      fs[n] = T(feq[0]);
      for (int i = 1; i < 9; i += 2) {
        fs[size_t(i + 1) * N + nb[i]] = T(feq[i]);
        fs[size_t(i) * N + n] = T(feq[i + 1]);
      }
    }
}

int validate(const flb_options* o, uint32_t nx, uint32_t ny, std::string& why) {
  if (!(o->tau > 0.5f) || !std::isfinite(o->tau)) return why = "tau must be > 0.5", 1;
  if (o->storage < 0 || o->storage > 3) return why = "unknown storage format", 1;
  if (o->collision < 0 || o->collision > 2) return why = "collision must be SRT, TRT or TRT_GUO", 1;
  if (o->problem < 0 || o->problem > 1) return why = "problem must be monofluid or free surface", 1;
  if (o->mass_mode < 0 || o->mass_mode > 1) return why = "unknown mass mode", 1;
  if (o->block_x != 0 && (o->block_x < 64 || o->block_x > 1024 || o->block_x % 64))
    return why = "block_x must be a multiple of 64 in [64, 1024]", 1;
  if (nx < 2 || ny < 2) return why = "grid must be at least 2 x 2", 1;
  if (o->whole_grid != 0 && o->whole_grid != 1) return why = "whole_grid must be 0 or 1", 1;
  if (o->row_pitch != 0 && (o->row_pitch < 0 || uint32_t(o->row_pitch) < nx))
    return why = "row_pitch must be 0 or at least nx", 1;
  if ((uint64_t(nx) + 128) * ny > 0xffffffffull / 2 ||
      uint64_t(o->row_pitch) * ny > 0xffffffffull / 2)
    return why = "grid too large for 32-bit cell indices", 1;
  // Kernels address every device plane (populations and cell fields) through
  // 32-bit byte offsets and buffer descriptors: a plane must stay under 4 GiB.
  const uint64_t pitch = o->row_pitch ? uint64_t(o->row_pitch) : (uint64_t(nx) + 63) / 64 * 64;
  if (flb::plane_stride(pitch * ny, 8) * 8 >= (1ull << 32))
    return why = "grid too large: a device plane must stay under 4 GiB (pitch * ny < 2^29)", 1;
  return 0;
}

// Row stride of the device planes: nx rounded up to 64 cells, so each
// wavefront row segment starts 256-byte aligned.  (A further 64 cells of
// padding for rows of a multiple of 4 KiB measured the same on the
// 8192 x 2048 dam break: stream-collide 170 vs 167 us.)
uint32_t row_pitch(const flb_options* o, uint32_t nx) {
  if (o->row_pitch > 0) return uint32_t(o->row_pitch);
  return (nx + 63u) / 64u * 64u;
}

}  // namespace

extern "C" {

void flb_default_options(flb_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->tau = 0.6f;
  o->collision = FLB_SRT;
  o->problem = FLB_MONOFLUID;
  o->storage = FLB_STORAGE_FP32;
  o->mass_mode = FLB_MASS_CONSERVATIVE;
}

const char* flb_status_string(int s) {
  switch (s) {
    case FLB_OK: return "ok";
    case FLB_ERR_ARGUMENT: return "invalid argument";
    case FLB_ERR_DEVICE: return "device error";
    case FLB_ERR_MEMORY: return "out of memory";
    case FLB_ERR_STATE: return "invalid state";
    case FLB_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

const char* flb_last_error(const flb_handle* h) { return h ? h->err.c_str() : ""; }

uint64_t flb_time(const flb_handle* h) { return h ? h->t : 0; }

int flb_get_shape(const flb_handle* h, uint32_t* nx, uint32_t* ny, int32_t* problem) {
  if (!h) return FLB_ERR_ARGUMENT;
  if (nx) *nx = h->nx;
  if (ny) *ny = h->ny;
  if (problem) *problem = h->opt.problem;
  return FLB_OK;
}

int flb_get_layout(const flb_handle* h, uint32_t* pitch, uint64_t* plane) {
  if (!h) return FLB_ERR_ARGUMENT;
  if (pitch) *pitch = h->pitch;
  if (plane) *plane = h->Nd;
  return FLB_OK;
}

void flb_destroy(flb_handle* h) {
  if (!h) return;
  if (h->stream) {
    (void)hipStreamSynchronize(h->stream);
    (void)hipStreamDestroy(h->stream);
  }
  for (void* v : h->owned) (void)hipFree(v);
  if (h->box_host) (void)hipHostFree(h->box_host);
  for (hipEvent_t e : h->in_flight)
    if (e) (void)hipEventDestroy(e);
  delete h;
}

int flb_create(const flb_options* opt, const uint8_t* flags, uint32_t nx, uint32_t ny,
               const float* rho0, const float* u0, flb_handle** out) {
  if (!out) return FLB_ERR_ARGUMENT;
  *out = nullptr;
  if (!opt || !flags) return FLB_ERR_ARGUMENT;
  std::string why;
  if (validate(opt, nx, ny, why)) return FLB_ERR_ARGUMENT;
  const uint64_t N = uint64_t(nx) * ny;
  // The reference reads rho[n+1] at inlets and u[n-1] at outlets
  // (lbmKernel2D.cu:336,344); those must exist.
  if (flags[N - 1] & FLB_INLET) return FLB_ERR_ARGUMENT;
  if ((flags[0] & FLB_OUTLET) && !(flags[0] & FLB_INLET)) return FLB_ERR_ARGUMENT;

  flb_handle* h = new (std::nothrow) flb_handle();
  if (!h) return FLB_ERR_MEMORY;
  h->opt = *opt;
  h->nx = nx;
  h->ny = ny;
  h->N = N;
  h->pitch = row_pitch(opt, nx);
  h->Nd = uint64_t(h->pitch) * ny;
  h->elem_bytes = opt->storage == FLB_STORAGE_FP32 ? 4 : (opt->storage == FLB_STORAGE_FP64 ? 8 : 2);
  h->Np = flb::plane_stride(h->Nd, uint32_t(h->elem_bytes));
  h->field_bytes = opt->storage == FLB_STORAGE_FP64 ? 8 : 4;
  h->inv_tau = 1.0f / opt->tau;
  auto bail = [&](int rc) {
    flb_destroy(h);
    return rc;
  };
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return bail(FLB_ERR_DEVICE);
  }
  if (opt->device < 0 || opt->device >= ndev) return bail(FLB_ERR_ARGUMENT);
  if (hipSetDevice(opt->device) != hipSuccess) return bail(FLB_ERR_DEVICE);
  // Only flb_flag_histogram_noalloc needs the scratch; it reports a failed
  // lookup itself, so the handle is usable either way.
  h->hist_scratch_err = scratch_address(opt->device, &h->hist_scratch);
  if (h->hist_scratch_err != hipSuccess) (void)hipGetLastError();
  if (hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, opt->device) !=
          hipSuccess || h->cus <= 0)
    h->cus = 256;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(FLB_ERR_DEVICE);

  int rc = FLB_OK;
  uint8_t* fbytes = nullptr;
  if ((rc = dev_alloc(h, &fbytes, 9 * h->Np * h->elem_bytes))) return bail(rc);
  h->f = fbytes;
  const bool fs = opt->problem == FLB_FREE_SURFACE;
  const uint64_t Nd = h->Nd;
  if ((rc = dev_alloc(h, &h->flags[0], Nd))) return bail(rc);
  if ((rc = dev_alloc(h, &h->flags[1], fs ? Nd : 1))) return bail(rc);
  if ((rc = field_alloc(h, &h->rho, Nd))) return bail(rc);
  if ((rc = field_alloc(h, &h->u, 2 * Nd))) return bail(rc);
  for (int k = 0; k < 2; ++k) {
    if ((rc = field_alloc(h, &h->pub_rho[k], Nd))) return bail(rc);
    if ((rc = field_alloc(h, &h->pub_u[k], 2 * Nd))) return bail(rc);
  }
  if (fs) {
    if ((rc = field_alloc(h, &h->mass, Nd))) return bail(rc);
    if ((rc = field_alloc(h, &h->excess[0], Nd))) return bail(rc);
    if ((rc = field_alloc(h, &h->excess[1], Nd))) return bail(rc);
    if ((rc = field_alloc(h, &h->phi, Nd))) return bail(rc);
    if ((rc = field_alloc(h, &h->rec, 9 * Nd))) return bail(rc);
    // Work words, lists and stamps of the free-surface kernels.
    h->work_x = (nx + flb::kFsTileW - 1) / flb::kFsTileW;
    h->tile_rows = (ny + flb::kFsTileH - 1) / flb::kFsTileH;
    const size_t tiles = size_t(h->work_x) * h->tile_rows;
    const size_t chunks = size_t(h->work_x) * ny;
    if ((rc = dev_alloc(h, &h->tile_work, tiles))) return bail(rc);
    if ((rc = dev_alloc(h, &h->tile_pub, tiles))) return bail(rc);
    if ((rc = dev_alloc(h, &h->tile_list, tiles))) return bail(rc);
    if ((rc = dev_alloc(h, &h->tile_body, tiles))) return bail(rc);
    if ((rc = dev_alloc(h, &h->tile_counts, 2))) return bail(rc);
    if ((rc = dev_alloc(h, &h->counts, 8))) return bail(rc);
    if ((rc = dev_alloc(h, &h->box, 4 * ((tiles + 63) / 64)))) return bail(rc);
    // A domain edge whose first or last row (column) is all wall: work cannot
    // spread across that periodic seam (walls never change), so the
    // stream-collide window need not wrap there.
    {
      bool row0 = true, row1 = true, col0 = true, col1 = true;
      for (uint32_t x = 0; x < nx; ++x) {
        row0 = row0 && (flags[x] & FLB_WALL);
        row1 = row1 && (flags[uint64_t(ny - 1) * nx + x] & FLB_WALL);
      }
      for (uint32_t y = 0; y < ny; ++y) {
        col0 = col0 && (flags[uint64_t(y) * nx] & FLB_WALL);
        col1 = col1 && (flags[uint64_t(y) * nx + nx - 1] & FLB_WALL);
      }
      h->seam_y_closed = row0 || row1;
      h->seam_x_closed = col0 || col1;
    }
    {
      void* hb = nullptr;
      if (hipHostMalloc(&hb, 2 * sizeof(unsigned long long),
                        hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess) {
        (void)hipGetLastError();
        return bail(FLB_ERR_MEMORY);
      }
      h->box_host = static_cast<unsigned long long*>(hb);
      h->box_host[0] = 0ull;
      h->box_host[1] = ~0ull;
      void* hd = nullptr;
      if (hipHostGetDevicePointer(&hd, hb, 0) != hipSuccess || !hd) {
        (void)hipGetLastError();
        return bail(FLB_ERR_DEVICE);
      }
      h->box_host_dev = static_cast<unsigned long long*>(hd);
      // The ring only tells the host that a step has finished (the box it
      // reads is published by the update kernel behind its own system-scope
      // fence), so its records skip the system-scope release (an L2
      // write-back per record).
      for (hipEvent_t& e : h->in_flight)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventDisableSystemFence) !=
            hipSuccess) {
          (void)hipGetLastError();
          e = nullptr;
          return bail(FLB_ERR_DEVICE);
        }
    }
    for (int k = 0; k < 2; ++k)
      if ((rc = dev_alloc(h, &h->fs1_list[k], chunks))) return bail(rc);
    std::vector<uint32_t> pub(tiles, 0u);
    for (uint64_t n = 0; n < N; ++n) {
      const bool pub_r = n > 0 && (flags[n - 1] & FLB_INLET);
      const bool pub_v = n + 1 < N && (flags[n + 1] & FLB_OUTLET) && !(flags[n] & FLB_INLET);
      const uint64_t x = n % nx, y = n / nx;
      if (pub_r || pub_v)
        pub[(y / flb::kFsTileH) * h->work_x + x / flb::kFsTileW] |=
            1u << (flb::kWorkSC + int(y % flb::kFsTileH));
    }
    if ((rc = upload(h, h->tile_pub, pub.data(), tiles * sizeof(uint32_t)))) return bail(rc);
    if ((rc = reset_work(h))) return bail(rc);
  }
  if ((rc = dev_alloc(h, &h->rho_init, Nd))) return bail(rc);
  if ((rc = dev_alloc(h, &h->u_init, 2 * Nd))) return bail(rc);
  const flb::Params pu = params(h, kUtilBlock);
  h->partial_len = pu.nblocks;
  if ((rc = dev_alloc(h, &h->partial, h->partial_len))) return bail(rc);

  // Fields: rho = rho0 (or 1), u = u0 (or 0); the publication buffers start
  // with the same values (what the reference's fields hold before step 0).
  std::vector<float> hr(N, 1.0f), hu(2 * N, 0.0f);
  if (rho0) std::memcpy(hr.data(), rho0, N * sizeof(float));
  if (u0) std::memcpy(hu.data(), u0, 2 * N * sizeof(float));
  if ((rc = cells_in(h, h->flags[0], flags, 1, 1))) return bail(rc);
  if (fs && (rc = cells_in(h, h->flags[1], flags, 1, 1))) return bail(rc);
  if ((rc = upload_field(h, h->rho, hr.data(), 1))) return bail(rc);
  if ((rc = upload_field(h, h->u, hu.data(), 2))) return bail(rc);
  if ((rc = cells_in(h, h->rho_init, hr.data(), sizeof(float), 1))) return bail(rc);
  if ((rc = cells_in(h, h->u_init, hu.data(), sizeof(float), 2))) return bail(rc);
  for (int k = 0; k < 2; ++k) {
    if ((rc = upload_field(h, h->pub_rho[k], hr.data(), 1))) return bail(rc);
    if ((rc = upload_field(h, h->pub_u[k], hu.data(), 2))) return bail(rc);
  }
  // Populations: shifted feq, converted to the storage format on the device
  // (from double for FP64 storage, from float otherwise).
  {
    const bool f64 = opt->storage == FLB_STORAGE_FP64;
    std::vector<double> fsd;
    std::vector<float> fsf;
    if (f64) initial_populations(nx, ny, rho0, u0, fsd);
    else initial_populations(nx, ny, rho0, u0, fsf);
    const size_t bytes = f64 ? fsd.size() * sizeof(double) : fsf.size() * sizeof(float);
    void* staging = nullptr;
    if (hipMalloc(&staging, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return bail(FLB_ERR_MEMORY);
    }
    int r2 = upload(h, staging, f64 ? static_cast<const void*>(fsd.data()) : fsf.data(), bytes);
    if (!r2) {
      const flb::Params p = params(h, kUtilBlock);
      if (f64)
        hipLaunchKernelGGL(flb::get_f64_to_pop(opt->storage), dim3(p.nblocks), dim3(kUtilBlock), 0,
                           h->stream, p, static_cast<const double*>(staging));
      else
        hipLaunchKernelGGL(flb::get_f32_to_pop(opt->storage), dim3(p.nblocks), dim3(kUtilBlock), 0,
                           h->stream, p, static_cast<const float*>(staging));
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
      if (e != hipSuccess) r2 = FLB_ERR_DEVICE;
    }
    (void)hipFree(staging);
    if (r2) return bail(r2);
  }
  *out = h;
  return FLB_OK;
}

}  // extern "C"

namespace {

// The 16-bit storage residual that stream-collide keeps in plane 0 of rec
// belongs to the stored populations, which new masses do not change, so it
// is kept here (flb_write_populations clears it).
template <class T>
int set_fs_fields(flb_handle* h, const T* mass, const T* phi, const T* excess) {
  if (!h || !mass || !phi) return h ? fail(h, FLB_ERR_ARGUMENT, "mass and phi required") : FLB_ERR_ARGUMENT;
  if (h->opt.problem != FLB_FREE_SURFACE) return fail(h, FLB_ERR_STATE, "not a free-surface handle");
  int rc;
  if ((rc = upload_field(h, h->mass, mass, 1))) return rc;
  if ((rc = upload_field(h, h->phi, phi, 1))) return rc;
  if ((rc = reset_work(h))) return rc;  // new masses: evaluate everything next step
  // Shares pulled by the next step; the other buffer receives the next shares.
  if ((rc = zero_field(h, h->excess[h->cur ^ 1], h->Nd))) return rc;
  if (excess) return upload_field(h, h->excess[h->cur], excess, 1);
  return zero_field(h, h->excess[h->cur], h->Nd);
}

}  // namespace

extern "C" {

int flb_set_fs_fields(flb_handle* h, const float* mass, const float* phi, const float* excess) {
  return set_fs_fields(h, mass, phi, excess);
}
int flb_set_fs_fields_f64(flb_handle* h, const double* mass, const double* phi,
                          const double* excess) {
  return set_fs_fields(h, mass, phi, excess);
}

// Window of the free-surface stream-collide launch for step h->t.  The last
// box the update published (step s <= t) bounds the chunks with
// stream-collide work at step s; each step can add work only one cell beyond
// it (a gas cell turns interface only next to an interface cell turning
// fluid), so at step t the work lies in the box widened by t - s cells.  A
// widened box that reaches the domain edge takes the whole extent in that
// direction (the lattice is periodic), unless a wall row (column) closes that
// seam; an empty box means no work at all (nblocks 0).  Without a published
// box the window is the whole grid.
void stream_collide_window(const flb_handle* h, flb::Params& p, int bx) {
  if (h->opt.whole_grid || !h->box_host || h->work_x >= 0xffffu || h->tile_rows >= 0xffffu)
    return;
  const volatile unsigned long long* bh = h->box_host;
  const unsigned long long s = bh[1];  // the step first, then the box
  if (s == ~0ull || s > h->t) return;
  const unsigned long long b = bh[0];
  const int64_t ylo_t = int64_t(b & 0xffffu), yhi_t = int64_t((b >> 16) & 0xffffu);
  const int64_t xlo_t = int64_t((b >> 32) & 0xffffu), xhi_t = int64_t((b >> 48) & 0xffffu);
  if (ylo_t > yhi_t || xlo_t > xhi_t) {
    p.nblocks = 0;
    return;
  }
  const int64_t k = int64_t(h->t - s);
  const int64_t ny = h->ny, nx = h->nx;
  const int64_t ylo = ylo_t * flb::kFsTileH - k, yhi = (yhi_t + 1) * flb::kFsTileH - 1 + k;
  const int64_t xlo = xlo_t * flb::kFsTileW - k, xhi = (xhi_t + 1) * flb::kFsTileW - 1 + k;
  int64_t band0 = 0, band1 = (ny - 1) / flb::kBandRows;
  if (h->seam_y_closed || (ylo >= 0 && yhi < ny)) {
    band0 = std::max<int64_t>(ylo, 0) / flb::kBandRows;
    band1 = std::min<int64_t>(yhi, ny - 1) / flb::kBandRows;
  }
  int64_t col0 = 0, col1 = int64_t(p.tiles_x) - 1;
  if (h->seam_x_closed || (xlo >= 0 && xhi < nx)) {
    col0 = std::max<int64_t>(xlo, 0) / bx;
    col1 = std::min<int64_t>(std::min<int64_t>(xhi, nx - 1) / bx, col1);
  }
  const int64_t rows = std::min<int64_t>(ny, (band1 + 1) * flb::kBandRows) - band0 * flb::kBandRows;
  p.band0 = uint32_t(band0);
  p.col0 = uint32_t(col0);
  p.ncols = uint32_t(col1 - col0 + 1);
  p.win_magic = flb::div_magic(uint64_t(flb::kBandRows) * p.ncols);
  p.nblocks = uint32_t(int64_t(p.ncols) * rows);
  flb::set_walk(p, uint32_t(rows), p.ncols);
}

// One lattice step on the handle's stream.  ev (optional) receives per-kernel
// event pairs: ev[2k], ev[2k+1] around kernel k.
static int enqueue_step(flb_handle* h, bool moments, hipEvent_t* ev) {
  const int bx = block_x(h);
  const int odd = int(h->t & 1u);
  const int st = h->opt.storage;
  const bool fs = h->opt.problem == FLB_FREE_SURFACE;
  flb::Params p = params(h, bx);
  auto run = [&](int k, flb::KernelFn fn, const flb::Params& q, int threads) -> int {
    if (ev) FLB_HIP(h, hipEventRecord(ev[2 * k], h->stream));
    int rc = launch(h, fn, q, threads);
    if (rc) return rc;
    if (ev) FLB_HIP(h, hipEventRecord(ev[2 * k + 1], h->stream));
    return FLB_OK;
  };
  int rc;
  if (fs) {
    // Pass-1 list of this step (built by the last update) and of the next.
    const int rd = int(h->t & 1u), wr = rd ^ 1;
    p.fs1_list = h->fs1_list[rd];
    p.counts_in = h->counts + 4 * rd;
    p.fs1_list_out = h->fs1_list[wr];
    p.counts_out = h->counts + 4 * wr;
    p.tile_count = h->tile_counts + rd;
    p.tile_count_next = h->tile_counts + wr;
    flb::Params q = p;  // pass 1: a grid of resident blocks walks the list
    q.nblocks = uint32_t(flb::kListBlocksPerCU * h->cus);
    if ((rc = run(FLB_K_FS1, flb::get_fs1(st, odd), q, flb::kListThreads))) return rc;
    if (h->opt.mass_mode != FLB_MASS_CONSERVATIVE) {
      if ((rc = run(FLB_K_STREAM_COLLIDE, flb::get_stream_collide(st, odd, true, moments), p, bx)))
        return rc;
    } else if (flb::KernelFn all = flb::get_stream_collide_all(st, odd, moments)) {
      // Surface chunks, interior fluid and bodies of fluid in one launch over
      // the window that can hold work.
      flb::Params w = p;
      stream_collide_window(h, w, bx);
      if (ev) FLB_HIP(h, hipEventRecord(ev[2 * FLB_K_STREAM_COLLIDE], h->stream));
      if (w.nblocks && (rc = launch(h, all, w, bx))) return rc;
      if (ev) FLB_HIP(h, hipEventRecord(ev[2 * FLB_K_STREAM_COLLIDE + 1], h->stream));
    } else {
      // FP64: bodies of fluid in one launch over the window, the surface
      // chunks (the pass-1 list) in a second one.
      flb::Params w = p;
      stream_collide_window(h, w, bx);
      if (ev) FLB_HIP(h, hipEventRecord(ev[2 * FLB_K_STREAM_COLLIDE], h->stream));
      if (w.nblocks && (rc = launch(h, flb::get_stream_collide_dense(st, odd, moments), w, bx)))
        return rc;
      if ((rc = launch(h, flb::get_stream_collide_surface(st, odd, moments), q, flb::kListThreads)))
        return rc;
      if (ev) FLB_HIP(h, hipEventRecord(ev[2 * FLB_K_STREAM_COLLIDE + 1], h->stream));
    }
    q = p;  // passes 2-4: blocks of kFsTileW x kFsTileH threads walk the tile list
    q.tiles_x = h->work_x;
    q.nblocks = uint32_t(flb::kUpdateBlocksPerCU * h->cus);
    if ((rc = run(FLB_K_FS_UPDATE, flb::get_fs_update(st, odd, h->nx, h->ny), q,
                  flb::kFsTileW * flb::kFsTileH)))
      return rc;
    h->cur ^= 1;  // the update wrote the other flag and share buffers
  } else {
    if ((rc = run(FLB_K_STREAM_COLLIDE, flb::get_stream_collide(st, odd, false, moments), p, bx)))
      return rc;
  }
  h->t += 1;
  return FLB_OK;
}

}  // extern "C"

namespace flb {
void host_initial_populations(uint32_t nx, uint32_t ny, const float* rho0, const float* u0,
                              std::vector<float>& fs) {
  initial_populations<float>(nx, ny, rho0, u0, fs);
}
int set_error(flb_handle* h, int code, const std::string& msg) { return fail(h, code, msg); }
Params make_params(const flb_handle* h, int bx) { return params(h, bx); }
int handle_block_x(const flb_handle* h) { return block_x(h); }
int enqueue_step_ex(flb_handle* h, bool moments, hipEvent_t* ev) {
  return enqueue_step(h, moments, ev);
}
int copy_cells_in(flb_handle* h, void* dev, const void* host, size_t elem, int planes) {
  return cells_in(h, dev, host, elem, planes);
}
int copy_cells_out(flb_handle* h, void* host, const void* dev, size_t elem, int planes) {
  return cells_out(h, host, dev, elem, planes);
}
}  // namespace flb

extern "C" {

int flb_step(flb_handle* h, uint64_t nsteps) {
  if (!h) return FLB_ERR_ARGUMENT;
  const bool throttle = h->box_host != nullptr;
  constexpr uint64_t G = flb_handle::kGroup, R = flb_handle::kRing;
  for (uint64_t k = 0; k < nsteps; ++k) {
    // Free surface: steps go in groups of G (by lattice time, so across calls
    // too).  Starting group g, wait for the event recorded after group g - 2
    // (a slot never recorded, or recorded by an older group, returns at once),
    // so that at most 2 G steps are queued and the published stream-collide
    // box is at most about 2 G steps old.  One event per group, not per step:
    // every event record is a queue marker the command processor drains, and
    // with one per step the next step's first kernel started 4.3 us after the
    // last one ended (8192 x 2048 dam break, kernel trace; no gap without).
    const uint64_t t = h->t, g = t / G;
    if (throttle && t % G == 0 && g >= 2) FLB_HIP(h, hipEventSynchronize(h->in_flight[(g - 2) % R]));
    const int rc = enqueue_step(h, k + 1 == nsteps, nullptr);
    if (rc) return rc;
    if (throttle && t % G == G - 1) FLB_HIP(h, hipEventRecord(h->in_flight[g % R], h->stream));
  }
  return FLB_OK;
}

int flb_sync(flb_handle* h) {
  if (!h) return FLB_ERR_ARGUMENT;
  FLB_HIP(h, hipStreamSynchronize(h->stream));
  return FLB_OK;
}

int flb_step_timed(flb_handle* h, uint64_t nsteps, float* total_ms, float* per_kernel_ms) {
  if (!h) return FLB_ERR_ARGUMENT;
  if (!per_kernel_ms) {
    hipEvent_t a, b;
    FLB_HIP(h, hipEventCreate(&a));
    FLB_HIP(h, hipEventCreate(&b));
    FLB_HIP(h, hipEventRecord(a, h->stream));
    int rc = flb_step(h, nsteps);
    if (!rc) {
      FLB_HIP(h, hipEventRecord(b, h->stream));
      FLB_HIP(h, hipEventSynchronize(b));
      float ms = 0.0f;
      FLB_HIP(h, hipEventElapsedTime(&ms, a, b));
      if (total_ms) *total_ms = ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
  }
  // Per-kernel mode: every launch of every step between two events of its
  // own (a free-surface handle waits for the step kInFlight back, keeping as
  // many steps queued as flb_step at most).  The events skip the system-scope release fence
  // (hipEventDisableSystemFence): with it each record wrote back and
  // invalidated the L2, and the next kernel started cold.
  const bool fs = h->opt.problem == FLB_FREE_SURFACE;
  const bool throttle = h->box_host != nullptr;
  constexpr uint64_t K = flb_handle::kInFlight;
  const size_t per = 2 * FLB_NUM_KERNELS;
  std::vector<hipEvent_t> ev(per * nsteps, nullptr);
  int rc = FLB_OK;
  for (auto& e : ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
      (void)hipGetLastError();
      e = nullptr;
      rc = fail(h, FLB_ERR_DEVICE, "hipEventCreateWithFlags failed");
      break;
    }
  const int last_k = fs ? FLB_K_FS_UPDATE : FLB_K_STREAM_COLLIDE;
  for (uint64_t s = 0; s < nsteps && !rc; ++s) {
    if (throttle && s >= K && hipEventSynchronize(ev[per * (s - K) + 2 * last_k + 1]) != hipSuccess)
      rc = fail(h, FLB_ERR_DEVICE, "hipEventSynchronize failed");
    if (!rc) rc = enqueue_step(h, s + 1 == nsteps, ev.data() + per * s);
  }
  for (int k = 0; k < FLB_NUM_KERNELS; ++k) per_kernel_ms[k] = 0.0f;
  float tot = 0.0f;
  if (!rc && nsteps && hipEventSynchronize(ev[per * (nsteps - 1) + 2 * last_k + 1]) != hipSuccess)
    rc = fail(h, FLB_ERR_DEVICE, "hipEventSynchronize failed");
  for (uint64_t s = 0; s < nsteps && !rc; ++s)
    for (int k = 0; k < FLB_NUM_KERNELS; ++k) {
      if (!fs && k != FLB_K_STREAM_COLLIDE) continue;
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, ev[per * s + 2 * k], ev[per * s + 2 * k + 1]) != hipSuccess) {
        rc = fail(h, FLB_ERR_DEVICE, "hipEventElapsedTime failed");
        break;
      }
      per_kernel_ms[k] += ms;
      tot += ms;
    }
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
  if (total_ms) *total_ms = tot;
  return rc;
}

}  // extern "C"

namespace {
template <class T>
int read_fields(flb_handle* h, T* u, T* rho, uint8_t* flags, T* phi) {
  if (!h) return FLB_ERR_ARGUMENT;
  int rc;
  if (u && (rc = download_field(h, u, h->u, 2))) return rc;
  if (rho && (rc = download_field(h, rho, h->rho, 1))) return rc;
  if (flags && (rc = cells_out(h, flags, h->flags[h->cur], 1, 1))) return rc;
  if (phi) {
    if (h->opt.problem != FLB_FREE_SURFACE) return fail(h, FLB_ERR_STATE, "phi needs free surface");
    if ((rc = download_field(h, phi, h->phi, 1))) return rc;
  }
  return FLB_OK;
}

template <class T>
int read_fs_fields(flb_handle* h, T* mass, T* excess) {
  if (!h) return FLB_ERR_ARGUMENT;
  if (h->opt.problem != FLB_FREE_SURFACE) return fail(h, FLB_ERR_STATE, "not a free-surface handle");
  int rc;
  if (mass && (rc = download_field(h, mass, h->mass, 1))) return rc;
  if (excess && (rc = download_field(h, excess, h->excess[h->cur], 1))) return rc;
  return FLB_OK;
}

template <class T>
int read_pops(flb_handle* h, T* f) {
  if (!h || !f) return FLB_ERR_ARGUMENT;
  T* tmp = nullptr;
  FLB_HIP(h, hipMalloc(&tmp, 9 * h->N * sizeof(T)));
  const flb::Params p = params(h, kUtilBlock);
  if (std::is_same<T, double>::value)
    hipLaunchKernelGGL(flb::get_pop_to_f64(h->opt.storage), dim3(p.nblocks), dim3(kUtilBlock), 0,
                       h->stream, p, reinterpret_cast<double*>(tmp));
  else
    hipLaunchKernelGGL(flb::get_pop_to_f32(h->opt.storage), dim3(p.nblocks), dim3(kUtilBlock), 0,
                       h->stream, p, reinterpret_cast<float*>(tmp));
  hipError_t e = hipGetLastError();
  int rc = e == hipSuccess ? download(h, f, tmp, 9 * h->N * sizeof(T)) : hip_fail(h, e, "pop_out");
  (void)hipFree(tmp);
  return rc;
}

}  // namespace

extern "C" {

int flb_read(flb_handle* h, float* u, float* rho, uint8_t* flags, float* phi) {
  return read_fields(h, u, rho, flags, phi);
}
int flb_read_f64(flb_handle* h, double* u, double* rho, uint8_t* flags, double* phi) {
  return read_fields(h, u, rho, flags, phi);
}
int flb_read_fs(flb_handle* h, float* mass, float* excess) { return read_fs_fields(h, mass, excess); }
int flb_read_fs_f64(flb_handle* h, double* mass, double* excess) {
  return read_fs_fields(h, mass, excess);
}
int flb_read_populations(flb_handle* h, float* f) { return read_pops(h, f); }
int flb_read_populations_f64(flb_handle* h, double* f) { return read_pops(h, f); }

}  // extern "C"

namespace {

// Shifted populations f - w in T, converted to the storage format on the
// device.  A free-surface handle's 16-bit storage residual (plane 0 of rec)
// belonged to the replaced populations and is cleared.
template <class T>
int write_pops(flb_handle* h, const T* f) {
  if (!h || !f) return FLB_ERR_ARGUMENT;
  const size_t N = h->N;
  std::vector<T> fsh(9 * N);
  for (int q = 0; q < 9; ++q) {
    const T w = q == 0 ? T(4) / T(9) : (q < 5 ? T(1) / T(9) : T(1) / T(36));
    for (size_t n = 0; n < N; ++n) fsh[q * N + n] = f[q * N + n] - w;
  }
  T* tmp = nullptr;
  FLB_HIP(h, hipMalloc(&tmp, 9 * N * sizeof(T)));
  int rc = upload(h, tmp, fsh.data(), 9 * N * sizeof(T));
  if (!rc) {
    const flb::Params p = params(h, kUtilBlock);
    if (std::is_same<T, double>::value)
      hipLaunchKernelGGL(flb::get_f64_to_pop(h->opt.storage), dim3(p.nblocks), dim3(kUtilBlock), 0,
                         h->stream, p, reinterpret_cast<const double*>(tmp));
    else
      hipLaunchKernelGGL(flb::get_f32_to_pop(h->opt.storage), dim3(p.nblocks), dim3(kUtilBlock), 0,
                         h->stream, p, reinterpret_cast<const float*>(tmp));
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = hip_fail(h, e, "write_populations");
  }
  (void)hipFree(tmp);
  if (!rc && h->opt.problem == FLB_FREE_SURFACE) rc = zero_field(h, h->rec, h->Nd);
  return rc;
}

}  // namespace

extern "C" {

int flb_write_populations(flb_handle* h, const float* f) { return write_pops(h, f); }
int flb_write_populations_f64(flb_handle* h, const double* f) { return write_pops(h, f); }

int flb_device_fields(flb_handle* h, void** d_u, void** d_rho, void** d_flags, void** d_phi) {
  if (!h) return FLB_ERR_ARGUMENT;
  if (d_u) *d_u = h->u;
  if (d_rho) *d_rho = h->rho;
  if (d_flags) *d_flags = h->flags[h->cur];
  if (d_phi) *d_phi = h->phi;
  return FLB_OK;
}

int flb_mass(flb_handle* h, double* total) {
  if (!h || !total) return FLB_ERR_ARGUMENT;
  const flb::Params p = params(h, kUtilBlock);
  const bool fs = h->opt.problem == FLB_FREE_SURFACE;
  hipLaunchKernelGGL(fs ? flb::get_fs_mass(h->opt.storage) : flb::get_pop_mass(h->opt.storage), dim3(p.nblocks),
                     dim3(kUtilBlock), 0, h->stream, p, h->partial);
  FLB_HIP(h, hipGetLastError());
  std::vector<double> part(h->partial_len);
  int rc = download(h, part.data(), h->partial, part.size() * sizeof(double));
  if (rc) return rc;
  double s = 0.0;
  for (double v : part) s += v;
  *total = fs ? s : s + double(h->N);  // monofluid: shifted sum + N * sum(w)
  return FLB_OK;
}

int flb_flag_histogram(flb_handle* h, uint64_t* counts) {
  if (!h || !counts) return FLB_ERR_ARGUMENT;
  unsigned long long* d = nullptr;
  FLB_HIP(h, hipMalloc(&d, 256 * sizeof(unsigned long long)));
  int rc = FLB_OK;
  hipError_t e = hipMemsetAsync(d, 0, 256 * sizeof(unsigned long long), h->stream);
  if (e == hipSuccess) {
    const flb::Params p = params(h, kUtilBlock);
    const uint32_t blocks = p.nblocks < 2048u ? p.nblocks : 2048u;  // grid-stride kernel
    hipLaunchKernelGGL(flb::get_flag_hist(), dim3(blocks), dim3(256), 0, h->stream, p, d);
    e = hipGetLastError();
  }
  if (e != hipSuccess) rc = hip_fail(h, e, "flag histogram");
  else rc = download(h, counts, d, 256 * sizeof(unsigned long long));
  (void)hipFree(d);
  return rc;
}

int flb_flag_histogram_noalloc(flb_handle* h, uint64_t* counts) {
  if (!h || !counts) return FLB_ERR_ARGUMENT;
  // The histogram has one counter per flag value, 256 whatever the grid, and
  // the scratch holds exactly that many; nothing a caller passes can exceed it.
  constexpr uint32_t kBins = 256;
  static_assert(kBins <= kScratchCounters, "histogram scratch too small");
  unsigned long long* d = h->hist_scratch;
  if (!d)
    return fail(h, FLB_ERR_DEVICE, std::string("hipGetSymbolAddress of the histogram scratch: ") +
                                       hipGetErrorString(h->hist_scratch_err));
  // Cleared by a kernel (not a memset node), as the last call left its counts.
  hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(kBins), 0, h->stream, d, kBins);
  const flb::Params p = params(h, kUtilBlock);
  const uint32_t blocks = p.nblocks < 2048u ? p.nblocks : 2048u;  // grid-stride kernel
  hipLaunchKernelGGL(flb::get_flag_hist(), dim3(blocks), dim3(256), 0, h->stream, p, d);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(h, e, "flag histogram");
  return download(h, counts, d, kBins * sizeof(unsigned long long));
}

}  // extern "C"
