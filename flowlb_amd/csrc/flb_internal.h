// flb_internal.h -- handle layout and helpers shared by the translation units
// of libflbcore.so (not part of the public ABI).
#pragma once

#include <string>
#include <vector>

#include "../../include/flb_core.h"
#include "flb_device.h"

namespace flb {
// Stride between population planes, in elements of elem_bytes: the plane
// rounded up to 8 KiB, plus 4 KiB.  A wavefront reads and writes the same
// offset of nine planes; with planes a multiple of 8 KiB apart (any
// power-of-two grid) those nine streams meet on the same HBM channels.  The
// nine-plane esoteric pattern in 16-bit at 8192 x 2048 moved 5.88 TB/s
// unpadded and 6.32-6.39 TB/s with the planes 4-6 KiB past a multiple of 8 KiB
// (FP32 moved 5.3-5.45 TB/s either way).  The stride stays a multiple of 256
// bytes, so every plane starts 256-byte aligned.
inline uint64_t plane_stride(uint64_t cells, uint32_t elem_bytes) {
  const uint64_t bytes = (cells * elem_bytes + 8191u) / 8192u * 8192u + 4096u;
  return bytes / elem_bytes;
}

using KernelFn = void (*)(const Params);
using MassFn = void (*)(const Params, double*);
using ConvOutFn = void (*)(const Params, float*);
using ConvInFn = void (*)(const Params, const float*);
using ConvOut64Fn = void (*)(const Params, double*);
using ConvIn64Fn = void (*)(const Params, const double*);
using HistFn = void (*)(const Params, unsigned long long*);
KernelFn get_stream_collide(int storage, int odd, bool fs, bool mom);
KernelFn get_stream_collide_dense(int storage, int odd, bool mom);
KernelFn get_stream_collide_surface(int storage, int odd, bool mom);
KernelFn get_stream_collide_all(int storage, int odd, bool mom);  // nullptr for FP64
KernelFn get_fs1(int storage, int odd);
KernelFn get_fs_update(int storage, int odd, uint32_t nx, uint32_t ny);
KernelFn get_stream_collide_trace(int storage, int odd, bool mom);  // FP32 or FP64
MassFn get_pop_mass(int storage);
MassFn get_fs_mass(int storage);
ConvOutFn get_pop_to_f32(int storage);
ConvOut64Fn get_pop_to_f64(int storage);
ConvInFn get_f32_to_pop(int storage);
ConvIn64Fn get_f64_to_pop(int storage);
HistFn get_flag_hist();
}  // namespace flb

struct flb_handle {
  flb_options opt{};
  uint32_t nx = 0, ny = 0;
  uint32_t pitch = 0;  // row stride of the device planes (cells)
  uint64_t N = 0;      // cells, nx * ny (host arrays)
  uint64_t Nd = 0;     // device plane of a cell field, pitch * ny
  uint64_t Np = 0;     // population plane stride (elements), plane_stride(Nd, elem_bytes)
  int elem_bytes = 4;   // bytes per stored population
  int field_bytes = 4;  // bytes per element of the cell fields (8 for FP64)
  hipStream_t stream = nullptr;
  void* f = nullptr;
  uint8_t* flags[2] = {nullptr, nullptr};
  int cur = 0;
  float* rho = nullptr;
  float* u = nullptr;
  float* pub_rho[2] = {nullptr, nullptr};
  float* pub_u[2] = {nullptr, nullptr};
  float* mass = nullptr;
  float* excess[2] = {nullptr, nullptr};  // shares by step parity (index cur: read)
  float* phi = nullptr;
  float* rec = nullptr;
  uint32_t* tile_work = nullptr;  // free surface: per-tile work words (Params::tile_work)
  uint32_t* tile_pub = nullptr;   // free surface: chunks holding an inlet/outlet publisher
  uint32_t work_x = 0, tile_rows = 0;
  // Free-surface pass-1 lists by step parity (Params::fs1_list), their
  // lengths (4 words per parity) and the update's per-tile skip marks.
  uint32_t* fs1_list[2] = {nullptr, nullptr};
  uint32_t* counts = nullptr;
  uint32_t* tile_list = nullptr;  // tiles the update of this step evaluates
  uint32_t* tile_body = nullptr;  // tiles stream-collide updates (bodies of fluid)
  uint32_t* tile_counts = nullptr;  // its length, by step parity
  // Stream-collide launch window: the box of chunks with stream-collide work
  // (device, by step parity) and its copy in mapped host memory (box, step).
  uint32_t* box = nullptr;
  unsigned long long* box_host = nullptr;      // host view
  unsigned long long* box_host_dev = nullptr;  // its device address (Params::box_host)
  bool seam_y_closed = false, seam_x_closed = false;  // a wall row / column on the edge
  // flb_step keeps at most 2 kGroup steps queued, so that the box the host
  // sizes a launch from is at most a few steps old: it records an event after
  // every kGroup-th step (ring of kRing) and, starting a group, waits for the
  // event two groups back.
  static constexpr int kGroup = 4;
  static constexpr int kRing = 4;
  static constexpr int kInFlight = 2 * kGroup;  // queued steps at most
  hipEvent_t in_flight[kRing] = {};
  int cus = 256;  // compute units of the device (grid size of the list-driven pass 1)
  float* alpha = nullptr;     // Brinkman solid fraction (allocated on first use)
  float* rho_init = nullptr;  // rho0 and u0 as given to flb_create
  float* u_init = nullptr;
  double* partial = nullptr;
  // Device address of the file-scope histogram scratch on this handle's GPU
  // (flb_flag_histogram_noalloc); not owned.
  unsigned long long* hist_scratch = nullptr;
  hipError_t hist_scratch_err = hipSuccess;
  size_t partial_len = 0;
  uint64_t t = 0;
  float inv_tau = 0.0f;
  std::vector<void*> owned;
  std::string err;
};


namespace flb {
// Helpers defined in flb_core.hip.
int set_error(flb_handle* h, int code, const std::string& msg);
Params make_params(const flb_handle* h, int bx);
int handle_block_x(const flb_handle* h);
int enqueue_step_ex(flb_handle* h, bool moments, hipEvent_t* ev);
// Copies `planes` planes of cells of `elem` bytes between a host array
// (planes of nx * ny, row-major) and a pitched device array (planes of Nd),
// on the handle's stream, and waits for the copy.
int copy_cells_in(flb_handle* h, void* dev, const void* host, size_t elem, int planes);
int copy_cells_out(flb_handle* h, void* host, const void* dev, size_t elem, int planes);
// Shifted initial populations feq(rho0, u0) in the odd-step slot pattern.
void host_initial_populations(uint32_t nx, uint32_t ny, const float* rho0, const float* u0,
                              std::vector<float>& fs);
}  // namespace flb
