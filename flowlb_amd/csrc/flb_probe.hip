// flb_probe.hip -- attainable-bandwidth reference for the stream-collide
// kernel: the same memory pattern with no arithmetic, on a buffer of its own.
//
// FLB_PROBE_ESOTERIC moves every cell's nine populations through the
// esoteric-pull slots exactly as stream-collide does (nine loads from the
// cell's load slots, nine stores to its store slots, in place, parity
// alternating per launch, and the odd launches taking the bands last to first,
// as stream-collide does), with a multiply by 1 as "collision";
// FLB_PROBE_ESOTERIC16 does the same with 16-bit values (the FP16S / BF16S
// pattern).  FLB_PROBE_COPY copies nine planes to nine other planes,
// FLB_PROBE_INPLACE adds 1 to nine planes in place.  One thread per cell,
// 256-cell row blocks in stream-collide's band order, planes laid out as a
// handle's (row pitch, plane stride flb::plane_stride); 18 values moved per
// cell: 72 bytes in FP32, 36 in 16-bit.  No reference counterpart (the
// reference reports no bandwidth).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "flb_cell.h"
#include "flb_internal.h"

namespace {

constexpr int kProbeBlock = 256;

struct ProbeArgs {
  flb::Params p;  // nx, ny, pitch, tiles_x, Np used
  void* f;
  void* g;        // second buffer (copy)
  // 1.0, multiplied into every value moved.  A value stored back unchanged
  // to the address it was loaded from is dropped by the compiler (the rest
  // population of the esoteric pattern: one ninth of the traffic), so the
  // "collision" is a multiply by a factor it cannot see.
  float one;
};

template <class T, int KIND, int ODD>
__global__ void __launch_bounds__(kProbeBlock) k_probe(const ProbeArgs a) {
  uint32_t tx, y;
  // Block order as stream-collide's (walk_block): odd launches take the bands
  // last to first.
  const uint32_t b = ODD ? flb::band_reversed(blockIdx.x, a.p.walk_last, a.p.walk_bands,
                                              flb::kBandRows * a.p.tiles_x, a.p.band_magic)
                         : blockIdx.x;
  flb::block_tile(b, a.p.tiles_x, a.p.ny, a.p.band_magic, a.p.last_magic, tx, y);
  const uint32_t x = tx * kProbeBlock + threadIdx.x;
  if (x >= a.p.nx) return;
  const uint64_t Np = a.p.Np;
  T* const f = static_cast<T*>(a.f);
  if (KIND == FLB_PROBE_ESOTERIC || KIND == FLB_PROBE_ESOTERIC16) {
    uint32_t nb[9];
    flb::neighbours(a.p, x, y, nb);
    float v[9];
    v[0] = float(f[nb[0]]);
#pragma unroll
    for (int i = 1; i < 9; i += 2) {
      v[i] = float(f[(ODD ? i : i + 1) * Np + nb[0]]);
      v[i + 1] = float(f[(ODD ? i + 1 : i) * Np + nb[i]]);
    }
#pragma unroll
    for (int i = 0; i < 9; i++) v[i] *= a.one;
    f[nb[0]] = T(v[0]);
#pragma unroll
    for (int i = 1; i < 9; i += 2) {
      f[(ODD ? i + 1 : i) * Np + nb[i]] = T(v[i]);
      f[(ODD ? i : i + 1) * Np + nb[0]] = T(v[i + 1]);
    }
  } else {
    T* const g = static_cast<T*>(a.g);
    const uint64_t n = x + uint64_t(y) * a.p.pitch;
    float v[9];
#pragma unroll
    for (int q = 0; q < 9; q++) v[q] = float(f[q * Np + n]) * a.one;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      if (KIND == FLB_PROBE_COPY) g[q * Np + n] = T(v[q]);
      else f[q * Np + n] = T(v[q] + 1.0f);
    }
  }
}

template <class T, int KIND>
void launch_probe(const ProbeArgs& a, uint32_t blocks, int odd, hipStream_t s) {
  if (odd) hipLaunchKernelGGL((k_probe<T, KIND, 1>), dim3(blocks), dim3(kProbeBlock), 0, s, a);
  else hipLaunchKernelGGL((k_probe<T, KIND, 0>), dim3(blocks), dim3(kProbeBlock), 0, s, a);
}

}  // namespace

extern "C" int flb_bandwidth_probe(int32_t device, uint32_t nx, uint32_t ny, int32_t kind,
                                   uint32_t launches, float* ms_per_launch,
                                   double* bytes_per_launch) {
  if (nx < 2 || ny < 2 || launches == 0 || !ms_per_launch ||
      (kind != FLB_PROBE_ESOTERIC && kind != FLB_PROBE_COPY && kind != FLB_PROBE_INPLACE &&
       kind != FLB_PROBE_ESOTERIC16))
    return FLB_ERR_ARGUMENT;
  const uint32_t elem = kind == FLB_PROBE_ESOTERIC16 ? 2 : 4;
  ProbeArgs a{};
  a.p.nx = nx;
  a.p.ny = ny;
  a.p.pitch = (nx + 63u) / 64u * 64u;
  a.p.tiles_x = (nx + kProbeBlock - 1) / kProbeBlock;
  a.p.Np = flb::plane_stride(uint64_t(a.p.pitch) * ny, elem);
  a.p.band_magic = flb::div_magic(uint64_t(flb::kBandRows) * a.p.tiles_x);
  a.p.last_magic = flb::div_magic(flb::last_band_rows(ny));
  flb::set_walk(a.p, ny, a.p.tiles_x);
  a.one = 1.0f;
  if ((uint64_t(a.p.pitch) + 128) * ny > 0xffffffffull / 2) return FLB_ERR_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return FLB_ERR_DEVICE;
  const size_t bytes = 9 * a.p.Np * elem;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FLB_OK;
  if (hipMalloc(&a.f, bytes) != hipSuccess) return FLB_ERR_MEMORY;
  if (kind == FLB_PROBE_COPY && hipMalloc(&a.g, bytes) != hipSuccess) rc = FLB_ERR_MEMORY;
  if (!rc && (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
              hipEventCreate(&e1) != hipSuccess))
    rc = FLB_ERR_DEVICE;
  if (!rc && hipMemsetAsync(a.f, 0, bytes, s) != hipSuccess) rc = FLB_ERR_DEVICE;
  if (!rc) {
    const uint32_t blocks = a.p.tiles_x * ny;
    auto run = [&](uint32_t k) {
      for (uint32_t i = 0; i < k; ++i) {
        const int odd = int(i & 1u);
        switch (kind) {
          case FLB_PROBE_ESOTERIC: launch_probe<float, FLB_PROBE_ESOTERIC>(a, blocks, odd, s); break;
          case FLB_PROBE_COPY: launch_probe<float, FLB_PROBE_COPY>(a, blocks, odd, s); break;
          case FLB_PROBE_INPLACE: launch_probe<float, FLB_PROBE_INPLACE>(a, blocks, odd, s); break;
          default: launch_probe<_Float16, FLB_PROBE_ESOTERIC16>(a, blocks, odd, s); break;
        }
      }
    };
    run(4);  // warm-up
    hipEventRecord(e0, s);
    run(launches);
    hipEventRecord(e1, s);
    float ms = 0.0f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = FLB_ERR_DEVICE;
    *ms_per_launch = ms / float(launches);
    if (bytes_per_launch) *bytes_per_launch = 18.0 * elem * double(nx) * double(ny);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(a.f);
  if (a.g) (void)hipFree(a.g);
  return rc;
}
