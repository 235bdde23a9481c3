// flb_view.hip -- device-side output for an interactive viewer: the velocity
// field as float2 texels and an isoline raster (marching squares) into an
// 8-bit image, both written into device buffers the caller owns (a GL pixel
// buffer mapped through HIP-GL interop, or any device allocation).
//
// Reference: CudaUtils::save2DDataToOpenGL (src/cuda/cudaUtils.cu:107-118)
// and d_CreateIsoSurfaceVector2D with its helpers
// (src/cuda/cudaIsosurface.cuh:33-198, src/cuda/cudaIsosurface.cu:16-86),
// launched per frame by the graphics loops (src/cuda/runCuda.cu:198-260,
// 396-430).  The reference writes through CUDA surface objects bound to GL
// textures; here the destinations are pitched linear buffers.
//
// Both kernels are HBM-bound and tiny next to a lattice step: the export
// reads 8 (16 for FP64) bytes and writes 8 bytes per cell; the isoline raster
// reads the two vertex rows of each square (8 bytes per cell and row,
// shared through L2 by neighbouring rows) and writes a few bytes per square
// the line crosses.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "flb_internal.h"

namespace {

constexpr int kViewBlock = 256;

// Texel (x, y) = (u_x, u_y) of cell (x, y); rows of dst_pitch bytes.  No
// flip: the reference writes surface row y from lattice row y.
template <class T>
__global__ void __launch_bounds__(kViewBlock)
    k_export_velocity(const T* __restrict__ u, uint64_t plane, uint32_t pitch, uint32_t nx,
                      float2* __restrict__ dst, size_t dst_pitch) {
  const uint32_t x = blockIdx.x * kViewBlock + threadIdx.x;
  const uint32_t y = blockIdx.y;
  if (x >= nx) return;
  const size_t n = x + size_t(y) * pitch;
  const float2 v = make_float2(float(u[n]), float(u[n + plane]));
  reinterpret_cast<float2*>(reinterpret_cast<char*>(dst) + size_t(y) * dst_pitch)[x] = v;
}

__global__ void __launch_bounds__(kViewBlock)
    k_fill_u8(uint8_t* __restrict__ dst, size_t pitch, uint32_t width) {
  const uint32_t x = blockIdx.x * kViewBlock + threadIdx.x;
  if (x < width) dst[size_t(blockIdx.y) * pitch + x] = 0;
}

struct IsoArgs {
  const void* field;  // u (two planes) or phi
  uint64_t plane;
  uint32_t pitch, nx, ny;
  float iso;
  uint8_t* tex;
  size_t tex_pitch;
  uint32_t width, height;
  uint32_t x_min, x_max, y_min, y_max;  // camera bounds as given
  uint32_t x0, y0;                      // first square of the launch
  float ppn_x, ppn_y;                   // texture pixels per lattice node
};

// The arithmetic below is kept in the reference's order with no fused
// multiply-adds, so that the rasterised pixels are reproducible.
#pragma clang fp contract(off)

// |u| of a vertex.  The reference calls hypotf; here the magnitude is
// rounded once from double (what C libraries' hypotf computes), so that it is
// reproducible bit for bit on the host.
__device__ inline float magnitude(float ux, float uy) {
  const double x = ux, y = uy;
  return float(sqrt(x * x + y * y));
}

// getInterpolatedPoint (cudaIsosurface.cu:16-21).
__device__ inline float2 interp(float2 p0, float2 p1, float v0, float v1, float iso) {
  const float a = (iso - v0) / (v1 - v0);
  const float b = 1.0f - a;
  return make_float2(b * p0.x + a * p1.x, b * p0.y + a * p1.y);
}

// drawLine (cudaIsosurface.cu:23-74): square coordinates to texture pixels
// (the y axis flipped once more, GL's origin being the lower left corner),
// then Bresenham's integer line from the first point up to, not including,
// the second (Zingl 2016).  Pixels outside the texture are skipped.
__device__ void draw_line(float2 p0, float2 p1, const IsoArgs& a, float xo, float yo) {
  const float xp0 = p0.x + xo, xp1 = p1.x + xo;
  const float yp0 = p0.y + yo, yp1 = p1.y + yo;
  int x0 = int(a.ppn_x * xp0 - 1.0f * float(int(xp0) != 0) + 0.5f);
  const int x1 = int(a.ppn_x * xp1 - 1.0f * float(int(xp1) != 0) + 0.5f);
  const float h = float(a.height);
  int y0 = int(h - a.ppn_y * yp0 - 0.5f);
  const int y1 = int(h - a.ppn_y * yp1 - 0.5f);
  const int dx = abs(x1 - x0);
  const int dy = -abs(y1 - y0);
  int err = dx + dy;
  const int sx = 2 * (x0 < x1) - 1;
  const int sy = 2 * (y0 < y1) - 1;
  while (x0 != x1 || y0 != y1) {
    if (x0 >= 0 && y0 >= 0 && x0 < int(a.width) && y0 < int(a.height))
      a.tex[size_t(y0) * a.tex_pitch + size_t(x0)] = 255;
    const int e2 = 2 * err;
    if (e2 >= dy) {
      err += dy;
      x0 += sx;
    }
    if (e2 <= dx) {
      err += dx;
      y0 += sy;
    }
  }
}

#pragma clang fp contract(on)

// Segments of each marching-squares case (cudaIsosurface.cuh:93-196), as
// pairs of edges in drawing order.  Edge 0: vertices 0-1 (top), 1: 1-2
// (right), 2: 3-2 (bottom), 3: 0-3 (left).  Cases 5 and 10 (saddles) draw
// two segments; 0 and 15 none.  Packed: a segment per byte (low byte
// first), its first edge in the low nibble; a byte 0xFF is no segment.
__constant__ uint16_t kSegments[16] = {
    0xFFFF, 0xFF30, 0xFF10, 0xFF13, 0xFF21, 0x1230, 0xFF20, 0xFF23,
    0xFF23, 0xFF20, 0x1023, 0xFF21, 0xFF13, 0xFF10, 0xFF30, 0xFFFF};

__device__ inline float2 edge_point(int e, const float* v, float iso) {
  const float2 p0 = make_float2(0.0f, 0.0f), p1 = make_float2(1.0f, 0.0f);
  const float2 p2 = make_float2(1.0f, 1.0f), p3 = make_float2(0.0f, 1.0f);
  switch (e) {
    case 0: return interp(p0, p1, v[0], v[1], iso);
    case 1: return interp(p1, p2, v[1], v[2], iso);
    case 2: return interp(p3, p2, v[3], v[2], iso);
    default: return interp(p0, p3, v[0], v[3], iso);
  }
}

// One thread per square of the camera window.  Display row y (counted from
// the top) has its upper vertices on lattice row ny - 1 - y and its lower
// ones on the row below.  VEC: vertex value |u| (magnitude() of the two velocity
// planes); otherwise the scalar field itself -- the reference runs its
// velocity kernel on phi and reads a second plane past phi's end
// (runCuda.cu:424); here phi has its own scalar path.
template <class T, bool VEC>
__global__ void __launch_bounds__(kViewBlock) k_isoline(const IsoArgs a) {
  const uint32_t x = a.x0 + blockIdx.x * kViewBlock + threadIdx.x;
  const uint32_t y = a.y0 + blockIdx.y;
  if (x >= a.nx - 1 || x >= a.x_max) return;
  const T* f = static_cast<const T*>(a.field);
  const size_t n0 = x + size_t(a.ny - 1 - y) * a.pitch;
  const size_t n3 = n0 - a.pitch;
  float v[4];
  if (VEC) {
    v[0] = magnitude(float(f[n0]), float(f[n0 + a.plane]));
    v[1] = magnitude(float(f[n0 + 1]), float(f[n0 + 1 + a.plane]));
    v[2] = magnitude(float(f[n3 + 1]), float(f[n3 + 1 + a.plane]));
    v[3] = magnitude(float(f[n3]), float(f[n3 + a.plane]));
  } else {
    v[0] = float(f[n0]);
    v[1] = float(f[n0 + 1]);
    v[2] = float(f[n3 + 1]);
    v[3] = float(f[n3]);
  }
  const unsigned idx = unsigned(v[0] > a.iso) | unsigned(v[1] > a.iso) << 1 |
                       unsigned(v[2] > a.iso) << 2 | unsigned(v[3] > a.iso) << 3;
  const unsigned seg = kSegments[idx];
  if ((seg & 0xFF) == 0xFF) return;
  const float xo = float(x - a.x_min), yo = float(y - a.y_min);
  draw_line(edge_point(seg & 0xF, v, a.iso), edge_point((seg >> 4) & 0xF, v, a.iso), a, xo, yo);
  if ((seg >> 8) != 0xFF)
    draw_line(edge_point((seg >> 8) & 0xF, v, a.iso), edge_point(seg >> 12, v, a.iso), a, xo, yo);
}

}  // namespace

extern "C" {

int flb_export_velocity(flb_handle* h, void* d_dst, size_t dst_pitch_bytes) {
  if (!h || !d_dst) return h ? flb::set_error(h, FLB_ERR_ARGUMENT, "null destination") : FLB_ERR_ARGUMENT;
  if (dst_pitch_bytes < size_t(h->nx) * sizeof(float2) || dst_pitch_bytes % sizeof(float2) ||
      reinterpret_cast<uintptr_t>(d_dst) % sizeof(float2))
    return flb::set_error(h, FLB_ERR_ARGUMENT,
                          "destination rows must hold nx float2 texels, 8-byte aligned");
  const dim3 grid((h->nx + kViewBlock - 1) / kViewBlock, h->ny);
  if (h->field_bytes == 8)
    hipLaunchKernelGGL(k_export_velocity<double>, grid, dim3(kViewBlock), 0, h->stream,
                       reinterpret_cast<const double*>(h->u), h->Nd, h->pitch, h->nx,
                       static_cast<float2*>(d_dst), dst_pitch_bytes);
  else
    hipLaunchKernelGGL(k_export_velocity<float>, grid, dim3(kViewBlock), 0, h->stream,
                       static_cast<const float*>(h->u), h->Nd, h->pitch, h->nx,
                       static_cast<float2*>(d_dst), dst_pitch_bytes);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return flb::set_error(h, FLB_ERR_DEVICE, std::string("export_velocity: ") + hipGetErrorString(e));
  return FLB_OK;
}

int flb_isosurface(flb_handle* h, const flb_iso_view* view, void* d_tex, size_t tex_pitch_bytes) {
  if (!h || !view || !d_tex)
    return h ? flb::set_error(h, FLB_ERR_ARGUMENT, "null view or texture") : FLB_ERR_ARGUMENT;
  const flb_iso_view& v = *view;
  if (v.field != FLB_ISO_VELOCITY && v.field != FLB_ISO_PHI)
    return flb::set_error(h, FLB_ERR_ARGUMENT, "field must be FLB_ISO_VELOCITY or FLB_ISO_PHI");
  if (v.field == FLB_ISO_PHI && h->opt.problem != FLB_FREE_SURFACE)
    return flb::set_error(h, FLB_ERR_STATE, "phi isolines need a free-surface handle");
  if (v.width == 0 || v.height == 0 || tex_pitch_bytes < v.width)
    return flb::set_error(h, FLB_ERR_ARGUMENT, "texture rows must hold width bytes");
  if (v.x_min >= v.x_max || v.y_min >= v.y_max)
    return flb::set_error(h, FLB_ERR_ARGUMENT, "empty camera bounds");
  uint8_t* tex = static_cast<uint8_t*>(d_tex);
  if (v.clear) {
    const dim3 g((v.width + kViewBlock - 1) / kViewBlock, v.height);
    hipLaunchKernelGGL(k_fill_u8, g, dim3(kViewBlock), 0, h->stream, tex, tex_pitch_bytes, v.width);
  }
  IsoArgs a{};
  a.field = v.field == FLB_ISO_PHI ? static_cast<const void*>(h->phi) : static_cast<const void*>(h->u);
  a.plane = h->Nd;
  a.pitch = h->pitch;
  a.nx = h->nx;
  a.ny = h->ny;
  a.iso = v.iso_value;
  a.tex = tex;
  a.tex_pitch = tex_pitch_bytes;
  a.width = v.width;
  a.height = v.height;
  a.x_min = v.x_min;
  a.x_max = v.x_max;
  a.y_min = v.y_min;
  a.y_max = v.y_max;
  a.ppn_x = float(v.width) / float(v.x_max - v.x_min);
  a.ppn_y = float(v.height) / float(v.y_max - v.y_min);
  // Squares exist for x < nx - 1, y < ny - 1; only the camera window is launched.
  const uint32_t x1 = std::min(v.x_max, h->nx - 1), y1 = std::min(v.y_max, h->ny - 1);
  if (h->nx >= 2 && h->ny >= 2 && v.x_min < x1 && v.y_min < y1) {
    a.x0 = v.x_min;
    a.y0 = v.y_min;
    const dim3 g((x1 - v.x_min + kViewBlock - 1) / kViewBlock, y1 - v.y_min);
    const bool f64 = h->field_bytes == 8;
    if (v.field == FLB_ISO_VELOCITY) {
      if (f64) hipLaunchKernelGGL((k_isoline<double, true>), g, dim3(kViewBlock), 0, h->stream, a);
      else hipLaunchKernelGGL((k_isoline<float, true>), g, dim3(kViewBlock), 0, h->stream, a);
    } else {
      if (f64) hipLaunchKernelGGL((k_isoline<double, false>), g, dim3(kViewBlock), 0, h->stream, a);
      else hipLaunchKernelGGL((k_isoline<float, false>), g, dim3(kViewBlock), 0, h->stream, a);
    }
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return flb::set_error(h, FLB_ERR_DEVICE, std::string("isosurface: ") + hipGetErrorString(e));
  return FLB_OK;
}

}  // extern "C"
