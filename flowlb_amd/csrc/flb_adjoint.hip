// flb_adjoint.hip -- hand-written discrete adjoint of the single-phase step
// (SRT, Guo forcing, Brinkman drag, velocity inlet, density outlet, implicit
// full-way bounce-back) for gradients of the time-averaged discharge through a
// column with respect to the Brinkman solid fraction and the inlet velocity.
// No reference counterpart: FlowLB has no adjoint.
//
// Reverse of one step, per cell and in place: the forward step of a cell
// loads nine slots and stores nine results into the same slot set (esoteric
// pull), so the adjoint kernel reads the adjoints of the stored results from
// the store slots and writes the adjoints of the loaded populations to the
// load slots.  The collision depends on the loaded populations only through
// rho and m = sum c_i f_i, so the forward recompute records three numbers per
// cell and step (the trace) instead of nine populations.
//
// Memory: the trace of every step (3 N floats per step) is kept from the
// forward sweep for as many of the last checkpoint segments as fit in the
// trace budget (by default half the free device memory: on a 288 GB device
// the whole 2000-step, 1024 x 256 rollout, 6.3 GB); the earlier segments get a
// checkpoint of the populations and the inlet/outlet publication buffers every
// `ckpt` steps and are recomputed, with their trace, during the reverse sweep.
// The adjoint population array has 9 planes; dJ/dalpha accumulates in FP64.
#include <algorithm>
#include <vector>

#include "flb_cell.h"
#include "flb_internal.h"

namespace flb {

struct AdjParams {
  float* abar;          // adjoint populations, 9 planes of Np floats
  float* rbar[2];       // adjoint of published densities, by parity
  float* ubar[2];       // adjoint of published velocities, by parity
  double* galpha;       // dJ/dalpha accumulator (N), FP64 over all steps, or nullptr
  const float* trace;   // (rho, m_x, m_y) or inlet/outlet reads of this step
  const float* rho_init;
  const float* u_init;
  double* duin;         // dJ/du_inlet accumulator
  uint32_t column;
  int in_window;
  float qbar;           // 1 / number of averaged steps
};

__device__ __forceinline__ void feq_grad(float rho, float ux, float uy, float* fe, float* dfx,
                                         float* dfy) {
  const float usq = ux * ux + uy * uy;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const float cu = float(cx(i)) * ux + float(cy(i)) * uy;
    fe[i] = wgt(i) * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * usq);
    dfx[i] = wgt(i) * rho * (3.0f * float(cx(i)) + 9.0f * cu * float(cx(i)) - 3.0f * ux);
    dfy[i] = wgt(i) * rho * (3.0f * float(cy(i)) + 9.0f * cu * float(cy(i)) - 3.0f * uy);
  }
}

__device__ __forceinline__ uint32_t cell_neighbour(const Params& p, uint32_t x, uint32_t y, int i) {
  const uint32_t nx = p.nx, ny = p.ny;
  uint32_t xx = x, yy = y;
  if (cx(i) == 1) xx = x + 1 == nx ? 0 : x + 1;
  if (cx(i) == -1) xx = x == 0 ? nx - 1 : x - 1;
  if (cy(i) == 1) yy = y + 1 == ny ? 0 : y + 1;
  if (cy(i) == -1) yy = y == 0 ? ny - 1 : y - 1;
  return xx + yy * p.pitch;
}

// Slot (plane, cell) where the forward step of parity ODD stores population i.
template <int ODD>
__device__ __forceinline__ uint64_t store_index(const Params& p, uint32_t n, const uint32_t* nb,
                                                int i) {
  if (i == 0) return n;
  if (i & 1) return (ODD ? i + 1 : i) * p.Np + nb[i];
  return (ODD ? i - 1 : i) * p.Np + n;
}
// Slot where it loads population i.
template <int ODD>
__device__ __forceinline__ uint64_t load_index(const Params& p, uint32_t n, const uint32_t* nb,
                                               int i) {
  if (i == 0) return n;
  if (i & 1) return (ODD ? i : i + 1) * p.Np + n;
  return (ODD ? i : i - 1) * p.Np + nb[i - 1];
}

__device__ __forceinline__ void block_add(double v, double* target) {
  __shared__ double lds[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int k = 0; k < int(blockDim.x >> 6); k++) s += lds[k];
    if (s != 0.0) atomicAdd(target, s);
  }
}

template <int ODD>
__global__ void __launch_bounds__(256) k_adjoint_step(const Params p, const AdjParams a) {
  uint32_t tx, y;
  block_tile(blockIdx.x, p.tiles_x, p.ny, tx, y);
  const uint32_t x = tx * blockDim.x + threadIdx.x;
  double duin = 0.0;
  if (x < p.nx) {
    const uint32_t n = x + y * p.pitch;
    const uint64_t N = p.N;
    const uint8_t fl = p.flags[n];
    if (!(fl & (kGas | kWall))) {
      uint32_t nb[9];
      nb[0] = n;
#pragma unroll
      for (int i = 1; i < 9; i++) nb[i] = cell_neighbour(p, x, y, i);
      float v[9];
#pragma unroll
      for (int i = 0; i < 9; i++) v[i] = a.abar[store_index<ODD>(p, n, nb, i)];
      const Linear lin = linear_neighbours(p, x, y, n);
      const uint8_t fl_prev = p.flags[lin.prev];  // unconditional: one round trip
      const uint8_t fl_next = p.flags[lin.next];
      const bool pub_r = lin.has_prev && (fl_prev & kInlet);
      const bool pub_v = lin.has_next && (fl_next & kOutlet) && !(fl & kInlet);
      float rb = 0.0f, ubx = 0.0f, uby = 0.0f;
      if (pub_r) {
        rb = a.rbar[ODD][n];
        a.rbar[ODD][n] = 0.0f;
      }
      if (pub_v) {
        ubx = a.ubar[ODD][n];
        uby = a.ubar[ODD][N + n];
        a.ubar[ODD][n] = 0.0f;
        a.ubar[ODD][N + n] = 0.0f;
      }
      const float Fx = p.use_force ? p.fx : 0.0f, Fy = p.use_force ? p.fy : 0.0f;
      const float t0 = a.trace[n], t1 = a.trace[N + n], t2 = a.trace[2 * N + n];
      const uint8_t io = fl & (kInlet | kOutlet);
      float gb[9];
      if (io) {
        // populations = feq(rho, u_read + F/(2 rho)): no dependence on the loaded ones
        const float rho = t0;
        const float ux = t1 + 0.5f / rho * Fx, uy = t2 + 0.5f / rho * Fy;
        float fe[9], dfx[9], dfy[9];
        feq_grad(rho, ux, uy, fe, dfx, dfy);
#pragma unroll
        for (int i = 0; i < 9; i++) {
          rb += v[i] * fe[i] / rho;
          ubx += v[i] * dfx[i];
          uby += v[i] * dfy[i];
          gb[i] = 0.0f;
        }
        rb += -(ubx * Fx + uby * Fy) * 0.5f / (rho * rho);
        if (io & kInlet) {
          a.rbar[ODD ^ 1][lin.next] += rb;
          duin = ubx;
        } else {
          a.ubar[ODD ^ 1][lin.prev] += ubx;
          a.ubar[ODD ^ 1][N + lin.prev] += uby;
        }
      } else {
        const float rho = t0, mx = t1, my = t2;
        const float al = p.alpha ? p.alpha[n] : 0.0f;
        const float sa = 1.0f / (1.0f + 0.5f * al);
        const float ux = (mx + 0.5f * Fx) * sa / rho, uy = (my + 0.5f * Fy) * sa / rho;
        const float Gx = Fx - al * rho * ux, Gy = Fy - al * rho * uy;
        const bool forced = p.use_force || p.alpha;
        const float om = p.inv_tau, tl = 1.0f - 0.5f * p.inv_tau, k3 = 0.33333334f;
        float fe[9], dfx[9], dfy[9];
        feq_grad(rho, ux, uy, fe, dfx, dfy);
        float Gbx = 0.0f, Gby = 0.0f;
#pragma unroll
        for (int i = 0; i < 9; i++) {
          gb[i] = (1.0f - om) * v[i];
          const float feb = om * v[i];
          rb += feb * fe[i] / rho;
          ubx += feb * dfx[i];
          uby += feb * dfy[i];
          if (forced) {
            const float Fb = tl * v[i] * 9.0f * wgt(i);
            const float cG = float(cx(i)) * Gx + float(cy(i)) * Gy;
            const float cu = float(cx(i)) * ux + float(cy(i)) * uy;
            ubx += Fb * (cG * float(cx(i)) - k3 * Gx);
            uby += Fb * (cG * float(cy(i)) - k3 * Gy);
            Gbx += Fb * (float(cx(i)) * (cu + k3) - k3 * ux);
            Gby += Fb * (float(cy(i)) * (cu + k3) - k3 * uy);
          }
        }
        ubx += -al * rho * Gbx;
        uby += -al * rho * Gby;
        rb += -al * (ux * Gbx + uy * Gby);
        float ab = -rho * (ux * Gbx + uy * Gby);
        if (a.in_window && x == a.column) {  // J += qbar * rho * u_x
          rb += a.qbar * ux;
          ubx += a.qbar * rho;
        }
        const float mbx = ubx * sa / rho, mby = uby * sa / rho;
        const float ubu = ubx * ux + uby * uy;
        rb += -ubu / rho;
        ab += -0.5f * sa * ubu;
#pragma unroll
        for (int i = 0; i < 9; i++) gb[i] += rb + float(cx(i)) * mbx + float(cy(i)) * mby;
        if (a.galpha) a.galpha[n] += double(ab);
      }
#pragma unroll
      for (int i = 0; i < 9; i++) a.abar[load_index<ODD>(p, n, nb, i)] = gb[i];
    }
  }
  block_add(duin, a.duin);
}

// The initial populations of inlet cells, feq(rho0, u0) stored with the odd
// pattern, depend on the inlet velocity.
__global__ void __launch_bounds__(256) k_adjoint_inlet_init(const Params p, const AdjParams a) {
  uint32_t tx, y;
  block_tile(blockIdx.x, p.tiles_x, p.ny, tx, y);
  const uint32_t x = tx * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (x < p.nx) {
    const uint32_t n = x + y * p.pitch;
    const uint8_t fl = p.flags[n];
    if ((fl & kInlet) && !(fl & (kGas | kWall))) {
      uint32_t nb[9];
      nb[0] = n;
#pragma unroll
      for (int i = 1; i < 9; i++) nb[i] = cell_neighbour(p, x, y, i);
      float fe[9], dfx[9], dfy[9];
      feq_grad(a.rho_init[n], a.u_init[n], a.u_init[p.N + n], fe, dfx, dfy);
#pragma unroll
      for (int i = 0; i < 9; i++) acc += double(a.abar[store_index<1>(p, n, nb, i)]) * dfx[i];
    }
  }
  block_add(acc, a.duin);
}

// Discharge through a column: sum of rho * u_x over its non-wall cells, from the
// readback fields of the last step, accumulated as scale * Q into acc[0].
__global__ void __launch_bounds__(256) k_discharge(const Params p, uint32_t column, double scale,
                                                   double* acc) {
  __shared__ double lds[4];
  double s = 0.0;
  for (uint32_t y = threadIdx.x; y < p.ny; y += blockDim.x) {
    const uint32_t n = column + y * p.pitch;
    if (!(p.flags[n] & (kWall | kGas))) s += double(p.rho[n]) * double(p.u[n]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) acc[0] += scale * (lds[0] + lds[1] + lds[2] + lds[3]);
}

}  // namespace flb

namespace {

#define ADJ_HIP(h, call)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return flb::set_error((h), FLB_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  std::vector<void*> ptrs;
  ~DevBuf() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  T* get(size_t count) {
    void* v = nullptr;
    if (hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    ptrs.push_back(v);
    return static_cast<T*>(v);
  }
};

// Copies the step state (populations and the inlet/outlet publication buffers,
// plus the readback fields when with_fields) between the handle and buf.
// Layout of buf (device planes of Nd cells): f (9 Np) | pub_rho[0] |
// pub_rho[1] (N each) | pub_u[0] | pub_u[1] (2N each) | rho (N) | u (2N).
size_t state_floats(const flb_handle* h, bool with_fields) {
  return 9 * h->Np + 6 * h->Nd + (with_fields ? 3 * h->Nd : 0);
}
int copy_state(flb_handle* h, float* buf, bool to_handle, bool with_fields) {
  const size_t N = h->Nd, Np = h->Np;
  struct Part {
    void* dev;
    size_t off, bytes;
  } parts[7] = {{h->f, 0, 9 * Np * sizeof(float)},
                {h->pub_rho[0], 9 * Np, N * sizeof(float)},
                {h->pub_rho[1], 9 * Np + N, N * sizeof(float)},
                {h->pub_u[0], 9 * Np + 2 * N, 2 * N * sizeof(float)},
                {h->pub_u[1], 9 * Np + 4 * N, 2 * N * sizeof(float)},
                {h->rho, 9 * Np + 6 * N, N * sizeof(float)},
                {h->u, 9 * Np + 7 * N, 2 * N * sizeof(float)}};
  const int count = with_fields ? 7 : 5;
  for (int k = 0; k < count; ++k) {
    void* a = to_handle ? parts[k].dev : static_cast<void*>(buf + parts[k].off);
    const void* b = to_handle ? static_cast<const void*>(buf + parts[k].off) : parts[k].dev;
    ADJ_HIP(h, hipMemcpyAsync(a, b, parts[k].bytes, hipMemcpyDeviceToDevice, h->stream));
  }
  return FLB_OK;
}

int check_adjoint_state(flb_handle* h, const flb_adjoint_spec* sp) {
  if (!h || !sp) return FLB_ERR_ARGUMENT;
  if (h->opt.problem != FLB_MONOFLUID || h->opt.collision != FLB_SRT || h->opt.smagorinsky ||
      h->opt.storage != FLB_STORAGE_FP32)
    return flb::set_error(h, FLB_ERR_UNSUPPORTED,
                          "adjoint needs a monofluid SRT handle with FP32 storage and no subgrid model");
  if (h->t != 0) return flb::set_error(h, FLB_ERR_STATE, "adjoint runs from the initial state (t = 0)");
  if (sp->nsteps == 0 || sp->t1 < sp->t0 || sp->t1 >= sp->nsteps || sp->column >= h->nx ||
      sp->ckpt == 0)
    return flb::set_error(h, FLB_ERR_ARGUMENT, "invalid adjoint window");
  return FLB_OK;
}

// Forward steps [t, t+count) with the objective accumulated into d_j; with a
// trace buffer, step t + k records its trace at trace + k * 3 N.
int run_forward(flb_handle* h, uint64_t count, const flb_adjoint_spec* sp, double qscale,
                double* d_j, float* trace = nullptr) {
  const int bx = flb::handle_block_x(h);
  for (uint64_t k = 0; k < count; ++k) {
    const uint64_t t = h->t;
    const bool win = d_j && t >= sp->t0 && t <= sp->t1;
    flb::Params p = flb::make_params(h, bx);
    p.trace = trace ? trace + k * 3 * h->Nd : nullptr;
    hipLaunchKernelGGL(trace ? flb::get_stream_collide_trace(int(t & 1), win)
                             : flb::get_stream_collide(FLB_STORAGE_FP32, int(t & 1), false, win),
                       dim3(p.nblocks), dim3(bx), 0, h->stream, p);
    ADJ_HIP(h, hipGetLastError());
    if (win) {
      hipLaunchKernelGGL(flb::k_discharge, dim3(1), dim3(256), 0, h->stream, p, sp->column, qscale,
                         d_j);
      ADJ_HIP(h, hipGetLastError());
    }
    h->t += 1;
  }
  return FLB_OK;
}

}  // namespace

extern "C" {

int flb_set_brinkman(flb_handle* h, const float* alpha) {
  if (!h) return FLB_ERR_ARGUMENT;
  if (!alpha) {
    h->alpha = nullptr;  // buffer stays owned by the handle
    return FLB_OK;
  }
  if (h->opt.problem != FLB_MONOFLUID)
    return flb::set_error(h, FLB_ERR_UNSUPPORTED, "Brinkman drag needs a monofluid handle");
  if (!h->alpha) {
    void* v = nullptr;
    ADJ_HIP(h, hipMalloc(&v, h->Nd * sizeof(float)));
    h->owned.push_back(v);
    h->alpha = static_cast<float*>(v);
    ADJ_HIP(h, hipMemsetAsync(v, 0, h->Nd * sizeof(float), h->stream));
  }
  return flb::copy_cells_in(h, h->alpha, alpha, sizeof(float), 1);
}

int flb_objective(flb_handle* h, const flb_adjoint_spec* sp, double* J) {
  int rc = check_adjoint_state(h, sp);
  if (rc) return rc;
  DevBuf mem;
  float* saved = mem.get<float>(state_floats(h, true));
  double* d_j = mem.get<double>(1);
  if (!saved || !d_j) return flb::set_error(h, FLB_ERR_MEMORY, "objective buffers");
  if ((rc = copy_state(h, saved, false, true))) return rc;
  ADJ_HIP(h, hipMemsetAsync(d_j, 0, sizeof(double), h->stream));
  rc = run_forward(h, sp->nsteps, sp, 1.0 / double(sp->t1 - sp->t0 + 1), d_j);
  const int rc2 = copy_state(h, saved, true, true);
  h->t = 0;
  if (rc) return rc;
  if (rc2) return rc2;
  double j = 0.0;
  ADJ_HIP(h, hipMemcpyAsync(&j, d_j, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  ADJ_HIP(h, hipStreamSynchronize(h->stream));
  if (J) *J = j;
  return FLB_OK;
}

int flb_adjoint(flb_handle* h, const flb_adjoint_spec* sp, double* J, float* dj_dalpha,
                double* dj_duin, float* phase_ms) {
  int rc = check_adjoint_state(h, sp);
  if (rc) return rc;
  const size_t N = h->Nd, Np = h->Np;  // device planes
  const uint64_t K = sp->ckpt, T = sp->nsteps;
  const uint64_t nseg = (T + K - 1) / K;
  const size_t ck_floats = state_floats(h, false);
  const size_t seg_bytes = size_t(std::min<uint64_t>(K, T)) * 3 * N * sizeof(float);
  // Segments whose trace stays resident from the forward sweep: the last
  // `resident` ones, as many as the budget holds.
  size_t budget = sp->trace_budget_bytes;
  if (budget == 0) {
    size_t free_b = 0, total_b = 0;
    ADJ_HIP(h, hipMemGetInfo(&free_b, &total_b));
    budget = free_b / 2;
  }
  const uint64_t resident = std::min<uint64_t>(nseg, budget / seg_bytes);
  const uint64_t first_res = nseg - resident;  // segments [0, first_res) are recomputed
  DevBuf mem;
  float* ckpt = first_res ? mem.get<float>(first_res * ck_floats) : nullptr;
  float* fields0 = mem.get<float>(state_floats(h, true));
  float* trace_res = resident ? mem.get<float>(resident * (seg_bytes / sizeof(float))) : nullptr;
  float* trace_tmp = first_res ? mem.get<float>(seg_bytes / sizeof(float)) : nullptr;
  float* abar = mem.get<float>(9 * Np);
  float* rbar = mem.get<float>(2 * N);
  float* ubar = mem.get<float>(4 * N);
  double* galpha = h->alpha ? mem.get<double>(N) : nullptr;
  double* d_acc = mem.get<double>(2);  // J, dJ/du_in
  if ((first_res && (!ckpt || !trace_tmp)) || (resident && !trace_res) || !fields0 || !abar ||
      !rbar || !ubar || (h->alpha && !galpha) || !d_acc)
    return flb::set_error(h, FLB_ERR_MEMORY, "adjoint buffers (checkpoints, trace, adjoint state)");
  auto seg_trace = [&](uint64_t s) {
    return s >= first_res ? trace_res + (s - first_res) * (seg_bytes / sizeof(float)) : trace_tmp;
  };

  if ((rc = copy_state(h, fields0, false, true))) return rc;
  ADJ_HIP(h, hipMemsetAsync(d_acc, 0, 2 * sizeof(double), h->stream));
  hipEvent_t ev[4];
  for (auto& e : ev) ADJ_HIP(h, hipEventCreate(&e));
  ADJ_HIP(h, hipEventRecord(ev[0], h->stream));

  // Forward sweep: the objective, checkpoints of the segments to recompute,
  // the trace of the resident ones.
  const double qs = 1.0 / double(sp->t1 - sp->t0 + 1);
  for (uint64_t s = 0; s < nseg && !rc; ++s) {
    const uint64_t count = std::min(K, T - s * K);
    if (s < first_res) {
      if ((rc = copy_state(h, ckpt + s * ck_floats, false, false))) break;
      rc = run_forward(h, count, sp, qs, d_acc);
    } else {
      rc = run_forward(h, count, sp, qs, d_acc, seg_trace(s));
    }
  }
  ADJ_HIP(h, hipEventRecord(ev[1], h->stream));

  // Reverse sweep, one checkpoint segment at a time.
  ADJ_HIP(h, hipMemsetAsync(abar, 0, 9 * Np * sizeof(float), h->stream));
  ADJ_HIP(h, hipMemsetAsync(rbar, 0, 2 * N * sizeof(float), h->stream));
  ADJ_HIP(h, hipMemsetAsync(ubar, 0, 4 * N * sizeof(float), h->stream));
  if (galpha) ADJ_HIP(h, hipMemsetAsync(galpha, 0, N * sizeof(double), h->stream));
  const int abx = 256;
  float recompute_ms = 0.0f;
  for (uint64_t s = nseg; s-- > 0 && !rc;) {
    const uint64_t ts = s * K, te = std::min(ts + K, T);
    float* trace = seg_trace(s);
    if (s < first_res) {
      if ((rc = copy_state(h, ckpt + s * ck_floats, true, false))) break;
      h->t = ts;
      ADJ_HIP(h, hipEventRecord(ev[2], h->stream));
      if ((rc = run_forward(h, te - ts, sp, qs, nullptr, trace))) break;
      ADJ_HIP(h, hipEventRecord(ev[3], h->stream));
    }
    for (uint64_t t = te; t-- > ts;) {
      flb::Params p = flb::make_params(h, abx);
      flb::AdjParams a{};
      a.abar = abar;
      a.rbar[0] = rbar;
      a.rbar[1] = rbar + N;
      a.ubar[0] = ubar;
      a.ubar[1] = ubar + 2 * N;
      a.galpha = galpha;
      a.trace = trace + (t - ts) * 3 * N;
      a.duin = d_acc + 1;
      a.column = sp->column;
      a.in_window = t >= sp->t0 && t <= sp->t1;
      a.qbar = float(qs);
      if (t & 1)
        hipLaunchKernelGGL(flb::k_adjoint_step<1>, dim3(p.nblocks), dim3(abx), 0, h->stream, p, a);
      else
        hipLaunchKernelGGL(flb::k_adjoint_step<0>, dim3(p.nblocks), dim3(abx), 0, h->stream, p, a);
      ADJ_HIP(h, hipGetLastError());
    }
    if (s < first_res) {
      ADJ_HIP(h, hipEventSynchronize(ev[3]));
      float ms = 0.0f;
      ADJ_HIP(h, hipEventElapsedTime(&ms, ev[2], ev[3]));
      recompute_ms += ms;
    }
  }
  if (!rc) {
    flb::Params p = flb::make_params(h, abx);
    flb::AdjParams a{};
    a.abar = abar;
    a.rho_init = h->rho_init;
    a.u_init = h->u_init;
    a.duin = d_acc + 1;
    hipLaunchKernelGGL(flb::k_adjoint_inlet_init, dim3(p.nblocks), dim3(abx), 0, h->stream, p, a);
    ADJ_HIP(h, hipGetLastError());
  }
  // Leave the handle in its initial state.
  hipEvent_t end;
  ADJ_HIP(h, hipEventCreate(&end));
  const int rc_restore = copy_state(h, fields0, true, true);
  if (!rc) rc = rc_restore;
  ADJ_HIP(h, hipEventRecord(end, h->stream));
  h->t = 0;
  double acc[2] = {0.0, 0.0};
  ADJ_HIP(h, hipMemcpyAsync(acc, d_acc, 2 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  ADJ_HIP(h, hipStreamSynchronize(h->stream));
  if (dj_dalpha && galpha && !rc) {
    std::vector<double> g(h->N);
    rc = flb::copy_cells_out(h, g.data(), galpha, sizeof(double), 1);
    for (size_t n = 0; n < h->N && !rc; ++n) dj_dalpha[n] = float(g[n]);
  }
  if (phase_ms) {
    float fwd = 0.0f, tot = 0.0f;
    ADJ_HIP(h, hipEventElapsedTime(&fwd, ev[0], ev[1]));
    ADJ_HIP(h, hipEventElapsedTime(&tot, ev[0], end));
    phase_ms[0] = fwd;                   // forward sweep (checkpoints or trace)
    phase_ms[1] = recompute_ms;          // forward recompute of non-resident segments
    phase_ms[2] = tot - fwd - recompute_ms;  // reverse sweep and restore
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(end);
  if (rc) return rc;
  if (dj_dalpha && !galpha)
    for (size_t n = 0; n < h->N; ++n) dj_dalpha[n] = 0.0f;
  if (J) *J = acc[0];
  if (dj_duin) *dj_duin = acc[1];
  return FLB_OK;
}

}  // extern "C"
