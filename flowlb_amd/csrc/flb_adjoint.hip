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
// Memory: the trace of every step (3 N values per step) is kept from the
// forward sweep for as many of the last checkpoint segments as fit in the
// trace budget (by default half the free device memory: on a 288 GB device
// the whole 2000-step, 1024 x 256 rollout, 6.3 GB); the earlier segments get a
// checkpoint of the populations and the inlet/outlet publication buffers every
// `ckpt` steps and are recomputed, with their trace, during the reverse sweep.
// The adjoint state (9 population planes, the adjoints of the published
// inlet/outlet values) and its arithmetic are FP64: over 2000 reverse steps an
// FP32 adjoint lost about 2e-4 (relative) on cells with small gradients.  The
// trace is what the forward computed: FP32 for an FP32 handle, FP64 for an
// FP64 handle (FP64 forward and trace: the gradient of the FP64 discrete
// model).  The forward's float constants enter as their exact double values.
#include <algorithm>
#include <vector>

#include "flb_cell.h"
#include "flb_internal.h"

namespace flb {

struct AdjParams {
  double* abar;         // adjoint populations, 9 planes of Np
  double* rbar[2];      // adjoint of published densities, by parity
  double* ubar[2];      // adjoint of published velocities, by parity
  double* galpha;       // dJ/dalpha accumulator (N), FP64 over all steps, or nullptr
  const void* trace;    // (rho, m_x, m_y) or inlet/outlet reads of this step, T
  const float* rho_init;
  const float* u_init;
  double* duin;         // dJ/du_inlet accumulator
  uint32_t column;
  int in_window;
  double qbar;          // 1 / number of averaged steps
};

// Equilibrium and its velocity derivatives with the weights of the forward
// arithmetic T (float weights for FP32, exact ones for FP64).
template <class T>
__device__ __forceinline__ void feq_grad(double rho, double ux, double uy, double* fe,
                                         double* dfx, double* dfy) {
  const double usq = ux * ux + uy * uy;
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const double w = double(wgt_r<T>(i)), ex = double(cx(i)), ey = double(cy(i));
    const double cu = ex * ux + ey * uy;
    fe[i] = w * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
    dfx[i] = w * rho * (3.0 * ex + 9.0 * cu * ex - 3.0 * ux);
    dfy[i] = w * rho * (3.0 * ey + 9.0 * cu * ey - 3.0 * uy);
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

template <class T, int ODD>
__global__ void __launch_bounds__(256) k_adjoint_step(const Params p, const AdjParams a) {
  uint32_t tx, y;
  block_tile(blockIdx.x, p.tiles_x, p.ny, p.band_magic, p.last_magic, tx, y);
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
      double v[9];
#pragma unroll
      for (int i = 0; i < 9; i++) v[i] = a.abar[store_index<ODD>(p, n, nb, i)];
      const Linear lin = linear_neighbours(p, x, y, n);
      const uint8_t fl_prev = p.flags[lin.prev];  // unconditional: one round trip
      const uint8_t fl_next = p.flags[lin.next];
      const bool pub_r = lin.has_prev && (fl_prev & kInlet);
      const bool pub_v = lin.has_next && (fl_next & kOutlet) && !(fl & kInlet);
      double rb = 0.0, ubx = 0.0, uby = 0.0;
      if (pub_r) {
        rb = a.rbar[ODD][n];
        a.rbar[ODD][n] = 0.0;
      }
      if (pub_v) {
        ubx = a.ubar[ODD][n];
        uby = a.ubar[ODD][N + n];
        a.ubar[ODD][n] = 0.0;
        a.ubar[ODD][N + n] = 0.0;
      }
      const double Fx = p.use_force ? p.fx : 0.0f, Fy = p.use_force ? p.fy : 0.0f;
      const T* const tr = static_cast<const T*>(a.trace);
      const double t0 = tr[n], t1 = tr[N + n], t2 = tr[2 * N + n];
      const uint8_t io = fl & (kInlet | kOutlet);
      double gb[9];
      if (io) {
        // populations = feq(rho, u_read + F/(2 rho)): no dependence on the loaded ones
        const double rho = t0;
        const double ux = t1 + 0.5 / rho * Fx, uy = t2 + 0.5 / rho * Fy;
        double fe[9], dfx[9], dfy[9];
        feq_grad<T>(rho, ux, uy, fe, dfx, dfy);
#pragma unroll
        for (int i = 0; i < 9; i++) {
          rb += v[i] * fe[i] / rho;
          ubx += v[i] * dfx[i];
          uby += v[i] * dfy[i];
          gb[i] = 0.0;
        }
        rb += -(ubx * Fx + uby * Fy) * 0.5 / (rho * rho);
        if (io & kInlet) {
          a.rbar[ODD ^ 1][lin.next] += rb;
          duin = ubx;
        } else {
          a.ubar[ODD ^ 1][lin.prev] += ubx;
          a.ubar[ODD ^ 1][N + lin.prev] += uby;
        }
      } else {
        const double rho = t0, mx = t1, my = t2;
        const double al = p.alpha ? p.alpha[n] : 0.0f;
        const double sa = 1.0 / (1.0 + 0.5 * al);
        const double ux = (mx + 0.5 * Fx) * sa / rho, uy = (my + 0.5 * Fy) * sa / rho;
        const double Gx = Fx - al * rho * ux, Gy = Fy - al * rho * uy;
        const bool forced = p.use_force || p.alpha;
        // The forward's SRT weights are float constants (also in FP64).
        const double om = p.inv_tau, omc = 1.0f - p.inv_tau, tl = 1.0f - 0.5f * p.inv_tau;
        const double k3 = double(0.33333334f);
        double fe[9], dfx[9], dfy[9];
        feq_grad<T>(rho, ux, uy, fe, dfx, dfy);
        double Gbx = 0.0, Gby = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
          const double ex = double(cx(i)), ey = double(cy(i));
          gb[i] = omc * v[i];
          const double feb = om * v[i];
          rb += feb * fe[i] / rho;
          ubx += feb * dfx[i];
          uby += feb * dfy[i];
          if (forced) {
            const double Fb = tl * v[i] * 9.0 * double(wgt_r<T>(i));
            const double cG = ex * Gx + ey * Gy;
            const double cu = ex * ux + ey * uy;
            ubx += Fb * (cG * ex - k3 * Gx);
            uby += Fb * (cG * ey - k3 * Gy);
            Gbx += Fb * (ex * (cu + k3) - k3 * ux);
            Gby += Fb * (ey * (cu + k3) - k3 * uy);
          }
        }
        ubx += -al * rho * Gbx;
        uby += -al * rho * Gby;
        rb += -al * (ux * Gbx + uy * Gby);
        double ab = -rho * (ux * Gbx + uy * Gby);
        if (a.in_window && x == a.column) {  // J += qbar * rho * u_x
          rb += a.qbar * ux;
          ubx += a.qbar * rho;
        }
        const double mbx = ubx * sa / rho, mby = uby * sa / rho;
        const double ubu = ubx * ux + uby * uy;
        rb += -ubu / rho;
        ab += -0.5 * sa * ubu;
#pragma unroll
        for (int i = 0; i < 9; i++) gb[i] += rb + double(cx(i)) * mbx + double(cy(i)) * mby;
        if (a.galpha) a.galpha[n] += ab;
      }
#pragma unroll
      for (int i = 0; i < 9; i++) a.abar[load_index<ODD>(p, n, nb, i)] = gb[i];
    }
  }
  block_add(duin, a.duin);
}

// The initial populations of inlet cells, feq(rho0, u0) stored with the odd
// pattern, depend on the inlet velocity.  The set-up evaluates them in double
// with exact weights for every storage format.
__global__ void __launch_bounds__(256) k_adjoint_inlet_init(const Params p, const AdjParams a) {
  uint32_t tx, y;
  block_tile(blockIdx.x, p.tiles_x, p.ny, p.band_magic, p.last_magic, tx, y);
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
      double fe[9], dfx[9], dfy[9];
      feq_grad<double>(a.rho_init[n], a.u_init[n], a.u_init[p.N + n], fe, dfx, dfy);
#pragma unroll
      for (int i = 0; i < 9; i++) acc += a.abar[store_index<1>(p, n, nb, i)] * dfx[i];
    }
  }
  block_add(acc, a.duin);
}

// Discharge through a column: sum of rho * u_x over its non-wall cells, from the
// readback fields of the last step, accumulated as scale * Q into acc[0].
template <class R>
__global__ void __launch_bounds__(256) k_discharge(const Params p, uint32_t column, double scale,
                                                   double* acc) {
  __shared__ double lds[4];
  double s = 0.0;
  for (uint32_t y = threadIdx.x; y < p.ny; y += blockDim.x) {
    const uint32_t n = column + y * p.pitch;
    if (!(p.flags[n] & (kWall | kGas))) s += double(fld<R>(p.rho)[n]) * double(fld<R>(p.u)[n]);
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
// Layout of buf (device planes of Nd cells, bytes): f (9 Np populations) |
// pub_rho[0] | pub_rho[1] (N each) | pub_u[0] | pub_u[1] (2N each) | rho (N) |
// u (2N), the fields in the handle's field type.
size_t state_bytes(const flb_handle* h, bool with_fields) {
  return 9 * h->Np * h->elem_bytes + (6 * h->Nd + (with_fields ? 3 * h->Nd : 0)) * h->field_bytes;
}
int copy_state(flb_handle* h, uint8_t* buf, bool to_handle, bool with_fields) {
  const size_t N = h->Nd * h->field_bytes, o = 9 * h->Np * h->elem_bytes;
  struct Part {
    void* dev;
    size_t off, bytes;
  } parts[7] = {{h->f, 0, o},
                {h->pub_rho[0], o, N},
                {h->pub_rho[1], o + N, N},
                {h->pub_u[0], o + 2 * N, 2 * N},
                {h->pub_u[1], o + 4 * N, 2 * N},
                {h->rho, o + 6 * N, N},
                {h->u, o + 7 * N, 2 * N}};
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
      (h->opt.storage != FLB_STORAGE_FP32 && h->opt.storage != FLB_STORAGE_FP64))
    return flb::set_error(h, FLB_ERR_UNSUPPORTED,
                          "adjoint needs a monofluid SRT handle with FP32 or FP64 storage and no "
                          "subgrid model");
  if (h->t != 0) return flb::set_error(h, FLB_ERR_STATE, "adjoint runs from the initial state (t = 0)");
  if (sp->nsteps == 0 || sp->t1 < sp->t0 || sp->t1 >= sp->nsteps || sp->column >= h->nx ||
      sp->ckpt == 0)
    return flb::set_error(h, FLB_ERR_ARGUMENT, "invalid adjoint window");
  return FLB_OK;
}

// Forward steps [t, t+count) with the objective accumulated into d_j; with a
// trace buffer, step t + k records its trace at trace + k * 3 N values.
int run_forward(flb_handle* h, uint64_t count, const flb_adjoint_spec* sp, double qscale,
                double* d_j, uint8_t* trace = nullptr) {
  const bool fp64 = h->opt.storage == FLB_STORAGE_FP64;
  const int bx = flb::handle_block_x(h);
  for (uint64_t k = 0; k < count; ++k) {
    const uint64_t t = h->t;
    const bool win = d_j && t >= sp->t0 && t <= sp->t1;
    flb::Params p = flb::make_params(h, bx);
    p.trace = trace ? reinterpret_cast<float*>(trace + k * 3 * h->Nd * h->field_bytes) : nullptr;
    hipLaunchKernelGGL(trace ? flb::get_stream_collide_trace(h->opt.storage, int(t & 1), win)
                             : flb::get_stream_collide(h->opt.storage, int(t & 1), false, win),
                       dim3(p.nblocks), dim3(bx), 0, h->stream, p);
    ADJ_HIP(h, hipGetLastError());
    if (win) {
      hipLaunchKernelGGL(fp64 ? flb::k_discharge<double> : flb::k_discharge<float>, dim3(1),
                         dim3(256), 0, h->stream, p, sp->column, qscale, d_j);
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
  uint8_t* saved = mem.get<uint8_t>(state_bytes(h, true));
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
  const size_t ck_bytes = state_bytes(h, false);
  const size_t seg_bytes = size_t(std::min<uint64_t>(K, T)) * 3 * N * h->field_bytes;
  const bool fp64 = h->opt.storage == FLB_STORAGE_FP64;
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
  uint8_t* ckpt = first_res ? mem.get<uint8_t>(first_res * ck_bytes) : nullptr;
  uint8_t* fields0 = mem.get<uint8_t>(state_bytes(h, true));
  uint8_t* trace_res = resident ? mem.get<uint8_t>(resident * seg_bytes) : nullptr;
  uint8_t* trace_tmp = first_res ? mem.get<uint8_t>(seg_bytes) : nullptr;
  double* abar = mem.get<double>(9 * Np);
  double* rbar = mem.get<double>(2 * N);
  double* ubar = mem.get<double>(4 * N);
  double* galpha = h->alpha ? mem.get<double>(N) : nullptr;
  double* d_acc = mem.get<double>(2);  // J, dJ/du_in
  if ((first_res && (!ckpt || !trace_tmp)) || (resident && !trace_res) || !fields0 || !abar ||
      !rbar || !ubar || (h->alpha && !galpha) || !d_acc)
    return flb::set_error(h, FLB_ERR_MEMORY, "adjoint buffers (checkpoints, trace, adjoint state)");
  auto seg_trace = [&](uint64_t s) {
    return s >= first_res ? trace_res + (s - first_res) * seg_bytes : trace_tmp;
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
      if ((rc = copy_state(h, ckpt + s * ck_bytes, false, false))) break;
      rc = run_forward(h, count, sp, qs, d_acc);
    } else {
      rc = run_forward(h, count, sp, qs, d_acc, seg_trace(s));
    }
  }
  ADJ_HIP(h, hipEventRecord(ev[1], h->stream));

  // Reverse sweep, one checkpoint segment at a time.
  ADJ_HIP(h, hipMemsetAsync(abar, 0, 9 * Np * sizeof(double), h->stream));
  ADJ_HIP(h, hipMemsetAsync(rbar, 0, 2 * N * sizeof(double), h->stream));
  ADJ_HIP(h, hipMemsetAsync(ubar, 0, 4 * N * sizeof(double), h->stream));
  if (galpha) ADJ_HIP(h, hipMemsetAsync(galpha, 0, N * sizeof(double), h->stream));
  const int abx = 256;
  float recompute_ms = 0.0f;
  for (uint64_t s = nseg; s-- > 0 && !rc;) {
    const uint64_t ts = s * K, te = std::min(ts + K, T);
    uint8_t* trace = seg_trace(s);
    if (s < first_res) {
      if ((rc = copy_state(h, ckpt + s * ck_bytes, true, false))) break;
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
      a.trace = trace + (t - ts) * 3 * N * h->field_bytes;
      a.duin = d_acc + 1;
      a.column = sp->column;
      a.in_window = t >= sp->t0 && t <= sp->t1;
      a.qbar = qs;
      const auto kernel = fp64 ? ((t & 1) ? flb::k_adjoint_step<double, 1> : flb::k_adjoint_step<double, 0>)
                               : ((t & 1) ? flb::k_adjoint_step<float, 1> : flb::k_adjoint_step<float, 0>);
      hipLaunchKernelGGL(kernel, dim3(p.nblocks), dim3(abx), 0, h->stream, p, a);
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
