// flb_setup.cpp -- host-side set-up of the FlowLB core: cell flags of a
// channel (initData, src/initData.h:220-333), free-surface seeding
// (initFreeSurfaceFlags, src/initData.cpp:13-37), free-surface fields
// (initFreeSurfaceFields, src/initData.h:175-206), a synthetic culvert builder
// and the SI <-> lattice unit conversion (FLB::Units, src/utils.cpp:7-81).
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/flb_core.h"

namespace {

constexpr double kShapeEps = 1e-12;  // Shape::m_Epsilon (src/geometry/shapes.h:44)

bool inside(const flb_rect& r, double xs, double ys) {
  if (xs > r.x0 + 0.5 * r.width + kShapeEps) return false;
  if (xs < r.x0 - 0.5 * r.width - kShapeEps) return false;
  if (ys > r.y0 + 0.5 * r.height + kShapeEps) return false;
  if (ys < r.y0 - 0.5 * r.height - kShapeEps) return false;
  return true;
}

bool valid_boundary(uint8_t b) {
  return b == FLB_INLET || b == FLB_OUTLET || b == FLB_WALL || b == FLB_FLUID;
}

}  // namespace

extern "C" int flb_build_channel_flags(const flb_channel* g, uint8_t* flags) {
  if (!g || !flags || g->nx < 2 || g->ny < 2 || !(g->dx > 0.0)) return FLB_ERR_ARGUMENT;
  if (!valid_boundary(g->bc_left) || !valid_boundary(g->bc_right) || !valid_boundary(g->bc_up) ||
      !valid_boundary(g->bc_down))
    return FLB_ERR_ARGUMENT;
  if ((g->n_obstacles && !g->obstacles) || (g->n_initial && !g->initial)) return FLB_ERR_ARGUMENT;
  const uint8_t node = g->free_surface ? FLB_GAS : FLB_FLUID;
  const double y_top = g->y_min + double(g->ny - 1) * g->dx;
  for (uint32_t y = 0; y < g->ny; ++y) {
    const double ys = g->y_min + y * g->dx;
    const bool bed = ys <= g->y_min + kShapeEps;
    const bool top = ys >= y_top - kShapeEps;
    for (uint32_t x = 0; x < g->nx; ++x) {
      const double xs = g->x_min + x * g->dx;
      // Boundary rows first (bed, then top), then the side columns.
      uint8_t v = node;
      if (bed) v = g->bc_down;
      else if (top) v = g->bc_up;
      else if (x == 0) v = g->bc_left;
      else if (x == g->nx - 1) v = g->bc_right;
      for (uint32_t k = 0; k < g->n_obstacles; ++k)
        if (inside(g->obstacles[k], xs, ys)) v = FLB_WALL;
      for (uint32_t k = 0; k < g->n_initial; ++k)
        if (inside(g->initial[k], xs, ys)) v = FLB_FLUID;
      flags[size_t(y) * g->nx + x] = v;
    }
  }
  if (g->free_surface) return flb_seed_free_surface(flags, g->nx, g->ny);
  return FLB_OK;
}

extern "C" int flb_seed_free_surface(uint8_t* flags, uint32_t nx, uint32_t ny) {
  if (!flags || nx < 1 || ny < 1) return FLB_ERR_ARGUMENT;
  const uint8_t seeders = FLB_FLUID | FLB_INLET | FLB_OUTLET;
  for (uint32_t y = 0; y < ny; ++y)
    for (uint32_t x = 0; x < nx; ++x) {
      uint8_t& c = flags[size_t(y) * nx + x];
      if (c & uint8_t(~seeders)) continue;  // only fluid, inlet, outlet cells act
      if (c & FLB_INLET) c |= FLB_FLUID;
      else if (c & FLB_OUTLET) c |= FLB_GAS;
      if (!(c & FLB_FLUID)) continue;
      // Bounded (non-periodic) 8-neighbourhood: gas neighbours become interface.
      const uint32_t y0 = y ? y - 1 : 0, y1 = y + 1 < ny ? y + 1 : y;
      const uint32_t x0 = x ? x - 1 : 0, x1 = x + 1 < nx ? x + 1 : x;
      for (uint32_t yy = y0; yy <= y1; ++yy)
        for (uint32_t xx = x0; xx <= x1; ++xx) {
          if (xx == x && yy == y) continue;
          uint8_t& nbr = flags[size_t(yy) * nx + xx];
          if (nbr == FLB_GAS) nbr = FLB_INTERFACE;
        }
    }
  return FLB_OK;
}

extern "C" int flb_init_free_surface_fields(const uint8_t* flags, size_t n, float* mass,
                                            float* phi, float* excess) {
  if (!flags || !mass || !phi) return FLB_ERR_ARGUMENT;
  for (size_t i = 0; i < n; ++i) {
    float level = 0.0f;
    switch (flags[i]) {
      case FLB_FLUID:
      case FLB_FLUID | FLB_INLET:
      case FLB_FLUID | FLB_OUTLET: level = 1.0f; break;
      case FLB_INTERFACE: level = 0.5f; break;
      default: break;  // gas, gas outlet and walls start empty
    }
    mass[i] = level;
    phi[i] = level;
    if (excess) excess[i] = 0.0f;
  }
  return FLB_OK;
}

extern "C" int flb_build_culvert(const flb_culvert* c, uint8_t* flags, float* rho0, float* u0) {
  if (!c || !flags || c->nx < 3 || c->ny < 3) return FLB_ERR_ARGUMENT;
  if (c->emb_x0 > c->emb_x1 || c->emb_x1 >= c->nx) return FLB_ERR_ARGUMENT;
  if (c->opening + 1 > c->ny) return FLB_ERR_ARGUMENT;
  // Lattice-unit geometry (spacing 1): the embankment spans the rows above the
  // opening up to the top, the water block sits on the bed from column 1.
  flb_rect emb{0.5 * (c->emb_x0 + c->emb_x1), 0.5 * (double(c->opening + 1) + (c->ny - 1)),
               double(c->ny - 1) - (c->opening + 1), double(c->emb_x1 - c->emb_x0)};
  std::vector<flb_rect> water;
  if (c->water_x1 >= 1 && c->water_depth >= 1)
    water.push_back(flb_rect{0.5 * (1.0 + c->water_x1), 0.5 * (1.0 + c->water_depth),
                             double(c->water_depth) - 1.0, double(c->water_x1) - 1.0});
  flb_channel g{};
  g.nx = c->nx;
  g.ny = c->ny;
  g.dx = 1.0;
  g.free_surface = c->free_surface;
  g.bc_left = c->bc_left;
  g.bc_right = c->bc_right;
  g.bc_up = c->bc_up;
  g.bc_down = c->bc_down;
  g.n_obstacles = c->opening + 1 < c->ny - 1 ? 1 : 0;
  g.obstacles = &emb;
  g.n_initial = uint32_t(water.size());
  g.initial = water.empty() ? nullptr : water.data();
  const int rc = flb_build_channel_flags(&g, flags);
  if (rc != FLB_OK) return rc;
  const size_t N = size_t(c->nx) * c->ny;
  if (rho0) {
    if (c->rho_noise > 0.0) {
      std::mt19937_64 gen(c->seed);
      std::uniform_real_distribution<double> dist(-c->rho_noise, c->rho_noise);
      for (size_t i = 0; i < N; ++i) rho0[i] = float(1.0 + dist(gen));
    } else {
      for (size_t i = 0; i < N; ++i) rho0[i] = 1.0f;
    }
  }
  if (u0) {
    for (size_t i = 0; i < N; ++i) {
      u0[i] = (flags[i] & FLB_INLET) ? c->u_inlet : 0.0f;
      u0[N + i] = 0.0f;
    }
  }
  return FLB_OK;
}

extern "C" int flb_units(const flb_units_in* in, flb_units_out* out) {
  if (!in || !out || in->nx < 2 || !(in->length_si > 0) || !(in->ref_velocity_si > 0) ||
      !(in->ref_velocity_lb > 0))
    return FLB_ERR_ARGUMENT;
  // Units::setConversionParameters (src/utils.cpp:7-13), LB density 1.
  const double L = in->length_si / double(in->nx - 1);
  const double T = L * in->ref_velocity_lb / in->ref_velocity_si;
  const double M = in->density_si * L * L * L;
  out->length_scale = L;
  out->time_scale = T;
  out->mass_scale = M;
  const bool has_tau = std::abs(in->relaxation_time) > 1e-6;
  const float nu = has_tau ? float((in->relaxation_time - 0.5) / 3.0)
                           : float(float(in->kinematic_viscosity) * T / (L * L));
  out->tau = has_tau ? float(in->relaxation_time) : 3.0f * nu + 0.5f;
  if (!(out->tau > 0.5f)) return FLB_ERR_ARGUMENT;
  out->inv_tau = 1.0f / out->tau;
  auto force = [&](double acc, double vol) -> float {
    if (std::abs(acc) > 1e-6) return float(in->density_si * acc * L * L * T * T / M);
    return float(vol * T * T * L * L / M);
  };
  out->fx = in->use_force ? force(in->acceleration_x, in->volumetric_force_x) : 0.0f;
  out->fy = in->use_force ? force(in->acceleration_y, in->volumetric_force_y) : 0.0f;
  const double vscale = L / T;
  out->velocity_lb_x = in->velocity_si_x / vscale;
  out->velocity_lb_y = in->velocity_si_y / vscale;
  return FLB_OK;
}
