// femhip_solver.hpp -- header-only C++ mirror of the reference's Solver
// interface (include/FEM/Solver.hpp:8-15 of tvumcc/fea-visualizer) over the
// C ABI of include/femhip.h.
//
// femhip::HIPSolver keeps the three virtuals of Solver -- advance_time(),
// clear_values(), has_numerical_instability() -- and the data flow of the
// reference's CPUSolver::advance_time (src/FEM/CPUSolver.cpp:64-147): each
// step gathers the surface values through idx_map, advances one implicit
// step, and scatters the result back to the surface (boundary values 0).
// Which field the surface carries follows the reference: u for heat,
// advection-diffusion and wave, v for Gray-Scott (CPUSolver.cpp:130,144).
//
// A reference build derives its third backend from this class (see the
// integration notes); here it is used directly by a headless driver.
// Failures of the library are reported as std::runtime_error carrying
// fem_last_error(); numerical blow-up is reported, as in the reference, by
// has_numerical_instability().
#ifndef FEMHIP_SOLVER_HPP
#define FEMHIP_SOLVER_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "femhip.h"

namespace femhip {

inline int check(int rc, const char* what) {
  if (rc < 0) throw std::runtime_error(std::string(what) + ": " + fem_last_error());
  return rc;
}

// Parameters of include/FEM/EquationParameters.hpp:4-49 with the reference's
// defaults.
struct HeatParameters { double time_step = 0.01, conductivity = 0.05; };
struct AdvectionDiffusionParameters {
  double time_step = 0.001, c = 0.25;
  double velocity[3] = {1.0, 0.0, 0.0};
};
struct WaveParameters { double time_step = 0.05, c = 0.05; };
struct ReactionDiffusionParameters {
  double time_step = 0.001, Du = 0.08, Dv = 0.04, feed_rate = 0.035, kill_rate = 0.06;
  int reaction_weighting = FEM_REACTION_NODAL;
};

// Mesh + boundary condition + assembled matrices (FEMContext).
class Context {
 public:
  Context(const float* xyz, int32_t num_vertices, const uint32_t* triangles, int32_t num_triangles,
          const uint8_t* on_boundary, int32_t boundary_condition) {
    fem_mesh m{num_vertices, num_triangles, xyz, triangles, on_boundary};
    check(fem_ctx_create(&m, boundary_condition, &ctx_), "fem_ctx_create");
  }
  ~Context() {
    if (ctx_) fem_ctx_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  // FEMContext::assemble_matrices with the advection velocity of
  // AdvectionDiffusionParameters (FEMContext.cpp:57-64).
  void assemble(const double velocity[3], bool normalize_velocity = true) {
    fem_assembly_params p{};
    for (int q = 0; q < 3; q++) p.velocity[q] = velocity[q];
    p.normalize_velocity = normalize_velocity ? 1 : 0;
    check(fem_ctx_assemble(ctx_, &p), "fem_ctx_assemble");
  }
  fem_ctx_info info() const {
    fem_ctx_info i{};
    check(fem_ctx_info_get(ctx_, &i), "fem_ctx_info_get");
    return i;
  }
  unsigned num_nodes() const { return (unsigned)info().num_nodes; }
  unsigned num_unknowns() const { return (unsigned)info().num_unknowns; }
  unsigned num_max_nonzeros_per_row() const { return (unsigned)info().max_row_nonzeros; }
  fem_ctx* handle() const { return ctx_; }

 private:
  fem_ctx* ctx_ = nullptr;
};

// The third backend: Solver's three virtuals plus the parameter setters the
// reference reads from FEMContext::parameters.
class HIPSolver {
 public:
  explicit HIPSolver(Context& ctx, const fem_solver_opts* opts = nullptr) : ctx_(ctx) {
    check(fem_solver_create(ctx.handle(), opts, &s_), "fem_solver_create");
  }
  virtual ~HIPSolver() {
    if (s_) fem_solver_destroy(s_);
  }
  HIPSolver(const HIPSolver&) = delete;
  HIPSolver& operator=(const HIPSolver&) = delete;

  void set_equation(const HeatParameters& p) {
    fem_heat_params c{p.time_step, p.conductivity};
    set(FEM_HEAT, &c, sizeof(c));
  }
  void set_equation(const AdvectionDiffusionParameters& p) {
    fem_advdiff_params c{p.time_step, p.c};
    set(FEM_ADVECTION_DIFFUSION, &c, sizeof(c));
  }
  void set_equation(const WaveParameters& p) {
    fem_wave_params c{p.time_step, p.c};
    set(FEM_WAVE, &c, sizeof(c));
  }
  void set_equation(const ReactionDiffusionParameters& p) {
    fem_rd_params c{p.time_step, p.Du, p.Dv, p.feed_rate, p.kill_rate, p.reaction_weighting};
    set(FEM_REACTION_DIFFUSION, &c, sizeof(c));
  }

  // Surface::values (nv floats) the solver reads before and writes after
  // every step, as the reference's gather / scatter (CPUSolver.cpp:13-37).
  void attach_surface_values(float* values_nv) { surface_ = values_nv; }

  // Solver::advance_time
  virtual void advance_time() {
    if (!surface_) throw std::runtime_error("HIPSolver: no surface values attached");
    const int field = surface_field();
    check(fem_solver_set_state(s_, field, surface_), "fem_solver_set_state");
    last_ = check(fem_solver_advance(s_, 1, &stats_), "fem_solver_advance");
    check(fem_solver_get_state(s_, field, surface_), "fem_solver_get_state");
  }
  // Solver::clear_values
  virtual void clear_values() { check(fem_solver_clear(s_), "fem_solver_clear"); }
  // Solver::has_numerical_instability
  virtual bool has_numerical_instability() {
    int32_t u = 0;
    check(fem_solver_unstable(s_, &u), "fem_solver_unstable");
    return u != 0;
  }

  // direct state access (surface space, nv values)
  void set_state(int field, const float* values_nv) { check(fem_solver_set_state(s_, field, values_nv), "set_state"); }
  void get_state(int field, float* values_nv) { check(fem_solver_get_state(s_, field, values_nv), "get_state"); }

  const fem_step_stats& last_stats() const { return stats_; }
  int last_outcome() const { return last_; }
  fem_solver* handle() const { return s_; }

 private:
  void set(int eq, const void* p, size_t n) {
    check(fem_solver_set_equation(s_, eq, p, n), "fem_solver_set_equation");
    eq_ = eq;
  }
  int surface_field() const { return eq_ == FEM_REACTION_DIFFUSION ? FEM_FIELD_V : FEM_FIELD_U; }

  Context& ctx_;
  fem_solver* s_ = nullptr;
  float* surface_ = nullptr;
  int eq_ = -1;
  int last_ = FEM_OK;
  fem_step_stats stats_{};
};

}  // namespace femhip

#endif  // FEMHIP_SOLVER_HPP
