// solver_driver.cpp -- headless replay of the reference's frame loop
// (src/Application.cpp:530-540) through the C++ adapter femhip::HIPSolver:
// every frame gathers the surface values, advances one step, checks the
// instability rule (clearing and stopping on it, as the reference pauses),
// and scatters the values back.
//
// usage: solver_driver <in.bin> <out.bin> <equation> <steps> <p0> <p1> [p2 p3 p4 [p5]]
//                      [--velocity-at FRAME X Y Z] [--bc-at FRAME BC]
//   in.bin : int32 nv, int32 ntri, int32 bc, float xyz[3 nv], uint32 tri[3 ntri],
//            uint8 on_boundary[nv], float values[nv]
//   out.bin: int32 frames_run, int32 unstable, float values[nv]
//   equation: heat | advdiff | wave | grayscott  (parameters in the order of
//             EquationParameters.hpp; grayscott's p5 is the reaction
//             weighting, 0 nodal as the reference, 1 lumped; advdiff starts
//             with the default velocity)
//   --velocity-at: before frame FRAME the velocity changes to (X, Y, Z) and
//             the matrices are re-assembled, as the reference does when its
//             velocity slider moves (Application.cpp:394)
//   --bc-at:  before frame FRAME the boundary condition changes to BC (0
//             Dirichlet, 1 Neumann), as the reference's combo box does
//             (Application.cpp:313-315: clear_solver, then
//             update_boundary_conditions, which re-numbers the unknowns); the
//             adapter rebuilds its context and solver, and the surface is then
//             painted again with the initial values
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "femhip_solver.hpp"

template <typename T>
static void rd(FILE* f, T* p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short read");
}

int main(int argc, char** argv) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s in.bin out.bin equation steps p0 p1 [p2 p3 p4]\n", argv[0]);
    return 2;
  }
  try {
    FILE* f = fopen(argv[1], "rb");
    if (!f) throw std::runtime_error("cannot open input");
    int32_t nv = 0, nt = 0, bc = 0;
    rd(f, &nv, 1);
    rd(f, &nt, 1);
    rd(f, &bc, 1);
    std::vector<float> xyz(3 * (size_t)nv), values(nv);
    std::vector<uint32_t> tri(3 * (size_t)nt);
    std::vector<uint8_t> onb(nv);
    rd(f, xyz.data(), xyz.size());
    rd(f, tri.data(), tri.size());
    rd(f, onb.data(), onb.size());
    rd(f, values.data(), values.size());
    fclose(f);

    const std::string eq = argv[3];
    const int steps = atoi(argv[4]);
    double p[6] = {0, 0, 0, 0, 0, 0};
    int np = 0, switch_frame = -1, bc_frame = -1, bc2 = 0;
    double vel2[3] = {0, 0, 0};
    for (int i = 5; i < argc; i++) {
      if (std::string(argv[i]) == "--velocity-at") {
        if (i + 4 >= argc) throw std::runtime_error("--velocity-at needs FRAME X Y Z");
        switch_frame = atoi(argv[i + 1]);
        for (int q = 0; q < 3; q++) vel2[q] = atof(argv[i + 2 + q]);
        i += 4;
      } else if (std::string(argv[i]) == "--bc-at") {
        if (i + 2 >= argc) throw std::runtime_error("--bc-at needs FRAME BC");
        bc_frame = atoi(argv[i + 1]);
        bc2 = atoi(argv[i + 2]);
        i += 2;
      } else if (np < 6) {
        p[np++] = atof(argv[i]);
      }
    }

    auto ctx = std::make_unique<femhip::Context>(xyz.data(), nv, tri.data(), nt, onb.data(), bc);
    std::unique_ptr<femhip::HIPSolver> solver;
    const std::vector<float> initial = values;
    auto setup = [&]() {
      solver = std::make_unique<femhip::HIPSolver>(*ctx);
      if (eq == "heat") {
        solver->set_equation(femhip::HeatParameters{p[0], p[1]});
      } else if (eq == "advdiff") {
        femhip::AdvectionDiffusionParameters a;
        a.time_step = p[0];
        a.c = p[1];
        ctx->assemble(a.velocity);
        solver->set_equation(a);
      } else if (eq == "wave") {
        solver->set_equation(femhip::WaveParameters{p[0], p[1]});
      } else if (eq == "grayscott") {
        femhip::ReactionDiffusionParameters r;
        r.time_step = p[0];
        r.Du = p[1];
        r.Dv = p[2];
        r.feed_rate = p[3];
        r.kill_rate = p[4];
        r.reaction_weighting = (int)p[5];
        solver->set_equation(r);
      } else {
        throw std::runtime_error("unknown equation " + eq);
      }
      solver->attach_surface_values(values.data());
    };
    setup();
    int frames = 0, unstable = 0;
    for (; frames < steps; frames++) {
      if (frames == switch_frame) ctx->assemble(vel2);  // FEMContext::assemble_matrices
      if (frames == bc_frame) {
        // clear_solver, then a context with the new numbering; the solver
        // goes first, since it refers to the context
        solver.reset();
        ctx = std::make_unique<femhip::Context>(xyz.data(), nv, tri.data(), nt, onb.data(), bc2);
        values = initial;
        setup();
      }
      solver->advance_time();
      if (solver->has_numerical_instability()) {
        // Application::clear_solver (Application.cpp:579-583): solver and surface values
        solver->clear_values();
        std::fill(values.begin(), values.end(), 0.0f);
        unstable = 1;
        frames++;
        break;
      }
    }
    FILE* o = fopen(argv[2], "wb");
    if (!o) throw std::runtime_error("cannot open output");
    fwrite(&frames, sizeof(int32_t), 1, o);
    fwrite(&unstable, sizeof(int32_t), 1, o);
    fwrite(values.data(), sizeof(float), values.size(), o);
    fclose(o);
    printf("frames %d unstable %d unknowns %u\n", frames, unstable, ctx->num_unknowns());
  } catch (const std::exception& e) {
    fprintf(stderr, "solver_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
