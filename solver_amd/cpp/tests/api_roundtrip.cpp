// Exercises the C++ interface end to end for the parity tests: reads u, v, w, p (reference
// storage layout, FP64, concatenated) from a file, applies the operators, the ADI predictor and
// the projection through the C++ functions, and writes every result (concatenated) to a file.
// usage: api_roundtrip nx ny nz dx dy dz periodic|cavity in.f64 out.f64
#include "bc/boundary_conditions.hpp"
#include "linsolve/poisson_solver.hpp"
#include "operators/discrete_operators.hpp"
#include "solver/momentum_terms.hpp"
#include "solver/projection.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

int main(int argc, char** argv) {
  using namespace solver;
  try {
    if(argc != 10) throw std::invalid_argument("usage: api_roundtrip nx ny nz dx dy dz periodic|cavity in out");
    const Grid g{std::stoi(argv[1]), std::stoi(argv[2]), std::stoi(argv[3]), std::stod(argv[4]), std::stod(argv[5]),
                 std::stod(argv[6]), 1};
    const std::string kind = argv[7];
    BoundaryConditionSet bcs = BoundaryConditionSet::all(PhysicalBoundaryType::periodic);
    if(kind == "cavity") {
      bcs = BoundaryConditionSet::cavity();
      bcs[BoundaryFace::y_max].type = PhysicalBoundaryType::prescribed_velocity;
      bcs[BoundaryFace::y_max].velocity = {1.0, 0.0, 0.0};
      bcs[BoundaryFace::z_min].type = PhysicalBoundaryType::symmetry;
      bcs[BoundaryFace::z_max].type = PhysicalBoundaryType::symmetry;
    } else if(kind != "periodic") {
      throw std::invalid_argument("unknown boundary kind");
    }
    VelocityField vel{g};
    PressureField p{g};
    std::ifstream in(argv[8], std::ios::binary);
    for(StructuredField* f : {static_cast<StructuredField*>(&vel.x), static_cast<StructuredField*>(&vel.y),
                              static_cast<StructuredField*>(&vel.z), static_cast<StructuredField*>(&p)})
      in.read(reinterpret_cast<char*>(f->data()), static_cast<std::streamsize>(f->size() * sizeof(double)));
    if(!in) throw std::runtime_error("short input file");
    std::ofstream out(argv[9], std::ios::binary);
    auto put = [&](const StructuredField& f) {
      out.write(reinterpret_cast<const char*>(f.data()), static_cast<std::streamsize>(f.size() * sizeof(double)));
    };
    auto put_vel = [&](const VelocityField& v) {
      put(v.x);
      put(v.y);
      put(v.z);
    };
    VelocityField adv{g}, dif{g}, grad{g};
    compute_advection_term(vel, AdvectionOptions{}, adv);
    put_vel(adv);
    compute_diffusion_term(vel, 0.01, dif);
    put_vel(dif);
    operators::compute_gradient(p, grad);
    put_vel(grad);
    ScalarField div{g};
    operators::compute_divergence(vel, div);
    put(div);
    VelocityField bc = vel;
    apply_velocity_boundary_conditions(bcs, bc);
    put_vel(bc);
    VelocityField star{g};
    const HelmholtzDiagnostics h = solve_predictor_adi(bc, 0.003, bcs, star);
    put_vel(star);
    PressureField phi{g};
    VelocityField corrected{g};
    ScalarField rhs{g};
    const ProjectionOptions opts{0.01, 1.0, 400, 1.0e-11};
    const ProjectionDiagnostics d = project_velocity(star, bcs, opts, phi, corrected, &rhs);
    put(phi);
    put_vel(corrected);
    put(rhs);
    std::printf("line_solves: %d\niterations: %d\nconverged: %d\nrelative_residual: %.17g\n"
                "divergence_l2_before: %.17g\ndivergence_l2_after: %.17g\n",
                h.line_solves, d.pressure_solve.iterations, d.pressure_solve.converged ? 1 : 0,
                d.pressure_solve.relative_residual, d.divergence_l2_before, d.divergence_l2_after);
    return 0;
  } catch(const std::exception& e) {
    std::fprintf(stderr, "api_roundtrip failed: %s\n", e.what());
    return 1;
  }
}
