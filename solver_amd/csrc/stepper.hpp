// Device-resident stepper state (see stepper.hip).
#pragma once

#include "ofh.hpp"

namespace ofh {

of_bc_set stepper_bcs(const of_stepper_config& c);
of_pressure_bc_set derive_pressure_bcs(const of_bc_set& b);

struct Stepper {
  Context& c;
  of_stepper_config cfg;
  of_grid g{};
  of_bc_set bcs{};
  of_pressure_bc_set pbcs{};
  bool all_periodic = false;
  double* vel[3]{};
  double* adv[3]{};
  double* rhs[3]{};
  double* pred[3]{};
  double* corr[3]{};
  double* dif[3]{};
  double* p_total = nullptr;
  double* phi = nullptr;
  double* prhs = nullptr;
  double* diag = nullptr;
  double* force[3]{};  // optional body force f: rhs += dt f on the active range (null: none)
  double* force_buf[3]{};
  PoissonSolver* poisson = nullptr;
  // Helmholtz predictor: OF_PREDICTOR_ADI, or OF_PREDICTOR_CG with its work arrays
  int predictor_kind = OF_PREDICTOR_ADI;
  double predictor_tol = 1e-12;
  int predictor_maxit = 80;
  HelmholtzCG* hcg = nullptr;
  int predictor_iterations = 0;  // CG iterations of the last step (largest component)
  int has_prev = 0;
  of_step_metrics metrics{};
  std::vector<double*> owned;

  Stepper(Context& ctx, const of_stepper_config& cfg);
  ~Stepper();
  void initialize(int convention);
  void step();
  double kinetic_energy();
  void total_pressure_bc(const double* const* v);
  // Copies a device forcing field (three face arrays in the velocity layout) into the stepper;
  // null clears it.
  void set_forcing(const double* const* f);
  // Selects the predictor solver (the CG one needs every axis periodic).
  void set_predictor(int kind, double tol, int maxit);
};

}  // namespace ofh
