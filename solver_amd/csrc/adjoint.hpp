// Discrete adjoint and tangent-linear model of the periodic Taylor-Green step (see adjoint.hip).
#pragma once

#include "stepper.hpp"

namespace ofh {

// transposed ghost fills (periodic faces only); in place on full ghosted storage
void launch_velocity_bc_adjoint(Context& c, const of_grid& g, double* const* v);
void launch_component_bc_adjoint(Context& c, const of_grid& g, int comp, double* f);
void launch_pressure_bc_adjoint(Context& c, const of_grid& g, double* p);
// pbar += scale * Grad^T(gbar) over every cell of the ghosted storage
void launch_gradient_adjoint(Context& c, const of_grid& g, double scale, const double* const* gbar, double* pbar);
// ubar_a += scale * Div^T(dbar) over every face of the ghosted storage
void launch_divergence_adjoint(Context& c, const of_grid& g, double scale, const double* dbar, double* const* ubar);
// linearised TVD advection with the branches frozen at `vel` (velocity with ghosts):
// out = dA(vel)[dvel] on the active range
void launch_advection_tangent(Context& c, const of_grid& g, int scheme, const double* const* vel,
                              const double* const* dvel, double* const* out);
// bar += dA(vel)^T[abar] (abar read on the active range; bar accumulates over ghosted storage)
void launch_advection_adjoint(Context& c, const of_grid& g, int scheme, const double* const* vel,
                              const double* const* abar, double* const* bar);

struct AdjointRun;
AdjointRun* adjoint_create(Context& c, const of_stepper_config& cfg, int checkpoint_every);
void adjoint_destroy(AdjointRun* a);
Stepper& adjoint_stepper(AdjointRun* a);
void adjoint_set_forcing(AdjointRun* a, const double* const* f);
// Forward run from u0 (three face arrays) and p0 (cell array, null: zero); only active entries
// are read. Returns the final kinetic energy.
double adjoint_forward(AdjointRun* a, const double* const* u0, const double* p0, int steps, of_step_metrics* m);
void adjoint_final_velocity(AdjointRun* a, double* const* out);
// Reverse sweep. seed: adjoint of the final velocity storage (null: gradient of the final
// kinetic energy). Outputs (null to skip): gradients w.r.t. the active entries of u0, p0 and
// the forcing field; inactive entries are zero.
void adjoint_backward(AdjointRun* a, const double* const* seed, double* const* gu, double* gp, double* const* gf);
// Tangent-linear run along (du0, dp0, df) (null: zero) from the forward run's initial state:
// du_out = final velocity perturbation (ghosted storage), dke = directional derivative of the
// final kinetic energy.
void adjoint_tangent(AdjointRun* a, const double* const* du0, const double* dp0, const double* const* df,
                     double* const* du_out, double* dke);

}  // namespace ofh
