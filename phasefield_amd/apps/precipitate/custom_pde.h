// Coupled Allen-Cahn / Cahn-Hilliard precipitate model, explicit in time, written
// against PRISMS-PF's application API (block structure as in the reference's
// applications/coupled_allen_cahn_cahn_hilliard/{custom_pde.h:63-176,main.cc}, with the
// free energy of the benchmark configuration):
//
//   f(c, eta) = (c-0.1)^2 (1-h(eta)) + (c-0.9)^2 h(eta) + W eta^2 (1-eta)^2,
//   h(eta)    = eta^3 (6 eta^2 - 15 eta + 10)
//
//   block 0 (eta): value  eta - dt L df/deta(c_old, eta_old), gradient -dt L kappa_eta grad(eta_old)
//   block 1 (c):   value  c_old,                              gradient -dt M grad(mu_old)
//   block 2 (mu):  value  df/dc(c, eta),                      gradient  kappa_c grad(c)
//
// Model constants: W, M, L, kappa_c, kappa_eta, c0, noise_amplitude, n_nuclei,
// nucleus_radius.
#include <prismspf/core/pde_operator_base.h>

#include "../common/seeded_noise.h"

#include <cmath>

PRISMS_PF_BEGIN_NAMESPACE

template <unsigned int dim, unsigned int degree, typename number>
class CustomPDE : public PDEOperatorBase<dim, degree, number>
{
public:
  using ScalarValue = dealii::VectorizedArray<number>;
  using ScalarGrad  = dealii::Tensor<1, dim, ScalarValue>;
  using PDEOperatorBase<dim, degree, number>::get_user_inputs;

  CustomPDE(const UserInputParameters<dim> &_user_inputs, PhaseFieldTools<dim> &_pf_tools)
    : PDEOperatorBase<dim, degree, number>(_user_inputs, _pf_tools)
  {
    const auto &uc = get_user_inputs().user_constants;
    W              = uc.get_double("W");
    M              = uc.get_double("M");
    L              = uc.get_double("L");
    kappa_c        = uc.get_double("kappa_c");
    kappa_eta      = uc.get_double("kappa_eta");
    c0             = uc.get_double("c0");
    noise          = uc.get_double("noise_amplitude");
    n_nuclei       = uc.get_int("n_nuclei");
    radius         = uc.get_double("nucleus_radius");
    seed           = get_user_inputs().misc_parameters.random_seed;
    const auto &mesh = get_user_inputs().spatial_discretization.rectangular_mesh;
    for (unsigned int d = 0; d < dim; ++d)
      {
        lower[d]    = mesh.lower_bound[d];
        size[d]     = mesh.size[d];
        cells[d]    = get_user_inputs().spatial_discretization.n_cells(d);
        periodic[d] = mesh.periodic_directions.count(d) != 0;
      }
    for (int k = 0; k < n_nuclei && k < max_nuclei; ++k)
      for (unsigned int d = 0; d < dim; ++d)
        centre[k][d] = ppf_noise::uniform(seed, 5, k * 4 + d, lower[d], size[d]);
  }

  template <typename T>
  PPF_HD static T
  interp(const T &e)
  {
    return e * e * e * (6.0 * e * e - 15.0 * e + 10.0);
  }

  template <typename T>
  PPF_HD static T
  interp_prime(const T &e)
  {
    return 30.0 * e * e * (1.0 - e) * (1.0 - e);
  }

private:
  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {
    double x[3] = {0, 0, 0};
    for (unsigned int d = 0; d < dim; ++d)
      x[d] = point[d];
    if (index == 0)
      {
        const uint64_t key = ppf_noise::node_key<dim>(x, lower, size, cells, degree, periodic);
        scalar_value       = number(c0 + ppf_noise::uniform(seed, 6, key, -noise, noise));
      }
    else if (index == 2)
      {
        double eta = 0.0;
        for (int k = 0; k < n_nuclei && k < max_nuclei; ++k)
          {
            double r2 = 0.0;
            for (unsigned int d = 0; d < dim; ++d)
              {
                double dx = x[d] - centre[k][d];
                if (periodic[d])
                  {
                    const double len = size[d] - lower[d];
                    dx -= len * std::round(dx / len);
                  }
                r2 += dx * dx;
              }
            eta += 0.5 * (1.0 - std::tanh(std::sqrt(r2) - radius));
          }
        scalar_value = number(eta < 1.0 ? eta : 1.0);
      }
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    const double dt = sim_timer.get_timestep();
    if (solve_block_id == 0)
      {
        const ScalarValue eta  = variable_list.template get_value<Scalar, OldOne>(2);
        const ScalarGrad  etax = variable_list.template get_gradient<Scalar, OldOne>(2);
        const ScalarValue c    = variable_list.template get_value<Scalar, OldOne>(0);
        const ScalarValue dfde = ((c - 0.9) * (c - 0.9) - (c - 0.1) * (c - 0.1)) * interp_prime(eta) +
                                 2.0 * W * eta * (1.0 - eta) * (1.0 - 2.0 * eta);
        variable_list.set_value_term(2, eta - dt * L * dfde);
        variable_list.set_gradient_term(2, -dt * L * kappa_eta * etax);
      }
    else if (solve_block_id == 1)
      {
        const ScalarValue c   = variable_list.template get_value<Scalar, OldOne>(0);
        const ScalarGrad  mux = variable_list.template get_gradient<Scalar, OldOne>(1);
        variable_list.set_value_term(0, c);
        variable_list.set_gradient_term(0, -dt * M * mux);
      }
    else if (solve_block_id == 2)
      {
        const ScalarValue c   = variable_list.template get_value<Scalar, Current>(0);
        const ScalarValue eta = variable_list.template get_value<Scalar, Current>(2);
        const ScalarGrad  cx  = variable_list.template get_gradient<Scalar, Current>(0);
        const ScalarValue h   = interp(eta);
        variable_list.set_value_term(1, 2.0 * (c - 0.1) * (1.0 - h) + 2.0 * (c - 0.9) * h);
        variable_list.set_gradient_term(1, kappa_c * cx);
      }
  }

  static constexpr int max_nuclei = 64;
  number               W = 1, M = 1, L = 1, kappa_c = 1, kappa_eta = 1;
  double               c0 = 0.3, noise = 0.01, radius = 4.0;
  int                  n_nuclei = 20;
  uint64_t             seed     = 2025;
  double               centre[max_nuclei][3] {};
  double               lower[3] {};
  double               size[3] {};
  unsigned             cells[3] {};
  bool                 periodic[3] {};
};

PRISMS_PF_END_NAMESPACE
