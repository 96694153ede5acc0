// Cahn-Hilliard equation split into concentration c and chemical potential mu, explicit
// in time, written against PRISMS-PF's application API (same block structure as the
// reference's applications/cahn_hilliard/explicit/custom_pde.h:91-131 and main.cc:29-50)
// with the free energy f(c) = c^2 (1-c)^2 of the benchmark configuration.
//
//   block 0 (c):  value term c_old,        gradient term -dt M grad(mu_old)
//   block 1 (mu): value term f'(c),        gradient term  kappa grad(c)
//   block 2 (f_tot, postprocess): f(c) + kappa/2 |grad c|^2
//
// Model constants: M, kappa, c0, noise_amplitude.
#include <prismspf/core/pde_operator_base.h>

#include "../common/seeded_noise.h"

PRISMS_PF_BEGIN_NAMESPACE

/// f(c) = c^2 (1-c)^2 and its derivatives f'(c) = 2c(1-c)(1-2c), f''(c) = 2 - 12c + 12c^2.
struct CahnHilliardWell
{
  template <typename T>
  PPF_HD static T
  fprime(const T &c)
  {
    return 2.0 * c * (1.0 - c) * (1.0 - 2.0 * c);
  }

  template <typename T>
  PPF_HD static T
  fprime2(const T &c)
  {
    return 2.0 - 12.0 * c + 12.0 * c * c;
  }
};

template <unsigned int dim, unsigned int degree, typename number>
class CustomPDE : public PDEOperatorBase<dim, degree, number>
{
public:
  using ScalarValue = dealii::VectorizedArray<number>;
  using ScalarGrad  = dealii::Tensor<1, dim, ScalarValue>;
  using PDEOperatorBase<dim, degree, number>::get_user_inputs;

  CustomPDE(const UserInputParameters<dim> &_user_inputs, PhaseFieldTools<dim> &_pf_tools)
    : PDEOperatorBase<dim, degree, number>(_user_inputs, _pf_tools)
    , M(get_user_inputs().user_constants.get_double("M"))
    , kappa(get_user_inputs().user_constants.get_double("kappa"))
    , c0(get_user_inputs().user_constants.get_double("c0"))
    , noise(get_user_inputs().user_constants.get_double("noise_amplitude"))
    , seed(get_user_inputs().misc_parameters.random_seed)
  {
    const auto &mesh = get_user_inputs().spatial_discretization.rectangular_mesh;
    for (unsigned int d = 0; d < dim; ++d)
      {
        lower[d]    = mesh.lower_bound[d];
        size[d]     = mesh.size[d];
        cells[d]    = get_user_inputs().spatial_discretization.n_cells(d);
        periodic[d] = mesh.periodic_directions.count(d) != 0;
      }
  }

private:
  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {
    if (index != 0)
      return;
    double x[3] = {0, 0, 0};
    for (unsigned int d = 0; d < dim; ++d)
      x[d] = point[d];
    const uint64_t key = ppf_noise::node_key<dim>(x, lower, size, cells, degree, periodic);
    scalar_value       = number(c0 + ppf_noise::uniform(seed, 3, key, -noise, noise));
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    if (solve_block_id == 0)
      {
        ScalarValue c   = variable_list.template get_value<Scalar, OldOne>(0);
        ScalarGrad  mux = variable_list.template get_gradient<Scalar, OldOne>(1);
        variable_list.set_value_term(0, c);
        variable_list.set_gradient_term(0, -M * sim_timer.get_timestep() * mux);
      }
    else if (solve_block_id == 1)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, Current>(0);
        ScalarGrad  cx = variable_list.template get_gradient<Scalar, Current>(0);
        variable_list.set_value_term(1, CahnHilliardWell::fprime(c));
        variable_list.set_gradient_term(1, kappa * cx);
      }
    else if (solve_block_id == 2)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, Current>(0);
        ScalarGrad  cx = variable_list.template get_gradient<Scalar, Current>(0);
        ScalarValue f  = c * c * (1.0 - c) * (1.0 - c);
        variable_list.set_value_term(2, f + 0.5 * kappa * cx.norm_square());
      }
  }

  number   M;
  number   kappa;
  double   c0;
  double   noise;
  uint64_t seed;
  double   lower[3] {};
  double   size[3] {};
  unsigned cells[3] {};
  bool     periodic[3] {};
};

PRISMS_PF_END_NAMESPACE
