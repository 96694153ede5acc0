// Allen-Cahn equation with a backward-Euler step solved by Newton's method, written
// against PRISMS-PF's application API: the equations, constants and initial condition
// of the reference's applications/allen_cahn/implicit/custom_pde.h:29-127.
//
//   d n / dt = -m_well (f'(n) - kappa Laplace(n)),  f'(n) = 4 n (n - 1)(n - 1/2)
//
// Newton residual (block 1 rhs), for the unknown n at the new time:
//   value term    n_old - n - dt m_well f'(n)
//   gradient term -dt kappa m_well grad(n)
// Jacobian applied to the update dn (block 1 lhs):
//   value term    dn (1 + dt m_well f''(n)),  f''(n) = 12 n (n - 1) + 2
//   gradient term dt kappa m_well grad(dn)
// Post-processing (block 2): |grad n| and the free-energy density
//   n^4 - 2 n^3 + n^2 + kappa/2 |grad n|^2.
//
// Model constants: m_well, kappa.
#include <prismspf/core/pde_operator_base.h>

#include <algorithm>
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
    , m_well(get_user_inputs().user_constants.get_double("m_well"))
    , kappa(get_user_inputs().user_constants.get_double("kappa"))
  {
    const auto &mesh = get_user_inputs().spatial_discretization.rectangular_mesh;
    for (unsigned int d = 0; d < dim; ++d)
      size[d] = mesh.size[d];
  }

  /// Twelve grains, 0.5 (1 - tanh((r - R_i) / 1.5)) each, summed and capped at 1; the
  /// centres are fractions of the domain size (reference custom_pde.h:37-70).
  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {
    static constexpr double fx[12]     = {0.1, 0.8, 0.5, 0.4, 0.3, 0.8, 0.9, 0.0, 0.1, 0.5, 1.0, 0.7};
    static constexpr double fy[12]     = {0.3, 0.7, 0.2, 0.4, 0.9, 0.1, 0.5, 0.1, 0.6, 0.6, 1.0, 0.95};
    static constexpr double radius[12] = {12, 14, 19, 16, 11, 12, 17, 15, 20, 10, 11, 14};
    double                  value      = 0.0;
    for (int k = 0; k < 12; ++k)
      {
        // The third centre coordinate is 0 for every grain.
        const double c[3] = {fx[k] * size[0], fy[k] * size[1], 0.0};
        double       r2   = 0.0;
        for (unsigned int d = 0; d < dim; ++d)
          r2 += (point[d] - c[d]) * (point[d] - c[d]);
        value += 0.5 * (1.0 - std::tanh((std::sqrt(r2) - radius[k]) / 1.5));
      }
    scalar_value = number(std::min(value, 1.0));
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    const number dt = number(sim_timer.get_timestep());
    if (solve_block_id == 1)
      {
        const ScalarValue n      = variable_list.template get_value<Scalar, Current>(0);
        const ScalarGrad  n_grad = variable_list.template get_gradient<Scalar, Current>(0);
        const ScalarValue n_old  = variable_list.template get_value<Scalar, OldOne>(0);
        const ScalarValue fprime = 4.0 * n * (n - 1.0) * (n - 0.5);
        variable_list.set_value_term(0, n_old - n - dt * m_well * fprime);
        variable_list.set_gradient_term(0, -dt * kappa * m_well * n_grad);
      }
    else if (solve_block_id == 2)
      {
        const ScalarValue n      = variable_list.template get_value<Scalar, Current>(0);
        const ScalarGrad  n_grad = variable_list.template get_gradient<Scalar, Current>(0);
        const ScalarValue f_chem = n * n * n * n - 2.0 * n * n * n + n * n;
        variable_list.set_value_term(1, n_grad.norm());
        variable_list.set_value_term(2, f_chem + 0.5 * kappa * n_grad.norm_square());
      }
  }

  void
  compute_lhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    if (solve_block_id == 1)
      {
        const number      dt      = number(sim_timer.get_timestep());
        const ScalarValue n       = variable_list.template get_value<Scalar, Current>(0);
        const ScalarValue dn      = variable_list.template get_value<Scalar, Change>(0);
        const ScalarGrad  dn_grad = variable_list.template get_gradient<Scalar, Change>(0);
        const ScalarValue fsecond = 12.0 * n * (n - 1.0) + 2.0;
        variable_list.set_value_term(0, dn * (1.0 + dt * m_well * fsecond));
        variable_list.set_gradient_term(0, dt * kappa * m_well * dn_grad);
      }
  }

private:
  number m_well;
  number kappa;
  double size[3] {};
};

PRISMS_PF_END_NAMESPACE
