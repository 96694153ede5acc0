// Cahn-Hilliard equation with a backward-Euler step for (c, mu) solved by Newton's
// method, written against PRISMS-PF's application API: the equations and constants of
// the reference's applications/cahn_hilliard/implicit/custom_pde.h:93-167.
//
//   dc/dt = div(M grad mu),  mu = W c (c - 1)(c - 1/2) - K Laplace(c)
//
// Newton residual (block 0 rhs) at the new (c, mu):
//   c:  value term  c_old - c,                    gradient term -dt M grad(mu)
//   mu: value term  mu - W c (c - 1)(c - 1/2),    gradient term -K grad(c)
// Jacobian applied to the update (dc, dmu) (block 0 lhs), a non-symmetric operator:
//   c:  value term  dc,                           gradient term  dt M grad(dmu)
//   mu: value term  W (1/2 + 3 c (c - 1)) dc - dmu, gradient term  K grad(dc)
// Post-processing (block 2): c^4 - 2 c^3 + c^2 + K/2 |grad c|^2.
//
// Model constants: McV (M), KcV (K), WcV (W), ic_type, c0, icamplitude.
// Initial condition: ic_type 0 is c0 + icamplitude * U[-1, 1] at every node.  The
// reference draws U from a std::mt19937_64 in deal.II's DoF traversal order; this
// application keys each draw by (random seed, node) instead (order-independent).
// ic_type 1 is 0.5 (1 - tanh(d / 1.5)) of the signed distance d to the nearest of
// twelve grains (reference custom_pde.h:56-86).  mu starts at 0.
#include <prismspf/core/pde_operator_base.h>

#include "../common/seeded_noise.h"

#include <algorithm>
#include <cmath>
#include <limits>

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
    , McV(get_user_inputs().user_constants.get_double("McV"))
    , KcV(get_user_inputs().user_constants.get_double("KcV"))
    , WcV(get_user_inputs().user_constants.get_double("WcV"))
    , ic_type(get_user_inputs().user_constants.get_int("ic_type"))
    , c0(get_user_inputs().user_constants.get_double("c0"))
    , icamplitude(get_user_inputs().user_constants.get_double("icamplitude"))
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

  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {
    if (index != 0)
      return;
    if (ic_type == 0)
      {
        double x[3] = {0, 0, 0};
        for (unsigned int d = 0; d < dim; ++d)
          x[d] = point[d];
        const uint64_t key = ppf_noise::node_key<dim>(x, lower, size, cells, degree, periodic);
        scalar_value = number(c0 + icamplitude * ppf_noise::uniform(seed, 4, key, -1.0, 1.0));
      }
    else
      {
        static constexpr double fx[12] = {0.1, 0.8, 0.5, 0.4, 0.3, 0.8, 0.9, 0.0, 0.1, 0.5, 1.0, 0.7};
        static constexpr double fy[12] = {0.3, 0.7, 0.2, 0.4, 0.9, 0.1, 0.5, 0.1, 0.6, 0.6, 1.0, 0.95};
        static constexpr double radius[12] = {12, 14, 19, 16, 11, 12, 17, 15, 20, 10, 11, 14};
        double                  sdf        = std::numeric_limits<double>::max();
        for (int k = 0; k < 12; ++k)
          {
            const double c[3] = {fx[k] * size[0], fy[k] * size[1], 0.0};
            double       r2   = 0.0;
            for (unsigned int d = 0; d < dim; ++d)
              r2 += (point[d] - c[d]) * (point[d] - c[d]);
            sdf = std::min(sdf, std::sqrt(r2) - radius[k]);
          }
        scalar_value = number(0.5 * (1.0 - std::tanh(sdf / 1.5)));
      }
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    const number dt = number(sim_timer.get_timestep());
    if (solve_block_id == 0)
      {
        const ScalarValue c       = variable_list.template get_value<Scalar, Current>(0);
        const ScalarGrad  c_grad  = variable_list.template get_gradient<Scalar, Current>(0);
        const ScalarValue mu      = variable_list.template get_value<Scalar, Current>(1);
        const ScalarGrad  mu_grad = variable_list.template get_gradient<Scalar, Current>(1);
        const ScalarValue c_old   = variable_list.template get_value<Scalar, OldOne>(0);
        variable_list.set_value_term(0, c_old - c);
        variable_list.set_gradient_term(0, -(dt * McV * mu_grad));
        variable_list.set_value_term(1, mu - WcV * c * (c - 1.0) * (c - 0.5));
        variable_list.set_gradient_term(1, -(KcV * c_grad));
      }
    else if (solve_block_id == 2)
      {
        const ScalarValue c      = variable_list.template get_value<Scalar, Current>(0);
        const ScalarGrad  c_grad = variable_list.template get_gradient<Scalar, Current>(0);
        const ScalarValue f_chem = c * c * c * c - 2.0 * c * c * c + c * c;
        variable_list.set_value_term(2, f_chem + 0.5 * KcV * c_grad.norm_square());
      }
  }

  void
  compute_lhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    if (solve_block_id == 0)
      {
        const number      dt       = number(sim_timer.get_timestep());
        const ScalarValue c        = variable_list.template get_value<Scalar, Current>(0);
        const ScalarValue dc       = variable_list.template get_value<Scalar, Change>(0);
        const ScalarGrad  dc_grad  = variable_list.template get_gradient<Scalar, Change>(0);
        const ScalarValue dmu      = variable_list.template get_value<Scalar, Change>(1);
        const ScalarGrad  dmu_grad = variable_list.template get_gradient<Scalar, Change>(1);
        variable_list.set_value_term(0, dc);
        variable_list.set_gradient_term(0, dt * McV * dmu_grad);
        variable_list.set_value_term(1, WcV * (0.5 + 3.0 * (c * (c - 1.0))) * dc - dmu);
        variable_list.set_gradient_term(1, KcV * dc_grad);
      }
  }

private:
  number   McV;
  number   KcV;
  number   WcV;
  int      ic_type;
  double   c0;
  double   icamplitude;
  uint64_t seed = 2025;
  double   lower[3] {};
  double   size[3] {};
  unsigned cells[3] {};
  bool     periodic[3] {};
};

PRISMS_PF_END_NAMESPACE
