// Test-bed equations for the generic cell path, written against PRISMS-PF's
// application API.  None of them has the collocation-split structure:
//
//   block 0 (c, explicit):   nonlinear diffusion, value c_old,
//                            gradient -dt D0 (1 + c_old^2) grad(c_old)
//   block 1 (v, linear):     alpha v - div((1 + c^2) grad v) = c
//                            (lhs value alpha lhs(v), gradient (1 + c^2) grad(lhs(v)))
//   block 2 (g, postprocess): |grad c|
//
// Model constants: D0, alpha, amplitude.
#include <prismspf/core/pde_operator_base.h>

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
    , D0(get_user_inputs().user_constants.get_double("D0"))
    , alpha(get_user_inputs().user_constants.get_double("alpha"))
    , amplitude(get_user_inputs().user_constants.get_double("amplitude"))
  {}

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
    double s = 0.0;
    for (unsigned int d = 0; d < dim; ++d)
      s += std::sin(0.7 * (d + 1) * point[d] + 0.3 * d);
    scalar_value = number(amplitude * s);
  }

  void
  set_dirichlet([[maybe_unused]] const unsigned int       &index,
                [[maybe_unused]] const unsigned int       &boundary_id,
                [[maybe_unused]] const unsigned int       &component,
                [[maybe_unused]] const dealii::Point<dim> &point,
                [[maybe_unused]] const SimulationTimer    &sim_timer,
                [[maybe_unused]] number                   &scalar_value,
                [[maybe_unused]] number &vector_component_value) const override
  {
    // The time term changes the values only where the boundary conditions are declared
    // time dependent (they are otherwise evaluated once, at time 0).
    scalar_value = number(0.25 * (boundary_id + 1) + 0.1 * point[0] + 2.0 * sim_timer.get_time());
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    if (solve_block_id == 0)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, OldOne>(0);
        ScalarGrad  cx = variable_list.template get_gradient<Scalar, OldOne>(0);
        variable_list.set_value_term(0, c);
        variable_list.set_gradient_term(0, -sim_timer.get_timestep() * D0 * (1.0 + c * c) * cx);
      }
    else if (solve_block_id == 1)
      variable_list.set_value_term(1, variable_list.template get_value<Scalar, Current>(0));
    else if (solve_block_id == 2)
      {
        ScalarGrad cx = variable_list.template get_gradient<Scalar, Current>(0);
        variable_list.set_value_term(2, cx.norm());
      }
  }

  void
  compute_lhs(FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer &sim_timer,
              unsigned int                             solve_block_id) const override
  {
    if (solve_block_id == 1)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, Current>(0);
        ScalarValue v  = variable_list.template get_value<Scalar, LHS>(1);
        ScalarGrad  vx = variable_list.template get_gradient<Scalar, LHS>(1);
        variable_list.set_value_term(1, alpha * v);
        variable_list.set_gradient_term(1, (1.0 + c * c) * vx);
      }
  }

  number D0;
  number alpha;
  double amplitude;
};

PRISMS_PF_END_NAMESPACE
