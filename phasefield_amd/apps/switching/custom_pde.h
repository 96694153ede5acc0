// Test-bed equations for the routing of hooks between the split (stencil) path and the
// generic cell path, written against PRISMS-PF's application API.  Three explicit
// diffusion blocks whose flux coefficients are constant on some inputs only:
//
//   block 0 (c): value c_old, gradient -dt D grad(c_old)             (constant: split)
//   block 1 (s): value s_old, gradient -dt D b(c_old) grad(s_old),
//                b = 1 for c_old <= 1, 2 above                        (value branch)
//   block 2 (r): value r_old, gradient -dt D a(n) grad(r_old),
//                a = 1 before increment 3, 3 from it on                (increment branch)
//   block 3 (u): value u_old, gradient -dt D e(c_old) grad(u_old),
//                e = 1 for c_old <= 60, 2 above                        (value branch
//                                                                       beyond the probe)
//
// Model constants: D, c_low, c_high (the initial c ramps from c_low at x = 0 to c_high).
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
    , D(get_user_inputs().user_constants.get_double("D"))
    , c_low(get_user_inputs().user_constants.get_double("c_low"))
    , c_high(get_user_inputs().user_constants.get_double("c_high"))
    , width(get_user_inputs().spatial_discretization.rectangular_mesh.size[0])
  {}

private:
  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {
    const double x = point[0] / width;
    const double y = dim > 1 ? point[1] / width : 0.0;
    if (index == 0)
      scalar_value = number(c_low + (c_high - c_low) * x + 0.3 * std::sin(6.0 * y));
    else
      scalar_value = number(std::sin(3.0 * x + 1.0 * index) * std::cos(4.0 * y));
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    const number dt = number(sim_timer.get_timestep());
    if (solve_block_id == 0)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, OldOne>(0);
        ScalarGrad  cx = variable_list.template get_gradient<Scalar, OldOne>(0);
        variable_list.set_value_term(0, c);
        variable_list.set_gradient_term(0, -dt * D * cx);
      }
    else if (solve_block_id == 1)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, OldOne>(0);
        ScalarValue s  = variable_list.template get_value<Scalar, OldOne>(1);
        ScalarGrad  sx = variable_list.template get_gradient<Scalar, OldOne>(1);
        const ScalarValue b = dealii::compare_and_apply_mask<dealii::SIMDComparison::greater_than>(
          c, ScalarValue(1.0), ScalarValue(2.0), ScalarValue(1.0));
        variable_list.set_value_term(1, s);
        variable_list.set_gradient_term(1, -dt * D * b * sx);
      }
    else if (solve_block_id == 2)
      {
        ScalarValue  r  = variable_list.template get_value<Scalar, OldOne>(2);
        ScalarGrad   rx = variable_list.template get_gradient<Scalar, OldOne>(2);
        const number a  = sim_timer.get_increment() >= 3 ? number(3) : number(1);
        variable_list.set_value_term(2, r);
        variable_list.set_gradient_term(2, -dt * D * a * rx);
      }
    else if (solve_block_id == 3)
      {
        ScalarValue c  = variable_list.template get_value<Scalar, OldOne>(0);
        ScalarValue u  = variable_list.template get_value<Scalar, OldOne>(3);
        ScalarGrad  ux = variable_list.template get_gradient<Scalar, OldOne>(3);
        const ScalarValue e = dealii::compare_and_apply_mask<dealii::SIMDComparison::greater_than>(
          c, ScalarValue(60.0), ScalarValue(2.0), ScalarValue(1.0));
        variable_list.set_value_term(3, u);
        variable_list.set_gradient_term(3, -dt * D * e * ux);
      }
  }

  number D;
  double c_low;
  double c_high;
  double width;
};

PRISMS_PF_END_NAMESPACE
