// Linear elasticity boundary-value problem, written against PRISMS-PF's application API
// (the equations of the reference's applications/mechanics/boundary_value_problem/
// custom_pde.h:41-83): one vector field u (displacement), one linear solve block with a
// zero right-hand side and the lhs gradient term sigma(eps(lhs(u))), where eps is the
// symmetric gradient and sigma = C eps with C from the `stiffness` elastic constants.
// Dirichlet values are zero except u_x = -1 on the x = 0 face.
#include <prismspf/core/pde_operator_base.h>

#include <prismspf/utilities/mechanics.h>

PRISMS_PF_BEGIN_NAMESPACE

template <unsigned int dim, unsigned int degree, typename number>
class CustomPDE : public PDEOperatorBase<dim, degree, number>
{
public:
  using ScalarValue = dealii::VectorizedArray<number>;
  using VectorValue = dealii::Tensor<1, dim, ScalarValue>;
  using VectorGrad  = dealii::Tensor<2, dim, ScalarValue>;
  using PDEOperatorBase<dim, degree, number>::get_user_inputs;

  CustomPDE(const UserInputParameters<dim> &_user_inputs, PhaseFieldTools<dim> &_pf_tools)
    : PDEOperatorBase<dim, degree, number>(_user_inputs, _pf_tools)
    , stiffness(get_user_inputs().user_constants.get_elasticity_tensor("stiffness"))
  {}

  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {}

  void
  set_dirichlet([[maybe_unused]] const unsigned int       &index,
                [[maybe_unused]] const unsigned int       &boundary_id,
                [[maybe_unused]] const unsigned int       &component,
                [[maybe_unused]] const dealii::Point<dim> &point,
                [[maybe_unused]] const SimulationTimer    &sim_timer,
                [[maybe_unused]] number                   &scalar_value,
                [[maybe_unused]] number &vector_component_value) const override
  {
    // Boundary id 0 is the x = 0 face.
    if (index == 0 && boundary_id == 0 && component == 0)
      vector_component_value = number(-1);
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer &sim_timer,
              unsigned int                             solve_block_id) const override
  {
    if (solve_block_id == 1)
      variable_list.set_value_term(0, VectorValue());
  }

  void
  compute_lhs(FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer &sim_timer,
              unsigned int                             solve_block_id) const override
  {
    if (solve_block_id == 1)
      {
        VectorGrad eps = variable_list.template get_symmetric_gradient<Vector, LHS>(0);
        VectorGrad stress;
        Mechanics::compute_stress<dim>(stiffness, eps, stress);
        variable_list.set_gradient_term(0, stress);
      }
  }

private:
  dealii::Tensor<2, Mechanics::voigt_tensor_size<dim>, double> stiffness;
};

PRISMS_PF_END_NAMESPACE
