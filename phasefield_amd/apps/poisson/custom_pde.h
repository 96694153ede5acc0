// Poisson problem -Laplace(u) = f with a linear solve block, written against PRISMS-PF's
// application API (same structure as the reference's applications/laplace/custom_pde.h:
// 66-87 and main.cc:28-36).  The right-hand side f is a Constant field initialized with
// seeded N(0,1) values per node, so the right-hand side vector is b = M f.
//
//   block 0 (f):  Constant, Primary (initial condition only)
//   block 1 (u):  Linear;  rhs value term f,  lhs gradient term grad(lhs(u))
#include <prismspf/core/pde_operator_base.h>

#include "../common/seeded_noise.h"

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
    if (index != 1)
      return;
    double x[3] = {0, 0, 0};
    for (unsigned int d = 0; d < dim; ++d)
      x[d] = point[d];
    const uint64_t key = ppf_noise::node_key<dim>(x, lower, size, cells, degree, periodic);
    scalar_value       = number(ppf_noise::normal(seed, 4, key));
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
    scalar_value = number(0);
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer &sim_timer,
              unsigned int                             solve_block_id) const override
  {
    if (solve_block_id == 1)
      variable_list.set_value_term(0, variable_list.template get_value<Scalar, Current>(1));
  }

  void
  compute_lhs(FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer &sim_timer,
              unsigned int                             solve_block_id) const override
  {
    if (solve_block_id == 1)
      {
        ScalarGrad ux = variable_list.template get_gradient<Scalar, LHS>(0);
        variable_list.set_gradient_term(0, ux);
      }
  }

private:
  uint64_t seed;
  double   lower[3] {};
  double   size[3] {};
  unsigned cells[3] {};
  bool     periodic[3] {};
};

PRISMS_PF_END_NAMESPACE
