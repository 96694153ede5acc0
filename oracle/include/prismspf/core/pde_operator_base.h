// Oracle PDEOperatorBase (reference: include/prismspf/core/pde_operator_base.h:24-163).
// TEST INFRASTRUCTURE.
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/field_container.h>
#include <prismspf/core/phase_field_tools.h>
#include <prismspf/core/simulation_timer.h>
#include <prismspf/core/types.h>
#include <prismspf/user_inputs/user_input_parameters.h>

PRISMS_PF_BEGIN_NAMESPACE

template <unsigned int dim, unsigned int degree, typename number>
class SolveContext;

template <unsigned int dim, unsigned int degree, typename number>
class PDEOperatorBase
{
public:
  explicit PDEOperatorBase(const UserInputParameters<dim> &in, const PhaseFieldTools<dim> &t)
    : user_inputs(&in)
    , pf_tools(&t)
  {}
  virtual ~PDEOperatorBase() = default;

  virtual void
  set_initial_condition(const unsigned int &, const unsigned int &, const dealii::Point<dim> &,
                        number &, number &) const
  {}

  virtual void
  set_dirichlet(const unsigned int &index, const unsigned int &, const unsigned int &component,
                const dealii::Point<dim> &point, const SimulationTimer &, number &scalar_value,
                number &vector_component_value) const
  {
    set_initial_condition(index, component, point, scalar_value, vector_component_value);
  }

  virtual void
  compute_rhs(FieldContainer<dim, degree, number> &, const SimulationTimer &, unsigned int) const
  {}

  virtual void
  compute_lhs(FieldContainer<dim, degree, number> &, const SimulationTimer &, unsigned int) const
  {}

  virtual void
  pre_solve_block(SolveContext<dim, degree, number> &, unsigned int)
  {}

  virtual void
  post_solve_block(SolveContext<dim, degree, number> &, unsigned int)
  {}

  const UserInputParameters<dim> &
  get_user_inputs() const
  {
    return *user_inputs;
  }

  const PhaseFieldTools<dim> &
  get_pf_tools() const
  {
    return *pf_tools;
  }

  void
  set_user_stop(bool f)
  {
    user_stop = f;
  }

  bool
  get_user_stop()
  {
    return user_stop;
  }

private:
  const UserInputParameters<dim> *user_inputs;
  const PhaseFieldTools<dim>     *pf_tools;
  bool                            user_stop = false;
};

PRISMS_PF_END_NAMESPACE
