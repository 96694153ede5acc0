// Base class of application equations (reference:
// include/prismspf/core/pde_operator_base.h:24-163).
//
// Applications derive `CustomPDE<dim, degree, number>` from this class and override
// the hooks exactly as for the reference.  The hooks are host+device functions: the
// library compiles the application header inside a `force_cuda_host_device` region,
// so unmodified application code becomes callable from the fused HIP kernels.  A
// kernel receives the application object by value, copies it into a local object and
// calls the hook through this base; because the local object's dynamic type is known,
// the call is devirtualized and inlined into the quadrature-point loop.
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
  using SizeType = dealii::VectorizedArray<number>;

  explicit PDEOperatorBase(const UserInputParameters<dim> &_user_inputs,
                           const PhaseFieldTools<dim>     &_pf_tools)
    : user_inputs(&_user_inputs)
    , pf_tools(&_pf_tools)
  {}

  PPF_HD PDEOperatorBase(const PDEOperatorBase &) = default;

  PPF_HD virtual ~PDEOperatorBase() = default;

  /// Initial condition at a support point (reference pde_operator_base.h:46-52).
  PPF_HD virtual void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const
  {}

  /// Dirichlet value; defaults to the initial condition (reference :58-72).
  PPF_HD virtual void
  set_dirichlet([[maybe_unused]] const unsigned int       &index,
                [[maybe_unused]] const unsigned int       &boundary_id,
                [[maybe_unused]] const unsigned int       &component,
                [[maybe_unused]] const dealii::Point<dim> &point,
                [[maybe_unused]] const SimulationTimer    &sim_timer,
                [[maybe_unused]] number                   &scalar_value,
                [[maybe_unused]] number                   &vector_component_value) const
  {
    this->set_initial_condition(index, component, point, scalar_value,
                                vector_component_value);
  }

  /// Right-hand side of explicit equations and of linear/Newton systems (reference :77-81).
  PPF_HD virtual void
  compute_rhs([[maybe_unused]] FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer               &sim_timer,
              [[maybe_unused]] unsigned int                         solver_id) const
  {}

  /// Left-hand side (the linear operator) of nonexplicit equations (reference :86-90).
  PPF_HD virtual void
  compute_lhs([[maybe_unused]] FieldContainer<dim, degree, number> &variable_list,
              [[maybe_unused]] const SimulationTimer               &sim_timer,
              [[maybe_unused]] unsigned int                         solver_id) const
  {}

  /// Called on the host before and after a solve block (reference :96-108).
  PPF_HD virtual void
  pre_solve_block([[maybe_unused]] SolveContext<dim, degree, number> &solve_context,
                  [[maybe_unused]] unsigned int                       solver_id)
  {}

  PPF_HD virtual void
  post_solve_block([[maybe_unused]] SolveContext<dim, degree, number> &solve_context,
                   [[maybe_unused]] unsigned int                       solver_id)
  {}

  [[nodiscard]] const UserInputParameters<dim> &
  get_user_inputs() const
  {
    return *user_inputs;
  }

  [[nodiscard]] const PhaseFieldTools<dim> &
  get_pf_tools() const
  {
    return *pf_tools;
  }

  void
  set_user_stop(bool flag)
  {
    user_stop = flag;
  }

  bool
  get_user_stop()
  {
    return user_stop;
  }

private:
  const UserInputParameters<dim> *user_inputs = nullptr;
  const PhaseFieldTools<dim>     *pf_tools    = nullptr;
  bool                            user_stop   = false;
};

PRISMS_PF_END_NAMESPACE
