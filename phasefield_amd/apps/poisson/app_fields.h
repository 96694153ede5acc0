// Fields and solve blocks of the Poisson application (the body of a reference main.cc,
// cf. applications/laplace/main.cc:26-46).
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
poisson_define(std::vector<prismspf::FieldAttributes> &fields,
               std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("u"), FieldAttributes("f")};

  SolveBlock rhs_field;
  rhs_field.id            = 0;
  rhs_field.solve_type    = Constant;
  rhs_field.solve_timing  = Initialized;
  rhs_field.field_indices = {1};

  SolveBlock linear_solve;
  linear_solve.id               = 1;
  linear_solve.field_indices    = {0};
  linear_solve.solve_type       = Linear;
  linear_solve.solve_timing     = Uninitialized;
  linear_solve.dependencies_rhs = make_dependency_set(fields, {"f"});
  linear_solve.dependencies_lhs = make_dependency_set(fields, {"grad(lhs(u))"});

  blocks = {rhs_field, linear_solve};
}
