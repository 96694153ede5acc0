// Fields and solve blocks of the linear elasticity application (the body of the
// reference's applications/mechanics/boundary_value_problem/main.cc:24-33).
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
mechanics_define(std::vector<prismspf::FieldAttributes> &fields,
                 std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("u", Vector)};

  SolveBlock linear_solve;
  linear_solve.id               = 1;
  linear_solve.solve_type       = Linear;
  linear_solve.solve_timing     = Uninitialized;
  linear_solve.field_indices    = {0};
  linear_solve.dependencies_lhs = make_dependency_set(fields, {"grad(lhs(u))"});

  blocks = {linear_solve};
}
