// Fields and solve blocks of the test-bed application.
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
testbed_define(std::vector<prismspf::FieldAttributes> &fields,
               std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("c"), FieldAttributes("v"), FieldAttributes("g")};

  SolveBlock diffusion;
  diffusion.id               = 0;
  diffusion.solve_type       = Explicit;
  diffusion.solve_timing     = Primary;
  diffusion.field_indices    = {0};
  diffusion.dependencies_rhs = make_dependency_set(fields, {"old_1(c)", "grad(old_1(c))"});

  SolveBlock helmholtz;
  helmholtz.id               = 1;
  helmholtz.solve_type       = Linear;
  helmholtz.solve_timing     = Secondary;
  helmholtz.field_indices    = {1};
  helmholtz.dependencies_rhs = make_dependency_set(fields, {"c"});
  helmholtz.dependencies_lhs = make_dependency_set(fields, {"c", "lhs(v)", "grad(lhs(v))"});

  SolveBlock post;
  post.id               = 2;
  post.solve_type       = Explicit;
  post.solve_timing     = PostProcess;
  post.field_indices    = {2};
  post.dependencies_rhs = make_dependency_set(fields, {"grad(c)"});

  blocks = {diffusion, helmholtz, post};
}
