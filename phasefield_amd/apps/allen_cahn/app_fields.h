// Fields and solve blocks of the Allen-Cahn application (the body of a reference
// main.cc, cf. applications/allen_cahn/explicit/main.cc:27-45).
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
allen_cahn_define(std::vector<prismspf::FieldAttributes> &fields,
                  std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("n", Scalar), FieldAttributes("f_tot", Scalar)};

  SolveBlock exp_block;
  exp_block.id               = 1;
  exp_block.solve_type       = Explicit;
  exp_block.solve_timing     = Primary;
  exp_block.field_indices    = {0};
  exp_block.dependencies_rhs = make_dependency_set(fields, {"old_1(n)", "grad(old_1(n))"});

  SolveBlock pp_block;
  pp_block.id               = 2;
  pp_block.solve_type       = Explicit;
  pp_block.solve_timing     = PostProcess;
  pp_block.field_indices    = {1};
  pp_block.dependencies_rhs = make_dependency_set(fields, {"n", "grad(n)"});

  blocks = {exp_block, pp_block};
}
