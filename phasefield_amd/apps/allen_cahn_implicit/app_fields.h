// Fields and solve blocks of the implicit Allen-Cahn application (the body of the
// reference's applications/allen_cahn/implicit/main.cc:24-46): a Newton block for n
// and a post-processing block for |grad n| and the free-energy density.
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
allen_cahn_implicit_define(std::vector<prismspf::FieldAttributes> &fields,
                           std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("n", Scalar), FieldAttributes("mg_n", Scalar),
            FieldAttributes("f_tot", Scalar)};

  SolveBlock newton_block;
  newton_block.id               = 1;
  newton_block.solve_type       = Newton;
  newton_block.solve_timing     = Primary;
  newton_block.field_indices    = {0};
  newton_block.dependencies_rhs = make_dependency_set(fields, {"n", "grad(n)", "old_1(n)"});
  newton_block.dependencies_lhs =
    make_dependency_set(fields, {"n", "change(n)", "grad(change(n))"});

  SolveBlock pp_block;
  pp_block.id               = 2;
  pp_block.solve_type       = Explicit;
  pp_block.solve_timing     = PostProcess;
  pp_block.field_indices    = {1, 2};
  pp_block.dependencies_rhs = make_dependency_set(fields, {"n", "grad(n)"});

  blocks = {newton_block, pp_block};
}
