// Fields and solve blocks of the Cahn-Hilliard application (the body of the reference
// main.cc, applications/cahn_hilliard/explicit/main.cc:29-50).
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
cahn_hilliard_define(std::vector<prismspf::FieldAttributes> &fields,
                     std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("c"), FieldAttributes("mu"), FieldAttributes("f_tot")};

  SolveBlock c_block;
  c_block.id               = 0;
  c_block.solve_type       = Explicit;
  c_block.solve_timing     = Initialized;
  c_block.field_indices    = {0};
  c_block.dependencies_rhs = make_dependency_set(fields, {"old_1(c)", "grad(old_1(mu))"});

  SolveBlock mu_block;
  mu_block.id               = 1;
  mu_block.solve_type       = Explicit;
  mu_block.solve_timing     = Uninitialized;
  mu_block.field_indices    = {1};
  mu_block.dependencies_rhs = make_dependency_set(fields, {"c", "grad(c)"});

  SolveBlock pp_block;
  pp_block.id               = 2;
  pp_block.solve_type       = Explicit;
  pp_block.solve_timing     = PostProcess;
  pp_block.field_indices    = {2};
  pp_block.dependencies_rhs = make_dependency_set(fields, {"c", "grad(c)"});

  blocks = {c_block, mu_block, pp_block};
}
