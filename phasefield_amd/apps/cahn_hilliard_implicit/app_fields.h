// Fields and solve blocks of the implicit Cahn-Hilliard application (the body of the
// reference's applications/cahn_hilliard/implicit/main.cc:25-58): one Newton block for
// (c, mu) and a post-processing block for the free-energy density.
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
cahn_hilliard_implicit_define(std::vector<prismspf::FieldAttributes> &fields,
                              std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("c"), FieldAttributes("mu"), FieldAttributes("f_tot")};

  SolveBlock c_block;
  c_block.id            = 0;
  c_block.solve_type    = Newton;
  c_block.solve_timing  = Initialized;
  c_block.field_indices = {0, 1};
  c_block.dependencies_rhs =
    make_dependency_set(fields, {"c", "grad(c)", "mu", "grad(mu)", "old_1(c)"});
  c_block.dependencies_lhs = make_dependency_set(
    fields,
    {"c", "grad(c)", "mu", "grad(mu)", "change(c)", "grad(change(c))", "change(mu)",
     "grad(change(mu))"});

  SolveBlock pp_block;
  pp_block.id               = 2;
  pp_block.solve_type       = Explicit;
  pp_block.solve_timing     = PostProcess;
  pp_block.field_indices    = {2};
  pp_block.dependencies_rhs = make_dependency_set(fields, {"c", "grad(c)"});

  blocks = {c_block, pp_block};
}
