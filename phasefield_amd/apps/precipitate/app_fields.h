// Fields and solve blocks of the precipitate application (block structure of the
// reference's applications/coupled_allen_cahn_cahn_hilliard/main.cc, split into one
// block per field as the benchmark configuration describes).
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <vector>

inline void
precipitate_define(std::vector<prismspf::FieldAttributes> &fields,
                   std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("c"), FieldAttributes("mu"), FieldAttributes("eta")};

  SolveBlock eta_block;
  eta_block.id               = 0;
  eta_block.solve_type       = Explicit;
  eta_block.solve_timing     = Primary;
  eta_block.field_indices    = {2};
  eta_block.dependencies_rhs =
    make_dependency_set(fields, {"old_1(eta)", "grad(old_1(eta))", "old_1(c)"});

  SolveBlock c_block;
  c_block.id               = 1;
  c_block.solve_type       = Explicit;
  c_block.solve_timing     = Primary;
  c_block.field_indices    = {0};
  c_block.dependencies_rhs = make_dependency_set(fields, {"old_1(c)", "grad(old_1(mu))"});

  SolveBlock mu_block;
  mu_block.id               = 2;
  mu_block.solve_type       = Explicit;
  mu_block.solve_timing     = Secondary;
  mu_block.field_indices    = {1};
  mu_block.dependencies_rhs = make_dependency_set(fields, {"c", "eta", "grad(c)"});

  blocks = {eta_block, c_block, mu_block};
}
