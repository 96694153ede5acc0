// Fields and solve blocks of the routing test bed (four explicit diffusion blocks).
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/solve_block.h>

#include <set>
#include <string>
#include <vector>

inline void
switching_define(std::vector<prismspf::FieldAttributes> &fields,
                 std::vector<prismspf::SolveBlock>      &blocks)
{
  using namespace prismspf;
  fields = {FieldAttributes("c"), FieldAttributes("s"), FieldAttributes("r"),
            FieldAttributes("u")};
  auto block = [&](Types::Index id, std::set<std::string> deps) {
    SolveBlock b;
    b.id               = id;
    b.solve_type       = Explicit;
    b.solve_timing     = Primary;
    b.field_indices    = {id};
    b.dependencies_rhs = make_dependency_set(fields, deps);
    return b;
  };
  blocks = {block(0, {"old_1(c)", "grad(old_1(c))"}),
            block(1, {"old_1(c)", "old_1(s)", "grad(old_1(s))"}),
            block(2, {"old_1(r)", "grad(old_1(r))"}),
            block(3, {"old_1(c)", "old_1(u)", "grad(old_1(u))"})};
}
