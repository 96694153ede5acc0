// Solve blocks: a set of fields advanced together, their solve type and timing, and
// the dependencies of their right- and left-hand sides
// (reference: include/prismspf/core/solve_block.h:28-341).
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/dependencies.h>
#include <prismspf/core/type_enums.h>
#include <prismspf/core/types.h>
#include <prismspf/user_inputs/solve_parameters.h>

#include <set>
#include <stdexcept>
#include <string>
#include <vector>

PRISMS_PF_BEGIN_NAMESPACE

/// When a block's fields are produced (reference solve_block.h:28-52).
enum SolveTiming
{
  Primary,
  Initialized = Primary,
  Secondary,
  Uninitialized = Secondary,
  PostProcess,
  NucleationRate
};

class SolveBlock
{
public:
  explicit SolveBlock(int                    _id               = -1,
                      SolveType              _solve_type       = Explicit,
                      SolveTiming            _solve_timing     = Primary,
                      std::set<Types::Index> _field_indices    = {},
                      DependencyMap          _dependencies_rhs = {},
                      DependencyMap          _dependencies_lhs = {})
    : id(_id)
    , solve_type(_solve_type)
    , solve_timing(_solve_timing)
    , field_indices(std::move(_field_indices))
    , dependencies_rhs(std::move(_dependencies_rhs))
    , dependencies_lhs(std::move(_dependencies_lhs))
  {}

  int                       id;
  SolveType                 solve_type;
  SolveTiming               solve_timing;
  std::set<Types::Index>    field_indices;
  DependencyMap             dependencies_rhs;
  DependencyMap             dependencies_lhs;
  LinearSolverParameters    linear_solver_parameters;
  NonlinearSolverParameters nonlinear_solver_parameters;

  bool
  operator<(const SolveBlock &other) const
  {
    return id < other.id;
  }

  /// Consistency rules of one block (reference solve_block.h:130-230).
  void
  validate() const
  {
    const std::string where = "solve block " + std::to_string(id) + ": ";
    if (field_indices.empty())
      throw std::invalid_argument(where + "solve blocks must manage at least 1 field");
    if (solve_type != Newton)
      for (Types::Index f : field_indices)
        {
          const auto it = dependencies_rhs.find(f);
          if (it != dependencies_rhs.end() && it->second.flag != EvalFlags::nothing)
            throw std::invalid_argument(
              where + "the current value of a field should never appear in the RHS of "
                      "a solve that is not type Newton");
        }
    if (solve_type == Linear)
      for (Types::Index f : field_indices)
        {
          const auto it = dependencies_lhs.find(f);
          if (it == dependencies_lhs.end() || it->second.src_flag == EvalFlags::nothing)
            throw std::invalid_argument(
              where + "every field in a linear solve should appear in the (LHS) "
                      "expression; be sure to use the src_flag");
        }
    if (solve_type == Explicit && !dependencies_lhs.empty())
      throw std::invalid_argument(where + "explicit solves should have no LHS dependencies");
    if (solve_type == Constant && (!dependencies_rhs.empty() || !dependencies_lhs.empty()))
      throw std::invalid_argument(where + "constant solves should have no dependencies");
    for (const auto &[f, dep] : dependencies_rhs)
      if (dep.src_flag != EvalFlags::nothing)
        throw std::invalid_argument(where +
                                    "trial/change terms should not appear in RHS expressions");
  }
};

/// Cross-block rules (reference solve_block.h:232-339): unique ids and field names,
/// every field in exactly one block, dependencies in range, and no block reads the
/// current value of a field that an earlier block has not produced yet.
inline const std::vector<SolveBlock> &
validate_solve_blocks(const std::vector<SolveBlock>      &solve_blocks,
                      const std::vector<FieldAttributes> &field_attributes)
{
  field_index_map(field_attributes);
  std::set<int> ids;
  for (const SolveBlock &b : solve_blocks)
    {
      b.validate();
      if (!ids.insert(b.id).second)
        throw std::invalid_argument("Each solve block must have a unique id");
    }
  std::vector<int> owner(field_attributes.size(), -1);
  for (const SolveBlock &b : solve_blocks)
    for (Types::Index f : b.field_indices)
      {
        if (f >= field_attributes.size() || owner[f] != -1)
          throw std::invalid_argument("Each field should be managed by exactly one "
                                      "solve block (field " + std::to_string(f) + ")");
        owner[f] = b.id;
      }
  for (Types::Index f = 0; f < field_attributes.size(); ++f)
    if (owner[f] == -1)
      throw std::invalid_argument("Field " + field_attributes[f].name +
                                  " is not managed by any solve block");
  std::set<Types::Index> solved;
  for (const SolveBlock &b : solve_blocks)
    {
      for (const DependencyMap *deps : {&b.dependencies_rhs, &b.dependencies_lhs})
        for (const auto &[f, dep] : *deps)
          {
            if (f >= field_attributes.size())
              throw std::invalid_argument("Dependencies of solve block " +
                                          std::to_string(b.id) +
                                          " refer to a field index that is out of range");
            if (dep.flag != EvalFlags::nothing && b.solve_type != Newton && !solved.count(f))
              throw std::invalid_argument(
                "Solve block " + std::to_string(b.id) +
                " depends on the current value of field \"" + field_attributes[f].name +
                "\" which is not solved in a previous solve block");
          }
      solved.insert(b.field_indices.begin(), b.field_indices.end());
    }
  return solve_blocks;
}

/// Declared field and component of a stored (scalar) field.
struct StoredComponent
{
  Types::Index declared  = 0;
  unsigned int component = 0;
  bool         vector    = false;
};

/// Vector fields (TensorRank::Vector, reference field_attributes.h:25-101) are stored as
/// `dim` scalar component fields at consecutive indices: field k of rank Vector becomes
/// the stored fields k .. k+dim-1, named <name>_x, <name>_y, <name>_z, each carrying the
/// boundary conditions of its component (`u:x` in the parameter file).  Only one vector
/// field, declared last, is supported, so that the declared index of every field is also
/// the stored index of its first component (the index the equations use).  Solve blocks
/// and dependency sets that name the vector field are widened to all its components.
/// Returns the declared field and component of every stored field.
inline std::vector<StoredComponent>
expand_vector_fields(std::vector<FieldAttributes> &fields, std::vector<SolveBlock> &blocks,
                     unsigned int dim)
{
  std::vector<StoredComponent> layout;
  for (Types::Index f = 0; f < fields.size(); ++f)
    if (fields[f].field_type == Vector)
      {
        if (f + 1 != fields.size())
          throw std::invalid_argument("field " + fields[f].name +
                                      ": a vector field must be the last declared field");
      }
    else if (fields[f].field_type != Scalar)
      throw std::invalid_argument("field " + fields[f].name + ": rank-2 fields are not "
                                  "supported");
  for (Types::Index f = 0; f < fields.size(); ++f)
    layout.push_back({f, 0u, false});
  if (fields.empty() || fields.back().field_type != Vector)
    return layout;
  const Types::Index   v    = fields.size() - 1;
  const FieldAttributes base = fields.back();
  static const char   *axes = "xyz";
  fields.pop_back();
  layout.pop_back();
  for (unsigned int c = 0; c < dim; ++c)
    {
      FieldAttributes fc = base;
      fc.name            = base.name + "_" + axes[c];
      fc.field_type      = Scalar;
      fc.boundary_conditions.component_constraints.clear();
      const auto it = base.boundary_conditions.component_constraints.find(c);
      if (it != base.boundary_conditions.component_constraints.end())
        fc.boundary_conditions.component_constraints[0] = it->second;
      fields.push_back(fc);
      layout.push_back({v, c, true});
    }
  for (SolveBlock &b : blocks)
    {
      if (b.field_indices.count(v))
        for (unsigned int c = 1; c < dim; ++c)
          b.field_indices.insert(v + c);
      for (DependencyMap *deps : {&b.dependencies_rhs, &b.dependencies_lhs})
        {
          const auto it = deps->find(v);
          if (it != deps->end())
            for (unsigned int c = 1; c < dim; ++c)
              (*deps)[v + c] = it->second;
        }
    }
  return layout;
}

PRISMS_PF_END_NAMESPACE
