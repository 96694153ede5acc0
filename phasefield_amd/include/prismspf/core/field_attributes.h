// Attributes of a field: name, rank and boundary conditions
// (reference: include/prismspf/core/field_attributes.h:25-101,
// include/prismspf/user_inputs/constraint_parameters.h).
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/type_enums.h>
#include <prismspf/core/types.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

PRISMS_PF_BEGIN_NAMESPACE

/// Boundary condition kinds (reference: constraint_parameters.h, enum Condition).
enum class Condition
{
  Natural,
  Dirichlet,
  Neumann,
  Periodic
};

inline Condition
condition_from_string(const std::string &s)
{
  if (s == "Natural" || s == "NATURAL" || s == "natural")
    return Condition::Natural;
  if (s == "Dirichlet" || s == "DIRICHLET" || s == "dirichlet")
    return Condition::Dirichlet;
  if (s == "Neumann" || s == "NEUMANN" || s == "neumann")
    return Condition::Neumann;
  if (s == "Periodic" || s == "PERIODIC" || s == "periodic")
    return Condition::Periodic;
  throw std::invalid_argument("Unknown boundary condition: " + s);
}

/// Conditions per boundary id (0..2*dim-1, deal.II colorized ids: x-, x+, y-, y+, z-, z+).
using ComponentConditions = std::map<unsigned int, Condition>;

struct BoundaryConditionSet
{
  std::map<unsigned int, ComponentConditions> component_constraints;
  bool                                        time_dependent = false;
};

struct FieldAttributes
{
  explicit FieldAttributes(std::string               _name                        = "",
                           TensorRank                _field_type                  = Scalar,
                           BoundaryConditionSet      _boundary_conditions         = {},
                           bool                      _is_nucleation_rate_variable = false,
                           std::vector<Types::Index> _nucleating_field_indices    = {})
    : name(std::move(_name))
    , field_type(_field_type)
    , boundary_conditions(std::move(_boundary_conditions))
    , is_nucleation_rate_variable(_is_nucleation_rate_variable)
    , nucleating_field_indices(std::move(_nucleating_field_indices))
  {}

  std::string               name;
  TensorRank                field_type = Scalar;
  BoundaryConditionSet      boundary_conditions;
  bool                      is_nucleation_rate_variable = false;
  std::vector<Types::Index> nucleating_field_indices;
};

inline std::map<std::string, Types::Index>
field_index_map(const std::vector<FieldAttributes> &fields)
{
  std::map<std::string, Types::Index> map;
  for (Types::Index i = 0; i < fields.size(); ++i)
    {
      if (map.count(fields[i].name))
        throw std::invalid_argument("The names of the fields must be unique: " +
                                    fields[i].name);
      map[fields[i].name] = i;
    }
  return map;
}

PRISMS_PF_END_NAMESPACE
