// Dependency declarations of solve blocks: which fields, at which ages, with values
// and/or gradients, an equation hook may read.
//
// Semantics follow the reference (include/prismspf/core/dependencies.h:30-213): a
// dependency string is an optional operator (`grad(`, `hess(`, `hessdiag(`, `lap(`,
// `div(`, `symgrad(`, `curl(`) around an optional state (`src(`, `change(`, `trial(`,
// `lhs(`, `old_k(` with k = 1..max_checked_age) around a field name.  Unknown strings
// are an error.
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/types.h>

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

PRISMS_PF_BEGIN_NAMESPACE

/// Evaluation flags of one field: for its current value, for the linear-operator
/// operand (src), and for each old age (index 0 = old_1).
struct Dependency
{
  EvalFlags              flag     = EvalFlags::nothing;
  EvalFlags              src_flag = EvalFlags::nothing;
  std::vector<EvalFlags> old_flags;

  explicit Dependency(EvalFlags                     _flag      = EvalFlags::nothing,
                      EvalFlags                     _src_flag  = EvalFlags::nothing,
                      const std::vector<EvalFlags> &_old_flags = {})
    : flag(_flag)
    , src_flag(_src_flag)
    , old_flags(_old_flags)
  {}

  Dependency
  operator|(const Dependency &other) const
  {
    Dependency r(flag | other.flag, src_flag | other.src_flag);
    const std::size_t n = std::max(old_flags.size(), other.old_flags.size());
    r.old_flags.assign(n, EvalFlags::nothing);
    for (std::size_t i = 0; i < n; ++i)
      {
        if (i < old_flags.size())
          r.old_flags[i] |= old_flags[i];
        if (i < other.old_flags.size())
          r.old_flags[i] |= other.old_flags[i];
      }
    return r;
  }

  Dependency &
  operator|=(const Dependency &other)
  {
    *this = *this | other;
    return *this;
  }

  /// Flags of a dependency type (Current, OldOne.., SRC).
  [[nodiscard]] EvalFlags
  flags_of(DependencyType type) const
  {
    if (type == SRC)
      return src_flag;
    if (type == Current)
      return flag;
    const unsigned int age = static_cast<unsigned int>(type);
    return age <= old_flags.size() ? old_flags[age - 1] : EvalFlags::nothing;
  }

  /// Oldest age read (0 if no old state is read).
  [[nodiscard]] unsigned int
  oldest_age() const
  {
    for (std::size_t i = old_flags.size(); i > 0; --i)
      if (old_flags[i - 1] != EvalFlags::nothing)
        return static_cast<unsigned int>(i);
    return 0;
  }
};

using DependencyMap = std::map<Types::Index, Dependency>;

constexpr unsigned int default_max_checked_age = 6;

namespace internal
{
  /// Strips `prefix` ... `)` around `s`; returns false if `s` does not have that shape.
  inline bool
  strip_wrapper(std::string &s, const std::string &prefix)
  {
    if (s.size() > prefix.size() + 1 && s.compare(0, prefix.size(), prefix) == 0 &&
        s.back() == ')')
      {
        s = s.substr(prefix.size(), s.size() - prefix.size() - 1);
        return true;
      }
    return false;
  }
} // namespace internal

/// Builds the dependency map of a solve block from dependency strings.
inline DependencyMap
make_dependency_set(const std::vector<FieldAttributes> &field_attributes,
                    std::set<std::string>               dependency_strings,
                    unsigned int max_checked_age = default_max_checked_age)
{
  static const std::vector<std::pair<std::string, EvalFlags>> operators = {
    {"grad(",     EvalFlags::gradients},
    {"hessdiag(", EvalFlags::hessians },
    {"hess(",     EvalFlags::hessians },
    {"lap(",      EvalFlags::hessians },
    {"div(",      EvalFlags::gradients},
    {"symgrad(",  EvalFlags::gradients},
    {"curl(",     EvalFlags::gradients},
  };
  static const std::vector<std::string> src_states = {"src(", "change(", "trial(", "lhs("};

  std::map<std::string, Types::Index> index_of;
  for (Types::Index i = 0; i < field_attributes.size(); ++i)
    index_of[field_attributes[i].name] = i;

  DependencyMap            result;
  std::vector<std::string> unknown;
  for (const std::string &original : dependency_strings)
    {
      std::string s    = original;
      EvalFlags   flag = EvalFlags::values;
      for (const auto &[prefix, op_flag] : operators)
        if (internal::strip_wrapper(s, prefix))
          {
            flag = op_flag;
            break;
          }

      enum class State
      {
        current,
        src,
        old
      } state              = State::current;
      unsigned int old_age = 0;
      for (const std::string &prefix : src_states)
        if (internal::strip_wrapper(s, prefix))
          {
            state = State::src;
            break;
          }
      if (state == State::current)
        for (unsigned int age = 1; age <= max_checked_age; ++age)
          if (internal::strip_wrapper(s, "old_" + std::to_string(age) + "("))
            {
              state   = State::old;
              old_age = age;
              break;
            }

      const auto it = index_of.find(s);
      if (it == index_of.end())
        {
          unknown.push_back(original);
          continue;
        }
      Dependency &dep = result[it->second];
      if (state == State::current)
        dep.flag |= flag;
      else if (state == State::src)
        dep.src_flag |= flag;
      else
        {
          if (dep.old_flags.size() < old_age)
            dep.old_flags.resize(old_age, EvalFlags::nothing);
          dep.old_flags[old_age - 1] |= flag;
        }
    }
  if (!unknown.empty())
    {
      std::string msg = "The following dependencies were not recognized: ";
      for (std::size_t i = 0; i < unknown.size(); ++i)
        msg += (i ? ", " : "") + unknown[i];
      throw std::invalid_argument(msg);
    }
  return result;
}

PRISMS_PF_END_NAMESPACE
