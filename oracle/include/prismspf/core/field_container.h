// Oracle FieldContainer: the equation hook's view of one quadrature point, filled by the
// oracle's cell loop from FEEvaluation-style sum factorization
// (reference: include/prismspf/core/field_container.h:45-553, 1065-1314).
// Reads of undeclared dependencies and submissions for fields outside the block are
// recorded so the cell loop can reject them.  TEST INFRASTRUCTURE.
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/lane_types.h>
#include <prismspf/core/type_enums.h>
#include <prismspf/core/types.h>

#include <stdexcept>
#include <type_traits>

PRISMS_PF_BEGIN_NAMESPACE

template <unsigned int dim, unsigned int degree, typename number>
class FieldContainer
{
public:
  using ScalarValue = dealii::VectorizedArray<number>;
  using VectorValue = dealii::Tensor<1, dim, ScalarValue>;
  template <TensorRank Rank>
  using Value = std::conditional_t<Rank == Vector, VectorValue, ScalarValue>;
  template <TensorRank Rank>
  using Gradient = std::conditional_t<Rank == Vector, dealii::Tensor<2, dim, ScalarValue>,
                                      dealii::Tensor<1, dim, ScalarValue>>;

  static constexpr unsigned int n_q_points =
    dim == 2 ? (degree + 1) * (degree + 1) : (degree + 1) * (degree + 1) * (degree + 1);

  static int
  slot_of(DependencyType t)
  {
    if (t == DST)
      throw std::logic_error("the destination of a solve cannot be read");
    return int(t) + 1;
  }

  // A vector field occupies dim consecutive stored fields, component c at f + c.
  template <TensorRank Rank, DependencyType type>
  Value<Rank>
  get_value(Types::Index f) const
  {
    return get_value<Rank>(f, type);
  }

  template <TensorRank Rank>
  Value<Rank>
  get_value(Types::Index f, DependencyType type) const
  {
    if constexpr (Rank == Vector)
      {
        VectorValue v;
        for (unsigned int c = 0; c < dim; ++c)
          v[c] = get_value<Scalar>(f + c, type);
        return v;
      }
    else
      return component_value(f, type);
  }

  ScalarValue
  component_value(Types::Index f, DependencyType type) const
  {
    const int s = slot_of(type);
    if (!(avail_values[f][s]))
      throw std::logic_error("value of an undeclared dependency was read");
    return values[f][s];
  }

  template <TensorRank Rank, DependencyType type>
  Gradient<Rank>
  get_gradient(Types::Index f) const
  {
    return get_gradient<Rank>(f, type);
  }

  template <TensorRank Rank>
  Gradient<Rank>
  get_gradient(Types::Index f, DependencyType type) const
  {
    if constexpr (Rank == Vector)
      {
        Gradient<Rank> g; // g[c][d] = d u_c / d x_d
        for (unsigned int c = 0; c < dim; ++c)
          g[c] = get_gradient<Scalar>(f + c, type);
        return g;
      }
    else
      {
        const int s = slot_of(type);
        if (!(avail_gradients[f][s]))
          throw std::logic_error("gradient of an undeclared dependency was read");
        Gradient<Rank> g;
        for (unsigned int d = 0; d < dim; ++d)
          g[d] = gradients[f][s][d];
        return g;
      }
  }

  template <TensorRank Rank, DependencyType type>
  Gradient<Rank>
  get_symmetric_gradient(Types::Index f) const
  {
    static_assert(Rank == Vector, "get_symmetric_gradient needs a vector field");
    const Gradient<Rank> g = get_gradient<Rank>(f, type);
    Gradient<Rank>       e;
    for (unsigned int i = 0; i < dim; ++i)
      for (unsigned int j = 0; j < dim; ++j)
        e[i][j] = ScalarValue(0.5 * (g[i][j].v + g[j][i].v));
    return e;
  }

  dealii::Point<dim, ScalarValue>
  get_q_point_location() const
  {
    dealii::Point<dim, ScalarValue> p;
    for (unsigned int d = 0; d < dim; ++d)
      p[d] = q_point[d];
    return p;
  }

  ScalarValue
  get_element_volume() const
  {
    return element_volume;
  }

  unsigned int
  get_n_q_points() const
  {
    return n_q_points;
  }

  void
  set_q_point(unsigned int)
  {}

  template <typename ValType>
  void
  set_value_term(Types::Index f, const ValType &val)
  {
    if (!(writable & (1u << f)))
      throw std::logic_error("submission for a field not part of this solve block");
    if constexpr (dealii::is_tensor<ValType>::value)
      {
        for (unsigned int c = 0; c < dim; ++c)
          set_value_term(f + c, val[c]);
      }
    else if constexpr (dealii::is_va<ValType>::value)
      value_term[f] = val.v;
    else
      value_term[f] = number(val);
  }

  template <typename GradType>
  void
  set_gradient_term(Types::Index f, const GradType &val)
  {
    if constexpr (GradType::rank == 2)
      {
        for (unsigned int c = 0; c < dim; ++c)
          set_gradient_term(f + c, val[c]);
      }
    else
      set_component_gradient_term(f, val);
  }

  template <typename GradType>
  void
  set_component_gradient_term(Types::Index f, const GradType &val)
  {
    if (!(writable & (1u << f)))
      throw std::logic_error("submission for a field not part of this solve block");
    for (unsigned int d = 0; d < dim; ++d)
      {
        if constexpr (dealii::is_va<typename GradType::value_type>::value)
          gradient_term[f][d] = val[d].v;
        else
          gradient_term[f][d] = number(val[d]);
      }
  }

  // Filled by the oracle's cell loop.
  number       values[PPF_MAX_FIELDS][PPF_N_SLOTS] {};
  number       gradients[PPF_MAX_FIELDS][PPF_N_SLOTS][dim] {};
  bool         avail_values[PPF_MAX_FIELDS][PPF_N_SLOTS] {};
  bool         avail_gradients[PPF_MAX_FIELDS][PPF_N_SLOTS] {};
  number       q_point[dim] {};
  number       element_volume = 0;
  unsigned int writable       = 0;
  number       value_term[PPF_MAX_FIELDS] {};
  number       gradient_term[PPF_MAX_FIELDS][dim] {};
};

PRISMS_PF_END_NAMESPACE
