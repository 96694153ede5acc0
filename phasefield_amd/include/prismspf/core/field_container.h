// FieldContainer: what an equation hook sees at one quadrature point.
//
// The reference wraps one deal.II FEEvaluation per (field, state) and serves
// get_value/get_gradient/set_value_term/set_gradient_term from it
// (reference: include/prismspf/core/field_container.h:45-553, 1065-1314).  On the
// MI355X path a kernel either fills the point view (values and gradients at one node,
// used by the node kernels) or the cell view (the cell's nodal values, differentiated
// on demand at the current quadrature point), calls the inlined hook, and reads the
// submitted terms back.  All field and slot indices seen by the hook are compile-time
// constants after inlining, so everything lives in registers.
//
// The same class is used on the host to probe the hook (structure detection and
// dependency validation); the `read_*` / `*_term_mask` bookkeeping serves that probe and
// is dead code in the kernels.
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/gll_basis.h>
#include <prismspf/core/lane_types.h>
#include <prismspf/core/type_enums.h>
#include <prismspf/core/types.h>

#include <type_traits>

PRISMS_PF_BEGIN_NAMESPACE

template <unsigned int dim, unsigned int degree, typename number>
class FieldContainer
{
  template <typename T>
  struct dependent_false : std::false_type
  {};

public:
  using ScalarValue = dealii::VectorizedArray<number>;
  using VectorValue = dealii::Tensor<1, dim, ScalarValue>;

  template <TensorRank Rank>
  using Value = std::conditional_t<Rank == Scalar || dim == 1,
                                   ScalarValue,
                                   dealii::Tensor<int(Rank), dim, ScalarValue>>;

  template <TensorRank Rank>
  using Gradient = dealii::Tensor<int(Rank) + 1, dim, ScalarValue>;

  static constexpr unsigned int dofs_per_component =
    dim == 1 ? degree + 1 : (dim == 2 ? (degree + 1) * (degree + 1) :
                                        (degree + 1) * (degree + 1) * (degree + 1));
  static constexpr unsigned int n_q_points = dofs_per_component;

  // ------------------------------------------------------------------ accessors
  template <TensorRank Rank, DependencyType type>
  [[nodiscard]] PPF_HD PPF_INLINE Value<Rank>
  get_value(Types::Index field_index) const
  {
    static_assert(Rank == Scalar || Rank == Vector, "Unknown field rank");
    static_assert(type != DST, "The destination of a solve cannot be read");
    return get_value<Rank>(field_index, type);
  }

  // A vector field is stored as dim component fields starting at its own index
  // (field_index + c holds component c), so its accessors gather the components.
  template <TensorRank Rank>
  [[nodiscard]] PPF_HD PPF_INLINE Value<Rank>
  get_value(Types::Index field_index, DependencyType type) const
  {
    static_assert(Rank == Scalar || Rank == Vector, "Unknown field rank");
    if constexpr (Rank == Vector && dim > 1)
      {
        Value<Rank> v;
        for (unsigned int c = 0; c < dim; ++c)
          v[c] = get_value<Scalar>(field_index + c, type);
        return v;
      }
    else
      return scalar_value(field_index, type);
  }

  template <TensorRank Rank, DependencyType type>
  [[nodiscard]] PPF_HD PPF_INLINE Gradient<Rank>
  get_gradient(Types::Index field_index) const
  {
    static_assert(Rank == Scalar || Rank == Vector, "Unknown field rank");
    static_assert(type != DST, "The destination of a solve cannot be read");
    return get_gradient<Rank>(field_index, type);
  }

  /// Gradient; for a vector field g[c][d] = d u_c / d x_d.
  template <TensorRank Rank>
  [[nodiscard]] PPF_HD PPF_INLINE Gradient<Rank>
  get_gradient(Types::Index field_index, DependencyType type) const
  {
    static_assert(Rank == Scalar || Rank == Vector, "Unknown field rank");
    if constexpr (Rank == Vector)
      {
        Gradient<Rank> g;
        for (unsigned int c = 0; c < dim; ++c)
          g[c] = scalar_gradient(field_index + c, type);
        return g;
      }
    else
      return scalar_gradient(field_index, type);
  }

  /// Symmetric part of a vector field's gradient, (g + g^T) / 2 (reference
  /// field_container.h get_symmetric_gradient).
  template <TensorRank Rank, DependencyType type>
  [[nodiscard]] PPF_HD PPF_INLINE Gradient<Rank>
  get_symmetric_gradient(Types::Index field_index) const
  {
    static_assert(Rank == Vector, "get_symmetric_gradient needs a vector field");
    static_assert(type != DST, "The destination of a solve cannot be read");
    const Gradient<Rank> g = get_gradient<Rank>(field_index, type);
    Gradient<Rank>       e;
    for (unsigned int i = 0; i < dim; ++i)
      for (unsigned int j = 0; j < dim; ++j)
        e[i][j] = ScalarValue(number(0.5) * (g[i][j].data + g[j][i].data));
    return e;
  }

  [[nodiscard]] PPF_HD PPF_INLINE ScalarValue
  scalar_value(Types::Index field_index, DependencyType type) const
  {
    const int slot = dependency_slot(type);
    read_value_mask[slot] |= 1u << field_index;
    if (block_view)
      return ScalarValue(block_ptr[field_index][slot][block_offset(int(cell_q))]);
    if (cell_view)
      return ScalarValue(cell_values[field_index][slot][cell_q]);
    return ScalarValue(value[field_index][slot]);
  }

  [[nodiscard]] PPF_HD PPF_INLINE Gradient<Scalar>
  scalar_gradient(Types::Index field_index, DependencyType type) const
  {
    const int slot = dependency_slot(type);
    read_gradient_mask[slot] |= 1u << field_index;
    Gradient<Scalar> g;
    if (block_view)
      {
        // As the cell view, with the cell's nodes read from the staged node block.
        using Basis     = GLLBasis<degree>;
        const number *u = block_ptr[field_index][slot];
        for (unsigned int d = 0; d < dim; ++d)
          {
            const int qd  = local_coord<degree>(int(cell_q), int(d));
            number    sum = number(0);
            for (int k = 0; k <= int(degree); ++k)
              if (Basis::D[qd][k] != 0.0)
                sum += number(Basis::D[qd][k]) *
                       u[block_offset(with_coord<degree>(int(cell_q), int(d), k))];
            g[d] = ScalarValue(sum * inv_h[d]);
          }
        return g;
      }
    if (cell_view)
      {
        // Collocation: the gradient at node q of the cell is the 1-D derivative matrix
        // applied along each direction through q.
        using Basis     = GLLBasis<degree>;
        const number *u = cell_values[field_index][slot];
        for (unsigned int d = 0; d < dim; ++d)
          {
            const int qd  = local_coord<degree>(int(cell_q), int(d));
            number    sum = number(0);
            for (int k = 0; k <= int(degree); ++k)
              if (Basis::D[qd][k] != 0.0)
                sum += number(Basis::D[qd][k]) * u[with_coord<degree>(int(cell_q), int(d), k)];
            g[d] = ScalarValue(sum * inv_h[d]);
          }
        return g;
      }
    for (unsigned int d = 0; d < dim; ++d)
      g[d] = ScalarValue(gradient[field_index][slot][d]);
    return g;
  }

  // Second derivatives are not evaluated by the Q1/Q2 collocation kernels.  Using them
  // in an equation is a compile-time error rather than a silent zero.
  template <TensorRank Rank, DependencyType type>
  PPF_HD void
  get_hessian(Types::Index) const
  {
    static_assert(dependent_false<number>::value,
                  "get_hessian is not supported on the MI355X path");
  }

  template <TensorRank Rank, DependencyType type>
  PPF_HD void
  get_laplacian(Types::Index) const
  {
    static_assert(dependent_false<number>::value,
                  "get_laplacian is not supported on the MI355X path");
  }

  template <TensorRank Rank, DependencyType type>
  PPF_HD void
  get_hessian_diagonal(Types::Index) const
  {
    static_assert(dependent_false<number>::value,
                  "get_hessian_diagonal is not supported on the MI355X path");
  }

  /// Location of the current quadrature point (a mesh node: GLL collocation).
  [[nodiscard]] PPF_HD PPF_INLINE dealii::Point<dim, ScalarValue>
  get_q_point_location() const
  {
    geometry_reads |= 1u;
    dealii::Point<dim, ScalarValue> p;
    for (unsigned int d = 0; d < dim; ++d)
      p[d] = ScalarValue(q_point[d]);
    return p;
  }

  /// Volume of the current cell: JxW / w = prod(h) on the uniform Cartesian mesh
  /// (reference field_container.h:1244-1249).
  [[nodiscard]] PPF_HD PPF_INLINE ScalarValue
  get_element_volume() const
  {
    geometry_reads |= 2u;
    return ScalarValue(element_volume);
  }

  [[nodiscard]] PPF_HD PPF_INLINE unsigned int
  get_n_q_points() const
  {
    return n_q_points;
  }

  /// Quadrature points are walked by the kernel; kept for API compatibility.
  PPF_HD PPF_INLINE void
  set_q_point(unsigned int)
  {}

  // ---------------------------------------------------------------- submissions
  template <typename ValType>
  PPF_HD PPF_INLINE void
  set_value_term(Types::Index field_index, const ValType &val)
  {
    if constexpr (dealii::is_lane_value<ValType>::value)
      value_term[field_index] = static_cast<number>(val.data);
    else if constexpr (std::is_arithmetic_v<ValType>)
      value_term[field_index] = static_cast<number>(val);
    else if constexpr (dealii::is_tensor<ValType>::value && ValType::rank == 1)
      {
        // Vector field: one value term per component field.
        for (unsigned int c = 0; c < dim; ++c)
          set_value_term(field_index + c, val[c]);
        return;
      }
    else
      static_assert(dependent_false<ValType>::value,
                    "Only scalar and vector value terms are supported on the MI355X path");
    value_term_mask |= 1u << field_index;
  }

  template <typename GradType>
  PPF_HD PPF_INLINE void
  set_gradient_term(Types::Index field_index, const GradType &val)
  {
    static_assert(dealii::is_tensor<GradType>::value && (GradType::rank == 1 ||
                                                         GradType::rank == 2),
                  "Gradient terms are rank 1 (scalar field) or rank 2 (vector field)");
    if constexpr (GradType::rank == 2)
      {
        // Vector field: row c is the gradient term of component c.
        for (unsigned int c = 0; c < dim; ++c)
          set_gradient_term(field_index + c, val[c]);
        return;
      }
    else
      set_scalar_gradient_term(field_index, val);
  }

  template <typename GradType>
  PPF_HD PPF_INLINE void
  set_scalar_gradient_term(Types::Index field_index, const GradType &val)
  {
    for (unsigned int d = 0; d < dim; ++d)
      {
        if constexpr (dealii::is_lane_value<typename GradType::value_type>::value)
          gradient_term[field_index][d] = static_cast<number>(val[d].data);
        else
          gradient_term[field_index][d] = static_cast<number>(val[d]);
      }
    gradient_term_mask |= 1u << field_index;
  }

  // ------------------------------------------------------------ kernel interface
  /// Clears the submitted terms before the next quadrature point.
  PPF_HD PPF_INLINE void
  clear_terms()
  {
    for (unsigned int f = 0; f < PPF_MAX_FIELDS; ++f)
      {
        value_term[f] = number(0);
        for (unsigned int d = 0; d < dim; ++d)
          gradient_term[f][d] = number(0);
      }
    value_term_mask    = 0;
    gradient_term_mask = 0;
  }

  /// Clears the probe bookkeeping.
  PPF_HD PPF_INLINE void
  clear_reads()
  {
    for (unsigned int s = 0; s < PPF_N_SLOTS; ++s)
      {
        read_value_mask[s]    = 0;
        read_gradient_mask[s] = 0;
      }
    geometry_reads = 0;
  }

  /// Point view: values and gradients of every slot at the current point.
  number value[PPF_MAX_FIELDS][PPF_N_SLOTS];
  number gradient[PPF_MAX_FIELDS][PPF_N_SLOTS][dim];

  /// Cell view: with `cell_view` set, accessors read the cell's nodal values and
  /// differentiate them at local node `cell_q` on demand, so only what the hook reads
  /// is ever computed.  (A flag rather than a null test of a pointer keeps every
  /// access a compile-time constant once the kernel is inlined.)
  bool         cell_view = false;
  unsigned int cell_q    = 0;
  number       inv_h[dim];
  number       cell_values[PPF_MAX_FIELDS][PPF_N_SLOTS][dofs_per_component];

  /// Block view: the cell's nodal values are read from a staged block of nodes (a
  /// workgroup's shared memory): local node k of the cell is at
  /// block_ptr[f][s][block_base + kx + block_stride[0] ky + block_stride[1] kz].
  bool          block_view = false;
  const number *block_ptr[PPF_MAX_FIELDS][PPF_N_SLOTS];
  int           block_base      = 0;
  int           block_stride[2] = {0, 0};

  [[nodiscard]] PPF_HD PPF_INLINE int
  block_offset(int k) const
  {
    int o = block_base + local_coord<degree>(k, 0);
    if constexpr (dim > 1)
      o += block_stride[0] * local_coord<degree>(k, 1);
    if constexpr (dim > 2)
      o += block_stride[1] * local_coord<degree>(k, 2);
    return o;
  }

  number value_term[PPF_MAX_FIELDS];
  number gradient_term[PPF_MAX_FIELDS][dim];
  number q_point[dim];
  number element_volume;

  unsigned int          value_term_mask    = 0;
  unsigned int          gradient_term_mask = 0;
  mutable unsigned int  read_value_mask[PPF_N_SLOTS]    = {};
  mutable unsigned int  read_gradient_mask[PPF_N_SLOTS] = {};
  mutable unsigned int  geometry_reads                  = 0;
};

PRISMS_PF_END_NAMESPACE
