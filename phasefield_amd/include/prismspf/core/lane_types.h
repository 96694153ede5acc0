// Value types of the application API, one lane wide.
//
// PRISMS-PF application code spells its quadrature-point types with deal.II names:
// `dealii::VectorizedArray<number>`, `dealii::Tensor<1, dim, ScalarValue>` and
// `dealii::Point<dim>` (reference: applications/*/custom_pde.h, e.g.
// applications/allen_cahn/explicit/custom_pde.h:13-18).  On the CPU deal.II packs
// several cells into one SIMD register; on CDNA4 each lane of a wavefront owns its own
// quadrature point, so here a "vectorized array" holds exactly one number and the
// wavefront provides the vectorization.  Only the operations the application API uses
// are provided: arithmetic, comparisons through std::min/max, elementary functions,
// tensor contraction, norm and norm_square.
#pragma once

#include <prismspf/config.h>

#include <cmath>
#include <cstddef>
#include <type_traits>

namespace dealii
{
  template <typename Number, std::size_t width = 1>
  class VectorizedArray
  {
  public:
    using value_type = Number;

    static constexpr std::size_t
    size()
    {
      return 1;
    }

    VectorizedArray() = default;

    PPF_HD constexpr VectorizedArray(Number scalar)
      : data(scalar)
    {}

    template <typename Other,
              typename = std::enable_if_t<std::is_arithmetic_v<Other> &&
                                          !std::is_same_v<Other, Number>>>
    PPF_HD constexpr VectorizedArray(Other scalar)
      : data(static_cast<Number>(scalar))
    {}

    PPF_HD VectorizedArray &
    operator=(Number scalar)
    {
      data = scalar;
      return *this;
    }

    PPF_HD Number &
    operator[](unsigned int)
    {
      return data;
    }

    PPF_HD const Number &
    operator[](unsigned int) const
    {
      return data;
    }

    PPF_HD VectorizedArray &
    operator+=(const VectorizedArray &o)
    {
      data += o.data;
      return *this;
    }

    PPF_HD VectorizedArray &
    operator-=(const VectorizedArray &o)
    {
      data -= o.data;
      return *this;
    }

    PPF_HD VectorizedArray &
    operator*=(const VectorizedArray &o)
    {
      data *= o.data;
      return *this;
    }

    PPF_HD VectorizedArray &
    operator/=(const VectorizedArray &o)
    {
      data /= o.data;
      return *this;
    }

    PPF_HD Number
    sum() const
    {
      return data;
    }

    Number data;
  };

  template <typename T>
  struct is_lane_value : std::false_type
  {};
  template <typename N, std::size_t w>
  struct is_lane_value<VectorizedArray<N, w>> : std::true_type
  {};

#define PPF_LANE_BINARY(op)                                                         \
  template <typename N, std::size_t w>                                              \
  PPF_HD PPF_INLINE VectorizedArray<N, w> operator op(const VectorizedArray<N, w> &a, \
                                                      const VectorizedArray<N, w> &b) \
  {                                                                                 \
    return VectorizedArray<N, w>(a.data op b.data);                                 \
  }                                                                                 \
  template <typename N, std::size_t w, typename S,                                  \
            typename = std::enable_if_t<std::is_arithmetic_v<S>>>                   \
  PPF_HD PPF_INLINE VectorizedArray<N, w> operator op(const VectorizedArray<N, w> &a, \
                                                      const S                     &s) \
  {                                                                                 \
    return VectorizedArray<N, w>(a.data op static_cast<N>(s));                      \
  }                                                                                 \
  template <typename N, std::size_t w, typename S,                                  \
            typename = std::enable_if_t<std::is_arithmetic_v<S>>>                   \
  PPF_HD PPF_INLINE VectorizedArray<N, w> operator op(const S                     &s, \
                                                      const VectorizedArray<N, w> &a) \
  {                                                                                 \
    return VectorizedArray<N, w>(static_cast<N>(s) op a.data);                      \
  }

  PPF_LANE_BINARY(+)
  PPF_LANE_BINARY(-)
  PPF_LANE_BINARY(*)
  PPF_LANE_BINARY(/)
#undef PPF_LANE_BINARY

  template <typename N, std::size_t w>
  PPF_HD PPF_INLINE VectorizedArray<N, w>
  operator-(const VectorizedArray<N, w> &a)
  {
    return VectorizedArray<N, w>(-a.data);
  }

  template <typename N, std::size_t w>
  PPF_HD PPF_INLINE VectorizedArray<N, w>
  operator+(const VectorizedArray<N, w> &a)
  {
    return a;
  }

  template <typename Number>
  PPF_HD PPF_INLINE VectorizedArray<Number>
  make_vectorized_array(const Number &u)
  {
    return VectorizedArray<Number>(u);
  }

  // ------------------------------------------------------------------ tensors
  template <int rank, int dim, typename Number = double>
  class Tensor;

  template <typename T>
  struct is_tensor : std::false_type
  {};
  template <int r, int d, typename N>
  struct is_tensor<Tensor<r, d, N>> : std::true_type
  {};

  /// Rank-1 tensor (a gradient or a vector field value).  Default construction
  /// zero-initializes, as deal.II's Tensor does.
  template <int dim, typename Number>
  class Tensor<1, dim, Number>
  {
  public:
    static constexpr unsigned int rank                     = 1;
    static constexpr unsigned int dimension                = dim;
    static constexpr unsigned int n_independent_components = dim;
    using value_type                                       = Number;

    PPF_HD constexpr Tensor()
      : values {}
    {
      for (int d = 0; d < dim; ++d)
        values[d] = Number(0.0);
    }

    /// Conversion from another number type (as deal.II's converting constructor; e.g.
    /// a double stiffness constant held by an equation instantiated in float).
    template <typename Other, typename = std::enable_if_t<!std::is_same_v<Other, Number>>>
    PPF_HD constexpr Tensor(const Tensor<1, dim, Other> &o)
      : values {}
    {
      for (int d = 0; d < dim; ++d)
        values[d] = Number(o[d]);
    }

    PPF_HD Number &
    operator[](unsigned int d)
    {
      return values[d];
    }

    PPF_HD const Number &
    operator[](unsigned int d) const
    {
      return values[d];
    }

    PPF_HD Tensor &
    operator+=(const Tensor &o)
    {
      for (int d = 0; d < dim; ++d)
        values[d] += o.values[d];
      return *this;
    }

    PPF_HD Tensor &
    operator-=(const Tensor &o)
    {
      for (int d = 0; d < dim; ++d)
        values[d] -= o.values[d];
      return *this;
    }

    template <typename S>
    PPF_HD Tensor &
    operator*=(const S &s)
    {
      for (int d = 0; d < dim; ++d)
        values[d] = values[d] * s;
      return *this;
    }

    template <typename S>
    PPF_HD Tensor &
    operator/=(const S &s)
    {
      for (int d = 0; d < dim; ++d)
        values[d] = values[d] / s;
      return *this;
    }

    PPF_HD Number
    norm_square() const
    {
      Number s = values[0] * values[0];
      for (int d = 1; d < dim; ++d)
        s += values[d] * values[d];
      return s;
    }

    PPF_HD Number
    norm() const;

    Number values[dim];
  };

  /// Rank-2 tensor, stored as dim rank-1 rows.
  template <int dim, typename Number>
  class Tensor<2, dim, Number>
  {
  public:
    static constexpr unsigned int rank                     = 2;
    static constexpr unsigned int dimension                = dim;
    static constexpr unsigned int n_independent_components = dim * dim;
    using value_type                                       = Number;

    PPF_HD constexpr Tensor() = default;

    /// Conversion from another number type (see the rank-1 tensor).
    template <typename Other, typename = std::enable_if_t<!std::is_same_v<Other, Number>>>
    PPF_HD constexpr Tensor(const Tensor<2, dim, Other> &o)
    {
      for (int d = 0; d < dim; ++d)
        values[d] = Tensor<1, dim, Number>(o[d]);
    }

    PPF_HD Tensor<1, dim, Number> &
    operator[](unsigned int d)
    {
      return values[d];
    }

    PPF_HD const Tensor<1, dim, Number> &
    operator[](unsigned int d) const
    {
      return values[d];
    }

    PPF_HD Tensor &
    operator+=(const Tensor &o)
    {
      for (int d = 0; d < dim; ++d)
        values[d] += o.values[d];
      return *this;
    }

    PPF_HD Tensor &
    operator-=(const Tensor &o)
    {
      for (int d = 0; d < dim; ++d)
        values[d] -= o.values[d];
      return *this;
    }

    PPF_HD Number
    norm_square() const
    {
      Number s = values[0].norm_square();
      for (int d = 1; d < dim; ++d)
        s += values[d].norm_square();
      return s;
    }

    Tensor<1, dim, Number> values[dim];
  };

  template <int r, int dim, typename N>
  PPF_HD PPF_INLINE Tensor<r, dim, N>
  operator+(Tensor<r, dim, N> a, const Tensor<r, dim, N> &b)
  {
    a += b;
    return a;
  }

  template <int r, int dim, typename N>
  PPF_HD PPF_INLINE Tensor<r, dim, N>
  operator-(Tensor<r, dim, N> a, const Tensor<r, dim, N> &b)
  {
    a -= b;
    return a;
  }

  template <int dim, typename N>
  PPF_HD PPF_INLINE Tensor<1, dim, N>
  operator-(const Tensor<1, dim, N> &a)
  {
    Tensor<1, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = -a[d];
    return r;
  }

  template <int dim, typename N>
  PPF_HD PPF_INLINE Tensor<2, dim, N>
  operator-(const Tensor<2, dim, N> &a)
  {
    Tensor<2, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = -a[d];
    return r;
  }

  /// Tensor times scalar, where the scalar is a lane value or a plain number.
  template <int dim, typename N, typename S,
            typename = std::enable_if_t<!is_tensor<S>::value>>
  PPF_HD PPF_INLINE Tensor<1, dim, N>
  operator*(const Tensor<1, dim, N> &a, const S &s)
  {
    Tensor<1, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = a[d] * s;
    return r;
  }

  template <int dim, typename N, typename S,
            typename = std::enable_if_t<!is_tensor<S>::value>>
  PPF_HD PPF_INLINE Tensor<1, dim, N>
  operator*(const S &s, const Tensor<1, dim, N> &a)
  {
    Tensor<1, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = s * a[d];
    return r;
  }

  template <int dim, typename N, typename S,
            typename = std::enable_if_t<!is_tensor<S>::value>>
  PPF_HD PPF_INLINE Tensor<1, dim, N>
  operator/(const Tensor<1, dim, N> &a, const S &s)
  {
    Tensor<1, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = a[d] / s;
    return r;
  }

  template <int dim, typename N, typename S,
            typename = std::enable_if_t<!is_tensor<S>::value>>
  PPF_HD PPF_INLINE Tensor<2, dim, N>
  operator*(const Tensor<2, dim, N> &a, const S &s)
  {
    Tensor<2, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = a[d] * s;
    return r;
  }

  template <int dim, typename N, typename S,
            typename = std::enable_if_t<!is_tensor<S>::value>>
  PPF_HD PPF_INLINE Tensor<2, dim, N>
  operator*(const S &s, const Tensor<2, dim, N> &a)
  {
    Tensor<2, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = s * a[d];
    return r;
  }

  /// Single contraction of two rank-1 tensors (the scalar product).
  template <int dim, typename N>
  PPF_HD PPF_INLINE N
  operator*(const Tensor<1, dim, N> &a, const Tensor<1, dim, N> &b)
  {
    N s = a[0] * b[0];
    for (int d = 1; d < dim; ++d)
      s += a[d] * b[d];
    return s;
  }

  /// Matrix-vector contraction.
  template <int dim, typename N>
  PPF_HD PPF_INLINE Tensor<1, dim, N>
  operator*(const Tensor<2, dim, N> &a, const Tensor<1, dim, N> &b)
  {
    Tensor<1, dim, N> r;
    for (int d = 0; d < dim; ++d)
      r[d] = a[d] * b;
    return r;
  }

  template <int dim, typename N>
  PPF_HD PPF_INLINE N
  scalar_product(const Tensor<1, dim, N> &a, const Tensor<1, dim, N> &b)
  {
    return a * b;
  }

  /// A point in space with plain-number coordinates (or lane values for quadrature
  /// point locations).
  template <int dim, typename Number = double>
  class Point : public Tensor<1, dim, Number>
  {
  public:
    PPF_HD constexpr Point() = default;

    PPF_HD Point(const Tensor<1, dim, Number> &t)
      : Tensor<1, dim, Number>(t)
    {}

    PPF_HD Point(Number x, Number y)
    {
      static_assert(dim == 2, "two coordinates need dim == 2");
      this->values[0] = x;
      this->values[1] = y;
    }

    PPF_HD Point(Number x, Number y, Number z)
    {
      static_assert(dim == 3, "three coordinates need dim == 3");
      this->values[0] = x;
      this->values[1] = y;
      this->values[2] = z;
    }

    PPF_HD Number
    operator()(unsigned int d) const
    {
      return this->values[d];
    }

    PPF_HD Number &
    operator()(unsigned int d)
    {
      return this->values[d];
    }

    PPF_HD Number
    distance_square(const Point &p) const
    {
      return (Tensor<1, dim, Number>(*this) - Tensor<1, dim, Number>(p)).norm_square();
    }

    PPF_HD Number
    distance(const Point &p) const;
  };
} // namespace dealii

// Elementary functions on lane values, found through `std::` as the application code
// spells them (deal.II provides the same overloads for its VectorizedArray).
namespace std
{
#define PPF_LANE_UNARY(fn)                                                         \
  template <typename N, std::size_t w>                                             \
  PPF_HD PPF_INLINE ::dealii::VectorizedArray<N, w> fn(                           \
    const ::dealii::VectorizedArray<N, w> &a)                                      \
  {                                                                                \
    using std::fn;                                                                 \
    return ::dealii::VectorizedArray<N, w>(fn(a.data));                            \
  }
  PPF_LANE_UNARY(sqrt)
  PPF_LANE_UNARY(abs)
  PPF_LANE_UNARY(exp)
  PPF_LANE_UNARY(log)
  PPF_LANE_UNARY(tanh)
  PPF_LANE_UNARY(sin)
  PPF_LANE_UNARY(cos)
  PPF_LANE_UNARY(tan)
  PPF_LANE_UNARY(atan)
  PPF_LANE_UNARY(erf)
  PPF_LANE_UNARY(erfc)
#undef PPF_LANE_UNARY

  template <typename N, std::size_t w, typename E>
  PPF_HD PPF_INLINE ::dealii::VectorizedArray<N, w>
  pow(const ::dealii::VectorizedArray<N, w> &a, const E &e)
  {
    using std::pow;
    return ::dealii::VectorizedArray<N, w>(pow(a.data, static_cast<N>(e)));
  }

  template <typename N, std::size_t w>
  PPF_HD PPF_INLINE ::dealii::VectorizedArray<N, w>
  max(const ::dealii::VectorizedArray<N, w> &a, const ::dealii::VectorizedArray<N, w> &b)
  {
    return ::dealii::VectorizedArray<N, w>(a.data > b.data ? a.data : b.data);
  }

  template <typename N, std::size_t w>
  PPF_HD PPF_INLINE ::dealii::VectorizedArray<N, w>
  min(const ::dealii::VectorizedArray<N, w> &a, const ::dealii::VectorizedArray<N, w> &b)
  {
    return ::dealii::VectorizedArray<N, w>(a.data < b.data ? a.data : b.data);
  }
} // namespace std

namespace dealii
{
  /// Lane-wise comparisons (deal.II vectorization.h, SIMDComparison and
  /// compare_and_apply_mask): for each lane, true_value where `left predicate right`
  /// holds, false_value otherwise.
  enum class SIMDComparison : int
  {
    equal,
    not_equal,
    less_than,
    less_than_or_equal,
    greater_than,
    greater_than_or_equal
  };

  template <SIMDComparison predicate, typename Number>
  PPF_HD PPF_INLINE Number
  compare_and_apply_mask(const Number &left, const Number &right, const Number &true_value,
                         const Number &false_value)
  {
    bool c = false;
    if constexpr (predicate == SIMDComparison::equal)
      c = left == right;
    else if constexpr (predicate == SIMDComparison::not_equal)
      c = left != right;
    else if constexpr (predicate == SIMDComparison::less_than)
      c = left < right;
    else if constexpr (predicate == SIMDComparison::less_than_or_equal)
      c = left <= right;
    else if constexpr (predicate == SIMDComparison::greater_than)
      c = left > right;
    else
      c = left >= right;
    return c ? true_value : false_value;
  }

  template <SIMDComparison predicate, typename Number, std::size_t width>
  PPF_HD PPF_INLINE VectorizedArray<Number, width>
  compare_and_apply_mask(const VectorizedArray<Number, width> &left,
                         const VectorizedArray<Number, width> &right,
                         const VectorizedArray<Number, width> &true_value,
                         const VectorizedArray<Number, width> &false_value)
  {
    return compare_and_apply_mask<predicate, Number>(left[0], right[0], true_value[0],
                                                     false_value[0]);
  }

  template <int dim, typename Number>
  PPF_HD PPF_INLINE Number
  Tensor<1, dim, Number>::norm() const
  {
    using std::sqrt;
    return sqrt(norm_square());
  }

  template <int dim, typename Number>
  PPF_HD PPF_INLINE Number
  Point<dim, Number>::distance(const Point &p) const
  {
    using std::sqrt;
    return sqrt(distance_square(p));
  }
} // namespace dealii
