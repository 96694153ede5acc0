// Oracle value types: a one-lane `dealii::VectorizedArray`, `dealii::Tensor` of rank 1
// and 2, and `dealii::Point`, written independently of the product's versions so that a
// mistake in one does not hide in the other.  TEST INFRASTRUCTURE.
#pragma once

#include <prismspf/config.h>

#include <cmath>
#include <cstddef>
#include <type_traits>

namespace dealii
{
  template <typename Number, std::size_t width = 1>
  struct VectorizedArray
  {
    using value_type = Number;
    Number v;
    VectorizedArray() = default;
    VectorizedArray(Number x) : v(x) {}
    template <typename O, typename = std::enable_if_t<std::is_arithmetic_v<O> && !std::is_same_v<O, Number>>>
    VectorizedArray(O x) : v(Number(x)) {}
    Number &operator[](unsigned) { return v; }
    const Number &operator[](unsigned) const { return v; }
    VectorizedArray &operator+=(VectorizedArray o) { v += o.v; return *this; }
    VectorizedArray &operator-=(VectorizedArray o) { v -= o.v; return *this; }
    VectorizedArray &operator*=(VectorizedArray o) { v *= o.v; return *this; }
    VectorizedArray &operator/=(VectorizedArray o) { v /= o.v; return *this; }
  };

  template <typename T> struct is_va : std::false_type {};
  template <typename N, std::size_t w> struct is_va<VectorizedArray<N, w>> : std::true_type {};

  template <typename N, std::size_t w> VectorizedArray<N, w> operator+(VectorizedArray<N, w> a, VectorizedArray<N, w> b) { return a.v + b.v; }
  template <typename N, std::size_t w> VectorizedArray<N, w> operator-(VectorizedArray<N, w> a, VectorizedArray<N, w> b) { return a.v - b.v; }
  template <typename N, std::size_t w> VectorizedArray<N, w> operator*(VectorizedArray<N, w> a, VectorizedArray<N, w> b) { return a.v * b.v; }
  template <typename N, std::size_t w> VectorizedArray<N, w> operator/(VectorizedArray<N, w> a, VectorizedArray<N, w> b) { return a.v / b.v; }
  template <typename N, std::size_t w> VectorizedArray<N, w> operator-(VectorizedArray<N, w> a) { return -a.v; }
  template <typename N, std::size_t w> VectorizedArray<N, w> operator+(VectorizedArray<N, w> a) { return a; }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator+(VectorizedArray<N, w> a, S s) { return a.v + N(s); }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator+(S s, VectorizedArray<N, w> a) { return N(s) + a.v; }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator-(VectorizedArray<N, w> a, S s) { return a.v - N(s); }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator-(S s, VectorizedArray<N, w> a) { return N(s) - a.v; }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator*(VectorizedArray<N, w> a, S s) { return a.v * N(s); }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator*(S s, VectorizedArray<N, w> a) { return N(s) * a.v; }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator/(VectorizedArray<N, w> a, S s) { return a.v / N(s); }
  template <typename N, std::size_t w, typename S, typename = std::enable_if_t<std::is_arithmetic_v<S>>>
  VectorizedArray<N, w> operator/(S s, VectorizedArray<N, w> a) { return N(s) / a.v; }

  template <typename Number>
  VectorizedArray<Number> make_vectorized_array(Number u) { return u; }

  // deal.II's lane-wise select (vectorization.h).
  enum class SIMDComparison : int { equal, not_equal, less_than, less_than_or_equal, greater_than, greater_than_or_equal };

  template <SIMDComparison p, typename N>
  N compare_and_apply_mask(const N &l, const N &r, const N &t, const N &f)
  {
    switch (p)
      {
        case SIMDComparison::equal: return l == r ? t : f;
        case SIMDComparison::not_equal: return l != r ? t : f;
        case SIMDComparison::less_than: return l < r ? t : f;
        case SIMDComparison::less_than_or_equal: return l <= r ? t : f;
        case SIMDComparison::greater_than: return l > r ? t : f;
        default: return l >= r ? t : f;
      }
  }

  template <SIMDComparison p, typename N, std::size_t w>
  VectorizedArray<N, w> compare_and_apply_mask(const VectorizedArray<N, w> &l, const VectorizedArray<N, w> &r,
                                               const VectorizedArray<N, w> &t, const VectorizedArray<N, w> &f)
  {
    return compare_and_apply_mask<p, N>(l.v, r.v, t.v, f.v);
  }

} // namespace dealii

namespace std
{
#define ORACLE_LANE_FN(fn) \
  template <typename N, std::size_t w> ::dealii::VectorizedArray<N, w> fn(::dealii::VectorizedArray<N, w> a) { using std::fn; return fn(a.v); }
  ORACLE_LANE_FN(sqrt)
  ORACLE_LANE_FN(abs)
  ORACLE_LANE_FN(exp)
  ORACLE_LANE_FN(log)
  ORACLE_LANE_FN(tanh)
  ORACLE_LANE_FN(sin)
  ORACLE_LANE_FN(cos)
#undef ORACLE_LANE_FN
} // namespace std

namespace dealii
{
  template <int rank, int dim, typename Number = double> struct Tensor;

  template <typename T> struct is_tensor : std::false_type {};
  template <int r, int d, typename N> struct is_tensor<Tensor<r, d, N>> : std::true_type {};

  template <int dim, typename Number>
  struct Tensor<1, dim, Number>
  {
    static constexpr unsigned int rank = 1;
    using value_type = Number;
    Number c[dim];
    Tensor() { for (int i = 0; i < dim; ++i) c[i] = Number(0.0); }
    Number &operator[](unsigned i) { return c[i]; }
    const Number &operator[](unsigned i) const { return c[i]; }
    Tensor &operator+=(const Tensor &o) { for (int i = 0; i < dim; ++i) c[i] += o.c[i]; return *this; }
    Tensor &operator-=(const Tensor &o) { for (int i = 0; i < dim; ++i) c[i] -= o.c[i]; return *this; }
    Number norm_square() const { Number s = c[0] * c[0]; for (int i = 1; i < dim; ++i) s += c[i] * c[i]; return s; }
    Number norm() const { using std::sqrt; return sqrt(norm_square()); }
  };

  template <int dim, typename N> Tensor<1, dim, N> operator+(Tensor<1, dim, N> a, const Tensor<1, dim, N> &b) { return a += b; }
  template <int dim, typename N> Tensor<1, dim, N> operator-(Tensor<1, dim, N> a, const Tensor<1, dim, N> &b) { return a -= b; }
  template <int dim, typename N> Tensor<1, dim, N> operator-(const Tensor<1, dim, N> &a) { Tensor<1, dim, N> r; for (int i = 0; i < dim; ++i) r[i] = -a[i]; return r; }
  template <int dim, typename N, typename S, typename = std::enable_if_t<!is_tensor<S>::value>>
  Tensor<1, dim, N> operator*(const Tensor<1, dim, N> &a, const S &s) { Tensor<1, dim, N> r; for (int i = 0; i < dim; ++i) r[i] = a[i] * s; return r; }
  template <int dim, typename N, typename S, typename = std::enable_if_t<!is_tensor<S>::value>>
  Tensor<1, dim, N> operator*(const S &s, const Tensor<1, dim, N> &a) { Tensor<1, dim, N> r; for (int i = 0; i < dim; ++i) r[i] = s * a[i]; return r; }
  template <int dim, typename N, typename S, typename = std::enable_if_t<!is_tensor<S>::value>>
  Tensor<1, dim, N> operator/(const Tensor<1, dim, N> &a, const S &s) { Tensor<1, dim, N> r; for (int i = 0; i < dim; ++i) r[i] = a[i] / s; return r; }
  template <int dim, typename N>
  N operator*(const Tensor<1, dim, N> &a, const Tensor<1, dim, N> &b) { N s = a[0] * b[0]; for (int i = 1; i < dim; ++i) s += a[i] * b[i]; return s; }

  template <int dim, typename Number>
  struct Tensor<2, dim, Number>
  {
    static constexpr unsigned int rank = 2;
    using value_type = Number;
    Tensor<1, dim, Number> r[dim];
    Tensor<1, dim, Number> &operator[](unsigned i) { return r[i]; }
    const Tensor<1, dim, Number> &operator[](unsigned i) const { return r[i]; }
  };

  template <int dim, typename Number = double>
  struct Point : Tensor<1, dim, Number>
  {
    Point() = default;
    Point(const Tensor<1, dim, Number> &t) : Tensor<1, dim, Number>(t) {}
    Number operator()(unsigned i) const { return this->c[i]; }
    Number distance(const Point &p) const { Number s = 0; for (int i = 0; i < dim; ++i) s += (this->c[i] - p.c[i]) * (this->c[i] - p.c[i]); using std::sqrt; return sqrt(s); }
  };
} // namespace dealii

