// One-dimensional GLL collocation data for degrees 1 and 2.
//
// The reference discretizes with FE_Q on Gauss-Lobatto support points and the matching
// Gauss-Lobatto quadrature (reference: include/prismspf/core/system_wide.h:40-61), so
// quadrature points coincide with nodes.  On the unit interval:
//   w[q]     quadrature weights;
//   D[q][k]  derivative of the k-th Lagrange basis function at node q;
//   K[j][k]  element stiffness sum_q w[q] D[q][j] D[q][k].
#pragma once

namespace prismspf
{
  template <int degree>
  struct GLLBasis;

  template <>
  struct GLLBasis<1>
  {
    static constexpr double w[2]    = {0.5, 0.5};
    static constexpr double D[2][2] = {
      {-1.0, 1.0},
      {-1.0, 1.0}
    };
    static constexpr double K[2][2] = {
      { 1.0, -1.0},
      {-1.0,  1.0}
    };
  };

  template <>
  struct GLLBasis<2>
  {
    static constexpr double w[3]    = {1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0};
    static constexpr double D[3][3] = {
      {-3.0,  4.0, -1.0},
      {-1.0,  0.0,  1.0},
      { 1.0, -4.0,  3.0}
    };
    static constexpr double K[3][3] = {
      { 7.0 / 3.0, -8.0 / 3.0,  1.0 / 3.0},
      {-8.0 / 3.0, 16.0 / 3.0, -8.0 / 3.0},
      { 1.0 / 3.0, -8.0 / 3.0,  7.0 / 3.0}
    };
  };

  /// Tensor-product index helpers: local dof k of a (degree+1)^dim cell.
  template <int degree>
  constexpr int
  local_coord(int k, int d)
  {
    return d == 0 ? k % (degree + 1) : (d == 1 ? (k / (degree + 1)) % (degree + 1) :
                                                 k / ((degree + 1) * (degree + 1)));
  }

  /// Local dof obtained from k by replacing its coordinate in direction d with v.
  template <int degree>
  constexpr int
  with_coord(int k, int d, int v)
  {
    const int stride = d == 0 ? 1 : (d == 1 ? degree + 1 : (degree + 1) * (degree + 1));
    return k + (v - local_coord<degree>(k, d)) * stride;
  }
} // namespace prismspf
