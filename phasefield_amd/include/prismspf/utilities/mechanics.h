// Linear-elastic stress from a strain tensor and a Voigt stiffness matrix
// (reference include/prismspf/utilities/mechanics.h: voigt_tensor_size, compute_stress).
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/lane_types.h>

PRISMS_PF_BEGIN_NAMESPACE

namespace Mechanics
{
  /// Number of independent components of a symmetric dim x dim tensor: 1, 3, 6.
  template <unsigned int dim>
  constexpr unsigned int voigt_tensor_size = dim * (dim + 1) / 2;

  /// Voigt position of tensor entry (i, j): the diagonal first, then the off-diagonal
  /// pairs ordered yz, xz, xy (3-D) or xy (2-D).
  template <unsigned int dim>
  PPF_HD constexpr unsigned int
  voigt_index(unsigned int i, unsigned int j)
  {
    if (i == j)
      return i;
    if constexpr (dim == 2)
      return 2;
    else
      return 6 - (i + j); // yz -> 3, xz -> 4, xy -> 5
  }

  /// sigma = C epsilon with epsilon in Voigt form (engineering shear strains
  /// eps_ij + eps_ji), returned as the symmetric stress tensor.
  template <unsigned int dim, typename T, typename Number>
  PPF_HD PPF_INLINE void
  compute_stress(const dealii::Tensor<2, voigt_tensor_size<dim>, Number> &stiffness,
                 const dealii::Tensor<2, dim, T>                         &strain,
                 dealii::Tensor<2, dim, T>                               &stress)
  {
    constexpr unsigned int nv = voigt_tensor_size<dim>;
    T                      eps[nv];
    for (unsigned int i = 0; i < dim; ++i)
      for (unsigned int j = i; j < dim; ++j)
        eps[voigt_index<dim>(i, j)] = i == j ? strain[i][i] : strain[i][j] + strain[j][i];
    T sigma[nv];
    for (unsigned int a = 0; a < nv; ++a)
      {
        T s = T(typename T::value_type(stiffness[a][0])) * eps[0];
        for (unsigned int b = 1; b < nv; ++b)
          s += T(typename T::value_type(stiffness[a][b])) * eps[b];
        sigma[a] = s;
      }
    for (unsigned int i = 0; i < dim; ++i)
      for (unsigned int j = 0; j < dim; ++j)
        stress[i][j] = sigma[voigt_index<dim>(i, j)];
  }
} // namespace Mechanics

PRISMS_PF_END_NAMESPACE
