// Uniform Cartesian Q1/Q2 mesh as seen by the kernels.
//
// The reference builds a deal.II triangulation with `subdivided_hyper_rectangle`,
// refines it globally, numbers DoFs cell-wise and adds periodicity and Dirichlet
// constraints (reference: src/core/triangulation_manager.cc:105-118,
// src/user_inputs/spatial_discretization.cc:133-177, src/core/constraint_manager.cc).
// On this path the mesh is a tensor-product lattice of GLL nodes and every geometric
// quantity is analytic, so nothing but this descriptor is stored:
//
//   * cells[d]    cells per direction (subdivisions * 2^refinement);
//   * nodes[d]    stored nodes per direction = degree*cells[d] + (periodic ? 0 : 1).
//                 A periodic direction keeps only the masters; the image plane of
//                 deal.II's numbering (its N_d counts it) is the wrap-around;
//   * node index  i0 + nodes[0]*(i1 + nodes[1]*i2), x fastest;
//   * Dirichlet   per-field bit mask over the faces 2d (lower) / 2d+1 (upper),
//                 ignored in periodic directions.
// Quadrature is GLL collocated with the nodes (reference system_wide.h:40-61), so the
// lumped mass of node i is the product of 1-D assembled masses (see lumped_mass_1d).
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/gll_basis.h>

#include <cstdint>

namespace ppf
{
  struct GridDesc
  {
    int     dim;
    int     degree;
    int     cells[3];
    int     nodes[3];
    int     periodic[3];
    double  lower[3];
    double  h[3];
    int64_t n_nodes;
    int64_t n_cells;
  };

  /// Number of nodes per direction in one cell.
  template <int degree>
  constexpr int points_1d = degree + 1;

  template <int dim, int degree>
  constexpr int dofs_per_cell = dim == 2 ? (degree + 1) * (degree + 1)
                                         : (degree + 1) * (degree + 1) * (degree + 1);

  template <int degree>
  using GLL = prismspf::GLLBasis<degree>;

  PPF_HD PPF_INLINE int64_t
  node_index(const GridDesc &g, int i0, int i1, int i2)
  {
    return int64_t(i0) + int64_t(g.nodes[0]) * (int64_t(i1) + int64_t(g.nodes[1]) * i2);
  }

  /// Lattice coordinate degree*c + k of a cell-local node, wrapped in periodic
  /// directions.
  PPF_HD PPF_INLINE int
  wrap_lattice(const GridDesc &g, int d, int i)
  {
    return (g.periodic[d] && i >= g.nodes[d]) ? i - g.nodes[d] : i;
  }

  /// True if lattice coordinate i of direction d sits on a non-periodic boundary face
  /// whose bit is set in `mask`.
  PPF_HD PPF_INLINE bool
  on_masked_face(const GridDesc &g, int d, int i, unsigned int mask)
  {
    if (g.periodic[d])
      return false;
    return (i == 0 && (mask & (1u << (2 * d)))) ||
           (i == g.nodes[d] - 1 && (mask & (1u << (2 * d + 1))));
  }

  template <int dim>
  PPF_HD PPF_INLINE bool
  is_dirichlet(const GridDesc &g, const int (&i)[3], unsigned int mask)
  {
    if (mask == 0)
      return false;
    bool r = false;
    for (int d = 0; d < dim; ++d)
      r = r || on_masked_face(g, d, i[d], mask);
    return r;
  }

  /// Assembled 1-D lumped mass of lattice coordinate i in direction d: the sum over the
  /// (one or two) cells touching the node of w_k * h.
  template <int degree>
  PPF_HD PPF_INLINE double
  lumped_mass_1d(const GridDesc &g, int d, int i)
  {
    const double h        = g.h[d];
    const bool   boundary = !g.periodic[d] && (i == 0 || i == g.nodes[d] - 1);
    if constexpr (degree == 1)
      return boundary ? 0.5 * h : h;
    else
      {
        if (i & 1)
          return (4.0 / 6.0) * h;
        return boundary ? (1.0 / 6.0) * h : (2.0 / 6.0) * h;
      }
  }

  template <int dim, int degree>
  PPF_HD PPF_INLINE double
  lumped_mass(const GridDesc &g, const int (&i)[3])
  {
    double m = lumped_mass_1d<degree>(g, 0, i[0]);
    for (int d = 1; d < dim; ++d)
      m *= lumped_mass_1d<degree>(g, d, i[d]);
    return m;
  }

  /// Diagonal scaling applied after assembly (reference mf_operator.cc:25-31).
  enum Scaling : int
  {
    scale_none      = 0,
    scale_invm      = 1,
    scale_invm_sqrt = 2
  };

  /// What a kernel writes at a Dirichlet node of an output field.
  enum DirichletWrite : int
  {
    /// Operator semantics: constrained rows are dropped (value 0).
    dirichlet_zero = 0,
    /// Explicit-update semantics: the constraint value g is distributed.
    dirichlet_value = 1
  };

  /// Bits of a dependency slot descriptor.
  enum SlotFlags : unsigned int
  {
    slot_values      = 0x1,
    slot_gradients   = 0x2,
    /// Read with homogeneous constraints applied (deal.II read_dof_values, as for the
    /// operand of a linear operator): Dirichlet nodes read as zero.
    slot_constrained = 0x4
  };
} // namespace ppf
