// Matrix-free operator kernels for uniform Q1/Q2 meshes with GLL collocation.
//
// The reference evaluates an operator with deal.II's MatrixFree::cell_loop: per cell
// batch it gathers the declared dependencies, evaluates values and gradients at the
// quadrature points, calls the equation hook at every point and integrates/scatters
// the submitted terms (reference: src/solvers/mf_operator.cc:17-80,
// include/prismspf/core/field_container.h:753-882).  Two device realizations of that
// operator live here; the runtime picks one per solve block and both produce the
// reference's result:
//
// 1. Generic fused cell path (any hook).
//    cell_operator_kernel: one thread per cell gathers the cell's (p+1)^d node values
//    of every dependency, walks the (p+1)^d collocated quadrature points with the hook
//    inlined, and integrates the submitted value and gradient terms into per-cell
//    nodal contributions, written cell-contiguous to a scratch buffer.
//    node_gather_kernel: one thread per node sums the contributions of the <= 2^d
//    cells around it in a fixed order (deterministic, no atomics), applies the
//    diagonal scaling and the Dirichlet rule, and writes the destination.
//
// 2. Collocation-split path (hooks whose value term depends only on values and whose
//    gradient term is a constant multiple of dependency gradients).
//    With GLL collocation the value term integrates node-locally to m_i r(u_i), and
//    the gradient term assembles to the tensor-product stiffness
//    K = sum_d (prod_{e!=d} M_e) (x) K_d (M_e assembled 1-D lumped masses, K_d the
//    assembled 1-D stiffness), i.e. a (2p+1)-point line stencil per direction.
//    split_operator_kernel: one thread per node evaluates the hook once with the
//    node's values and adds beta * (K u) for each gradient term.  The runtime decides
//    eligibility by probing the hook on the host and cross-checks the two paths.
#pragma once

#include "grid.hpp"

#include <prismspf/core/field_container.h>
#include <prismspf/core/pde_operator_base.h>
#include <prismspf/core/simulation_timer.h>

namespace ppf
{
  constexpr int block_size = 256;

  /// Dependencies of an operator: one array per (field, slot) with its flags.
  template <typename number>
  struct DepTable
  {
    const number *ptr[PPF_MAX_FIELDS][PPF_N_SLOTS];
    unsigned int  flags[PPF_MAX_FIELDS][PPF_N_SLOTS];
    unsigned int  dirichlet_mask[PPF_MAX_FIELDS];
  };

  /// Outputs of an operator, indexed by global field index.
  template <typename number>
  struct OutTable
  {
    number       *dst[PPF_MAX_FIELDS];
    const number *dirichlet_values[PPF_MAX_FIELDS];
    unsigned int  dirichlet_mask[PPF_MAX_FIELDS];
    int           scratch_slot[PPF_MAX_FIELDS];
    unsigned int  out_mask;
    int           scaling;
    int           dirichlet_write;
    /// Device flag that, once nonzero, turns the launch into a no-op (an increment
    /// enqueued ahead of the previous increment's NaN probe); null: always run.
    const int    *halt;
  };

  __device__ inline bool
  halted(const int *halt)
  {
    return halt && *halt;
  }

  /// One gradient-stencil term of the split path: dst[out] += beta * K u[field][slot].
  struct SplitTerm
  {
    int          out_field;
    int          in_field;
    int          in_slot;
    double       beta;
  };

  constexpr int max_split_terms = 8;

  struct SplitTerms
  {
    int       n_terms;
    SplitTerm term[max_split_terms];
  };

  template <int dim>
  PPF_HD PPF_INLINE void
  decompose(const GridDesc &g, int64_t linear, const int (&extent)[3], int (&i)[3])
  {
    i[0] = int(linear % extent[0]);
    i[1] = int((linear / extent[0]) % extent[1]);
    i[2] = dim == 3 ? int(linear / (int64_t(extent[0]) * extent[1])) : 0;
  }

  /// Calls the application's hook with a qualified (non-virtual) call on the object the
  /// kernel received by value, so the hook is inlined into the quadrature loop and its
  /// constants are read straight from the kernel arguments.  Reference applications
  /// declare their hooks `private` (e.g. applications/allen_cahn/explicit/
  /// custom_pde.h:70-101) and reach them through the base's public virtuals; the
  /// application libraries are therefore compiled with clang's -fno-access-control.
  /// The virtual table is never used on the device.
  template <class PDE, int dim, int degree, typename number, bool lhs, int block_id>
  PPF_HD PPF_INLINE void
  call_hook(const PDE                                     &pde,
            prismspf::FieldContainer<dim, degree, number> &fc,
            const prismspf::SimulationTimer               &timer)
  {
    if constexpr (lhs)
      pde.PDE::compute_lhs(fc, timer, block_id);
    else
      pde.PDE::compute_rhs(fc, timer, block_id);
  }

  // ------------------------------------------------------------------------------
  // Generic fused cell path
  // ------------------------------------------------------------------------------

  /// Per-cell evaluation: gather, hook at every collocated quadrature point, integrate.
  /// `scratch` holds n_out * dofs_per_cell * n_cells numbers, laid out
  /// [slot][local dof][cell] so that neighbouring threads store contiguously.
  ///
  /// With `diagonal` set the kernel instead computes the operator diagonal the way the
  /// reference probes it (mf_operator.cc:134-218): for every local dof j of field
  /// `probe_field` the src values are the unit vector e_j, and only the j-th integrated
  /// entry is kept.
  template <class PDE, int dim, int degree, typename number, bool lhs, int block_id,
            bool diagonal, int n_fields>
  __global__ void __launch_bounds__(block_size)
  cell_operator_kernel(const PDE                     pde,
                       const GridDesc                g,
                       const DepTable<number>        deps,
                       const OutTable<number>        out,
                       number *__restrict__          scratch,
                       const prismspf::SimulationTimer timer,
                       const int                     probe_field)
  {
    constexpr int NPC = dofs_per_cell<dim, degree>;
    constexpr int P1  = degree + 1;
    using Basis       = GLL<degree>;

    const int64_t cell = int64_t(blockIdx.x) * block_size + threadIdx.x;
    if (cell >= g.n_cells || halted(out.halt))
      return;
    int c[3];
    decompose<dim>(g, cell, g.cells, c);

    // Node lattice coordinates of the cell's local nodes, per direction.
    int lat[3][P1];
    for (int d = 0; d < 3; ++d)
      for (int k = 0; k < P1; ++k)
        lat[d][k] = d < dim ? wrap_lattice(g, d, degree * c[d] + k) : 0;

    prismspf::FieldContainer<dim, degree, number> fc;
    fc.cell_view = true;
    auto &u      = fc.cell_values;
#pragma unroll
    for (int f = 0; f < n_fields; ++f)
#pragma unroll
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        {
          const unsigned int fl = deps.flags[f][s];
          if (fl & (slot_values | slot_gradients))
            {
#pragma unroll
              for (int k = 0; k < NPC; ++k)
                {
                  int i[3] = {lat[0][prismspf::local_coord<degree>(k, 0)],
                              lat[1][prismspf::local_coord<degree>(k, 1)],
                              dim == 3 ? lat[2][prismspf::local_coord<degree>(k, 2)] : 0};
                  number v = deps.ptr[f][s][node_index(g, i[0], i[1], i[2])];
                  if ((fl & slot_constrained) && is_dirichlet<dim>(g, i, deps.dirichlet_mask[f]))
                    v = number(0);
                  u[f][s][k] = v;
                }
            }
          else
            {
#pragma unroll
              for (int k = 0; k < NPC; ++k)
                u[f][s][k] = number(0);
            }
        }

    number inv_h[dim];
    number cell_volume = number(1);
    for (int d = 0; d < dim; ++d)
      {
        inv_h[d] = number(1.0 / g.h[d]);
        cell_volume *= number(g.h[d]);
      }

    fc.element_volume = cell_volume;
    for (int d = 0; d < dim; ++d)
      fc.inv_h[d] = inv_h[d];

    constexpr int src = prismspf::dependency_slot(prismspf::SRC);
    const int     n_probe = diagonal ? NPC : 1;
    number        diag_acc[PPF_MAX_FIELDS][NPC];
    for (int j = 0; j < n_probe; ++j)
      {
        if constexpr (diagonal)
          {
#pragma unroll
            for (int f = 0; f < n_fields; ++f)
#pragma unroll
              for (int k = 0; k < NPC; ++k)
                u[f][src][k] = (f == probe_field && k == j) ? number(1) : number(0);
          }

        number acc[PPF_MAX_FIELDS][NPC];
#pragma unroll
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
          for (int k = 0; k < NPC; ++k)
            acc[f][k] = number(0);

#pragma unroll
        for (int q = 0; q < NPC; ++q)
          {
            number weight = cell_volume;
#pragma unroll
            for (int d = 0; d < dim; ++d)
              weight *= number(Basis::w[prismspf::local_coord<degree>(q, d)]);

            fc.cell_q = q;
#pragma unroll
            for (int d = 0; d < dim; ++d)
              fc.q_point[d] =
                number(g.lower[d] + (degree * c[d] + prismspf::local_coord<degree>(q, d)) *
                                      (g.h[d] / degree));

            fc.clear_terms();
            call_hook<PDE, dim, degree, number, lhs, block_id>(pde, fc, timer);

#pragma unroll
            for (int f = 0; f < PPF_MAX_FIELDS; ++f)
              {
                if (fc.value_term_mask & (1u << f))
                  acc[f][q] += weight * fc.value_term[f];
                if (fc.gradient_term_mask & (1u << f))
                  {
#pragma unroll
                    for (int d = 0; d < dim; ++d)
                      {
                        const number t  = weight * inv_h[d] * fc.gradient_term[f][d];
                        const int    qd = prismspf::local_coord<degree>(q, d);
#pragma unroll
                        for (int k = 0; k < P1; ++k)
                          if (Basis::D[qd][k] != 0.0)
                            acc[f][prismspf::with_coord<degree>(q, d, k)] +=
                              number(Basis::D[qd][k]) * t;
                      }
                  }
              }
          }

        if constexpr (diagonal)
          {
#pragma unroll
            for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
              for (int k = 0; k < NPC; ++k)
                if (k == j)
                  diag_acc[f][k] = acc[f][k];
          }
        else
          {
#pragma unroll
            for (int f = 0; f < PPF_MAX_FIELDS; ++f)
              {
                const int slot = out.scratch_slot[f];
                if (slot >= 0)
                  {
#pragma unroll
                    for (int k = 0; k < NPC; ++k)
                      scratch[(int64_t(slot) * NPC + k) * g.n_cells + cell] = acc[f][k];
                  }
              }
          }
      }

    if constexpr (diagonal)
      {
        const int slot = out.scratch_slot[probe_field];
#pragma unroll
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
          if (f == probe_field)
            {
#pragma unroll
              for (int k = 0; k < NPC; ++k)
                scratch[(int64_t(slot) * NPC + k) * g.n_cells + cell] = diag_acc[f][k];
            }
      }
  }

  /// Cells touching lattice coordinate i in direction d, as (cell coordinate, local
  /// index) pairs: slot 0 is the left cell when it exists, slot 1 the right cell.
  /// Returns how many slots are valid (1 or 2).  Fixed slots keep the arrays in
  /// registers.
  template <int degree>
  PPF_HD PPF_INLINE int
  adjacent_cells_1d(const GridDesc &g, int d, int i, int (&cell)[2], int (&local)[2])
  {
    if (degree == 2 && (i & 1))
      {
        cell[0] = cell[1] = (i - 1) / 2;
        local[0] = local[1] = 1;
        return 1;
      }
    const int  right     = i / degree;
    const bool has_left  = right > 0 || g.periodic[d];
    const bool has_right = right < g.cells[d];
    const int  left      = right > 0 ? right - 1 : g.cells[d] - 1;
    cell[0]              = has_left ? left : right;
    local[0]             = has_left ? degree : 0;
    cell[1]              = right;
    local[1]             = 0;
    return int(has_left) + int(has_right);
  }

  template <int dim, int degree>
  PPF_HD PPF_INLINE double
  node_scaling(const GridDesc &g, const int (&i)[3], int scaling)
  {
    if (scaling == scale_none)
      return 1.0;
    const double m = lumped_mass<dim, degree>(g, i);
    return scaling == scale_invm ? 1.0 / m : 1.0 / sqrt(m);
  }

  /// Node-ordered assembly of the per-cell contributions (see cell_operator_kernel).
  template <int dim, int degree, typename number>
  __global__ void __launch_bounds__(block_size)
  node_gather_kernel(const GridDesc g, const OutTable<number> out,
                     const number *__restrict__ scratch)
  {
    constexpr int NPC  = dofs_per_cell<dim, degree>;
    const int64_t node = int64_t(blockIdx.x) * block_size + threadIdx.x;
    if (node >= g.n_nodes || halted(out.halt))
      return;
    int i[3];
    decompose<dim>(g, node, g.nodes, i);

    int cells[3][2], locs[3][2], counts[3] = {1, 1, 1};
    for (int d = 0; d < dim; ++d)
      counts[d] = adjacent_cells_1d<degree>(g, d, i[d], cells[d], locs[d]);
    for (int d = dim; d < 3; ++d)
      {
        cells[d][0] = cells[d][1] = 0;
        locs[d][0] = locs[d][1] = 0;
      }

    const double scale = node_scaling<dim, degree>(g, i, out.scaling);
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
      {
        const int slot = out.scratch_slot[f];
        if (slot < 0)
          continue;
        number result;
        if (is_dirichlet<dim>(g, i, out.dirichlet_mask[f]))
          result = out.dirichlet_write == dirichlet_value ? out.dirichlet_values[f][node]
                                                          : number(0);
        else
          {
            number sum = number(0);
#pragma unroll
            for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
              for (int a1 = 0; a1 < 2; ++a1)
#pragma unroll
                for (int a0 = 0; a0 < 2; ++a0)
                  {
                    if (a0 >= counts[0] || a1 >= counts[1] || a2 >= counts[2])
                      continue;
                    const int64_t c = cells[0][a0] +
                                      int64_t(g.cells[0]) *
                                        (cells[1][a1] + int64_t(g.cells[1]) * cells[2][a2]);
                    const int k = locs[0][a0] + (degree + 1) * (locs[1][a1] +
                                                                (degree + 1) * locs[2][a2]);
                    sum += scratch[(int64_t(slot) * NPC + k) * g.n_cells + c];
                  }
            result = number(scale * double(sum));
          }
        out.dst[f][node] = result;
      }
  }

  // ------------------------------------------------------------------------------
  // Collocation-split path
  // ------------------------------------------------------------------------------

  /// (K u)_i for the tensor-product stiffness: sum over directions of the transverse
  /// lumped mass times the assembled 1-D stiffness row, reading along the node's line.
  template <int dim, int degree, typename number>
  PPF_HD PPF_INLINE number
  stiffness_row(const GridDesc &g, const int (&i)[3], const number *__restrict__ u,
                bool constrained, unsigned int dmask)
  {
    using Basis = GLL<degree>;
    number acc  = number(0);
#pragma unroll
    for (int d = 0; d < dim; ++d)
      {
        double transverse = 1.0;
#pragma unroll
        for (int e = 0; e < dim; ++e)
          if (e != d)
            transverse *= lumped_mass_1d<degree>(g, e, i[e]);
        int       cells[2], locs[2];
        const int n    = adjacent_cells_1d<degree>(g, d, i[d], cells, locs);
        number    line = number(0);
#pragma unroll
        for (int a = 0; a < 2; ++a)
          {
            if (a >= n)
              continue;
#pragma unroll
            for (int k = 0; k <= degree; ++k)
              {
                int j[3] = {i[0], i[1], i[2]};
                j[d]     = wrap_lattice(g, d, degree * cells[a] + k);
                number v = u[node_index(g, j[0], j[1], j[2])];
                if (constrained && is_dirichlet<dim>(g, j, dmask))
                  v = number(0);
                line += number(Basis::K[locs[a]][k]) * v;
              }
          }
        acc += number(transverse / g.h[d]) * line;
      }
    return acc;
  }

  template <class PDE, int dim, int degree, typename number, bool lhs, int block_id>
  __global__ void __launch_bounds__(block_size)
  split_operator_kernel(const PDE                       pde,
                        const GridDesc                  g,
                        const DepTable<number>          deps,
                        const OutTable<number>          out,
                        const SplitTerms                terms,
                        const prismspf::SimulationTimer timer)
  {
    const int64_t node = int64_t(blockIdx.x) * block_size + threadIdx.x;
    if (node >= g.n_nodes || halted(out.halt))
      return;
    int i[3];
    decompose<dim>(g, node, g.nodes, i);

    prismspf::FieldContainer<dim, degree, number> fc;
    double                                        cell_volume = 1.0;
#pragma unroll
    for (int d = 0; d < dim; ++d)
      {
        cell_volume *= g.h[d];
        fc.q_point[d] = number(g.lower[d] + i[d] * (g.h[d] / degree));
      }
    fc.element_volume = number(cell_volume);
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        {
          number v = number(0);
          const unsigned int fl = deps.flags[f][s];
          if (fl & slot_values)
            {
              v = deps.ptr[f][s][node];
              if ((fl & slot_constrained) && is_dirichlet<dim>(g, i, deps.dirichlet_mask[f]))
                v = number(0);
            }
          fc.value[f][s] = v;
#pragma unroll
          for (int d = 0; d < dim; ++d)
            fc.gradient[f][s][d] = number(0);
        }
    fc.clear_terms();
    call_hook<PDE, dim, degree, number, lhs, block_id>(pde, fc, timer);

    const double mass  = lumped_mass<dim, degree>(g, i);
    const double scale = node_scaling<dim, degree>(g, i, out.scaling);
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
      {
        if (!(out.out_mask & (1u << f)))
          continue;
        number result;
        if (is_dirichlet<dim>(g, i, out.dirichlet_mask[f]))
          result = out.dirichlet_write == dirichlet_value ? out.dirichlet_values[f][node]
                                                          : number(0);
        else
          {
            number acc = number(0);
            if (fc.value_term_mask & (1u << f))
              acc = number(mass) * fc.value_term[f];
            for (int t = 0; t < terms.n_terms; ++t)
              if (terms.term[t].out_field == f)
                {
                  const int in = terms.term[t].in_field, s = terms.term[t].in_slot;
                  acc += number(terms.term[t].beta) *
                         stiffness_row<dim, degree, number>(
                           g, i, deps.ptr[in][s], (deps.flags[in][s] & slot_constrained) != 0,
                           deps.dirichlet_mask[in]);
                }
            result = number(scale) * acc;
          }
        out.dst[f][node] = result;
      }
  }

  /// Diagonal of the split operator: value-term derivative (probed on the host as
  /// `alpha`, constant) times the lumped mass, plus beta times the stiffness diagonal.
  template <int dim, int degree, typename number>
  __global__ void __launch_bounds__(block_size)
  split_diagonal_kernel(const GridDesc g, number *__restrict__ diag, double alpha, double beta,
                        unsigned int dmask, int scaling)
  {
    using Basis        = GLL<degree>;
    const int64_t node = int64_t(blockIdx.x) * block_size + threadIdx.x;
    if (node >= g.n_nodes)
      return;
    int i[3];
    decompose<dim>(g, node, g.nodes, i);
    if (is_dirichlet<dim>(g, i, dmask))
      {
        diag[node] = number(0);
        return;
      }
    double kii = 0.0;
    for (int d = 0; d < dim; ++d)
      {
        double transverse = 1.0;
        for (int e = 0; e < dim; ++e)
          if (e != d)
            transverse *= lumped_mass_1d<degree>(g, e, i[e]);
        int       cells[2], locs[2];
        const int n    = adjacent_cells_1d<degree>(g, d, i[d], cells, locs);
        double    line = 0.0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
          if (a < n)
            line += Basis::K[locs[a]][locs[a]];
        kii += transverse * line / g.h[d];
      }
    // The reference scales the probed diagonal like the operator output
    // (mf_operator.cc:91-97): left scaling only.
    const double m = lumped_mass<dim, degree>(g, i);
    diag[node]     = number(node_scaling<dim, degree>(g, i, scaling) * (alpha * m + beta * kii));
  }
} // namespace ppf
