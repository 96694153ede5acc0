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
  // Generic path, one launch: cell blocks in shared memory
  // ------------------------------------------------------------------------------
  //
  // A workgroup evaluates a box of cells and writes the final values of the nodes it
  // owns.  Per evaluated box:
  //   1. every input (field, slot) of the operator is staged in shared memory over the
  //      box's nodes (per-axis tables give each position's global offset, periodic wrap
  //      and Dirichlet faces; constrained inputs read as zero on their Dirichlet faces);
  //   2. one thread per cell walks the (p+1)^d collocated quadrature points with the
  //      hook inlined, reading values and gradients from the staged box (1-D derivative
  //      contractions), and integrates the submitted terms into the cell's (p+1)^d
  //      nodal contributions, kept in registers;
  //   3. per output field the contributions go to shared memory, and every owned node
  //      sums those of its <= 2^d cells in the fixed order of node_gather_kernel (lower
  //      cell before upper cell, z outermost), applies the scaling and the Dirichlet
  //      rule, and is written once.
  // Two kernels use these steps:
  //   * cell_block_kernel (1-D, 2-D): a block of E cells per axis owning the nodes of
  //     its last E - 1; the first layer is the neighbour whose upper faces carry owned
  //     nodes, evaluated again;
  //   * cell_column_kernel (3-D): a 9 x 9 column of cells owning 8 x 8 marches up a
  //     chunk of z layers in slabs of 3.  The nodes on a slab's top plane keep their
  //     lower-cell sums in shared memory until the next slab adds the upper cells, so
  //     only the chunk's first layer and the x/y halo are evaluated twice.
  // No global scratch, one launch, deterministic: the sum of every node is formed in
  // the same order as by the two-pass kernels, so both give the same bits.

  constexpr int cb_threads    = 256;
  constexpr int cb_max_inputs = PPF_MAX_FIELDS * PPF_N_SLOTS;

  /// Evaluated cells per axis of a cell_block_kernel block (their product is cb_threads).
  template <int dim>
  constexpr int cb_extent(int d)
  {
    return dim == 3 ? (d == 2 ? 4 : 8) : (dim == 2 ? (d < 2 ? 16 : 1) : (d == 0 ? 256 : 1));
  }

  /// Evaluated cells per slab of cell_column_kernel: 9 x 9 x 3 (243 of the 256 threads).
  constexpr int cc_ex = 9, cc_ey = 9, cc_ez = 3;

  /// One owned lattice position of a block along one axis.
  struct GatherAxis
  {
    int          a[2];  ///< shared-memory offsets of the slot 0 / slot 1 cell contributions
    unsigned int slots; ///< bit s: slot s exists
    unsigned int faces; ///< Dirichlet face bits of the position
    int64_t      goff;  ///< global node index part (position * stride); -1: not owned
    double       mass;  ///< 1-D lumped mass (1 beyond dim)
  };

  struct CellBlockArgs
  {
    int n_in;
    int in_field[cb_max_inputs];
    int in_slot[cb_max_inputs];
    int nblk[3];  ///< blocks (cell_block_kernel) or x/y tiles and z chunks (column kernel)
    int bcell[3]; ///< owned cells per block and axis (column kernel: 8, 8, chunk layers)
    /// Staged position of each (field, slot), 0 for those not staged (never read).
    int pos[PPF_MAX_FIELDS][PPF_N_SLOTS];
    /// cell_column_kernel: a second contribution buffer, output fields gathered in pairs.
    int pair;
  };

  namespace cellblock
  {
    constexpr unsigned long long off_mask = (1ull << 48) - 1;
    constexpr unsigned long long outside  = 1ull << 63;

    /// Staging table entry of lattice position `gi` along axis d (d < dim): global
    /// offset (wrapped when periodic), Dirichlet face bits << 48, or the outside bit.
    PPF_HD PPF_INLINE unsigned long long
    staging_entry(const GridDesc &g, int d, int gi)
    {
      const int nd = g.nodes[d];
      unsigned long long e;
      if (g.periodic[d])
        {
          gi = ((gi % nd) + nd) % nd;
          e  = 0;
        }
      else if (gi < 0 || gi >= nd)
        return outside;
      else
        e = (gi == 0 ? 1ull << (48 + 2 * d) : 0ull) |
            (gi == nd - 1 ? 1ull << (48 + 2 * d + 1) : 0ull);
      const int64_t stride =
        d == 0 ? 1 : (d == 1 ? int64_t(g.nodes[0]) : int64_t(g.nodes[0]) * g.nodes[1]);
      return e | (unsigned long long)(int64_t(gi) * stride);
    }

    /// Gather table entry of owned lattice position n along axis d (d < dim) of a box
    /// whose first evaluated cell is c0: slot offsets in the contribution array (cell
    /// stride cstr, local-node stride kstr), face bits, global offset, 1-D mass.
    template <int degree>
    PPF_HD PPF_INLINE GatherAxis
    gather_entry(const GridDesc &g, int d, int n, bool owned, int c0, int cstr, int kstr)
    {
      GatherAxis e {};
      const int64_t stride =
        d == 0 ? 1 : (d == 1 ? int64_t(g.nodes[0]) : int64_t(g.nodes[0]) * g.nodes[1]);
      e.goff       = owned ? int64_t(n) * stride : -1;
      const int lp = n - degree * c0; // lattice position in the box, >= degree
      if (degree == 2 && (n & 1))
        {
          e.a[0]  = ((lp - 1) / 2) * cstr + kstr;
          e.a[1]  = 0;
          e.slots = 1;
        }
      else
        {
          const int  right     = n / degree;
          const bool has_left  = right > 0 || g.periodic[d];
          const bool has_right = right < g.cells[d];
          e.a[0]  = has_left ? (lp / degree - 1) * cstr + degree * kstr : (lp / degree) * cstr;
          e.a[1]  = (lp / degree) * cstr;
          e.slots = (has_left && has_right) ? 3u : 1u;
        }
      e.faces = g.periodic[d] ? 0u
                              : ((n == 0 ? 1u << (2 * d) : 0u) |
                                 (n == g.nodes[d] - 1 ? 1u << (2 * d + 1) : 0u));
      e.mass = owned ? lumped_mass_1d<degree>(g, d, n) : 1.0;
      return e;
    }

    /// Stages every input over a box of bx x by x bz nodes whose per-axis staging
    /// entries are in `table`; per batch of up to four inputs every load is issued
    /// before the first shared-memory store, so their latencies overlap.
    template <int bx, int by, int bz, int ax_max, typename number>
    __device__ PPF_INLINE void
    stage(number *lds, const unsigned long long (*table)[ax_max], const DepTable<number> &deps,
          const CellBlockArgs &cb, int tid)
    {
      constexpr int box_n      = bx * by * bz;
      constexpr int per_thread = (box_n + cb_threads - 1) / cb_threads;
      constexpr int batch      = 4;
      for (int s0 = 0; s0 < cb.n_in; s0 += batch)
        {
          number v[batch][per_thread];
#pragma unroll
          for (int u = 0; u < batch; ++u)
            {
              if (s0 + u >= cb.n_in)
                break;
              const int                f   = cb.in_field[s0 + u];
              const int                sl  = cb.in_slot[s0 + u];
              const number            *src = deps.ptr[f][sl];
              const unsigned long long dm =
                (deps.flags[f][sl] & slot_constrained) ? deps.dirichlet_mask[f] : 0u;
              const unsigned long long reject = (dm << 48) | outside;
#pragma unroll
              for (int r = 0; r < per_thread; ++r)
                {
                  const int i  = tid + r * cb_threads;
                  const int ii = i < box_n ? i : 0;
                  const unsigned long long e0 = table[0][ii % bx];
                  const unsigned long long e1 = table[1][(ii / bx) % by];
                  const unsigned long long e2 = table[2][ii / (bx * by)];
                  const bool keep = i < box_n && !((e0 | e1 | e2) & reject);
                  const unsigned long long o =
                    (e0 & off_mask) + (e1 & off_mask) + (e2 & off_mask);
                  // rejected nodes read entry 0 (in bounds) and keep zero
                  const number x = src[keep ? o : 0ull];
                  v[u][r]        = keep ? x : number(0);
                }
            }
#pragma unroll
          for (int u = 0; u < batch; ++u)
            {
              if (s0 + u >= cb.n_in)
                break;
#pragma unroll
              for (int r = 0; r < per_thread; ++r)
                {
                  const int i = tid + r * cb_threads;
                  if (i < box_n)
                    lds[(s0 + u) * box_n + i] = v[u][r];
                }
            }
        }
    }

    /// Step 2 for one cell: the hook at every collocated point of cell `gc` (global
    /// cell coordinates), whose local node (0,0,0) is at `base` of every staged box
    /// (strides bx, bx by), integrated into acc.
    template <class PDE, int dim, int degree, typename number, bool lhs, int block_id>
    __device__ PPF_INLINE void
    integrate_cell(const PDE &pde, const GridDesc &g, const prismspf::SimulationTimer &timer,
                   const number *lds, int box_n, const CellBlockArgs &cb, int base, int sx,
                   int sxy, const int (&gc)[3],
                   number (&acc)[PPF_MAX_FIELDS][dofs_per_cell<dim, degree>])
    {
      constexpr int NPC = dofs_per_cell<dim, degree>;
      constexpr int P1  = degree + 1;
      using Basis       = GLL<degree>;
      prismspf::FieldContainer<dim, degree, number> fc;
      fc.block_view = true;
      // Constant indices into the container only: a runtime index would place the
      // whole container in private memory.
#pragma unroll
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
        for (int sl = 0; sl < PPF_N_SLOTS; ++sl)
          fc.block_ptr[f][sl] = lds + cb.pos[f][sl] * box_n;
      fc.block_base      = base;
      fc.block_stride[0] = sx;
      fc.block_stride[1] = sxy;

      number inv_h[dim];
      number cell_volume = number(1);
#pragma unroll
      for (int d = 0; d < dim; ++d)
        {
          inv_h[d] = number(1.0 / g.h[d]);
          cell_volume *= number(g.h[d]);
        }
      fc.element_volume = cell_volume;
#pragma unroll
      for (int d = 0; d < dim; ++d)
        fc.inv_h[d] = inv_h[d];

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
              number(g.lower[d] + (degree * gc[d] + prismspf::local_coord<degree>(q, d)) *
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
    }

    /// Global cell coordinate of evaluated cell c0 + e along axis d: wrapped when
    /// periodic; `valid` cleared when outside a non-periodic axis.
    PPF_HD PPF_INLINE int
    cell_coordinate(const GridDesc &g, int d, int c, bool &valid)
    {
      if (g.periodic[d])
        return ((c % g.cells[d]) + g.cells[d]) % g.cells[d];
      valid = valid && c >= 0 && c < g.cells[d];
      return c;
    }

    /// What the write-out of one output field needs, read from the kernel arguments
    /// once per field (a load from the arguments inside the node loop would also wait
    /// for every shared-memory read in flight).
    template <typename number>
    struct FieldOut
    {
      number      *dst;
      unsigned int mask;  ///< Dirichlet faces
      bool         keep;  ///< Dirichlet nodes are left to dirichlet_faces_kernel
      int          scaling;

      PPF_HD PPF_INLINE
      FieldOut(const OutTable<number> &out, int f)
        : dst(out.dst[f])
        , mask(out.dirichlet_mask[f])
        , keep(out.dirichlet_write == dirichlet_value)
        , scaling(out.scaling)
      {}

      /// Writes an owned node: the scaled sum, zero at a Dirichlet node, nothing at a
      /// Dirichlet node of an explicit update (dirichlet_faces_kernel writes those
      /// values afterwards; the kernels read no global memory while storing, so no
      /// store waits for an earlier one).
      PPF_HD PPF_INLINE void
      write(number sum, unsigned int faces, double mass, int64_t node) const
      {
        const bool dir   = (faces & mask) != 0;
        double     scale = 1.0;
        if (scaling != scale_none)
          scale = scaling == scale_invm ? 1.0 / mass : 1.0 / sqrt(mass);
        const number v = dir ? number(0) : number(scale * double(sum));
        if (!(dir && keep))
          dst[node] = v;
      }
    };
  } // namespace cellblock

  template <class PDE, int dim, int degree, typename number, bool lhs, int block_id>
  __global__ void __launch_bounds__(cb_threads)
  cell_block_kernel(const PDE pde, const GridDesc g, const DepTable<number> deps,
                    const OutTable<number> out, const prismspf::SimulationTimer timer,
                    const CellBlockArgs cb)
  {
    constexpr int NPC = dofs_per_cell<dim, degree>;
    constexpr int P1  = degree + 1;
    constexpr int E0 = cb_extent<dim>(0), E1 = cb_extent<dim>(1), E2 = cb_extent<dim>(2);
    static_assert(E0 * E1 * E2 == cb_threads, "one evaluated cell per thread");
    extern __shared__ unsigned char cb_shared[];
    number *lds = reinterpret_cast<number *>(cb_shared);

    if (halted(out.halt))
      return;
    const int tid = threadIdx.x;
    int       bc[3];
    {
      int b = blockIdx.x;
      bc[0] = b % cb.nblk[0];
      b /= cb.nblk[0];
      bc[1] = b % cb.nblk[1];
      bc[2] = b / cb.nblk[1];
    }
    // First evaluated cell per axis: one layer below the owned cells.
    int c0[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      c0[d] = d < dim ? bc[d] * cb.bcell[d] - 1 : 0;
    constexpr int bx    = degree * E0 + 1;
    constexpr int by    = dim > 1 ? degree * E1 + 1 : 1;
    constexpr int bz    = dim > 2 ? degree * E2 + 1 : 1;
    constexpr int box_n = bx * by * bz;
    constexpr int ax_max = bx > by ? (bx > bz ? bx : bz) : (by > bz ? by : bz);
    // Owned lattice positions per axis: p B (+1: the closing boundary node).
    constexpr int ox     = degree * (E0 - 1) + 1;
    constexpr int oy     = dim > 1 ? degree * (E1 - 1) + 1 : 1;
    constexpr int oz     = dim > 2 ? degree * (E2 - 1) + 1 : 1;
    constexpr int ow_max = ox > oy ? (ox > oz ? ox : oz) : (oy > oz ? oy : oz);
    constexpr int n_cand = (ox * oy * oz + cb_threads - 1) / cb_threads;
    __shared__ unsigned long long axis_table[3][ax_max];
    __shared__ GatherAxis         gather_table[3][ow_max];
    {
      constexpr int ext[3] = {bx, by, bz};
      int           d = 0, l = tid;
      while (d < 3 && l >= ext[d])
        l -= ext[d++];
      if (d < 3)
        axis_table[d][l] = d < dim ? cellblock::staging_entry(g, d, degree * c0[d] + l) : 0ull;
    }
    {
      constexpr int ext[3]  = {ox, oy, oz};
      constexpr int cstr[3] = {1, E0, E0 * E1};
      constexpr int kstr[3] = {cb_threads, P1 * cb_threads, P1 * P1 * cb_threads};
      int           d = 0, j = tid;
      while (d < 3 && j >= ext[d])
        j -= ext[d++];
      if (d < 3)
        {
          GatherAxis e {};
          if (d < dim)
            {
              const int n0 = degree * bc[d] * cb.bcell[d];
              const int n1 = (bc[d] + 1) * cb.bcell[d] >= g.cells[d]
                               ? g.nodes[d]
                               : degree * (bc[d] + 1) * cb.bcell[d];
              e = cellblock::gather_entry<degree>(g, d, n0 + j, n0 + j < n1, c0[d], cstr[d],
                                                  kstr[d]);
            }
          else
            {
              e.slots = 1;
              e.mass  = 1.0;
            }
          gather_table[d][j] = e;
        }
    }
    __syncthreads();

    // 1. Stage the inputs over the node box.
    cellblock::stage<bx, by, bz, ax_max>(lds, axis_table, deps, cb, tid);
    __syncthreads();

    // 2. This thread's cell.
    const int ec[3] = {tid % E0, (tid / E0) % E1, tid / (E0 * E1)};
    int       gc[3];
    bool      valid = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      gc[d] = d < dim ? cellblock::cell_coordinate(g, d, c0[d] + ec[d], valid) : 0;
    number acc[PPF_MAX_FIELDS][NPC];
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
      for (int k = 0; k < NPC; ++k)
        acc[f][k] = number(0);
    if (valid)
      cellblock::integrate_cell<PDE, dim, degree, number, lhs, block_id>(
        pde, g, timer, lds, box_n, cb,
        degree * ec[0] + bx * (degree * ec[1] + by * degree * ec[2]), bx, bx * by, gc, acc);

    // 3. Per output field: contributions to shared memory, owned nodes gathered.  Per
    //    axis a table gives the two cell slots' shared-memory offsets (an absent slot
    //    reads a zero entry, which leaves the sum unchanged), the 1-D lumped mass, the
    //    Dirichlet face bits and the global index part.
    __syncthreads(); // every cell evaluated: the staged inputs are no longer read
    constexpr int zero_at = NPC * cb_threads;
#pragma unroll
    for (int f = 0; f < PPF_MAX_FIELDS; ++f)
      {
        if (out.scratch_slot[f] < 0)
          continue;
#pragma unroll
        for (int k = 0; k < NPC; ++k)
          lds[k * cb_threads + tid] = acc[f][k];
        if (tid == 0)
          lds[zero_at] = number(0);
        __syncthreads();
        const cellblock::FieldOut<number> fo(out, f);
        for (int c = 0; c < n_cand; ++c)
          {
            const int        i  = tid + c * cb_threads;
            const int        ii = i < ox * oy * oz ? i : 0;
            const GatherAxis t0 = gather_table[0][ii % ox];
            const GatherAxis t1 = gather_table[1][(ii / ox) % oy];
            const GatherAxis t2 = gather_table[2][ii / (ox * oy)];
            if (!((i < ox * oy * oz) & (t0.goff >= 0) & (t1.goff >= 0) & (t2.goff >= 0)))
              continue;
            number sum = number(0);
#pragma unroll
            for (int a2 = 0; a2 < 2; ++a2)
#pragma unroll
              for (int a1 = 0; a1 < 2; ++a1)
#pragma unroll
                for (int a0 = 0; a0 < 2; ++a0)
                  {
                    const unsigned int present = (t0.slots >> a0) & (t1.slots >> a1) &
                                                 (t2.slots >> a2) & 1u;
                    sum += lds[present ? t0.a[a0] + t1.a[a1] + t2.a[a2] : zero_at];
                  }
            fo.write(sum, t0.faces | t1.faces | t2.faces, t0.mass * t1.mass * t2.mass,
                     t0.goff + t1.goff + t2.goff);
          }
        __syncthreads(); // contributions of this field no longer read
      }
  }

  /// 3-D generic kernel: a column of 9 x 9 cells (owning the nodes of its last 8 x 8)
  /// marching up a chunk of z layers [z0, z0 + chunk) in slabs of 3 layers.  The first
  /// slab starts one layer below the chunk.  Per slab, node planes p zs .. p zs + 3p:
  /// the bottom plane adds the upper cells to the sums kept from the previous slab (in
  /// shared memory, `carry`), the top plane's lower-cell sums are kept for the next
  /// slab, and the planes between are complete.  Planes outside the chunk's
  /// [p z0, p z1) (plus the closing boundary plane of a non-periodic last chunk) are
  /// not written.  Grid: x tiles * y tiles * z chunks.
  template <class PDE, int degree, typename number, bool lhs, int block_id>
  __global__ void __launch_bounds__(cb_threads)
  cell_column_kernel(const PDE pde, const GridDesc g, const DepTable<number> deps,
                     const OutTable<number> out, const prismspf::SimulationTimer timer,
                     const CellBlockArgs cb)
  {
    constexpr int dim = 3;
    constexpr int NPC = dofs_per_cell<dim, degree>;
    constexpr int P1  = degree + 1;
    constexpr int bx = degree * cc_ex + 1, by = degree * cc_ey + 1, bz = degree * cc_ez + 1;
    constexpr int box_n  = bx * by * bz;
    constexpr int ax_max = bx > bz ? bx : bz;
    constexpr int ox = degree * (cc_ex - 1) + 1, oy = degree * (cc_ey - 1) + 1;
    constexpr int n_cols  = ox * oy;
    constexpr int zero_at = NPC * cb_threads;
    constexpr int kz      = P1 * P1 * cb_threads;  // local-node stride along z
    constexpr int cz      = cc_ex * cc_ey;         // cell stride along z
    extern __shared__ unsigned char cb_shared[];
    number *lds = reinterpret_cast<number *>(cb_shared);
    // buffer A: the staged box, then contributions; buffer B (cb.pair): contributions
    const int region = (cb.n_in * box_n > zero_at + 1 ? cb.n_in * box_n : zero_at + 1);
    number   *lds_b  = lds + region;
    number   *carry  = lds_b + (cb.pair ? zero_at + 1 : 0); // [output slot][column]

    if (halted(out.halt))
      return;
    const int tid   = threadIdx.x;
    const int tile  = blockIdx.x % (cb.nblk[0] * cb.nblk[1]);
    const int chunk = blockIdx.x / (cb.nblk[0] * cb.nblk[1]);
    const int tx = tile % cb.nblk[0], ty = tile / cb.nblk[0];
    const int c0x = tx * (cc_ex - 1) - 1, c0y = ty * (cc_ey - 1) - 1;
    const int z0  = chunk * cb.bcell[2];
    const int z1  = min(z0 + cb.bcell[2], g.cells[2]);
    const int pz0 = degree * z0, pz1 = degree * z1;
    // the closing boundary plane belongs to the last chunk of a non-periodic z axis
    const int  pz_close = (!g.periodic[2] && z1 == g.cells[2]) ? g.nodes[2] - 1 : INT_MIN;
    const int64_t zstride = int64_t(g.nodes[0]) * g.nodes[1];

    __shared__ unsigned long long axis_table[3][ax_max];
    __shared__ GatherAxis         gather_table[2][ox > oy ? ox : oy];
    {
      // x/y staging and gather entries, fixed for the whole column
      if (tid < bx)
        axis_table[0][tid] = cellblock::staging_entry(g, 0, degree * c0x + tid);
      else if (tid < bx + by)
        axis_table[1][tid - bx] = cellblock::staging_entry(g, 1, degree * c0y + tid - bx);
      const int j = tid - (bx + by);
      if (j >= 0 && j < ox + oy)
        {
          const int d  = j < ox ? 0 : 1;
          const int jj = j < ox ? j : j - ox;
          const int t  = d == 0 ? tx : ty;
          const int n0 = degree * t * (cc_ex - 1);
          const int n1 = (t + 1) * (cc_ex - 1) >= g.cells[d] ? g.nodes[d]
                                                             : degree * (t + 1) * (cc_ex - 1);
          gather_table[d][jj] = cellblock::gather_entry<degree>(
            g, d, n0 + jj, n0 + jj < n1, d == 0 ? c0x : c0y, d == 0 ? 1 : cc_ex,
            d == 0 ? cb_threads : P1 * cb_threads);
        }
    }

    const int ec[3] = {tid % cc_ex, (tid / cc_ex) % cc_ey, tid / (cc_ex * cc_ey)};
    for (int zs = z0 - 1; zs < z1; zs += cc_ez)
      {
        const bool first = zs == z0 - 1;
        if (tid < bz)
          axis_table[2][tid] = cellblock::staging_entry(g, 2, degree * zs + tid);
        __syncthreads(); // tables written; the previous slab's contributions read

        // 1. Stage the slab's node box.
        cellblock::stage<bx, by, bz, ax_max>(lds, axis_table, deps, cb, tid);
        __syncthreads();

        // 2. This thread's cell.
        number acc[PPF_MAX_FIELDS][NPC];
#pragma unroll
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
#pragma unroll
          for (int k = 0; k < NPC; ++k)
            acc[f][k] = number(0);
        {
          bool valid = tid < cc_ex * cc_ey * cc_ez;
          int  gc[3];
          gc[0] = cellblock::cell_coordinate(g, 0, c0x + ec[0], valid);
          gc[1] = cellblock::cell_coordinate(g, 1, c0y + ec[1], valid);
          gc[2] = cellblock::cell_coordinate(g, 2, zs + ec[2], valid);
          if (valid)
            cellblock::integrate_cell<PDE, dim, degree, number, lhs, block_id>(
              pde, g, timer, lds, box_n, cb,
              degree * ec[0] + bx * (degree * ec[1] + by * degree * ec[2]), bx, bx * by, gc,
              acc);
        }
        __syncthreads(); // every cell evaluated: the staged inputs are no longer read

        // 3. Per output field: contributions to a shared-memory buffer, then every owned
        //    column of the tile walks the slab's node planes.  With a second buffer
        //    (cb.pair) the fields go two at a time: fewer barriers, two independent
        //    gathers per column, and the integrated contributions leave registers early.
        const auto gather = [&](const cellblock::FieldOut<number> &fo, number *kept_base,
                                const number *buf, int col, const GatherAxis &t0,
                                const GatherAxis &t1) {
          int          xy[4];
          unsigned int pres[4];
#pragma unroll
          for (int a1 = 0; a1 < 2; ++a1)
#pragma unroll
            for (int a0 = 0; a0 < 2; ++a0)
              {
                xy[2 * a1 + a0]   = t0.a[a0] + t1.a[a1];
                pres[2 * a1 + a0] = (t0.slots >> a0) & (t1.slots >> a1) & 1u;
              }
          number &kept = kept_base[col];
          // The sums of all planes first (every read issued without a branch), each in
          // node_gather_kernel's order: lower cell (or the kept sum), then upper cell.
          constexpr int NZ = degree * cc_ez + 1;
          number        part[NZ];
#pragma unroll
          for (int lz = 0; lz < NZ; ++lz)
            {
              number sum = number(0);
              if (lz == 0)
                sum = first ? number(0) : kept;
              else
                {
                  const int o = (degree == 2 && (lz & 1)) ? ((lz - 1) / 2) * cz + kz
                                                          : (lz / degree - 1) * cz + degree * kz;
#pragma unroll
                  for (int s = 0; s < 4; ++s)
                    sum += buf[pres[s] ? xy[s] + o : zero_at];
                }
              if (lz < NZ - 1 && !(degree == 2 && (lz & 1)))
                {
                  const int o = (lz / degree) * cz;
#pragma unroll
                  for (int s = 0; s < 4; ++s)
                    sum += buf[pres[s] ? xy[s] + o : zero_at];
                }
              part[lz] = sum;
            }
          const int64_t      goff_xy  = t0.goff + t1.goff;
          const double       mass_xy  = t0.mass * t1.mass;
          const unsigned int faces_xy = t0.faces | t1.faces;
#pragma unroll
          for (int lz = 0; lz < NZ; ++lz)
            {
              const int  P     = degree * zs + lz;
              const bool owned = (P >= pz0 && P < pz1) || P == pz_close;
              if (!owned)
                continue;
              if (lz == NZ - 1 && P != pz_close)
                {
                  kept = part[lz]; // completed by the next slab
                  continue;
                }
              unsigned int faces = faces_xy;
              if (!g.periodic[2])
                faces |= (P == 0 ? 1u << 4 : 0u) | (P == g.nodes[2] - 1 ? 1u << 5 : 0u);
              fo.write(part[lz], faces, mass_xy * lumped_mass_1d<degree>(g, 2, P),
                       goff_xy + int64_t(P) * zstride);
            }
        };
        // fa: field whose contributions wait in buffer A (-1: none)
        int fa = -1;
#pragma unroll
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
          {
            if (out.scratch_slot[f] < 0)
              continue;
            const bool to_b = fa >= 0; // A is taken: this field goes to B, gather both
            number    *buf  = to_b ? lds_b : lds;
#pragma unroll
            for (int k = 0; k < NPC; ++k)
              buf[k * cb_threads + tid] = acc[f][k];
            if (tid == 0)
              buf[zero_at] = number(0);
            const bool more = (out.out_mask >> (f + 1)) != 0u;
            if (!to_b && cb.pair && more)
              {
                fa = f; // gathered together with the next output field
                continue;
              }
            __syncthreads();
            const cellblock::FieldOut<number> fo(out, f);
            const cellblock::FieldOut<number> fo_a(out, to_b ? fa : f);
            number *const kept   = carry + out.scratch_slot[f] * n_cols;
            number *const kept_a = carry + out.scratch_slot[to_b ? fa : f] * n_cols;
            for (int col = tid; col < n_cols; col += cb_threads)
              {
                const GatherAxis t0 = gather_table[0][col % ox];
                const GatherAxis t1 = gather_table[1][col / ox];
                if (t0.goff < 0 || t1.goff < 0)
                  continue;
                if (to_b)
                  gather(fo_a, kept_a, lds, col, t0, t1);
                gather(fo, kept, buf, col, t0, t1);
              }
            fa = -1;
            __syncthreads(); // the buffers are free again
          }
      }
  }

  /// Shared memory of a cell_block_kernel (dim < 3) or cell_column_kernel (dim 3)
  /// launch in bytes, for `n_in` staged inputs and `n_out` output fields.
  template <int dim, int degree, typename number>
  inline int64_t
  cell_block_shared_bytes(int n_in, int n_out, bool pair)
  {
    int64_t box = 1;
    for (int d = 0; d < dim; ++d)
      box *= int64_t(degree) * (dim == 3 ? (d == 2 ? cc_ez : cc_ex) : cb_extent<dim>(d)) + 1;
    const int64_t contrib = int64_t(cb_threads) * dofs_per_cell<dim, degree> + 1; // + zero entry
    int64_t       n       = std::max(int64_t(n_in) * box, contrib);
    if (dim == 3)
      n += (pair ? contrib : 0) +
           int64_t(n_out) * (degree * (cc_ex - 1) + 1) * (degree * (cc_ey - 1) + 1);
    return n * int64_t(sizeof(number));
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
