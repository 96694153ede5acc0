// CPU oracle: an FP64 restatement of PRISMS-PF's matrix-free path on a uniform mesh.
//
// TEST INFRASTRUCTURE.  Only the tests, the smoke check and the benchmark's CPU baseline
// use this code; the product never links it.
//
// It follows the reference's semantics on deal.II's full DoF set (periodic images are
// separate DoFs constrained to their masters, Dirichlet DoFs carry inhomogeneous
// constraints), independently of the HIP implementation:
//   * 1-D GLL data (points, weights, Lagrange values and derivatives) are derived
//     numerically from the Lagrange polynomials, not tabulated;
//   * cell loop = FEEvaluation sum factorization: gather (plain for current/old states,
//     constraint-resolving for the src operand unless read_plain), evaluate values and
//     gradients at the Gauss-Lobatto points, call the hook, integrate, scatter through
//     the constraints (images forward to masters, Dirichlet rows are dropped)
//     (reference: src/solvers/mf_operator.cc:17-80, include/prismspf/core/
//     field_container.h:753-882);
//   * invM = 1/jxw from a cell loop integrating 1, zero on constrained DoFs
//     (reference include/prismspf/core/invm_manager.h:121-239);
//   * explicit update, distribute (images := masters, Dirichlet := g), increment loop,
//     old-state rotation (reference explicit_solver.h:63-77, solver_base.h:106-130,
//     src/core/problem.cc:340-558, src/core/group_solution_handler.cc:333-403);
//   * linear solve with the inhomogeneous-Dirichlet lift, deal.II SolverCG recurrences,
//     Chebyshev (Lanczos eigenvalue estimate) and a geometric multigrid V-cycle
//     (reference linear_solver.h:47-427);
//   * the Newton loop of Newton blocks on the same operators and CG (reference
//     newton_solver.h:73-156);
//   * the explicit Allen-Cahn adjoint (no reference counterpart).
// Cells are processed in 2^dim parity colours with OpenMP, so scatters never race.
#pragma once

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/pde_operator_base.h>
#include <prismspf/core/solve_block.h>
#include <prismspf/user_inputs/user_input_parameters.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#  include <omp.h>
#endif

namespace oracle
{
  template <class A, class = void>
  struct has_energy : std::false_type
  {};
  template <class A>
  struct has_energy<A, std::void_t<typename A::FreeEnergy>> : std::true_type
  {};
  /// Applications whose adjoint is the explicit Cahn-Hilliard step.
  template <class A, class = void>
  struct is_cahn_hilliard_adjoint : std::false_type
  {};
  template <class A>
  struct is_cahn_hilliard_adjoint<A, std::void_t<decltype(A::cahn_hilliard_adjoint)>>
    : std::bool_constant<A::cahn_hilliard_adjoint>
  {};

  using prismspf::DependencyMap;
  using prismspf::SolveBlock;

  // -------------------------------------------------------------------- 1-D basis
  struct Basis1D
  {
    int                              p;
    std::vector<double>              xi, w;
    std::vector<std::vector<double>> val, der; // [q][k]

    explicit Basis1D(int p_)
      : p(p_)
    {
      // Gauss-Lobatto points on [0,1]: the end points plus the roots of P'_p (p <= 2).
      if (p == 1)
        xi = {0.0, 1.0};
      else if (p == 2)
        xi = {0.0, 0.5, 1.0};
      else
        throw std::invalid_argument("oracle supports degree 1 and 2");
      // Weights: exact integrals of the Lagrange polynomials (5-point Gauss-Legendre).
      static const double gx[5] = {-0.9061798459386640, -0.5384693101056831, 0.0,
                                   0.5384693101056831, 0.9061798459386640};
      static const double gw[5] = {0.2369268850561891, 0.4786286704993665, 0.5688888888888889,
                                   0.4786286704993665, 0.2369268850561891};
      w.assign(p + 1, 0.0);
      for (int k = 0; k <= p; ++k)
        for (int g = 0; g < 5; ++g)
          w[k] += 0.5 * gw[g] * lagrange(k, 0.5 * (gx[g] + 1.0));
      val.assign(p + 1, std::vector<double>(p + 1));
      der.assign(p + 1, std::vector<double>(p + 1));
      for (int q = 0; q <= p; ++q)
        for (int k = 0; k <= p; ++k)
          {
            val[q][k] = lagrange(k, xi[q]);
            der[q][k] = lagrange_derivative(k, xi[q]);
          }
    }

    double
    lagrange(int k, double x) const
    {
      double r = 1.0;
      for (int j = 0; j <= p; ++j)
        if (j != k)
          r *= (x - xi[j]) / (xi[k] - xi[j]);
      return r;
    }

    double
    lagrange_derivative(int k, double x) const
    {
      double s = 0.0;
      for (int m = 0; m <= p; ++m)
        {
          if (m == k)
            continue;
          double t = 1.0 / (xi[k] - xi[m]);
          for (int j = 0; j <= p; ++j)
            if (j != k && j != m)
              t *= (x - xi[j]) / (xi[k] - xi[j]);
          s += t;
        }
      return s;
    }
  };

  // ------------------------------------------------------------------------ mesh
  struct Mesh
  {
    int     dim = 0, p = 0;
    int     cells[3] = {1, 1, 1};
    int     N[3]     = {1, 1, 1}; // p*cells+1 lattice nodes per direction (images included)
    bool    periodic[3] = {false, false, false};
    double  lower[3] = {0, 0, 0}, h[3] = {1, 1, 1};
    int64_t ndofs = 1, ncells = 1;

    int64_t
    index(int i0, int i1, int i2) const
    {
      return i0 + int64_t(N[0]) * (i1 + int64_t(N[1]) * i2);
    }

    void
    coords(int64_t dof, int (&i)[3]) const
    {
      i[0] = int(dof % N[0]);
      i[1] = int((dof / N[0]) % N[1]);
      i[2] = int(dof / (int64_t(N[0]) * N[1]));
    }
  };

  enum Kind : int8_t
  {
    free_dof  = 0,
    image_dof = 1,
    dirichlet = 2
  };

  /// Constraints of one field on one mesh: images point at their master (the final
  /// target after resolving chains), Dirichlet DoFs carry their value.
  struct Constraints
  {
    std::vector<int8_t>  kind;
    std::vector<int64_t> target;
    std::vector<double>  g;
  };

  // ------------------------------------------------------------------- the model
  template <class App>
  class Model
  {
  public:
    using AppType               = App;
    static constexpr int dim    = App::dim;
    static constexpr int degree = App::degree;
    using PDE                   = typename App::template PDE<double>;
    using FC                    = prismspf::FieldContainer<dim, degree, double>;
    static constexpr int NPC    = dim == 2 ? (degree + 1) * (degree + 1)
                                           : (degree + 1) * (degree + 1) * (degree + 1);

    explicit Model(const std::string &prm)
      : inputs(prismspf::UserInputParameters<dim>::from_text(prm))
      , basis(degree)
    {
      App::define(fields, blocks);
      prismspf::validate_solve_blocks(blocks, fields);
      for (auto &b : blocks)
        {
          auto it = inputs.linear_solve_parameters.find(b.id);
          if (it != inputs.linear_solve_parameters.end())
            b.linear_solver_parameters = it->second;
          const auto nit = inputs.nonlinear_solve_parameters.find(b.id);
          if (nit != inputs.nonlinear_solve_parameters.end())
            b.nonlinear_solver_parameters = nit->second;
        }
      for (auto &fa : fields)
        {
          auto it = inputs.boundary_parameters.find(fa.name);
          if (it != inputs.boundary_parameters.end())
            fa.boundary_conditions = it->second;
        }
      // A vector field is stored as dim scalar component fields (declared last).
      layout = prismspf::expand_vector_fields(fields, blocks, dim);
      pde   = std::make_unique<PDE>(inputs, tools);
      timer = prismspf::SimulationTimer(inputs.temporal_discretization.dt,
                                        inputs.temporal_discretization.initial_time);
      top   = make_mesh(inputs.spatial_discretization.global_refinement);
      for (std::size_t f = 0; f < fields.size(); ++f)
        cons.push_back(make_constraints(top, int(f), false));
      jxw = lumped_mass(top);
      // storage: 1 + oldest age read
      std::vector<unsigned> max_age(fields.size(), 0);
      for (const auto &b : blocks)
        for (const auto *deps : {&b.dependencies_rhs, &b.dependencies_lhs})
          for (const auto &[f, dep] : *deps)
            max_age[f] = std::max(max_age[f], dep.oldest_age());
      storage.resize(fields.size());
      for (std::size_t f = 0; f < fields.size(); ++f)
        storage[f].assign(1 + max_age[f], std::vector<double>(top.ndofs, 0.0));
    }

    /// The application object seen through its base class: hooks are dispatched
    /// virtually, as the reference does (mf_operator.cc:66).
    const prismspf::PDEOperatorBase<dim, degree, double> &
    hooks() const
    {
      return *pde;
    }

    // ------------------------------------------------------------ construction
    Mesh
    make_mesh(unsigned refinement) const
    {
      const auto &rm = inputs.spatial_discretization.rectangular_mesh;
      Mesh        m;
      m.dim = dim;
      m.p   = degree;
      for (int d = 0; d < dim; ++d)
        {
          m.cells[d]    = int(rm.subdivisions[d] << refinement);
          m.N[d]        = degree * m.cells[d] + 1;
          m.periodic[d] = rm.periodic_directions.count(d) != 0;
          m.lower[d]    = rm.lower_bound[d];
          m.h[d]        = (rm.size[d] - rm.lower_bound[d]) / m.cells[d];
        }
      m.ndofs  = int64_t(m.N[0]) * m.N[1] * m.N[2];
      m.ncells = int64_t(m.cells[0]) * m.cells[1] * m.cells[2];
      return m;
    }

    unsigned
    dirichlet_mask(int f) const
    {
      unsigned m = 0;
      for (const auto &[comp, conds] : fields[f].boundary_conditions.component_constraints)
        for (const auto &[bid, c] : conds)
          if (c == prismspf::Condition::Dirichlet && bid < 2u * dim)
            m |= 1u << bid;
      return m;
    }

    /// Periodicity first, then Dirichlet on faces in ascending boundary id, an already
    /// constrained DoF keeping its first constraint (deal.II semantics).
    Constraints
    make_constraints(const Mesh &m, int f, bool homogeneous) const
    {
      Constraints c;
      c.kind.assign(m.ndofs, free_dof);
      c.target.assign(m.ndofs, -1);
      c.g.assign(m.ndofs, 0.0);
      for (int64_t dof = 0; dof < m.ndofs; ++dof)
        {
          int i[3];
          m.coords(dof, i);
          bool img = false;
          int  j[3] = {i[0], i[1], i[2]};
          for (int d = 0; d < dim; ++d)
            if (m.periodic[d] && i[d] == m.N[d] - 1)
              {
                img  = true;
                j[d] = 0;
              }
          if (img)
            {
              c.kind[dof]   = image_dof;
              c.target[dof] = m.index(j[0], j[1], j[2]);
            }
        }
      const unsigned mask = dirichlet_mask(f);
      for (int64_t dof = 0; dof < m.ndofs; ++dof)
        {
          if (c.kind[dof] != free_dof)
            continue;
          int i[3];
          m.coords(dof, i);
          for (unsigned bid = 0; bid < 2u * dim; ++bid)
            {
              const int d = int(bid / 2);
              if (!(mask & (1u << bid)) || m.periodic[d])
                continue;
              const bool on = (bid % 2 == 0) ? i[d] == 0 : i[d] == m.N[d] - 1;
              if (!on)
                continue;
              c.kind[dof] = dirichlet;
              if (!homogeneous)
                {
                  dealii::Point<dim> pt;
                  for (int e = 0; e < dim; ++e)
                    pt[e] = m.lower[e] + i[e] * m.h[e] / degree;
                  double      v = 0.0, vec = 0.0;
                  const auto &lc = layout[f];
                  hooks().set_dirichlet(unsigned(lc.declared), bid, lc.component, pt, timer, v,
                                        vec);
                  c.g[dof] = lc.vector ? vec : v;
                }
              break;
            }
        }
      // Images of Dirichlet masters are Dirichlet with the master's value.
      for (int64_t dof = 0; dof < m.ndofs; ++dof)
        if (c.kind[dof] == image_dof && c.kind[c.target[dof]] == dirichlet)
          {
            c.kind[dof] = dirichlet;
            c.g[dof]    = c.g[c.target[dof]];
          }
      return c;
    }

    // --------------------------------------------------------------- cell loop
    using QFunc = std::function<void(FC &)>;

    struct Inputs
    {
      // per (field, slot): vector (full lattice) and flags; src is read through
      // constraints unless `plain`.
      const double *vec[PPF_MAX_FIELDS][PPF_N_SLOTS] = {};
      unsigned      flags[PPF_MAX_FIELDS][PPF_N_SLOTS] = {};
    };

    static constexpr unsigned f_values = 1, f_gradients = 2, f_resolve = 4;

    /// Tensor-product contraction of a (p+1)^dim array along direction d with M[q][k].
    void
    contract(const double *in, double *out, const std::vector<std::vector<double>> &M, int d,
             bool transpose) const
    {
      const int P1 = degree + 1;
      int       stride = 1;
      for (int e = 0; e < d; ++e)
        stride *= P1;
      for (int idx = 0; idx < NPC; ++idx)
        {
          const int kd   = (idx / stride) % P1;
          const int base = idx - kd * stride;
          double    s    = 0.0;
          for (int k = 0; k < P1; ++k)
            s += (transpose ? M[k][kd] : M[kd][k]) * in[base + k * stride];
          out[idx] = s;
        }
    }

    /// Gathers the cell's local values of every declared input (plain, or through the
    /// constraints for the src operand).
    void
    gather(const Mesh &m, const std::vector<Constraints> &cs, const Inputs &in,
           const int64_t (&dof)[NPC], std::vector<double> &u) const
    {
      u.assign(PPF_MAX_FIELDS * PPF_N_SLOTS * NPC, 0.0);
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        for (int s = 0; s < PPF_N_SLOTS; ++s)
          {
            const unsigned fl = in.flags[f][s];
            if (!(fl & (f_values | f_gradients)))
              continue;
            for (int k = 0; k < NPC; ++k)
              {
                const int64_t gdof = dof[k];
                double        v;
                if (!(fl & f_resolve))
                  v = in.vec[f][s][gdof];
                else
                  {
                    const auto &cf = cs[f];
                    if (cf.kind[gdof] == free_dof)
                      v = in.vec[f][s][gdof];
                    else if (cf.kind[gdof] == image_dof)
                      v = in.vec[f][s][cf.target[gdof]];
                    else
                      v = 0.0;
                  }
                u[(f * PPF_N_SLOTS + s) * NPC + k] = v;
              }
          }
      (void) m;
    }

    /// FEEvaluation evaluate -> hook at every quadrature point -> integrate, for one
    /// cell with local input values u; acc[o][k] receives the integrated terms of
    /// output o.
    void
    eval_cell(const Mesh &m, const int (&cc)[3], const Inputs &in, const std::vector<double> &u,
              const std::vector<int> &outs, unsigned writable, const QFunc &qf,
              std::vector<double> &acc) const
    {
      const int           P1 = degree + 1;
      std::vector<double> val(PPF_MAX_FIELDS * PPF_N_SLOTS * NPC, 0.0);
      std::vector<double> grad(PPF_MAX_FIELDS * PPF_N_SLOTS * dim * NPC, 0.0);
      double              t1[NPC], t2[NPC];
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        for (int s = 0; s < PPF_N_SLOTS; ++s)
          {
            if (!(in.flags[f][s] & (f_values | f_gradients)))
              continue;
            const double *uu = &u[(f * PPF_N_SLOTS + s) * NPC];
            // values: interpolation (identity at the GLL points, applied anyway)
            std::copy(uu, uu + NPC, t1);
            for (int d = 0; d < dim; ++d)
              {
                contract(t1, t2, basis.val, d, false);
                std::copy(t2, t2 + NPC, t1);
              }
            std::copy(t1, t1 + NPC, &val[(f * PPF_N_SLOTS + s) * NPC]);
            for (int gd = 0; gd < dim; ++gd)
              {
                std::copy(uu, uu + NPC, t1);
                for (int d = 0; d < dim; ++d)
                  {
                    contract(t1, t2, d == gd ? basis.der : basis.val, d, false);
                    std::copy(t2, t2 + NPC, t1);
                  }
                for (int q = 0; q < NPC; ++q)
                  grad[((f * PPF_N_SLOTS + s) * dim + gd) * NPC + q] = t1[q] / m.h[gd];
              }
          }
      std::vector<double> qv(outs.size() * NPC, 0.0), qg(outs.size() * dim * NPC, 0.0);
      double              vol = 1.0;
      for (int d = 0; d < dim; ++d)
        vol *= m.h[d];
      for (int q = 0; q < NPC; ++q)
        {
          int qq[3] = {q % P1, (q / P1) % P1, q / (P1 * P1)};
          FC  fc;
          fc.writable       = writable;
          fc.element_volume = vol;
          for (int d = 0; d < dim; ++d)
            fc.q_point[d] = m.lower[d] + (cc[d] + basis.xi[qq[d]]) * m.h[d];
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
            for (int s = 0; s < PPF_N_SLOTS; ++s)
              {
                const unsigned fl        = in.flags[f][s];
                fc.avail_values[f][s]    = fl & f_values;
                fc.avail_gradients[f][s] = fl & f_gradients;
                fc.values[f][s]          = val[(f * PPF_N_SLOTS + s) * NPC + q];
                for (int d = 0; d < dim; ++d)
                  fc.gradients[f][s][d] = grad[((f * PPF_N_SLOTS + s) * dim + d) * NPC + q];
              }
          qf(fc);
          double wq = vol;
          for (int d = 0; d < dim; ++d)
            wq *= basis.w[qq[d]];
          for (std::size_t o = 0; o < outs.size(); ++o)
            {
              qv[o * NPC + q] = wq * fc.value_term[outs[o]];
              for (int d = 0; d < dim; ++d)
                qg[(o * dim + d) * NPC + q] = wq * fc.gradient_term[outs[o]][d] / m.h[d];
            }
        }
      acc.assign(outs.size() * NPC, 0.0);
      for (std::size_t o = 0; o < outs.size(); ++o)
        {
          std::copy(&qv[o * NPC], &qv[o * NPC] + NPC, t1);
          for (int d = 0; d < dim; ++d)
            {
              contract(t1, t2, basis.val, d, true);
              std::copy(t2, t2 + NPC, t1);
            }
          for (int k = 0; k < NPC; ++k)
            acc[o * NPC + k] = t1[k];
          for (int gd = 0; gd < dim; ++gd)
            {
              std::copy(&qg[(o * dim + gd) * NPC], &qg[(o * dim + gd) * NPC] + NPC, t1);
              for (int d = 0; d < dim; ++d)
                {
                  contract(t1, t2, d == gd ? basis.der : basis.val, d, true);
                  std::copy(t2, t2 + NPC, t1);
                }
              for (int k = 0; k < NPC; ++k)
                acc[o * NPC + k] += t1[k];
            }
        }
    }

    /// dst_o = assembled terms of qf for the outputs `outs`.  With `probe_src >= 0`
    /// computes instead the operator diagonal of output `probe_src` the way the
    /// reference probes it (mf_operator.cc:134-218): per cell, unit local values on the
    /// src slot of that field, keeping the j-th integrated entry, distributed through
    /// the constraints.
    void
    cell_loop(const Mesh &m, const std::vector<Constraints> &cs, const Inputs &in,
              const std::vector<int> &outs, std::vector<double *> &dst, const QFunc &qf,
              int probe_src = -1) const
    {
      for (std::size_t o = 0; o < outs.size(); ++o)
        std::fill(dst[o], dst[o] + m.ndofs, 0.0);
      const int P1       = degree + 1;
      unsigned  writable = 0;
      for (int f : outs)
        writable |= 1u << f;
      for (int d = 0; d < dim; ++d)
        if (m.periodic[d] && (m.cells[d] % 2))
          throw std::invalid_argument("oracle: periodic directions need an even cell count");
      for (int colour = 0; colour < (1 << dim); ++colour)
        {
          std::vector<int64_t> cells;
          for (int64_t c = 0; c < m.ncells; ++c)
            {
              const int c0 = int(c % m.cells[0]), c1 = int((c / m.cells[0]) % m.cells[1]),
                        c2 = int(c / (int64_t(m.cells[0]) * m.cells[1]));
              if (((c0 & 1) | ((c1 & 1) << 1) | ((c2 & 1) << 2)) == colour)
                cells.push_back(c);
            }
#pragma omp parallel for schedule(static)
          for (int64_t ci = 0; ci < int64_t(cells.size()); ++ci)
            {
              const int64_t c     = cells[ci];
              int           cc[3] = {int(c % m.cells[0]), int((c / m.cells[0]) % m.cells[1]),
                                     int(c / (int64_t(m.cells[0]) * m.cells[1]))};
              int64_t       dof[NPC];
              for (int k = 0; k < NPC; ++k)
                {
                  int kk[3] = {k % P1, (k / P1) % P1, k / (P1 * P1)};
                  int i[3];
                  for (int d = 0; d < 3; ++d)
                    i[d] = d < dim ? degree * cc[d] + kk[d] : 0;
                  dof[k] = m.index(i[0], i[1], i[2]);
                }
              std::vector<double> u, acc;
              gather(m, cs, in, dof, u);
              std::vector<double> local(outs.size() * NPC, 0.0);
              if (probe_src < 0)
                {
                  eval_cell(m, cc, in, u, outs, writable, qf, acc);
                  local = acc;
                }
              else
                {
                  const int o = int(std::find(outs.begin(), outs.end(), probe_src) - outs.begin());
                  for (int j = 0; j < NPC; ++j)
                    {
                      for (int f : outs)
                        for (int k = 0; k < NPC; ++k)
                          u[(f * PPF_N_SLOTS + 0) * NPC + k] =
                            (f == probe_src && k == j) ? 1.0 : 0.0;
                      eval_cell(m, cc, in, u, outs, writable, qf, acc);
                      local[o * NPC + j] = acc[o * NPC + j];
                    }
                }
              for (std::size_t o = 0; o < outs.size(); ++o)
                {
                  const auto &cf = cs[outs[o]];
                  for (int k = 0; k < NPC; ++k)
                    {
                      const int64_t gdof = dof[k];
                      if (cf.kind[gdof] == free_dof)
                        dst[o][gdof] += local[o * NPC + k];
                      else if (cf.kind[gdof] == image_dof)
                        dst[o][cf.target[gdof]] += local[o * NPC + k];
                    }
                }
            }
        }
    }

    /// jxw of every DoF: the integral of its basis function, assembled with the
    /// periodicity constraints only (the reference's "generic" constraints,
    /// src/core/constraint_manager.cc:152-173), so images carry zero and Dirichlet DoFs
    /// keep their mass.
    std::vector<double>
    lumped_mass(const Mesh &m) const
    {
      Constraints periodic;
      periodic.kind.assign(m.ndofs, free_dof);
      periodic.target.assign(m.ndofs, -1);
      periodic.g.assign(m.ndofs, 0.0);
      for (int64_t dof = 0; dof < m.ndofs; ++dof)
        {
          int i[3];
          m.coords(dof, i);
          bool img = false;
          for (int d = 0; d < dim; ++d)
            if (m.periodic[d] && i[d] == m.N[d] - 1)
              {
                img  = true;
                i[d] = 0;
              }
          if (img)
            {
              periodic.kind[dof]   = image_dof;
              periodic.target[dof] = m.index(i[0], i[1], i[2]);
            }
        }
      std::vector<double>      out(m.ndofs);
      std::vector<double *>    dst {out.data()};
      std::vector<Constraints> cs(PPF_MAX_FIELDS, periodic);
      Inputs                   in;
      cell_loop(m, cs, in, {0}, dst, [](FC &fc) { fc.value_term[0] = 1.0; });
      return out;
    }

    static std::vector<double>
    scaling_vector(const std::vector<double> &jx, int scaling)
    {
      std::vector<double> s(jx.size(), 1.0);
      for (std::size_t i = 0; i < jx.size(); ++i)
        {
          if (scaling == 0)
            s[i] = 1.0;
          else
            {
              const double inv = jx[i] == 0.0 ? 0.0 : 1.0 / jx[i];
              s[i]             = scaling == 1 ? inv : std::sqrt(inv);
            }
        }
      return s;
    }

    // ------------------------------------------------------------ operators
    struct Level
    {
      Mesh                                          mesh;
      std::vector<Constraints>                      cons;
      std::vector<double>                           jxw;
      std::vector<std::vector<std::vector<double>>> fields; // [f][age] for non-src deps
    };

    std::vector<int>
    block_fields(int b) const
    {
      return {blocks[b].field_indices.begin(), blocks[b].field_indices.end()};
    }

    /// Operator of block b (rhs or lhs) on a level; src/dst are block vectors
    /// (concatenated fields) on the level's full lattice.
    void
    apply(int b, bool lhs, const Mesh &m, const std::vector<Constraints> &cs,
          const std::vector<double> &jx,
          const std::vector<std::vector<std::vector<double>>> &level_fields,
          const double *src, double *dst, int scaling, bool read_plain) const
    {
      const SolveBlock &blk  = blocks[b];
      const auto        bf   = block_fields(b);
      const auto       &deps = lhs ? blk.dependencies_lhs : blk.dependencies_rhs;
      Inputs            in;
      for (const auto &[f, dep] : deps)
        {
          auto add = [&](int s, prismspf::EvalFlags ef, const double *v, bool resolve) {
            unsigned fl = 0;
            if (ef & dealii::EvaluationFlags::values)
              fl |= f_values;
            if (ef & dealii::EvaluationFlags::gradients)
              fl |= f_gradients;
            if (!fl)
              return;
            if (resolve)
              fl |= f_resolve;
            in.vec[f][s]   = v;
            in.flags[f][s] = fl;
          };
          if (dep.src_flag != prismspf::EvalFlags::nothing)
            {
              const auto pos = std::find(bf.begin(), bf.end(), int(f)) - bf.begin();
              add(0, dep.src_flag, src + pos * m.ndofs, !read_plain);
            }
          if (dep.flag != prismspf::EvalFlags::nothing)
            add(1, dep.flag, level_fields[f][0].data(), false);
          for (std::size_t a = 0; a < dep.old_flags.size(); ++a)
            if (dep.old_flags[a] != prismspf::EvalFlags::nothing)
              add(int(a) + 2, dep.old_flags[a], level_fields[f][a + 1].data(), false);
        }
      std::vector<double *> d;
      for (std::size_t k = 0; k < bf.size(); ++k)
        d.push_back(dst + k * m.ndofs);
      const int id = blk.id;
      cell_loop(m, cs, in, bf, d, [&](FC &fc) {
        if (lhs)
          hooks().compute_lhs(fc, timer, id);
        else
          hooks().compute_rhs(fc, timer, id);
      });
      const auto sc = scaling_vector(jx, scaling);
      for (std::size_t k = 0; k < bf.size(); ++k)
        for (int64_t i = 0; i < m.ndofs; ++i)
          d[k][i] *= sc[i];
    }

    std::vector<std::vector<std::vector<double>>> &
    top_fields()
    {
      return storage;
    }

    void
    apply_top(int b, bool lhs, const double *src, double *dst, int scaling, bool plain)
    {
      apply(b, lhs, top, cons, jxw, storage, src, dst, scaling, plain);
    }

    /// Sets images to their masters and Dirichlet DoFs to g (AffineConstraints::distribute).
    static void
    distribute(const Constraints &c, double *v, bool homogeneous = false)
    {
      for (std::size_t i = 0; i < c.kind.size(); ++i)
        if (c.kind[i] == dirichlet)
          v[i] = homogeneous ? 0.0 : c.g[i];
      for (std::size_t i = 0; i < c.kind.size(); ++i)
        if (c.kind[i] == image_dof)
          v[i] = v[c.target[i]];
    }

    // ------------------------------------------------------------------ solvers
    void
    solve_explicit(int b)
    {
      const auto          bf = block_fields(b);
      std::vector<double> out(bf.size() * top.ndofs);
      apply_top(b, false, nullptr, out.data(), 1, false);
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          auto &sol = storage[bf[k]][0];
          std::copy(out.begin() + k * top.ndofs, out.begin() + (k + 1) * top.ndofs, sol.begin());
          distribute(cons[bf[k]], sol.data());
        }
    }

    struct Stats
    {
      int    iterations = 0;
      double r0 = 0, r = 0, tol = 0;
      bool   converged = false;
    };

    static double
    dot(const std::vector<double> &a, const std::vector<double> &b)
    {
      // Fixed chunks summed in chunk order: the same result for any thread count and
      // every run (an OpenMP reduction combines thread partials in arrival order).
      constexpr int64_t   chunks = 64;
      const int64_t       n      = int64_t(a.size());
      std::vector<double> part(chunks, 0.0);
#pragma omp parallel for schedule(static)
      for (int64_t c = 0; c < chunks; ++c)
        {
          double        s  = 0.0;
          const int64_t lo = n * c / chunks, hi = n * (c + 1) / chunks;
          for (int64_t i = lo; i < hi; ++i)
            s += a[i] * b[i];
          part[c] = s;
        }
      double s = 0.0;
      for (double v : part)
        s += v;
      return s;
    }

    int
    scaling_of(int b) const
    {
      return blocks[b].linear_solver_parameters.tolerance_type != prismspf::AbsoluteResidual ? 2 : 0;
    }

    double
    normalization(int b) const
    {
      const auto t = blocks[b].linear_solver_parameters.tolerance_type;
      double     v = 1.0;
      if (t == prismspf::RMSEPerField || t == prismspf::RMSETotal)
        {
          double vol = 1.0;
          for (int d = 0; d < dim; ++d)
            vol *= top.h[d] * top.cells[d];
          v *= std::sqrt(vol);
        }
      if (t == prismspf::RMSEPerField || t == prismspf::IntegratedPerField)
        {
          // Declared fields, not stored components: a vector field counts once,
          // as in the reference's solve_block.field_indices.
          std::set<prismspf::Types::Index> declared;
          for (const auto f : blocks[b].field_indices)
            declared.insert(layout[f].declared);
          v *= std::sqrt(double(declared.size()));
        }
      return v;
    }

    /// Deterministic start vector of the eigenvalue estimate, keyed by (block field,
    /// stored node index with periodic images folded onto masters).
    static double
    start_value(uint64_t key)
    {
      uint64_t x = key + 0x9e3779b97f4a7c15ull;
      x          = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
      x          = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
      x ^= x >> 31;
      return double(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }

    std::vector<double>
    start_vector(const Mesh &m, int nblock) const
    {
      std::vector<double> v(nblock * m.ndofs);
      int                 stored[3];
      for (int d = 0; d < 3; ++d)
        stored[d] = d < dim ? (m.periodic[d] ? m.N[d] - 1 : m.N[d]) : 1;
      const int64_t nstored = int64_t(stored[0]) * stored[1] * stored[2];
      for (int k = 0; k < nblock; ++k)
        for (int64_t dof = 0; dof < m.ndofs; ++dof)
          {
            int i[3];
            m.coords(dof, i);
            for (int d = 0; d < dim; ++d)
              if (m.periodic[d] && i[d] == m.N[d] - 1)
                i[d] = 0;
            const int64_t key = i[0] + int64_t(stored[0]) * (i[1] + int64_t(stored[1]) * i[2]);
            v[k * m.ndofs + dof] = start_value(uint64_t(k * nstored + key));
          }
      return v;
    }

    struct Smoother
    {
      std::vector<double> inv_diag;
      double              lmax = 0, lmin = 0;
    };

    using Op = std::function<void(const std::vector<double> &, std::vector<double> &)>;

    /// CG-Lanczos estimate of the largest eigenvalue of P^{-1}A (safety factor 1.2).
    void
    estimate(const Op &A, Smoother &sm, const std::vector<double> &x0,
             const prismspf::ChebyshevParameters &cp) const
    {
      const std::size_t   n = x0.size();
      std::vector<double> r(n), z(n), p(n), ap(n);
      A(x0, r);
      for (std::size_t i = 0; i < n; ++i)
        z[i] = sm.inv_diag[i] * r[i];
      p         = z;
      double rz = dot(r, z);
      std::vector<double> al, be;
      for (unsigned it = 0; it < cp.eig_cg_n_iterations && rz > 0.0; ++it)
        {
          A(p, ap);
          const double pap = dot(p, ap);
          if (!(pap > 0.0))
            break;
          const double a = rz / pap;
          al.push_back(a);
          for (std::size_t i = 0; i < n; ++i)
            {
              r[i] -= a * ap[i];
              z[i] = sm.inv_diag[i] * r[i];
            }
          const double rzn = dot(r, z);
          const double bt  = rzn / rz;
          be.push_back(bt);
          rz = rzn;
          for (std::size_t i = 0; i < n; ++i)
            p[i] = z[i] + bt * p[i];
        }
      // Largest eigenvalue of the Lanczos matrix by power iteration on the dense copy.
      const int                        mlen = int(al.size());
      std::vector<std::vector<double>> T(mlen, std::vector<double>(mlen, 0.0));
      for (int i = 0; i < mlen; ++i)
        {
          T[i][i] = 1.0 / al[i] + (i > 0 ? be[i - 1] / al[i - 1] : 0.0);
          if (i + 1 < mlen)
            T[i][i + 1] = T[i + 1][i] = std::sqrt(std::max(0.0, be[i])) / al[i];
        }
      // Largest eigenvalue of the symmetric Lanczos matrix by cyclic Jacobi rotations.
      double lam = 1.0;
      if (mlen > 0)
        {
          for (int sweep = 0; sweep < 100; ++sweep)
            {
              double off = 0.0;
              for (int i = 0; i < mlen; ++i)
                for (int j = i + 1; j < mlen; ++j)
                  off += T[i][j] * T[i][j];
              if (off < 1e-30)
                break;
              for (int pi = 0; pi < mlen; ++pi)
                for (int qi = pi + 1; qi < mlen; ++qi)
                  {
                    if (std::abs(T[pi][qi]) < 1e-300)
                      continue;
                    const double theta = (T[qi][qi] - T[pi][pi]) / (2.0 * T[pi][qi]);
                    const double t     = (theta >= 0 ? 1.0 : -1.0) /
                                     (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                    const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                    for (int r = 0; r < mlen; ++r)
                      {
                        const double a1 = T[r][pi], a2 = T[r][qi];
                        T[r][pi]        = cs * a1 - sn * a2;
                        T[r][qi]        = sn * a1 + cs * a2;
                      }
                    for (int r = 0; r < mlen; ++r)
                      {
                        const double a1 = T[pi][r], a2 = T[qi][r];
                        T[pi][r]        = cs * a1 - sn * a2;
                        T[qi][r]        = sn * a1 + cs * a2;
                      }
                  }
            }
          lam = T[0][0];
          for (int i = 1; i < mlen; ++i)
            lam = std::max(lam, T[i][i]);
        }
      sm.lmax = 1.2 * lam;
      sm.lmin = sm.lmax / cp.smoothing_range;
    }

    void
    chebyshev(const Op &A, const Smoother &sm, unsigned deg, std::vector<double> &x,
              const std::vector<double> &rhs, bool zero_start) const
    {
      const std::size_t   n     = rhs.size();
      const double        theta = 0.5 * (sm.lmax + sm.lmin), delta = 0.5 * (sm.lmax - sm.lmin);
      const double        sigma = theta / delta;
      double              rho   = 1.0 / sigma;
      std::vector<double> d(n, 0.0), ax(n, 0.0);
      if (zero_start)
        std::fill(x.begin(), x.end(), 0.0);
      for (unsigned k = 0; k < deg; ++k)
        {
          if (k == 0 && zero_start)
            std::fill(ax.begin(), ax.end(), 0.0);
          else
            A(x, ax);
          double c1 = 0.0, c2;
          if (k == 0)
            c2 = 1.0 / theta;
          else
            {
              const double rn = 1.0 / (2.0 * sigma - rho);
              c1              = rn * rho;
              c2              = 2.0 * rn / delta;
              rho             = rn;
            }
          for (std::size_t i = 0; i < n; ++i)
            {
              d[i] = c1 * d[i] + c2 * sm.inv_diag[i] * (rhs[i] - ax[i]);
              x[i] += d[i];
            }
        }
    }

    /// Inverse of the lhs diagonal of block b, scaled like the operator, zero entries
    /// replaced by one (reference mf_operator.cc:82-255).
    std::vector<double>
    inverse_diagonal(int b, const Mesh &m, const std::vector<Constraints> &cs,
                     const std::vector<double> &jx,
                     const std::vector<std::vector<std::vector<double>>> &lf, int scaling) const
    {
      const SolveBlock   &blk = blocks[b];
      const auto          bf  = block_fields(b);
      std::vector<double> diag(bf.size() * m.ndofs, 0.0);
      Inputs              in;
      for (const auto &[f, dep] : blk.dependencies_lhs)
        {
          auto add = [&](int s, prismspf::EvalFlags ef, const double *v) {
            unsigned fl = 0;
            if (ef & dealii::EvaluationFlags::values)
              fl |= f_values;
            if (ef & dealii::EvaluationFlags::gradients)
              fl |= f_gradients;
            if (!fl)
              return;
            in.vec[f][s]   = v;
            in.flags[f][s] = fl;
          };
          // The probe overwrites the src slot's local values.
          if (dep.src_flag != prismspf::EvalFlags::nothing)
            add(0, dep.src_flag, diag.data());
          if (dep.flag != prismspf::EvalFlags::nothing)
            add(1, dep.flag, lf[f][0].data());
          for (std::size_t a = 0; a < dep.old_flags.size(); ++a)
            if (dep.old_flags[a] != prismspf::EvalFlags::nothing)
              add(int(a) + 2, dep.old_flags[a], lf[f][a + 1].data());
        }
      const auto sc = scaling_vector(jx, scaling);
      const int  id = blk.id;
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          std::vector<double>   out(bf.size() * m.ndofs, 0.0);
          std::vector<double *> d;
          for (std::size_t o = 0; o < bf.size(); ++o)
            d.push_back(out.data() + o * m.ndofs);
          cell_loop(m, cs, in, bf, d, [&](FC &fc) { hooks().compute_lhs(fc, timer, id); }, bf[k]);
          for (int64_t i = 0; i < m.ndofs; ++i)
            {
              const double v             = out[k * m.ndofs + i] * sc[i];
              diag[k * m.ndofs + i]      = v == 0.0 ? 1.0 : 1.0 / v;
            }
        }
      return diag;
    }

    // ------------------------------------------------------------ multigrid
    struct MGLevel
    {
      Level    lvl;
      Smoother sm;
    };

    /// Prolongation weights of fine lattice coordinate I from the coarse mesh (1-D).
    std::vector<std::pair<int, double>>
    prolong_1d(const Mesh &fine, const Mesh &coarse, int d, int I) const
    {
      const double x  = I * fine.h[d] / degree; // relative to the lower bound
      int          c  = int(std::floor(x / coarse.h[d] + 1e-12));
      c               = std::min(std::max(c, 0), coarse.cells[d] - 1);
      const double xi = (x - c * coarse.h[d]) / coarse.h[d];
      std::vector<std::pair<int, double>> w;
      for (int k = 0; k <= degree; ++k)
        {
          const double v = basis.lagrange(k, xi);
          if (std::abs(v) > 1e-14)
            w.push_back({degree * c + k, v});
        }
      return w;
    }

    /// fine += P coarse (coarse read through homogeneous constraints; fine constrained
    /// DoFs untouched).
    void
    prolongate_add(const Level &f, const Level &c, int field_pos, int field, const double *cv,
                   double *fv) const
    {
      const Mesh &fm = f.mesh, &cm = c.mesh;
      for (int64_t dof = 0; dof < fm.ndofs; ++dof)
        {
          if (f.cons[field].kind[dof] != free_dof)
            continue;
          int I[3];
          fm.coords(dof, I);
          std::vector<std::pair<int, double>> w[3];
          for (int d = 0; d < 3; ++d)
            w[d] = d < dim ? prolong_1d(fm, cm, d, I[d]) : std::vector<std::pair<int, double>> {{0, 1.0}};
          double s = 0.0;
          for (auto &a2 : w[2])
            for (auto &a1 : w[1])
              for (auto &a0 : w[0])
                {
                  const int64_t cd = cm.index(a0.first, a1.first, a2.first);
                  const auto   &cc = c.cons[field];
                  double        v  = 0.0;
                  if (cc.kind[cd] == free_dof)
                    v = cv[cd];
                  else if (cc.kind[cd] == image_dof)
                    v = cv[cc.target[cd]];
                  s += a0.second * a1.second * a2.second * v;
                }
          fv[dof] += s;
        }
      (void) field_pos;
    }

    /// coarse = P^T fine (exact transpose of prolongate_add).
    void
    restrict_to(const Level &f, const Level &c, int field, const double *fv, double *cv) const
    {
      const Mesh &fm = f.mesh, &cm = c.mesh;
      std::fill(cv, cv + cm.ndofs, 0.0);
      for (int64_t dof = 0; dof < fm.ndofs; ++dof)
        {
          if (f.cons[field].kind[dof] != free_dof)
            continue;
          int I[3];
          fm.coords(dof, I);
          std::vector<std::pair<int, double>> w[3];
          for (int d = 0; d < 3; ++d)
            w[d] = d < dim ? prolong_1d(fm, cm, d, I[d]) : std::vector<std::pair<int, double>> {{0, 1.0}};
          for (auto &a2 : w[2])
            for (auto &a1 : w[1])
              for (auto &a0 : w[0])
                {
                  const int64_t cd = cm.index(a0.first, a1.first, a2.first);
                  const auto   &cc = c.cons[field];
                  const double  ww = a0.second * a1.second * a2.second * fv[dof];
                  if (cc.kind[cd] == free_dof)
                    cv[cd] += ww;
                  else if (cc.kind[cd] == image_dof)
                    cv[cc.target[cd]] += ww;
                }
        }
    }

    std::vector<MGLevel> mg;

    void
    setup_mg(int b)
    {
      const auto &lp    = blocks[b].linear_solver_parameters;
      const int   topr  = int(inputs.spatial_discretization.global_refinement);
      const int   depth = int(lp.mg_depth);
      mg.clear();
      mg.resize(depth);
      const auto bf = block_fields(b);
      for (int i = 0; i < depth; ++i)
        {
          auto &L = mg[i].lvl;
          L.mesh  = make_mesh(unsigned(topr - depth + 1 + i));
          for (std::size_t f = 0; f < fields.size(); ++f)
            L.cons.push_back(make_constraints(L.mesh, int(f), true));
          L.jxw = lumped_mass(L.mesh);
          // inject non-src fields
          L.fields.resize(fields.size());
          for (std::size_t f = 0; f < fields.size(); ++f)
            {
              L.fields[f].resize(storage[f].size());
              for (std::size_t a = 0; a < storage[f].size(); ++a)
                {
                  auto &v = L.fields[f][a];
                  v.assign(L.mesh.ndofs, 0.0);
                  const int stride = 1 << (depth - 1 - i);
                  for (int64_t dof = 0; dof < L.mesh.ndofs; ++dof)
                    {
                      int J[3];
                      L.mesh.coords(dof, J);
                      v[dof] = storage[f][a][top.index(J[0] * stride, dim > 1 ? J[1] * stride : 0,
                                                       dim > 2 ? J[2] * stride : 0)];
                    }
                }
            }
          mg[i].sm.inv_diag =
            inverse_diagonal(b, L.mesh, L.cons, L.jxw, L.fields, scaling_of(b));
          Op A = [&, i](const std::vector<double> &x, std::vector<double> &y) {
            auto &LL = mg[i].lvl;
            y.assign(x.size(), 0.0);
            apply(b, true, LL.mesh, LL.cons, LL.jxw, LL.fields, x.data(), y.data(), scaling_of(b),
                  false);
          };
          // The smoothers are initialized once (reference linear_solver.h:135-147) and
          // deal.II's PreconditionChebyshev estimates its eigenvalues at its first
          // application only: the estimates of the first solve are kept.
          auto &kept = mg_eigenvalues[b];
          if (kept.size() == std::size_t(depth))
            {
              mg[i].sm.lmax = kept[i].first;
              mg[i].sm.lmin = kept[i].second;
            }
          else
            estimate(A, mg[i].sm, start_vector(L.mesh, int(bf.size())), lp.chebyshev_parameters);
        }
      auto &kept = mg_eigenvalues[b];
      if (kept.size() != std::size_t(depth))
        {
          kept.clear();
          for (int i = 0; i < depth; ++i)
            kept.emplace_back(mg[i].sm.lmax, mg[i].sm.lmin);
        }
    }
    std::map<int, std::vector<std::pair<double, double>>> mg_eigenvalues; // per block

    void
    v_cycle(int b, int li, std::vector<double> &u, const std::vector<double> &defect) const
    {
      const auto &lp = blocks[b].linear_solver_parameters;
      const auto &L  = mg[li].lvl;
      const auto  bf = block_fields(b);
      Op          A  = [&](const std::vector<double> &x, std::vector<double> &y) {
        y.assign(x.size(), 0.0);
        apply(b, true, L.mesh, L.cons, L.jxw, L.fields, x.data(), y.data(), scaling_of(b), false);
      };
      const unsigned deg = lp.chebyshev_parameters.degree;
      if (li == 0)
        {
          if (lp.coarse_solver == prismspf::CoarseSolver::Smoother)
            chebyshev(A, mg[0].sm, deg, u, defect, true);
          else
            {
              const std::size_t   n = defect.size();
              std::vector<double> r = defect, z(n), p(n), ap(n);
              std::fill(u.begin(), u.end(), 0.0);
              const double r0 = std::sqrt(dot(r, r));
              if (r0 > 0.0)
                {
                  for (std::size_t i = 0; i < n; ++i)
                    z[i] = mg[0].sm.inv_diag[i] * r[i];
                  p         = z;
                  double rz = dot(r, z);
                  for (unsigned it = 0; it < lp.coarse_cg_max_iterations; ++it)
                    {
                      A(p, ap);
                      const double a = rz / dot(p, ap);
                      for (std::size_t i = 0; i < n; ++i)
                        {
                          u[i] += a * p[i];
                          r[i] -= a * ap[i];
                        }
                      if (std::sqrt(dot(r, r)) <= lp.coarse_cg_tolerance * r0)
                        break;
                      for (std::size_t i = 0; i < n; ++i)
                        z[i] = mg[0].sm.inv_diag[i] * r[i];
                      const double rzn = dot(r, z);
                      for (std::size_t i = 0; i < n; ++i)
                        p[i] = z[i] + (rzn / rz) * p[i];
                      rz = rzn;
                    }
                }
            }
          return;
        }
      chebyshev(A, mg[li].sm, deg, u, defect, true);
      std::vector<double> t;
      A(u, t);
      for (std::size_t i = 0; i < t.size(); ++i)
        t[i] = defect[i] - t[i];
      const auto         &C = mg[li - 1].lvl;
      std::vector<double> cd(bf.size() * C.mesh.ndofs), cu(bf.size() * C.mesh.ndofs, 0.0);
      for (std::size_t k = 0; k < bf.size(); ++k)
        restrict_to(L, C, bf[k], t.data() + k * L.mesh.ndofs, cd.data() + k * C.mesh.ndofs);
      v_cycle(b, li - 1, cu, cd);
      for (std::size_t k = 0; k < bf.size(); ++k)
        prolongate_add(L, C, int(k), bf[k], cu.data() + k * C.mesh.ndofs,
                       u.data() + k * L.mesh.ndofs);
      chebyshev(A, mg[li].sm, deg, u, defect, false);
    }

    Smoother                                     fine_cheb;
    std::map<int, std::pair<double, double>>     cheb_eigenvalues; // per block
    std::map<int, std::vector<double>> homog; // persistent initial guesses per block
    std::map<int, Stats>               last_stats;

    /// deal.II SolverBicgstab (deal.II >= 9.6, lac/solver_bicgstab.h: solve() and
    /// iterate(); selected by `solver type = bicgstab`, reference linear_solver.h:461-490),
    /// restated.  Convergence checks (SolverControl::check: success at res <= tol,
    /// failure from step >= max on): at step 0 on |b - A x|, and in each step after the
    /// half update on |r| (early success adds alpha y) and after the full update on
    /// |A x - b| (exact residual) or |r|.  |r . rbar| or |rbar . v| below `breakdown`
    /// restarts from the current x, keeping the step count.
    template <class AOp, class POp>
    static Stats
    bicgstab(AOp &A, POp &P, std::vector<double> &x, const std::vector<double> &b, double tol,
             unsigned max_steps, bool exact_residual, double breakdown)
    {
      const std::size_t   n = x.size();
      std::vector<double> r(n), rbar(n), p(n), y(n), z(n), t(n), v(n);
      Stats               st;
      st.tol        = tol;
      unsigned step = 0;
      double   res  = 0.0;
      bool     first = true;
      auto     check = [&](double value) { return value <= tol ? 1 : (step >= max_steps ? -1 : 0); };
      while (true)
        {
          A(x, r);
          for (std::size_t i = 0; i < n; ++i)
            r[i] = b[i] - r[i];
          res = std::sqrt(dot(r, r));
          if (first)
            st.r0 = res;
          first     = false;
          int state = check(res);
          if (state != 0)
            {
              st.converged = state > 0;
              break;
            }
          double alpha = 1.0, omega = 1.0, rho = 1.0;
          rbar         = r;
          bool startup = true, restart = false;
          do
            {
              ++step;
              double rhobar = dot(r, rbar);
              if (std::fabs(rhobar) < breakdown)
                {
                  restart = true;
                  break;
                }
              const double beta = rhobar * alpha / (rho * omega);
              rho               = rhobar;
              if (startup)
                {
                  p       = r;
                  startup = false;
                }
              else
                for (std::size_t i = 0; i < n; ++i)
                  p[i] = (beta * p[i] + r[i]) - beta * omega * v[i];
              P(p, y);
              A(y, v);
              rhobar = dot(rbar, v);
              if (std::fabs(rhobar) < breakdown)
                {
                  restart = true;
                  break;
                }
              alpha = rho / rhobar;
              for (std::size_t i = 0; i < n; ++i)
                r[i] -= alpha * v[i];
              res = std::sqrt(dot(r, r));
              if (check(res) > 0)
                {
                  for (std::size_t i = 0; i < n; ++i)
                    x[i] += alpha * y[i];
                  state = 1;
                  break;
                }
              P(r, z);
              A(z, t);
              rhobar            = dot(t, r);
              const double t_sq = dot(t, t);
              omega = t_sq < std::numeric_limits<double>::min() ? 0.0 : rhobar / t_sq;
              for (std::size_t i = 0; i < n; ++i)
                {
                  x[i] += alpha * y[i] + omega * z[i];
                  r[i] -= omega * t[i];
                }
              if (exact_residual)
                {
                  A(x, t);
                  for (std::size_t i = 0; i < n; ++i)
                    t[i] -= b[i];
                  res = std::sqrt(dot(t, t));
                }
              else
                res = std::sqrt(dot(r, r));
              state = check(res);
            }
          while (state == 0);
          if (!restart)
            {
              st.converged = state > 0;
              break;
            }
        }
      st.iterations = int(step);
      st.r          = res;
      return st;
    }

    /// deal.II SolverRichardson (lac/solver_richardson.h), restated: r = b - A x,
    /// d = P^{-1} r, check |r| (or |d|) at the step count, x += omega d.
    template <class AOp, class POp>
    static Stats
    richardson(AOp &A, POp &P, std::vector<double> &x, const std::vector<double> &b, double tol,
               unsigned max_steps, double omega, bool preconditioned_residual)
    {
      const std::size_t   n = x.size();
      std::vector<double> r(n), d(n);
      Stats               st;
      st.tol        = tol;
      unsigned step = 0;
      for (;; ++step)
        {
          A(x, r);
          for (std::size_t i = 0; i < n; ++i)
            r[i] = b[i] - r[i];
          P(r, d);
          const double res = std::sqrt(preconditioned_residual ? dot(d, d) : dot(r, r));
          if (step == 0)
            st.r0 = res;
          st.r = res;
          if (res <= tol || step >= max_steps)
            {
              st.converged = res <= tol;
              break;
            }
          for (std::size_t i = 0; i < n; ++i)
            x[i] += omega * d[i];
        }
      st.iterations = int(step);
      return st;
    }

    /// deal.II SolverMinRes (lac/solver_minres.h), restated as the published
    /// Paige-Saunders MINRES (SIAM J. Numer. Anal. 12, 1975): preconditioned Lanczos
    /// three-term recurrence, one Givens rotation per step, direction vectors
    /// d_k = (v_k - eps_k d_{k-2} - delta_k d_{k-1}) / gamma_k.  The checked value is
    /// the residual estimate phibar (|r| in the P^{-1} norm; |b - A x| without a
    /// preconditioner) at step 0 and after each step.
    template <class AOp, class POp>
    static Stats
    minres(AOp &A, POp &P, std::vector<double> &x, const std::vector<double> &b, double tol,
           unsigned max_steps)
    {
      const std::size_t   n = x.size();
      std::vector<double> u_prev(n, 0.0), u_cur(n), z(n), q(n), Aq(n);
      std::vector<double> d_prev2(n, 0.0), d_prev(n, 0.0), d(n);
      Stats               st;
      st.tol = tol;
      A(x, Aq);
      for (std::size_t i = 0; i < n; ++i)
        u_cur[i] = b[i] - Aq[i];
      P(u_cur, z);
      double beta_k = std::sqrt(std::max(0.0, dot(u_cur, z)));
      double beta_prev = 0.0;
      double phibar    = beta_k;
      st.r0            = phibar;
      // Rotation of the previous two steps: (c_{k-1}, s_{k-1}) and the pieces of the
      // tridiagonal column they leave behind.
      double   c = -1.0, s = 0.0, dbar = 0.0, eps_next = 0.0;
      unsigned step = 0;
      int      state = phibar <= tol ? 1 : (step >= max_steps ? -1 : 0);
      while (state == 0)
        {
          ++step;
          for (std::size_t i = 0; i < n; ++i)
            q[i] = z[i] / beta_k;
          A(q, Aq);
          // Next Lanczos vector (unpreconditioned form): u = A q - beta_k u_{k-1}/beta_{k-1}
          // - alpha u_k/beta_k, alpha = q.(A q - beta_k u_{k-1}/beta_{k-1}) (the modified
          // Gram-Schmidt order of the published algorithm).
          if (step >= 2)
            for (std::size_t i = 0; i < n; ++i)
              Aq[i] -= (beta_k / beta_prev) * u_prev[i];
          const double alpha = dot(q, Aq);
          for (std::size_t i = 0; i < n; ++i)
            {
              u_prev[i] = u_cur[i];
              u_cur[i]  = Aq[i] - (alpha / beta_k) * u_cur[i];
            }
          P(u_cur, z);
          beta_prev               = beta_k;
          beta_k                  = std::sqrt(std::max(0.0, dot(u_cur, z)));
          const double eps_k      = eps_next;
          const double delta_k    = c * dbar + s * alpha;
          const double gamma_bar  = s * dbar - c * alpha;
          eps_next                = s * beta_k;
          dbar                    = -c * beta_k;
          const double gamma_k    = std::max(std::hypot(gamma_bar, beta_k),
                                          std::numeric_limits<double>::epsilon());
          c                       = gamma_bar / gamma_k;
          s                       = beta_k / gamma_k;
          const double tau        = c * phibar;
          phibar                  = s * phibar;
          for (std::size_t i = 0; i < n; ++i)
            {
              d[i] = (q[i] - eps_k * d_prev2[i] - delta_k * d_prev[i]) / gamma_k;
              x[i] += tau * d[i];
            }
          std::swap(d_prev2, d_prev);
          std::swap(d_prev, d);
          if (!std::isfinite(phibar))
            throw std::runtime_error("MINRES produced a non-finite residual");
          state = phibar <= tol ? 1 : (step >= max_steps ? -1 : 0);
          if (state == 0 && beta_k == 0.0)
            state = -1;
        }
      st.iterations = int(step);
      st.r          = phibar;
      st.converged  = state > 0;
      return st;
    }

    /// deal.II SolverGMRES (lac/solver_gmres.h), restated: restarted GMRES, basis of
    /// `max basis size` vectors, left (default) or right preconditioning, modified
    /// Gram-Schmidt, Givens rotations; each cycle checks the recomputed (preconditioned,
    /// for left preconditioning) residual at the accumulated step count and each Arnoldi
    /// step the least-squares estimate.  `flexible`: SolverFGMRES, right
    /// preconditioning keeping z_j = P^{-1} v_j and updating x += Z y.
    template <class AOp, class POp>
    static Stats
    gmres(AOp &A, POp &P, std::vector<double> &x, const std::vector<double> &b, double tol,
          unsigned max_steps, int m, bool right, bool flexible = false)
    {
      right                              = right || flexible;
      const std::size_t                n = x.size();
      std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
      std::vector<std::vector<double>> Z(flexible ? m : 0, std::vector<double>(n));
      std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
      std::vector<double>              w(n), t(n), cs(m), sn(m), g(m + 1), y(m);
      Stats                            st;
      st.tol        = tol;
      unsigned its  = 0;
      double   rho  = 0.0;
      int      state = 0;
      bool     first = true;
      auto     check = [&](double v) { return v <= tol ? 1 : (its >= max_steps ? -1 : 0); };
      while (true)
        {
          if (!right)
            {
              A(x, t);
              for (std::size_t i = 0; i < n; ++i)
                t[i] = b[i] - t[i];
              P(t, V[0]);
            }
          else
            {
              A(x, V[0]);
              for (std::size_t i = 0; i < n; ++i)
                V[0][i] = b[i] - V[0][i];
            }
          rho = std::sqrt(dot(V[0], V[0]));
          if (first)
            st.r0 = rho;
          first = false;
          state = check(rho);
          if (state != 0)
            break;
          for (auto &v : V[0])
            v *= 1.0 / rho;
          std::fill(g.begin(), g.end(), 0.0);
          g[0]  = rho;
          int j = 0;
          while (j < m)
            {
              ++its;
              if (!right)
                {
                  A(V[j], t);
                  P(t, w);
                }
              else
                {
                  auto &z = flexible ? Z[j] : t;
                  P(V[j], z);
                  A(z, w);
                }
              for (int i = 0; i <= j; ++i)
                {
                  const double h = dot(w, V[i]);
                  H[i][j]        = h;
                  for (std::size_t k = 0; k < n; ++k)
                    w[k] -= h * V[i][k];
                }
              const double hn = std::sqrt(dot(w, w));
              H[j + 1][j]     = hn;
              if (hn > 0.0)
                for (std::size_t k = 0; k < n; ++k)
                  V[j + 1][k] = w[k] / hn;
              for (int i = 0; i < j; ++i)
                {
                  const double a = H[i][j], c = H[i + 1][j];
                  H[i][j]        = cs[i] * a + sn[i] * c;
                  H[i + 1][j]    = -sn[i] * a + cs[i] * c;
                }
              const double a = H[j][j], c = H[j + 1][j];
              const double r = std::sqrt(a * a + c * c);
              cs[j]          = r > 0.0 ? a / r : 1.0;
              sn[j]          = r > 0.0 ? c / r : 0.0;
              H[j][j]        = r;
              H[j + 1][j]    = 0.0;
              g[j + 1]       = -sn[j] * g[j];
              g[j]           = cs[j] * g[j];
              rho            = std::fabs(g[j + 1]);
              ++j;
              state = check(rho);
              if (state != 0 || hn == 0.0)
                break;
            }
          for (int i = j - 1; i >= 0; --i)
            {
              double s = g[i];
              for (int k = i + 1; k < j; ++k)
                s -= H[i][k] * y[k];
              y[i] = s / H[i][i];
            }
          if (!right || flexible)
            {
              const auto &B = flexible ? Z : V;
              for (int i = 0; i < j; ++i)
                for (std::size_t k = 0; k < n; ++k)
                  x[k] += y[i] * B[i][k];
            }
          else
            {
              std::fill(w.begin(), w.end(), 0.0);
              for (int i = 0; i < j; ++i)
                for (std::size_t k = 0; k < n; ++k)
                  w[k] += y[i] * V[i][k];
              P(w, t);
              for (std::size_t k = 0; k < n; ++k)
                x[k] += t[k];
            }
          if (state != 0)
            break;
        }
      st.iterations = int(its);
      st.r          = rho;
      st.converged  = state > 0;
      return st;
    }

    /// The block's preconditioner set up for a solve (reference linear_solver.h:348-379):
    /// Chebyshev on the fine level (inverse diagonal re-evaluated every solve, reference
    /// linear_solver.h:362-367; eigenvalues estimated at the preconditioner's first
    /// application only) or the multigrid levels.
    void
    prepare_preconditioner(int b)
    {
      const auto &lp = blocks[b].linear_solver_parameters;
      const int   sc = scaling_of(b);
      if (lp.preconditioner == prismspf::Chebyshev)
        {
          fine_cheb.inv_diag = inverse_diagonal(b, top, cons, jxw, storage, sc);
          auto it            = cheb_eigenvalues.find(b);
          if (it == cheb_eigenvalues.end())
            {
              estimate(top_operator(b), fine_cheb, start_vector(top, int(block_fields(b).size())),
                       lp.chebyshev_parameters);
              cheb_eigenvalues[b] = {fine_cheb.lmax, fine_cheb.lmin};
            }
          else
            {
              fine_cheb.lmax = it->second.first;
              fine_cheb.lmin = it->second.second;
            }
        }
      else if (lp.preconditioner == prismspf::GMG)
        setup_mg(b);
    }

    /// h = P^{-1} g with the preconditioner prepared by prepare_preconditioner.
    void
    precondition(int b, const std::vector<double> &g, std::vector<double> &h) const
    {
      const auto &lp = blocks[b].linear_solver_parameters;
      if (lp.preconditioner == prismspf::None)
        h = g;
      else if (lp.preconditioner == prismspf::Chebyshev)
        {
          h.assign(g.size(), 0.0);
          chebyshev(top_operator(b), fine_cheb, lp.chebyshev_parameters.degree, h, g, true);
        }
      else
        {
          h.assign(g.size(), 0.0);
          v_cycle(b, int(mg.size()) - 1, h, g);
        }
    }

    /// The lhs operator of block b on the finest level, scaled for the linear solver.
    Op
    top_operator(int b) const
    {
      const int sc = scaling_of(b);
      return [this, b, sc](const std::vector<double> &x, std::vector<double> &y) {
        y.assign(x.size(), 0.0);
        apply(b, true, top, cons, jxw, storage, x.data(), y.data(), sc, false);
      };
    }

    /// The lhs operator of block b on multigrid level li (homogeneous level constraints).
    Op
    level_operator(int b, int li) const
    {
      const int sc = scaling_of(b);
      return [this, b, li, sc](const std::vector<double> &x, std::vector<double> &y) {
        const auto &L = mg[li].lvl;
        y.assign(x.size(), 0.0);
        apply(b, true, L.mesh, L.cons, L.jxw, L.fields, x.data(), y.data(), sc, false);
      };
    }

    /// do_linear_solve (reference linear_solver.h:348-402): the block's preconditioner,
    /// set up for this solve, and deal.II SolverCG's recurrences from the initial guess x
    /// (absolute residual tolerance = tolerance value * normalization).
    Stats
    cg_solve(int b, std::vector<double> &x, const std::vector<double> &rhs)
    {
      const auto    bf = block_fields(b);
      const int64_t n  = int64_t(bf.size()) * top.ndofs;
      const int     sc = scaling_of(b);
      const auto   &lp = blocks[b].linear_solver_parameters;
      Op A = [&](const std::vector<double> &x, std::vector<double> &y) {
        y.assign(x.size(), 0.0);
        apply_top(b, true, x.data(), y.data(), sc, false);
      };
      prepare_preconditioner(b);
      auto P = [&](const std::vector<double> &g, std::vector<double> &h) {
        precondition(b, g, h);
      };

      Stats st;
      st.tol = lp.tolerance * normalization(b);
      if (lp.solver_type == "bicgstab")
        return bicgstab(A, P, x, rhs, st.tol, lp.max_iterations, lp.bicgstab_exact_residual,
                        lp.bicgstab_breakdown);
      if (lp.solver_type == "gmres" || lp.solver_type == "fgmres")
        return gmres(A, P, x, rhs, st.tol, lp.max_iterations, int(lp.gmres_max_basis_size),
                     lp.gmres_right_preconditioning, lp.solver_type == "fgmres");
      if (lp.solver_type == "richardson")
        return richardson(A, P, x, rhs, st.tol, lp.max_iterations, lp.richardson_omega,
                          lp.richardson_use_preconditioned_residual);
      if (lp.solver_type == "minres")
        return minres(A, P, x, rhs, st.tol, lp.max_iterations);
      if (lp.solver_type != "cg")
        throw std::invalid_argument("linear solver type " + lp.solver_type +
                                    " is not part of the oracle");
      std::vector<double> g(n), h(n), d(n), ad(n);
      A(x, g);
      for (int64_t i = 0; i < n; ++i)
        g[i] -= rhs[i];
      double res = std::sqrt(dot(g, g));
      st.r0      = res;
      unsigned step = 0;
      if (!(res <= st.tol))
        {
          P(g, h);
          for (int64_t i = 0; i < n; ++i)
            d[i] = -h[i];
          double gh = dot(g, h);
          while (true)
            {
              ++step;
              A(d, ad);
              const double alpha = gh / dot(d, ad);
              for (int64_t i = 0; i < n; ++i)
                {
                  x[i] += alpha * d[i];
                  g[i] += alpha * ad[i];
                }
              res = std::sqrt(dot(g, g));
              if (res <= st.tol || step >= lp.max_iterations)
                break;
              P(g, h);
              const double ghn = dot(g, h);
              const double bt  = ghn / gh;
              gh               = ghn;
              for (int64_t i = 0; i < n; ++i)
                d[i] = -h[i] + bt * d[i];
            }
        }
      st.iterations = int(step);
      st.r          = res;
      st.converged  = res <= st.tol;
      return st;
    }

    Stats
    solve_linear(int b)
    {
      const auto    bf = block_fields(b);
      const int64_t n  = int64_t(bf.size()) * top.ndofs;
      const int     sc = scaling_of(b);
      std::vector<double> rhs(n), inh(n, 0.0), tmp(n);
      apply_top(b, false, nullptr, rhs.data(), sc, false);
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          const auto &c = cons[bf[k]];
          for (int64_t i = 0; i < top.ndofs; ++i)
            if (c.kind[i] == dirichlet)
              inh[k * top.ndofs + i] = c.g[i];
        }
      apply_top(b, true, inh.data(), tmp.data(), sc, true);
      for (int64_t i = 0; i < n; ++i)
        rhs[i] -= tmp[i];

      auto &x = homog[b];
      if (x.size() != std::size_t(n))
        x.assign(n, 0.0);
      Stats st = cg_solve(b, x, rhs);
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          auto &sol = storage[bf[k]][0];
          for (int64_t i = 0; i < top.ndofs; ++i)
            sol[i] = x[k * top.ndofs + i] + inh[k * top.ndofs + i];
          distribute(cons[bf[k]], sol.data());
        }
      last_stats[b] = st;
      return st;
    }

    struct NewtonStats
    {
      int    iterations = 0, linear_iterations = 0;
      double residual = 0.0, tol = 0.0;
      bool   converged = false;
    };
    std::map<int, std::vector<double>> newton_update; // persistent per block
    std::map<int, NewtonStats>         last_newton;

    /// NewtonSolver::solve_impl (reference newton_solver.h:73-156), step for step:
    /// start from old_1 when the block's fields keep old states; each iteration applies
    /// the constraints to the solution, evaluates the (scaled) rhs operator as the
    /// residual, stops when its l2 norm is <= tolerance value * normalization, and
    /// otherwise solves lhs(du) = r with CG from the previous update (the update vector
    /// is a member of the solver and is never reset) and adds step_length * du.
    NewtonStats
    solve_newton(int b)
    {
      const auto    bf = block_fields(b);
      const int64_t n  = int64_t(bf.size()) * top.ndofs;
      const int     sc = scaling_of(b);
      const auto   &np = blocks[b].nonlinear_solver_parameters;
      NewtonStats   ns;
      ns.tol = np.tolerance_value * normalization(b);
      for (int f : bf)
        if (storage[f].size() > 1)
          storage[f][0] = storage[f][1];
      auto &du = newton_update[b];
      if (du.size() != std::size_t(n))
        du.assign(n, 0.0);
      std::vector<double> r(n);
      bool                unconverged = true;
      unsigned            iter        = 0;
      while (unconverged && iter < np.max_iterations)
        {
          for (int f : bf)
            distribute(cons[f], storage[f][0].data());
          std::fill(r.begin(), r.end(), 0.0);
          apply_top(b, false, nullptr, r.data(), sc, false);
          ns.residual = std::sqrt(dot(r, r));
          unconverged = ns.residual > ns.tol;
          if (!unconverged)
            continue;
          const Stats st = cg_solve(b, du, r);
          last_stats[b]  = st;
          ns.linear_iterations += st.iterations;
          for (std::size_t k = 0; k < bf.size(); ++k)
            {
              auto &sol = storage[bf[k]][0];
              for (int64_t i = 0; i < top.ndofs; ++i)
                sol[i] += np.step_length * du[k * top.ndofs + i];
            }
          ++iter;
        }
      ns.iterations  = int(iter);
      ns.converged   = !unconverged;
      last_newton[b] = ns;
      return ns;
    }

    // ------------------------------------------------------------ time loop
    void
    apply_initial_conditions(int b)
    {
      for (int f : block_fields(b))
        {
          auto &v = storage[f][0];
          for (int64_t dof = 0; dof < top.ndofs; ++dof)
            {
              int i[3];
              top.coords(dof, i);
              // Periodic images take the master's value.
              for (int d = 0; d < dim; ++d)
                if (top.periodic[d] && i[d] == top.N[d] - 1)
                  i[d] = 0;
              dealii::Point<dim> p;
              for (int d = 0; d < dim; ++d)
                p[d] = top.lower[d] + i[d] * top.h[d] / degree;
              double      val = 0.0, vec = 0.0;
              const auto &lc  = layout[f];
              hooks().set_initial_condition(unsigned(lc.declared), lc.component, p, val, vec);
              v[dof] = lc.vector ? vec : val;
            }
        }
    }

    void
    solve_block(int b)
    {
      const auto &blk = blocks[b];
      if (timer.get_increment() == 0 && blk.solve_timing == prismspf::Primary)
        apply_initial_conditions(b);
      else if (blk.solve_type == prismspf::Explicit)
        solve_explicit(b);
      else if (blk.solve_type == prismspf::Linear)
        solve_linear(b);
      else if (blk.solve_type == prismspf::Newton)
        solve_newton(b);
      if (timer.get_increment() == 0)
        for (int f : block_fields(b))
          for (std::size_t a = 1; a < storage[f].size(); ++a)
            storage[f][a] = storage[f][0];
      for (int f : block_fields(b))
        rotated[f] = false;
    }

    void
    rotate(int b)
    {
      for (int f : block_fields(b))
        {
          for (int a = int(storage[f].size()) - 1; a >= 1; --a)
            std::swap(storage[f][a], storage[f][a - 1]);
          rotated[f] = storage[f].size() > 1;
        }
    }

    /// Age holding the newest values of field f: after a rotation (the end of every
    /// increment) that is old_1, as in the reference's solution handler.
    int
    latest_age(int f) const
    {
      return rotated.count(f) && rotated.at(f) ? 1 : 0;
    }

    std::map<int, bool> rotated;

    std::pair<double, double>
    norms(int f, int age) const
    {
      const auto &v  = storage[f][age];
      double      s2 = 0.0, si = 0.0;
      for (int64_t i = 0; i < top.ndofs; ++i)
        {
          s2 += v[i] * v[i];
          si += v[i] * jxw[i];
        }
      return {std::sqrt(s2), si};
    }

    /// ConstraintManager::update_time_dependent_constraints (reference
    /// constraint_manager.cc:380-423, called first in problem.cc:428-430): the
    /// constraints of fields whose boundary conditions are time dependent are rebuilt
    /// from set_dirichlet at the current time.  Level constraints are homogeneous and
    /// do not change.
    void
    update_time_dependent_constraints()
    {
      for (std::size_t f = 0; f < fields.size(); ++f)
        if (fields[f].boundary_conditions.time_dependent)
          cons[f] = make_constraints(top, int(f), false);
    }

    int
    solve_increment()
    {
      const unsigned inc = timer.get_increment();
      const bool     out = inputs.output_parameters.should_output(inc);
      update_time_dependent_constraints();
      for (std::size_t b = 0; b < blocks.size(); ++b)
        {
          const auto t = blocks[b].solve_timing;
          if ((t == prismspf::PostProcess || t == prismspf::NucleationRate) && !out)
            continue;
          solve_block(int(b));
        }
      int status = 0;
      for (std::size_t f = 0; f < fields.size(); ++f)
        if (!std::isfinite(storage[f][0][0]))
          {
            status = 2;
            break;
          }
      if (status == 0 && pde->get_user_stop())
        status = 3;
      if (out || status)
        {
          std::ostringstream os;
          os.precision(12);
          os << "Iteration " << inc << " Time " << timer.get_time() << "\n";
          for (std::size_t f = 0; f < fields.size(); ++f)
            {
              auto [l2, in] = norms(int(f), 0);
              os << "  Field index " << f << " name " << fields[f].name << " l2-norm " << l2
                 << " integrated value " << in << "\n";
            }
          log += os.str();
        }
      for (std::size_t b = 0; b < blocks.size(); ++b)
        rotate(int(b));
      return status;
    }

    int
    block_index(int id) const
    {
      for (std::size_t b = 0; b < blocks.size(); ++b)
        if (blocks[b].id == id)
          return int(b);
      throw std::invalid_argument("no solve block with id " + std::to_string(id));
    }

    // ------------------------------------------------------------- adjoint
    /// (K u) with constraint-resolving reads and scatters (images folded, no Dirichlet).
    std::vector<double>
    stiffness(const std::vector<double> &u) const
    {
      std::vector<double>   out(top.ndofs);
      std::vector<double *> dst {out.data()};
      Inputs                in;
      in.vec[0][0]   = u.data();
      in.flags[0][0] = f_gradients | f_resolve;
      cell_loop(top, cons, in, {0}, dst, [](FC &fc) {
        for (int d = 0; d < dim; ++d)
          fc.gradient_term[0][d] = fc.gradients[0][0][d];
      });
      return out;
    }

    template <class E>
    std::vector<double>
    ac_step(const std::vector<double> &phi, double L, double kappa) const
    {
      const double dt = timer.get_timestep();
      auto         kp = stiffness(phi);
      std::vector<double> out(top.ndofs);
      for (int64_t i = 0; i < top.ndofs; ++i)
        {
          if (cons[0].kind[i] != free_dof)
            continue;
          out[i] = phi[i] - dt * L * E::fprime(phi[i]) - dt * L * kappa * kp[i] / jxw[i];
        }
      distribute(cons[0], out.data(), true);
      return out;
    }

    template <class E>
    std::vector<double>
    ac_forward(const std::vector<double> &phi0, double L, double kappa, unsigned n) const
    {
      auto phi = phi0;
      for (unsigned s = 0; s < n; ++s)
        phi = ac_step<E>(phi, L, kappa);
      return phi;
    }

    /// Reverse sweep storing every forward state (the oracle needs no checkpointing).
    template <class E>
    void
    ac_vjp(const std::vector<double> &phi0, std::vector<double> lam, double L, double kappa,
           unsigned n, double &dL, double &dk, std::vector<double> &dphi0) const
    {
      const double                     dt = timer.get_timestep();
      std::vector<std::vector<double>> states {phi0};
      for (unsigned s = 0; s < n; ++s)
        states.push_back(ac_step<E>(states.back(), L, kappa));
      // Fold image entries of the cotangent onto masters.
      for (int64_t i = 0; i < top.ndofs; ++i)
        if (cons[0].kind[i] == image_dof)
          {
            lam[cons[0].target[i]] += lam[i];
            lam[i] = 0.0;
          }
      dL = dk = 0.0;
      for (int s = int(n) - 1; s >= 0; --s)
        {
          const auto         &phi = states[s];
          std::vector<double> lom(top.ndofs, 0.0);
          for (int64_t i = 0; i < top.ndofs; ++i)
            if (cons[0].kind[i] == free_dof)
              lom[i] = lam[i] / jxw[i];
          distribute(cons[0], lom.data(), true);
          const auto          klm = stiffness(lom);
          const auto          kp  = stiffness(phi);
          std::vector<double> prev(top.ndofs, 0.0);
          for (int64_t i = 0; i < top.ndofs; ++i)
            {
              if (cons[0].kind[i] != free_dof)
                continue;
              prev[i] = lam[i] - dt * L * E::fprime2(phi[i]) * lam[i] - dt * L * kappa * klm[i];
              dL += lam[i] * (-dt * E::fprime(phi[i]) - dt * kappa * kp[i] / jxw[i]);
              dk += lam[i] * (-dt * L * kp[i] / jxw[i]);
            }
          lam = prev;
        }
      dphi0 = lam;
    }

    /// Explicit Cahn-Hilliard step (the c and mu blocks, reference applications/
    /// cahn_hilliard/explicit/custom_pde.h:92-118, each an invm-scaled assembled
    /// operator): mu = f'(c) + kappa M^-1 K c, c' = c - dt M_mob M^-1 K mu.
    template <class E>
    std::vector<double>
    ch_potential(const std::vector<double> &c, double kappa) const
    {
      const auto          kc = stiffness(c);
      std::vector<double> mu(top.ndofs, 0.0);
      for (int64_t i = 0; i < top.ndofs; ++i)
        if (cons[0].kind[i] == free_dof)
          mu[i] = E::fprime(c[i]) + kappa * kc[i] / jxw[i];
      distribute(cons[0], mu.data(), true);
      return mu;
    }

    template <class E>
    std::vector<double>
    ch_step(const std::vector<double> &c, double mob, double kappa) const
    {
      const double        dt  = timer.get_timestep();
      const auto          mu  = ch_potential<E>(c, kappa);
      const auto          kmu = stiffness(mu);
      std::vector<double> out(top.ndofs, 0.0);
      for (int64_t i = 0; i < top.ndofs; ++i)
        if (cons[0].kind[i] == free_dof)
          out[i] = c[i] - dt * mob * kmu[i] / jxw[i];
      distribute(cons[0], out.data(), true);
      return out;
    }

    template <class E>
    std::vector<double>
    ch_forward(const std::vector<double> &c0, double mob, double kappa, unsigned n) const
    {
      auto c = c0;
      for (unsigned s = 0; s < n; ++s)
        c = ch_step<E>(c, mob, kappa);
      return c;
    }

    /// Reverse sweep of the Cahn-Hilliard step storing every forward state:
    /// lam_n = lam - dt M_mob (f''(c) w + kappa K(M^-1 w)) with w = K(M^-1 lam);
    /// dJ/dM_mob += -dt lam . M^-1 K mu, dJ/dkappa += -dt M_mob w . M^-1 K c.
    template <class E>
    void
    ch_vjp(const std::vector<double> &c0, std::vector<double> lam, double mob, double kappa,
           unsigned n, double &dmob, double &dk, std::vector<double> &dc0) const
    {
      const double                     dt = timer.get_timestep();
      std::vector<std::vector<double>> states {c0};
      for (unsigned s = 0; s < n; ++s)
        states.push_back(ch_step<E>(states.back(), mob, kappa));
      for (int64_t i = 0; i < top.ndofs; ++i)
        if (cons[0].kind[i] == image_dof)
          {
            lam[cons[0].target[i]] += lam[i];
            lam[i] = 0.0;
          }
      auto over_mass = [&](const std::vector<double> &v) {
        std::vector<double> o(top.ndofs, 0.0);
        for (int64_t i = 0; i < top.ndofs; ++i)
          if (cons[0].kind[i] == free_dof)
            o[i] = v[i] / jxw[i];
        distribute(cons[0], o.data(), true);
        return o;
      };
      dmob = dk = 0.0;
      for (int s = int(n) - 1; s >= 0; --s)
        {
          const auto &c   = states[s];
          const auto  mu  = ch_potential<E>(c, kappa);
          const auto  kmu = stiffness(mu);
          const auto  kc  = stiffness(c);
          auto        w   = stiffness(over_mass(lam));
          for (int64_t i = 0; i < top.ndofs; ++i)
            if (cons[0].kind[i] != free_dof)
              w[i] = 0.0;
          const auto          kw = stiffness(over_mass(w));
          std::vector<double> prev(top.ndofs, 0.0);
          for (int64_t i = 0; i < top.ndofs; ++i)
            {
              if (cons[0].kind[i] != free_dof)
                continue;
              prev[i] = lam[i] - dt * mob * (E::fprime2(c[i]) * w[i] + kappa * kw[i]);
              dmob += lam[i] * (-dt * kmu[i] / jxw[i]);
              dk += w[i] * (-dt * mob * kc[i] / jxw[i]);
            }
          lam = prev;
        }
      dc0 = lam;
    }

    // ----------------------------------------------------------------- state
    prismspf::UserInputParameters<dim>               inputs;
    prismspf::PhaseFieldTools<dim>                   tools;
    Basis1D                                          basis;
    std::vector<prismspf::FieldAttributes>           fields;
    std::vector<prismspf::StoredComponent>           layout;
    std::vector<SolveBlock>                          blocks;
    std::unique_ptr<PDE>                             pde;
    prismspf::SimulationTimer                        timer;
    Mesh                                             top;
    std::vector<Constraints>                         cons;
    std::vector<double>                              jxw;
    std::vector<std::vector<std::vector<double>>>    storage;
    std::string                                      log;
  };
} // namespace oracle
