// Host runtime of one application on one GPU: field storage, operator application,
// the explicit and linear solvers (CG with Chebyshev or geometric-multigrid
// preconditioning, optional FP32 levels), and the increment loop.
//
// It is the device counterpart of the reference's solver layer
// (reference: include/prismspf/solvers/{explicit_solver,linear_solver,solver_base}.h,
// src/solvers/mf_operator.cc, src/core/{problem,group_solution_handler}.cc) and keeps
// its ordering and constraint semantics; everything heavy runs as HIP kernels on the
// runtime's stream.  `App` supplies the dimension, degree, the equation class and the
// field/solve-block declarations (the body of a reference `main.cc`).
#pragma once

#include "device.hpp"
#include "grid.hpp"
#include "operator_kernels.hpp"
#include "operator_plan.hpp"
#include "pair_kernels.hpp"
#include "stencil_kernels.hpp"
#include "vector_kernels.hpp"
#include "vtu_writer.hpp"

#include <prismspf/core/field_attributes.h>
#include <prismspf/core/phase_field_tools.h>
#include <prismspf/core/simulation_timer.h>
#include <prismspf/core/solve_block.h>
#include <prismspf/user_inputs/user_input_parameters.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <functional>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <chrono>
#include <deque>
#include <set>
#include <vector>

namespace ppf
{
  enum class OperatorPath : int
  {
    Auto      = 0,
    Generic   = 1,
    /// Split operator on the tiled stencil kernel (per-node kernel beyond two
    /// stencil operands).
    Split     = 2,
    /// Split operator on the per-node kernel.
    SplitNode = 3
  };

  /// Cache policy of the stencil kernel's output stores.
  enum class StorePolicy : int
  {
    Auto        = 0,
    Temporal    = 1,
    NonTemporal = 2
  };

  /// Statistics of the last linear solve of a block.
  struct SolveStats
  {
    int    iterations       = 0;
    double initial_residual = 0.0;
    double final_residual   = 0.0;
    double tolerance        = 0.0;
    int    converged        = 0;
  };

  /// Statistics of the last Newton solve of a block (the quantities of the reference's
  /// log line, newton_solver.h:140-147).
  struct NewtonStats
  {
    int    iterations        = 0;   ///< Newton updates applied
    int    linear_iterations = 0;   ///< CG steps over all updates
    double residual          = 0.0; ///< l2 norm of the last residual evaluated
    double tolerance         = 0.0; ///< tolerance value * normalization
    int    converged         = 0;
  };

  template <int... ids, class F>
  void
  dispatch_block_id(std::integer_sequence<int, ids...>, int id, F &&f)
  {
    const bool found = ((id == ids ? (f(std::integral_constant<int, ids> {}), true) : false) ||
                        ...);
    if (!found)
      throw Error(status_invalid, "solve block id " + std::to_string(id) +
                                    " is not compiled into this application library");
  }

  /// Most increments the run loop enqueues ahead of the oldest unread NaN probe.
  constexpr int max_pipeline_depth = 3;

  /// Whether id is one of the compiled solve block ids.
  template <int id, int... ids>
  constexpr bool
  has_block_id(std::integer_sequence<int, ids...>)
  {
    return ((id == ids) || ...);
  }

  /// Tridiagonal symmetric eigenvalues by bisection (Sturm sequences); used for the
  /// Lanczos estimate of the Chebyshev eigenvalue range.
  inline double
  tridiagonal_max_eigenvalue(const std::vector<double> &a, const std::vector<double> &b)
  {
    const int n = int(a.size());
    if (n == 0)
      return 0.0;
    double lo = a[0], hi = a[0];
    for (int i = 0; i < n; ++i)
      {
        const double r = (i > 0 ? std::abs(b[i - 1]) : 0.0) + (i + 1 < n ? std::abs(b[i]) : 0.0);
        lo             = std::min(lo, a[i] - r);
        hi             = std::max(hi, a[i] + r);
      }
    auto count_below = [&](double x) {
      int    c = 0;
      double q = a[0] - x;
      if (q < 0)
        ++c;
      for (int i = 1; i < n; ++i)
        {
          const double denom = q == 0.0 ? 1e-300 : q;
          q                  = a[i] - x - b[i - 1] * b[i - 1] / denom;
          if (q < 0)
            ++c;
        }
      return c;
    };
    for (int it = 0; it < 200; ++it)
      {
        const double mid = 0.5 * (lo + hi);
        if (count_below(mid) >= n)
          hi = mid;
        else
          lo = mid;
      }
    return hi;
  }

  template <class App>
  class Runtime
  {
  public:
    static constexpr int dim    = App::dim;
    static constexpr int degree = App::degree;
    template <typename number>
    using PDE = typename App::template PDE<number>;

    explicit Runtime(const std::string &prm_text)
      : inputs(prismspf::UserInputParameters<dim>::from_text(prm_text))
    {
      App::define(fields, blocks);
      prismspf::validate_solve_blocks(blocks, fields);

      // Parameter-file overrides of the declared solver settings and boundary
      // conditions (reference problem.cc:152-224).
      for (auto &b : blocks)
        {
          const auto it = inputs.linear_solve_parameters.find(b.id);
          if (it != inputs.linear_solve_parameters.end())
            b.linear_solver_parameters = it->second;
          const auto nit = inputs.nonlinear_solve_parameters.find(b.id);
          if (nit != inputs.nonlinear_solve_parameters.end())
            b.nonlinear_solver_parameters = nit->second;
        }
      for (auto &fa : fields)
        {
          const auto it = inputs.boundary_parameters.find(fa.name);
          if (it != inputs.boundary_parameters.end())
            fa.boundary_conditions = it->second;
        }
      try
        {
          layout = prismspf::expand_vector_fields(fields, blocks, dim);
        }
      catch (const std::invalid_argument &e)
        {
          throw Error(status_unsupported, e.what());
        }
      if (int(fields.size()) != App::n_fields)
        throw Error(status_internal, "App::n_fields does not match the stored fields");
      if (fields.size() > PPF_MAX_FIELDS)
        throw Error(status_unsupported, "more than PPF_MAX_FIELDS fields");

      pde_d = std::make_unique<PDE<double>>(inputs, tools);
      pde_f = std::make_unique<PDE<float>>(inputs, tools);
      timer = prismspf::SimulationTimer(inputs.temporal_discretization.dt,
                                        inputs.temporal_discretization.initial_time);

      fine = make_grid(inputs.spatial_discretization.global_refinement);
      PPF_HIP_CHECK(hipStreamCreate(&stream));
      reduce_buf.resize(2 * reduce_blocks + 8);
      scalar_buf.resize(8);
      coarse_slots.resize(8);
      coarse_flag.resize(1);
      halt_flag.resize(max_pipeline_depth + 1);
      setup_storage();
      compute_dirichlet_values(fine, dirichlet_values);
    }

    ~Runtime()
    {
      for (auto &entry : linear_states)
        for (auto &e : entry.second.vg_exec)
          if (e)
            (void) hipGraphExecDestroy(e);
      if (stream)
        (void) hipStreamDestroy(stream);
    }

    Runtime(const Runtime &) = delete;
    Runtime &operator=(const Runtime &) = delete;

    // ---------------------------------------------------------------- geometry
    GridDesc
    make_grid(unsigned int refinement) const
    {
      const auto &mesh = inputs.spatial_discretization.rectangular_mesh;
      GridDesc    g {};
      g.dim     = dim;
      g.degree  = degree;
      g.n_nodes = 1;
      g.n_cells = 1;
      for (int d = 0; d < 3; ++d)
        {
          if (d < dim)
            {
              g.cells[d]    = int(mesh.subdivisions[d] << refinement);
              g.periodic[d] = mesh.periodic_directions.count(d) ? 1 : 0;
              g.nodes[d]    = degree * g.cells[d] + (g.periodic[d] ? 0 : 1);
              g.lower[d]    = mesh.lower_bound[d];
              g.h[d]        = (mesh.size[d] - mesh.lower_bound[d]) / g.cells[d];
              if (!(g.h[d] > 0.0))
                throw Error(status_invalid, "mesh size must exceed the lower bound");
            }
          else
            {
              g.cells[d]    = 1;
              g.nodes[d]    = 1;
              g.periodic[d] = 0;
              g.lower[d]    = 0.0;
              g.h[d]        = 1.0;
            }
          g.n_nodes *= g.nodes[d];
          g.n_cells *= g.cells[d];
        }
      return g;
    }

    /// deal.II's node count with periodic images, (p n + 1)^d.
    int64_t
    full_size(const GridDesc &g) const
    {
      int64_t n = 1;
      for (int d = 0; d < dim; ++d)
        n *= int64_t(degree) * g.cells[d] + 1;
      return n;
    }

    /// Dirichlet face mask of a field (faces in periodic directions never count).
    unsigned int
    dirichlet_mask(int f) const
    {
      unsigned int m = 0;
      for (const auto &[comp, conds] : fields[f].boundary_conditions.component_constraints)
        for (const auto &[bid, c] : conds)
          if (c == prismspf::Condition::Dirichlet && bid < 2u * dim &&
              !fine.periodic[bid / 2])
            m |= 1u << bid;
      return m;
    }

    // ---------------------------------------------------------------- storage
    void
    setup_storage()
    {
      // Oldest age read of each field, over every block (reference
      // dependency_extents.h:75-133 computes it per block group; the union is the same
      // for rotation because each field belongs to exactly one block).
      std::vector<unsigned int> max_age(fields.size(), 0);
      for (const auto &b : blocks)
        for (const auto *deps : {&b.dependencies_rhs, &b.dependencies_lhs})
          for (const auto &[f, dep] : *deps)
            max_age[f] = std::max(max_age[f], dep.oldest_age());
      storage.resize(fields.size());
      for (std::size_t f = 0; f < fields.size(); ++f)
        {
          if (max_age[f] > PPF_MAX_AGE)
            throw Error(status_unsupported, "dependencies older than old_4 are not supported");
          storage[f].resize(1 + max_age[f]);
          for (auto &buf : storage[f])
            {
              buf.resize(fine.n_nodes);
              fill(buf.get(), 0.0, fine.n_nodes);
            }
        }
      sync();
    }

    /// Evaluates set_dirichlet at every Dirichlet node of each field; lower boundary ids
    /// win at shared edges, as deal.II keeps the first constraint of a DoF.
    void
    compute_dirichlet_values(const GridDesc &g, std::vector<DeviceBuffer<double>> &out)
    {
      out.clear();
      out.resize(fields.size());
      dirichlet_nodes.assign(fields.size(), {});
      dirichlet_bids.assign(fields.size(), {});
      for (std::size_t f = 0; f < fields.size(); ++f)
        {
          const unsigned int mask = dirichlet_mask(int(f));
          if (!mask)
            continue;
          std::vector<double> host(g.n_nodes, 0.0);
          for (int64_t node = 0; node < g.n_nodes; ++node)
            {
              int i[3];
              i[0] = int(node % g.nodes[0]);
              i[1] = int((node / g.nodes[0]) % g.nodes[1]);
              i[2] = dim == 3 ? int(node / (int64_t(g.nodes[0]) * g.nodes[1])) : 0;
              for (unsigned int bid = 0; bid < 2u * dim; ++bid)
                if (on_masked_face(g, int(bid / 2), i[bid / 2], mask & (1u << bid)))
                  {
                    host[node] = dirichlet_at(int(f), bid, i);
                    dirichlet_nodes[f].push_back(node);
                    dirichlet_bids[f].push_back(bid);
                    break;
                  }
            }
          out[f].resize(g.n_nodes);
          PPF_HIP_CHECK(hipMemcpyAsync(out[f].get(), host.data(), sizeof(double) * g.n_nodes,
                                       hipMemcpyHostToDevice, stream));
        }
      sync();
    }

    /// set_dirichlet of field f at lattice node i (boundary id bid) at the current time.
    double
    dirichlet_at(int f, unsigned int bid, const int (&i)[3])
    {
      dealii::Point<dim> p;
      for (int d = 0; d < dim; ++d)
        p[d] = fine.lower[d] + i[d] * fine.h[d] / degree;
      double     value = 0.0, vec = 0.0;
      const auto &lc    = layout[f];
      pde_d->set_dirichlet(unsigned(lc.declared), bid, lc.component, p, timer, value, vec);
      return lc.vector ? vec : value;
    }

    bool
    any_time_dependent_dirichlet() const
    {
      for (std::size_t f = 0; f < fields.size(); ++f)
        if (fields[f].boundary_conditions.time_dependent && !dirichlet_nodes[f].empty())
          return true;
      return false;
    }

    /// ConstraintManager::update_time_dependent_constraints (reference
    /// constraint_manager.cc:380-423, called first in every increment, problem.cc:
    /// 428-430): the Dirichlet values of fields whose boundary conditions are time
    /// dependent are re-evaluated at the current time, on their boundary nodes only,
    /// and scattered into the device arrays (and their FP32 copies).  Multigrid levels
    /// use homogeneous constraints and do not change.
    void
    update_time_dependent_dirichlet()
    {
      dirichlet_idx.resize(fields.size());
      dirichlet_stage.resize(fields.size());
      dirichlet_host.resize(fields.size());
      for (std::size_t f = 0; f < fields.size(); ++f)
        {
          const auto &nodes = dirichlet_nodes[f];
          if (!fields[f].boundary_conditions.time_dependent || nodes.empty())
            continue;
          const int64_t count = int64_t(nodes.size());
          auto         &host  = dirichlet_host[f];
          host.resize(nodes.size());
          for (int64_t k = 0; k < count; ++k)
            {
              const int64_t node = nodes[k];
              int           i[3];
              i[0]    = int(node % fine.nodes[0]);
              i[1]    = int((node / fine.nodes[0]) % fine.nodes[1]);
              i[2]    = dim == 3 ? int(node / (int64_t(fine.nodes[0]) * fine.nodes[1])) : 0;
              host[k] = dirichlet_at(int(f), dirichlet_bids[f][k], i);
            }
          if (dirichlet_idx[f].size() != count)
            {
              dirichlet_idx[f].resize(count);
              PPF_HIP_CHECK(hipMemcpyAsync(dirichlet_idx[f].get(), nodes.data(),
                                           sizeof(int64_t) * count, hipMemcpyHostToDevice,
                                           stream));
            }
          dirichlet_stage[f].resize(count);
          PPF_HIP_CHECK(hipMemcpyAsync(dirichlet_stage[f].get(), host.data(),
                                       sizeof(double) * count, hipMemcpyHostToDevice, stream));
          hipLaunchKernelGGL(scatter_kernel<double>, dim3(grid_for(count, 256)), dim3(256), 0,
                             stream, dirichlet_values[f].get(), dirichlet_idx[f].get(),
                             dirichlet_stage[f].get(), count);
          PPF_HIP_CHECK(hipGetLastError());
          if (f < dirichlet_values_f.size() && dirichlet_values_f[f].get())
            convert(dirichlet_values_f[f].get(), dirichlet_values[f].get(), fine.n_nodes);
        }
      // The host staging vectors are reused by the next increment.
      sync();
    }

    double *
    field_ptr(int f, int age)
    {
      if (age == -1 && f >= 0 && f < int(fields.size()))
        age = latest_age(f);
      if (f < 0 || f >= int(fields.size()) || age < 0 || age >= int(storage[f].size()))
        throw Error(status_invalid, "field/age out of range");
      return storage[f][age].get();
    }

    /// Copies a field from host memory given over deal.II's full node set (periodic
    /// images included, lexicographic, x fastest); image entries are ignored.
    void
    set_field(int f, int age, const double *host, int64_t n)
    {
      const GridDesc &g = fine;
      if (n != full_size(g))
        throw Error(status_invalid, "set_field: expected " + std::to_string(full_size(g)) +
                                      " values");
      std::vector<double> uniq(g.n_nodes);
      const int64_t       f0 = int64_t(degree) * g.cells[0] + 1;
      const int64_t       f1 = dim > 1 ? int64_t(degree) * g.cells[1] + 1 : 1;
      for (int64_t node = 0; node < g.n_nodes; ++node)
        {
          const int64_t i0 = node % g.nodes[0];
          const int64_t i1 = (node / g.nodes[0]) % g.nodes[1];
          const int64_t i2 = dim == 3 ? node / (int64_t(g.nodes[0]) * g.nodes[1]) : 0;
          uniq[node]       = host[i0 + f0 * (i1 + f1 * i2)];
        }
      PPF_HIP_CHECK(hipMemcpyAsync(field_ptr(f, age), uniq.data(), sizeof(double) * g.n_nodes,
                                   hipMemcpyHostToDevice, stream));
      sync();
    }

    void
    get_field(int f, int age, double *host, int64_t n)
    {
      const GridDesc &g = fine;
      if (n != full_size(g))
        throw Error(status_invalid, "get_field: expected " + std::to_string(full_size(g)) +
                                      " values");
      std::vector<double> uniq(g.n_nodes);
      PPF_HIP_CHECK(hipMemcpyAsync(uniq.data(), field_ptr(f, age), sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToHost, stream));
      sync();
      const int64_t f0 = int64_t(degree) * g.cells[0] + 1;
      const int64_t f1 = dim > 1 ? int64_t(degree) * g.cells[1] + 1 : 1;
      const int64_t f2 = dim > 2 ? int64_t(degree) * g.cells[2] + 1 : 1;
      for (int64_t j2 = 0; j2 < f2; ++j2)
        for (int64_t j1 = 0; j1 < f1; ++j1)
          for (int64_t j0 = 0; j0 < f0; ++j0)
            {
              const int64_t i0 = j0 % g.nodes[0], i1 = j1 % g.nodes[1], i2 = j2 % g.nodes[2];
              host[j0 + f0 * (j1 + f1 * j2)] = uniq[node_index(g, int(i0), int(i1), int(i2))];
            }
    }

    /// SolutionOutput at an output increment (reference problem.cc:500-515 and
    /// solution_output.h:53-126, VTU branch): every field's current state, to
    /// `<directory>/<file name>_<increment>.vtu`, the increment zero-padded to the digits
    /// of the number of increments.  Written only when the option write_output is set.
    void
    write_output_files()
    {
      const auto &op = inputs.output_parameters;
      if (op.file_type != "vtu")
        throw Error(status_unsupported, "output file type '" + op.file_type +
                                          "' is not supported (vtu)");
      if (op.subdivisions != 0 && op.subdivisions != unsigned(degree))
        throw Error(status_unsupported,
                    "output subdivisions other than the element degree are not supported");
      const unsigned int n_inc  = inputs.temporal_discretization.n_increments;
      const int          digits = n_inc == 0 ? 1 : int(std::floor(std::log10(double(n_inc)))) + 1;
      std::ostringstream name;
      name << op.file_name << "_" << std::setw(digits) << std::setfill('0')
           << timer.get_increment() << ".vtu";
      const std::filesystem::path path = std::filesystem::path(op.directory) / name.str();
      std::filesystem::create_directories(path.parent_path().empty() ? "." : path.parent_path());
      int64_t n[3] = {1, 1, 1};
      double  lower[3] = {0, 0, 0}, step[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d)
        {
          n[d]     = int64_t(degree) * fine.cells[d] + 1;
          lower[d] = fine.lower[d];
          step[d]  = fine.h[d] / degree;
        }
      // One output array per declared field: the dim component fields of a vector field
      // (stored as <name>_x, _y, _z) form one 3-component array named <name>.
      std::vector<std::string>         names;
      std::vector<std::vector<double>> values;
      std::vector<int>                 groups;
      for (std::size_t f = 0; f < fields.size(); ++f)
        {
          if (!layout[f].vector)
            {
              names.push_back(fields[f].name);
              groups.push_back(1);
            }
          else if (layout[f].component == 0)
            {
              names.push_back(fields[f].name.substr(0, fields[f].name.size() - 2));
              groups.push_back(dim);
            }
          values.emplace_back(full_size(fine));
          get_field(int(f), 0, values.back().data(), int64_t(values.back().size()));
        }
      write_vtu(path.string(), dim, n, lower, step, names, values, groups, timer.get_time(),
                timer.get_increment());
    }

    // ---------------------------------------------------------------- utilities
    void
    sync()
    {
      ++sync_count_;
      PPF_HIP_CHECK(hipStreamSynchronize(stream));
    }

    template <typename T>
    void
    fill(T *x, T v, int64_t n)
    {
      if (n <= 0)
        return;
      hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, stream, x, v, n);
    }

    template <typename T>
    void
    copy(T *dst, const T *src, int64_t n)
    {
      PPF_HIP_CHECK(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToDevice, stream));
    }

    /// dst = src converted on the stream; skipped on the device once *halt is set.
    template <typename T, typename U>
    void
    convert(T *dst, const U *src, int64_t n, const int *halt = nullptr)
    {
      hipLaunchKernelGGL((convert_kernel<T, U>), dim3(grid_for(n, 256)), dim3(256), 0, stream,
                         dst, src, n, halt);
    }

    template <typename T>
    void
    axpby(T *y, T a, const T *x, T b, int64_t n)
    {
      hipLaunchKernelGGL(axpby_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, stream, y, a,
                         x, b, n);
    }

    template <typename T>
    void
    scale(T *y, T b, int64_t n)
    {
      hipLaunchKernelGGL(scale_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, stream, y, b, n);
    }

    /// Deterministic dot product; the result is copied to the host.
    template <typename T>
    double
    dot(const T *x, const T *y, int64_t n)
    {
      hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(reduce_blocks), dim3(reduce_threads), 0,
                         stream, reduce_buf.get(), x, y, n);
      return finish_reduction(reduce_blocks);
    }

    /// Deterministic dot product into a device slot (no host wait).
    template <typename T>
    void
    dot_to(const T *x, const T *y, int64_t n, double *slot)
    {
      hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(reduce_blocks), dim3(reduce_threads), 0,
                         stream, reduce_buf.get(), x, y, n);
      hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(reduce_threads), 0, stream, slot,
                         reduce_buf.get(), reduce_blocks);
    }

    double
    finish_reduction(int count, int offset = 0)
    {
      hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(reduce_threads), 0, stream,
                         scalar_buf.get(), reduce_buf.get() + offset, count);
      double r = 0.0;
      PPF_HIP_CHECK(hipMemcpyAsync(&r, scalar_buf.get(), sizeof(double), hipMemcpyDeviceToHost,
                                   stream));
      sync();
      return r;
    }

    // ------------------------------------------------------------ operator plans
    struct OpKey
    {
      int  block;
      bool lhs;
      int  precision; // 0 double, 1 float
      bool
      operator<(const OpKey &o) const
      {
        return std::tie(block, lhs, precision) < std::tie(o.block, o.lhs, o.precision);
      }
    };

    struct PlanEntry
    {
      SplitPlan    plan;
      double       time      = 0.0;
      double       dt        = 0.0;
      unsigned int increment = 0;
      /// The split path was compared with the generic path on real fields (first use).
      bool verified = false;
    };

    /// The split plan of block b's rhs or lhs, re-derived when the time step changes and,
    /// for coefficients that depend on the time or the increment, when either changes.
    template <typename number>
    PlanEntry &
    split_plan(int b, bool lhs)
    {
      const OpKey key {b, lhs, std::is_same_v<number, float> ? 1 : 0};
      auto        it = plans.find(key);
      if (it != plans.end() && it->second.dt == timer.get_timestep() &&
          !(it->second.plan.time_dependent &&
            (it->second.time != timer.get_time() ||
             it->second.increment != timer.get_increment())))
        return it->second;
      const auto &pde = *pde_of<number>();
      HookProbe<PDE<number>, dim, degree, number> probe(pde, blocks[b], lhs);
      PlanEntry entry;
      entry.plan      = probe.analyse(timer.get_time(), timer.get_timestep(), timer.get_increment());
      ++graph_epoch;
      entry.time      = timer.get_time();
      entry.dt        = timer.get_timestep();
      entry.increment = timer.get_increment();
      // A plan that was rejected by the runtime comparison stays rejected.
      if (it != plans.end() && !it->second.plan.eligible && it->second.verified)
        {
          entry.plan.eligible = false;
          entry.plan.reason   = it->second.plan.reason;
          entry.verified      = true;
        }
      else if (it != plans.end())
        entry.verified = it->second.verified && entry.plan.eligible;
      return plans[key] = entry, plans[key];
    }

    template <typename number>
    const PDE<number> *
    pde_of() const
    {
      if constexpr (std::is_same_v<number, float>)
        return pde_f.get();
      else
        return pde_d.get();
    }

    /// Level-local view of the non-operand fields an operator reads.
    template <typename number>
    struct LevelFields
    {
      const GridDesc *grid = nullptr;
      /// ptr[f][slot] for current/old slots; nullptr where not provided.
      const number *ptr[PPF_MAX_FIELDS][PPF_N_SLOTS] = {};
    };

    LevelFields<double>
    fine_fields()
    {
      LevelFields<double> lf;
      lf.grid = &fine;
      for (std::size_t f = 0; f < storage.size(); ++f)
        for (std::size_t a = 0; a < storage[f].size(); ++a)
          lf.ptr[f][a + 1] = storage[f][a].get();
      return lf;
    }

    /// FP32 copies of the field states that the rhs or lhs of block `b` reads (its
    /// declared non-src dependencies), for FP32 operator applications on the fine grid.
    LevelFields<float>
    fp32_fields(int b, bool lhs, std::vector<DeviceBuffer<float>> &shadow)
    {
      LevelFields<float> lf {};
      lf.grid               = &fine;
      const auto &blk       = blocks[b];
      constexpr int src_slot = prismspf::dependency_slot(prismspf::SRC);
      shadow.clear();
      shadow.reserve(PPF_MAX_FIELDS * (PPF_MAX_AGE + 1));
      for (const auto &in : declared_inputs(lhs ? blk.dependencies_lhs : blk.dependencies_rhs))
        {
          if (in.slot == src_slot || lf.ptr[in.field][in.slot])
            continue;
          const int age = in.slot - 1;
          if (age < 0 || age >= int(storage[in.field].size()))
            throw Error(status_internal, "dependency storage not available");
          float *copy = nullptr;
          if (borrowed.base)
            copy = carve<float>(fine.n_nodes, "the FP32 copies of the operator's inputs");
          else
            {
              shadow.emplace_back(fine.n_nodes);
              copy = shadow.back().get();
            }
          convert(copy, storage[in.field][age].get(), fine.n_nodes);
          lf.ptr[in.field][in.slot] = copy;
        }
      return lf;
    }

    /// Applies the rhs or lhs operator of block `b`:
    ///   dst = scaling * (assembled hook terms) with the Dirichlet rule `dwrite`.
    /// `src[k]`/`dst[k]` belong to the k-th field of the block (ascending field index).
    template <typename number>
    void
    apply_operator(int b, bool lhs, const LevelFields<number> &lf,
                   const std::vector<const number *> &src, const std::vector<number *> &dst,
                   int scaling, int dwrite, bool read_plain,
                   const std::vector<DeviceBuffer<number>> *dvalues = nullptr)
    {
      const prismspf::SolveBlock &blk = blocks[b];
      const GridDesc             &g   = *lf.grid;
      DepTable<number>            deps {};
      OutTable<number>            out {};
      std::vector<int>            bf(blk.field_indices.begin(), blk.field_indices.end());
      if (dst.size() != bf.size() || (!src.empty() && src.size() != bf.size()))
        throw Error(status_invalid, "operator: block vector size mismatch");
      for (std::size_t f = 0; f < fields.size(); ++f)
        deps.dirichlet_mask[f] = dirichlet_mask(int(f));
      constexpr int src_slot = prismspf::dependency_slot(prismspf::SRC);
      for (const auto &in : declared_inputs(lhs ? blk.dependencies_lhs : blk.dependencies_rhs))
        {
          unsigned int  fl  = in.flags;
          const number *ptr = nullptr;
          if (in.slot == src_slot)
            {
              const auto pos = std::find(bf.begin(), bf.end(), in.field);
              if (pos == bf.end())
                throw Error(status_invalid, "src dependency on a field outside the block");
              if (src.empty())
                throw Error(status_invalid, "operator needs a src vector");
              ptr = src[pos - bf.begin()];
              if (!read_plain)
                fl |= slot_constrained;
            }
          else
            ptr = lf.ptr[in.field][in.slot];
          if (!ptr)
            throw Error(status_internal, "dependency storage not available on this level");
          deps.ptr[in.field][in.slot]   = ptr;
          deps.flags[in.field][in.slot] = fl;
        }
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        out.scratch_slot[f] = -1;
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          const int f            = bf[k];
          out.dst[f]             = dst[k];
          out.dirichlet_mask[f]  = deps.dirichlet_mask[f];
          out.scratch_slot[f]    = int(k);
          out.out_mask          |= 1u << f;
          if (dwrite == dirichlet_value && out.dirichlet_mask[f])
            {
              if (!dvalues || !(*dvalues)[f].get())
                throw Error(status_internal, "missing Dirichlet values");
              out.dirichlet_values[f] = (*dvalues)[f].get();
            }
        }
      out.scaling         = scaling;
      out.dirichlet_write = dwrite;
      out.halt            = enqueue_halt;

      PlanEntry       &entry = split_plan<number>(b, lhs);
      const SplitPlan &plan  = entry.plan;
      bool use_split = path == OperatorPath::Split || path == OperatorPath::SplitNode ||
                       (path == OperatorPath::Auto && plan.eligible);
      if (use_split && !plan.eligible)
        throw Error(status_unsupported, "split path requested but the hook does not "
                                        "qualify: " + plan.reason);
      const auto        &pde = *pde_of<number>();
      const unsigned int nb  = grid_for(g.n_nodes, block_size);
      auto launch = [&](bool split, const OutTable<number> &o) {
        dispatch_block_id(typename App::BlockIds {}, blk.id, [&](auto bid) {
          constexpr int id = decltype(bid)::value;
          if (split && path != OperatorPath::SplitNode &&
              launch_stencil<number, id>(b, lhs, g, deps, o, plan))
            return;
          if (split)
            {
              if (lhs)
                hipLaunchKernelGGL((split_operator_kernel<PDE<number>, dim, degree, number, true, id>),
                                   dim3(nb), dim3(block_size), 0, stream, pde, g, deps, o,
                                   plan.terms, timer);
              else
                hipLaunchKernelGGL((split_operator_kernel<PDE<number>, dim, degree, number, false, id>),
                                   dim3(nb), dim3(block_size), 0, stream, pde, g, deps, o,
                                   plan.terms, timer);
            }
          else if (!launch_cell_block<number, id>(lhs, g, deps, o))
            {
              number *scratch = ensure_scratch<number>(int64_t(bf.size()) *
                                                       dofs_per_cell<dim, degree> * g.n_cells);
              const unsigned int nc = grid_for(g.n_cells, block_size);
              if (lhs)
                hipLaunchKernelGGL(
                  (cell_operator_kernel<PDE<number>, dim, degree, number, true, id, false, App::n_fields>),
                  dim3(nc), dim3(block_size), 0, stream, pde, g, deps, o, scratch, timer, -1);
              else
                hipLaunchKernelGGL(
                  (cell_operator_kernel<PDE<number>, dim, degree, number, false, id, false, App::n_fields>),
                  dim3(nc), dim3(block_size), 0, stream, pde, g, deps, o, scratch, timer, -1);
              hipLaunchKernelGGL((node_gather_kernel<dim, degree, number>), dim3(nb),
                                 dim3(block_size), 0, stream, g, o, scratch);
            }
        });
        PPF_HIP_CHECK(hipGetLastError());
      };

      // First use of an automatically chosen split plan on the fine lattice: the split
      // path's result is compared with the generic path's on the real fields, and the
      // plan is dropped (for good, with a line in the output log) if they disagree.  This
      // catches coefficients that branch on values or times the structure probe did not
      // visit.
      if (use_split && path == OperatorPath::Auto && !entry.verified && lf.grid == &fine &&
          !std::any_of(std::begin(dst), std::end(dst),
                       [&](const number *d) {
                         for (const number *sp : src)
                           if (sp == d)
                             return true;
                         return false;
                       }))
        {
          const int64_t n   = g.n_nodes;
          number       *ref = nullptr;
          if (borrowed.base)
            {
              if (borrowed.used + int64_t(bf.size()) * n * int64_t(sizeof(number)) + 256 +
                    int64_t(bf.size()) * dofs_per_cell<dim, degree> * g.n_cells *
                      int64_t(sizeof(number)) + 256 <=
                  borrowed.bytes)
                ref = carve<number>(int64_t(bf.size()) * n, "the split-path comparison");
            }
          else
            {
              auto &buf = [&]() -> DeviceBuffer<number> & {
                if constexpr (std::is_same_v<number, float>)
                  return verify_f;
                else
                  return verify_d;
              }();
              if (buf.size() < int64_t(bf.size()) * n)
                buf.resize(int64_t(bf.size()) * n);
              ref = buf.get();
            }
          if (ref)
            {
              OutTable<number> oref = out;
              for (std::size_t k = 0; k < bf.size(); ++k)
                oref.dst[bf[k]] = ref + int64_t(k) * n;
              launch(false, oref);
              launch(true, out);
              double worst = 0.0, scale = 0.0;
              for (std::size_t k = 0; k < bf.size(); ++k)
                {
                  hipLaunchKernelGGL(max_diff_partial_kernel<number>, dim3(reduce_blocks),
                                     dim3(reduce_threads), 0, stream, reduce_buf.get(),
                                     dst[k], oref.dst[bf[k]], n);
                  std::vector<double> part(2 * reduce_blocks);
                  PPF_HIP_CHECK(hipMemcpyAsync(part.data(), reduce_buf.get(),
                                               sizeof(double) * part.size(),
                                               hipMemcpyDeviceToHost, stream));
                  sync();
                  for (int i = 0; i < reduce_blocks; ++i)
                    {
                      worst = std::max(worst, part[i]);
                      scale = std::max(scale, part[reduce_blocks + i]);
                    }
                }
              int halted_now = 0;
              if (out.halt)
                PPF_HIP_CHECK(hipMemcpy(&halted_now, out.halt, sizeof(int),
                                        hipMemcpyDeviceToHost));
              if (!halted_now)
                {
                  const double tol = std::is_same_v<number, float> ? 1e-4 : 1e-11;
                  entry.verified   = true;
                  if (!(worst <= tol * scale) && !(scale == 0.0 && worst == 0.0))
                    {
                      std::ostringstream msg;
                      msg << "solve block " << blk.id << " (" << (lhs ? "lhs" : "rhs")
                          << "): the split operator differs from the generic path by "
                          << worst / std::max(scale, 1e-300)
                          << " (relative) on the current fields; using the generic path";
                      entry.plan.eligible = false;
                      entry.plan.reason   = msg.str();
                      log += msg.str() + "\n";
                      if (std::getenv("PPF_VERBOSE"))
                        std::fprintf(stderr, "%s\n", msg.str().c_str());
                      // the generic result is the operator: copy it into place
                      for (std::size_t k = 0; k < bf.size(); ++k)
                        copy(dst[k], oref.dst[bf[k]], n);
                    }
                }
              return;
            }
        }
      launch(use_split, out);
    }

    /// Launches the one-pass generic kernel (cell_column_kernel in 3-D,
    /// cell_block_kernel otherwise); false if it is switched off or its staged inputs
    /// exceed the shared memory of a workgroup, in which case the two-pass cell/gather
    /// kernels run instead.
    template <typename number, int id>
    bool
    launch_cell_block(bool lhs, const GridDesc &g, const DepTable<number> &deps,
                      const OutTable<number> &out)
    {
      if (!generic_cell_block)
        return false;
      CellBlockArgs cb {};
      for (int f = 0; f < App::n_fields; ++f)
        for (int s = 0; s < PPF_N_SLOTS; ++s)
          if (deps.ptr[f][s] && (deps.flags[f][s] & (slot_values | slot_gradients)))
            {
              if (cb.n_in == cb_max_inputs)
                return false;
              cb.in_field[cb.n_in] = f;
              cb.in_slot[cb.n_in]  = s;
              cb.pos[f][s]         = cb.n_in;
              ++cb.n_in;
            }
      int n_out = 0;
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        n_out += out.scratch_slot[f] >= 0 ? 1 : 0;
      int64_t bytes = cell_block_shared_bytes<dim, degree, number>(cb.n_in, n_out, false);
      if (bytes > cell_block_lds_limit())
        return false;
      if (dim == 3 && n_out > 1 &&
          cell_block_shared_bytes<dim, degree, number>(cb.n_in, n_out, true) <=
            cell_block_lds_limit())
        {
          cb.pair = 1;
          bytes   = cell_block_shared_bytes<dim, degree, number>(cb.n_in, n_out, true);
        }
      int64_t n_blocks = 1;
      if constexpr (dim == 3)
        {
          cb.nblk[0]  = (g.cells[0] + cc_ex - 2) / (cc_ex - 1);
          cb.nblk[1]  = (g.cells[1] + cc_ey - 2) / (cc_ey - 1);
          cb.bcell[0] = cc_ex - 1;
          cb.bcell[1] = cc_ey - 1;
          // z chunks: enough columns to fill every compute unit to the kernel's
          // occupancy, each chunk as long as that allows
          const void *kernel =
            lhs ? reinterpret_cast<const void *>(
                    &cell_column_kernel<PDE<number>, degree, number, true, id>)
                : reinterpret_cast<const void *>(
                    &cell_column_kernel<PDE<number>, degree, number, false, id>);
          int resident = 1;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, cb_threads,
                                                           size_t(bytes)) != hipSuccess)
            {
              (void)hipGetLastError();
              resident = 1;
            }
          const int64_t target = device_cus() * std::max(1, resident);
          const int64_t tiles  = int64_t(cb.nblk[0]) * cb.nblk[1];
          const int64_t chunks = std::max<int64_t>(1, (target + tiles - 1) / tiles);
          cb.bcell[2]            = column_chunk > 0 ? std::min(column_chunk, g.cells[2])
                                                    : int((g.cells[2] + chunks - 1) / chunks);
          cb.nblk[2]             = (g.cells[2] + cb.bcell[2] - 1) / cb.bcell[2];
          n_blocks               = tiles * cb.nblk[2];
        }
      else
        for (int d = 0; d < 3; ++d)
          {
            if (d < dim)
              {
                cb.bcell[d] = cb_extent<dim>(d) - 1;
                cb.nblk[d]  = (g.cells[d] + cb.bcell[d] - 1) / cb.bcell[d];
              }
            else
              {
                cb.bcell[d] = 1;
                cb.nblk[d]  = 1;
              }
            n_blocks *= cb.nblk[d];
          }
      if (n_blocks > int64_t(std::numeric_limits<int>::max()))
        return false;
      const auto &pde = *pde_of<number>();
      auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(unsigned(n_blocks)), dim3(cb_threads),
                           unsigned(bytes), stream, pde, g, deps, out, timer, cb);
      };
      if constexpr (dim == 3)
        {
          if (lhs)
            go(cell_column_kernel<PDE<number>, degree, number, true, id>);
          else
            go(cell_column_kernel<PDE<number>, degree, number, false, id>);
        }
      else
        {
          if (lhs)
            go(cell_block_kernel<PDE<number>, dim, degree, number, true, id>);
          else
            go(cell_block_kernel<PDE<number>, dim, degree, number, false, id>);
        }
      dirichlet_faces(g, out);
      return true;
    }

    /// Writes the constraint values on the Dirichlet faces of every output of an
    /// explicit update (the operator kernels leave those nodes).
    template <typename number>
    void
    dirichlet_faces(const GridDesc &g, const OutTable<number> &out)
    {
      if (out.dirichlet_write != dirichlet_value)
        return;
      int n[3], per[3];
      for (int d = 0; d < 3; ++d)
        {
          n[d]   = d < dim ? g.nodes[d] : 1;
          per[d] = d < dim ? g.periodic[d] : 1;
        }
      int64_t face = 0;
      for (int d = 0; d < 3; ++d)
        face = std::max(face, int64_t(n[(d + 1) % 3]) * n[(d + 2) % 3]);
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        if ((out.out_mask & (1u << f)) && out.dirichlet_mask[f])
          hipLaunchKernelGGL(dirichlet_faces_kernel<number>,
                             dim3(unsigned((face + 255) / 256), 6), dim3(256), 0, stream,
                             out.dst[f], out.dirichlet_values[f], n[0], n[1], n[2], per[0],
                             per[1], per[2], out.dirichlet_mask[f], out.halt);
    }

    /// Compute units of the current device.
    int64_t
    device_cus()
    {
      if (n_cus_ == 0)
        {
          int dev = 0, v = 0;
          PPF_HIP_CHECK(hipGetDevice(&dev));
          PPF_HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
          n_cus_ = std::max(1, v);
        }
      return n_cus_;
    }

    /// Dynamic shared memory one cell-block launch may use: the device's per-workgroup
    /// limit less room for the kernels' static tables.
    int64_t
    cell_block_lds_limit()
    {
      if (lds_limit_ == 0)
        {
          int dev = 0, v = 0;
          PPF_HIP_CHECK(hipGetDevice(&dev));
          PPF_HIP_CHECK(
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
          lds_limit_ = std::max<int64_t>(int64_t(v) - 4096, 0);
        }
      return lds_limit_;
    }

    /// Launches the tiled stencil kernel for a split operator; false if the operator
    /// has more stencil operands than the kernel stages (the per-node kernel runs it).
    /// The kernel computes in `number` and loads and stores arrays of `ST` (ST = double
    /// with number = float: FP32 arithmetic on the FP64 state, no converted copies).
    template <typename number, int id, typename ST = number>
    bool
    launch_stencil(int b, bool lhs, const GridDesc &g, const DepTable<ST> &deps,
                   const OutTable<ST> &out, const SplitPlan &plan)
    {
      // Small lattices (multigrid's coarse levels) go to the one-thread-per-node kernel:
      // a tile's march over its z planes would take longer than the whole lattice node
      // by node.  The fused coarse solve rides on this launch and keeps it.
      if (g.n_nodes < stencil_min_nodes && !coarse_pending &&
          !(probe_req.pending && probe_req.block == b))
        return false;
      StencilArgs<number, ST> a {};
      // Distinct (field, slot) operands of the gradient terms.
      int op_field[stencil_max_ops], op_slot[stencil_max_ops];
      a.n_ops = 0;
      for (int t = 0; t < plan.terms.n_terms; ++t)
        {
          const SplitTerm &tm = plan.terms.term[t];
          int              k  = 0;
          while (k < a.n_ops && !(op_field[k] == tm.in_field && op_slot[k] == tm.in_slot))
            ++k;
          if (k == a.n_ops)
            {
              if (a.n_ops == stencil_max_ops)
                return false;
              op_field[k] = tm.in_field;
              op_slot[k]  = tm.in_slot;
              a.op[k]     = deps.ptr[tm.in_field][tm.in_slot];
              a.op_dmask[k] =
                (deps.flags[tm.in_field][tm.in_slot] & slot_constrained)
                  ? deps.dirichlet_mask[tm.in_field]
                  : 0u;
              if (!a.op[k])
                throw Error(status_internal, "stencil operand without storage");
              ++a.n_ops;
            }
          a.beta[tm.out_field][k] += number(tm.beta);
        }
      // Unused operand and value entries point at an array of the lattice size so the
      // kernel loads them without a branch (their values are never used).
      const ST *any_array = nullptr;
      for (int f = 0; f < PPF_MAX_FIELDS && !any_array; ++f)
        if (out.out_mask & (1u << f))
          any_array = out.dst[f];
      if (!any_array)
        throw Error(status_internal, "operator without an output field");
      // An operator without gradient terms still runs with one operand, which reads
      // zeros so that its (unused) sums stay finite; operands beyond the instantiated
      // count are never read and point at any lattice-size array.
      for (int k = a.n_ops; k < stencil_max_ops; ++k)
        a.op[k] = k == 0 ? zero_array<ST>(g.n_nodes) : any_array;
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        for (int s = 0; s < PPF_N_SLOTS; ++s)
          {
            a.val_alias[f][s] = -1;
            a.val[f][s]       = any_array;
            if (!(deps.flags[f][s] & slot_values))
              continue;
            a.val[f][s]       = deps.ptr[f][s];
            a.val_dmask[f][s] = (deps.flags[f][s] & slot_constrained) ? deps.dirichlet_mask[f] : 0u;
            for (int k = 0; k < a.n_ops; ++k)
              if (op_field[k] == f && op_slot[k] == s)
                a.val_alias[f][s] = static_cast<signed char>(k);
          }

      double volume = 1.0;
      for (int d = 0; d < 3; ++d)
        {
          a.n[d]        = d < dim ? g.nodes[d] : 1;
          a.periodic[d] = d < dim ? g.periodic[d] : 0;
          if (a.periodic[d] && a.n[d] < 2)
            return false;
          // Interior taps (centre, +-1, +-2) in units of 1/h^2 per coordinate parity;
          // see stencil_kernels.hpp for their derivation from the GLL stiffness.
          const double tq2[2][3] = {{14.0, -8.0, 1.0}, {8.0, -4.0, 0.0}};
          const double tq1[3]    = {2.0, -1.0, 0.0};
          for (int par = 0; par < 2; ++par)
            for (int k = 0; k < 3; ++k)
              {
                const double t = degree == 2 ? tq2[par][k] : tq1[k];
                a.tap[d][par][k] = d < dim ? number(t / (g.h[d] * g.h[d])) : number(0);
              }
          for (int c = 0; c < 3; ++c)
            {
              const double m = d < dim ? class_mass_units<degree>(c) * g.h[d] : 1.0;
              a.sfac[d][c]   = number(out.scaling == scale_none ? m
                                      : out.scaling == scale_invm ? 1.0 : std::sqrt(m));
            }
          if (d < dim)
            {
              a.q_lower[d] = g.lower[d];
              a.q_step[d]  = g.h[d] / degree;
              volume *= g.h[d];
            }
        }
      a.element_volume = volume;
      a.any_dmask      = 0;
      for (int k = 0; k < a.n_ops; ++k)
        a.any_dmask |= a.op_dmask[k];
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        {
          for (int s = 0; s < PPF_N_SLOTS; ++s)
            a.any_dmask |= a.val_dmask[f][s];
          if (out.out_mask & (1u << f))
            a.any_dmask |= out.dirichlet_mask[f];
        }
      // 32-bit byte offsets in the kernel: arrays up to 4 GiB.
      if (int64_t(a.n[0]) * a.n[1] * a.n[2] * int64_t(sizeof(ST)) > (int64_t(1) << 32))
        return false;
      // Store policy: non-temporal when the distinct arrays the launch reads and writes
      // exceed 192 MiB, i.e. cannot stay in the 256 MB Infinity Cache (the output would
      // not be read from the caches by the next launch anyway).  Measured: the FP64 Q2
      // Poisson apply at 257^3 nodes (271 MB) is 11 % faster with non-temporal stores,
      // its FP32 form (136 MB) 15 % slower, and the Cahn-Hilliard increment at 129^3
      // (51 MB per launch) 6 % slower, although each of its launches alone is faster:
      // the next block reads the output from the caches.
      {
        const void *arrays[stencil_max_ops + PPF_MAX_FIELDS * (PPF_N_SLOTS + 1)];
        int         n_arrays = 0;
        auto        add      = [&](const void *p) {
          for (int k = 0; k < n_arrays; ++k)
            if (arrays[k] == p)
              return;
          arrays[n_arrays++] = p;
        };
        for (int k = 0; k < a.n_ops; ++k)
          add(a.op[k]);
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
          {
            for (int s = 0; s < PPF_N_SLOTS; ++s)
              if (deps.flags[f][s] & slot_values)
                add(a.val[f][s]);
            if (out.out_mask & (1u << f))
              add(out.dst[f]);
          }
        const int64_t footprint = int64_t(n_arrays) * g.n_nodes * int64_t(sizeof(ST));
        a.nt_store = stencil_store_policy == StorePolicy::NonTemporal ||
                     (stencil_store_policy == StorePolicy::Auto &&
                      footprint > stencil_nt_min_bytes);
      }
      // A pending coarse-grid solve takes this operator (src = p, dst = ap) and runs the
      // whole CG in one workgroup instead of one launch per vector operation.
      if constexpr (std::is_same_v<number, double> && std::is_same_v<ST, double>)
        if (coarse_pending && lhs)
          {
            coarse_pending = false;
            auto launch    = [&](auto ops_c) {
              constexpr int NS = decltype(ops_c)::value;
              hipLaunchKernelGGL((coarse_cg_kernel<PDE<double>, dim, degree, id, App::n_fields, NS>),
                                 dim3(1), dim3(reduce_threads), 0, stream, *pde_of<double>(), a,
                                 out, timer, coarse_req);
            };
            if (a.n_ops > 1)
              launch(std::integral_constant<int, 2> {});
            else
              launch(std::integral_constant<int, 1> {});
            PPF_HIP_CHECK(hipGetLastError());
            return true;
          }

      // Tile shape: waves per workgroup x rows per wave.  Default 4 x 2 (measured
      // faster than 4 x 4 at 129^3 and 257^3: fewer registers); two staged operands
      // take 4 x 2 or 8 x 1 (more rows would spill in FP64).
      int rows = stencil_rows, waves = stencil_waves;
      if (rows == 0)
        rows = waves == 8 ? 1 : 2;
      if (waves == 0)
        waves = 4;
      const bool shape_ok = (waves == 4 && (rows == 2 || rows == 4)) ||
                            (waves == 8 && (rows == 1 || rows == 2));
      if (!shape_ok)
        throw Error(status_invalid, "stencil tile: waves x rows must be 4x2, 4x4, 8x1 or 8x2");
      if (a.n_ops > 1 && rows * waves > 8)
        {
          waves = 4;
          rows  = 2;
        }
      const int ty = waves * rows;
      // Full tiles only where a few columns or rows would otherwise occupy a whole tile
      // row (e.g. 257 = 4 x 64 + 1): those are left to the remainder workgroups.
      int       tx_n = a.n[0], ty_n = a.n[1];
      const int rx   = a.n[0] % stencil_tx, ry = a.n[1] % ty;
      if (a.n[0] >= 2 * stencil_tx && rx != 0 && rx <= 8)
        tx_n = a.n[0] - rx;
      if (a.n[1] >= 2 * ty && ry != 0 && ry <= std::max(1, ty / 4))
        ty_n = a.n[1] - ry;
      a.tile_n[0] = tx_n;
      a.tile_n[1] = ty_n;
      a.rem_a     = int64_t(a.n[0] - tx_n) * a.n[1] * a.n[2];
      a.rem_b     = int64_t(tx_n) * (a.n[1] - ty_n) * a.n[2];
      a.tiles_x   = (tx_n + stencil_tx - 1) / stencil_tx;
      a.tiles_y   = (ty_n + ty - 1) / ty;
      if constexpr (dim == 3)
        {
          int zc = stencil_z_chunk;
          if (zc <= 0)
            {
              const int64_t xy = int64_t(a.tiles_x) * a.tiles_y;
              zc = int((int64_t(a.n[2]) * xy + stencil_target_blocks - 1) / stencil_target_blocks);
              zc = std::max(zc, 4 * degree);
            }
          // Q2 chunks start at even z (the kernel takes the z taps' parity from the
          // step's phase).
          if (degree == 2 && (zc & 1))
            ++zc;
          a.z_chunk  = std::min(zc, a.n[2]);
          a.chunks_z = (a.n[2] + a.z_chunk - 1) / a.z_chunk;
        }
      else
        {
          a.z_chunk  = 1;
          a.chunks_z = 1;
        }
      int64_t nblocks = int64_t(a.tiles_x) * a.tiles_y * a.chunks_z;
      a.map_mode      = stencil_map;
      a.ty_per_xcd    = (a.tiles_y + 7) / 8;
      if (a.map_mode == 1)
        nblocks = 8 * int64_t(a.tiles_x) * a.ty_per_xcd * a.chunks_z;
      const int     wg_threads = stencil_tx * waves;
      const int64_t nrem       = (a.rem_a + a.rem_b + wg_threads - 1) / wg_threads;
      if (nblocks + nrem >= (int64_t(1) << 31))
        return false;
      a.n_blocks = int(nblocks);

      const auto &pde = *pde_of<number>();
      // The increment's NaN probe rides on this launch if it is the block's last kernel
      // (no Dirichlet-value faces kernel follows) and node 0 lies in tile 0.
      bool faces_follow = false;
      if (out.dirichlet_write == dirichlet_value)
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
          faces_follow = faces_follow || ((out.out_mask & (1u << f)) && out.dirichlet_mask[f]);
      if constexpr (std::is_same_v<ST, double>)
        if (probe_req.pending && probe_req.block == b && !faces_follow && a.n_blocks > 0)
          {
            a.probe_result = probe_req.result;
            a.probe_halt   = probe_req.halt;
            a.probe_hmask  = probe_req.hmask;
            a.probe_tag    = probe_req.tag;
            a.probe_n      = probe_req.fields.n;
            for (int f = 0; f < PPF_MAX_FIELDS; ++f)
              a.probe_ptr[f] = probe_req.fields.ptr[f];
            probe_req.pending = false;
            probe_req.taken   = true;
          }
      const dim3  grid(unsigned(nblocks + nrem)), threads(stencil_tx, unsigned(waves));
      auto        go = [&](auto waves_c, auto rows_c, auto ops_c) {
        constexpr int WV = decltype(waves_c)::value;
        constexpr int RR = decltype(rows_c)::value;
        constexpr int NS = decltype(ops_c)::value;
        if (lhs)
          hipLaunchKernelGGL(
            (stencil_operator_kernel<PDE<number>, dim, degree, number, true, id, App::n_fields, WV, RR, NS, ST>),
            grid, threads, 0, stream, pde, a, out, timer);
        else
          hipLaunchKernelGGL(
            (stencil_operator_kernel<PDE<number>, dim, degree, number, false, id, App::n_fields, WV, RR, NS, ST>),
            grid, threads, 0, stream, pde, a, out, timer);
      };
      using I1 = std::integral_constant<int, 1>;
      using I2 = std::integral_constant<int, 2>;
      using I4 = std::integral_constant<int, 4>;
      using I8 = std::integral_constant<int, 8>;
      if (a.n_ops > 1)
        {
          if (waves == 8)
            go(I8 {}, I1 {}, I2 {});
          else
            go(I4 {}, I2 {}, I2 {});
        }
      else if (waves == 8)
        {
          if (rows == 1)
            go(I8 {}, I1 {}, I1 {});
          else
            go(I8 {}, I2 {}, I1 {});
        }
      else if (rows == 2)
        go(I4 {}, I2 {}, I1 {});
      else
        go(I4 {}, I4 {}, I1 {});
      dirichlet_faces(g, out);
      return true;
    }

    /// Operator diagonal of the lhs of block `b` for block field k, scaled like the
    /// operator; zero entries (Dirichlet rows) become one and the inverse is returned
    /// (reference mf_operator.cc:82-255).
    template <typename number>
    void
    inverse_diagonal(int b, const LevelFields<number> &lf, int scaling, number *inv_diag)
    {
      const prismspf::SolveBlock &blk = blocks[b];
      const GridDesc             &g   = *lf.grid;
      std::vector<int>            bf(blk.field_indices.begin(), blk.field_indices.end());
      const SplitPlan            &plan = split_plan<number>(b, true).plan;
      const bool                  use_split =
        path != OperatorPath::Generic && plan.eligible && plan.alpha_constant;
      const unsigned int nb = grid_for(g.n_nodes, block_size);
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          const int f   = bf[k];
          number   *out = inv_diag + k * g.n_nodes;
          if (use_split)
            hipLaunchKernelGGL((split_diagonal_kernel<dim, degree, number>), dim3(nb),
                               dim3(block_size), 0, stream, g, out, plan.alpha[f],
                               plan.beta_src[f], dirichlet_mask(f), scaling);
          else
            generic_diagonal<number>(b, lf, int(k), scaling, out);
        }
      for (std::size_t k = 0; k < bf.size(); ++k)
        {
          number *out = inv_diag + k * g.n_nodes;
          hipLaunchKernelGGL(invert_diagonal_kernel<number>, dim3(nb), dim3(block_size), 0,
                             stream, out, out, g.n_nodes);
        }
      PPF_HIP_CHECK(hipGetLastError());
    }

    template <typename number>
    void
    generic_diagonal(int b, const LevelFields<number> &lf, int k, int scaling, number *out_ptr)
    {
      const prismspf::SolveBlock &blk = blocks[b];
      const GridDesc             &g   = *lf.grid;
      std::vector<int>            bf(blk.field_indices.begin(), blk.field_indices.end());
      DepTable<number>            deps {};
      OutTable<number>            out {};
      constexpr int               src_slot = prismspf::dependency_slot(prismspf::SRC);
      for (std::size_t f = 0; f < fields.size(); ++f)
        deps.dirichlet_mask[f] = dirichlet_mask(int(f));
      for (const auto &in : declared_inputs(blk.dependencies_lhs))
        {
          if (in.slot == src_slot)
            {
              // Values come from the unit probe inside the kernel.
              deps.ptr[in.field][in.slot]   = nullptr;
              deps.flags[in.field][in.slot] = 0;
              continue;
            }
          deps.ptr[in.field][in.slot]   = lf.ptr[in.field][in.slot];
          deps.flags[in.field][in.slot] = in.flags;
        }
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        out.scratch_slot[f] = -1;
      const int f           = bf[k];
      out.scratch_slot[f]   = 0;
      out.dst[f]            = out_ptr;
      out.dirichlet_mask[f] = deps.dirichlet_mask[f];
      out.out_mask          = 1u << f;
      out.scaling           = scaling;
      out.dirichlet_write   = dirichlet_zero;
      number *scratch       = ensure_scratch<number>(int64_t(dofs_per_cell<dim, degree>) * g.n_cells);
      const auto &pde       = *pde_of<number>();
      dispatch_block_id(typename App::BlockIds {}, blk.id, [&](auto bid) {
        constexpr int id = decltype(bid)::value;
        hipLaunchKernelGGL((cell_operator_kernel<PDE<number>, dim, degree, number, true, id, true, App::n_fields>),
                           dim3(grid_for(g.n_cells, block_size)), dim3(block_size), 0, stream,
                           pde, g, deps, out, scratch, timer, f);
      });
      hipLaunchKernelGGL((node_gather_kernel<dim, degree, number>),
                         dim3(grid_for(g.n_nodes, block_size)), dim3(block_size), 0, stream, g,
                         out, scratch);
    }

    /// Device region that replaces every allocation of the operator path while it is
    /// set (ppf_apply_operator_noalloc): scratch, the zero operand and FP32 copies are
    /// carved from it in call order; a request that does not fit throws.
    struct BorrowedScratch
    {
      unsigned char *base  = nullptr;
      int64_t        bytes = 0;
      int64_t        used  = 0;
    };
    BorrowedScratch borrowed;

    template <typename T>
    T *
    carve(int64_t n, const char *what)
    {
      const int64_t off  = (borrowed.used + 255) / 256 * 256;
      const int64_t need = n * int64_t(sizeof(T));
      if (off + need > borrowed.bytes)
        throw Error(status_invalid, std::string(what) + " needs " + std::to_string(off + need) +
                                      " bytes of device scratch; the allocation-free entry "
                                      "point has " + std::to_string(borrowed.bytes));
      borrowed.used = off + need;
      return reinterpret_cast<T *>(borrowed.base + off);
    }

    /// Device array of at least n zeros of type number (kept for the context).
    template <typename number>
    const number *
    zero_array(int64_t n)
    {
      if (borrowed.base)
        {
          // The borrowed region is shared by every call on the GPU: clear it each time.
          number *z = carve<number>(n, "the zero operand");
          hipLaunchKernelGGL(fill_kernel<number>, dim3(grid_for(n, block_size)),
                             dim3(block_size), 0, stream, z, number(0), n);
          PPF_HIP_CHECK(hipGetLastError());
          return z;
        }
      auto &buf = [&]() -> DeviceBuffer<number> & {
        if constexpr (std::is_same_v<number, float>)
          return zeros_f;
        else
          return zeros_d;
      }();
      if (buf.size() < n)
        {
          buf.resize(n);
          hipLaunchKernelGGL(fill_kernel<number>, dim3(grid_for(n, block_size)),
                             dim3(block_size), 0, stream, buf.get(), number(0), n);
          PPF_HIP_CHECK(hipGetLastError());
        }
      return buf.get();
    }

    template <typename number>
    number *
    ensure_scratch(int64_t n)
    {
      if (borrowed.base)
        return carve<number>(n, "the generic cell path's per-cell results");
      if constexpr (std::is_same_v<number, float>)
        {
          if (scratch_f.size() < n)
            scratch_f.resize(n);
          return scratch_f.get();
        }
      else
        {
          if (scratch_d.size() < n)
            scratch_d.resize(n);
          return scratch_d.get();
        }
    }

    // ----------------------------------------------------------- explicit solve
    /// ExplicitSolver::solve_impl (reference explicit_solver.h:63-77): the block's fields
    /// become invm * (assembled rhs), with Dirichlet nodes set to their values.
    void
    solve_explicit(int b)
    {
      if (explicit_fp32)
        {
          solve_explicit_fp32(b);
          return;
        }
      std::vector<double *> dst;
      for (auto f : blocks[b].field_indices)
        dst.push_back(field_ptr(int(f), 0));
      apply_operator<double>(b, false, fine_fields(), {}, dst, scale_invm, dirichlet_value,
                             false, &dirichlet_values);
    }

    /// Explicit update with FP32 operator arithmetic and FP64 state (option
    /// explicit_precision = single): the rhs inputs are rounded to FP32 copies, the FP32
    /// instantiation of the equations computes the new values, and those are stored
    /// back into the FP64 field state.
    void
    solve_explicit_fp32(int b)
    {
      if (!explicit_fp32_copies && solve_explicit_mixed(b))
        return;
      const auto        &blk = blocks[b];
      const int64_t      n   = fine.n_nodes;
      LevelFields<float> lf {};
      lf.grid                = &fine;
      constexpr int src_slot = prismspf::dependency_slot(prismspf::SRC);
      explicit_in_f.resize(fields.size() * (PPF_MAX_AGE + 1));
      for (const auto &in : declared_inputs(blk.dependencies_rhs))
        {
          if (in.slot == src_slot || lf.ptr[in.field][in.slot])
            continue;
          const int age = in.slot - 1;
          if (age < 0 || age >= int(storage[in.field].size()))
            throw Error(status_internal, "dependency storage not available");
          auto &buf = explicit_in_f[in.field * (PPF_MAX_AGE + 1) + age];
          buf.resize(n);
          convert(buf.get(), storage[in.field][age].get(), n);
          lf.ptr[in.field][in.slot] = buf.get();
        }
      if (dirichlet_values_f.size() != dirichlet_values.size())
        {
          dirichlet_values_f.clear();
          dirichlet_values_f.resize(dirichlet_values.size());
          for (std::size_t f = 0; f < dirichlet_values.size(); ++f)
            if (dirichlet_values[f].get())
              {
                dirichlet_values_f[f].resize(n);
                convert(dirichlet_values_f[f].get(), dirichlet_values[f].get(), n);
              }
        }
      explicit_out_f.resize(fields.size());
      std::vector<float *> dst;
      for (auto f : blk.field_indices)
        {
          explicit_out_f[f].resize(n);
          dst.push_back(explicit_out_f[f].get());
        }
      apply_operator<float>(b, false, lf, {}, dst, scale_invm, dirichlet_value, false,
                            &dirichlet_values_f);
      for (auto f : blk.field_indices)
        convert(field_ptr(int(f), 0), explicit_out_f[f].get(), n, enqueue_halt);
    }

    /// Dependency and output tables of the explicit update of block b on the FP64 state:
    /// the rhs inputs from the field storage, the outputs into the current slots, M^-1
    /// scaling and Dirichlet values; false if an input has no storage.
    bool
    explicit_tables(int b, DepTable<double> &deps, OutTable<double> &out)
    {
      const prismspf::SolveBlock &blk = blocks[b];
      const LevelFields<double>   lf  = fine_fields();
      for (std::size_t f = 0; f < fields.size(); ++f)
        deps.dirichlet_mask[f] = dirichlet_mask(int(f));
      for (const auto &in : declared_inputs(blk.dependencies_rhs))
        {
          const double *ptr = lf.ptr[in.field][in.slot];
          if (!ptr)
            return false;
          deps.ptr[in.field][in.slot]   = ptr;
          deps.flags[in.field][in.slot] = in.flags;
        }
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        out.scratch_slot[f] = -1;
      int k = 0;
      for (auto fi : blk.field_indices)
        {
          const int f           = int(fi);
          out.dst[f]            = field_ptr(f, 0);
          out.dirichlet_mask[f] = deps.dirichlet_mask[f];
          out.scratch_slot[f]   = k++;
          out.out_mask         |= 1u << f;
          if (out.dirichlet_mask[f])
            {
              if (!dirichlet_values[f].get())
                throw Error(status_internal, "missing Dirichlet values");
              out.dirichlet_values[f] = dirichlet_values[f].get();
            }
        }
      out.scaling         = scale_invm;
      out.dirichlet_write = dirichlet_value;
      out.halt            = enqueue_halt;
      return true;
    }

    /// Explicit blocks ba and bb (consecutive in the solve order) in one launch
    /// (pair_kernel): bb's stencil operand is ba's output.  Taken for 3-D FP64 updates
    /// when both blocks update one field each through split plans that have been checked
    /// against the generic path (the first increment runs the blocks one by one, which
    /// does that check), ba has one stencil operand, no field involved has Dirichlet
    /// faces, and neither block reads what the other writes except bb reading ba's
    /// output; false, with nothing launched, otherwise.
    /// A pair that does not qualify; with PPF_VERBOSE set, says where (once per place).
    bool
    pair_skip(int ba, int line)
    {
      if (std::getenv("PPF_VERBOSE") && !explicit_pair_skipped.count(line))
        {
          explicit_pair_skipped.insert(line);
          std::fprintf(stderr, "explicit pair at block %d: not fused (runtime.hpp:%d)\n",
                       blocks[ba].id, line);
        }
      return false;
    }

    bool
    solve_explicit_pair(int ba, int bb)
    {
      if constexpr (dim != 3)
        return pair_skip(ba, __LINE__);
      else
        {
          if (!explicit_pairs || explicit_fp32 || timer.get_increment() == 0)
            return pair_skip(ba, __LINE__);
          if (path == OperatorPath::Generic || path == OperatorPath::SplitNode)
            return pair_skip(ba, __LINE__);
          const prismspf::SolveBlock &A = blocks[ba], &B = blocks[bb];
          if (A.solve_type != prismspf::Explicit || B.solve_type != prismspf::Explicit ||
              A.field_indices.size() != 1 || B.field_indices.size() != 1)
            return pair_skip(ba, __LINE__);
          const int fa = int(*A.field_indices.begin()), fb = int(*B.field_indices.begin());
          if (fa == fb || B.id != A.id + 1)
            return pair_skip(ba, __LINE__);
          const PlanEntry ea = split_plan<double>(ba, false);
          const PlanEntry eb = split_plan<double>(bb, false);
          if (!ea.plan.eligible || !eb.plan.eligible)
            return pair_skip(ba, __LINE__);
          if (path == OperatorPath::Auto && (!ea.verified || !eb.verified))
            return pair_skip(ba, __LINE__);
          DepTable<double> da {}, db {};
          OutTable<double> oa {}, ob {};
          if (!explicit_tables(ba, da, oa) || !explicit_tables(bb, db, ob))
            return pair_skip(ba, __LINE__);
          const double *dst_a = oa.dst[fa];
          double       *dst_b = ob.dst[fb];
          // One stencil operand per block; bb's is ba's output.
          auto operand = [](const SplitPlan &pl, int out_f, const DepTable<double> &d,
                            const double *&ptr, double &beta) {
            if (pl.terms.n_terms < 1)
              return false;
            const SplitTerm &t0 = pl.terms.term[0];
            beta                = 0.0;
            for (int t = 0; t < pl.terms.n_terms; ++t)
              {
                const SplitTerm &tm = pl.terms.term[t];
                if (tm.out_field != out_f || tm.in_field != t0.in_field ||
                    tm.in_slot != t0.in_slot)
                  return false;
                beta += tm.beta;
              }
            ptr = d.ptr[t0.in_field][t0.in_slot];
            return ptr != nullptr;
          };
          const double *op_a = nullptr, *op_b = nullptr;
          double        beta_a = 0.0, beta_b = 0.0;
          if (!operand(ea.plan, fa, da, op_a, beta_a) || !operand(eb.plan, fb, db, op_b, beta_b))
            return pair_skip(ba, __LINE__);
          if (op_b != dst_a || op_a == dst_a || op_a == dst_b)
            return pair_skip(ba, __LINE__);
          // No Dirichlet faces on the outputs or on anything read.
          if (oa.dirichlet_mask[fa] || ob.dirichlet_mask[fb])
            return pair_skip(ba, __LINE__);
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
            for (int s = 0; s < PPF_N_SLOTS; ++s)
              if ((da.ptr[f][s] || db.ptr[f][s]) && da.dirichlet_mask[f])
                return pair_skip(ba, __LINE__);

          PairArgs<double> a {};
          for (int d = 0; d < 3; ++d)
            {
              a.n[d]        = fine.nodes[d];
              a.periodic[d] = fine.periodic[d];
              // grown-tile positions reach 2p beyond the lattice: one wrap / mirror
              if (a.n[d] < 2 * degree + 1)
                return pair_skip(ba, __LINE__);
              const double tq2[2][3] = {{14.0, -8.0, 1.0}, {8.0, -4.0, 0.0}};
              const double tq1[3]    = {2.0, -1.0, 0.0};
              for (int par = 0; par < 2; ++par)
                for (int k = 0; k < 3; ++k)
                  a.tap[d][par][k] =
                    (degree == 2 ? tq2[par][k] : tq1[k]) / (fine.h[d] * fine.h[d]);
              for (int c = 0; c < 3; ++c)
                a.sfac[d][c] = 1.0; // M^-1 scaling
              a.q_lower[d] = fine.lower[d];
              a.q_step[d]  = fine.h[d] / degree;
            }
          a.element_volume = fine.h[0] * fine.h[1] * fine.h[2];
          if (int64_t(a.n[0]) * a.n[1] * a.n[2] * int64_t(sizeof(double)) > (int64_t(1) << 32))
            return pair_skip(ba, __LINE__);
          a.op_a   = op_a;
          a.beta_a = beta_a;
          a.fa     = fa;
          a.dst_a  = const_cast<double *>(dst_a);
          a.beta_b = beta_b;
          a.fb     = fb;
          a.dst_b  = dst_b;
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
            for (int s = 0; s < PPF_N_SLOTS; ++s)
              {
                // Entries without a value input read nothing (alias 0).
                a.alias_a[f][s] = 0;
                a.alias_b[f][s] = 0;
                a.val_a[f][s]   = op_a;
                a.val_b[f][s]   = op_a;
                if ((da.flags[f][s] & slot_values) && da.ptr[f][s] != op_a)
                  {
                    if (da.ptr[f][s] == dst_a || da.ptr[f][s] == dst_b)
                      return pair_skip(ba, __LINE__);
                    a.alias_a[f][s] = -1;
                    a.val_a[f][s]   = da.ptr[f][s];
                  }
                if ((db.flags[f][s] & slot_values) && db.ptr[f][s] != dst_a)
                  {
                    if (db.ptr[f][s] == dst_b)
                      return pair_skip(ba, __LINE__);
                    a.alias_b[f][s] = -1;
                    a.val_b[f][s]   = db.ptr[f][s];
                  }
              }
          a.halt = enqueue_halt;

          // Tiles: 64 columns (the last few columns node by node when they are at most
          // 8), TY = 4 BR rows (the last tile row partial), z chunks for ~2 workgroups
          // per compute unit.
          const int br = pair_rows;
          const int ty = pair_main_waves * br;
          const int rx = a.n[0] % stencil_tx;
          a.tile_n0    = (rx != 0 && rx <= 8 && a.n[0] >= 2 * stencil_tx) ? a.n[0] - rx : a.n[0];
          a.tiles_x    = (a.tile_n0 + stencil_tx - 1) / stencil_tx;
          a.tiles_y    = (a.n[1] + ty - 1) / ty;
          a.rem        = int64_t(a.n[0] - a.tile_n0) * a.n[1] * a.n[2];
          int zc       = pair_z_chunk;
          if (zc <= 0)
            {
              const int64_t xy     = int64_t(a.tiles_x) * a.tiles_y;
              const int64_t target = 2 * device_cus();
              zc = int((int64_t(a.n[2]) * xy + target - 1) / target);
              zc = std::max(zc, 4 * degree);
            }
          a.z_chunk  = std::min(zc, a.n[2]);
          a.chunks_z = (a.n[2] + a.z_chunk - 1) / a.z_chunk;
          const int64_t nblocks = int64_t(a.tiles_x) * a.tiles_y * a.chunks_z;
          const int64_t nrem    = (a.rem + 255) / 256;
          if (nblocks >= (int64_t(1) << 31) || nrem >= (int64_t(1) << 31))
            return pair_skip(ba, __LINE__);
          a.n_blocks = int(nblocks);
          {
            const int64_t footprint = 4 * fine.n_nodes * int64_t(sizeof(double));
            a.nt_store = stencil_store_policy == StorePolicy::NonTemporal ||
                         (stencil_store_policy == StorePolicy::Auto &&
                          footprint > stencil_nt_min_bytes);
          }
          if (probe_req.pending && probe_req.block == bb && a.n_blocks > 0)
            {
              a.probe_result = probe_req.result;
              a.probe_halt   = probe_req.halt;
              a.probe_hmask  = probe_req.hmask;
              a.probe_tag    = probe_req.tag;
              a.probe_n      = probe_req.fields.n;
              for (int f = 0; f < PPF_MAX_FIELDS; ++f)
                a.probe_ptr[f] = probe_req.fields.ptr[f];
              probe_req.pending = false;
              probe_req.taken   = true;
            }
          bool launched = false;
          dispatch_block_id(typename App::BlockIds {}, A.id, [&](auto aid) {
            constexpr int id_a = decltype(aid)::value;
            if constexpr (has_block_id<id_a + 1>(typename App::BlockIds {}))
              if (B.id == id_a + 1)
                {
                  const dim3 grid {unsigned(nblocks)}, threads {stencil_tx, pair_waves};
                  const auto &pde = *pde_of<double>();
                  if (nrem > 0)
                    hipLaunchKernelGGL((pair_node_kernel<PDE<double>, degree, double, id_a, id_a + 1, App::n_fields>),
                                       dim3(unsigned(nrem)), dim3(256), 0, stream, pde, a, timer);
                  if (br == 2)
                    hipLaunchKernelGGL((pair_kernel<PDE<double>, degree, double, id_a, id_a + 1, App::n_fields, 2>),
                                       grid, threads, 0, stream, pde, a, timer);
                  else
                    hipLaunchKernelGGL((pair_kernel<PDE<double>, degree, double, id_a, id_a + 1, App::n_fields, 3>),
                                       grid, threads, 0, stream, pde, a, timer);
                  launched = true;
                }
          });
          if (!launched)
            {
              if (a.probe_result)
                {
                  probe_req.pending = true;
                  probe_req.taken   = false;
                }
              return pair_skip(ba, __LINE__);
            }
          PPF_HIP_CHECK(hipGetLastError());
          if (!explicit_pair_noted.count(ba))
            {
              explicit_pair_noted.insert(ba);
              log += "solve blocks " + std::to_string(A.id) + " and " + std::to_string(B.id) +
                     " (rhs): one launch\n";
            }
          for (auto f : A.field_indices)
            rotated[f] = false;
          for (auto f : B.field_indices)
            rotated[f] = false;
          return true;
        }
    }

    /// The FP32 explicit update straight on the FP64 state: the tiled stencil kernel,
    /// instantiated with FP32 arithmetic, loads the FP64 inputs, rounds them in
    /// registers (as the FP32 copies would be rounded) and stores the FP64 outputs,
    /// and the constraint values are written in FP64.  Taken when the block's FP32 split
    /// plan qualifies and, on the automatic path, has been checked against the generic
    /// path on the fields (the first increment runs through the FP32 copies, which does
    /// that check); false, with nothing launched, otherwise.
    bool
    solve_explicit_mixed(int b)
    {
      if (path == OperatorPath::Generic || path == OperatorPath::SplitNode)
        return false;
      const PlanEntry &entry = split_plan<float>(b, false);
      if (!entry.plan.eligible || (path == OperatorPath::Auto && !entry.verified))
        return false;
      const prismspf::SolveBlock &blk = blocks[b];
      DepTable<double>            deps {};
      OutTable<double>            out {};
      if (!explicit_tables(b, deps, out))
        return false;
      bool launched = false;
      dispatch_block_id(typename App::BlockIds {}, blk.id, [&](auto bid) {
        constexpr int id = decltype(bid)::value;
        launched = launch_stencil<float, id, double>(b, false, fine, deps, out, entry.plan);
      });
      PPF_HIP_CHECK(hipGetLastError());
      if (launched && !explicit_mixed_noted.count(b))
        {
          explicit_mixed_noted.insert(b);
          log += "solve block " + std::to_string(blk.id) +
                 " (rhs): FP32 stencil arithmetic on the FP64 state\n";
        }
      return launched;
    }

    // ------------------------------------------------------------- linear solve
    struct Level
    {
      GridDesc                            grid;
      int                                 precision = 0;
      DeviceBuffer<double>                vd[6];
      DeviceBuffer<float>                 vf[6];
      DeviceBuffer<double>                inv_diag_d;
      DeviceBuffer<float>                 inv_diag_f;
      /// inverse diagonal by node class (diag_table_build_kernel) and its check flag
      DeviceBuffer<double>                table_d;
      DeviceBuffer<float>                 table_f;
      DeviceBuffer<int>                   table_ok;
      double                              lambda_max = 0.0, lambda_min = 0.0;
      bool                                eigenvalues_estimated = false;
      std::vector<DeviceBuffer<double>>   field_d[PPF_MAX_FIELDS];
      std::vector<DeviceBuffer<float>>    field_f[PPF_MAX_FIELDS];
      DeviceBuffer<double>                tmp_d[2]; // separable transfers to the coarser level
      DeviceBuffer<float>                 tmp_f[2];
    };

    struct LinearState
    {
      bool                 initialized = false;
      int64_t              n           = 0; // block vector length
      DeviceBuffer<double> x_homog, inhom, rhs, g, h, d, ad, tmp;
      DeviceBuffer<double> inv_diag; // Chebyshev preconditioner on the fine level
      DeviceBuffer<double> inv_table;
      DeviceBuffer<int>    inv_table_ok;
      double               cheb_lmax = 0.0, cheb_lmin = 0.0;
      bool                 cheb_eigenvalues_estimated = false;
      DeviceBuffer<double> cheb_d, cheb_ax;
      DeviceBuffer<double> bs_r, bs_rbar, bs_p, bs_y, bs_z, bs_t, bs_v; // BiCGStab
      std::vector<DeviceBuffer<double>> gm_v, gm_z;                    // GMRES bases
      DeviceBuffer<double>              gm_w, gm_t;
      std::vector<Level>   levels; // coarsest first
      SolveStats           stats;
      NewtonStats          newton;
      /// Captured V-cycles (run_vcycle): two slots, keyed by the halt flag the launches
      /// carry, the capture epoch, the allocation count and the top level's precision.
      hipGraphExec_t vg_exec[2]  = {nullptr, nullptr};
      const int     *vg_halt[2]  = {nullptr, nullptr};
      long           vg_epoch[2] = {-1, -1};
      long           vg_alloc[2] = {-1, -1};
      int            vg_prec[2]  = {-1, -1};
      int            vg_next     = 0;
      long           vg_warm     = -1;  ///< epoch of the last eager V-cycle found capturable
      bool           vg_off      = false; ///< the V-cycle synchronizes with the host
    };

    const prismspf::LinearSolverParameters &
    lin_params(int b) const
    {
      return blocks[b].linear_solver_parameters;
    }

    int
    linear_scaling(int b) const
    {
      return lin_params(b).tolerance_type != prismspf::AbsoluteResidual ? scale_invm_sqrt
                                                                        : scale_none;
    }

    /// normalization_value (reference linear_solver.h:414-427).
    double
    normalization(int b) const
    {
      const auto type = lin_params(b).tolerance_type;
      double     v    = 1.0;
      if (type == prismspf::RMSEPerField || type == prismspf::RMSETotal)
        {
          double vol = 1.0;
          for (int d = 0; d < dim; ++d)
            vol *= fine.h[d] * fine.cells[d];
          v *= std::sqrt(vol);
        }
      if (type == prismspf::RMSEPerField || type == prismspf::IntegratedPerField)
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

    std::vector<double *>
    split_block(double *base, int b, const GridDesc &g) const
    {
      std::vector<double *> v;
      for (std::size_t k = 0; k < blocks[b].field_indices.size(); ++k)
        v.push_back(base + k * g.n_nodes);
      return v;
    }

    template <typename T>
    std::vector<T *>
    split_level(T *base, int b, const GridDesc &g) const
    {
      std::vector<T *> v;
      for (std::size_t k = 0; k < blocks[b].field_indices.size(); ++k)
        v.push_back(base + k * g.n_nodes);
      return v;
    }

    LinearState &
    linear_state(int b)
    {
      LinearState &ls = linear_states[b];
      if (ls.initialized)
        return ls;
      const int64_t n = int64_t(blocks[b].field_indices.size()) * fine.n_nodes;
      ls.n            = n;
      for (auto *buf : {&ls.x_homog, &ls.inhom, &ls.rhs, &ls.g, &ls.h, &ls.d, &ls.ad, &ls.tmp})
        {
          buf->resize(n);
          fill(buf->get(), 0.0, n);
        }
      ls.initialized = true;
      return ls;
    }

    /// Full-block vmult with the fine FP64 lhs operator (homogeneous constraints).
    void
    vmult_fine(int b, double *dst, const double *src)
    {
      auto s = split_block(const_cast<double *>(src), b, fine);
      apply_operator<double>(b, true, fine_fields(), {s.begin(), s.end()},
                             split_block(dst, b, fine), linear_scaling(b), dirichlet_zero,
                             false);
    }

    /// LinearSolver::solve_impl (reference linear_solver.h:305-345).
    SolveStats
    solve_linear(int b)
    {
      LinearState &ls = linear_state(b);
      const auto  &bf = blocks[b].field_indices;
      const int    sc = linear_scaling(b);

      // rhs = scaling * rhs operator
      apply_operator<double>(b, false, fine_fields(), {}, split_block(ls.rhs.get(), b, fine), sc,
                             dirichlet_zero, false);
      // inhomogeneous values: g at Dirichlet nodes, 0 elsewhere
      {
        int k = 0;
        for (auto f : bf)
          {
            double *dstp = ls.inhom.get() + k * fine.n_nodes;
            fill(dstp, 0.0, fine.n_nodes);
            if (dirichlet_mask(int(f)))
              hipLaunchKernelGGL((dirichlet_compose_kernel<dim, double>),
                                 dim3(grid_for(fine.n_nodes, 256)), dim3(256), 0, stream, dstp,
                                 fine, dirichlet_values[f].get(), dirichlet_mask(int(f)), 0.0, 1);
            ++k;
          }
      }
      // rhs -= A_plain(inhom)
      {
        auto s = split_block(ls.inhom.get(), b, fine);
        apply_operator<double>(b, true, fine_fields(), {s.begin(), s.end()},
                               split_block(ls.tmp.get(), b, fine), sc, dirichlet_zero, true);
        axpby(ls.rhs.get(), -1.0, ls.tmp.get(), 1.0, ls.n);
      }

      prepare_preconditioner(b);
      if (zero_initial_guess)
        fill(ls.x_homog.get(), 0.0, ls.n);
      SolveStats st = krylov(b, ls.x_homog.get(), ls.rhs.get(),
                             lin_params(b).tolerance * normalization(b),
                             lin_params(b).max_iterations);

      // solution = x + inhomogeneous values
      int k = 0;
      for (auto f : bf)
        {
          double *sol = field_ptr(int(f), 0);
          copy(sol, ls.x_homog.get() + k * fine.n_nodes, fine.n_nodes);
          axpby(sol, 1.0, ls.inhom.get() + k * fine.n_nodes, 1.0, fine.n_nodes);
          ++k;
        }
      sync();
      ls.stats = st;
      return st;
    }

    /// NewtonSolver::solve_impl (reference newton_solver.h:73-156).  The residual is
    /// the block's rhs operator at the current solution and the Jacobian its lhs
    /// operator, applied to the update (the "change" field) with homogeneous
    /// constraints; both carry the linear solver's scaling.  The solution starts from
    /// old_1 when the field keeps old states.  The update vector persists between
    /// Newton iterations and increments: each linear solve starts from the previous
    /// update, as the reference's newton_update member does.
    NewtonStats
    solve_newton(int b)
    {
      LinearState &ls = linear_state(b);
      const auto  &bf = blocks[b].field_indices;
      const auto  &np = blocks[b].nonlinear_solver_parameters;
      const int    sc = linear_scaling(b);
      NewtonStats  ns;
      ns.tolerance = np.tolerance_value * normalization(b);

      for (auto f : bf)
        if (storage[f].size() > 1)
          copy(field_ptr(int(f), 0), field_ptr(int(f), 1), fine.n_nodes);

      bool     unconverged = true;
      unsigned iter        = 0;
      while (unconverged && iter < np.max_iterations)
        {
          // Constraints on the solution: Dirichlet values.
          for (auto f : bf)
            if (dirichlet_mask(int(f)))
              hipLaunchKernelGGL((dirichlet_compose_kernel<dim, double>),
                                 dim3(grid_for(fine.n_nodes, 256)), dim3(256), 0, stream,
                                 field_ptr(int(f), 0), fine, dirichlet_values[f].get(),
                                 dirichlet_mask(int(f)), 0.0, 1);
          // Residual r = (scaled) rhs operator at the current solution.
          apply_operator<double>(b, false, fine_fields(), {}, split_block(ls.rhs.get(), b, fine),
                                 sc, dirichlet_zero, false);
          ns.residual = std::sqrt(dot(ls.rhs.get(), ls.rhs.get(), ls.n));
          if (!std::isfinite(ns.residual))
            throw Error(status_nan, "Newton solver produced a non-finite residual");
          unconverged = ns.residual > ns.tolerance;
          if (!unconverged)
            continue;
          // J du = r, from the previous update (level operators see the current
          // solution through the injected fields and re-evaluated diagonals).
          prepare_preconditioner(b);
          if (zero_initial_guess)
            fill(ls.x_homog.get(), 0.0, ls.n);
          ls.stats = krylov(b, ls.x_homog.get(), ls.rhs.get(),
                            lin_params(b).tolerance * normalization(b),
                            lin_params(b).max_iterations);
          ns.linear_iterations += ls.stats.iterations;
          int k = 0;
          for (auto f : bf)
            {
              axpby(field_ptr(int(f), 0), np.step_length, ls.x_homog.get() + k * fine.n_nodes,
                    1.0, fine.n_nodes);
              ++k;
            }
          ++iter;
        }
      sync();
      ns.iterations = int(iter);
      ns.converged  = unconverged ? 0 : 1;
      ls.newton     = ns;
      return ns;
    }

    /// The block's linear solver (SolverSelector by `solver type`, reference
    /// linear_solver.h:461-490): every name the parser accepts (cg, bicgstab, gmres,
    /// fgmres, richardson, minres).
    SolveStats
    krylov(int b, double *x, const double *rhs, double tol, unsigned int max_steps)
    {
      const std::string &type = lin_params(b).solver_type;
      if (type == "cg")
        return cg(b, x, rhs, tol, max_steps);
      if (type == "bicgstab")
        return bicgstab(b, x, rhs, tol, max_steps);
      if (type == "gmres")
        return gmres(b, x, rhs, tol, max_steps, false);
      if (type == "fgmres")
        return gmres(b, x, rhs, tol, max_steps, true);
      if (type == "richardson")
        return richardson(b, x, rhs, tol, max_steps);
      if (type == "minres")
        return minres(b, x, rhs, tol, max_steps);
      throw Error(status_unsupported, "unknown linear solver type '" + type + "'");
    }

    /// MINRES for a symmetric operator with a symmetric positive definite
    /// preconditioner (deal.II's SolverMinRes, lac/solver_minres.h; `solver type =
    /// minres`), restated as the published Paige-Saunders recurrence (Lanczos on
    /// P^{-1} A, Givens rotations, three-term direction update).  The checked residual is
    /// the recurrence's estimate phibar, |r|_P (|b - A x| without a preconditioner),
    /// at step 0 and after every step.
    SolveStats
    minres(int b, double *x, const double *rhs, double tol, unsigned int max_steps)
    {
      LinearState &ls = linear_state(b);
      const int64_t n = ls.n;
      for (auto *buf : {&ls.bs_r, &ls.bs_rbar, &ls.bs_p, &ls.bs_y, &ls.bs_z, &ls.bs_t, &ls.bs_v})
        buf->resize(n);
      double *r1 = ls.bs_r.get(), *r2 = ls.bs_rbar.get(), *y = ls.bs_y.get(), *v = ls.bs_v.get(),
             *w = ls.bs_p.get(), *w1 = ls.bs_z.get(), *w2 = ls.bs_t.get();
      SolveStats st;
      st.tolerance = tol;
      vmult_fine(b, r1, x);
      axpby(r1, 1.0, rhs, -1.0, n); // r1 = b - A x
      precondition(b, y, r1);
      const double beta1 = std::sqrt(std::max(0.0, dot(r1, y, n)));
      double       phibar = beta1;
      st.initial_residual = phibar;
      unsigned step  = 0;
      int      state = control_check(phibar, tol, step, max_steps);
      if (state == 0)
        {
          copy(r2, r1, n);
          fill(w, 0.0, n);
          fill(w2, 0.0, n);
          double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, cs = -1.0, sn = 0.0;
          while (state == 0)
            {
              ++step;
              axpby(v, 1.0 / beta, y, 0.0, n); // v = y / beta
              vmult_fine(b, y, v);
              if (step >= 2)
                axpby(y, -beta / oldb, r1, 1.0, n);
              const double alfa = dot(v, y, n);
              axpby(y, -alfa / beta, r2, 1.0, n);
              std::swap(r1, r2); // r1 = old r2
              copy(r2, y, n);
              precondition(b, y, r2);
              oldb                = beta;
              beta                = std::sqrt(std::max(0.0, dot(r2, y, n)));
              const double oldeps = epsln;
              const double delta  = cs * dbar + sn * alfa;
              const double gbar   = sn * dbar - cs * alfa;
              epsln               = sn * beta;
              dbar                = -cs * beta;
              const double gamma =
                std::max(std::sqrt(gbar * gbar + beta * beta), std::numeric_limits<double>::epsilon());
              cs               = gbar / gamma;
              sn               = beta / gamma;
              const double phi = cs * phibar;
              phibar           = sn * phibar;
              // w = (v - oldeps w1 - delta w2) / gamma, with (w1, w2) <- (w2, w).
              std::swap(w1, w2);
              std::swap(w2, w);
              copy(w, v, n);
              axpby(w, -oldeps, w1, 1.0, n);
              axpby(w, -delta, w2, 1.0, n);
              scale(w, 1.0 / gamma, n);
              axpby(x, phi, w, 1.0, n);
              state = control_check(phibar, tol, step, max_steps);
              if (beta == 0.0 && state == 0)
                state = phibar <= tol ? 1 : -1;
            }
        }
      st.iterations     = int(step);
      st.final_residual = phibar;
      st.converged      = state > 0 ? 1 : 0;
      return st;
    }

    /// SolverControl::check: 1 success (value <= tol), -1 failure (step >= max_steps or
    /// not finite), 0 iterate.
    static int
    control_check(double value, double tol, unsigned int step, unsigned int max_steps)
    {
      if (!std::isfinite(value))
        throw Error(status_nan, "linear solver produced a non-finite residual");
      if (value <= tol)
        return 1;
      return step >= max_steps ? -1 : 0;
    }

    /// deal.II SolverRichardson (lac/solver_richardson.h; `solver type = richardson`),
    /// restated: each step forms r = b - A x and d = P^{-1} r, checks |r| (or |d| with
    /// `use preconditioned residual`) at the step count so far, then x += omega d.
    SolveStats
    richardson(int b, double *x, const double *rhs, double tol, unsigned int max_steps)
    {
      LinearState &ls = linear_state(b);
      const int64_t n = ls.n;
      const auto   &lp = lin_params(b);
      SolveStats    st;
      st.tolerance = tol;
      double   res  = 0.0;
      int      state = 0;
      unsigned step = 0;
      for (;; ++step)
        {
          vmult_fine(b, ls.g.get(), x);
          axpby(ls.g.get(), 1.0, rhs, -1.0, n); // r = b - A x
          precondition(b, ls.h.get(), ls.g.get());
          const double *crit = lp.richardson_use_preconditioned_residual ? ls.h.get() : ls.g.get();
          res = std::sqrt(dot(crit, crit, n));
          if (step == 0)
            st.initial_residual = res;
          state = control_check(res, tol, step, max_steps);
          if (state != 0)
            break;
          axpby(x, lp.richardson_omega, ls.h.get(), 1.0, n);
        }
      st.iterations     = int(step);
      st.final_residual = res;
      st.converged      = state > 0 ? 1 : 0;
      return st;
    }

    /// deal.II SolverGMRES (lac/solver_gmres.h; `solver type = gmres`), restated:
    /// restarted GMRES with a basis of `max basis size` vectors, left preconditioning
    /// (the residual checked is |P^{-1}(b - A x)|) or right preconditioning.  Each cycle
    /// checks the recomputed residual at the accumulated step count, then each Arnoldi
    /// step checks the least-squares estimate |g_{j+1}| (`use default residual`).  The
    /// basis is orthogonalized by modified Gram-Schmidt with the inner products and
    /// updates on the device; the Hessenberg column (j + 2 values) is read once per step
    /// and the Givens rotations run on the host.
    ///
    /// `flexible`: deal.II's SolverFGMRES, right preconditioning that keeps the
    /// preconditioned basis z_j = P^{-1} v_j and updates x += Z y (the same iterates as
    /// right-preconditioned GMRES for a fixed preconditioner).
    SolveStats
    gmres(int b, double *x, const double *rhs, double tol, unsigned int max_steps,
          bool flexible)
    {
      LinearState &ls = linear_state(b);
      const int64_t n = ls.n;
      const auto   &lp = lin_params(b);
      if (!flexible && !lp.gmres_use_default_residual)
        throw Error(status_unsupported, "GMRES with use default residual = false");
      const int  m     = int(lp.gmres_max_basis_size);
      const bool right = flexible || lp.gmres_right_preconditioning;
      if (int(ls.gm_v.size()) != m + 1)
        ls.gm_v.resize(m + 1);
      for (auto &v : ls.gm_v)
        v.resize(n);
      if (flexible)
        {
          if (int(ls.gm_z.size()) != m)
            ls.gm_z.resize(m);
          for (auto &z : ls.gm_z)
            z.resize(n);
        }
      ls.gm_w.resize(n);
      ls.gm_t.resize(n);
      krylov_slots.resize(std::max<int64_t>(8, m + 2));
      double                          *ks = krylov_slots.get();
      std::vector<double>              col(m + 2), cs(m), sn(m), g(m + 1), y(m);
      std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
      SolveStats                       st;
      st.tolerance   = tol;
      unsigned its   = 0;
      double   rho   = 0.0;
      int      state = 0;
      bool     first = true;
      double  *w = ls.gm_w.get(), *t = ls.gm_t.get();
      while (true)
        {
          double *v0 = ls.gm_v[0].get();
          if (!right)
            {
              vmult_fine(b, t, x);
              axpby(t, 1.0, rhs, -1.0, n); // b - A x
              precondition(b, v0, t);
            }
          else
            {
              vmult_fine(b, v0, x);
              axpby(v0, 1.0, rhs, -1.0, n);
            }
          rho = std::sqrt(dot(v0, v0, n));
          if (first)
            st.initial_residual = rho;
          first = false;
          state = control_check(rho, tol, its, max_steps);
          if (state != 0)
            break;
          scale(v0, 1.0 / rho, n);
          std::fill(g.begin(), g.end(), 0.0);
          g[0]  = rho;
          int j = 0;
          while (j < m)
            {
              ++its;
              double *vj = ls.gm_v[j].get();
              if (!right)
                {
                  vmult_fine(b, t, vj);
                  precondition(b, w, t);
                }
              else
                {
                  double *z = flexible ? ls.gm_z[j].get() : t;
                  precondition(b, z, vj);
                  vmult_fine(b, w, z);
                }
              for (int i = 0; i <= j; ++i)
                {
                  dot_to(w, ls.gm_v[i].get(), n, ks + i);
                  hipLaunchKernelGGL(axpy_slot_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                                     stream, w, ls.gm_v[i].get(), ks, i, -1.0, n);
                }
              dot_to(w, w, n, ks + j + 1);
              PPF_HIP_CHECK(hipMemcpyAsync(col.data(), ks, sizeof(double) * (j + 2),
                                           hipMemcpyDeviceToHost, stream));
              sync();
              for (int i = 0; i <= j; ++i)
                H[i][j] = col[i];
              const double hn = std::sqrt(col[j + 1]);
              H[j + 1][j]     = hn;
              if (hn > 0.0 && j + 1 <= m)
                axpby(ls.gm_v[j + 1].get(), 1.0 / hn, w, 0.0, n);
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
              state = control_check(rho, tol, its, max_steps);
              if (state != 0 || hn == 0.0)
                break;
            }
          // x += V y (left) or P V y (right), R y = g over the j steps of this cycle.
          for (int i = j - 1; i >= 0; --i)
            {
              double s = g[i];
              for (int k = i + 1; k < j; ++k)
                s -= H[i][k] * y[k];
              y[i] = s / H[i][i];
            }
          if (!right)
            for (int i = 0; i < j; ++i)
              axpby(x, y[i], ls.gm_v[i].get(), 1.0, n);
          else if (flexible)
            for (int i = 0; i < j; ++i)
              axpby(x, y[i], ls.gm_z[i].get(), 1.0, n);
          else
            {
              fill(w, 0.0, n);
              for (int i = 0; i < j; ++i)
                axpby(w, y[i], ls.gm_v[i].get(), 1.0, n);
              precondition(b, t, w);
              axpby(x, 1.0, t, 1.0, n);
            }
          if (state != 0)
            break;
        }
      st.iterations     = int(its);
      st.final_residual = rho;
      st.converged      = state > 0 ? 1 : 0;
      return st;
    }

    /// deal.II SolverBicgstab (deal.II >= 9.6, lac/solver_bicgstab.h; the reference
    /// selects it by `solver type = bicgstab`), restated: residual r = b - A x checked at
    /// step 0; each step checks |r| after the half update (early success adds alpha y)
    /// and again after the full one, where `exact residual` recomputes |A x - b|; a
    /// step whose |r . rbar| or |rbar . v| falls below `breakdown` restarts the
    /// iteration from the current x with the step counter kept.
    SolveStats
    bicgstab(int b, double *x, const double *rhs, double tol, unsigned int max_steps)
    {
      LinearState &ls = linear_state(b);
      const int64_t n = ls.n;
      for (auto *buf : {&ls.bs_r, &ls.bs_rbar, &ls.bs_p, &ls.bs_y, &ls.bs_z, &ls.bs_t, &ls.bs_v})
        buf->resize(n);
      double *r = ls.bs_r.get(), *rbar = ls.bs_rbar.get(), *p = ls.bs_p.get(),
             *y = ls.bs_y.get(), *z = ls.bs_z.get(), *t = ls.bs_t.get(), *v = ls.bs_v.get();
      const auto  &lp        = lin_params(b);
      const double breakdown = lp.bicgstab_breakdown;
      SolveStats   st;
      st.tolerance  = tol;
      unsigned step = 0;
      double   res  = 0.0;
      bool     first = true, success = false;
      // control.check: success at res <= tol, failure from step >= max_steps on.
      auto check = [&](double value) {
        if (!std::isfinite(value))
          throw Error(status_nan, "linear solver produced a non-finite residual");
        if (value <= tol)
          return 1;
        return step >= max_steps ? -1 : 0;
      };
      while (true) // one pass per (re)start
        {
          vmult_fine(b, r, x);
          axpby(r, 1.0, rhs, -1.0, n); // r = b - A x
          res = std::sqrt(dot(r, r, n));
          if (first)
            st.initial_residual = res;
          first     = false;
          int state = check(res);
          if (state != 0)
            {
              success = state > 0;
              break;
            }
          // Scalars in device slots (see bicgstab_scalar_kernel); the host reads the
          // residual norms only, and |r.rbar|, |rbar.v| when a breakdown threshold is set.
          krylov_slots.resize(8);
          double *ks = krylov_slots.get();
          hipLaunchKernelGGL(fill_kernel<double>, dim3(1), dim3(64), 0, stream, ks, 1.0,
                             int64_t(3));
          const unsigned nb      = grid_for(n, 256);
          const bool     read_bd = breakdown > 0.0;
          auto           slot    = [&](int k) {
            double val = 0.0;
            PPF_HIP_CHECK(hipMemcpyAsync(&val, ks + k, sizeof(double), hipMemcpyDeviceToHost,
                                         stream));
            sync();
            return val;
          };
          copy(rbar, r, n);
          bool startup = true, restart = false;
          do
            {
              ++step;
              dot_to(r, rbar, n, ks + 3);
              if (read_bd && std::fabs(slot(3)) < breakdown)
                {
                  restart = true;
                  break;
                }
              hipLaunchKernelGGL(bicgstab_scalar_kernel, dim3(1), dim3(1), 0, stream, ks, 0);
              if (startup)
                {
                  copy(p, r, n);
                  startup = false;
                }
              else
                hipLaunchKernelGGL(bicgstab_p_kernel, dim3(nb), dim3(256), 0, stream, p, r, v,
                                   ks, n);
              precondition(b, y, p);
              vmult_fine(b, v, y);
              dot_to(rbar, v, n, ks + 4);
              if (read_bd && std::fabs(slot(4)) < breakdown)
                {
                  restart = true;
                  break;
                }
              hipLaunchKernelGGL(bicgstab_scalar_kernel, dim3(1), dim3(1), 0, stream, ks, 1);
              hipLaunchKernelGGL(axpy_slot_norm_kernel, dim3(reduce_blocks), dim3(reduce_threads),
                                 0, stream, reduce_buf.get(), r, v, ks, 1, -1.0, n);
              res = std::sqrt(finish_reduction(reduce_blocks));
              if (check(res) > 0)
                {
                  hipLaunchKernelGGL(axpy_slot_kernel, dim3(nb), dim3(256), 0, stream, x, y, ks,
                                     1, 1.0, n);
                  state = 1;
                  break;
                }
              precondition(b, z, r);
              vmult_fine(b, t, z);
              dot_to(t, r, n, ks + 5);
              dot_to(t, t, n, ks + 6);
              hipLaunchKernelGGL(bicgstab_scalar_kernel, dim3(1), dim3(1), 0, stream, ks, 2);
              hipLaunchKernelGGL(bicgstab_x_kernel, dim3(nb), dim3(256), 0, stream, x, y, z, ks,
                                 n);
              hipLaunchKernelGGL(axpy_slot_kernel, dim3(nb), dim3(256), 0, stream, r, t, ks, 2,
                                 -1.0, n);
              if (lp.bicgstab_exact_residual)
                {
                  vmult_fine(b, t, x);
                  axpby(t, -1.0, rhs, 1.0, n); // t = A x - b
                  res = std::sqrt(dot(t, t, n));
                }
              else
                res = std::sqrt(dot(r, r, n));
              state = check(res);
            }
          while (state == 0);
          PPF_HIP_CHECK(hipGetLastError());
          if (!restart)
            {
              success = state > 0;
              break;
            }
        }
      st.iterations     = int(step);
      st.final_residual = res;
      st.converged      = success ? 1 : 0;
      return st;
    }

    /// Preconditioned CG with deal.II SolverCG's recurrences and stopping rule
    /// (absolute residual tolerance; step 0 checks the initial residual).
    SolveStats
    cg(int b, double *x, const double *rhs, double tol, unsigned int max_steps)
    {
      LinearState &ls = linear_state(b);
      const int64_t n = ls.n;
      SolveStats    st;
      st.tolerance = tol;
      // g = A x - b
      vmult_fine(b, ls.g.get(), x);
      axpby(ls.g.get(), -1.0, rhs, 1.0, n);
      double res          = std::sqrt(dot(ls.g.get(), ls.g.get(), n));
      st.initial_residual = res;
      unsigned int step   = 0;
      if (!(res <= tol))
        {
          // The step scalars stay on the device: g.h alternates between slots 0 and 1,
          // d.Ad is slot 2; alpha and beta are formed by the kernels that use them.  The
          // host reads only the residual norm, once per step.
          krylov_slots.resize(8);
          double *ks = krylov_slots.get();
          int     cur = 0;
          precondition(b, ls.h.get(), ls.g.get());
          axpby(ls.d.get(), -1.0, ls.h.get(), 0.0, n);
          dot_to(ls.g.get(), ls.h.get(), n, ks + cur);
          while (true)
            {
              ++step;
              vmult_fine(b, ls.ad.get(), ls.d.get());
              dot_to(ls.d.get(), ls.ad.get(), n, ks + 2);
              hipLaunchKernelGGL(cg_update_slots_kernel, dim3(reduce_blocks),
                                 dim3(reduce_threads), 0, stream, reduce_buf.get(), x,
                                 ls.g.get(), ls.d.get(), ls.ad.get(), ks, cur, 2, n);
              res = std::sqrt(finish_reduction(reduce_blocks));
              if (!std::isfinite(res))
                throw Error(status_nan, "linear solver produced a non-finite residual");
              if (res <= tol || step >= max_steps)
                break;
              precondition(b, ls.h.get(), ls.g.get());
              dot_to(ls.g.get(), ls.h.get(), n, ks + (1 - cur));
              hipLaunchKernelGGL(cg_direction_slots_kernel, dim3(grid_for(n, 256)), dim3(256),
                                 0, stream, ls.d.get(), ls.h.get(), ks, 1 - cur, cur, n);
              cur = 1 - cur;
            }
        }
      st.iterations     = int(step);
      st.final_residual = res;
      st.converged      = res <= tol ? 1 : 0;
      return st;
    }

    // ------------------------------------------------------------ preconditioners
    void
    prepare_preconditioner(int b)
    {
      LinearState &ls   = linear_state(b);
      const auto   kind = lin_params(b).preconditioner;
      if (kind == prismspf::Chebyshev)
        {
          // The reference re-evaluates the diagonal every solve (linear_solver.h:362-378)
          // into the matrix its PreconditionChebyshev was initialized with once
          // (initialize_chebyshev, :446-458).  deal.II estimates the eigenvalues at that
          // preconditioner's first application and keeps them: so does this path.
          ls.inv_diag.resize(ls.n);
          ls.cheb_d.resize(ls.n);
          ls.cheb_ax.resize(ls.n);
          inverse_diagonal<double>(b, fine_fields(), linear_scaling(b), ls.inv_diag.get());
          build_diag_table<double>(ls.inv_diag.get(), fine, int(blocks[b].field_indices.size()),
                                   ls.inv_table, ls.inv_table_ok);
          if (!ls.cheb_eigenvalues_estimated)
            estimate_eigenvalues<double>(
              [&](double *dst, const double *src) { vmult_fine(b, dst, src); },
              ls.inv_diag.get(), ls.n, lin_params(b).chebyshev_parameters, ls.cheb_lmax,
              ls.cheb_lmin, ls.tmp.get(), ls.cheb_d.get(), ls.cheb_ax.get(), ls.h.get());
          ls.cheb_eigenvalues_estimated = true;
        }
      else if (kind == prismspf::GMG)
        setup_multigrid(b);
    }

    void
    precondition(int b, double *h, const double *g)
    {
      LinearState &ls   = linear_state(b);
      const auto   kind = lin_params(b).preconditioner;
      if (kind == prismspf::None)
        copy(h, g, ls.n);
      else if (kind == prismspf::Chebyshev)
        {
          DiagTable<double> tab;
          if (ls.inv_table.get() && ls.inv_table_ok.get())
            {
              tab.table = ls.inv_table.get();
              tab.ok    = ls.inv_table_ok.get();
              tab.grid  = class_grid(fine, int(blocks[b].field_indices.size()));
            }
          chebyshev<double>(
            [&](double *dst, const double *src) { vmult_fine(b, dst, src); },
            ls.inv_diag.get(), ls.n, lin_params(b).chebyshev_parameters.degree, ls.cheb_lmax,
            ls.cheb_lmin, h, g, ls.cheb_d.get(), ls.cheb_ax.get(), true, tab);
        }
      else
        multigrid_vmult(b, h, g);
    }

    /// Lanczos estimate of the largest eigenvalue of P^{-1} A from `its` CG steps (the
    /// approach of deal.II's PreconditionChebyshev); lmax gets a 1.2 safety factor and
    /// lmin = lmax / smoothing_range.  The start vector is a fixed pseudo-random vector
    /// (parity with deal.II's own start vector is not claimed).
    template <typename T, class Vmult>
    void
    estimate_eigenvalues(Vmult &&A, const T *inv_diag, int64_t n,
                         const prismspf::ChebyshevParameters &cp, double &lmax, double &lmin,
                         T *r, T *p, T *ap, T *z)
    {
      // Start vector: splitmix64 of the block-vector index (field-major, stored node
      // order), uniform in [-0.5, 0.5), generated on the device.
      hipLaunchKernelGGL(splitmix_fill_kernel<T>, dim3(std::min<int64_t>(grid_for(n, 256), 4096)),
                         dim3(256), 0, stream, z, n);
      PPF_HIP_CHECK(hipGetLastError());
      // CG for A e = 0 from the pseudo-random guess: the initial residual is A x0, which
      // also has zero constrained rows.  z = P^{-1} r.
      A(r, z);
      hipLaunchKernelGGL(diag_mult_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0, stream, z,
                         inv_diag, r, n);
      copy(p, z, n);
      // The CG scalars stay on the device (rz_i in slot i, pAp_i in slot its + 1 + i);
      // the host reads them once, after the last step.
      const int its = int(cp.eig_cg_n_iterations);
      eig_slots.resize(2 * its + 2);
      double *rz  = eig_slots.get();
      double *pap = rz + its + 1;
      fill(rz, 0.0, 2 * its + 2);
      dot_to(r, z, n, rz);
      const unsigned int nb = grid_for(n, 256);
      for (int it = 0; it < its; ++it)
        {
          A(ap, p);
          dot_to(p, ap, n, pap + it);
          hipLaunchKernelGGL(eig_r_update_kernel<T>, dim3(nb), dim3(256), 0, stream, r, ap, rz,
                             pap, it, n);
          hipLaunchKernelGGL(diag_mult_kernel<T>, dim3(nb), dim3(256), 0, stream, z, inv_diag,
                             r, n);
          dot_to(r, z, n, rz + it + 1);
          hipLaunchKernelGGL(eig_p_update_kernel<T>, dim3(nb), dim3(256), 0, stream, p, z, rz,
                             pap, it, n);
        }
      PPF_HIP_CHECK(hipGetLastError());
      std::vector<double> host(2 * its + 2);
      PPF_HIP_CHECK(hipMemcpyAsync(host.data(), rz, sizeof(double) * host.size(),
                                   hipMemcpyDeviceToHost, stream));
      sync();
      // The host loop's coefficients: stop at the first non-positive rz or pAp.
      std::vector<double> alphas, betas;
      for (int it = 0; it < its; ++it)
        {
          const double rzi = host[it], papi = host[its + 1 + it];
          if (!(rzi > 0.0) || !(papi > 0.0))
            break;
          alphas.push_back(rzi / papi);
          betas.push_back(host[it + 1] / rzi);
        }
      // Lanczos tridiagonal from the CG coefficients.
      const int           m = int(alphas.size());
      std::vector<double> diag(m), off(std::max(0, m - 1));
      for (int i = 0; i < m; ++i)
        {
          diag[i] = 1.0 / alphas[i] + (i > 0 ? betas[i - 1] / alphas[i - 1] : 0.0);
          if (i + 1 < m)
            off[i] = std::sqrt(std::max(0.0, betas[i])) / alphas[i];
        }
      const double est = m > 0 ? tridiagonal_max_eigenvalue(diag, off) : 1.0;
      lmax             = 1.2 * est;
      lmin             = lmax / cp.smoothing_range;
    }

    /// Chebyshev iteration of the given degree for A x = rhs preconditioned by the
    /// inverse diagonal.  With `zero_start` x is overwritten (preconditioner use,
    /// deal.II vmult); otherwise x is the starting guess (smoother use, deal.II step).
    /// Inverse diagonal of a smoother by node class (see diag_table_build_kernel).
    template <typename T>
    struct DiagTable
    {
      const T   *table = nullptr;
      const int *ok    = nullptr;
      ClassGrid  grid {};
    };

    ClassGrid
    class_grid(const GridDesc &g, int n_fields) const
    {
      ClassGrid c {};
      for (int d = 0; d < 3; ++d)
        {
          c.n[d]        = d < dim ? g.nodes[d] : 1;
          c.periodic[d] = d < dim ? g.periodic[d] : 1;
        }
      c.degree   = degree;
      c.dim      = dim;
      c.n_fields = n_fields;
      c.n_nodes  = g.n_nodes;
      return c;
    }

    /// Builds the class table of `inv_diag` into `table` (n_fields * 64 entries) and
    /// checks every node against it; *ok ends 1 only if all match.
    template <typename T>
    DiagTable<T>
    build_diag_table(const T *inv_diag, const GridDesc &g, int n_fields, DeviceBuffer<T> &table,
                     DeviceBuffer<int> &ok)
    {
      DiagTable<T> t;
      t.grid = class_grid(g, n_fields);
      if (int64_t(t.grid.n[0]) * t.grid.n[1] >= (int64_t(1) << 31) || t.grid.n[2] > 65535 ||
          n_fields > 64)
        return DiagTable<T> {};
      if (table.size() < int64_t(n_fields) * diag_classes)
        table.resize(int64_t(n_fields) * diag_classes);
      if (ok.size() < 1)
        ok.resize(1);
      hipLaunchKernelGGL(diag_table_build_kernel<T>, dim3(1),
                         dim3(unsigned(n_fields * diag_classes)), 0, stream, table.get(),
                         ok.get(), inv_diag, t.grid);
      hipLaunchKernelGGL(diag_table_check_kernel<T>, class_launch(t.grid), dim3(256), 0, stream,
                         ok.get(), table.get(), inv_diag, t.grid);
      PPF_HIP_CHECK(hipGetLastError());
      t.table = table.get();
      t.ok    = ok.get();
      return t;
    }

    static dim3
    class_launch(const ClassGrid &g)
    {
      return dim3(grid_for(int64_t(g.n[0]) * g.n[1], 256), unsigned(g.n[2]),
                  unsigned(g.n_fields));
    }

    template <typename T, class Vmult>
    void
    chebyshev(Vmult &&A, const T *inv_diag, int64_t n, unsigned int deg, double lmax,
              double lmin, T *x, const T *rhs, T *d, T *ax, bool zero_start,
              const DiagTable<T> &tab = DiagTable<T> {})
    {
      const double theta = 0.5 * (lmax + lmin);
      const double delta = 0.5 * (lmax - lmin);
      const double sigma = theta / delta;
      double       rho   = 1.0 / sigma;
      for (unsigned int k = 0; k < deg; ++k)
        {
          if (k == 0 && zero_start)
            {
              // From x = 0 the residual is b: no operator application, no clearing.
              if (tab.table)
                hipLaunchKernelGGL(chebyshev_table_kernel<T>, class_launch(tab.grid),
                                   dim3(256), 0, stream, x, d, rhs, (const T *)nullptr,
                                   inv_diag, tab.table, tab.ok, T(0), T(1.0 / theta), 1, 1,
                                   tab.grid);
              else
                hipLaunchKernelGGL(chebyshev_first_zero_kernel<T>, dim3(grid_for(n, 256)),
                                   dim3(256), 0, stream, x, d, rhs, inv_diag, T(1.0 / theta),
                                   n);
              continue;
            }
          A(ax, x);
          double c1 = 0.0, c2;
          if (k == 0)
            c2 = 1.0 / theta;
          else
            {
              const double rho_new = 1.0 / (2.0 * sigma - rho);
              c1                   = rho_new * rho;
              c2                   = 2.0 * rho_new / delta;
              rho                  = rho_new;
            }
          if (tab.table)
            hipLaunchKernelGGL(chebyshev_table_kernel<T>, class_launch(tab.grid), dim3(256), 0,
                               stream, x, d, rhs, ax, inv_diag, tab.table, tab.ok, T(c1),
                               T(c2), k == 0 ? 1 : 0, 0, tab.grid);
          else
            hipLaunchKernelGGL(chebyshev_step_kernel<T>, dim3(grid_for(n, 256)), dim3(256), 0,
                               stream, x, d, rhs, ax, inv_diag, T(c1), T(c2), k == 0 ? 1 : 0,
                               n);
        }
    }

    // ----------------------------------------------------------------- multigrid
    template <typename T>
    T *
    lvec(Level &l, int i)
    {
      if constexpr (std::is_same_v<T, float>)
        return l.vf[i].get();
      else
        return l.vd[i].get();
    }

    /// The class table of level l's inverse diagonal (empty until setup_level built it).
    template <typename T>
    DiagTable<T>
    level_table(Level &l, int n_fields)
    {
      DiagTable<T> t;
      const T     *table = nullptr;
      if constexpr (std::is_same_v<T, float>)
        table = l.table_f.get();
      else
        table = l.table_d.get();
      if (!table || !l.table_ok.get())
        return t;
      t.table = table;
      t.ok    = l.table_ok.get();
      t.grid  = class_grid(l.grid, n_fields);
      return t;
    }

    template <typename T>
    T *
    linv(Level &l)
    {
      if constexpr (std::is_same_v<T, float>)
        return l.inv_diag_f.get();
      else
        return l.inv_diag_d.get();
    }

    template <typename T>
    LevelFields<T>
    level_fields(Level &l)
    {
      LevelFields<T> lf;
      lf.grid = &l.grid;
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        {
          if constexpr (std::is_same_v<T, float>)
            {
              for (std::size_t a = 0; a < l.field_f[f].size(); ++a)
                lf.ptr[f][a + 1] = l.field_f[f][a].get();
            }
          else
            {
              for (std::size_t a = 0; a < l.field_d[f].size(); ++a)
                lf.ptr[f][a + 1] = l.field_d[f][a].get();
            }
        }
      return lf;
    }

    template <typename T>
    void
    level_vmult(int b, Level &l, T *dst, const T *src)
    {
      auto s = split_level<T>(const_cast<T *>(src), b, l.grid);
      apply_operator<T>(b, true, level_fields<T>(l), {s.begin(), s.end()},
                        split_level<T>(dst, b, l.grid), level_scaling(b), dirichlet_zero, false);
    }

    int
    level_scaling(int b) const
    {
      return linear_scaling(b);
    }

    /// MGContext::init (reference linear_solver.h:86-165): level operators, level
    /// diagonals, Chebyshev smoothers, transfers.  Levels span
    /// [global_refinement - mg_depth + 1, global_refinement].  Called before every
    /// solve: the level fields and diagonals are refreshed each time (as
    /// do_linear_solve re-evaluates them, linear_solver.h:369-379); the smoothers'
    /// eigenvalue estimates are made once per level, at the first solve, because the
    /// reference initializes its smoothers once (:135-147) and deal.II's
    /// PreconditionChebyshev estimates at its first application only.
    void
    setup_multigrid(int b)
    {
      LinearState &ls    = linear_state(b);
      const auto  &lp    = lin_params(b);
      const int    top   = int(inputs.spatial_discretization.global_refinement);
      const int    depth = int(lp.mg_depth);
      if (depth < 1 || depth > top + 1)
        throw Error(status_invalid, "mg depth must be in [1, global refinement + 1]");
      const int  prec    = lp.level_precision == prismspf::LevelPrecision::Single ? 1 : 0;
      const bool rebuild = ls.levels.size() != std::size_t(depth) ||
                           (!ls.levels.empty() && ls.levels[0].precision != prec);
      if (rebuild)
        {
          ls.levels.clear();
          ls.levels.resize(depth);
          for (int i = 0; i < depth; ++i)
            {
              Level &l    = ls.levels[i];
              l.grid      = make_grid(unsigned(top - depth + 1 + i));
              l.precision = prec;
              const int64_t n = int64_t(blocks[b].field_indices.size()) * l.grid.n_nodes;
              for (int v = 0; v < 6; ++v)
                {
                  if (prec)
                    l.vf[v].resize(n);
                  else
                    l.vd[v].resize(n);
                }
              if (prec)
                l.inv_diag_f.resize(n);
              else
                l.inv_diag_d.resize(n);
            }
        }
      transfer_fields_to_levels(b, ls);
      for (auto &l : ls.levels)
        {
          if (prec)
            setup_level<float>(b, l);
          else
            setup_level<double>(b, l);
        }
      // FP64 coarse-grid CG storage lives on the coarsest level.
      if (lp.coarse_solver == prismspf::CoarseSolver::CG && prec)
        {
          Level        &c = ls.levels[0];
          const int64_t n = int64_t(blocks[b].field_indices.size()) * c.grid.n_nodes;
          for (int v = 0; v < 6; ++v)
            c.vd[v].resize(n);
          c.inv_diag_d.resize(n);
          inverse_diagonal<double>(b, level_fields<double>(c), level_scaling(b),
                                   c.inv_diag_d.get());
        }
    }

    /// Non-operand dependencies of the lhs are injected to every level
    /// (reference group_solution_handler.h:356-417).
    void
    transfer_fields_to_levels(int b, LinearState &ls)
    {
      // The level copies keep their buffers from one solve to the next (the block's
      // dependencies do not change), so no solve allocates or frees device memory.
      const auto &deps = blocks[b].dependencies_lhs;
      for (const auto &in : declared_inputs(deps))
        {
          if (in.slot == prismspf::dependency_slot(prismspf::SRC))
            continue;
          const int age = in.slot - 1;
          for (auto &l : ls.levels)
            {
              const bool want_f = l.precision == 1;
              const int  stride = fine.cells[0] / l.grid.cells[0];
              auto      &vd     = l.field_d[in.field];
              auto      &vf     = l.field_f[in.field];
              if (vd.size() <= std::size_t(age))
                vd.resize(age + 1);
              if (vf.size() <= std::size_t(age))
                vf.resize(age + 1);
              vd[age].resize(l.grid.n_nodes);
              hipLaunchKernelGGL((inject_kernel<dim, double, double>),
                                 dim3(grid_for(l.grid.n_nodes, 256)), dim3(256), 0, stream,
                                 vd[age].get(), l.grid, field_ptr(in.field, age), fine, stride);
              if (want_f)
                {
                  vf[age].resize(l.grid.n_nodes);
                  convert(vf[age].get(), vd[age].get(), l.grid.n_nodes);
                }
            }
        }
    }

    template <typename T>
    void
    setup_level(int b, Level &l)
    {
      const auto   &lp = lin_params(b);
      const int64_t n  = int64_t(blocks[b].field_indices.size()) * l.grid.n_nodes;
      inverse_diagonal<T>(b, level_fields<T>(l), level_scaling(b), linv<T>(l));
      if constexpr (std::is_same_v<T, float>)
        build_diag_table<float>(linv<T>(l), l.grid, int(blocks[b].field_indices.size()),
                                l.table_f, l.table_ok);
      else
        build_diag_table<double>(linv<T>(l), l.grid, int(blocks[b].field_indices.size()),
                                 l.table_d, l.table_ok);
      if (l.eigenvalues_estimated)
        return;
      estimate_eigenvalues<T>([&](T *dst, const T *src) { level_vmult<T>(b, l, dst, src); },
                              linv<T>(l), n, lp.chebyshev_parameters, l.lambda_max, l.lambda_min,
                              lvec<T>(l, 2), lvec<T>(l, 3), lvec<T>(l, 4), lvec<T>(l, 5));
      l.eigenvalues_estimated = true;
    }

    /// PreconditionMG::vmult: one V-cycle on the residual g.
    void
    multigrid_vmult(int b, double *h, const double *g)
    {
      LinearState &ls  = linear_state(b);
      Level       &top = ls.levels.back();
      if (top.precision)
        convert(lvec<float>(top, 1), g, ls.n);
      else
        copy(lvec<double>(top, 1), g, ls.n);
      run_vcycle(b);
      if (top.precision)
        convert(h, lvec<float>(top, 0), ls.n);
      else
        copy(h, lvec<double>(top, 0), ls.n);
    }

    /// One V-cycle on the top level's defect vector.  With `mg_graphs` set, the launch
    /// sequence is captured into a HIP graph once an eager V-cycle has shown that it
    /// needs no host synchronization and allocates nothing, and then replayed: the
    /// V-cycle's ~150 launches (7 levels) go to the GPU as one submission.  The graph is
    /// captured again after anything that changes the launches' arguments: a new split
    /// plan or option (capture epoch), a reallocated buffer (allocation count), another
    /// halt flag (increment parity) or another level precision.  V-cycles whose coarse
    /// solve is driven from the host stay eager.
    void
    run_vcycle(int b)
    {
      LinearState &ls    = linear_state(b);
      Level       &top   = ls.levels.back();
      const int    li    = int(ls.levels.size()) - 1;
      auto         eager = [&] {
        if (top.precision)
          v_cycle<float>(b, li);
        else
          v_cycle<double>(b, li);
      };
      if (!mg_graphs || ls.vg_off || (probe_req.pending && probe_req.block == b))
        {
          eager();
          return;
        }
      const long allocs = device_allocations();
      for (int s = 0; s < 2; ++s)
        if (ls.vg_exec[s] && ls.vg_halt[s] == enqueue_halt && ls.vg_epoch[s] == graph_epoch &&
            ls.vg_alloc[s] == allocs && ls.vg_prec[s] == top.precision)
          {
            PPF_HIP_CHECK(hipGraphLaunch(ls.vg_exec[s], stream));
            return;
          }
      if (ls.vg_warm != graph_epoch)
        {
          const long syncs = sync_count_;
          eager();
          if (sync_count_ != syncs)
            ls.vg_off = true;
          else if (device_allocations() == allocs)
            ls.vg_warm = graph_epoch;
          if (std::getenv("PPF_VERBOSE"))
            std::fprintf(stderr, "block %d: eager V-cycle (epoch %ld, syncs %ld, allocations "
                                 "%ld -> %ld)\n",
                         b, graph_epoch, sync_count_ - syncs, allocs, long(device_allocations()));
          return;
        }
      const int  slot  = ls.vg_next;
      hipGraph_t graph = nullptr;
      if (std::getenv("PPF_VERBOSE"))
        std::fprintf(stderr, "block %d: capturing the V-cycle (epoch %ld, allocations %ld)\n",
                     b, graph_epoch, allocs);
      PPF_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
      bool captured = true;
      try
        {
          eager();
        }
      catch (...)
        {
          captured = false;
        }
      const hipError_t end = hipStreamEndCapture(stream, &graph);
      hipGraphExec_t   exec = nullptr;
      if (captured && end == hipSuccess && graph && ls.vg_exec[slot])
        {
          // same launch sequence with new arguments: update the instantiated graph
          hipGraphNode_t              bad = nullptr;
          hipGraphExecUpdateResult    res {};
          if (hipGraphExecUpdate(ls.vg_exec[slot], graph, &bad, &res) == hipSuccess &&
              res == hipGraphExecUpdateSuccess)
            exec = ls.vg_exec[slot];
          else
            {
              (void) hipGetLastError();
              (void) hipGraphExecDestroy(ls.vg_exec[slot]);
              ls.vg_exec[slot] = nullptr;
            }
        }
      if (captured && end == hipSuccess && graph &&
          (exec || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess))
        {
          ls.vg_exec[slot]  = exec;
          ls.vg_halt[slot]  = enqueue_halt;
          ls.vg_epoch[slot] = graph_epoch;
          ls.vg_alloc[slot] = allocs;
          ls.vg_prec[slot]  = top.precision;
          ls.vg_next        = slot ^ 1;
        }
      if (graph)
        (void) hipGraphDestroy(graph);
      if (!exec)
        {
          // not capturable after all: this block's V-cycles stay eager
          (void) hipGetLastError();
          ls.vg_off = true;
          eager();
          return;
        }
      PPF_HIP_CHECK(hipGraphLaunch(exec, stream));
    }

    /// Multigrid::level_v_step.  Level vectors: 0 solution, 1 defect, 2 scratch t,
    /// 3-4 smoother temporaries, 5 spare.
    template <typename T>
    void
    v_cycle(int b, int li)
    {
      LinearState  &ls = linear_state(b);
      Level        &l  = ls.levels[li];
      const int64_t n  = int64_t(blocks[b].field_indices.size()) * l.grid.n_nodes;
      const auto   &lp = lin_params(b);
      auto A = [&](T *dst, const T *src) { level_vmult<T>(b, l, dst, src); };
      if (li == 0)
        {
          coarse_solve<T>(b, l);
          return;
        }
      const unsigned int deg = lp.chebyshev_parameters.degree;
      const DiagTable<T> tab = level_table<T>(l, int(blocks[b].field_indices.size()));
      chebyshev<T>(A, linv<T>(l), n, deg, l.lambda_max, l.lambda_min, lvec<T>(l, 0),
                   lvec<T>(l, 1), lvec<T>(l, 3), lvec<T>(l, 4), true, tab);
      // t = defect - A u
      A(lvec<T>(l, 2), lvec<T>(l, 0));
      axpby(lvec<T>(l, 2), T(1), lvec<T>(l, 1), T(-1), n);
      Level &c = ls.levels[li - 1];
      for (std::size_t k = 0; k < blocks[b].field_indices.size(); ++k)
        {
          const int f = int(*std::next(blocks[b].field_indices.begin(), k));
          restrict_separable<T>(lvec<T>(c, 1) + k * c.grid.n_nodes, c.grid,
                                lvec<T>(l, 2) + k * l.grid.n_nodes, l, dirichlet_mask(f));
        }
      v_cycle<T>(b, li - 1);
      for (std::size_t k = 0; k < blocks[b].field_indices.size(); ++k)
        {
          const int f = int(*std::next(blocks[b].field_indices.begin(), k));
          prolongate_add_separable<T>(lvec<T>(l, 0) + k * l.grid.n_nodes, l,
                                      lvec<T>(c, 0) + k * c.grid.n_nodes, c.grid,
                                      dirichlet_mask(f));
        }
      chebyshev<T>(A, linv<T>(l), n, deg, l.lambda_max, l.lambda_min, lvec<T>(l, 0),
                   lvec<T>(l, 1), lvec<T>(l, 3), lvec<T>(l, 4), false, tab);
    }

    /// Scratch lattices of the separable transfers between level l and the next
    /// coarser one: [0] coarse x-extent by the fine y, z extents; [1] coarse x, y by fine z.
    template <typename T>
    T *
    transfer_tmp(Level &l, int i, int64_t n)
    {
      auto &buf = [&]() -> auto & {
        if constexpr (std::is_same_v<T, float>)
          return l.tmp_f[i];
        else
          return l.tmp_d[i];
      }();
      if (buf.size() < n)
        buf.resize(n);
      return buf.get();
    }

    /// fine += P coarse (fine Dirichlet nodes untouched), one pass per axis.
    template <typename T>
    void
    prolongate_add_separable(T *fine_v, Level &l, const T *coarse_v, const GridDesc &gc,
                             unsigned int mask)
    {
      const GridDesc &gf = l.grid;
      const int c0 = gc.nodes[0], c1 = gc.nodes[1], c2 = gc.nodes[2];
      const int f0 = gf.nodes[0], f1 = gf.nodes[1], f2 = gf.nodes[2];
      const T  *in = coarse_v;
      if constexpr (dim == 3)
        {
          // z: (c0, c1, c2) -> (c0, c1, f2)
          T *t1 = transfer_tmp<T>(l, 1, int64_t(c0) * c1 * f2);
          hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, false, 0, 2>),
                             transfer_grid(c0, c1, f2), dim3(256), 0, stream, t1, in, c0, c1,
                             c0, c1, gc, gf, 0u);
          in = t1;
        }
      // y: (c0, c1, f2) -> (c0, f1, f2)
      T *t0 = transfer_tmp<T>(l, 0, int64_t(c0) * f1 * f2);
      hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, false, 0, 1>),
                         transfer_grid(c0, f1, f2), dim3(256), 0, stream, t0, in, c0, f1, c0,
                         c1, gc, gf, 0u);
      // x: (c0, f1, f2) -> fine, added
      hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, false, 1, 0>),
                         transfer_grid(f0, f1, f2), dim3(256), 0, stream, fine_v, t0, f0, f1,
                         c0, f1, gc, gf, mask);
    }

    /// Launch grid of transfer_axis_kernel for an e0 x e1 x e2 output lattice.
    static dim3
    transfer_grid(int e0, int e1, int e2)
    {
      const int64_t plane = int64_t(e0) * e1;
      if (plane >= (int64_t(1) << 31) || e2 > 65535)
        throw Error(status_unsupported, "transfer: lattice too large for one launch");
      return dim3(grid_for(plane, 256), unsigned(e2));
    }

    /// coarse = P^T fine (coarse Dirichlet nodes zero), one pass per axis.
    template <typename T>
    void
    restrict_separable(T *coarse_v, const GridDesc &gc, const T *fine_v, Level &l,
                       unsigned int mask)
    {
      const GridDesc &gf = l.grid;
      const int c0 = gc.nodes[0], c1 = gc.nodes[1], c2 = gc.nodes[2];
      const int f0 = gf.nodes[0], f1 = gf.nodes[1], f2 = gf.nodes[2];
      // x: (f0, f1, f2) -> (c0, f1, f2)
      T *t0 = transfer_tmp<T>(l, 0, int64_t(c0) * f1 * f2);
      hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, true, 0, 0>),
                         transfer_grid(c0, f1, f2), dim3(256), 0, stream, t0, fine_v, c0, f1,
                         f0, f1, gf, gc, 0u);
      if constexpr (dim == 3)
        {
          // y: (c0, f1, f2) -> (c0, c1, f2)
          T *t1 = transfer_tmp<T>(l, 1, int64_t(c0) * c1 * f2);
          hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, true, 0, 1>),
                             transfer_grid(c0, c1, f2), dim3(256), 0, stream, t1, t0, c0, c1,
                             c0, f1, gf, gc, 0u);
          // z: (c0, c1, f2) -> coarse, Dirichlet nodes zero
          hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, true, 2, 2>),
                             transfer_grid(c0, c1, c2), dim3(256), 0, stream, coarse_v, t1,
                             c0, c1, c0, c1, gf, gc, mask);
        }
      else
        // y: (c0, f1) -> coarse, Dirichlet nodes zero
        hipLaunchKernelGGL((transfer_axis_kernel<dim, degree, T, true, 2, 1>),
                             transfer_grid(c0, c1, c2), dim3(256), 0, stream, coarse_v, t0,
                             c0, c1, c0, f1, gf, gc, mask);
    }

    template <typename T>
    void
    coarse_solve(int b, Level &l)
    {
      const auto   &lp = lin_params(b);
      const int64_t n  = int64_t(blocks[b].field_indices.size()) * l.grid.n_nodes;
      if (lp.coarse_solver == prismspf::CoarseSolver::Smoother)
        {
          // MGCoarseGridApplySmoother (reference linear_solver.h:65,145).
          chebyshev<T>([&](T *dst, const T *src) { level_vmult<T>(b, l, dst, src); },
                       linv<T>(l), n, lp.chebyshev_parameters.degree, l.lambda_max,
                       l.lambda_min, lvec<T>(l, 0), lvec<T>(l, 1), lvec<T>(l, 3), lvec<T>(l, 4),
                       true);
          return;
        }
      // FP64 CG with the diagonal preconditioner.
      double *x = l.vd[0].get(), *r = l.vd[1].get(), *p = l.vd[2].get(), *ap = l.vd[3].get(),
             *z = l.vd[4].get();
      if constexpr (std::is_same_v<T, float>)
        convert(r, l.vf[1].get(), n);
      auto A = [&](double *dst, const double *src) { level_vmult<double>(b, l, dst, src); };
      if (l.grid.n_nodes <= coarse_fused_max_nodes)
        {
          // One workgroup runs the whole solve when the level operator is a stencil
          // operator (the operator launch below is taken over); otherwise that launch
          // is an ordinary application and the host-driven loop follows.
          coarse_req     = CoarseCgArgs {x, r, p, ap, z, l.inv_diag_d.get(), n,
                                     lp.coarse_cg_tolerance, int(lp.coarse_cg_max_iterations)};
          coarse_pending = true;
          A(ap, p);
          const bool fused = !coarse_pending;
          coarse_pending   = false;
          if (fused)
            {
              if constexpr (std::is_same_v<T, float>)
                convert(l.vf[0].get(), x, n);
              return;
            }
        }
      fill(x, 0.0, n);
      // Scalars and the stop flag stay on the device; the host looks at the flag every
      // few steps (steps after convergence change nothing: their updates are skipped).
      double *cs   = coarse_slots.get();
      int    *done = coarse_flag.get();
      dot_to(r, r, n, cs + cs_r0sq);
      hipLaunchKernelGGL(coarse_cg_start_kernel, dim3(1), dim3(1), 0, stream, cs, done);
      const unsigned int nb = grid_for(n, 256);
      hipLaunchKernelGGL(diag_mult_kernel<double>, dim3(nb), dim3(256), 0, stream, z,
                         l.inv_diag_d.get(), r, n);
      copy(p, z, n);
      dot_to(r, z, n, cs + cs_rz);
      constexpr unsigned int check_every = 4;
      for (unsigned int it = 0; it < lp.coarse_cg_max_iterations; ++it)
        {
          const int par = int(it & 1u);
          A(ap, p);
          dot_to(p, ap, n, cs + cs_pap);
          hipLaunchKernelGGL(coarse_cg_xr_kernel, dim3(nb), dim3(256), 0, stream, x, r, p, ap,
                             cs, done, par, n);
          dot_to(r, r, n, cs + cs_rr);
          hipLaunchKernelGGL(coarse_cg_check_kernel, dim3(1), dim3(1), 0, stream, cs, done,
                             lp.coarse_cg_tolerance);
          hipLaunchKernelGGL(diag_mult_kernel<double>, dim3(nb), dim3(256), 0, stream, z,
                             l.inv_diag_d.get(), r, n);
          dot_to(r, z, n, cs + cs_rz + (par ^ 1));
          hipLaunchKernelGGL(coarse_cg_p_kernel, dim3(nb), dim3(256), 0, stream, p, z, cs, done,
                             par, n);
          if (it % check_every == check_every - 1)
            {
              int h = 0;
              PPF_HIP_CHECK(hipMemcpyAsync(&h, done, sizeof(int), hipMemcpyDeviceToHost, stream));
              sync();
              if (h)
                break;
            }
        }
      PPF_HIP_CHECK(hipGetLastError());
      if constexpr (std::is_same_v<T, float>)
        convert(l.vf[0].get(), x, n);
    }

    // -------------------------------------------------------------- time loop
    /// Initial conditions at the nodes (reference solver_base.h:179-229 interpolates the
    /// user function at the support points).  Nodes are visited in lexicographic order;
    /// periodic images take the master's value.
    void
    apply_initial_conditions(int b)
    {
      for (auto f : blocks[b].field_indices)
        {
          std::vector<double> host(fine.n_nodes);
          for (int64_t node = 0; node < fine.n_nodes; ++node)
            {
              int i[3];
              i[0] = int(node % fine.nodes[0]);
              i[1] = int((node / fine.nodes[0]) % fine.nodes[1]);
              i[2] = dim == 3 ? int(node / (int64_t(fine.nodes[0]) * fine.nodes[1])) : 0;
              dealii::Point<dim> p;
              for (int d = 0; d < dim; ++d)
                p[d] = fine.lower[d] + i[d] * fine.h[d] / degree;
              double     value = 0.0, vec = 0.0;
              const auto &lc    = layout[f];
              pde_d->set_initial_condition(unsigned(lc.declared), lc.component, p, value, vec);
              host[node] = lc.vector ? vec : value;
            }
          PPF_HIP_CHECK(hipMemcpyAsync(field_ptr(int(f), 0), host.data(),
                                       sizeof(double) * fine.n_nodes, hipMemcpyHostToDevice,
                                       stream));
          sync();
        }
    }

    void
    copy_current_to_old(int b)
    {
      for (auto f : blocks[b].field_indices)
        for (std::size_t a = 1; a < storage[f].size(); ++a)
          copy(storage[f][a].get(), storage[f][0].get(), fine.n_nodes);
    }

    /// SolverBase::solve (reference solver_base.h:106-130).
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
        copy_current_to_old(b);
      for (auto f : blocks[b].field_indices)
        rotated[f] = false;
    }

    /// Old-state rotation by pointer swaps (reference group_solution_handler.cc:369-403).
    void
    rotate(int b)
    {
      for (auto f : blocks[b].field_indices)
        {
          for (int age = int(storage[f].size()) - 1; age >= 1; --age)
            storage[f][age].swap(storage[f][age - 1]);
          rotated[f] = storage[f].size() > 1;
        }
    }

    /// Age holding the newest values of field f.  Every increment ends with the
    /// rotation, after which the reference's solution handler keeps the newest values
    /// in old_1 (the current slot is then overwritten by the next solve).
    int
    latest_age(int f) const
    {
      return rotated[f] ? 1 : 0;
    }

    void
    rotate_all()
    {
      for (std::size_t b = 0; b < blocks.size(); ++b)
        rotate(int(b));
    }

    /// Inverse of rotate_all (ages shifted back by one).
    void
    unrotate_all()
    {
      for (const auto &blk : blocks)
        for (auto f : blk.field_indices)
          for (int age = 1; age < int(storage[f].size()); ++age)
            storage[f][age].swap(storage[f][age - 1]);
    }

    /// Enqueues the solve blocks of the current increment and its NaN probe of entry 0
    /// of every field, whose result goes to pinned slot `slot`; no host wait.  The probe
    /// rides on the last operator launch when that launch is the stencil kernel of an
    /// explicit block, else it is a kernel of its own.
    ///
    /// With `halt` (a ring of `ring` words) given, increment k's launches read word
    /// halt[slot] (slot = k mod ring) and do nothing once it is set; increment k's probe
    /// sets every other word on a non-finite entry, so the up to ring - 1 increments
    /// enqueued after k become no-ops, and a probe that shares a launch with its
    /// increment never stops that increment's own workgroups.
    void
    enqueue_increment(bool do_output, int slot, int *halt, int ring = 2)
    {
      const unsigned int inc = timer.get_increment();
      if (any_time_dependent_dirichlet())
        update_time_dependent_dirichlet();
      std::vector<int>   solve;
      for (std::size_t b = 0; b < blocks.size(); ++b)
        {
          const auto timing = blocks[b].solve_timing;
          if ((timing == prismspf::PostProcess || timing == prismspf::NucleationRate) &&
              !do_output)
            continue;
          solve.push_back(int(b));
        }
      ProbeFields pf {};
      pf.n = int(fields.size());
      for (std::size_t f = 0; f < fields.size(); ++f)
        pf.ptr[f] = field_ptr(int(f), 0);
      probe_tag[slot] = int(++probe_count & 0x3ffffffu);
      int *result     = probe_flag.get(max_pipeline_depth + 1) + slot;
      *reinterpret_cast<volatile int *>(result) = -1;
      probe_req         = {};
      probe_req.result  = result;
      probe_req.halt    = halt;
      probe_req.hmask   = halt ? ((1u << ring) - 1u) & ~(1u << slot) : 0u;
      probe_req.tag     = probe_tag[slot];
      probe_req.fields  = pf;
      probe_req.block   = solve.empty() ? -1 : solve.back();
      probe_req.pending = !solve.empty() && inc > 0 &&
                          blocks[solve.back()].solve_type == prismspf::Explicit;
      enqueue_halt = halt ? halt + slot : nullptr;
      for (std::size_t i = 0; i < solve.size(); ++i)
        {
          if (i + 1 < solve.size() && solve_explicit_pair(solve[i], solve[i + 1]))
            {
              ++i;
              continue;
            }
          solve_block(solve[i]);
        }
      enqueue_halt = nullptr;
      if (!probe_req.taken)
        {
          hipLaunchKernelGGL(finite_probe_fields_kernel, dim3(1), dim3(1), 0, stream, result,
                             pf, halt, probe_req.hmask, probe_tag[slot]);
          PPF_HIP_CHECK(hipGetLastError());
        }
      probe_req = {};
    }

    /// Waits for the probe result in `slot` by polling its pinned word (no event or
    /// stream synchronization, whose completion markers would cost the GPU several
    /// microseconds per increment); true if the probe found a non-finite entry.
    bool
    probe_found_nan(int slot)
    {
      const volatile int *word = probe_flag.get(max_pipeline_depth + 1) + slot;
      const int           tag  = probe_tag[slot];
      for (unsigned long spin = 1;; ++spin)
        {
          const int v = *word;
          if ((v >> 5) == tag)
            return (v & 31) != 0;
          if (spin % 4096 == 0)
            {
              // Ends the wait if the stream has failed or drained without the result.
              const hipError_t q = hipStreamQuery(stream);
              if (q == hipSuccess)
                {
                  const int w = *word;
                  if ((w >> 5) != tag)
                    throw Error(status_internal, "NaN probe result missing");
                  return (w & 31) != 0;
                }
              if (q != hipErrorNotReady)
                PPF_HIP_CHECK(q);
            }
        }
    }

    /// Problem::solve_increment (reference problem.cc:414-558).  Returns the exit status:
    /// 0 normal, 2 NaN detected, 3 user stop.  The probe's host wait is the increment's
    /// only one.
    int
    solve_increment()
    {
      const bool do_output = inputs.output_parameters.should_output(timer.get_increment());
      enqueue_increment(do_output, 0, nullptr);
      int status = probe_found_nan(0) ? 2 : 0;
      if (status == 0 && pde_d->get_user_stop())
        status = 3;
      if (do_output || status != 0)
        {
          if (write_output)
            write_output_files();
          log_norms();
        }
      rotate_all();
      return status;
    }

    /// Whether increment `inc` may be enqueued before the probe of the increment
    /// before it has been read: it solves explicit blocks only (no host-side
    /// convergence test), writes no output, and no user stop is set.
    bool
    speculable(unsigned int inc) const
    {
      if (inc == 0 || inputs.output_parameters.should_output(inc) || pde_d->get_user_stop() ||
          any_time_dependent_dirichlet())
        return false;
      for (const auto &blk : blocks)
        {
          const auto timing = blk.solve_timing;
          if (timing == prismspf::PostProcess || timing == prismspf::NucleationRate)
            continue; // skipped on increments without output
          if (blk.solve_type != prismspf::Explicit)
            return false;
        }
      return true;
    }

    /// Host-side state that an enqueued increment changes: the timer and the
    /// rotation flags (the age order of the storage is undone by unrotate_all).
    struct HostState
    {
      prismspf::SimulationTimer             timer;
      std::array<bool, PPF_MAX_FIELDS>      rotated;
    };

    /// Problem::solve (reference problem.cc:340-412): increments until n_increments,
    /// stopping after the first increment whose probe finds a non-finite entry
    /// (status 2) or on a user stop (3).
    ///
    /// Runs of explicit increments without output are pipelined `pipeline_depth`
    /// deep: increments k+1 .. k+depth are enqueued before increment k's probe is
    /// read, so the GPU neither idles on the host round trip nor on a late host
    /// thread.  Their launches carry device flags that increment k's probe sets on a
    /// NaN, which makes them no-ops; the host then undoes their rotations and timer
    /// steps.  Results, the log and the final state are those of the one-at-a-time
    /// loop.
    int
    run(unsigned int n_increments)
    {
      int       status = 0;
      int      *halt   = halt_flag.get();
      const int depth  = std::clamp(pipeline_depth, 1, max_pipeline_depth);
      const int ring   = depth + 1;
      hipLaunchKernelGGL(fill_kernel<int>, dim3(1), dim3(64), 0, stream, halt, 0,
                         int64_t(ring));
      PPF_HIP_CHECK(hipGetLastError());
      struct Pending
      {
        int       slot;
        HostState before;
      };
      std::deque<Pending> pending; // enqueued increments whose probes are unread
      // NaN in the pending increment: back to the state before its rotation, log its
      // norms, then rotate and step the timer as the one-at-a-time loop does.
      auto settle_nan = [&](const HostState &before) {
        unrotate_all();
        timer   = before.timer;
        rotated = before.rotated;
        if (write_output)
          write_output_files();
        log_norms();
        rotate_all();
        timer.increment();
        return 2;
      };
      // The oldest pending increment found a NaN: the increments enqueued after it did
      // nothing on the device; undo their rotations and timer steps (newest first),
      // then settle the NaN increment.
      auto settle_oldest = [&](const Pending &p) {
        for (auto it = pending.rbegin(); it != pending.rend(); ++it)
          {
            unrotate_all();
            timer   = it->before.timer;
            rotated = it->before.rotated;
          }
        pending.clear();
        return settle_nan(p.before);
      };
      // Reads the oldest pending probe; nonzero status on a NaN.
      auto read_oldest = [&]() {
        const Pending p = pending.front();
        pending.pop_front();
        return probe_found_nan(p.slot) ? settle_oldest(p) : 0;
      };
      while (status == 0 && timer.get_increment() <= n_increments)
        {
          const unsigned int inc = timer.get_increment();
          if (!speculable(inc))
            {
              if (!pending.empty())
                {
                  status = read_oldest();
                  continue;
                }
              status = solve_increment();
              timer.increment();
              continue;
            }
          const HostState before {timer, rotated};
          const int       slot = int(inc % unsigned(ring));
          const auto      t0   = std::chrono::steady_clock::now();
          enqueue_increment(false, slot, halt, ring);
          rotate_all();
          timer.increment();
          pending.push_back({slot, before});
          enqueue_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
          ++enqueue_count;
          if (int(pending.size()) > depth)
            status = read_oldest();
        }
      while (status == 0 && !pending.empty())
        status = read_oldest();
      sync();
      if (std::getenv("PPF_VERBOSE") && enqueue_count)
        std::fprintf(stderr, "run: %ld increments enqueued ahead, %.2f us of host time each\n",
                     enqueue_count, 1e6 * enqueue_seconds / double(enqueue_count));
      return status;
    }

    /// l2-norm over deal.II's node set and the integral sum(u * jxw) of a field
    /// (reference problem.cc:517-549).
    std::pair<double, double>
    norms(int f, int age = 0)
    {
      const GridDesc &g = fine;
      hipLaunchKernelGGL((norms_partial_kernel<dim, degree, double>), dim3(reduce_blocks),
                         dim3(reduce_threads), 0, stream, reduce_buf.get(), g,
                         field_ptr(f, age));
      const double s2 = finish_reduction(reduce_blocks, 0);
      const double si = finish_reduction(reduce_blocks, reduce_blocks);
      return {std::sqrt(s2), si};
    }

    void
    log_norms()
    {
      std::ostringstream os;
      os.precision(12);
      os << "Iteration " << timer.get_increment() << " Time " << timer.get_time() << "\n";
      for (std::size_t f = 0; f < fields.size(); ++f)
        {
          const auto [l2, integral] = norms(int(f));
          os << "  Field index " << f << " name " << fields[f].name << " l2-norm " << l2
             << " integrated value " << integral << "\n";
        }
      log += os.str();
      if (std::getenv("PPF_VERBOSE"))
        std::fputs(os.str().c_str(), stdout);
    }

    int
    block_index(int id) const
    {
      for (std::size_t b = 0; b < blocks.size(); ++b)
        if (blocks[b].id == id)
          return int(b);
      throw Error(status_invalid, "no solve block with id " + std::to_string(id));
    }

    // ------------------------------------------------------------------- state
    prismspf::UserInputParameters<dim>    inputs;
    prismspf::PhaseFieldTools<dim>        tools;
    std::vector<prismspf::FieldAttributes> fields;
    std::vector<prismspf::StoredComponent>  layout;
    std::vector<prismspf::SolveBlock>      blocks;
    std::unique_ptr<PDE<double>>           pde_d;
    std::unique_ptr<PDE<float>>            pde_f;
    prismspf::SimulationTimer              timer;
    GridDesc                               fine {};
    hipStream_t                            stream = nullptr;
    OperatorPath                           path   = OperatorPath::Auto;
    /// Tile shape of the stencil kernel: rows per wave and waves per workgroup (0:
    /// defaults), and z-planes per workgroup (0: chosen to give about
    /// `stencil_target_blocks` groups).
    int                                    stencil_rows    = 0;
    int                                    stencil_waves   = 0;
    int                                    stencil_z_chunk = 0;
    /// Workgroup -> tile order of the stencil kernel (StencilArgs::map_mode).
    int                                    stencil_map     = 0;
    /// Generic path kernel: the one-pass cell-block kernel (true) or the two-pass
    /// cell/gather kernels through a global scratch array (false).
    bool                                   generic_cell_block = true;
    /// z layers per column of cell_column_kernel (0: enough columns to fill the device).
    int                                    column_chunk       = 0;
    int64_t                                n_cus_     = 0;
    int64_t                                lds_limit_ = 0;
    int                                    stencil_target_blocks = 2048;
    /// Cache policy of the stencil kernel's output stores (Auto: non-temporal for
    /// launches whose arrays exceed stencil_nt_min_bytes).
    StorePolicy                            stencil_store_policy = StorePolicy::Auto;
    int64_t                                stencil_nt_min_bytes = int64_t(192) << 20;
    /// Lattices with fewer nodes run split operators node by node (launch_stencil).
    int64_t                                stencil_min_nodes    = 100000;
    /// Start every linear solve from zero instead of the previous solution (which the
    /// reference keeps as the next initial guess, linear_solver.h:330-336).
    bool                                   zero_initial_guess    = false;
    /// FP32 operator arithmetic for explicit blocks (FP64 state), with its buffers;
    /// explicit_fp32_copies: always through FP32 copies of the inputs and outputs (the
    /// stencil kernel otherwise loads and stores the FP64 state itself).
    bool                                   explicit_fp32         = false;
    bool                                   explicit_fp32_copies  = false;
    std::set<int>                          explicit_mixed_noted;
    /// Consecutive explicit blocks in one launch (solve_explicit_pair): on/off (off by
    /// default: at 128^3 Q2 the fused launch takes 45 us against 2 x 13 us for the two
    /// stencil launches; its 5-wave workgroups hold 230 registers a lane, one
    /// workgroup per compute unit), B rows per main wave of a tile (2 or 3), z chunk
    /// (0: automatic).
    bool                                   explicit_pairs = false;
    int                                    pair_rows      = 3;
    int                                    pair_z_chunk   = 0;
    std::set<int>                          explicit_pair_noted, explicit_pair_skipped;
    std::vector<DeviceBuffer<float>>       explicit_in_f, explicit_out_f, dirichlet_values_f;

    std::vector<std::vector<DeviceBuffer<double>>> storage;
    std::vector<DeviceBuffer<double>>              dirichlet_values;
    DeviceBuffer<double>                           krylov_slots; // CG / BiCGStab scalars
    bool write_output = false; // option write_output: solution files at output increments
    /// Boundary nodes (and the boundary id that set each) of every field's Dirichlet
    /// faces, for re-evaluating time-dependent values.
    std::vector<std::vector<int64_t>>              dirichlet_nodes;
    std::vector<std::vector<unsigned int>>         dirichlet_bids;
    std::vector<DeviceBuffer<int64_t>>             dirichlet_idx;
    std::vector<DeviceBuffer<double>>              dirichlet_stage;
    std::vector<std::vector<double>>               dirichlet_host;
    std::map<int, LinearState>                     linear_states;
    std::map<OpKey, PlanEntry>                     plans;
    DeviceBuffer<double>                           scratch_d;
    DeviceBuffer<float>                            scratch_f;
    /// Reference output of the split-path comparison (first use of a plan).
    DeviceBuffer<double>                           verify_d;
    DeviceBuffer<float>                            verify_f;
    DeviceBuffer<double>                           zeros_d;
    DeviceBuffer<float>                            zeros_f;
    DeviceBuffer<double>                           reduce_buf;
    DeviceBuffer<double>                           scalar_buf;
    DeviceBuffer<double>                           eig_slots;
    DeviceBuffer<double>                           coarse_slots;
    /// Coarse-grid solve handed to the next lhs stencil launch (coarse_cg_kernel).
    CoarseCgArgs                                   coarse_req {};
    bool                                           coarse_pending = false;
    /// Largest coarse lattice solved by one workgroup.
    int64_t                                        coarse_fused_max_nodes = 4096;
    DeviceBuffer<int>                              coarse_flag;
    PinnedBuffer<int>                              probe_flag; // probe slots (pipeline ring)
    int                                            probe_tag[max_pipeline_depth + 1] = {};
    /// Increments enqueued ahead of the oldest unread NaN probe (option pipeline_depth).
    int                                            pipeline_depth = 2;
    /// Host time spent enqueueing pipelined increments (reported with PPF_VERBOSE).
    double                                         enqueue_seconds = 0.0;
    long                                           enqueue_count   = 0;
    unsigned int                                   probe_count  = 0;
    /// The NaN probe of the increment being enqueued, until a launch takes it.
    struct ProbeRequest
    {
      int        *result  = nullptr;
      int        *halt    = nullptr;
      unsigned    hmask   = 0;
      int         tag     = 0;
      ProbeFields fields  = {};
      int         block   = -1;
      bool        pending = false;
      bool        taken   = false;
    } probe_req;
    DeviceBuffer<int>                              halt_flag;
    const int                                     *enqueue_halt = nullptr;
    /// Capture V-cycles into HIP graphs (run_vcycle); graph_epoch changes whenever a
    /// captured launch's arguments may have changed.
    bool                                           mg_graphs   = true;
    long                                           graph_epoch = 0;
    long                                           sync_count_ = 0;
    std::string                                    log;
    std::array<bool, PPF_MAX_FIELDS>               rotated {};
  };
} // namespace ppf
