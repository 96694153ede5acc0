// Implementation of the C-ABI declared in include/ppf.h for one application.
//
// An application translation unit defines its `App` description, then
//     #define PPF_APP App
//     #include "capi_impl.hpp"
// which instantiates the runtime and kernels for that application and exports the
// `ppf_*` symbols.  Exceptions never cross the boundary: each entry point catches,
// records the message on the context, and returns a status code.
#pragma once

#ifndef PPF_APP
#  error "define PPF_APP to the application description before including capi_impl.hpp"
#endif

#include "adjoint.hpp"
#include "runtime.hpp"

#include <ppf.h>

#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>

namespace ppf
{
  /// Size of the device scratch of ppf_apply_operator_noalloc (per GPU, per application
  /// library).
  constexpr std::size_t noalloc_scratch_bytes = std::size_t(PPF_NOALLOC_SCRATCH_BYTES);
} // namespace ppf

/// Device scratch of ppf_apply_operator_noalloc.  Each GPU the library runs on holds its
/// own copy until the process ends.  In host code this name is only a stand-in for the
/// symbol: kernels and copies take the device address from noalloc_scratch_ptr().
static __device__ __align__(256) unsigned char scratch[ppf::noalloc_scratch_bytes];

namespace ppf
{
  /// Device address of `scratch` on the current device, looked up once per device (with
  /// that device current) and kept.
  inline unsigned char *
  noalloc_scratch_ptr()
  {
    constexpr int         max_devices = 64;
    static unsigned char *ptr[max_devices] = {};
    static std::mutex     lock;
    int                   dev = 0;
    PPF_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= max_devices)
      throw Error(status_invalid, "device index out of range");
    std::lock_guard<std::mutex> g(lock);
    if (!ptr[dev])
      {
        void *p = nullptr;
        PPF_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = static_cast<unsigned char *>(p);
      }
    return ptr[dev];
  }

  template <class A, class = void>
  struct has_free_energy : std::false_type
  {};
  template <class A>
  struct has_free_energy<A, std::void_t<typename A::FreeEnergy>> : std::true_type
  {};

  struct Context
  {
    explicit Context(const std::string &prm)
      : rt(prm)
      , scratch(noalloc_scratch_ptr())
    {}
    Runtime<PPF_APP> rt;
    /// Device address of the scratch array on the context's device (set up at creation).
    unsigned char   *scratch = nullptr;
    std::string      error;
    bool             synchronize = true;
    /// The adjoint driver of the application (created on first use), whose checkpoint
    /// and segment buffers stay allocated for the following calls.
    std::shared_ptr<void> adjoint;
  };

  inline std::string &
  global_error()
  {
    static thread_local std::string e;
    return e;
  }

  template <class F>
  int
  guarded(ppf_ctx *ctx, F &&f)
  {
    auto *c = reinterpret_cast<Context *>(ctx);
    if (!c)
      {
        global_error() = "null context";
        return PPF_ERR_INVALID;
      }
    try
      {
        f(*c);
        c->error.clear();
        return PPF_OK;
      }
    catch (const Error &e)
      {
        c->error = e.what();
        return e.code;
      }
    catch (const std::exception &e)
      {
        c->error = e.what();
        return PPF_ERR_INVALID;
      }
    catch (...)
      {
        c->error = "unknown error";
        return PPF_ERR_INTERNAL;
      }
  }

  /// Applications whose explicit step is the Cahn-Hilliard pair of blocks declare
  /// `static constexpr bool cahn_hilliard_adjoint = true`; the others get the Allen-Cahn
  /// adjoint.
  template <class A, class = void>
  struct is_cahn_hilliard_adjoint : std::false_type
  {};
  template <class A>
  struct is_cahn_hilliard_adjoint<A, std::void_t<decltype(A::cahn_hilliard_adjoint)>>
    : std::bool_constant<A::cahn_hilliard_adjoint>
  {};

  template <class A>
  auto &
  adjoint_of(Context &c)
  {
    using Adjoint =
      std::conditional_t<is_cahn_hilliard_adjoint<A>::value,
                         CahnHilliardAdjoint<A::dim, A::degree, typename A::FreeEnergy>,
                         AllenCahnAdjoint<A::dim, A::degree, typename A::FreeEnergy>>;
    if (!c.adjoint)
      c.adjoint = std::make_shared<Adjoint>(c.rt.fine, c.rt.stream);
    return *static_cast<Adjoint *>(c.adjoint.get());
  }

  /// Adjoint entry points for applications that declare a free energy (the discarded
  /// branch must depend on a template parameter).
  template <class A>
  void
  adjoint_forward(Context &c, const double *phi0, double L, double kappa, unsigned n,
                  int precision, double *phi_final)
  {
    if constexpr (has_free_energy<A>::value)
      {
        auto &rt = c.rt;
        adjoint_of<A>(c).forward(phi0, phi_final, rt.timer.get_timestep(), L, kappa, n,
                                 precision);
        rt.sync();
      }
    else
      throw Error(status_unsupported, "this application declares no free energy for the adjoint");
  }

  template <class A>
  void
  adjoint_vjp(Context &c, const double *phi0, const double *lam, double L, double kappa,
              unsigned n, unsigned every, int precision, double &dL, double &dk, double *dphi0)
  {
    if constexpr (has_free_energy<A>::value)
      {
        auto &rt = c.rt;
        adjoint_of<A>(c).vjp(phi0, lam, rt.timer.get_timestep(), L, kappa, n, every, precision,
                             dL, dk, dphi0);
      }
    else
      throw Error(status_unsupported, "this application declares no free energy for the adjoint");
  }
} // namespace ppf

namespace ppf
{
  /// Sets the runtime's borrowed device region for the duration of one call.
  template <class R>
  struct BorrowScope
  {
    BorrowScope(R &rt_, unsigned char *base, int64_t bytes)
      : rt(rt_)
    {
      rt.borrowed = {base, bytes, 0};
    }
    ~BorrowScope()
    {
      rt.borrowed = {};
    }
    R &rt;
  };
} // namespace ppf

namespace ppf
{
  /// ppf_apply_operator's work on a context (shared by the allocation-free variant).
  inline void
  apply_operator_abi(Context &c, int block_id, int which, int scaling, int precision,
                     const void *d_src, void *d_dst)
  {
    auto     &rt  = c.rt;
    const int b   = rt.block_index(block_id);
    const int nbf = int(rt.blocks[b].field_indices.size());
    if (which < PPF_OP_RHS || which > PPF_OP_LHS_PLAIN || scaling < 0 || scaling > 2)
      throw Error(status_invalid, "bad operator selector or scaling");
    if (!d_dst || (which != PPF_OP_RHS && !d_src))
      throw Error(status_invalid, "null device array");
    auto run = [&](auto tag) {
      using T = decltype(tag);
      auto lf = typename Runtime<PPF_APP>::template LevelFields<T> {};
      std::vector<DeviceBuffer<T>> shadow;
      if constexpr (std::is_same_v<T, double>)
        lf = rt.fine_fields();
      else
        lf = rt.fp32_fields(b, which != PPF_OP_RHS, shadow); // FP32 copies of what it reads
      std::vector<const T *> src;
      std::vector<T *>       dst;
      for (int k = 0; k < nbf; ++k)
        {
          if (d_src)
            src.push_back(static_cast<const T *>(d_src) + k * rt.fine.n_nodes);
          dst.push_back(static_cast<T *>(d_dst) + k * rt.fine.n_nodes);
        }
      if (which == PPF_OP_RHS)
        src.clear();
      rt.template apply_operator<T>(b, which != PPF_OP_RHS, lf, src, dst, scaling,
                                    dirichlet_zero, which == PPF_OP_LHS_PLAIN);
      // FP32 copies in context-owned buffers must outlive the launch.
      if (c.synchronize || !shadow.empty())
        rt.sync();
    };
    if (precision == PPF_FP32)
      run(float {});
    else
      run(double {});
  }
} // namespace ppf

extern "C" {

const char *
ppf_app_name(void)
{
  return PPF_APP::name();
}

int
ppf_create(const char *prm_text, ppf_ctx **out)
{
  if (!prm_text || !out)
    {
      ppf::global_error() = "ppf_create: null argument";
      return PPF_ERR_INVALID;
    }
  *out = nullptr;
  try
    {
      *out = reinterpret_cast<ppf_ctx *>(new ppf::Context(prm_text));
      return PPF_OK;
    }
  catch (const ppf::Error &e)
    {
      ppf::global_error() = e.what();
      return e.code;
    }
  catch (const std::exception &e)
    {
      ppf::global_error() = e.what();
      return PPF_ERR_INVALID;
    }
}

void
ppf_destroy(ppf_ctx *ctx)
{
  delete reinterpret_cast<ppf::Context *>(ctx);
}

const char *
ppf_last_error(const ppf_ctx *ctx)
{
  if (!ctx)
    return ppf::global_error().c_str();
  return reinterpret_cast<const ppf::Context *>(ctx)->error.c_str();
}

int
ppf_get_info(ppf_ctx *ctx, ppf_info *info)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    if (!info)
      throw ppf::Error(ppf::status_invalid, "null info");
    const auto &g = c.rt.fine;
    std::memset(info, 0, sizeof(*info));
    info->dim    = g.dim;
    info->degree = g.degree;
    for (int d = 0; d < 3; ++d)
      {
        info->cells[d]    = g.cells[d];
        info->nodes[d]    = g.nodes[d];
        info->periodic[d] = g.periodic[d];
        info->lower[d]    = g.lower[d];
        info->h[d]        = g.h[d];
      }
    info->n_nodes      = g.n_nodes;
    info->n_dofs       = c.rt.full_size(g);
    info->n_fields     = int(c.rt.fields.size());
    info->n_blocks     = int(c.rt.blocks.size());
    info->dt           = c.rt.timer.get_timestep();
    info->time         = c.rt.timer.get_time();
    info->increment    = c.rt.timer.get_increment();
    info->n_increments = c.rt.inputs.temporal_discretization.n_increments;
  });
}

int
ppf_field_name(ppf_ctx *ctx, int field, char *buf, int len)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    if (field < 0 || field >= int(c.rt.fields.size()) || !buf || len <= 0)
      throw ppf::Error(ppf::status_invalid, "bad field or buffer");
    std::strncpy(buf, c.rt.fields[field].name.c_str(), size_t(len) - 1);
    buf[len - 1] = '\0';
  });
}

int
ppf_set_option(ppf_ctx *ctx, const char *key, const char *value)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const std::string k = key ? key : "", v = value ? value : "";
    if (k == "operator_path")
      {
        if (v == "auto")
          c.rt.path = ppf::OperatorPath::Auto;
        else if (v == "generic")
          c.rt.path = ppf::OperatorPath::Generic;
        else if (v == "split")
          c.rt.path = ppf::OperatorPath::Split;
        else if (v == "split_node")
          c.rt.path = ppf::OperatorPath::SplitNode;
        else
          throw ppf::Error(ppf::status_invalid,
                           "operator_path must be auto|generic|split|split_node");
      }
    else if (k == "stencil_rows")
      c.rt.stencil_rows = std::stoi(v);
    else if (k == "stencil_waves")
      c.rt.stencil_waves = std::stoi(v);
    else if (k == "stencil_z_chunk")
      c.rt.stencil_z_chunk = std::stoi(v);
    else if (k == "stencil_map")
      {
        if (v != "linear" && v != "band")
          throw ppf::Error(ppf::status_invalid, "stencil_map: linear or band");
        c.rt.stencil_map = v == "band" ? 1 : 0;
      }
    else if (k == "generic_kernel")
      {
        if (v != "cell_block" && v != "two_pass")
          throw ppf::Error(ppf::status_invalid, "generic_kernel: cell_block or two_pass");
        c.rt.generic_cell_block = v == "cell_block";
      }
    else if (k == "generic_column_chunk")
      c.rt.column_chunk = std::max(0, std::stoi(v));
    else if (k == "mg_graphs")
      {
        if (v != "0" && v != "1")
          throw ppf::Error(ppf::status_invalid, "mg_graphs: 0 or 1");
        c.rt.mg_graphs = v == "1";
      }
    else if (k == "stencil_min_nodes")
      c.rt.stencil_min_nodes = std::max<int64_t>(0, std::stoll(v));
    else if (k == "stencil_target_blocks")
      c.rt.stencil_target_blocks = std::max(1, std::stoi(v));
    else if (k == "stencil_store_policy")
      {
        if (v == "auto")
          c.rt.stencil_store_policy = ppf::StorePolicy::Auto;
        else if (v == "temporal")
          c.rt.stencil_store_policy = ppf::StorePolicy::Temporal;
        else if (v == "nontemporal")
          c.rt.stencil_store_policy = ppf::StorePolicy::NonTemporal;
        else
          throw ppf::Error(ppf::status_invalid,
                           "stencil_store_policy must be auto|temporal|nontemporal");
      }
    else if (k == "synchronize")
      c.synchronize = v != "0";
    else if (k == "mg_level_precision")
      {
        for (auto &b : c.rt.blocks)
          b.linear_solver_parameters.level_precision =
            (v == "single" || v == "float" || v == "fp32") ? prismspf::LevelPrecision::Single
                                                           : prismspf::LevelPrecision::Double;
      }
    else if (k == "mg_coarse_solver")
      {
        for (auto &b : c.rt.blocks)
          b.linear_solver_parameters.coarse_solver =
            v == "cg" ? prismspf::CoarseSolver::CG : prismspf::CoarseSolver::Smoother;
      }
    else if (k == "linear_tolerance")
      {
        for (auto &b : c.rt.blocks)
          b.linear_solver_parameters.tolerance = std::stod(v);
      }
    else if (k == "linear_max_iterations")
      {
        for (auto &b : c.rt.blocks)
          b.linear_solver_parameters.max_iterations = unsigned(std::stoul(v));
      }
    else if (k == "pipeline_depth")
      {
        const int d = std::stoi(v);
        if (d < 1 || d > ppf::max_pipeline_depth)
          throw ppf::Error(ppf::status_invalid, "pipeline_depth must be 1, 2 or 3");
        c.rt.pipeline_depth = d;
      }
    else if (k == "explicit_pairs")
      {
        if (v != "0" && v != "1")
          throw ppf::Error(ppf::status_invalid, "explicit_pairs must be 0|1");
        c.rt.explicit_pairs = v == "1";
      }
    else if (k == "pair_rows")
      {
        if (v != "2" && v != "3")
          throw ppf::Error(ppf::status_invalid, "pair_rows must be 2|3");
        c.rt.pair_rows = std::stoi(v);
      }
    else if (k == "pair_z_chunk")
      c.rt.pair_z_chunk = std::max(0, std::stoi(v));
    else if (k == "explicit_precision")
      {
        if (v != "double" && v != "single" && v != "single_copies")
          throw ppf::Error(ppf::status_invalid,
                           "explicit_precision must be double|single|single_copies");
        c.rt.explicit_fp32        = v != "double";
        c.rt.explicit_fp32_copies = v == "single_copies";
      }
    else if (k == "write_output")
      {
        if (v != "0" && v != "1")
          throw ppf::Error(ppf::status_invalid, "write_output must be 0|1");
        c.rt.write_output = v == "1";
      }
    else if (k == "linear_initial_guess")
      {
        if (v != "previous" && v != "zero")
          throw ppf::Error(ppf::status_invalid, "linear_initial_guess must be previous|zero");
        c.rt.zero_initial_guess = v == "zero";
      }
    else
      throw ppf::Error(ppf::status_invalid, "unknown option " + k);
    // captured V-cycles carry the kernels' launch settings
    if (k.rfind("stencil_", 0) == 0 || k.rfind("generic_", 0) == 0 || k == "operator_path" ||
        k == "mg_coarse_solver" || k == "mg_level_precision" || k == "mg_graphs")
      ++c.rt.graph_epoch;
  });
}

int
ppf_set_field(ppf_ctx *ctx, int field, int age, const double *host, long long n)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) { c.rt.set_field(field, age, host, n); });
}

int
ppf_get_field(ppf_ctx *ctx, int field, int age, double *host, long long n)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) { c.rt.get_field(field, age, host, n); });
}

int
ppf_field_device_ptr(ppf_ctx *ctx, int field, int age, void **ptr)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) { *ptr = c.rt.field_ptr(field, age); });
}

int
ppf_run(ppf_ctx *ctx, unsigned int n_increments, int *exit_status)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const int s = c.rt.run(n_increments);
    if (exit_status)
      *exit_status = s;
  });
}

int
ppf_solve_increment(ppf_ctx *ctx, int *exit_status)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const int s = c.rt.solve_increment();
    c.rt.timer.increment();
    c.rt.sync();
    if (exit_status)
      *exit_status = s;
  });
}

int
ppf_solve_block(ppf_ctx *ctx, int block_id)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    c.rt.solve_block(c.rt.block_index(block_id));
    c.rt.sync();
  });
}

int
ppf_rotate_old(ppf_ctx *ctx, int block_id)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) { c.rt.rotate(c.rt.block_index(block_id)); });
}

int
ppf_set_time(ppf_ctx *ctx, unsigned int increment, double time)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    c.rt.timer.set_increment(increment);
    c.rt.timer.set_time(time);
  });
}

int
ppf_apply_operator(ppf_ctx *ctx, int block_id, int which, int scaling, int precision,
                   const void *d_src, void *d_dst)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    ppf::apply_operator_abi(c, block_id, which, scaling, precision, d_src, d_dst);
  });
}

int
ppf_apply_operator_noalloc(ppf_ctx *ctx, int block_id, int which, int scaling, int precision,
                           const void *d_src, void *d_dst)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    if (!c.scratch)
      throw ppf::Error(ppf::status_internal, "device scratch not set up");
    ppf::BorrowScope<ppf::Runtime<PPF_APP>> scope(c.rt, c.scratch,
                                                  int64_t(ppf::noalloc_scratch_bytes));
    ppf::apply_operator_abi(c, block_id, which, scaling, precision, d_src, d_dst);
  });
}

int
ppf_inverse_diagonal(ppf_ctx *ctx, int block_id, int scaling, int precision, void *d_dst)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    auto     &rt = c.rt;
    const int b  = rt.block_index(block_id);
    if (precision == PPF_FP32)
      {
        std::vector<ppf::DeviceBuffer<float>> shadow;
        const auto lf = rt.fp32_fields(b, true, shadow);
        rt.template inverse_diagonal<float>(b, lf, scaling, static_cast<float *>(d_dst));
      }
    else
      rt.template inverse_diagonal<double>(b, rt.fine_fields(), scaling,
                                           static_cast<double *>(d_dst));
    rt.sync();
  });
}

int
ppf_linear_solve(ppf_ctx *ctx, int block_id, ppf_solve_stats *stats)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const auto st = c.rt.solve_linear(c.rt.block_index(block_id));
    if (stats)
      *stats = {st.iterations, st.initial_residual, st.final_residual, st.tolerance,
                st.converged};
  });
}

int
ppf_last_solve_stats(ppf_ctx *ctx, int block_id, ppf_solve_stats *stats)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const auto &st = c.rt.linear_state(c.rt.block_index(block_id)).stats;
    if (stats)
      *stats = {st.iterations, st.initial_residual, st.final_residual, st.tolerance,
                st.converged};
  });
}

int
ppf_last_newton_stats(ppf_ctx *ctx, int block_id, ppf_newton_stats *stats)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const auto &st = c.rt.linear_state(c.rt.block_index(block_id)).newton;
    if (stats)
      *stats = {st.iterations, st.linear_iterations, st.residual, st.tolerance, st.converged};
  });
}

int
ppf_field_norms(ppf_ctx *ctx, int field, int age, double *l2, double *integral)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    const auto r = c.rt.norms(field, age);
    if (l2)
      *l2 = r.first;
    if (integral)
      *integral = r.second;
  });
}

const char *
ppf_output_log(ppf_ctx *ctx)
{
  if (!ctx)
    return "";
  return reinterpret_cast<ppf::Context *>(ctx)->rt.log.c_str();
}

void *
ppf_stream(ppf_ctx *ctx)
{
  if (!ctx)
    return nullptr;
  return reinterpret_cast<ppf::Context *>(ctx)->rt.stream;
}

int
ppf_synchronize(ppf_ctx *ctx)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) { c.rt.sync(); });
}

int
ppf_adjoint_forward(ppf_ctx *ctx, const double *d_phi0, double L, double kappa,
                    unsigned int n_steps, int precision, double *d_phi_final)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    ppf::adjoint_forward<PPF_APP>(c, d_phi0, L, kappa, n_steps, precision, d_phi_final);
  });
}

int
ppf_adjoint_vjp(ppf_ctx *ctx, const double *d_phi0, const double *d_lam_final, double L,
                double kappa, unsigned int n_steps, unsigned int checkpoint_every,
                int precision, double *dL, double *dkappa, double *d_dphi0)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    double gl = 0.0, gk = 0.0;
    ppf::adjoint_vjp<PPF_APP>(c, d_phi0, d_lam_final, L, kappa, n_steps, checkpoint_every,
                              precision, gl, gk, d_dphi0);
    if (dL)
      *dL = gl;
    if (dkappa)
      *dkappa = gk;
  });
}

int
ppf_lumped_mass(ppf_ctx *ctx, double *d_mass)
{
  return ppf::guarded(ctx, [&](ppf::Context &c) {
    auto &rt = c.rt;
    hipLaunchKernelGGL((ppf::lumped_mass_kernel<PPF_APP::dim, PPF_APP::degree>),
                       dim3(ppf::grid_for(rt.fine.n_nodes, 256)), dim3(256), 0, rt.stream, d_mass,
                       rt.fine);
    rt.sync();
  });
}

} // extern "C"
