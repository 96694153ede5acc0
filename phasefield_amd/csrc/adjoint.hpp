// Adjoint of the explicit Allen-Cahn step.
//
// The reference has no adjoint.  The forward map is the explicit Allen-Cahn update of
// an explicit solve block (reference applications/allen_cahn/explicit/custom_pde.h:72-87,
// evaluated by ExplicitSolver, explicit_solver.h:63-77): with GLL collocation and the
// lumped mass M, for mobility L and gradient coefficient kappa
//
//     phi' = F(phi) = phi - dt L f'(phi) - dt L kappa M^{-1} K phi,
//
// K the assembled stiffness matrix.  Its transpose-Jacobian is
//
//     dF/dphi^T lam = lam - dt L f''(phi) lam - dt L kappa K (M^{-1} lam),
//
// and the parameter derivatives are dF/dL = -dt f'(phi) - dt kappa M^{-1} K phi and
// dF/dkappa = -dt L M^{-1} K phi.  A vector-Jacobian product over n steps runs the
// steps backwards, recomputing the forward states of each checkpoint segment.
#pragma once

#include "device.hpp"
#include "grid.hpp"
#include "operator_kernels.hpp"
#include "vector_kernels.hpp"

namespace ppf
{
  /// One forward step in arithmetic T; state stored in FP64.
  template <int dim, int degree, typename T, class Energy>
  __global__ void __launch_bounds__(block_size)
  ac_forward_kernel(double *__restrict__ out, const double *__restrict__ phi, const GridDesc g,
                    double dt, double L, double kappa)
  {
    const int64_t node = int64_t(blockIdx.x) * block_size + threadIdx.x;
    if (node >= g.n_nodes)
      return;
    int i[3];
    decompose<dim>(g, node, g.nodes, i);
    const T      p   = T(phi[node]);
    const T      kp  = T(stiffness_row<dim, degree, double>(g, i, phi, false, 0u));
    const double m   = lumped_mass<dim, degree>(g, i);
    const T      upd = p - T(dt * L) * Energy::template fprime<T>(p) - T(dt * L * kappa / m) * kp;
    out[node]        = double(upd);
  }

  /// One adjoint step: lam_prev = dF/dphi(phi)^T lam, plus per-block partial sums of
  /// lam . dF/dL and lam . dF/dkappa.  `lam_over_m` holds lam / m.
  template <int dim, int degree, class Energy>
  __global__ void __launch_bounds__(block_size)
  ac_adjoint_kernel(double *__restrict__ lam_prev, double *__restrict__ partial,
                    const double *__restrict__ lam, const double *__restrict__ lam_over_m,
                    const double *__restrict__ phi, const GridDesc g, double dt, double L,
                    double kappa)
  {
    const int64_t node = int64_t(blockIdx.x) * block_size + threadIdx.x;
    double        gL = 0.0, gk = 0.0;
    if (node < g.n_nodes)
      {
        int i[3];
        decompose<dim>(g, node, g.nodes, i);
        const double p   = phi[node];
        const double l   = lam[node];
        const double m   = lumped_mass<dim, degree>(g, i);
        const double kp  = stiffness_row<dim, degree, double>(g, i, phi, false, 0u);
        const double klm = stiffness_row<dim, degree, double>(g, i, lam_over_m, false, 0u);
        lam_prev[node]   = l - dt * L * Energy::template fprime2<double>(p) * l - dt * L * kappa * klm;
        gL               = l * (-dt * Energy::template fprime<double>(p) - dt * kappa * kp / m);
        gk               = l * (-dt * L * kp / m);
      }
    gL = block_sum(gL);
    __syncthreads();
    gk = block_sum(gk);
    if (threadIdx.x == 0)
      {
        partial[blockIdx.x]             = gL;
        partial[gridDim.x + blockIdx.x] = gk;
      }
  }

  template <int dim, int degree>
  __global__ void
  divide_by_mass_kernel(double *__restrict__ out, const double *__restrict__ in, const GridDesc g)
  {
    const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= g.n_nodes)
      return;
    int i[3];
    decompose<dim>(g, node, g.nodes, i);
    out[node] = in[node] / lumped_mass<dim, degree>(g, i);
  }

  template <int dim, int degree>
  __global__ void
  lumped_mass_kernel(double *__restrict__ out, const GridDesc g)
  {
    const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= g.n_nodes)
      return;
    int i[3];
    decompose<dim>(g, node, g.nodes, i);
    out[node] = lumped_mass<dim, degree>(g, i);
  }

  /// Sums `count` partials in fixed order, for two interleaved sums.
  __global__ void __launch_bounds__(reduce_threads)
  reduce_two_kernel(double *__restrict__ out, const double *__restrict__ partial, int count)
  {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x)
      {
        a += partial[i];
        b += partial[count + i];
      }
    a = block_sum(a);
    __syncthreads();
    b = block_sum(b);
    if (threadIdx.x == 0)
      {
        out[0] += a;
        out[1] += b;
      }
  }

  /// Forward runs and vector-Jacobian products of the Allen-Cahn step on one grid.
  template <int dim, int degree, class Energy>
  class AllenCahnAdjoint
  {
  public:
    AllenCahnAdjoint(const GridDesc &g_, hipStream_t s_)
      : g(g_)
      , stream(s_)
    {}

    void
    step(double *out, const double *phi, double dt, double L, double kappa, int precision)
    {
      const unsigned int nb = grid_for(g.n_nodes, block_size);
      if (precision == 1)
        hipLaunchKernelGGL((ac_forward_kernel<dim, degree, float, Energy>), dim3(nb),
                           dim3(block_size), 0, stream, out, phi, g, dt, L, kappa);
      else
        hipLaunchKernelGGL((ac_forward_kernel<dim, degree, double, Energy>), dim3(nb),
                           dim3(block_size), 0, stream, out, phi, g, dt, L, kappa);
    }

    /// phi_final = F^n(phi0).
    void
    forward(const double *phi0, double *phi_final, double dt, double L, double kappa,
            unsigned int n, int precision)
    {
      ensure(2);
      const double *cur = phi0;
      for (unsigned int s = 0; s < n; ++s)
        {
          double *next = (s + 1 == n) ? phi_final : work[s % 2].get();
          step(next, cur, dt, L, kappa, precision);
          cur = next;
        }
      if (n == 0)
        PPF_HIP_CHECK(hipMemcpyAsync(phi_final, phi0, sizeof(double) * g.n_nodes,
                                     hipMemcpyDeviceToDevice, stream));
      PPF_HIP_CHECK(hipGetLastError());
    }

    /// Given lam_N = dJ/dphi_N, returns dJ/dL, dJ/dkappa and dJ/dphi0 (into dphi0).
    /// Forward states are recomputed from checkpoints every `every` steps.
    void
    vjp(const double *phi0, const double *lam_final, double dt, double L, double kappa,
        unsigned int n, unsigned int every, int precision, double &dL, double &dkappa,
        double *dphi0)
    {
      if (n == 0)
        {
          PPF_HIP_CHECK(hipMemcpyAsync(dphi0, lam_final, sizeof(double) * g.n_nodes,
                                       hipMemcpyDeviceToDevice, stream));
          PPF_HIP_CHECK(hipStreamSynchronize(stream));
          dL = dkappa = 0.0;
          return;
        }
      if (every == 0)
        every = n;
      const unsigned int n_ckpt = (n + every - 1) / every;
      checkpoints.resize(n_ckpt);
      for (auto &c : checkpoints)
        c.resize(g.n_nodes);
      segment.resize(every);
      for (auto &s : segment)
        s.resize(g.n_nodes);
      ensure(3);
      partial.resize(2 * reduce_blocks_for());
      sums.resize(2);
      PPF_HIP_CHECK(hipMemsetAsync(sums.get(), 0, 2 * sizeof(double), stream));

      // Forward sweep storing the checkpoints phi_{k*every}, k = 0 .. n_ckpt-1.
      PPF_HIP_CHECK(hipMemcpyAsync(checkpoints[0].get(), phi0, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
      const double      *cur       = checkpoints[0].get();
      const unsigned int last_ckpt = (n_ckpt - 1) * every;
      for (unsigned int s = 0; s < last_ckpt; ++s)
        {
          double *next = (s + 1) % every == 0 ? checkpoints[(s + 1) / every].get()
                                              : work[s % 2].get();
          step(next, cur, dt, L, kappa, precision);
          cur = next;
        }

      // Reverse sweep.
      double *lam  = work[0].get();
      double *lamp = work[1].get();
      double *lom  = work[2].get();
      PPF_HIP_CHECK(hipMemcpyAsync(lam, lam_final, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
      const unsigned int nb = grid_for(g.n_nodes, block_size);
      for (int seg = int(n_ckpt) - 1; seg >= 0; --seg)
        {
          const unsigned int first = unsigned(seg) * every;
          const unsigned int last  = std::min(n, first + every); // exclusive
          // Recompute phi_first .. phi_{last-1} into the segment buffer.
          PPF_HIP_CHECK(hipMemcpyAsync(segment[0].get(), checkpoints[seg].get(),
                                       sizeof(double) * g.n_nodes, hipMemcpyDeviceToDevice,
                                       stream));
          for (unsigned int s = first + 1; s < last; ++s)
            step(segment[s - first].get(), segment[s - first - 1].get(), dt, L, kappa,
                 precision);
          for (int s = int(last) - 1; s >= int(first); --s)
            {
              hipLaunchKernelGGL((divide_by_mass_kernel<dim, degree>), dim3(nb),
                                 dim3(block_size), 0, stream, lom, lam, g);
              hipLaunchKernelGGL((ac_adjoint_kernel<dim, degree, Energy>), dim3(nb),
                                 dim3(block_size), 0, stream, lamp, partial.get(), lam, lom,
                                 segment[s - first].get(), g, dt, L, kappa);
              hipLaunchKernelGGL(reduce_two_kernel, dim3(1), dim3(reduce_threads), 0, stream,
                                 sums.get(), partial.get(), int(nb));
              std::swap(lam, lamp);
            }
        }
      PPF_HIP_CHECK(hipMemcpyAsync(dphi0, lam, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
      double host[2];
      PPF_HIP_CHECK(hipMemcpyAsync(host, sums.get(), 2 * sizeof(double), hipMemcpyDeviceToHost,
                                   stream));
      PPF_HIP_CHECK(hipStreamSynchronize(stream));
      dL     = host[0];
      dkappa = host[1];
    }

  private:
    int
    reduce_blocks_for() const
    {
      return int(grid_for(g.n_nodes, block_size));
    }

    void
    ensure(int k)
    {
      if (int(work.size()) < k)
        work.resize(k);
      for (auto &w : work)
        if (w.size() != g.n_nodes)
          w.resize(g.n_nodes);
    }

    GridDesc                          g;
    hipStream_t                       stream;
    std::vector<DeviceBuffer<double>> work, checkpoints, segment;
    DeviceBuffer<double>              partial, sums;
  };
} // namespace ppf
