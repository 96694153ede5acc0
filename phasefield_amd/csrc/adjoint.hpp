// Adjoints of the explicit Allen-Cahn step and of the explicit Cahn-Hilliard step.
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
#include "stencil_kernels.hpp"

namespace ppf
{
  /// Lattice of the column kernels below.  With GLL collocation the assembled
  /// stiffness row of a node is its lumped mass times L u = sum_d sum_o W_d(i_d, o)
  /// u_{i + o e_d} (the interior taps W of the stencil kernel, mirror ghosts at
  /// non-periodic faces), so M^{-1} K phi = L phi and K (M^{-1} lam) = M L (lam / M).
  struct AdjointLattice
  {
    int    n[3];
    int    periodic[3];
    int    plane;   ///< n[0] * n[1]
    int    z_chunk; ///< planes per column segment (one workgroup row of the grid)
    double tap[3][2][3];
    double mass[3][3];     ///< 1-D lumped mass per scale class
    double inv_mass[3][3]; ///< its inverse
  };

  template <int dim, int degree>
  inline AdjointLattice
  make_adjoint_lattice(const GridDesc &g)
  {
    AdjointLattice a {};
    const double tq2[2][3] = {{14.0, -8.0, 1.0}, {8.0, -4.0, 0.0}};
    const double tq1[3]    = {2.0, -1.0, 0.0};
    for (int d = 0; d < 3; ++d)
      {
        a.n[d]        = d < dim ? g.nodes[d] : 1;
        a.periodic[d] = d < dim ? g.periodic[d] : 0;
        for (int par = 0; par < 2; ++par)
          for (int k = 0; k < 3; ++k)
            a.tap[d][par][k] =
              d < dim ? (degree == 2 ? tq2[par][k] : tq1[k]) / (g.h[d] * g.h[d]) : 0.0;
        for (int c = 0; c < 3; ++c)
          {
            a.mass[d][c]     = d < dim ? class_mass_units<degree>(c) * g.h[d] : 1.0;
            a.inv_mass[d][c] = 1.0 / a.mass[d][c];
          }
      }
    a.plane = a.n[0] * a.n[1];
    // Enough column segments for ~2048 workgroups (the GPU holds 256 CUs).
    const int blocks_x = (a.plane + block_size - 1) / block_size;
    const int chunks   = std::max(1, std::min(a.n[2], (2048 + blocks_x - 1) / blocks_x));
    a.z_chunk          = (a.n[2] + chunks - 1) / chunks;
    return a;
  }

  /// Per-column constants: the in-plane offsets of the x and y neighbours (source nodes
  /// after wrap / mirror), their taps by the column's parity, and 1-D inverse masses.
  template <int degree, typename T>
  struct AdjointColumn
  {
    static constexpr int RAD = degree;
    int                  xoff[2 * RAD + 1], yoff[2 * RAD + 1];
    T                    tx[RAD + 1], ty[RAD + 1];
    double               ixm[2 * RAD + 1], iym[2 * RAD + 1];
    double               mx, my;

    __device__ PPF_INLINE
    AdjointColumn(const AdjointLattice &a, int x, int y)
    {
#pragma unroll
      for (int k = 0; k <= 2 * RAD; ++k)
        {
          const int xs = source_coord(x + k - RAD, a.n[0], a.periodic[0]);
          const int ys = source_coord(y + k - RAD, a.n[1], a.periodic[1]);
          xoff[k]      = y * a.n[0] + xs;
          yoff[k]      = ys * a.n[0] + x;
          ixm[k]       = a.inv_mass[0][scale_class<degree>(xs, a.n[0], a.periodic[0])];
          iym[k]       = a.inv_mass[1][scale_class<degree>(ys, a.n[1], a.periodic[1])];
        }
#pragma unroll
      for (int o = 0; o <= RAD; ++o)
        {
          tx[o] = T(a.tap[0][degree == 2 && (x & 1)][o]);
          ty[o] = T(a.tap[1][degree == 2 && (y & 1)][o]);
        }
      mx = a.mass[0][scale_class<degree>(x, a.n[0], a.periodic[0])];
      my = a.mass[1][scale_class<degree>(y, a.n[1], a.periodic[1])];
    }
  };

  /// One forward step in arithmetic T (state stored in FP64):
  /// phi' = phi - dt L f'(phi) - dt L kappa L phi.  A thread owns one (x, y) column
  /// of a z segment; z neighbours come from a register ring, x and y neighbours from
  /// the plane (cached: the neighbouring lanes load them as their centres).
  template <int dim, int degree, typename T, class Energy>
  __global__ void __launch_bounds__(block_size)
  ac_forward_kernel(double *__restrict__ out, const double *__restrict__ phi,
                    const AdjointLattice a, double dt, double L, double kappa)
  {
    constexpr int RAD = degree;
    const int     q   = int(blockIdx.x) * block_size + int(threadIdx.x);
    if (q >= a.plane)
      return;
    const int                   x = q % a.n[0], y = q / a.n[0];
    const AdjointColumn<degree, T> col(a, x, y);
    const int z0 = dim == 3 ? int(blockIdx.y) * a.z_chunk : 0;
    const int z1 = dim == 3 ? min(z0 + a.z_chunk, a.n[2]) : 1;
    const T   c1 = T(dt * L), c2 = T(dt * L * kappa);

    T ring[2 * RAD + 1];
    if constexpr (dim == 3)
#pragma unroll
      for (int k = 0; k < 2 * RAD; ++k)
        ring[k + 1] = T(phi[int64_t(source_coord(z0 + k - RAD, a.n[2], a.periodic[2])) * a.plane + q]);
    for (int z = z0; z < z1; ++z)
      {
        const double *pl = phi + int64_t(z) * a.plane;
        T             u;
        T             lap;
        if constexpr (dim == 3)
          {
#pragma unroll
            for (int k = 0; k < 2 * RAD; ++k)
              ring[k] = ring[k + 1];
            ring[2 * RAD] =
              T(phi[int64_t(source_coord(z + RAD, a.n[2], a.periodic[2])) * a.plane + q]);
            u                = ring[RAD];
            const double *tz = a.tap[2][degree == 2 && (z & 1)];
            lap              = (col.tx[0] + col.ty[0] + T(tz[0])) * u;
#pragma unroll
            for (int o = 1; o <= RAD; ++o)
              lap += T(tz[o]) * (ring[RAD - o] + ring[RAD + o]);
          }
        else
          {
            u   = T(pl[q]);
            lap = (col.tx[0] + col.ty[0]) * u;
          }
#pragma unroll
        for (int o = 1; o <= RAD; ++o)
          {
            lap += col.tx[o] * (T(pl[col.xoff[RAD - o]]) + T(pl[col.xoff[RAD + o]]));
            lap += col.ty[o] * (T(pl[col.yoff[RAD - o]]) + T(pl[col.yoff[RAD + o]]));
          }
        out[int64_t(z) * a.plane + q] = double(u - c1 * Energy::template fprime<T>(u) - c2 * lap);
      }
  }

  /// One adjoint step, lam_prev = dF/dphi(phi)^T lam
  ///   = lam - dt L f''(phi) lam - dt L kappa M L(lam / M),
  /// with lam . dF/dL and lam . dF/dkappa (dF/dL = -dt f'(phi) - dt kappa L phi,
  /// dF/dkappa = -dt L L phi) added to per-workgroup sums (partial[b], partial[count + b]),
  /// which the host reduces once after the sweep.  Same column traversal as the forward
  /// step; lam / M is formed as lam is loaded.
  template <int dim, int degree, class Energy>
  __global__ void __launch_bounds__(block_size)
  ac_adjoint_kernel(double *__restrict__ lam_prev, double *__restrict__ partial,
                    const double *__restrict__ lam, const double *__restrict__ phi,
                    const AdjointLattice a, double dt, double L, double kappa)
  {
    constexpr int RAD = degree;
    const int     q   = int(blockIdx.x) * block_size + int(threadIdx.x);
    double        gL = 0.0, gk = 0.0;
    if (q < a.plane)
      {
        const int x = q % a.n[0], y = q / a.n[0];
        const AdjointColumn<degree, double> col(a, x, y);
        const int z0 = dim == 3 ? int(blockIdx.y) * a.z_chunk : 0;
        const int z1 = dim == 3 ? min(z0 + a.z_chunk, a.n[2]) : 1;
        const double ixy = col.ixm[RAD] * col.iym[RAD];

        // Rings of phi and lam / M along z (lam / M uses the mass of the source node).
        double rp[2 * RAD + 1], rl[2 * RAD + 1];
        auto   load = [&](int k, int zz) {
          const int zs = source_coord(zz, a.n[2], a.periodic[2]);
          const int64_t i = int64_t(zs) * a.plane + q;
          rp[k] = phi[i];
          rl[k] = lam[i] * (ixy * a.inv_mass[2][scale_class<degree>(zs, a.n[2], a.periodic[2])]);
        };
        if constexpr (dim == 3)
#pragma unroll
          for (int k = 0; k < 2 * RAD; ++k)
            load(k + 1, z0 + k - RAD);
        for (int z = z0; z < z1; ++z)
          {
            const int64_t base = int64_t(z) * a.plane;
            const double *pp = phi + base, *pl = lam + base;
            double        p, lom, lp, ll, mz;
            if constexpr (dim == 3)
              {
#pragma unroll
                for (int k = 0; k < 2 * RAD; ++k)
                  {
                    rp[k] = rp[k + 1];
                    rl[k] = rl[k + 1];
                  }
                load(2 * RAD, z + RAD);
                p                = rp[RAD];
                lom              = rl[RAD];
                const double *tz = a.tap[2][degree == 2 && (z & 1)];
                lp               = (col.tx[0] + col.ty[0] + tz[0]) * p;
                ll               = (col.tx[0] + col.ty[0] + tz[0]) * lom;
#pragma unroll
                for (int o = 1; o <= RAD; ++o)
                  {
                    lp += tz[o] * (rp[RAD - o] + rp[RAD + o]);
                    ll += tz[o] * (rl[RAD - o] + rl[RAD + o]);
                  }
                mz = a.mass[2][scale_class<degree>(z, a.n[2], a.periodic[2])];
              }
            else
              {
                p   = pp[q];
                lom = pl[q] * ixy;
                lp  = (col.tx[0] + col.ty[0]) * p;
                ll  = (col.tx[0] + col.ty[0]) * lom;
                mz  = 1.0;
              }
            const double imz = 1.0 / mz;
            double       sx = 0.0, sy = 0.0;
#pragma unroll
            for (int o = 1; o <= RAD; ++o)
              {
                lp += col.tx[o] * (pp[col.xoff[RAD - o]] + pp[col.xoff[RAD + o]]);
                lp += col.ty[o] * (pp[col.yoff[RAD - o]] + pp[col.yoff[RAD + o]]);
                sx += col.tx[o] * (pl[col.xoff[RAD - o]] * col.ixm[RAD - o] +
                                   pl[col.xoff[RAD + o]] * col.ixm[RAD + o]);
                sy += col.ty[o] * (pl[col.yoff[RAD - o]] * col.iym[RAD - o] +
                                   pl[col.yoff[RAD + o]] * col.iym[RAD + o]);
              }
            ll += imz * (col.iym[RAD] * sx + col.ixm[RAD] * sy);
            const double m = col.mx * col.my * mz;
            const double l = pl[q];
            lam_prev[base + q] =
              l - dt * L * Energy::template fprime2<double>(p) * l - dt * L * kappa * m * ll;
            gL += l * (-dt * Energy::template fprime<double>(p) - dt * kappa * lp);
            gk += l * (-dt * L * lp);
          }
      }
    gL = block_sum(gL);
    __syncthreads();
    gk = block_sum(gk);
    if (threadIdx.x == 0)
      {
        const int count = int(gridDim.x * gridDim.y);
        const int b     = int(blockIdx.x + gridDim.x * blockIdx.y);
        partial[b] += gL;
        partial[count + b] += gk;
      }
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
        out[0] = a;
        out[1] = b;
      }
  }

  // ------------------------------------------------------------------------------
  // Cahn-Hilliard step
  // ------------------------------------------------------------------------------
  //
  // The explicit Cahn-Hilliard pair of blocks (reference applications/cahn_hilliard/
  // explicit/custom_pde.h:92-118: c block value c_old, gradient -dt M grad(mu_old); mu
  // block value f'(c), gradient kappa grad(c)) is, with L = M^{-1} K as above,
  //
  //     c' = c - dt M L mu,        mu' = g(c') = f'(c') + kappa L c',
  //
  // with mu_0 = g(c_0) (increment 0 computes mu from the initial condition).  So
  // c_{n+1} = F(c_n) = c_n - dt M L g(c_n) and, with L^T lam = K (M^{-1} lam) =
  // M L(lam / M) =: w,
  //
  //     dF/dc^T lam = lam - dt M (f''(c) w + kappa L^T w),
  //     lam . dF/dM = -dt lam . L mu,      lam . dF/dkappa = -dt M w . L c.

  /// One forward stage of the Cahn-Hilliard step in arithmetic T (state in FP64):
  /// STAGE 0: out = c - k L s with s = mu, k = dt M (the c block);
  /// STAGE 1: out = f'(s) + k L s with s = c', k = kappa (the mu block).
  template <int dim, int degree, typename T, class Energy, int STAGE>
  __global__ void __launch_bounds__(block_size)
  ch_forward_kernel(double *__restrict__ out, const double *__restrict__ c,
                    const double *__restrict__ s, const AdjointLattice a, double k)
  {
    constexpr int RAD = degree;
    const int     q   = int(blockIdx.x) * block_size + int(threadIdx.x);
    if (q >= a.plane)
      return;
    const int                      x = q % a.n[0], y = q / a.n[0];
    const AdjointColumn<degree, T> col(a, x, y);
    const int z0 = dim == 3 ? int(blockIdx.y) * a.z_chunk : 0;
    const int z1 = dim == 3 ? min(z0 + a.z_chunk, a.n[2]) : 1;
    const T   kt = T(k);

    T ring[2 * RAD + 1];
    if constexpr (dim == 3)
#pragma unroll
      for (int j = 0; j < 2 * RAD; ++j)
        ring[j + 1] = T(s[int64_t(source_coord(z0 + j - RAD, a.n[2], a.periodic[2])) * a.plane + q]);
    for (int z = z0; z < z1; ++z)
      {
        const double *pl = s + int64_t(z) * a.plane;
        T             u, lap;
        if constexpr (dim == 3)
          {
#pragma unroll
            for (int j = 0; j < 2 * RAD; ++j)
              ring[j] = ring[j + 1];
            ring[2 * RAD] =
              T(s[int64_t(source_coord(z + RAD, a.n[2], a.periodic[2])) * a.plane + q]);
            u                = ring[RAD];
            const double *tz = a.tap[2][degree == 2 && (z & 1)];
            lap              = (col.tx[0] + col.ty[0] + T(tz[0])) * u;
#pragma unroll
            for (int o = 1; o <= RAD; ++o)
              lap += T(tz[o]) * (ring[RAD - o] + ring[RAD + o]);
          }
        else
          {
            u   = T(pl[q]);
            lap = (col.tx[0] + col.ty[0]) * u;
          }
#pragma unroll
        for (int o = 1; o <= RAD; ++o)
          {
            lap += col.tx[o] * (T(pl[col.xoff[RAD - o]]) + T(pl[col.xoff[RAD + o]]));
            lap += col.ty[o] * (T(pl[col.yoff[RAD - o]]) + T(pl[col.yoff[RAD + o]]));
          }
        const int64_t i = int64_t(z) * a.plane + q;
        if constexpr (STAGE == 0)
          out[i] = double(T(c[i]) - kt * lap);
        else
          out[i] = double(Energy::template fprime<T>(u) + kt * lap);
      }
  }

  /// One pass of the Cahn-Hilliard adjoint step at state (c_n, mu_n):
  /// STAGE 0 (l = lam_{n+1}, p = mu_n): out = w = M L(lam / M), and -dt lam . L mu
  ///   is added to partial[b];
  /// STAGE 1 (l = w, p = c_n, v = lam_{n+1}): out = lam_n = v - dt M (f''(c) w +
  ///   kappa M L(w / M)), and -dt M w . L c is added to partial[count + b].
  /// Same column traversal as ac_adjoint_kernel; l / M is formed as l is loaded.
  template <int dim, int degree, class Energy, int STAGE>
  __global__ void __launch_bounds__(block_size)
  ch_adjoint_kernel(double *__restrict__ out, double *__restrict__ partial,
                    const double *__restrict__ l_in, const double *__restrict__ p_in,
                    const double *__restrict__ v_in, const AdjointLattice a, double dt,
                    double mob, double kappa)
  {
    constexpr int RAD = degree;
    const int     q   = int(blockIdx.x) * block_size + int(threadIdx.x);
    double        g   = 0.0;
    if (q < a.plane)
      {
        const int x = q % a.n[0], y = q / a.n[0];
        const AdjointColumn<degree, double> col(a, x, y);
        const int    z0  = dim == 3 ? int(blockIdx.y) * a.z_chunk : 0;
        const int    z1  = dim == 3 ? min(z0 + a.z_chunk, a.n[2]) : 1;
        const double ixy = col.ixm[RAD] * col.iym[RAD];

        double rp[2 * RAD + 1], rl[2 * RAD + 1];
        auto   load = [&](int j, int zz) {
          const int     zs = source_coord(zz, a.n[2], a.periodic[2]);
          const int64_t i  = int64_t(zs) * a.plane + q;
          rp[j] = p_in[i];
          rl[j] = l_in[i] * (ixy * a.inv_mass[2][scale_class<degree>(zs, a.n[2], a.periodic[2])]);
        };
        if constexpr (dim == 3)
#pragma unroll
          for (int j = 0; j < 2 * RAD; ++j)
            load(j + 1, z0 + j - RAD);
        for (int z = z0; z < z1; ++z)
          {
            const int64_t base = int64_t(z) * a.plane;
            const double *pp = p_in + base, *pl = l_in + base;
            double        pv, lom, lp, ll, mz;
            if constexpr (dim == 3)
              {
#pragma unroll
                for (int j = 0; j < 2 * RAD; ++j)
                  {
                    rp[j] = rp[j + 1];
                    rl[j] = rl[j + 1];
                  }
                load(2 * RAD, z + RAD);
                pv               = rp[RAD];
                lom              = rl[RAD];
                const double *tz = a.tap[2][degree == 2 && (z & 1)];
                lp               = (col.tx[0] + col.ty[0] + tz[0]) * pv;
                ll               = (col.tx[0] + col.ty[0] + tz[0]) * lom;
#pragma unroll
                for (int o = 1; o <= RAD; ++o)
                  {
                    lp += tz[o] * (rp[RAD - o] + rp[RAD + o]);
                    ll += tz[o] * (rl[RAD - o] + rl[RAD + o]);
                  }
                mz = a.mass[2][scale_class<degree>(z, a.n[2], a.periodic[2])];
              }
            else
              {
                pv  = pp[q];
                lom = pl[q] * ixy;
                lp  = (col.tx[0] + col.ty[0]) * pv;
                ll  = (col.tx[0] + col.ty[0]) * lom;
                mz  = 1.0;
              }
            const double imz = 1.0 / mz;
            double       sx = 0.0, sy = 0.0;
#pragma unroll
            for (int o = 1; o <= RAD; ++o)
              {
                lp += col.tx[o] * (pp[col.xoff[RAD - o]] + pp[col.xoff[RAD + o]]);
                lp += col.ty[o] * (pp[col.yoff[RAD - o]] + pp[col.yoff[RAD + o]]);
                sx += col.tx[o] * (pl[col.xoff[RAD - o]] * col.ixm[RAD - o] +
                                   pl[col.xoff[RAD + o]] * col.ixm[RAD + o]);
                sy += col.ty[o] * (pl[col.yoff[RAD - o]] * col.iym[RAD - o] +
                                   pl[col.yoff[RAD + o]] * col.iym[RAD + o]);
              }
            ll += imz * (col.iym[RAD] * sx + col.ixm[RAD] * sy);
            const double m  = col.mx * col.my * mz;
            const double lv = pl[q];
            if constexpr (STAGE == 0)
              {
                out[base + q] = m * ll;
                g += lv * (-dt * lp);
              }
            else
              {
                out[base + q] =
                  v_in[base + q] - dt * mob * (Energy::template fprime2<double>(pv) * lv + kappa * m * ll);
                g += lv * (-dt * mob * lp);
              }
          }
      }
    g = block_sum(g);
    if (threadIdx.x == 0)
      {
        const int count = int(gridDim.x * gridDim.y);
        const int b     = int(blockIdx.x + gridDim.x * blockIdx.y);
        partial[STAGE * count + b] += g;
      }
  }

  /// Forward runs and vector-Jacobian products of the Allen-Cahn step on one grid.
  template <int dim, int degree, class Energy>
  class AllenCahnAdjoint
  {
  public:
    AllenCahnAdjoint(const GridDesc &g_, hipStream_t s_)
      : g(g_)
      , lat(make_adjoint_lattice<dim, degree>(g_))
      , stream(s_)
    {
      // Column kernels: int in-plane offsets and a z-segment grid row per chunk.
      if (int64_t(g.nodes[0]) * (dim > 1 ? g.nodes[1] : 1) >= (int64_t(1) << 31))
        throw Error(status_unsupported, "adjoint: lattice plane exceeds 2^31 nodes");
      grid = dim3((lat.plane + block_size - 1) / block_size,
                  dim == 3 ? (lat.n[2] + lat.z_chunk - 1) / lat.z_chunk : 1);
    }

    void
    step(double *out, const double *phi, double dt, double L, double kappa, int precision)
    {
      if (precision == 1)
        hipLaunchKernelGGL((ac_forward_kernel<dim, degree, float, Energy>), grid,
                           dim3(block_size), 0, stream, out, phi, lat, dt, L, kappa);
      else
        hipLaunchKernelGGL((ac_forward_kernel<dim, degree, double, Energy>), grid,
                           dim3(block_size), 0, stream, out, phi, lat, dt, L, kappa);
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
      ensure(2);
      const int n_blocks = int(grid.x * grid.y);
      partial.resize(2 * n_blocks);
      sums.resize(2);
      PPF_HIP_CHECK(hipMemsetAsync(partial.get(), 0, 2 * n_blocks * sizeof(double), stream));

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
      PPF_HIP_CHECK(hipMemcpyAsync(lam, lam_final, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
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
              hipLaunchKernelGGL((ac_adjoint_kernel<dim, degree, Energy>), grid,
                                 dim3(block_size), 0, stream, lamp, partial.get(), lam,
                                 segment[s - first].get(), lat, dt, L, kappa);
              std::swap(lam, lamp);
            }
        }
      // Per-workgroup sums over all steps, then workgroups in index order.
      hipLaunchKernelGGL(reduce_two_kernel, dim3(1), dim3(reduce_threads), 0, stream, sums.get(),
                         partial.get(), n_blocks);
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
    AdjointLattice                    lat;
    dim3                              grid;
    hipStream_t                       stream;
    std::vector<DeviceBuffer<double>> work, checkpoints, segment;
    DeviceBuffer<double>              partial, sums;
  };

  /// Forward runs and vector-Jacobian products of the explicit Cahn-Hilliard step on one
  /// grid, with the same entry points as AllenCahnAdjoint: the parameters are the
  /// mobility M (in place of L) and kappa, the state is c.
  template <int dim, int degree, class Energy>
  class CahnHilliardAdjoint
  {
  public:
    CahnHilliardAdjoint(const GridDesc &g_, hipStream_t s_)
      : g(g_)
      , lat(make_adjoint_lattice<dim, degree>(g_))
      , stream(s_)
    {
      if (int64_t(g.nodes[0]) * (dim > 1 ? g.nodes[1] : 1) >= (int64_t(1) << 31))
        throw Error(status_unsupported, "adjoint: lattice plane exceeds 2^31 nodes");
      grid = dim3((lat.plane + block_size - 1) / block_size,
                  dim == 3 ? (lat.n[2] + lat.z_chunk - 1) / lat.z_chunk : 1);
    }

    /// mu = g(c) = f'(c) + kappa L c.
    void
    potential(double *mu, const double *c, double kappa, int precision)
    {
      if (precision == 1)
        hipLaunchKernelGGL((ch_forward_kernel<dim, degree, float, Energy, 1>), grid,
                           dim3(block_size), 0, stream, mu, c, c, lat, kappa);
      else
        hipLaunchKernelGGL((ch_forward_kernel<dim, degree, double, Energy, 1>), grid,
                           dim3(block_size), 0, stream, mu, c, c, lat, kappa);
    }

    /// c_out = c - dt M L mu.
    void
    update(double *c_out, const double *c, const double *mu, double dt, double mob, int precision)
    {
      if (precision == 1)
        hipLaunchKernelGGL((ch_forward_kernel<dim, degree, float, Energy, 0>), grid,
                           dim3(block_size), 0, stream, c_out, c, mu, lat, dt * mob);
      else
        hipLaunchKernelGGL((ch_forward_kernel<dim, degree, double, Energy, 0>), grid,
                           dim3(block_size), 0, stream, c_out, c, mu, lat, dt * mob);
    }

    /// c_final = F^n(c0).
    void
    forward(const double *c0, double *c_final, double dt, double mob, double kappa,
            unsigned int n, int precision)
    {
      ensure(3);
      if (n == 0)
        {
          PPF_HIP_CHECK(hipMemcpyAsync(c_final, c0, sizeof(double) * g.n_nodes,
                                       hipMemcpyDeviceToDevice, stream));
          return;
        }
      double       *mu  = work[2].get();
      const double *cur = c0;
      for (unsigned int s = 0; s < n; ++s)
        {
          potential(mu, cur, kappa, precision);
          double *next = (s + 1 == n) ? c_final : work[s % 2].get();
          update(next, cur, mu, dt, mob, precision);
          cur = next;
        }
      PPF_HIP_CHECK(hipGetLastError());
    }

    /// Given lam_N = dJ/dc_N, returns dJ/dM, dJ/dkappa and dJ/dc0 (into dc0).  Forward
    /// states (c_n and mu_n) are recomputed from checkpoints of c every `every` steps.
    void
    vjp(const double *c0, const double *lam_final, double dt, double mob, double kappa,
        unsigned int n, unsigned int every, int precision, double &dmob, double &dkappa,
        double *dc0)
    {
      if (n == 0)
        {
          PPF_HIP_CHECK(hipMemcpyAsync(dc0, lam_final, sizeof(double) * g.n_nodes,
                                       hipMemcpyDeviceToDevice, stream));
          PPF_HIP_CHECK(hipStreamSynchronize(stream));
          dmob = dkappa = 0.0;
          return;
        }
      if (every == 0)
        every = n;
      const unsigned int n_ckpt = (n + every - 1) / every;
      checkpoints.resize(n_ckpt);
      for (auto &c : checkpoints)
        c.resize(g.n_nodes);
      seg_c.resize(every);
      seg_mu.resize(every);
      for (auto &v : seg_c)
        v.resize(g.n_nodes);
      for (auto &v : seg_mu)
        v.resize(g.n_nodes);
      ensure(4);
      const int n_blocks = int(grid.x * grid.y);
      partial.resize(2 * n_blocks);
      sums.resize(2);
      hipLaunchKernelGGL(fill_kernel<double>, dim3(grid_for(2 * n_blocks, block_size)),
                         dim3(block_size), 0, stream, partial.get(), 0.0, int64_t(2 * n_blocks));

      // Forward sweep storing the checkpoints c_{k*every}, k = 0 .. n_ckpt-1.
      PPF_HIP_CHECK(hipMemcpyAsync(checkpoints[0].get(), c0, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
      double            *mu        = work[2].get();
      const double      *cur       = checkpoints[0].get();
      const unsigned int last_ckpt = (n_ckpt - 1) * every;
      for (unsigned int s = 0; s < last_ckpt; ++s)
        {
          double *next = (s + 1) % every == 0 ? checkpoints[(s + 1) / every].get()
                                              : work[s % 2].get();
          potential(mu, cur, kappa, precision);
          update(next, cur, mu, dt, mob, precision);
          cur = next;
        }

      // Reverse sweep: two passes per step (w = L^T lam, then lam_n).
      double *lam  = work[0].get();
      double *lamp = work[1].get();
      double *w    = work[3].get();
      PPF_HIP_CHECK(hipMemcpyAsync(lam, lam_final, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
      for (int seg = int(n_ckpt) - 1; seg >= 0; --seg)
        {
          const unsigned int first = unsigned(seg) * every;
          const unsigned int last  = std::min(n, first + every); // exclusive
          PPF_HIP_CHECK(hipMemcpyAsync(seg_c[0].get(), checkpoints[seg].get(),
                                       sizeof(double) * g.n_nodes, hipMemcpyDeviceToDevice,
                                       stream));
          for (unsigned int s = first; s < last; ++s)
            {
              potential(seg_mu[s - first].get(), seg_c[s - first].get(), kappa, precision);
              if (s + 1 < last)
                update(seg_c[s + 1 - first].get(), seg_c[s - first].get(),
                       seg_mu[s - first].get(), dt, mob, precision);
            }
          for (int s = int(last) - 1; s >= int(first); --s)
            {
              hipLaunchKernelGGL((ch_adjoint_kernel<dim, degree, Energy, 0>), grid,
                                 dim3(block_size), 0, stream, w, partial.get(), lam,
                                 seg_mu[s - first].get(), lam, lat, dt, mob, kappa);
              hipLaunchKernelGGL((ch_adjoint_kernel<dim, degree, Energy, 1>), grid,
                                 dim3(block_size), 0, stream, lamp, partial.get(), w,
                                 seg_c[s - first].get(), lam, lat, dt, mob, kappa);
              std::swap(lam, lamp);
            }
        }
      hipLaunchKernelGGL(reduce_two_kernel, dim3(1), dim3(reduce_threads), 0, stream, sums.get(),
                         partial.get(), n_blocks);
      PPF_HIP_CHECK(hipMemcpyAsync(dc0, lam, sizeof(double) * g.n_nodes,
                                   hipMemcpyDeviceToDevice, stream));
      double host[2];
      PPF_HIP_CHECK(hipMemcpyAsync(host, sums.get(), 2 * sizeof(double), hipMemcpyDeviceToHost,
                                   stream));
      PPF_HIP_CHECK(hipStreamSynchronize(stream));
      dmob   = host[0];
      dkappa = host[1];
    }

  private:
    void
    ensure(int k)
    {
      if (int(work.size()) < k)
        work.resize(k);
      for (auto &v : work)
        if (v.size() != g.n_nodes)
          v.resize(g.n_nodes);
    }

    GridDesc                          g;
    AdjointLattice                    lat;
    dim3                              grid;
    hipStream_t                       stream;
    std::vector<DeviceBuffer<double>> work, checkpoints, seg_c, seg_mu;
    DeviceBuffer<double>              partial, sums;
  };
} // namespace ppf
