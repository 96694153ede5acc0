// CPU baseline of the Cahn-Hilliard 3-D Q2 explicit increment (the benchmark's
// workload): the reference's matrix-free cell loop restated for the host cores, as
// deal.II runs it for PRISMS-PF.  TEST AND BENCHMARK INFRASTRUCTURE: the benchmark times
// it as the CPU baseline beside the GPU, and the tests check it against the oracle; no
// product code loads it.
//
// Per solve block (reference src/solvers/mf_operator.cc:17-80 with FEEvaluation,
// include/prismspf/core/field_container.h:753-882):
//   * cells are processed in batches of 8 neighbouring cells along x, one per SIMD lane
//     (deal.II's VectorizedArray<double> on AVX-512);
//   * read_dof_values: the 27 nodal values of each input field are gathered per lane;
//   * evaluate: GLL collocation, so values at the quadrature points are the nodal values
//     and gradients are the 1-D derivative matrix applied along each axis (sum
//     factorization), divided by the cell size;
//   * the equations at every quadrature point (the Cahn-Hilliard application's c block:
//     value c_old, gradient -dt M grad(mu_old); mu block: value f'(c), gradient
//     kappa grad(c), f = c^2 (1-c)^2);
//   * integrate: values times the quadrature weights, gradients by the transposed 1-D
//     derivative matrix; distribute_local_to_global adds the 27 results per lane into the
//     global vector, lane by lane;
//   * cell batches run on OpenMP threads in z-slabs of two cell layers, even slabs then
//     odd slabs, so no two threads add into the same node (deal.II colours its cell
//     partitions for the same reason);
//   * ExplicitSolver (reference explicit_solver.h:63-77) then scales by the inverse
//     lumped mass.
// The mesh is the uniform periodic lattice: node (i, j, k), 2 n nodes per axis, index
// i + N (j + N k); cell (x, y, z) holds the nodes 2x + a, 2y + b, 2z + c (a, b, c in
// {0, 1, 2}), wrapped.
#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace
{
  constexpr int W = 8; // lanes per cell batch
  typedef double vd __attribute__((vector_size(W * sizeof(double))));

  // Q2 on GLL points {0, 1/2, 1}: derivative of basis k at point q on the unit interval,
  // and the quadrature weights.
  constexpr double D1[3][3] = {{-3.0, 4.0, -1.0}, {-1.0, 0.0, 1.0}, {1.0, -4.0, 3.0}};
  constexpr double W1[3]    = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};

  struct Lattice
  {
    int    n; // cells per axis (a multiple of 8)
    int    N; // nodes per axis (2 n)
    double h; // cell size
  };

  inline vd
  splat(double x)
  {
    vd v;
    for (int l = 0; l < W; ++l)
      v[l] = x;
    return v;
  }

  inline int64_t
  node(const Lattice &m, int i, int j, int k)
  {
    i = i >= m.N ? i - m.N : i;
    j = j >= m.N ? j - m.N : j;
    k = k >= m.N ? k - m.N : k;
    return int64_t(i) + int64_t(m.N) * (int64_t(j) + int64_t(m.N) * k);
  }

  /// Gathers the 27 nodal values of the batch of cells (x0 .. x0+7, y, z).
  inline void
  gather(const Lattice &m, const double *src, int x0, int y, int z, vd (&u)[27])
  {
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          {
            vd v;
            for (int l = 0; l < W; ++l)
              v[l] = src[node(m, 2 * (x0 + l) + a, 2 * y + b, 2 * z + c)];
            u[a + 3 * (b + 3 * c)] = v;
          }
  }

  /// Gradient at the 27 collocation points by 1-D contractions along each axis.
  inline void
  gradients(const vd (&u)[27], double inv_h, vd (&g)[3][27])
  {
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          {
            vd gx = splat(0.0), gy = splat(0.0), gz = splat(0.0);
            for (int k = 0; k < 3; ++k)
              {
                gx += D1[a][k] * u[k + 3 * (b + 3 * c)];
                gy += D1[b][k] * u[a + 3 * (k + 3 * c)];
                gz += D1[c][k] * u[a + 3 * (b + 3 * k)];
              }
            const int q = a + 3 * (b + 3 * c);
            g[0][q]     = gx * inv_h;
            g[1][q]     = gy * inv_h;
            g[2][q]     = gz * inv_h;
          }
  }

  /// Integrates value terms r and gradient terms t (already times the quadrature
  /// weights and the cell volume) and adds the 27 results per lane into dst.
  inline void
  integrate_add(const Lattice &m, double *dst, int x0, int y, int z, const vd (&r)[27],
                const vd (&t)[3][27], double inv_h)
  {
    vd acc[27];
    for (int c = 0; c < 3; ++c)
      for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
          {
            const int k = a + 3 * (b + 3 * c);
            vd        s = r[k];
            for (int q = 0; q < 3; ++q)
              {
                s += (D1[q][a] * inv_h) * t[0][q + 3 * (b + 3 * c)];
                s += (D1[q][b] * inv_h) * t[1][a + 3 * (q + 3 * c)];
                s += (D1[q][c] * inv_h) * t[2][a + 3 * (b + 3 * q)];
              }
            acc[k] = s;
          }
    for (int l = 0; l < W; ++l)
      for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
          for (int a = 0; a < 3; ++a)
            dst[node(m, 2 * (x0 + l) + a, 2 * y + b, 2 * z + c)] += acc[a + 3 * (b + 3 * c)][l];
  }

  /// One solve block: dst = invM * sum_cells (integrated terms).  `hook(q, w, r, t)`
  /// fills the value and gradient terms of point q (weight w = quadrature weight times
  /// cell volume) from the batch's gathered inputs.
  template <class Eval>
  void
  block(const Lattice &m, double *dst, const Eval &eval)
  {
    const int64_t n_nodes = int64_t(m.N) * m.N * m.N;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_nodes; ++i)
      dst[i] = 0.0;
    const int    slabs = m.n / 2;
    const double inv_h = 1.0 / m.h;
    for (int colour = 0; colour < 2; ++colour)
      {
#pragma omp parallel for schedule(static)
        for (int s = colour; s < slabs; s += 2)
          for (int z = 2 * s; z < 2 * s + 2; ++z)
            for (int y = 0; y < m.n; ++y)
              for (int x0 = 0; x0 < m.n; x0 += W)
                {
                  vd r[27], t[3][27];
                  eval(x0, y, z, r, t);
                  integrate_add(m, dst, x0, y, z, r, t, inv_h);
                }
      }
    // inverse lumped mass: product of 1-D masses (h/3 at cell vertices, 2h/3 at midpoints)
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < int64_t(m.N); ++k)
      for (int j = 0; j < m.N; ++j)
        for (int i = 0; i < m.N; ++i)
          {
            auto m1 = [&](int c) { return (c & 1) ? 2.0 * m.h / 3.0 : m.h / 3.0; };
            dst[i + int64_t(m.N) * (j + int64_t(m.N) * k)] /= m1(i) * m1(int(j)) * m1(int(k));
          }
  }

  inline vd
  fprime(vd c)
  {
    return 2.0 * c * (1.0 - c) * (1.0 - 2.0 * c);
  }

  /// mu block: value f'(c), gradient kappa grad(c).
  void
  mu_block(const Lattice &m, double *mu, const double *c, double kappa)
  {
    const double vol = m.h * m.h * m.h, inv_h = 1.0 / m.h;
    block(m, mu, [&](int x0, int y, int z, vd (&r)[27], vd (&t)[3][27]) {
      vd u[27], g[3][27];
      gather(m, c, x0, y, z, u);
      gradients(u, inv_h, g);
      for (int q = 0; q < 27; ++q)
        {
          const double w = W1[q % 3] * W1[(q / 3) % 3] * W1[q / 9] * vol;
          r[q]           = w * fprime(u[q]);
          for (int d = 0; d < 3; ++d)
            t[d][q] = (w * kappa) * g[d][q];
        }
    });
  }

  /// c block: value c_old, gradient -dt M grad(mu_old).
  void
  c_block(const Lattice &m, double *c_new, const double *c_old, const double *mu_old,
          double dt, double mob)
  {
    const double vol = m.h * m.h * m.h, inv_h = 1.0 / m.h;
    block(m, c_new, [&](int x0, int y, int z, vd (&r)[27], vd (&t)[3][27]) {
      vd uc[27], um[27], g[3][27];
      gather(m, c_old, x0, y, z, uc);
      gather(m, mu_old, x0, y, z, um);
      gradients(um, inv_h, g);
      for (int q = 0; q < 27; ++q)
        {
          const double w = W1[q % 3] * W1[(q / 3) % 3] * W1[q / 9] * vol;
          r[q]           = w * uc[q];
          for (int d = 0; d < 3; ++d)
            t[d][q] = (-w * dt * mob) * g[d][q];
        }
    });
  }
} // namespace

extern "C" {

/// Cahn-Hilliard run on the periodic Q2 lattice of n_cells^3 cells of size h: mu from
/// the initial c (increment 0), then n_increments increments (c block, mu block).
/// c0, c_out, mu_out: (2 n_cells)^3 values in lattice order.  *seconds: wall time of the
/// increments (increment 0 excluded).  Returns 0, or -1 for a cell count that is not a
/// positive multiple of 8.
int
ch_baseline_run(int n_cells, double h, double dt, double mob, double kappa, int n_increments,
                int threads, const double *c0, double *c_out, double *mu_out, double *seconds)
{
  if (n_cells <= 0 || n_cells % W != 0)
    return -1;
  if (threads > 0)
    omp_set_num_threads(threads);
  const Lattice       m {n_cells, 2 * n_cells, h};
  const int64_t       n = int64_t(m.N) * m.N * m.N;
  std::vector<double> c(c0, c0 + n), cn(n), mu(n);
  mu_block(m, mu.data(), c.data(), kappa);
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < n_increments; ++s)
    {
      c_block(m, cn.data(), c.data(), mu.data(), dt, mob);
      c.swap(cn);
      mu_block(m, mu.data(), c.data(), kappa);
    }
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::memcpy(c_out, c.data(), sizeof(double) * n);
  std::memcpy(mu_out, mu.data(), sizeof(double) * n);
  return 0;
}

/// 1 if this build uses 512-bit vectors (AVX-512), 0 otherwise.
int
ch_baseline_isa_avx512(void)
{
#ifdef __AVX512F__
  return 1;
#else
  return 0;
#endif
}
}
