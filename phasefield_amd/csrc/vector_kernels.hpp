// Vector kernels of the solvers: BLAS-1 updates, deterministic reductions, constraint
// helpers and the geometric multigrid transfers.
//
// Reductions are two-stage with a fixed grid, so a dot product gives the same bits on
// every run (the reference's MPI_Allreduce-based dots are deterministic for a fixed
// rank count too).
#pragma once

#include "grid.hpp"

#include <cmath>

namespace ppf
{
  constexpr int reduce_threads = 256;
  constexpr int reduce_blocks  = 1024;

  template <typename T>
  __device__ PPF_INLINE double
  wave_sum(double v)
  {
    for (int off = warpSize / 2; off > 0; off /= 2)
      v += __shfl_down(v, off, warpSize);
    return v;
  }

  /// Block-level sum of one double per thread; valid in thread 0.
  __device__ PPF_INLINE double
  block_sum(double v)
  {
    __shared__ double partial[reduce_threads / 32];
    v             = wave_sum<double>(v);
    const int lane = threadIdx.x % warpSize;
    const int wave = threadIdx.x / warpSize;
    if (lane == 0)
      partial[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
      for (int w = 0; w < int(blockDim.x) / warpSize; ++w)
        r += partial[w];
    return r;
  }

  /// Block-level maximum of one double per thread; valid in thread 0.
  __device__ PPF_INLINE double
  block_max(double v)
  {
    __shared__ double partial[reduce_threads / 32];
    for (int o = warpSize / 2; o > 0; o /= 2)
      v = fmax(v, __shfl_xor(v, o));
    const int lane = threadIdx.x % warpSize;
    const int wave = threadIdx.x / warpSize;
    if (lane == 0)
      partial[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
      for (int w = 0; w < int(blockDim.x) / warpSize; ++w)
        r = fmax(r, partial[w]);
    return r;
  }

  /// Per-block max |a - b| (NaN or infinite counted as +inf) and max |b|, into
  /// partial[b] and partial[gridDim.x + b].
  template <typename T>
  __global__ void __launch_bounds__(reduce_threads)
  max_diff_partial_kernel(double *__restrict__ partial, const T *__restrict__ a,
                          const T *__restrict__ b, int64_t n)
  {
    double d = 0.0, m = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      {
        const double e = fabs(double(a[i]) - double(b[i]));
        d              = e <= d ? d : (e == e && e < INFINITY ? e : INFINITY);
        m              = fmax(m, fabs(double(b[i])));
      }
    d = block_max(d);
    __syncthreads();
    m = block_max(m);
    if (threadIdx.x == 0)
      {
        partial[blockIdx.x]             = d;
        partial[gridDim.x + blockIdx.x] = m;
      }
  }

  template <typename T>
  __global__ void
  fill_kernel(T *__restrict__ x, T value, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      x[i] = value;
  }

  /// dst = src converted; a no-op once *halt is set (halt may be null).
  template <typename T, typename U>
  __global__ void
  convert_kernel(T *__restrict__ dst, const U *__restrict__ src, int64_t n, const int *halt)
  {
    if (halt && *reinterpret_cast<const volatile int *>(halt))
      return;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      dst[i] = T(src[i]);
  }

  /// y = a*x + b*y
  template <typename T>
  __global__ void
  axpby_kernel(T *__restrict__ y, T a, const T *__restrict__ x, T b, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    // b = 0 overwrites y without reading it (y may hold anything, NaN included).
    if (i < n)
      y[i] = b == T(0) ? a * x[i] : a * x[i] + b * y[i];
  }

  /// y = b*y (in place; one array, so no aliasing of restrict-qualified operands)
  template <typename T>
  __global__ void
  scale_kernel(T *__restrict__ y, T b, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      y[i] = b * y[i];
  }

  /// z = x - y
  template <typename T>
  __global__ void
  sub_kernel(T *__restrict__ z, const T *__restrict__ x, const T *__restrict__ y, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      z[i] = x[i] - y[i];
  }

  /// y = d .* x (diagonal preconditioner application)
  template <typename T>
  __global__ void
  diag_mult_kernel(T *__restrict__ y, const T *__restrict__ d, const T *__restrict__ x,
                   int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      y[i] = d[i] * x[i];
  }

  /// inv = 1/d with zero entries mapped to 1 (reference mf_operator.cc:98-100,245-255).
  template <typename T>
  __global__ void
  invert_diagonal_kernel(T *__restrict__ inv, const T *__restrict__ d, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      inv[i] = d[i] == T(0) ? T(1) : T(1) / d[i];
  }

  /// Start vector of the eigenvalue estimate: splitmix64 of the index, uniform in
  /// [-0.5, 0.5) (integer mixing and exact 53-bit scaling: the same values as on a CPU).
  template <typename T>
  __global__ void
  splitmix_fill_kernel(T *__restrict__ x, int64_t n)
  {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      {
        uint64_t z = uint64_t(i) + 0x9e3779b97f4a7c15ull;
        z          = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z          = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        x[i] = T(double(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
      }
  }

  /// Stage 1 of a dot product: one partial per block (grid-stride, fixed grid).
  template <typename T>
  __global__ void __launch_bounds__(reduce_threads)
  dot_partial_kernel(double *__restrict__ partial, const T *__restrict__ x,
                     const T *__restrict__ y, int64_t n)
  {
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      s += double(x[i]) * double(y[i]);
    s = block_sum(s);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = s;
  }

  /// Stage 2: sums `count` partials in a fixed order into out[slot].
  __global__ void __launch_bounds__(reduce_threads)
  reduce_partials_kernel(double *__restrict__ out, const double *__restrict__ partial,
                         int count)
  {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += blockDim.x)
      s += partial[i];
    s = block_sum(s);
    if (threadIdx.x == 0)
      out[0] = s;
  }

  // ---------------------------------------------------- device-side CG scalars
  //
  // Small CG loops (eigenvalue estimates, coarse-grid solves) keep their scalars in
  // device memory: dot products land in slots, alpha and beta are formed in the update
  // kernels with the same expressions the host loop uses, and the host reads the slots
  // once per estimate or every few coarse iterations instead of three times per step.

  /// Iteration i of the eigenvalue estimate is valid if rz and pAp stayed positive up
  /// to and including i (the host loop stops at the first that does not).
  __device__ inline bool
  eig_valid(const double *rz, const double *pap, int i)
  {
    for (int j = 0; j <= i; ++j)
      if (!(rz[j] > 0.0) || !(pap[j] > 0.0))
        return false;
    return true;
  }

  /// r = (-alpha) ap + r with alpha = rz_i / pAp_i, if iteration i is valid.
  template <typename T>
  __global__ void
  eig_r_update_kernel(T *__restrict__ r, const T *__restrict__ ap, const double *rz,
                      const double *pap, int i, int64_t n)
  {
    if (!eig_valid(rz, pap, i))
      return;
    const double  alpha = rz[i] / pap[i];
    const int64_t k     = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n)
      r[k] = T(-alpha) * ap[k] + T(1) * r[k];
  }

  /// p = z + beta p with beta = rz_{i+1} / rz_i, if iteration i is valid.
  template <typename T>
  __global__ void
  eig_p_update_kernel(T *__restrict__ p, const T *__restrict__ z, const double *rz,
                      const double *pap, int i, int64_t n)
  {
    if (!eig_valid(rz, pap, i))
      return;
    const double  beta = rz[i + 1] / rz[i];
    const int64_t k    = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n)
      p[k] = T(1) * z[k] + T(beta) * p[k];
  }

  /// Slots of the coarse CG: |r0|^2, rz (two parities), pAp, |r|^2; flag[0] = done.
  enum CoarseSlot : int
  {
    cs_r0sq = 0,
    cs_rz   = 1,
    cs_pap  = 3,
    cs_rr   = 4
  };

  __global__ void
  coarse_cg_start_kernel(const double *s, int *flag)
  {
    flag[0] = s[cs_r0sq] > 0.0 ? 0 : 1;
  }

  /// x = alpha p + x, r = (-alpha) ap + r with alpha = rz / pAp, unless done.
  __global__ void
  coarse_cg_xr_kernel(double *__restrict__ x, double *__restrict__ r, const double *__restrict__ p,
                      const double *__restrict__ ap, const double *s, const int *flag, int par,
                      int64_t n)
  {
    if (flag[0])
      return;
    const double  alpha = s[cs_rz + par] / s[cs_pap];
    const int64_t k     = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n)
      {
        x[k] = alpha * p[k] + 1.0 * x[k];
        r[k] = (-alpha) * ap[k] + 1.0 * r[k];
      }
  }

  /// Marks the solve done once |r| <= tol |r0|.
  __global__ void
  coarse_cg_check_kernel(const double *s, int *flag, double tol)
  {
    if (!flag[0] && std::sqrt(s[cs_rr]) <= tol * std::sqrt(s[cs_r0sq]))
      flag[0] = 1;
  }

  /// p = z + beta p with beta = rz_new / rz, unless done.
  __global__ void
  coarse_cg_p_kernel(double *__restrict__ p, const double *__restrict__ z, const double *s,
                     const int *flag, int par, int64_t n)
  {
    if (flag[0])
      return;
    const double  beta = s[cs_rz + (par ^ 1)] / s[cs_rz + par];
    const int64_t k    = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n)
      p[k] = 1.0 * z[k] + beta * p[k];
  }

  /// CG update fused with the residual norm: x += alpha p, r += alpha Ap, partial |r|^2.
  template <typename T>
  __global__ void __launch_bounds__(reduce_threads)
  cg_update_kernel(double *__restrict__ partial, T *__restrict__ x, T *__restrict__ r,
                   const T *__restrict__ p, const T *__restrict__ ap, T alpha, int64_t n)
  {
    double s = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      {
        x[i] += alpha * p[i];
        const T ri = r[i] + alpha * ap[i];
        r[i]       = ri;
        s += double(ri) * double(ri);
      }
    s = block_sum(s);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = s;
  }

  /// p = -h + beta p
  template <typename T>
  __global__ void
  cg_direction_kernel(T *__restrict__ p, const T *__restrict__ h, T beta, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      p[i] = -h[i] + beta * p[i];
  }

  /// cg_update_kernel with alpha = s[num] / s[den] read from device slots (the same
  /// double division the host would do).
  __global__ void __launch_bounds__(reduce_threads)
  cg_update_slots_kernel(double *__restrict__ partial, double *__restrict__ x,
                         double *__restrict__ r, const double *__restrict__ p,
                         const double *__restrict__ ap, const double *__restrict__ s, int num,
                         int den, int64_t n)
  {
    const double alpha = s[num] / s[den];
    double       acc   = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      {
        x[i] += alpha * p[i];
        const double ri = r[i] + alpha * ap[i];
        r[i]            = ri;
        acc += ri * ri;
      }
    acc = block_sum(acc);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = acc;
  }

  /// BiCGStab scalars in device slots: 0 rho, 1 alpha, 2 omega, 3 r.rbar, 4 rbar.v,
  /// 5 t.r, 6 t.t, 7 beta.  One thread; `op` 0: beta = (r.rbar) alpha / (rho omega),
  /// rho = r.rbar; 1: alpha = rho / (rbar.v); 2: omega = (t.r) / (t.t), 0 when t.t is
  /// below the smallest normal double.
  __global__ void
  bicgstab_scalar_kernel(double *s, int op)
  {
    if (op == 0)
      {
        s[7] = s[3] * s[1] / (s[0] * s[2]);
        s[0] = s[3];
      }
    else if (op == 1)
      s[1] = s[0] / s[4];
    else
      s[2] = s[6] < 2.2250738585072014e-308 ? 0.0 : s[5] / s[6];
  }

  /// p = beta p + r - beta omega v (beta = s[7], omega = s[2]).
  __global__ void
  bicgstab_p_kernel(double *__restrict__ p, const double *__restrict__ r,
                    const double *__restrict__ v, const double *__restrict__ s, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      p[i] = (s[7] * p[i] + r[i]) - s[7] * s[2] * v[i];
  }

  /// y += sign s[k] x.
  __global__ void
  axpy_slot_kernel(double *__restrict__ y, const double *__restrict__ x,
                   const double *__restrict__ s, int k, double sign, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      y[i] += (sign * s[k]) * x[i];
  }

  /// y += sign s[k] x, with per-block partial sums of y.y (grid-stride).
  __global__ void __launch_bounds__(reduce_threads)
  axpy_slot_norm_kernel(double *__restrict__ partial, double *__restrict__ y,
                        const double *__restrict__ x, const double *__restrict__ s, int k,
                        double sign, int64_t n)
  {
    const double a   = sign * s[k];
    double       acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      {
        const double yi = y[i] + a * x[i];
        y[i]            = yi;
        acc += yi * yi;
      }
    acc = block_sum(acc);
    if (threadIdx.x == 0)
      partial[blockIdx.x] = acc;
  }

  /// x += alpha y + omega z (alpha = s[1], omega = s[2]).
  __global__ void
  bicgstab_x_kernel(double *__restrict__ x, const double *__restrict__ y,
                    const double *__restrict__ z, const double *__restrict__ s, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      x[i] += s[1] * y[i] + s[2] * z[i];
  }

  /// cg_direction_kernel with beta = s[num] / s[den] read from device slots.
  __global__ void
  cg_direction_slots_kernel(double *__restrict__ p, const double *__restrict__ h,
                            const double *__restrict__ s, int num, int den, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
      p[i] = -h[i] + (s[num] / s[den]) * p[i];
  }

  /// First Chebyshev step from x = 0 (A x = 0): d = c2 P^{-1} b, x = d.  Same values as
  /// chebyshev_step_kernel with x = ax = 0, without reading or clearing them.
  template <typename T>
  __global__ void
  chebyshev_first_zero_kernel(T *__restrict__ x, T *__restrict__ d, const T *__restrict__ b,
                              const T *__restrict__ inv_diag, T c2, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
      return;
    const T di = c2 * inv_diag[i] * b[i];
    d[i]       = di;
    x[i]       = di;
  }

  /// Chebyshev step (Varga / Saad 12.1 recurrence as used by deal.II's
  /// PreconditionChebyshev): d = c1 * d + c2 * P^{-1} (b - A x);  x += d.
  /// With `first` set, d = c2 * P^{-1}(b - A x).
  template <typename T>
  __global__ void
  chebyshev_step_kernel(T *__restrict__ x, T *__restrict__ d, const T *__restrict__ b,
                        const T *__restrict__ ax, const T *__restrict__ inv_diag, T c1,
                        T c2, int first, int64_t n)
  {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
      return;
    const T r  = b[i] - ax[i];
    const T di = first ? c2 * inv_diag[i] * r : c1 * d[i] + c2 * inv_diag[i] * r;
    d[i]       = di;
    x[i] += di;
  }

  // Inverse diagonal by node class.  For constant-coefficient split operators (the
  // Poisson and level operators) the diagonal is a function of each axis' node class
  // only: 0 odd (Q2), 1 even interior (or periodic), 2 lower and 3 upper non-periodic
  // boundary.  The smoother then reads a 4^dim table per field instead of a diagonal
  // array (one stream fewer per step).  The table is checked node by node against the
  // computed diagonal whenever it is rebuilt; where any node differs the kernels read
  // the diagonal array as before (the flag is read once per thread, no host sync).

  /// Lattice of a diagonal table: nodes and periodicity per axis, fields, degree.
  struct ClassGrid
  {
    int     n[3];
    int     periodic[3];
    int     degree;
    int     dim;
    int     n_fields;
    int64_t n_nodes;
  };

  constexpr int diag_classes = 64;

  PPF_HD PPF_INLINE int
  axis_class(const ClassGrid &g, int d, int i)
  {
    if (d >= g.dim)
      return 1;
    if (!g.periodic[d])
      {
        if (i == 0)
          return 2;
        if (i == g.n[d] - 1)
          return 3;
      }
    return (g.degree == 2 && (i & 1)) ? 0 : 1;
  }

  /// Representative coordinate of class c along axis d, -1 if the class does not occur.
  PPF_HD PPF_INLINE int
  class_representative(const ClassGrid &g, int d, int c)
  {
    if (d >= g.dim)
      return c == 1 ? 0 : -1;
    const int n = g.n[d];
    int       i = -1;
    if (c == 0)
      i = (g.degree == 2 && n >= 2) ? 1 : -1;
    else if (c == 1)
      i = g.periodic[d] ? 0 : (g.degree == 2 ? 2 : 1);
    else if (c == 2)
      i = g.periodic[d] ? -1 : 0;
    else
      i = (g.periodic[d] || n < 2) ? -1 : n - 1;
    return (i >= 0 && i < n && axis_class(g, d, i) == c) ? i : -1;
  }

  /// table[f * 64 + class] from the diagonal at one node of each occurring class; one
  /// workgroup of n_fields * 64 threads; sets *ok = 1.
  template <typename T>
  __global__ void
  diag_table_build_kernel(T *__restrict__ table, int *__restrict__ ok,
                          const T *__restrict__ inv_diag, const ClassGrid g)
  {
    const int t = threadIdx.x;
    if (t == 0)
      *ok = 1;
    if (t >= g.n_fields * diag_classes)
      return;
    const int f = t / diag_classes, c = t % diag_classes;
    const int i0 = class_representative(g, 0, c & 3);
    const int i1 = class_representative(g, 1, (c >> 2) & 3);
    const int i2 = class_representative(g, 2, c >> 4);
    T         v  = T(0);
    if (i0 >= 0 && i1 >= 0 && i2 >= 0)
      v = inv_diag[f * g.n_nodes + i0 + int64_t(g.n[0]) * (i1 + int64_t(g.n[1]) * i2)];
    table[t] = v;
  }

  PPF_HD PPF_INLINE int
  node_class(const ClassGrid &g, int i0, int i1, int i2)
  {
    return axis_class(g, 0, i0) + 4 * axis_class(g, 1, i1) + 16 * axis_class(g, 2, i2);
  }

  /// Clears *ok where a node's diagonal differs from its class entry.  Grid: x over the
  /// n0 * n1 plane, y over n2, z over the fields.
  template <typename T>
  __global__ void
  diag_table_check_kernel(int *__restrict__ ok, const T *__restrict__ table,
                          const T *__restrict__ inv_diag, const ClassGrid g)
  {
    const unsigned int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= unsigned(g.n[0]) * unsigned(g.n[1]))
      return;
    const int     i0 = int(r % unsigned(g.n[0])), i1 = int(r / unsigned(g.n[0]));
    const int     i2 = int(blockIdx.y), f = int(blockIdx.z);
    const int64_t i  = f * g.n_nodes + int64_t(i2) * g.n[0] * g.n[1] + r;
    if (inv_diag[i] != table[f * diag_classes + node_class(g, i0, i1, i2)])
      *ok = 0;
  }

  /// chebyshev_step_kernel / chebyshev_first_zero_kernel (first_zero set) with the
  /// inverse diagonal taken from the class table when *ok (grid as the check kernel).
  template <typename T>
  __global__ void
  chebyshev_table_kernel(T *__restrict__ x, T *__restrict__ d, const T *__restrict__ b,
                         const T *__restrict__ ax, const T *__restrict__ inv_diag,
                         const T *__restrict__ table, const int *__restrict__ ok, T c1, T c2,
                         int first, int first_zero, const ClassGrid g)
  {
    const unsigned int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= unsigned(g.n[0]) * unsigned(g.n[1]))
      return;
    const int     i0 = int(r % unsigned(g.n[0])), i1 = int(r / unsigned(g.n[0]));
    const int     i2 = int(blockIdx.y), f = int(blockIdx.z);
    const int64_t i  = f * g.n_nodes + int64_t(i2) * g.n[0] * g.n[1] + r;
    const T inv = *ok ? table[f * diag_classes + node_class(g, i0, i1, i2)] : inv_diag[i];
    if (first_zero)
      {
        const T di = c2 * inv * b[i];
        d[i]       = di;
        x[i]       = di;
        return;
      }
    const T rr = b[i] - ax[i];
    const T di = first ? c2 * inv * rr : c1 * d[i] + c2 * inv * rr;
    d[i]       = di;
    x[i] += di;
  }

  /// NaN/Inf probe of entry 0 (reference problem.cc:453-469 checks one entry per field).
  template <typename T>
  __global__ void
  finite_probe_kernel(int *__restrict__ flag, const T *__restrict__ x)
  {
    if (threadIdx.x == 0 && blockIdx.x == 0)
      flag[0] = isfinite(double(x[0])) ? 0 : 1;
  }

  /// Entry-0 NaN/Inf probe of every field at once: writes the index of the first field
  /// whose entry 0 is not finite, or -1.
  struct ProbeFields
  {
    const double *ptr[PPF_MAX_FIELDS];
    int           n;
  };

  /// Probe of entry 0 of every field.  Writes (tag << 5) | (bad + 1) to *result, where
  /// bad is the first field whose entry 0 is not finite (-1: none), in one store so that
  /// a host polling the word sees the tag and the result together.  When a field is bad
  /// and halt is given, sets *halt so that the launches carrying the flag that follow on
  /// the stream do nothing.
  __global__ void
  finite_probe_fields_kernel(int *__restrict__ result, const ProbeFields fields,
                             int *__restrict__ halt, unsigned int hmask, int tag)
  {
    if (threadIdx.x == 0 && blockIdx.x == 0)
      {
        int bad = -1;
        for (int f = 0; f < fields.n && f < PPF_MAX_FIELDS; ++f)
          if (bad < 0 && !isfinite(fields.ptr[f][0]))
            bad = f;
        if (bad >= 0 && halt)
          for (int i = 0; i < 32; ++i)
            if ((hmask >> i) & 1u)
              halt[i] = 1;
        *reinterpret_cast<volatile int *>(result) = (tag << 5) | (bad + 1);
      }
  }

  /// Output diagnostics: partial sums of u^2 over deal.II's full node set (periodic
  /// images counted with their multiplicity) and of u * jxw (images carry jxw = 0).
  template <int dim, int degree, typename T>
  __global__ void __launch_bounds__(reduce_threads)
  norms_partial_kernel(double *__restrict__ partial, const GridDesc g, const T *__restrict__ u)
  {
    double s2 = 0.0, si = 0.0;
    for (int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; node < g.n_nodes;
         node += int64_t(gridDim.x) * blockDim.x)
      {
        int i[3];
        i[0] = int(node % g.nodes[0]);
        i[1] = int((node / g.nodes[0]) % g.nodes[1]);
        i[2] = dim == 3 ? int(node / (int64_t(g.nodes[0]) * g.nodes[1])) : 0;
        double mult = 1.0;
        for (int d = 0; d < dim; ++d)
          if (g.periodic[d] && i[d] == 0)
            mult *= 2.0;
        const double v = double(u[node]);
        s2 += mult * v * v;
        si += v * lumped_mass<dim, degree>(g, i);
      }
    s2 = block_sum(s2);
    __syncthreads();
    si = block_sum(si);
    if (threadIdx.x == 0)
      {
        partial[blockIdx.x]             = s2;
        partial[gridDim.x + blockIdx.x] = si;
      }
  }

  /// dst[idx[k]] = vals[k] for k < count.
  template <typename T>
  __global__ void
  scatter_kernel(T *__restrict__ dst, const int64_t *__restrict__ idx, const T *__restrict__ vals,
                 int64_t count)
  {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < count)
      dst[idx[k]] = vals[k];
  }

  /// dst = value at non-Dirichlet nodes, g at Dirichlet nodes.
  template <int dim, typename T>
  __global__ void
  dirichlet_compose_kernel(T *__restrict__ dst, const GridDesc g, const T *__restrict__ gvals,
                           unsigned int mask, T value, int keep_interior)
  {
    const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= g.n_nodes)
      return;
    int i[3];
    i[0] = int(node % g.nodes[0]);
    i[1] = int((node / g.nodes[0]) % g.nodes[1]);
    i[2] = dim == 3 ? int(node / (int64_t(g.nodes[0]) * g.nodes[1])) : 0;
    if (is_dirichlet<dim>(g, i, mask))
      dst[node] = gvals[node];
    else if (!keep_interior)
      dst[node] = value;
  }

  /// Zeroes the Dirichlet nodes of x (homogeneous level constraints).
  template <int dim, typename T>
  __global__ void
  zero_dirichlet_kernel(T *__restrict__ x, const GridDesc g, unsigned int mask)
  {
    const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= g.n_nodes)
      return;
    int i[3];
    i[0] = int(node % g.nodes[0]);
    i[1] = int((node / g.nodes[0]) % g.nodes[1]);
    i[2] = dim == 3 ? int(node / (int64_t(g.nodes[0]) * g.nodes[1])) : 0;
    if (is_dirichlet<dim>(g, i, mask))
      x[node] = T(0);
  }

  // ------------------------------------------------------------------ transfers
  //
  // Nested uniform grids: the fine grid has twice the cells per direction.  Fine lattice
  // coordinate I sits at coarse lattice position I/2.  Prolongation evaluates the coarse
  // Q_p function at the fine nodes (exact embedding; this is the interpolation that
  // deal.II's MGTransferGlobalCoarsening uses for h-coarsening); restriction is its
  // transpose.  1-D weights:
  //   p = 1: even I -> coarse I/2 (1); odd I -> (I-1)/2, (I+1)/2 (1/2, 1/2)
  //   p = 2: even I -> coarse I/2 (1); odd I in coarse cell c = I/4 at xi = 1/4 or 3/4
  //          -> local nodes (0,1,2) with (3/8, 3/4, -1/8) or (-1/8, 3/4, 3/8).

  /// One axis of a separable transfer between nested lattices (prolongation and
  /// restriction are tensor products of the 1-D operators above, so they run as one
  /// pass per axis; far fewer loads than the 3^d / 5^d gathers).  `in` and `out` are
  /// lattices of extents ei / eo (x fastest), equal except along AXIS.  Prolongation
  /// passes (RESTRICT false) index the output by fine coordinates I and take the 1-D
  /// weights above on the coarse grid `g1d`; restriction passes index the output by
  /// coarse coordinates J and gather the transposed weights from fine positions 2J,
  /// 2J -+ 1 (and 2J -+ 3 for p = 2, even J) of the fine grid `g1d`, dropping positions
  /// outside a non-periodic lattice.  Every output takes a fixed number of taps, an
  /// absent one adding an exact zero.  The last pass of a prolongation adds into the
  /// fine vector except at its Dirichlet nodes (MODE 1); the last pass of a restriction
  /// zeroes the coarse Dirichlet nodes (MODE 2); intermediate passes store (MODE 0).
  /// `gout` is the output grid for MODE 1 and 2 (storage index and Dirichlet test).
  /// Grid: x over the output plane (eo0 * eo1 < 2^31), y over its planes.
  template <int dim, int degree, typename T, bool RESTRICT, int MODE, int AXIS>
  __global__ void
  transfer_axis_kernel(T *__restrict__ out, const T *__restrict__ in, int eo0, int eo1,
                       int ei0, int ei1, const GridDesc g1d, const GridDesc gout,
                       unsigned int mask)
  {
    const unsigned int r     = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned int plane = unsigned(eo0) * unsigned(eo1);
    if (r >= plane)
      return;
    int o[3];
    o[0]            = int(r % unsigned(eo0));
    o[1]            = int(r / unsigned(eo0));
    o[2]            = int(blockIdx.y);
    const int64_t t = int64_t(o[2]) * plane + r;
    if constexpr (MODE == 1)
      if (is_dirichlet<dim>(gout, o, mask))
        return;
    if constexpr (MODE == 2)
      if (is_dirichlet<dim>(gout, o, mask))
        {
          out[t] = T(0);
          return;
        }
    // input offset of the output's position with the AXIS coordinate zero
    int c[3] = {o[0], o[1], o[2]};
    c[AXIS]  = 0;
    const int64_t base   = c[0] + int64_t(ei0) * (c[1] + int64_t(ei1) * c[2]);
    const int64_t stride = AXIS == 0 ? 1 : (AXIS == 1 ? int64_t(ei0) : int64_t(ei0) * ei1);
    const int     n      = g1d.nodes[AXIS];
    const bool    per    = g1d.periodic[AXIS] != 0;
    // Tap positions stay within one period of the lattice (a finer periodic lattice has
    // at least 2 cells), so one wrap suffices.
    const auto    tap    = [&](int i, bool &ok) {
      if (per)
        i = i < 0 ? i + n : (i >= n ? i - n : i);
      else
        ok = ok && i >= 0 && i < n;
      return ok ? i : 0;
    };
    const int a = o[AXIS];
    T         sum = T(0);
    if constexpr (RESTRICT)
      {
        // taps in the order 2J, 2J - 1, 2J + 1, 2J - 3, 2J + 3
        constexpr int K = degree == 1 ? 3 : 5;
        int           off[K];
        T             w[K];
        if constexpr (degree == 1)
          {
            off[0] = 0, off[1] = -1, off[2] = 1;
            w[0] = T(1.0), w[1] = T(0.5), w[2] = T(0.5);
          }
        else
          {
            const bool even = (a & 1) == 0;
            off[0] = 0, off[1] = -1, off[2] = 1, off[3] = -3, off[4] = 3;
            w[0] = T(1.0);
            w[1] = w[2] = even ? T(3.0 / 8.0) : T(3.0 / 4.0);
            w[3] = w[4] = even ? T(-1.0 / 8.0) : T(0);
          }
#pragma unroll
        for (int k = 0; k < K; ++k)
          {
            bool      ok = degree == 1 || k < 3 || (a & 1) == 0;
            const int i  = tap(2 * a + off[k], ok);
            const T   x  = in[base + i * stride];
            sum += ok ? w[k] * x : T(0);
          }
      }
    else
      {
        // taps in the order of the coarse local nodes
        if ((a & 1) == 0)
          sum += T(1.0) * in[base + (a / 2) * stride];
        else if constexpr (degree == 1)
          {
            bool      ok = true;
            const int j1 = tap((a + 1) / 2, ok);
            sum += T(0.5) * in[base + ((a - 1) / 2) * stride];
            sum += T(0.5) * in[base + j1 * stride];
          }
        else
          {
            const int  b0 = 2 * (a / 4);
            bool       ok = true;
            const int  j2 = tap(b0 + 2, ok);
            const bool q1 = (a & 3) == 1;
            sum += (q1 ? T(3.0 / 8.0) : T(-1.0 / 8.0)) * in[base + b0 * stride];
            sum += T(3.0 / 4.0) * in[base + (b0 + 1) * stride];
            sum += (q1 ? T(-1.0 / 8.0) : T(3.0 / 8.0)) * in[base + j2 * stride];
          }
      }
    if constexpr (MODE == 1)
      out[t] += sum;
    else
      out[t] = sum;
  }


  /// Nodal injection fine -> coarse (coarse lattice J = fine lattice stride*J, stride a
  /// power of two); transfers non-operand fields to the levels
  /// (reference group_solution_handler.h:356-417).
  template <int dim, typename T, typename U>
  __global__ void
  inject_kernel(T *__restrict__ coarse, const GridDesc gc, const U *__restrict__ fine,
                const GridDesc gf, int stride)
  {
    const int64_t node = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (node >= gc.n_nodes)
      return;
    const int j0 = int(node % gc.nodes[0]);
    const int j1 = int((node / gc.nodes[0]) % gc.nodes[1]);
    const int j2 = dim == 3 ? int(node / (int64_t(gc.nodes[0]) * gc.nodes[1])) : 0;
    coarse[node] = T(fine[node_index(gf, stride * j0, stride * j1, stride * j2)]);
  }
} // namespace ppf
