// C-ABI of the oracle for one application (same conventions as include/ppf.h, with
// every array on deal.II's full DoF set).  TEST INFRASTRUCTURE.
//
// An oracle translation unit defines its App description, then
//     #define ORACLE_APP App
//     #include "oracle_capi.hpp"
#pragma once

#include "oracle.hpp"

#include <cstring>
#include <exception>
#include <string>

namespace oracle
{
  struct Handle
  {
    explicit Handle(const std::string &prm)
      : model(prm)
    {}
    Model<ORACLE_APP> model;
    std::string       error;
  };

  inline std::string &
  create_error()
  {
    static thread_local std::string e;
    return e;
  }

  template <class F>
  int
  guard(void *h, F &&f)
  {
    auto *c = static_cast<Handle *>(h);
    if (!c)
      return -1;
    try
      {
        f(c->model);
        c->error.clear();
        return 0;
      }
    catch (const std::exception &e)
      {
        c->error = e.what();
        return -1;
      }
  }
} // namespace oracle

extern "C" {

int
oracle_create(const char *prm, void **out)
{
  try
    {
      *out = new oracle::Handle(prm);
      return 0;
    }
  catch (const std::exception &e)
    {
      oracle::create_error() = e.what();
      *out                   = nullptr;
      return -1;
    }
}

void
oracle_destroy(void *h)
{
  delete static_cast<oracle::Handle *>(h);
}

const char *
oracle_last_error(void *h)
{
  return h ? static_cast<oracle::Handle *>(h)->error.c_str() : oracle::create_error().c_str();
}

int
oracle_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#endif
  return 0;
}

int
oracle_info(void *h, long long *ndofs, int *nfields, int *nblocks)
{
  return oracle::guard(h, [&](auto &m) {
    *ndofs   = m.top.ndofs;
    *nfields = int(m.fields.size());
    *nblocks = int(m.blocks.size());
  });
}

int
oracle_set_option(void *h, const char *key, const char *value)
{
  return oracle::guard(h, [&](auto &m) {
    const std::string k = key, v = value;
    for (auto &b : m.blocks)
      {
        auto &lp = b.linear_solver_parameters;
        if (k == "linear_tolerance")
          lp.tolerance = std::stod(v);
        else if (k == "linear_max_iterations")
          lp.max_iterations = unsigned(std::stoul(v));
        else if (k == "mg_coarse_solver")
          lp.coarse_solver =
            v == "cg" ? prismspf::CoarseSolver::CG : prismspf::CoarseSolver::Smoother;
        else if (k == "mg_level_precision")
          ; // the oracle is FP64 throughout
        else
          throw std::invalid_argument("unknown option " + k);
      }
  });
}

int
oracle_set_field(void *h, int f, int age, const double *v, long long n)
{
  return oracle::guard(h, [&](auto &m) {
    if (n != m.top.ndofs || f < 0 || f >= int(m.storage.size()) || age < 0 ||
        age >= int(m.storage[f].size()))
      throw std::invalid_argument("oracle_set_field: bad field, age or size");
    std::copy(v, v + n, m.storage[f][age].begin());
  });
}

int
oracle_get_field(void *h, int f, int age, double *v, long long n)
{
  return oracle::guard(h, [&](auto &m) {
    if (age < 0 && f >= 0 && f < int(m.storage.size()))
      age = m.latest_age(f);
    if (n != m.top.ndofs || f < 0 || f >= int(m.storage.size()) || age < 0 ||
        age >= int(m.storage[f].size()))
      throw std::invalid_argument("oracle_get_field: bad field, age or size");
    std::copy(m.storage[f][age].begin(), m.storage[f][age].end(), v);
  });
}

int
oracle_set_time(void *h, unsigned increment, double time)
{
  return oracle::guard(h, [&](auto &m) {
    m.timer.set_increment(increment);
    m.timer.set_time(time);
  });
}

int
oracle_apply_operator(void *h, int block_id, int which, int scaling, const double *src,
                      double *dst)
{
  return oracle::guard(h, [&](auto &m) {
    m.apply_top(m.block_index(block_id), which != 0, src, dst, scaling, which == 2);
  });
}

int
oracle_inverse_diagonal(void *h, int block_id, int scaling, double *dst)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    auto      d = m.inverse_diagonal(b, m.top, m.cons, m.jxw, m.storage, scaling);
    std::copy(d.begin(), d.end(), dst);
  });
}

int
oracle_solve_block(void *h, int block_id)
{
  return oracle::guard(h, [&](auto &m) { m.solve_block(m.block_index(block_id)); });
}

int
oracle_rotate_old(void *h, int block_id)
{
  return oracle::guard(h, [&](auto &m) { m.rotate(m.block_index(block_id)); });
}

int
oracle_solve_increment(void *h, int *status)
{
  return oracle::guard(h, [&](auto &m) {
    *status = m.solve_increment();
    m.timer.increment();
  });
}

int
oracle_run(void *h, unsigned n, int *status)
{
  return oracle::guard(h, [&](auto &m) {
    int s = 0;
    while (m.timer.get_increment() <= n && s == 0)
      {
        s = m.solve_increment();
        m.timer.increment();
      }
    *status = s;
  });
}

int
oracle_linear_solve(void *h, int block_id, int *its, double *r0, double *r, double *tol,
                    int *converged)
{
  return oracle::guard(h, [&](auto &m) {
    auto st    = m.solve_linear(m.block_index(block_id));
    *its       = st.iterations;
    *r0        = st.r0;
    *r         = st.r;
    *tol       = st.tol;
    *converged = st.converged ? 1 : 0;
  });
}

int
oracle_last_newton_stats(void *h, int block_id, int *its, int *lin_its, double *r, double *tol,
                         int *converged)
{
  return oracle::guard(h, [&](auto &m) {
    const int b  = m.block_index(block_id);
    const auto st = m.last_newton.count(b) ? m.last_newton.at(b)
                                           : typename std::decay_t<decltype(m)>::NewtonStats {};
    *its       = st.iterations;
    *lin_its   = st.linear_iterations;
    *r         = st.residual;
    *tol       = st.tol;
    *converged = st.converged ? 1 : 0;
  });
}

int
oracle_field_norms(void *h, int f, int age, double *l2, double *integral)
{
  return oracle::guard(h, [&](auto &m) {
    if (age < 0)
      age = m.latest_age(f);
    auto r    = m.norms(f, age);
    *l2       = r.first;
    *integral = r.second;
  });
}

const char *
oracle_output_log(void *h)
{
  return static_cast<oracle::Handle *>(h)->model.log.c_str();
}

int
oracle_lumped_mass(void *h, double *out)
{
  return oracle::guard(h, [&](auto &m) { std::copy(m.jxw.begin(), m.jxw.end(), out); });
}

int
oracle_adjoint_forward(void *h, const double *phi0, double L, double kappa, unsigned n,
                       double *out)
{
  return oracle::guard(h, [&](auto &m) {
    using A = typename std::decay_t<decltype(m)>::AppType;
    if constexpr (oracle::has_energy<A>::value)
      {
        std::vector<double> p(phi0, phi0 + m.top.ndofs), r;
        if constexpr (oracle::is_cahn_hilliard_adjoint<A>::value)
          r = m.template ch_forward<typename A::FreeEnergy>(p, L, kappa, n);
        else
          r = m.template ac_forward<typename A::FreeEnergy>(p, L, kappa, n);
        std::copy(r.begin(), r.end(), out);
      }
    else
      throw std::invalid_argument("no free energy declared");
  });
}

int
oracle_adjoint_vjp(void *h, const double *phi0, const double *lam, double L, double kappa,
                   unsigned n, double *dL, double *dk, double *dphi0)
{
  return oracle::guard(h, [&](auto &m) {
    using A = typename std::decay_t<decltype(m)>::AppType;
    if constexpr (oracle::has_energy<A>::value)
      {
        std::vector<double> p(phi0, phi0 + m.top.ndofs), l(lam, lam + m.top.ndofs), g;
        if constexpr (oracle::is_cahn_hilliard_adjoint<A>::value)
          m.template ch_vjp<typename A::FreeEnergy>(p, l, L, kappa, n, *dL, *dk, g);
        else
          m.template ac_vjp<typename A::FreeEnergy>(p, l, L, kappa, n, *dL, *dk, g);
        std::copy(g.begin(), g.end(), dphi0);
      }
    else
      throw std::invalid_argument("no free energy declared");
  });
}

/// Number of DoFs (full lattice) of multigrid level `level` of block `block_id`
/// (levels of the block's V-cycle, coarsest first).
int
oracle_mg_level_ndofs(void *h, int block_id, int level, long long *ndofs)
{
  return oracle::guard(h, [&](auto &m) {
    m.setup_mg(m.block_index(block_id));
    if (level < 0 || level >= int(m.mg.size()))
      throw std::invalid_argument("level out of range");
    *ndofs = m.mg[level].lvl.mesh.ndofs;
  });
}

/// Transfer of field position 0 of block `block_id` between level `level` (fine) and
/// level - 1 (coarse): direction 0 prolongates (out_fine = P in_coarse), direction 1
/// restricts (out_coarse = P^T in_fine).
int
oracle_mg_transfer(void *h, int block_id, int level, int direction, const double *in,
                   double *out)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.setup_mg(b);
    if (level < 1 || level >= int(m.mg.size()))
      throw std::invalid_argument("level out of range");
    const auto &F     = m.mg[level].lvl;
    const auto &C     = m.mg[level - 1].lvl;
    const int   field = m.block_fields(b)[0];
    if (direction == 0)
      {
        std::fill(out, out + F.mesh.ndofs, 0.0);
        m.prolongate_add(F, C, 0, field, in, out);
      }
    else
      m.restrict_to(F, C, field, in, out);
  });
}

/// The inverse diagonal of multigrid level `level` of block `block_id` (full level
/// lattice, zero diagonal entries replaced by one), as the level's smoother uses it.
int
oracle_mg_inv_diag(void *h, int block_id, int level, double *out)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.setup_mg(b);
    if (level < 0 || level >= int(m.mg.size()))
      throw std::invalid_argument("level out of range");
    const auto &d = m.mg[level].sm.inv_diag;
    std::copy(d.begin(), d.end(), out);
  });
}

/// One application of level `level`'s Chebyshev smoother (the block's degree, its
/// eigenvalue bounds) from a zero start: x = S rhs.  Returns the bounds it used.
int
oracle_mg_smooth(void *h, int block_id, int level, const double *rhs, double *x, double *lmax,
                 double *lmin)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.setup_mg(b);
    if (level < 0 || level >= int(m.mg.size()))
      throw std::invalid_argument("level out of range");
    const std::size_t   n = m.mg[level].sm.inv_diag.size();
    std::vector<double> r(rhs, rhs + n), u(n, 0.0);
    m.chebyshev(m.level_operator(b, level), m.mg[level].sm,
                m.blocks[b].linear_solver_parameters.chebyshev_parameters.degree, u, r, true);
    std::copy(u.begin(), u.end(), x);
    *lmax = m.mg[level].sm.lmax;
    *lmin = m.mg[level].sm.lmin;
  });
}

/// The smoother's eigenvalue estimate on level `level` from the start vector x0:
/// lmax (safety factor included) and lmin = lmax / smoothing range.
int
oracle_mg_estimate(void *h, int block_id, int level, const double *x0, double *lmax,
                   double *lmin)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.setup_mg(b);
    if (level < 0 || level >= int(m.mg.size()))
      throw std::invalid_argument("level out of range");
    auto                sm = m.mg[level].sm;
    std::vector<double> start(x0, x0 + sm.inv_diag.size());
    m.estimate(m.level_operator(b, level), sm, start,
               m.blocks[b].linear_solver_parameters.chebyshev_parameters);
    *lmax = sm.lmax;
    *lmin = sm.lmin;
  });
}

/// One V-cycle of block `block_id` (the GMG preconditioner): u = V(defect) on the
/// finest level.
int
oracle_mg_vcycle(void *h, int block_id, const double *defect, double *u)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.setup_mg(b);
    const std::size_t   n = m.mg.back().sm.inv_diag.size();
    std::vector<double> d(defect, defect + n), out(n, 0.0);
    m.v_cycle(b, int(m.mg.size()) - 1, out, d);
    std::copy(out.begin(), out.end(), u);
  });
}

/// Level operator of block `block_id` on level `level`: dst = A_level src.
int
oracle_mg_level_vmult(void *h, int block_id, int level, const double *src, double *dst)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.setup_mg(b);
    if (level < 0 || level >= int(m.mg.size()))
      throw std::invalid_argument("level out of range");
    const std::size_t   n = m.mg[level].sm.inv_diag.size();
    std::vector<double> x(src, src + n), y;
    m.level_operator(b, level)(x, y);
    std::copy(y.begin(), y.end(), dst);
  });
}

/// h = P^{-1} g with block `block_id`'s preconditioner, set up as for a solve.
int
oracle_precondition(void *h, int block_id, const double *g, double *out)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.prepare_preconditioner(b);
    const std::size_t   n = m.block_fields(b).size() * std::size_t(m.top.ndofs);
    std::vector<double> gv(g, g + n), hv;
    m.precondition(b, gv, hv);
    std::copy(hv.begin(), hv.end(), out);
  });
}

/// Exactly `steps` steps of block `block_id`'s Krylov solver (its solver type,
/// preconditioner and GMRES basis size) on lhs x = rhs from x (tolerance 0, so the
/// iteration stops at the step limit); x is updated in place.
int
oracle_krylov_steps(void *h, int block_id, double *x, const double *rhs, unsigned steps)
{
  return oracle::guard(h, [&](auto &m) {
    const int b = m.block_index(block_id);
    m.prepare_preconditioner(b);
    const auto         &lp = m.blocks[b].linear_solver_parameters;
    const std::size_t   n  = m.block_fields(b).size() * std::size_t(m.top.ndofs);
    std::vector<double> xv(x, x + n), bv(rhs, rhs + n);
    auto                A = m.top_operator(b);
    auto P = [&](const std::vector<double> &g, std::vector<double> &hv) {
      m.precondition(b, g, hv);
    };
    if (lp.solver_type == "gmres" || lp.solver_type == "fgmres")
      m.gmres(A, P, xv, bv, 0.0, steps, int(lp.gmres_max_basis_size),
              lp.gmres_right_preconditioning, lp.solver_type == "fgmres");
    else
      throw std::invalid_argument("oracle_krylov_steps: gmres or fgmres only");
    std::copy(xv.begin(), xv.end(), x);
  });
}

} // extern "C"
