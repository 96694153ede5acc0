// Host-side analysis of an equation hook: dependency validation and detection of the
// collocation-split structure.
//
// The reference validates, in debug builds, that a hook only reads declared
// dependencies and only submits terms for fields of its block
// (reference: include/prismspf/core/field_container.h:1020-1063, 1258-1314).  The probe
// below does that on the host for every (block, lhs/rhs) pair by calling the hook on a
// FieldContainer filled with pseudo-random values and checking which slots it touched.
//
// The same evaluations decide whether the operator has the split structure used by
// split_operator_kernel:
//   * no use of the quadrature-point location or the element volume;
//   * value terms that do not change when any gradient input changes;
//   * gradient terms equal to sum_in beta_in * grad(in) with scalar beta_in that do not
//     depend on the values (checked on random inputs to a relative 1e-12 in FP64,
//     1e-5 in FP32).
// A hook that fails any check runs on the generic cell path.
#pragma once

#include "device.hpp"
#include "operator_kernels.hpp"

#include <prismspf/core/dependencies.h>
#include <prismspf/core/solve_block.h>

#include <cmath>
#include <random>
#include <sstream>
#include <string>
#include <vector>

namespace ppf
{
  /// One declared input of an operator.
  struct InputSlot
  {
    int          field;
    int          slot;
    unsigned int flags; // slot_values | slot_gradients
  };

  inline std::vector<InputSlot>
  declared_inputs(const prismspf::DependencyMap &deps)
  {
    std::vector<InputSlot> in;
    auto add = [&](int f, int slot, prismspf::EvalFlags ef) {
      if (ef & dealii::EvaluationFlags::hessians)
        throw Error(status_unsupported,
                    "hessian/laplacian dependencies are not supported on the MI355X path");
      unsigned int fl = 0;
      if (ef & dealii::EvaluationFlags::values)
        fl |= slot_values;
      if (ef & dealii::EvaluationFlags::gradients)
        fl |= slot_gradients;
      if (fl)
        in.push_back({f, slot, fl});
    };
    for (const auto &[f, dep] : deps)
      {
        if (f >= PPF_MAX_FIELDS)
          throw Error(status_unsupported, "field index exceeds PPF_MAX_FIELDS");
        add(int(f), prismspf::dependency_slot(prismspf::SRC), dep.src_flag);
        add(int(f), prismspf::dependency_slot(prismspf::Current), dep.flag);
        for (std::size_t a = 0; a < dep.old_flags.size(); ++a)
          {
            if (a + 1 > PPF_MAX_AGE && dep.old_flags[a] != prismspf::EvalFlags::nothing)
              throw Error(status_unsupported, "dependencies older than old_4 are not supported");
            if (a + 1 <= PPF_MAX_AGE)
              add(int(f), prismspf::dependency_slot(prismspf::DependencyType(a + 1)),
                  dep.old_flags[a]);
          }
      }
    return in;
  }

  struct SplitPlan
  {
    bool        eligible       = false;
    bool        time_dependent = false;
    std::string reason;
    SplitTerms  terms {};
    /// Derivative of each field's value term with respect to its own src value, when
    /// that derivative is a constant (used for the operator diagonal).
    bool   alpha_constant = false;
    double alpha[PPF_MAX_FIELDS] {};
    double beta_src[PPF_MAX_FIELDS] {};
  };

  template <class PDE, int dim, int degree, typename number>
  class HookProbe
  {
  public:
    using FC = prismspf::FieldContainer<dim, degree, number>;

    HookProbe(const PDE &pde_, const prismspf::SolveBlock &block_, bool lhs_)
      : pde(pde_)
      , block(block_)
      , lhs(lhs_)
      , inputs(declared_inputs(lhs_ ? block_.dependencies_lhs : block_.dependencies_rhs))
    {
      for (auto f : block.field_indices)
        out_mask |= 1u << f;
    }

    struct State
    {
      number v[PPF_MAX_FIELDS][PPF_N_SLOTS] {};
      number g[PPF_MAX_FIELDS][PPF_N_SLOTS][dim] {};
      number q[dim] {};
      number volume = number(1);
    };

    /// Evaluates the hook at simulated time `time` and increment `inc`; returns the
    /// container with the submitted terms.
    FC
    eval(const State &st, double time, double dt, unsigned int inc) const
    {
      FC fc;
      for (int f = 0; f < PPF_MAX_FIELDS; ++f)
        for (int s = 0; s < PPF_N_SLOTS; ++s)
          {
            fc.value[f][s] = st.v[f][s];
            for (int d = 0; d < dim; ++d)
              fc.gradient[f][s][d] = st.g[f][s][d];
          }
      for (int d = 0; d < dim; ++d)
        fc.q_point[d] = st.q[d];
      fc.element_volume = st.volume;
      fc.clear_terms();
      fc.clear_reads();
      prismspf::SimulationTimer timer(dt, time);
      timer.set_increment(inc);
      const prismspf::PDEOperatorBase<dim, degree, number> &base = pde;
      if (lhs)
        base.compute_lhs(fc, timer, block.id);
      else
        base.compute_rhs(fc, timer, block.id);
      validate(fc);
      return fc;
    }

    /// A random input state whose values lie in [centre - width, centre + width].
    State
    random_state(std::mt19937_64 &rng, double centre, double width) const
    {
      std::uniform_real_distribution<double> u(-1.0, 1.0);
      State                                  st;
      for (const auto &in : inputs)
        {
          st.v[in.field][in.slot] = number(centre + width * u(rng));
          for (int d = 0; d < dim; ++d)
            st.g[in.field][in.slot][d] = number(u(rng));
        }
      for (int d = 0; d < dim; ++d)
        st.q[d] = number(3.0 + u(rng));
      st.volume = number(1.0 + 0.25 * u(rng));
      return st;
    }

    /// Structure of the hook at simulated time `time` and increment `inc`.  The random
    /// states cover values in [0.1, 0.9] (order parameters, concentrations), [-3, 3] and
    /// [-50, 50], so a coefficient that branches on a value outside the unit interval
    /// is seen as non-constant; the coefficients are compared at several times and
    /// increments (a plan whose coefficients change with them is re-derived whenever
    /// either changes).  A branch the probe still misses is caught by the runtime's
    /// comparison with the generic path at the plan's first use on real fields.
    SplitPlan
    analyse(double time, double dt, unsigned int inc) const
    {
      SplitPlan        plan;
      std::mt19937_64  rng(0x9e3779b97f4a7c15ull ^ (uint64_t(block.id) * 2 + lhs));
      const double     tol = std::is_same_v<number, float> ? 1e-5 : 1e-12;
      auto fail = [&](const std::string &why) {
        plan.eligible = false;
        plan.reason   = why;
        return plan;
      };

      std::vector<State> states;
      for (int k = 0; k < 3; ++k)
        states.push_back(random_state(rng, 0.5, 0.4));
      for (int k = 0; k < 3; ++k)
        states.push_back(random_state(rng, 0.0, 3.0));
      for (int k = 0; k < 2; ++k)
        states.push_back(random_state(rng, 0.0, 50.0));
      std::vector<FC> base;
      for (const auto &st : states)
        base.push_back(eval(st, time, dt, inc));
      for (const auto &fc : base)
        if (fc.geometry_reads)
          return fail("the hook reads the quadrature-point location or element volume");

      // Value terms must not depend on gradients; gradient terms must vanish for zero
      // gradients (no value-dependent or constant flux).
      for (std::size_t k = 0; k < states.size(); ++k)
        {
          State st2 = states[k];
          State st0 = states[k];
          std::uniform_real_distribution<double> u(-1.0, 1.0);
          for (const auto &in : inputs)
            for (int d = 0; d < dim; ++d)
              {
                st2.g[in.field][in.slot][d] = number(2.0 * u(rng));
                st0.g[in.field][in.slot][d] = number(0);
              }
          const FC a = eval(st2, time, dt, inc);
          const FC z = eval(st0, time, dt, inc);
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
            {
              if (a.value_term[f] != base[k].value_term[f])
                return fail("a value term depends on gradients");
              for (int d = 0; d < dim; ++d)
                if (z.gradient_term[f][d] != number(0))
                  return fail("a gradient term does not vanish for zero gradients");
            }
        }

      // Coefficients beta_{f,in} from unit gradients at one state.
      int n_terms = 0;
      for (const auto &in : inputs)
        {
          if (!(in.flags & slot_gradients))
            continue;
          State st = states[0];
          for (const auto &o : inputs)
            for (int d = 0; d < dim; ++d)
              st.g[o.field][o.slot][d] = number(0);
          st.g[in.field][in.slot][0] = number(1);
          const FC r                 = eval(st, time, dt, inc);
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
            {
              if (!(out_mask & (1u << f)))
                continue;
              for (int d = 1; d < dim; ++d)
                if (r.gradient_term[f][d] != number(0))
                  return fail("a gradient term is not isotropic in its gradient inputs");
              const double beta = double(r.gradient_term[f][0]);
              if (beta == 0.0)
                continue;
              if (n_terms >= max_split_terms)
                return fail("too many gradient terms for the split path");
              plan.terms.term[n_terms++] = {f, in.field, in.slot, beta};
              if (in.slot == prismspf::dependency_slot(prismspf::SRC) && in.field == f)
                plan.beta_src[f] = beta;
            }
        }
      plan.terms.n_terms = n_terms;

      // Verify linearity with constant coefficients on every random state.
      for (std::size_t k = 0; k < states.size(); ++k)
        {
          const FC &fc = base[k];
          for (int f = 0; f < PPF_MAX_FIELDS; ++f)
            for (int d = 0; d < dim; ++d)
              {
                double pred = 0.0, scale = 0.0;
                for (int t = 0; t < n_terms; ++t)
                  if (plan.terms.term[t].out_field == f)
                    {
                      const auto  &tm = plan.terms.term[t];
                      const double gv = double(states[k].g[tm.in_field][tm.in_slot][d]);
                      pred += tm.beta * gv;
                      scale += std::abs(tm.beta * gv);
                    }
                const double got = double(fc.gradient_term[f][d]);
                if (std::abs(got - pred) > tol * (scale + std::abs(got)) + 1e-300)
                  return fail("a gradient term is not linear with constant coefficients");
              }
        }

      // Dependence of the coefficients on the simulated time or the increment.
      {
        const double       times[] = {time + 0.731, time + 13.7, 2.0 * time + 101.0};
        const unsigned int incs[]  = {inc + 1, inc + 3, 3u, 1000u};
        for (const State *st : {&states[1], &states[4]})
          {
            const FC &ref = base[st == &states[1] ? 1 : 4];
            auto      cmp = [&](const FC &a) {
              for (int f = 0; f < PPF_MAX_FIELDS; ++f)
                for (int d = 0; d < dim; ++d)
                  if (a.gradient_term[f][d] != ref.gradient_term[f][d])
                    plan.time_dependent = true;
            };
            for (const double t : times)
              cmp(eval(*st, t, dt, inc));
            for (const unsigned int i : incs)
              cmp(eval(*st, time, dt, i));
          }
      }

      // Constant derivative of the value term with respect to the own src value.
      plan.alpha_constant = true;
      constexpr int src   = prismspf::dependency_slot(prismspf::SRC);
      for (int f = 0; f < PPF_MAX_FIELDS && plan.alpha_constant; ++f)
        {
          if (!(out_mask & (1u << f)))
            continue;
          double first = 0.0;
          for (int k = 0; k < 3; ++k)
            {
              State s0 = states[k], s1 = states[k], s2 = states[k];
              s0.v[f][src] = number(0);
              s1.v[f][src] = number(1);
              s2.v[f][src] = number(2);
              const double r0 = double(eval(s0, time, dt, inc).value_term[f]);
              const double r1 = double(eval(s1, time, dt, inc).value_term[f]);
              const double r2 = double(eval(s2, time, dt, inc).value_term[f]);
              const double a  = r1 - r0;
              if (std::abs((r2 - r1) - a) > tol * (std::abs(a) + std::abs(r2)) + 1e-300 ||
                  (k > 0 && std::abs(a - first) > tol * std::abs(first) + 1e-300))
                {
                  plan.alpha_constant = false;
                  break;
                }
              if (k == 0)
                first = a;
            }
          plan.alpha[f] = first;
        }

      plan.eligible = true;
      return plan;
    }

  private:
    void
    validate(const FC &fc) const
    {
      for (int s = 0; s < PPF_N_SLOTS; ++s)
        for (int f = 0; f < PPF_MAX_FIELDS; ++f)
          {
            const bool want_v = fc.read_value_mask[s] & (1u << f);
            const bool want_g = fc.read_gradient_mask[s] & (1u << f);
            if (!want_v && !want_g)
              continue;
            unsigned int have = 0;
            for (const auto &in : inputs)
              if (in.field == f && in.slot == s)
                have |= in.flags;
            if ((want_v && !(have & slot_values)) || (want_g && !(have & slot_gradients)))
              {
                static const char *names[PPF_N_SLOTS] = {"src",   "current", "old_1",
                                                         "old_2", "old_3",   "old_4"};
                std::ostringstream msg;
                msg << "solve block " << block.id << ": access to field " << f
                    << " (dependency type " << names[s] << ", "
                    << (want_g && !(have & slot_gradients) ? "gradient" : "value")
                    << ") that was not declared as a dependency";
                throw Error(status_invalid, msg.str());
              }
          }
      const unsigned int submitted = fc.value_term_mask | fc.gradient_term_mask;
      if (submitted & ~out_mask)
        throw Error(status_invalid, "solve block " + std::to_string(block.id) +
                                      ": submission for a field not part of this solve block");
    }

    const PDE                  &pde;
    const prismspf::SolveBlock &block;
    bool                        lhs;
    std::vector<InputSlot>      inputs;
    unsigned int                out_mask = 0;
  };
} // namespace ppf
