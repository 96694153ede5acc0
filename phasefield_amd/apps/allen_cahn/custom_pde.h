// Allen-Cahn equation with an explicit Euler step, written against PRISMS-PF's
// application API (same structure as the reference's
// applications/allen_cahn/explicit/custom_pde.h:72-101, with the free energy
// f(phi) = phi^4/4 - phi^2/2 of the benchmark configuration).
//
//   d phi / dt = -L (f'(phi) - kappa Laplace(phi))
//   weak explicit form:  value term  phi - dt L f'(phi)
//                        gradient term -dt L kappa grad(phi)
//
// Model constants: L (mobility), kappa (gradient-energy coefficient), n_circles,
// radius_min, radius_max, interface_width, noise_amplitude.
#include <prismspf/core/pde_operator_base.h>

#include "../common/seeded_noise.h"

#include <cmath>

PRISMS_PF_BEGIN_NAMESPACE

/// Double-well f(phi) = phi^4/4 - phi^2/2 and its derivatives.
struct QuarticWell
{
  template <typename T>
  PPF_HD static T
  fprime(const T &p)
  {
    return p * p * p - p;
  }

  template <typename T>
  PPF_HD static T
  fprime2(const T &p)
  {
    return 3.0 * p * p - 1.0;
  }
};

template <unsigned int dim, unsigned int degree, typename number>
class CustomPDE : public PDEOperatorBase<dim, degree, number>
{
public:
  using ScalarValue = dealii::VectorizedArray<number>;
  using ScalarGrad  = dealii::Tensor<1, dim, ScalarValue>;
  using PDEOperatorBase<dim, degree, number>::get_user_inputs;

  CustomPDE(const UserInputParameters<dim> &_user_inputs, PhaseFieldTools<dim> &_pf_tools)
    : PDEOperatorBase<dim, degree, number>(_user_inputs, _pf_tools)
    , L(get_user_inputs().user_constants.get_double("L"))
    , kappa(get_user_inputs().user_constants.get_double("kappa"))
  {
    const auto &uc       = get_user_inputs().user_constants;
    const auto &mesh     = get_user_inputs().spatial_discretization.rectangular_mesh;
    n_circles            = uc.contains("n_circles") ? uc.get_int("n_circles") : 8;
    const double rmin    = uc.contains("radius_min") ? uc.get_double("radius_min") : 5.0;
    const double rmax    = uc.contains("radius_max") ? uc.get_double("radius_max") : 15.0;
    width                = uc.contains("interface_width") ? uc.get_double("interface_width") : 1.0;
    noise                = uc.contains("noise_amplitude") ? uc.get_double("noise_amplitude") : 0.05;
    seed                 = get_user_inputs().misc_parameters.random_seed;
    for (unsigned int d = 0; d < dim; ++d)
      {
        lower[d]    = mesh.lower_bound[d];
        size[d]     = mesh.size[d];
        cells[d]    = get_user_inputs().spatial_discretization.n_cells(d);
        periodic[d] = mesh.periodic_directions.count(d) != 0;
      }
    for (int k = 0; k < n_circles && k < max_circles; ++k)
      {
        for (unsigned int d = 0; d < dim; ++d)
          centre[k][d] = ppf_noise::uniform(seed, 1, k * 4 + d, lower[d], size[d]);
        radius[k] = ppf_noise::uniform(seed, 1, k * 4 + 3, rmin, rmax);
      }
  }

  void
  set_initial_condition([[maybe_unused]] const unsigned int       &index,
                        [[maybe_unused]] const unsigned int       &component,
                        [[maybe_unused]] const dealii::Point<dim> &point,
                        [[maybe_unused]] number                   &scalar_value,
                        [[maybe_unused]] number &vector_component_value) const override
  {
    double value = 0.0;
    for (int k = 0; k < n_circles && k < max_circles; ++k)
      {
        double r2 = 0.0;
        for (unsigned int d = 0; d < dim; ++d)
          r2 += (point[d] - centre[k][d]) * (point[d] - centre[k][d]);
        value += 0.5 * (1.0 - std::tanh((std::sqrt(r2) - radius[k]) / width));
      }
    double x[3] = {0, 0, 0};
    for (unsigned int d = 0; d < dim; ++d)
      x[d] = point[d];
    const uint64_t key = ppf_noise::node_key<dim>(x, lower, size, cells, degree, periodic);
    value += ppf_noise::uniform(seed, 2, key, -noise, noise);
    scalar_value = number(value);
  }

  void
  compute_rhs(FieldContainer<dim, degree, number> &variable_list,
              const SimulationTimer               &sim_timer,
              unsigned int                         solve_block_id) const override
  {
    if (solve_block_id == 1)
      {
        ScalarValue n      = variable_list.template get_value<Scalar, OldOne>(0);
        ScalarGrad  n_grad = variable_list.template get_gradient<Scalar, OldOne>(0);
        ScalarValue eq_n   = n - sim_timer.get_timestep() * L * QuarticWell::fprime(n);
        ScalarGrad  eqx_n  = -sim_timer.get_timestep() * kappa * L * n_grad;
        variable_list.set_value_term(0, eq_n);
        variable_list.set_gradient_term(0, eqx_n);
      }
    else if (solve_block_id == 2)
      {
        // Free-energy density f(phi) + kappa/2 |grad phi|^2 (postprocess).
        ScalarValue n      = variable_list.template get_value<Scalar, Current>(0);
        ScalarGrad  n_grad = variable_list.template get_gradient<Scalar, Current>(0);
        ScalarValue f_chem = 0.25 * n * n * n * n - 0.5 * n * n;
        ScalarValue f_grad = 0.5 * kappa * n_grad.norm_square();
        variable_list.set_value_term(1, f_chem + f_grad);
      }
  }

private:
  static constexpr int max_circles = 32;
  number               L;
  number               kappa;
  int                  n_circles = 8;
  double               width     = 1.0;
  double               noise     = 0.05;
  uint64_t             seed      = 2025;
  double               centre[max_circles][3] {};
  double               radius[max_circles] {};
  double               lower[3] {};
  double               size[3] {};
  unsigned int         cells[3] {};
  bool                 periodic[3] {};
};

PRISMS_PF_END_NAMESPACE
