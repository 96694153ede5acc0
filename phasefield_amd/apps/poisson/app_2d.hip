// Poisson, 2-D, Q2: the 2-D build of the Poisson application, used by the small parity
// cases of the linear solver (the benchmark configuration is 3-D).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct Poisson2D
{
  static constexpr int dim      = 2;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 2;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 1>;

  static const char *
  name()
  {
    return "poisson_2d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    poisson_define(fields, blocks);
  }
};

#define PPF_APP Poisson2D
#include "../../csrc/capi_impl.hpp"
