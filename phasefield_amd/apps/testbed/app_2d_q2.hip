// Test-bed application, 2-D, Q2: exercises the generic cell path (see custom_pde.h).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct Testbed2DQ2
{
  static constexpr int dim      = 2;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 3;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 1, 2>;

  static const char *
  name()
  {
    return "testbed_2d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    testbed_define(fields, blocks);
  }
};

#define PPF_APP Testbed2DQ2
#include "../../csrc/capi_impl.hpp"
