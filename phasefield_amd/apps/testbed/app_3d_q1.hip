// Test-bed application, 3-D, Q1: exercises the generic cell path (see custom_pde.h).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct Testbed3DQ1
{
  static constexpr int dim      = 3;
  static constexpr int degree   = 1;
  static constexpr int n_fields = 3;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 1, 2>;

  static const char *
  name()
  {
    return "testbed_3d_q1";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    testbed_define(fields, blocks);
  }
};

#define PPF_APP Testbed3DQ1
#include "../../csrc/capi_impl.hpp"
