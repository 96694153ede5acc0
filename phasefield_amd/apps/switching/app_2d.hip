// Routing test bed, 2-D, Q2: hooks whose flux coefficients branch on a value or on the
// increment (see custom_pde.h).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct Switching2D
{
  static constexpr int dim      = 2;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 4;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 1, 2, 3>;

  static const char *
  name()
  {
    return "switching_2d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    switching_define(fields, blocks);
  }
};

#define PPF_APP Switching2D
#include "../../csrc/capi_impl.hpp"
