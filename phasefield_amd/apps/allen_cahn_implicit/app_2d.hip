// Implicit Allen-Cahn, 2-D, Q1 (backward Euler, Newton with CG): library of the
// reference's applications/allen_cahn/implicit (dim 2, degree 1, main.cc:21-22).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct AllenCahnImplicit2D
{
  static constexpr int dim      = 2;
  static constexpr int degree   = 1;
  static constexpr int n_fields = 3;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 1, 2>;

  static const char *
  name()
  {
    return "allen_cahn_implicit_2d_q1";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    allen_cahn_implicit_define(fields, blocks);
  }
};

#define PPF_APP AllenCahnImplicit2D
#include "../../csrc/capi_impl.hpp"
