// Implicit Cahn-Hilliard, 2-D, Q2 (backward Euler for (c, mu), Newton with BiCGStab):
// library of the reference's applications/cahn_hilliard/implicit (dim 2, degree 2,
// main.cc:22-23).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct CahnHilliardImplicit2D
{
  static constexpr int dim      = 2;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 3;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 2>;

  static const char *
  name()
  {
    return "cahn_hilliard_implicit_2d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    cahn_hilliard_implicit_define(fields, blocks);
  }
};

#define PPF_APP CahnHilliardImplicit2D
#include "../../csrc/capi_impl.hpp"
