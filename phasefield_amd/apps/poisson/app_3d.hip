// Poisson, 3-D, Q2: library of the benchmark configuration "Poisson/Laplace 3D linear
// solve" (128^3 cells on the unit cube, zero Dirichlet, CG + geometric multigrid).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct Poisson3D
{
  static constexpr int dim      = 3;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 2;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 1>;

  static const char *
  name()
  {
    return "poisson_3d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    poisson_define(fields, blocks);
  }
};

#define PPF_APP Poisson3D
#include "../../csrc/capi_impl.hpp"
