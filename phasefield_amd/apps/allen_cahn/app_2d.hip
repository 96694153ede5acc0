// Allen-Cahn, 2-D, Q1: library of the benchmark configuration "Allen-Cahn 2D explicit"
// (256x256 cells on [0,100]^2, periodic).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct AllenCahn2D
{
  static constexpr int dim    = 2;
  static constexpr int degree   = 1;
  static constexpr int n_fields = 2;
  template <typename number>
  using PDE        = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds   = std::integer_sequence<int, 1, 2>;
  using FreeEnergy = prismspf::QuarticWell;

  static const char *
  name()
  {
    return "allen_cahn_2d_q1";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    allen_cahn_define(fields, blocks);
  }
};

#define PPF_APP AllenCahn2D
#include "../../csrc/capi_impl.hpp"
