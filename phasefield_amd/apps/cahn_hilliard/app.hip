// Cahn-Hilliard, 3-D, Q2: library of the benchmark configuration "Cahn-Hilliard 3D
// explicit" (64^3 cells on [0,64]^3, periodic, fields c and mu) plus a postprocessed
// free-energy density.
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct CahnHilliard3D
{
  static constexpr int dim      = 3;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 3;
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 0, 1, 2>;
  /// Adjoint entry points: the explicit Cahn-Hilliard step with this free energy.
  using FreeEnergy                            = prismspf::CahnHilliardWell;
  static constexpr bool cahn_hilliard_adjoint = true;

  static const char *
  name()
  {
    return "cahn_hilliard_3d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    cahn_hilliard_define(fields, blocks);
  }
};

#define PPF_APP CahnHilliard3D
#include "../../csrc/capi_impl.hpp"
