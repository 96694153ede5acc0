// Linear elasticity, 3-D, Q2: the reference's mechanics boundary-value problem (one
// vector field, stored as three component fields; CG on the linear block).
#include "../../csrc/app_prelude.hpp"

#pragma clang force_cuda_host_device begin
#include "custom_pde.h"
#pragma clang force_cuda_host_device end

#include "app_fields.h"

struct Mechanics3D
{
  static constexpr int dim      = 3;
  static constexpr int degree   = 2;
  static constexpr int n_fields = 3; // stored component fields u_x, u_y, u_z
  template <typename number>
  using PDE      = prismspf::CustomPDE<dim, degree, number>;
  using BlockIds = std::integer_sequence<int, 1>;

  static const char *
  name()
  {
    return "mechanics_3d_q2";
  }

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    mechanics_define(fields, blocks);
  }
};

#define PPF_APP Mechanics3D
#include "../../csrc/capi_impl.hpp"
