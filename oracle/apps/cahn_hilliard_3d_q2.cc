// Oracle build of the cahn_hilliard_3d_q2 application: the application's own equations and solve
// blocks, compiled against the oracle.  TEST INFRASTRUCTURE.
#include <prismspf/core/pde_operator_base.h>

#include "../../phasefield_amd/apps/cahn_hilliard/custom_pde.h"
#include "../../phasefield_amd/apps/cahn_hilliard/app_fields.h"

struct OracleApp
{
  static constexpr int dim    = 3;
  static constexpr int degree = 2;
  template <typename number>
  using PDE = prismspf::CustomPDE<dim, degree, number>;
  using FreeEnergy                            = prismspf::CahnHilliardWell;
  static constexpr bool cahn_hilliard_adjoint = true;

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    cahn_hilliard_define(fields, blocks);
  }
};

#define ORACLE_APP OracleApp
#include "../src/oracle_capi.hpp"
