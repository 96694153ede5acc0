// Oracle build of the switching_2d_q2 application: the application's own equations and
// solve blocks, compiled against the oracle.  TEST INFRASTRUCTURE.
#include <prismspf/core/pde_operator_base.h>

#include "../../phasefield_amd/apps/switching/custom_pde.h"
#include "../../phasefield_amd/apps/switching/app_fields.h"

struct OracleApp
{
  static constexpr int dim    = 2;
  static constexpr int degree = 2;
  template <typename number>
  using PDE = prismspf::CustomPDE<dim, degree, number>;

  static void
  define(std::vector<prismspf::FieldAttributes> &fields, std::vector<prismspf::SolveBlock> &blocks)
  {
    switching_define(fields, blocks);
  }
};

#define ORACLE_APP OracleApp
#include "../src/oracle_capi.hpp"
