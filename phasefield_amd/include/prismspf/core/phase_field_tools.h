// Shared tools handed to the equations (reference: include/prismspf/core/phase_field_tools.h).
// Nucleation is outside the matrix-free path, so the nucleus list stays empty.
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/lane_types.h>

#include <vector>

PRISMS_PF_BEGIN_NAMESPACE

template <unsigned int dim>
struct Nucleus
{
  dealii::Point<dim> location;
  double             seed_time      = 0.0;
  unsigned int       seed_increment = 0;
};

template <unsigned int dim>
struct PhaseFieldTools
{
  std::vector<Nucleus<dim>> nuclei_list;
};

PRISMS_PF_END_NAMESPACE
