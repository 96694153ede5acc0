// First include of an application translation unit.
//
// Everything the library and the standard library provide is included here, before
// the application header is compiled inside a `force_cuda_host_device` region: that
// region turns the application's own functions (its equation hooks, free energy,
// initial conditions) into host+device functions, and must not re-declare library or
// standard headers.
#pragma once

#include "adjoint.hpp"
#include "runtime.hpp"

#include <prismspf/core/dependencies.h>
#include <prismspf/core/field_attributes.h>
#include <prismspf/core/pde_operator_base.h>
#include <prismspf/core/solve_block.h>
#include <prismspf/user_inputs/user_input_parameters.h>

#include <cmath>
#include <cstdint>
#include <random>
#include <utility>
#include <vector>
