// Enumerations of the application API.  Names and enumerator values follow the
// reference so that application code compiles unchanged
// (reference: include/prismspf/core/type_enums.h:16-147).
#pragma once

#include <prismspf/config.h>

PRISMS_PF_BEGIN_NAMESPACE

/// How a solve block advances its fields (reference type_enums.h:16-44).
enum SolveType
{
  Constant,
  Explicit,
  Linear,
  Newton
};

/// Rank of a field.  A vector field is stored as dim scalar component fields.
enum TensorRank
{
  Undefined = static_cast<unsigned int>(-1),
  Scalar    = 0,
  Vector    = 1
};

/// Which state of a field an accessor reads.  SRC is the operand of a linear operator
/// (aliases LHS, Trial, Change, Solution); Current is the latest solution; OldOne..OldFour
/// are earlier increments (reference type_enums.h:90-104).
enum DependencyType
{
  DST      = -2,
  SRC      = -1,
  LHS      = SRC,
  Solution = SRC,
  Trial    = SRC,
  Change   = SRC,
  Current  = 0,
  OldOne,
  OldTwo,
  OldThree,
  OldFour
};

/// Stopping rule of the linear solver (reference type_enums.h:109-135).
enum SolverToleranceType
{
  AbsoluteResidual,
  RMSEPerField,
  IntegratedPerField,
  RMSETotal,
  IntegratedTotal
};

/// Preconditioner of the linear solver (reference type_enums.h:140-145).
enum PreconditionerType
{
  None,
  Chebyshev,
  GMG
};

/// Register slot of a dependency type inside the device FieldContainer.
PPF_HD constexpr int
dependency_slot(DependencyType type)
{
  return type == DST ? -1 : int(type) + 1;
}

PRISMS_PF_END_NAMESPACE
