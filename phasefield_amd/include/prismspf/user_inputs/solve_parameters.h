// Linear and nonlinear solver parameters of a solve block
// (reference: include/prismspf/user_inputs/solve_parameters.h:20-130 and the defaults
// declared in src/user_inputs/solve_parameters.cc:10-188).
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/type_enums.h>

#include <string>

PRISMS_PF_BEGIN_NAMESPACE

/// Chebyshev smoother / preconditioner settings.
struct ChebyshevParameters
{
  unsigned int degree              = 5;
  double       smoothing_range     = 15.0;
  unsigned int eig_cg_n_iterations = 10;
};

/// Arithmetic of the multigrid hierarchy.  The outer Krylov iteration is always FP64.
enum class LevelPrecision
{
  Double,
  Single
};

/// Solver used on the coarsest multigrid level.
enum class CoarseSolver
{
  /// One application of the level smoother (the reference's MGCoarseGridApplySmoother,
  /// reference linear_solver.h:65,145).
  Smoother,
  /// FP64 conjugate gradients to a tight relative tolerance.
  CG
};

struct LinearSolverParameters
{
  std::string         solver_type    = "cg";
  double              tolerance      = 1.0e-10;
  SolverToleranceType tolerance_type = RMSEPerField;
  unsigned int        max_iterations = 100;
  unsigned int        mg_depth       = 1;
  PreconditionerType  preconditioner = None;
  ChebyshevParameters chebyshev_parameters;
  /// deal.II SolverBicgstab::AdditionalData as the reference fills it
  /// (src/user_inputs/solve_parameters.cc:213-230).  The declared default of
  /// "breakdown" is std::to_string(numeric_limits<double>::min()), i.e. the text
  /// "0.000000", so the threshold is 0 unless a parameter file sets it.
  bool   bicgstab_exact_residual = true;
  double bicgstab_breakdown      = 0.0;
  /// SolverRichardson::AdditionalData (solve_parameters.cc:191-210).
  double richardson_omega                       = 1.0;
  bool   richardson_use_preconditioned_residual = false;
  /// SolverGMRES::AdditionalData (solve_parameters.cc:233-284).  Every orthogonalization
  /// strategy name is accepted; this path orthogonalizes by modified Gram-Schmidt (the
  /// strategies agree in exact arithmetic).
  unsigned int gmres_max_basis_size       = 30;
  bool         gmres_right_preconditioning = false;
  bool         gmres_use_default_residual  = true;
  std::string  gmres_orthogonalization     = "delayed_classical_gram_schmidt";

  // Settings of the MI355X path with no reference counterpart.
  LevelPrecision level_precision        = LevelPrecision::Double;
  CoarseSolver   coarse_solver          = CoarseSolver::Smoother;
  double         coarse_cg_tolerance    = 1.0e-12;
  unsigned int   coarse_cg_max_iterations = 200;
};

/// Newton solver settings (reference include/prismspf/user_inputs/solve_parameters.h:142-175,
/// defaults of src/user_inputs/solve_parameters.cc:352-400).
struct NonlinearSolverParameters
{
  double       step_length     = 1.0;
  unsigned int max_iterations  = 100;
  double       tolerance_value = 1.0e-10;
};

PRISMS_PF_END_NAMESPACE
