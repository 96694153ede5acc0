/*
 * C-ABI of the MI355X matrix-free operator library.
 *
 * One shared library is built per application (its custom_pde.h and the field / solve
 * block declarations of its main.cc are compiled into the device kernels); every such
 * library exports exactly these symbols.  All pointers are borrowed for the duration
 * of a call; the context owns all device memory it allocates.  No C++ exception
 * crosses this boundary: every function returns PPF_OK (0) or a negative status and
 * leaves a message for ppf_last_error().
 *
 * Field arrays exchanged with the host use deal.II's node set of the uniform mesh:
 * (degree*cells_d + 1) nodes per direction, periodic images included, lexicographic
 * with x fastest.  Device arrays passed to ppf_apply_operator use the stored layout
 * reported by ppf_get_info (periodic images omitted), one block after another for
 * multi-field solve blocks.  A vector field appears as dim component fields named
 * <name>_x, <name>_y, <name>_z (counted in ppf_info.n_fields).
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * PRISMS-PF source tree).
 */
#ifndef PPF_H
#define PPF_H

#ifdef __cplusplus
extern "C" {
#endif

#define PPF_OK 0
#define PPF_ERR_INVALID -1
#define PPF_ERR_HIP -2
#define PPF_ERR_NOT_CONVERGED -3
#define PPF_ERR_NAN -4
#define PPF_ERR_UNSUPPORTED -5
#define PPF_ERR_INTERNAL -6

/* Operator selectors of ppf_apply_operator. */
#define PPF_OP_RHS 0       /* compute_rhs (reference pde_operator_base.h:77-81)        */
#define PPF_OP_LHS 1       /* compute_lhs, src read with homogeneous constraints       */
#define PPF_OP_LHS_PLAIN 2 /* compute_lhs, src read plain (MFOperator::read_plain)    */

/* Diagonal scalings (reference mf_operator.cc:25-31). */
#define PPF_SCALE_NONE 0
#define PPF_SCALE_INVM 1
#define PPF_SCALE_INVM_SQRT 2

#define PPF_FP64 0
#define PPF_FP32 1

typedef struct ppf_ctx ppf_ctx;

typedef struct ppf_info
{
  int          dim;
  int          degree;
  int          cells[3];
  int          nodes[3];    /* stored nodes per direction (periodic images omitted) */
  int          periodic[3];
  double       lower[3];
  double       h[3];
  long long    n_nodes;     /* stored nodes per field                              */
  long long    n_dofs;      /* deal.II n_dofs per field: prod(degree*cells_d + 1)  */
  int          n_fields;
  int          n_blocks;
  double       dt;
  double       time;
  unsigned int increment;
  unsigned int n_increments;
} ppf_info;

typedef struct ppf_solve_stats
{
  int    iterations;
  double initial_residual;
  double final_residual;
  double tolerance;
  int    converged;
} ppf_solve_stats;

/* Name of the application compiled into this library. */
const char *ppf_app_name(void);

/* Builds a context from the text of a .prm parameter file: parses it
 * (replaces UserInputParameters<dim>(filename), user_input_parameters.h:39-51),
 * constructs the mesh descriptor, the application's fields and solve blocks and the
 * CustomPDE object, allocates device storage and evaluates the Dirichlet values
 * (replaces Problem::init_system, src/core/problem.cc:227-338). */
int ppf_create(const char *prm_text, ppf_ctx **out);

void ppf_destroy(ppf_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed ppf_create if ctx is
 * NULL). */
const char *ppf_last_error(const ppf_ctx *ctx);

int ppf_get_info(ppf_ctx *ctx, ppf_info *info);

/* Copies the name of field `field` into buf (NUL-terminated, at most len bytes). */
int ppf_field_name(ppf_ctx *ctx, int field, char *buf, int len);

/* Options: "operator_path" = auto|generic|split|split_node (split: tiled stencil
 * kernel, split_node: one thread per node); "stencil_waves" = 0|4|8,
 * "stencil_rows" = 0|1|2|4 (waves x rows: 4x2, 4x4, 8x1 or 8x2) and
 * "stencil_z_chunk" = <planes> (0: automatic) tune the tiled kernel's workgroup
 * shape; "stencil_min_nodes" = <n> (default 100000): smaller lattices run split
 * operators with the one-thread-per-node kernel; "stencil_store_policy" = auto|temporal|nontemporal sets the cache policy of
 * its output stores (auto: non-temporal when the launch's arrays exceed 192 MiB);
 * "generic_kernel" = cell_block|two_pass selects the generic path's kernels (cell_block:
 * one launch, cell blocks staged in shared memory, no device scratch; two_pass: a cell
 * kernel into a device scratch array and a node gather; cell_block falls back to
 * two_pass when a block's staged inputs exceed the shared memory of a workgroup);
 * "generic_column_chunk" = <layers> sets the z layers each 3-D cell column marches
 * over (0: automatic, enough columns for every compute unit); "mg_graphs" = 1|0 (1:
 * multigrid V-cycles that need no host synchronization are captured into a HIP graph
 * and replayed);
 * "synchronize" = 1|0 (0 leaves ppf_apply_operator asynchronous on the
 * context's stream); "write_output" = 0|1 (1: at every output increment ppf_run and
 * ppf_solve_increment write the solution file of the `output` parameters,
 * <directory>/<file name>_<increment>.vtu, as SolutionOutput does, problem.cc:500-515);
 * "mg_level_precision" = double|single, "mg_coarse_solver" = smoother|cg,
 * "linear_tolerance" = <value> and "linear_max_iterations" = <n> (applied to every
 * linear block, overriding the parameter file); "linear_initial_guess" =
 * previous|zero (previous: the last solution, as the reference); "explicit_precision" =
 * double|single|single_copies (single: explicit blocks compute in FP32 on the FP64
 * state: the stencil kernel rounds its FP64 loads to FP32 in registers and stores FP64,
 * after a first increment through FP32 copies that checks the split operator; blocks
 * the stencil kernel does not take, and single_copies always, compute from FP32 copies
 * of their inputs and convert the results into the FP64 state); "explicit_pairs" = 0|1
 * (1: two consecutive explicit 3-D blocks, the second reading the first's output through
 * its stencil, in one launch, with results identical to the two launches; off by
 * default, being slower than the two launches), "pair_rows" = 2|3 and "pair_z_chunk" =
 * <n> (its tile height and z chunk, 0: automatic); "pipeline_depth" = 1|2|3 (explicit
 * increments ppf_run enqueues ahead of the oldest unread NaN probe, default 2). */
int ppf_set_option(ppf_ctx *ctx, const char *key, const char *value);

/* Field state `age` (0 = current, k = old_k, PPF_AGE_LATEST = the newest values) from/to
 * host memory over deal.II's node set (replaces SolutionIndexer / GroupSolutionHandler
 * access, src/core/solution_indexer.cc:12-104).  Every increment ends with the old-state
 * rotation, after which the newest values of a field with old states are in old_1, as in
 * the reference; PPF_AGE_LATEST resolves that. */
#define PPF_AGE_LATEST -1
int ppf_set_field(ppf_ctx *ctx, int field, int age, const double *host, long long n);
int ppf_get_field(ppf_ctx *ctx, int field, int age, double *host, long long n);

/* Device pointer of a field state in the stored layout. */
int ppf_field_device_ptr(ppf_ctx *ctx, int field, int age, void **ptr);

/* Problem::solve (src/core/problem.cc:340-412): increments 0..n_increments.
 * exit_status: 0 normal, 2 NaN detected, 3 user stop. */
int ppf_run(ppf_ctx *ctx, unsigned int n_increments, int *exit_status);

/* Problem::solve_increment (src/core/problem.cc:414-558) followed by the time
 * increment. */
int ppf_solve_increment(ppf_ctx *ctx, int *exit_status);

/* SolverBase::solve of one block at the current increment
 * (include/prismspf/solvers/solver_base.h:106-130): initial condition at increment 0
 * for Primary blocks, otherwise ExplicitSolver::solve_impl (explicit_solver.h:63-77)
 * or LinearSolver::solve_impl (linear_solver.h:305-345). */
int ppf_solve_block(ppf_ctx *ctx, int block_id);

/* GroupSolutionHandler::update (src/core/group_solution_handler.cc:369-403). */
int ppf_rotate_old(ppf_ctx *ctx, int block_id);

/* SimulationTimer state. */
int ppf_set_time(ppf_ctx *ctx, unsigned int increment, double time);

/* MFOperator::compute_operator (src/solvers/mf_operator.cc:17-32) of a block's rhs or
 * lhs on device arrays; other dependencies are read from the context's fields.
 * Constrained (Dirichlet) rows of dst are zero.  `precision` selects the FP64 or FP32
 * instantiation of the equations; arrays are then double or float. */
int ppf_apply_operator(ppf_ctx *ctx, int block_id, int which, int scaling, int precision,
                       const void *d_src, void *d_dst);

/* ppf_apply_operator without any device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results and status
 * codes as ppf_apply_operator; the scratch the operator path needs (the generic cell
 * path's per-cell results, the zero operand of an operator without gradient terms,
 * FP32 copies of the fields an FP32 application reads) comes from a static device array
 * of PPF_NOALLOC_SCRATCH_BYTES per GPU and application library, held until the process
 * ends.  A launch that needs more fails with PPF_ERR_INVALID and a message naming the
 * size (e.g. the generic path needs fields x (degree+1)^dim x 8 B per cell in FP64).
 * Every call on a GPU shares that array: calls that overlap on one GPU (other threads,
 * other streams) race.  The context itself (ppf_create) still allocates as usual. */
#define PPF_NOALLOC_SCRATCH_BYTES (64ll << 20)
int ppf_apply_operator_noalloc(ppf_ctx *ctx, int block_id, int which, int scaling,
                               int precision, const void *d_src, void *d_dst);

/* MFOperator::compute_diagonal + eval_matrix_diagonal (src/solvers/mf_operator.cc:82-255):
 * inverse of the lhs diagonal with zero entries replaced by one, on a device array. */
int ppf_inverse_diagonal(ppf_ctx *ctx, int block_id, int scaling, int precision, void *d_dst);

/* LinearSolver::solve_impl of a linear block, returning the solver statistics. */
int ppf_linear_solve(ppf_ctx *ctx, int block_id, ppf_solve_stats *stats);

int ppf_last_solve_stats(ppf_ctx *ctx, int block_id, ppf_solve_stats *stats);

/* Statistics of a Newton block's last solve (NewtonSolver::solve_impl,
 * include/prismspf/solvers/newton_solver.h:73-156; the quantities of its log line,
 * :140-147): Newton updates applied, CG steps over all of them, the l2 norm of the last
 * residual evaluated and the tolerance (tolerance value times the linear solver's
 * normalization).  ppf_last_solve_stats of the block reports its last linear solve. */
typedef struct ppf_newton_stats
{
  int    iterations;
  int    linear_iterations;
  double residual;
  double tolerance;
  int    converged;
} ppf_newton_stats;

int ppf_last_newton_stats(ppf_ctx *ctx, int block_id, ppf_newton_stats *stats);

/* Output diagnostics of problem.cc:517-549: l2-norm over deal.II's node set and the
 * integral sum(u_i * jxw_i). */
int ppf_field_norms(ppf_ctx *ctx, int field, int age, double *l2, double *integral);

/* Accumulated output log ("Iteration ... Field index ... l2-norm ..."). */
const char *ppf_output_log(ppf_ctx *ctx);

/* The HIP stream all work of the context is issued on (hipStream_t). */
void *ppf_stream(ppf_ctx *ctx);

int ppf_synchronize(ppf_ctx *ctx);

/* Adjoint of the application's explicit step.  No reference counterpart: the reference
 * has no adjoint.  Available in applications that declare a free energy for it;
 * otherwise PPF_ERR_UNSUPPORTED.  Arrays are device arrays in the stored layout; K is the
 * assembled stiffness matrix and M the lumped mass.
 *
 * Allen-Cahn applications (an explicit Allen-Cahn block, reference applications/
 * allen_cahn/explicit/custom_pde.h:72-87, advanced by ExplicitSolver,
 * explicit_solver.h:63-77): phi is the order parameter, L the mobility and
 *   phi' = F(phi) = phi - dt*L*f'(phi) - dt*L*kappa * M^{-1} K phi.
 * Cahn-Hilliard applications (the c and mu blocks, reference applications/cahn_hilliard/
 * explicit/custom_pde.h:92-118): phi is the concentration c, the parameter L is the
 * mobility M_c, and with mu = f'(c) + kappa * M^{-1} K c (computed from the initial c
 * and after every c update, as the mu block does)
 *   c' = F(c) = c - dt*M_c * M^{-1} K mu.
 *
 * ppf_adjoint_forward: phi_final = F^n_steps(phi0) with parameters (L, kappa);
 * precision PPF_FP32 evaluates the step in FP32 (state kept in FP64).
 *
 * ppf_adjoint_vjp: given lam_final = dJ/dphi_N for any loss J, returns dJ/dL,
 * dJ/dkappa and dJ/dphi0, running the adjoint steps backwards in time and recomputing
 * forward states from checkpoints stored every `checkpoint_every` steps (Cahn-Hilliard:
 * the segment keeps c and mu of every step).  The
 * checkpoints (ceil(n_steps / checkpoint_every) states) and the recomputed segment
 * (checkpoint_every states) stay allocated in the context for later calls, until
 * ppf_destroy.
 *
 * ppf_lumped_mass: the lumped mass m_i of every stored node (for losses such as
 * J = 1/2 sum_i m_i (phi_N,i - target_i)^2). */
int ppf_adjoint_forward(ppf_ctx *ctx, const double *d_phi0, double L, double kappa,
                        unsigned int n_steps, int precision, double *d_phi_final);
int ppf_adjoint_vjp(ppf_ctx *ctx, const double *d_phi0, const double *d_lam_final, double L,
                    double kappa, unsigned int n_steps, unsigned int checkpoint_every,
                    int precision, double *dL, double *dkappa, double *d_dphi0);
int ppf_lumped_mass(ppf_ctx *ctx, double *d_mass);

#ifdef __cplusplus
}
#endif

#endif /* PPF_H */
