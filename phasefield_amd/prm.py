"""Parameter files (PRISMS-PF ``.prm`` text) of the benchmark configurations.

`config_prm(name, **overrides)` returns the text for one of the five configurations
with optional size overrides (used by the tests to run the same model on small
meshes).  Keys follow the reference's parameter names
(src/user_inputs/{spatial_discretization,temporal_discretization,solve_parameters,
constraint_parameters,io_parameters,miscellaneous_parameters}.cc).
"""

from __future__ import annotations

# name -> (application library, description)
CONFIGS = {
    "allen_cahn_2d": ("allen_cahn_2d_q1", "Allen-Cahn 2D explicit, Q1, 256^2 cells on [0,100]^2"),
    "cahn_hilliard_3d": ("cahn_hilliard_3d_q2", "Cahn-Hilliard 3D explicit, Q2, 64^3 cells on [0,64]^3"),
    "poisson_3d": ("poisson_3d_q2", "Poisson 3D, Q2, 128^3 cells on the unit cube, CG + GMG"),
    "allen_cahn_adjoint_3d": ("allen_cahn_3d_q1", "Allen-Cahn 3D adjoint, Q1, 128^3 cells on [0,128]^3"),
    "precipitate_3d": ("precipitate_3d_q1", "Coupled AC/CH precipitate 3D, Q1, 192^3 cells on [0,192]^3"),
}


def mesh_block(dim, size, refinement, subdivisions=None, periodic=True, lower=None):
    axes = "xyz"[:dim]
    subdivisions = subdivisions or [1] * dim
    lower = lower or [0.0] * dim
    size = size if isinstance(size, (list, tuple)) else [size] * dim
    lines = ["subsection Rectangular mesh"]
    for d, a in enumerate(axes):
        lines.append(f"  set {a} size = {size[d]!r}")
        if lower[d]:
            lines.append(f"  set {a} lower bound = {lower[d]!r}")
        lines.append(f"  set {a} subdivisions = {subdivisions[d]}")
        lines.append(f"  set {a} periodic = {'true' if periodic else 'false'}")
    lines.append("end")
    lines.append(f"set global refinement = {refinement}")
    return "\n".join(lines)


def _common(steps, dt, outputs=1, seed=2025):
    return "\n".join([
        f"set time step = {dt!r}",
        f"set number steps = {steps}",
        "subsection output",
        "  set condition = EQUAL_SPACING",
        f"  set number = {outputs}",
        "end",
        f"set random seed = {seed}",
    ])


def allen_cahn(dim=2, size=100.0, refinement=8, steps=1000, dt=0.005, L=1.0, kappa=1.5,
               subdivisions=None, outputs=1, n_circles=8):
    return "\n".join([
        mesh_block(dim, size, refinement, subdivisions),
        _common(steps, dt, outputs),
        f"set Model constant L = {L!r}, DOUBLE",
        f"set Model constant kappa = {kappa!r}, DOUBLE",
        f"set Model constant n_circles = {n_circles}, INT",
        "set Model constant radius_min = 5.0, DOUBLE",
        "set Model constant radius_max = 15.0, DOUBLE",
        "set Model constant interface_width = 1.0, DOUBLE",
        "set Model constant noise_amplitude = 0.05, DOUBLE",
    ]) + "\n"


def allen_cahn_implicit(size=100.0, refinement=8, steps=10000, dt=1.0e-2, m_well=1.0, kappa=2.0,
                        linear_tolerance=1.0e-6, newton_tolerance=1.0e-6, newton_max_iterations=100,
                        step_size=1.0, preconditioner=None, subdivisions=None, outputs=5,
                        periodic=False, tolerance_type="RMSEPerField", mg_depth=None,
                        solver="cg", linear_extra=()):
    """The implicit Allen-Cahn application's parameters (defaults: the reference's
    applications/allen_cahn/implicit/parameters.prm on a uniform mesh, 2-D)."""
    lines = [
        mesh_block(2, size, refinement, subdivisions, periodic=periodic),
        _common(steps, dt, outputs),
        f"set Model constant m_well = {m_well!r}, DOUBLE",
        f"set Model constant kappa = {kappa!r}, DOUBLE",
        "subsection linear solver parameters: 0",
        "  set solver_ids = 1",
        f"  set solver type = {solver}",
        "  set max iterations = 1000",
        f"  set tolerance type = {tolerance_type}",
        f"  set tolerance value = {linear_tolerance!r}",
    ]
    if preconditioner:
        lines.append(f"  set preconditioner type = {preconditioner}")
        if mg_depth is not None:
            lines.append(f"  set mg depth = {mg_depth}")
    lines += list(linear_extra)
    lines += [
        "end",
        "subsection newton solver parameters: 0",
        "  set solver_ids = 1",
        f"  set max iterations = {newton_max_iterations}",
        f"  set step size = {step_size!r}",
        f"  set tolerance value = {newton_tolerance!r}",
        "end",
    ]
    return "\n".join(lines) + "\n"


def cahn_hilliard_implicit(size=100.0, refinement=7, steps=40000, dt=5.0e-4, McV=1.0, KcV=1.5,
                           WcV=16.0, ic_type=0, c0=0.5, icamplitude=0.1, solver="bicgstab",
                           linear_tolerance=1.0e-6, newton_tolerance=1.0e-6,
                           newton_max_iterations=100, exact_residual=True, preconditioner=None,
                           subdivisions=None, outputs=500, periodic=False,
                           tolerance_type="RMSEPerField", mg_depth=None, linear_extra=()):
    """The implicit Cahn-Hilliard application's parameters (defaults: the reference's
    applications/cahn_hilliard/implicit/parameters.prm on a uniform mesh, 2-D; end time
    20 = 40000 steps of 5e-4)."""
    lines = [
        mesh_block(2, size, refinement, subdivisions, periodic=periodic),
        _common(steps, dt, outputs),
        f"set Model constant McV = {McV!r}, DOUBLE",
        f"set Model constant KcV = {KcV!r}, DOUBLE",
        f"set Model constant WcV = {WcV!r}, DOUBLE",
        f"set Model constant ic_type = {ic_type}, INT",
        f"set Model constant c0 = {c0!r}, DOUBLE",
        f"set Model constant icamplitude = {icamplitude!r}, DOUBLE",
        "subsection linear solver parameters: 0",
        "  set solver_ids = 0",
        f"  set solver type = {solver}",
        "  set max iterations = 10000",
        f"  set tolerance type = {tolerance_type}",
        f"  set tolerance value = {linear_tolerance!r}",
    ]
    if preconditioner:
        lines.append(f"  set preconditioner type = {preconditioner}")
        if mg_depth is not None:
            lines.append(f"  set mg depth = {mg_depth}")
    lines += list(linear_extra)
    lines += [
        "  subsection BiCGStab",
        f"    set exact residual = {'true' if exact_residual else 'false'}",
        "  end",
        "end",
        "subsection newton solver parameters: 0",
        "  set solver_ids = 0",
        f"  set max iterations = {newton_max_iterations}",
        "  set step size = 1.0",
        f"  set tolerance value = {newton_tolerance!r}",
        "end",
    ]
    return "\n".join(lines) + "\n"


def cahn_hilliard(size=64.0, refinement=6, steps=200, dt=1e-4, M=1.0, kappa=1.0, outputs=1,
                  subdivisions=None):
    return "\n".join([
        mesh_block(3, size, refinement, subdivisions),
        _common(steps, dt, outputs),
        f"set Model constant M = {M!r}, DOUBLE",
        f"set Model constant kappa = {kappa!r}, DOUBLE",
        "set Model constant c0 = 0.5, DOUBLE",
        "set Model constant noise_amplitude = 0.01, DOUBLE",
    ]) + "\n"


def poisson(dim=3, refinement=7, mg_depth=7, tolerance=1e-10, preconditioner="GMG",
            smoother_degree=4, max_iterations=1000, level_precision="double",
            coarse_solver="cg", size=1.0, tolerance_type="AbsoluteResidual", subdivisions=None):
    faces = ", ".join(["Dirichlet"] * (2 * dim))
    return "\n".join([
        mesh_block(dim, size, refinement, subdivisions, periodic=False),
        _common(0, 0.0),
        "subsection boundary conditions: 0",
        "  set variables = u",
        f"  set conditions = {faces}",
        "end",
        "subsection linear solver parameters: 0",
        "  set solver_ids = 1",
        f"  set tolerance type = {tolerance_type}",
        f"  set tolerance value = {tolerance!r}",
        f"  set max iterations = {max_iterations}",
        f"  set preconditioner = {preconditioner}",
        f"  set mg depth = {mg_depth}",
        f"  set mg level precision = {level_precision}",
        f"  set mg coarse solver = {coarse_solver}",
        "  subsection Chebyshev",
        f"    set smoother degree = {smoother_degree}",
        "    set smoothing range = 15.0",
        "    set eigenvalue cg iterations = 10",
        "  end",
        "end",
    ]) + "\n"


def precipitate(size=192.0, refinement=6, subdivisions=(3, 3, 3), steps=1000, dt=5e-3,
                outputs=1, n_nuclei=20, radius=4.0):
    return "\n".join([
        mesh_block(3, size, refinement, list(subdivisions)),
        _common(steps, dt, outputs),
        "set Model constant W = 1.0, DOUBLE",
        "set Model constant M = 1.0, DOUBLE",
        "set Model constant L = 1.0, DOUBLE",
        "set Model constant kappa_c = 1.0, DOUBLE",
        "set Model constant kappa_eta = 1.0, DOUBLE",
        "set Model constant c0 = 0.3, DOUBLE",
        "set Model constant noise_amplitude = 0.01, DOUBLE",
        f"set Model constant n_nuclei = {n_nuclei}, INT",
        f"set Model constant nucleus_radius = {radius!r}, DOUBLE",
    ]) + "\n"


def testbed(dim=2, refinement=3, size=4.0, steps=2, dt=1e-3, dirichlet_faces=(0, 3),
            tolerance=1e-10, preconditioner="None", mg_depth=2, periodic=False,
            time_dependent=False):
    conds = ["Natural"] * (2 * dim)
    for f in dirichlet_faces:
        conds[f] = "Dirichlet"
    return "\n".join([
        mesh_block(dim, size, refinement, periodic=periodic),
        _common(steps, dt, 1),
        "subsection boundary conditions: 0",
        "  set variables = c, v",
        f"  set conditions = {', '.join(conds)}",
        f"  set time dependent = {'true' if time_dependent else 'false'}",
        "end",
        "subsection linear solver parameters: 0",
        "  set solver_ids = 1",
        "  set tolerance type = AbsoluteResidual",
        f"  set tolerance value = {tolerance!r}",
        "  set max iterations = 500",
        f"  set preconditioner = {preconditioner}",
        f"  set mg depth = {mg_depth}",
        "  set mg coarse solver = cg",
        "end",
        "set Model constant D0 = 0.7, DOUBLE",
        "set Model constant alpha = 0.5, DOUBLE",
        "set Model constant amplitude = 0.3, DOUBLE",
    ]) + "\n"


def mechanics(size=100.0, refinement=6, E=2.0, nu=0.3, tolerance=1e-10,
              tolerance_type="RMSEPerField", preconditioner=None, max_iterations=10000,
              mg_depth=None, conditions=("Dirichlet", "Dirichlet", "Natural", "Dirichlet", "Natural",
                          "Dirichlet")):
    """The linear elasticity application's parameters (defaults: the reference's
    applications/mechanics/boundary_value_problem/parameters.prm)."""
    lines = [
        mesh_block(3, size, refinement, periodic=False),
        _common(0, 0.0),
        "subsection linear solver parameters: 0",
        "  set solver_ids = 1",
        f"  set max iterations = {max_iterations}",
        f"  set tolerance type = {tolerance_type}",
        f"  set tolerance value = {tolerance!r}",
    ]
    if preconditioner:
        lines.append(f"  set preconditioner = {preconditioner}")
    if mg_depth is not None:
        lines.append(f"  set mg depth = {mg_depth}")
    lines += [
        "end",
        "subsection boundary conditions: 0",
        "  set variables = u:x, u:y, u:z",
        f"  set conditions = {', '.join(conditions)}",
        "end",
        f"set Model constant stiffness = ({E!r},{nu!r}), isotropic elastic constants",
    ]
    return "\n".join(lines) + "\n"


def switching(size=8.0, refinement=4, steps=6, dt=2e-3, D=1.0, c_low=-10.0, c_high=70.0,
              outputs=1):
    """Routing test bed (apps/switching): the initial c ramps from c_low to c_high along
    x, so the value branches of its flux coefficients (c > 1, c > 60) are taken in part of
    the domain."""
    return "\n".join([
        mesh_block(2, size, refinement, periodic=False),
        _common(steps, dt, outputs),
        f"set Model constant D = {D!r}, DOUBLE",
        f"set Model constant c_low = {c_low!r}, DOUBLE",
        f"set Model constant c_high = {c_high!r}, DOUBLE",
    ]) + "\n"


def config_prm(name: str, **kw) -> str:
    """Parameter text of a benchmark configuration (keyword overrides shrink it)."""
    if name == "allen_cahn_2d":
        return allen_cahn(dim=2, **kw)
    if name == "cahn_hilliard_3d":
        return cahn_hilliard(**kw)
    if name == "poisson_3d":
        return poisson(**kw)
    if name == "allen_cahn_adjoint_3d":
        kw.setdefault("size", 128.0)
        kw.setdefault("refinement", 7)
        kw.setdefault("steps", 500)
        kw.setdefault("dt", 0.01)
        kw.setdefault("kappa", 1.5)
        return allen_cahn(dim=3, **kw)
    if name == "precipitate_3d":
        return precipitate(**kw)
    raise KeyError(name)
