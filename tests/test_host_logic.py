"""Host-side logic on the CPU: the shared C++ headers (dependency strings, solve-block
rules, parameter files) and the C-ABI libraries' exported symbols (no GPU calls)."""

from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def host_logic(tmp_path_factory):
    exe = tmp_path_factory.mktemp("host") / "host_logic"
    subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT / 'phasefield_amd' / 'include'}",
                    str(ROOT / "tests" / "native" / "host_logic.cc"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict(line.split(": ", 1) for line in out.splitlines())


def test_dependency_strings(host_logic):
    # EvaluationFlags: values 1, gradients 2, hessians 4 (reference dependencies.h:30-213)
    assert host_logic["c.flag"] == "5"          # "c" + "hess(c)"
    assert host_logic["c.old1"] == "2"          # "grad(old_1(c))"
    assert host_logic["eta.old_size"] == "3"    # "old_3(eta)" extends the age list
    assert host_logic["eta.old3"] == "1"
    assert host_logic["v.src"] == "3"           # "lhs(v)" + "grad(lhs(v))"
    assert host_logic["eta.src"] == "2"         # "grad(change(eta))"
    assert host_logic["unknown"].endswith("not recognized: grad(q)")


def test_solve_block_rules(host_logic):
    assert "current value of a field should never appear in the RHS" in host_logic["own_current"]
    assert "explicit solves should have no LHS dependencies" in host_logic["explicit_lhs"]
    assert "src_flag" in host_logic["linear_no_src"]
    assert host_logic["linear_ok"] == "[]"


def test_parameter_file(host_logic):
    assert host_logic["size"] == "12.5 3"
    assert host_logic["periodic"] == "1 0"
    assert host_logic["refinement"] == "3"
    assert host_logic["n_cells"] == "16 8"          # subdivisions * 2^refinement
    assert host_logic["dt"] == "0.25"
    assert host_logic["n_increments"] == "10"
    assert host_logic["outputs"] == "0 3 6 9"       # EQUAL_SPACING, period max(1, 10/3)
    assert host_logic["D0"] == "0.7" and host_logic["n"] == "4"
    assert "not of type INT" in host_logic["const_type"]
    assert host_logic["bc_c"] == "0=1 1=0 2=0 3=1"  # Dirichlet=1, Natural=0 per face
    # RMSEPerField=1, tolerance, iterations, GMG=2, depth
    assert host_logic["lin"] == "1 1e-09 77 2 3"
    assert "Invalid preconditioner type: ILU" in host_logic["bad_precond"]
    assert "Missing 'end'" in host_logic["unterminated"]


def declared_symbols():
    text = (ROOT / "include" / "ppf.h").read_text()
    return sorted(set(re.findall(r"\b(ppf_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from phasefield_amd.capi import EXPORTED_SYMBOLS

    assert declared_symbols() == sorted(EXPORTED_SYMBOLS)


@pytest.fixture(scope="module")
def product_libs():
    from phasefield_amd import build

    missing = [a for a in build.APPS if not build.library_path(a).exists()]
    if missing:
        build.build(missing)
    return [build.library_path(a) for a in build.APPS]


def test_libraries_export_every_declared_symbol(product_libs):
    """Each application library loads without a GPU and exports the full C-ABI."""
    for path in product_libs:
        lib = ctypes.CDLL(str(path))
        missing = [s for s in declared_symbols() if not hasattr(lib, s)]
        assert not missing, f"{path.name} lacks {missing}"
        lib.ppf_app_name.restype = ctypes.c_char_p
        app = path.name[len("libppf_"):-len(".so")]
        assert lib.ppf_app_name().decode() == app


def test_vector_field_expansion_and_elastic_constants(host_logic):
    # A vector field u (declared last) becomes u_x, u_y, u_z; the block and its lhs
    # dependency are widened; component 1's boundary conditions go to u_y.
    assert host_logic["vec_names"] == "T u_x u_y u_z"
    assert host_logic["vec_layout"] == "0.0s 1.0v 1.1v 1.2v"
    assert host_logic["vec_block"] == "1:2 2:2 3:2"        # src gradients of every component
    assert host_logic["vec_bc"] == "0 1"                  # u_x: none; u_y: Dirichlet on face 0
    assert "must be the last declared field" in host_logic["vec_not_last"]
    # E = 2, nu = 0.25: lambda = 0.8, mu = 0.8 -> C11 = 2.4, C12 = 0.8, C44 = 0.8.
    assert host_logic["stiffness"] == "2.4 0.8 0.8 0"
    assert "only isotropic" in host_logic["ortho"]
    # 2-D (reference get_cij_matrix): plane strain by default and when named, lambda =
    # nu E/((1+nu)(1-2nu)) = 0.8; plane stress lambda = nu E/(1-nu^2) = 0.5/0.9375.
    assert host_logic["stiffness2d_s0"] == "2.4 0.8 0.8"
    assert host_logic["stiffness2d_s1"] == "2.4 0.8 0.8"
    lam = 0.5 / 0.9375
    c = [float(v) for v in host_logic["stiffness2d_s2"].split()]
    assert c == pytest.approx([lam + 1.6, lam, 0.8], rel=1e-5)
    assert "Wrong 2D stress state" in host_logic["stiffness2d_bad"]
    assert host_logic["stiffness1d"] == "2"                 # 1-D: C = E
