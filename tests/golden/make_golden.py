"""Writes the golden fixtures of tests/test_golden.py (run: python -m tests.golden.make_golden).

Each fixture is a small run of one application, produced by the CPU oracle after the
oracle has been checked against brute-force assembly, closed forms and invariants
(tests/test_oracle.py).  A fixture stores the parameter text, the inputs that are not
produced by the application itself, and the outputs, all over deal.II's full node set.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from phasefield_amd import prm

HERE = Path(__file__).resolve().parent
LATEST = -1


def cases():
    """name -> (application, parameter text, increments, extra description)."""
    return {
        "allen_cahn_2d_q1": ("allen_cahn_2d_q1",
                             prm.allen_cahn(dim=2, size=20.0, refinement=4, dt=0.01, steps=20,
                                            outputs=4, n_circles=3), 20),
        "cahn_hilliard_3d_q2": ("cahn_hilliard_3d_q2",
                                prm.cahn_hilliard(size=8.0, refinement=2, dt=1e-3, steps=10,
                                                  outputs=2), 10),
        "precipitate_3d_q1": ("precipitate_3d_q1",
                              prm.precipitate(size=12.0, refinement=1, dt=5e-3, steps=5,
                                              outputs=1, n_nuclei=2, radius=2.0), 5),
        "testbed_2d_q2": ("testbed_2d_q2", prm.testbed(dim=2, refinement=3, steps=2), 2),
        "poisson_2d_q2": ("poisson_2d_q2",
                          prm.poisson(dim=2, refinement=4, mg_depth=3, tolerance=1e-10), 0),
    }


def run_case(app, text, n):
    from oracle.binding import Oracle

    o = Oracle(app, text, threads=4)
    status = o.run(n)
    out = {"status": np.array(status)}
    for f in range(o.nfields):
        out[f"field_{f}"] = o.get_field(f, LATEST)
    out["log"] = np.frombuffer(o.output_log().encode(), dtype=np.uint8)
    return out


def adjoint_case():
    from oracle.binding import Oracle

    text = prm.allen_cahn(dim=2, size=16.0, refinement=3, dt=0.05, steps=12, outputs=0)
    o = Oracle("allen_cahn_2d_q1", text, threads=4)
    n = 9  # 8 cells of size 2 per direction, periodic image included
    x = np.arange(n) * 2.0
    X, Y = np.meshgrid(x, x, indexing="xy")
    assert n * n == o.ndofs
    phi0 = (0.6 * np.sin(2 * np.pi * X / 16.0) * np.cos(2 * np.pi * Y / 16.0) + 0.1).ravel()
    target = o.adjoint_forward(phi0, 1.2, 1.3, 12)
    final = o.adjoint_forward(phi0, 1.0, 1.5, 12)
    m = o.lumped_mass()
    lam = m * (final - target)
    dL, dk, dphi = o.adjoint_vjp(phi0, lam, 1.0, 1.5, 12)
    return text, {"phi0": phi0, "target": target, "final": final, "lam": lam,
                  "dL": np.array(dL), "dkappa": np.array(dk), "dphi0": dphi}


def main():
    meta = {}
    for name, (app, text, n) in cases().items():
        out = run_case(app, text, n)
        np.savez_compressed(HERE / f"{name}.npz", **out)
        meta[name] = {"app": app, "increments": n, "prm": text}
        print("wrote", name, flush=True)
    text, out = adjoint_case()
    np.savez_compressed(HERE / "allen_cahn_2d_q1_adjoint.npz", **out)
    meta["allen_cahn_2d_q1_adjoint"] = {"app": "allen_cahn_2d_q1", "L": 1.0, "kappa": 1.5,
                                        "steps": 12, "target_L": 1.2, "target_kappa": 1.3,
                                        "prm": text}
    (HERE / "fixtures.json").write_text(json.dumps(meta, indent=1, sort_keys=True) + "\n")
    print("wrote fixtures.json")


if __name__ == "__main__":
    main()
