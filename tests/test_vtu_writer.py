"""The solution-file writer (VTK XML UnstructuredGrid, the reference's default output
format, include/prismspf/core/solution_output.h:53-126) on the CPU: files written by
the library's header are read back with Python's XML parser and base64 decoder and
checked against the lattice they describe (points, linear sub-cells in VTK corner order,
point data, TIME/CYCLE field data)."""

from __future__ import annotations

import base64
import struct
import subprocess
import xml.etree.ElementTree as ET
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
DTYPES = {"Float64": "<f8", "Int64": "<i8", "Int32": "<i4", "UInt8": "u1"}


def read_vtu(path):
    """{array name: ndarray} plus the piece's point and cell counts."""
    root = ET.parse(path).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("header_type") == "UInt64"
    out = {}
    for arr in root.iter("DataArray"):
        raw = base64.b64decode(arr.text.strip())
        (nbytes,) = struct.unpack("<Q", raw[:8])
        data = np.frombuffer(raw[8:8 + nbytes], dtype=DTYPES[arr.get("type")])
        comps = int(arr.get("NumberOfComponents", "1"))
        out[arr.get("Name")] = data.reshape(-1, comps) if comps > 1 else data
    piece = root.find("UnstructuredGrid/Piece")
    return out, int(piece.get("NumberOfPoints")), int(piece.get("NumberOfCells"))


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    d = tmp_path_factory.mktemp("vtu")
    exe = d / "vtu_writer"
    subprocess.run(["g++", "-std=c++20", "-O1", f"-I{ROOT / 'phasefield_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "vtu_writer.cc"), "-o", str(exe)], check=True)
    subprocess.run([str(exe), str(d)], check=True, capture_output=True)
    return d


@pytest.mark.parametrize("dim,n", [(2, (4, 3, 1)), (3, (3, 2, 2))])
def test_vtu_round_trip(written, dim, n):
    arrs, n_points, n_cells = read_vtu(written / f"lattice_{dim}d.vtu")
    assert n_points == n[0] * n[1] * n[2]
    cells = (n[0] - 1) * (n[1] - 1) * (n[2] - 1 if dim == 3 else 1)
    assert n_cells == cells
    lower, step = np.array([1.0, -2.0, 0.5]), np.array([0.5, 0.25, 2.0])
    k, j, i = np.meshgrid(*[np.arange(m) for m in reversed(n)], indexing="ij")
    idx = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)  # x fastest
    pts = lower + idx * step
    if dim == 2:
        pts[:, 2] = 0.0
    np.testing.assert_array_equal(arrs["Points"], pts)
    np.testing.assert_array_equal(arrs["f"], pts[:, 0] + 10 * pts[:, 1] + 100 * pts[:, 2])
    np.testing.assert_array_equal(arrs["g"], np.arange(n_points, dtype=np.float64))
    corners = 8 if dim == 3 else 4
    assert np.array_equal(arrs["offsets"], corners * np.arange(1, cells + 1))
    assert np.all(arrs["types"] == (12 if dim == 3 else 9))
    conn = arrs["connectivity"].reshape(cells, corners)
    # Every cell is the unit box of lattice steps at its first corner, in VTK order.
    c = idx[conn]                                   # cells x corners x 3
    d = c - c[:, :1, :]
    unit = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)]
    if dim == 3:
        unit += [(0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    assert np.array_equal(d, np.broadcast_to(np.array(unit), d.shape))
    assert arrs["TIME"][0] == 0.125 * dim and arrs["CYCLE"][0] == 7 * dim


@pytest.mark.parametrize("dim,n", [(2, (4, 3, 1)), (3, (3, 2, 2))])
def test_vtu_vector_field_is_one_array(written, dim, n):
    """A vector field's component fields are written as ONE 3-component array named
    after the declared field (DataOut's component_is_part_of_vector, reference
    solution_output.h:80-84), the z component zero in 2-D."""
    arrs, n_points, _ = read_vtu(written / f"vector_{dim}d.vtu")
    assert "u" in arrs and "u_x" not in arrs
    u = arrs["u"]
    assert u.shape == (n_points, 3)
    np.testing.assert_array_equal(u, arrs["Points"])  # u = (x, y, z), z = 0 in 2-D
    np.testing.assert_array_equal(arrs["g"], np.arange(n_points, dtype=np.float64))
