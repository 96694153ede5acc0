"""Brute-force global assembly of the GLL-collocation finite-element operators.

An independent restatement used to pin the oracle on tiny meshes: no sum
factorization, no tabulated basis constants.  The 1-D Lagrange basis on the
Gauss-Lobatto-Legendre points, its derivatives and the quadrature weights are derived
here from their definitions (numpy polynomials); element matrices are explicit
quadrature sums over the (p+1)^d points; global matrices are assembled with scipy.

Lattice convention (deal.II's node set of the uniform mesh, as at the C-ABI): node
(i0, i1[, i2]) with i_d in [0, p*cells_d], index i0 + N0*(i1 + N1*i2), x fastest.
Periodic image nodes (i_d = p*cells_d in a periodic direction) fold onto their masters
(i_d = 0); Dirichlet nodes are those on the listed faces (2d = lower, 2d+1 = upper) of
non-periodic directions.
"""

from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
from numpy.polynomial import polynomial as P


def gll_points(p: int) -> np.ndarray:
    """Gauss-Lobatto-Legendre points on [0, 1]: the ends and the roots of L_p'."""
    if p == 1:
        return np.array([0.0, 1.0])
    leg = np.polynomial.legendre.Legendre.basis(p)
    inner = np.sort(leg.deriv().roots().real)
    return np.concatenate([[0.0], (inner + 1.0) / 2.0, [1.0]])


def lagrange_tables(p: int):
    """Basis values, derivatives at the points, and exact weights w_b = int_0^1 L_b."""
    xs = gll_points(p)
    n = p + 1
    coeffs = []
    for b in range(n):
        c = np.array([1.0])
        for m in range(n):
            if m != b:
                c = P.polymul(c, np.array([-xs[m], 1.0]) / (xs[b] - xs[m]))
        coeffs.append(c)
    D = np.array([[P.polyval(xs[a], P.polyder(coeffs[b])) for b in range(n)] for a in range(n)])
    W = np.array([P.polyval(1.0, P.polyint(c)) - P.polyval(0.0, P.polyint(c)) for c in coeffs])
    # GLL quadrature weights on n points are exact for degree 2n-3 and equal the
    # integrals of the Lagrange basis.
    return xs, D, W


class Mesh:
    def __init__(self, dim, degree, cells, size, periodic, dirichlet_faces=()):
        self.dim, self.p = dim, degree
        self.cells = list(cells)
        self.size = list(size)
        self.periodic = list(periodic)
        self.h = [self.size[d] / self.cells[d] for d in range(dim)]
        self.N = [degree * self.cells[d] + 1 for d in range(dim)]
        self.ndofs = int(np.prod(self.N))
        self.faces = set(dirichlet_faces)
        # master index of every full-lattice node
        coords = self.coords()
        master = coords.copy()
        for d in range(dim):
            if self.periodic[d]:
                master[:, d] = np.where(master[:, d] == self.N[d] - 1, 0, master[:, d])
        self.master = self.index(master)
        self.unique = np.unique(self.master)
        self.pos = -np.ones(self.ndofs, dtype=np.int64)
        self.pos[self.unique] = np.arange(self.unique.size)
        dir_mask = np.zeros(self.ndofs, dtype=bool)
        for f in self.faces:
            d, upper = f // 2, f % 2
            if d < dim and not self.periodic[d]:
                dir_mask |= coords[:, d] == (self.N[d] - 1 if upper else 0)
        self.dirichlet = dir_mask

    def coords(self):
        grids = np.meshgrid(*[np.arange(n) for n in self.N], indexing="ij")
        # x fastest: flatten in Fortran order over (i0, i1, i2)
        return np.stack([g.ravel(order="F") for g in grids], axis=1)

    def index(self, c):
        idx = c[:, 0].copy()
        stride = 1
        for d in range(1, self.dim):
            stride *= self.N[d - 1]
            idx = idx + stride * c[:, d]
        return idx

    def node_x(self):
        return self.coords() * (np.array(self.h) / self.p)

    def cells_iter(self):
        return itertools.product(*[range(c) for c in self.cells])

    def cell_dofs(self, cell):
        """Full-lattice indices of the (p+1)^d nodes of a cell, local index x fastest."""
        loc = list(itertools.product(*[range(self.p + 1)] * self.dim))
        loc = [tuple(reversed(l)) for l in loc]  # make x fastest
        c = np.array([[self.p * cell[d] + l[d] for d in range(self.dim)] for l in loc])
        return self.index(c), np.array(loc)


def element_gradients(mesh: Mesh):
    """grad phi_l at quadrature point q: array [q, l, d] (collocation: q runs over nodes)."""
    xs, D, W = lagrange_tables(mesh.p)
    n = mesh.p + 1
    loc = [tuple(reversed(l)) for l in itertools.product(*[range(n)] * mesh.dim)]
    nq = len(loc)
    G = np.zeros((nq, nq, mesh.dim))
    wq = np.zeros(nq)
    for q, kq in enumerate(loc):
        wq[q] = np.prod([W[kq[d]] * mesh.h[d] for d in range(mesh.dim)])
        for l, kl in enumerate(loc):
            for d in range(mesh.dim):
                val = D[kq[d], kl[d]] / mesh.h[d]
                for e in range(mesh.dim):
                    if e != d:
                        val *= 1.0 if kq[e] == kl[e] else 0.0
                G[q, l, d] = val
    return G, wq


def assemble(mesh: Mesh, coeff=None, mass_coeff=0.0):
    """Global matrices on the unique nodes: A = sum_cells int a(x) grad.grad + mass_coeff*M.

    coeff: optional array over the full lattice giving a(x) at every node (collocated
    quadrature), default 1.  Returns (A, lumped mass vector) on unique nodes.
    """
    G, wq = element_gradients(mesh)
    rows, cols, vals = [], [], []
    mass = np.zeros(mesh.unique.size)
    for cell in mesh.cells_iter():
        dofs, _ = mesh.cell_dofs(cell)
        u = mesh.pos[mesh.master[dofs]]
        a = np.ones(len(dofs)) if coeff is None else coeff[dofs]
        Ke = np.einsum("q,qid,qjd->ij", wq * a, G, G)
        Ke += np.diag(mass_coeff * wq)
        rows.append(np.repeat(u, len(u)))
        cols.append(np.tile(u, len(u)))
        vals.append(Ke.ravel())
        np.add.at(mass, u, wq)
    n = mesh.unique.size
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n)).tocsr()
    return A, mass


def read_unique(mesh: Mesh, full: np.ndarray, constrained: bool) -> np.ndarray:
    """Values on unique nodes from a full-lattice vector (Dirichlet -> 0 if constrained)."""
    v = np.asarray(full, dtype=np.float64)[mesh.unique].copy()
    if constrained:
        v[mesh.dirichlet[mesh.unique]] = 0.0
    return v


def scatter_full(mesh: Mesh, uvals: np.ndarray, dirichlet_zero=True) -> np.ndarray:
    """Operator output on the full lattice: master entries, images 0, Dirichlet rows 0."""
    out = np.zeros(mesh.ndofs)
    out[mesh.unique] = uvals
    if dirichlet_zero:
        out[mesh.dirichlet] = 0.0
    return out


def distribute(mesh: Mesh, full: np.ndarray) -> np.ndarray:
    """Images take their masters' values."""
    return np.asarray(full)[mesh.master]


def explicit_rhs(mesh: Mesh, c_full, value_fn, flux_fn):
    """b_i = sum_cells sum_q w_q [ r(c_q, grad c_q) phi_i(q) + g(c_q, grad c_q) . grad phi_i(q) ]
    on unique nodes, with r = value_fn(c, grad c) and g = flux_fn(c, grad c) (arrays over
    the quadrature points of a cell; grad c is the cell's own gradient)."""
    G, wq = element_gradients(mesh)
    b = np.zeros(mesh.unique.size)
    for cell in mesh.cells_iter():
        dofs, _ = mesh.cell_dofs(cell)
        cv = np.asarray(c_full)[mesh.master[dofs]]
        grad = np.einsum("qld,l->qd", G, cv)
        r = value_fn(cv, grad)
        g = flux_fn(cv, grad)
        be = wq * r + np.einsum("q,qd,qld->l", wq, g, G)
        np.add.at(b, mesh.pos[mesh.master[dofs]], be)
    return b
