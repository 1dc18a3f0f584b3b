"""ORACLE - test infrastructure only (imported by tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg, never by the product library).

CPU restatement of the reference's field solve:
  PoissonSolver (punc/src/poisson.cpp:260-376) on the forms of
  punc/ufl/Potential3D.ufl: a = eps0 grad(phi).grad(v) dx, L = rho v dx,
  with DOLFIN's DirichletBC.apply (rows of the boundary dofs replaced by unit
  rows, right-hand side set to the boundary value; exterior first, then the
  objects, poisson.cpp:318-331, 347-356) and, for remove_null_space, the
  right-hand side orthogonalised against the normalised constant vector
  (poisson.cpp:339-347, 357) and the zero-mean solution;
  ESolver (punc/src/efield.cpp:41-75) on punc/ufl/EField3D.ufl:
  a = u.v dx, L = -grad(phi).v dx on the vector CG1 space.
The element integrals are the exact P1 ones (assembled with numpy over all
cells at once); the linear systems are solved with a sparse direct solver
(scipy's SuperLU) where the reference runs a Krylov method to relative
1e-12, so the two agree to that tolerance.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_FACETS = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))  # local facet f: the vertices other than f


def _gradients(hm):
    """Per cell: barycentric gradients g (nc, 4, 3) and volume (nc)."""
    P = hm.xyz[hm.cells]                       # (nc, 4, 3)
    J = np.transpose(P[:, 1:, :] - P[:, :1, :], (0, 2, 1))  # columns P_k - P_0
    Jinv = np.linalg.inv(J)                    # rows: gradients of lambda_1..3
    g = np.empty((len(P), 4, 3))
    g[:, 1:, :] = Jinv
    g[:, 0, :] = -Jinv.sum(axis=1)
    vol = np.abs(np.linalg.det(J)) / 6.0
    return g, vol


class FieldOracle:
    def __init__(self, hm, eps0=1.0, adj=None):
        """hm: the host mesh; adj: (nc, 4) facet neighbours of the oracle's
        localizer (-1 exterior, -(2+j) object j)."""
        self.hm = hm
        self.eps0 = float(eps0)
        n = hm.ndof
        dofs = hm.vdof[hm.cells]                 # (nc, 4)
        g, vol = _gradients(hm)
        rows = np.repeat(dofs, 4, axis=1).ravel()
        cols = np.tile(dofs, (1, 4)).ravel()
        Ke = self.eps0 * vol[:, None, None] * np.einsum("cad,cbd->cab", g, g)
        Me = vol[:, None, None] * (np.ones((4, 4)) + np.eye(4))[None] / 20.0
        self.K = sp.csr_matrix((Ke.ravel(), (rows, cols)), shape=(n, n))
        self.M = sp.csr_matrix((Me.ravel(), (rows, cols)), shape=(n, n))
        # B_c[i, j] = sum over cells of -(V/4) (g_j)_c : rows test dofs, cols phi dofs
        self.B = []
        for c in range(3):
            Be = -(vol[:, None, None] / 4.0) * np.broadcast_to(g[:, None, :, c], (len(vol), 4, 4))
            self.B.append(sp.csr_matrix((Be.ravel(), (rows, cols)), shape=(n, n)))
        self.dofs = dofs
        self.adj = adj

    def boundary_dofs(self, bnd):
        """Dofs of the facets with facet neighbour -bnd (1 exterior, 2+j object j)."""
        c, f = np.nonzero(self.adj == -bnd)
        out = set()
        for ci, fi in zip(c, f):
            for k in _FACETS[fi]:
                out.add(int(self.dofs[ci, k]))
        return np.array(sorted(out), dtype=np.int64)

    def poisson(self, rho, exterior=True, exterior_values=None, object_potentials=(), remove_null_space=False):
        n = self.hm.ndof
        b = self.M @ np.asarray(rho, dtype=np.float64)
        bcs = []
        if exterior:
            d = self.boundary_dofs(1)
            v = np.zeros(len(d)) if exterior_values is None else np.asarray(exterior_values)[d]
            bcs.append((d, v))
        for j, V in enumerate(object_potentials):
            d = self.boundary_dofs(2 + j)
            bcs.append((d, np.full(len(d), float(V))))
        isd = np.zeros(n)
        for d, _ in bcs:  # DirichletBC::apply(A): the rows of the boundary dofs become unit rows
            isd[d] = 1.0
        A = (sp.diags(1.0 - isd) @ self.K + sp.diags(isd)).tocsr()
        for d, v in bcs:  # DirichletBC::apply(b), in application order
            b[d] = v
        if remove_null_space:
            e = np.full(n, np.sqrt(1.0 / n))
            b = b - (b @ e) * e
            Z = sp.bmat([[A, sp.csr_matrix(e[:, None])], [sp.csr_matrix(e[None, :]), None]], format="csc")
            x = spla.spsolve(Z, np.append(b, 0.0))[:n]
            return x - x.mean()
        return spla.spsolve(A.tocsc(), b)

    def efield(self, phi):
        """E (ndof, 3): M E_c = B_c phi."""
        Mc = self.M.tocsc()
        lu = spla.splu(Mc)
        return np.column_stack([lu.solve(self.B[c] @ np.asarray(phi, dtype=np.float64)) for c in range(3)])
