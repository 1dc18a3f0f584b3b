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
  a = u.v dx, L = -grad(phi).v dx on the vector CG1 space;
  the object charge form of punc/ufl/Charge.ufl (eps0 grad(phi).n ds over an
  object's facets, n the domain's outward normal) and the capacitance-matrix
  circuit CircuitCM (punc/src/object_CM.cpp:40-357, with the charge-sharing
  sets of punc/src/object.cpp:68-135): inverse capacitance from unit-potential
  Laplace solves, circuit matrix, pre_solve / post_solve and the corrected
  solve of PoissonSolver::solve_circuit (poisson.cpp:364-376).
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

    def object_charges(self, phi, n_objects):
        """Charge.ufl: eps0 grad(phi).n ds over each object's facets, n the
        outward normal of the domain (pointing into the object)."""
        g, vol = _gradients(self.hm)
        P = self.hm.xyz[self.hm.cells]
        phi = np.asarray(phi, dtype=np.float64)
        gphi = np.einsum("ckd,ck->cd", g, phi[self.dofs])  # grad phi per cell
        out = np.zeros(n_objects)
        for j in range(n_objects):
            c, f = np.nonzero(self.adj == -(2 + j))
            for ci, fi in zip(c, f):
                a, b, d = (P[ci, k] for k in _FACETS[fi])
                nn = np.cross(b - a, d - a)
                area = 0.5 * np.linalg.norm(nn)
                nn = nn / np.linalg.norm(nn)
                if nn @ (P[ci, fi] - a) > 0:  # point away from the opposite vertex
                    nn = -nn
                out[j] += self.eps0 * (gphi[ci] @ nn) * area
        return out

    def efield(self, phi):
        """E (ndof, 3): M E_c = B_c phi."""
        Mc = self.M.tocsc()
        lu = spla.splu(Mc)
        return np.column_stack([lu.solve(self.B[c] @ np.asarray(phi, dtype=np.float64)) for c in range(3)])


def _sharing_set(vs, node, group):
    """get_charge_sharing_set (object.cpp:68-88)."""
    group.append(node)
    i = 0
    while i < len(vs):
        a, b, _ = vs[i]
        if a == node:
            del vs[i]
            _sharing_set(vs, b, group)
        elif b == node:
            del vs[i]
            _sharing_set(vs, a, group)
        else:
            i += 1


def charge_sharing_sets(vsources, num_objects):
    """get_charge_sharing_sets (object.cpp:90-135)."""
    vs = list(vsources)
    nodes = list(range(num_objects))
    groups = []
    while vs:
        group = []
        _sharing_set(vs, vs[0][0], group)
        groups.append(group)
    for g in groups:
        for v in g:
            if v != -1 and v in nodes:
                nodes.remove(v)
    groups = [g for g in groups if -1 not in g]
    for v in nodes:
        groups.append([v])
    return groups


class CircuitOracle:
    """CircuitCM on a FieldOracle with n objects (exterior grounded)."""

    def __init__(self, field, n_objects, dt, vsources=(), isources=()):
        self.f = field
        self.n = n = n_objects
        self.dt = dt
        self.vsources = list(vsources)
        self.isources = list(isources)
        # inv_capacitance (object_CM.cpp:53-93)
        C = np.empty((n, n))
        for j in range(n):
            pot = [1.0 if i == j else 0.0 for i in range(n)]
            phi = field.poisson(np.zeros(field.hm.ndof), exterior=True, object_potentials=pot)
            C[:, j] = field.object_charges(phi, n)
        self.capacitance = C
        self.inv_c = np.linalg.inv(C)
        self.groups = charge_sharing_sets(self.vsources, n)
        self.fixed_voltage = []
        for a, b, _ in self.vsources:
            if a != -1:
                continue
            tmp = [b]
            node = b
            for a2, b2, _ in self.vsources:
                if a2 == node:
                    tmp.append(b2)
                    node = b2
            self.fixed_voltage.append(tmp)
        self.vector = np.zeros(2 * n)
        for a, b, v in self.vsources:
            self.vector[b] = v
        # assemble_matrix (object_CM.cpp:204-248)
        M = np.zeros((2 * n, 2 * n))
        iC = self.inv_c
        for a, b, _ in self.vsources:
            M[b, b] = 1.0
            if a != -1:
                M[b, a] = -1.0
        for g in self.groups:
            for j in g:
                M[n + g[0], n + j] = 1.0
            M[g[0], g[0]] = 1.0
            for j in g:
                M[g[0], n + j] = -1.0 * iC[g[0], j]
            for j in g[1:]:
                M[n + j, j] = 1.0
                for k in g:
                    M[n + j, n + k] = -1.0 * iC[j, k]
        for fv in self.fixed_voltage:
            for j in fv:
                M[n + j, j] = 1.0
                for k in fv:
                    M[n + j, n + k] = -1.0 * iC[j, k]
        self.inv_circuit = np.linalg.inv(M)

    def solve(self, rho, charges):
        """solve_circuit: returns (phi, potentials, charges)."""
        n, f, iC = self.n, self.f, self.inv_c
        q = np.array(charges, dtype=np.float64)
        for a, b, v in self.isources:  # apply_isources_to_object (object_CM.cpp:343-357)
            dQ = v * self.dt
            if a != -1:
                q[a] -= dQ
            if b != -1:
                q[b] += dQ
        phi0 = f.poisson(rho, exterior=True, object_potentials=[0.0] * n)
        image = f.object_charges(phi0, n)
        cv = self.vector
        for g in self.groups:  # assemble_vector (object_CM.cpp:250-297)
            cv[n + g[0]] = sum(q[j] for j in g)
        for g in self.groups:
            cv[g[0]] = sum(iC[g[0], j] * image[j] for j in g)
            for j in g[1:]:
                cv[n + j] = sum(iC[j, k] * image[k] for k in g)
        for fv in self.fixed_voltage:
            for j in fv:
                cv[n + j] = sum(iC[j, k] * image[k] for k in fv)
        pot = self.inv_circuit[:n] @ cv
        q_new = self.inv_circuit[n:] @ cv
        phi = f.poisson(rho, exterior=True, object_potentials=list(pot))
        return phi, pot, q_new
