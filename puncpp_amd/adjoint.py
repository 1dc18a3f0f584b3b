"""torch.autograd harness over the adjoint kernels (punc_gpu_adjoint_*).

``PushDeposit`` runs K steps of boris_cg1 + move (+ relocation) in a fixed CG1
field and deposits the final positions; ``push_deposit(v0, E, runner)`` is the
differentiable map (v0, E) -> w, w_j = sum_p q_p lambda_j(x_K), and
``push_deposit_collect(v0, E, runner)`` the map (v0, E) -> (w, Q), Q_b the
time average over the K steps of the charge boundary b has collected (the
collected charge's continuous form, punc_gpu_adjoint_collected).  Their
backward runs the hand-written reverse kernels.  ``efield_projection(phi,
field)`` is the device E projection E = M^-1 B phi with its reverse, so
objectives differentiate down to the vertex potentials.  Objectives are
ordinary torch expressions, e.g. ||rho||^2 = sum_j (dv_inv_j w_j)^2 or Q[2].
"""
import numpy as np
import torch

from . import _native as N
from ._native import dptr, i32ptr, i64ptr, vec3


class PushDeposit:
    def __init__(self, mesh, q, m, species, x0, cell0, B, dt, K, every=10, Ld=None):
        self.mesh = mesh
        self.n = int(len(species))
        self.q = np.ascontiguousarray(q, dtype=np.float64)
        self.m = np.ascontiguousarray(m, dtype=np.float64)
        self.species = np.ascontiguousarray(species, dtype=np.int32)
        self.x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1, 3)
        self.cell0 = np.ascontiguousarray(cell0, dtype=np.int64)
        self.B = vec3(B)
        self.dt = float(dt)
        self.K = int(K)
        self.every = int(every)
        self.Ld = None if Ld is None else vec3(Ld)
        h = N.vp()
        L = N.lib()
        self._check(L.punc_gpu_adjoint_create(mesh._h, len(self.q), dptr(self.q), dptr(self.m), self.n,
                                              i32ptr(self.species), N.C.byref(h)))
        self._h = h

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise N.PuncGpuError(N.lib().punc_gpu_adjoint_last_error().decode())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().punc_gpu_adjoint_destroy(h)
            self._h = None

    def forward(self, v0, E, final_state=False):
        v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1, 3)
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(-1, 3)
        w = np.empty(self.mesh.ndof)
        xK = np.empty((self.n, 3)) if final_state else None
        vK = np.empty((self.n, 3)) if final_state else None
        cK = np.empty(self.n, dtype=np.int64) if final_state else None
        self._check(N.lib().punc_gpu_adjoint_forward(
            self._h, dptr(self.x0), i64ptr(self.cell0), dptr(v0), dptr(E), dptr(self.B), self.dt,
            dptr(self.Ld) if self.Ld is not None else None, self.K, self.every, dptr(w), dptr(xK), dptr(vK),
            i64ptr(cK)))
        if final_state:
            return w, xK, vK, cK
        return w

    def collected(self):
        """Of the last forward: per particle the weight (K - n - s)/K of its
        charge in Q, the boundary it left through and the step (0 and -1 while
        in the mesh), and the number of irregular exits."""
        w = np.empty(self.n)
        b = np.empty(self.n, dtype=np.int32)
        st = np.empty(self.n, dtype=np.int32)
        irr = N.C.c_int64()
        self._check(N.lib().punc_gpu_adjoint_collected(self._h, dptr(w), i32ptr(b), i32ptr(st), N.C.byref(irr)))
        return w, b, st, irr.value

    def collected_charge(self):
        """Q (64,): Q_b = sum_p q_p weight_p over the particles that left
        through boundary b, summed in particle order."""
        w, b, _, _ = self.collected()
        return np.bincount(b, weights=self.q[self.species] * w, minlength=64)[:64]

    def backward(self, g, gq=None):
        g = np.ascontiguousarray(g, dtype=np.float64)
        dE = np.empty((self.mesh.ndof, 3))
        dv = np.empty((self.n, 3))
        dx = np.empty((self.n, 3))
        if gq is None:
            self._check(N.lib().punc_gpu_adjoint_backward(self._h, dptr(g), dptr(dE), dptr(dv), dptr(dx)))
        else:
            gq = np.ascontiguousarray(gq, dtype=np.float64)
            if gq.shape != (64,):
                raise ValueError("gq must hold 64 values (one per boundary id)")
            self._check(N.lib().punc_gpu_adjoint_backward_q(self._h, dptr(g), dptr(gq), dptr(dE), dptr(dv),
                                                            dptr(dx)))
        return dE, dv, dx


class _PushDepositFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, E, runner):
        v = v0.detach().cpu().numpy()
        e = E.detach().cpu().numpy()
        w = runner.forward(v, e)
        ctx.runner = runner
        ctx.inputs = (v, e)
        return torch.from_numpy(w)

    @staticmethod
    def backward(ctx, gw):
        # The reverse sweep replays the checkpoints of the runner's last
        # forward; redo this call's forward in case the runner ran others since.
        ctx.runner.forward(*ctx.inputs)
        dE, dv, _ = ctx.runner.backward(gw.detach().cpu().numpy())
        return torch.from_numpy(dv), torch.from_numpy(dE), None


class _PushDepositCollectFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, E, runner):
        v = v0.detach().cpu().numpy()
        e = E.detach().cpu().numpy()
        w = runner.forward(v, e)
        Q = runner.collected_charge()
        ctx.runner = runner
        ctx.inputs = (v, e)
        return torch.from_numpy(w), torch.from_numpy(Q)

    @staticmethod
    def backward(ctx, gw, gQ):
        ctx.runner.forward(*ctx.inputs)
        g = np.zeros(ctx.runner.mesh.ndof) if gw is None else gw.detach().cpu().numpy()
        gq = np.zeros(64) if gQ is None else gQ.detach().cpu().numpy()
        dE, dv, _ = ctx.runner.backward(g, gq)
        return torch.from_numpy(dv), torch.from_numpy(dE), None


class _EProjectionFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, field):
        ctx.field = field
        return torch.from_numpy(field.efield(phi.detach().cpu().numpy()))

    @staticmethod
    def backward(ctx, gE):
        return torch.from_numpy(ctx.field.efield_adjoint(gE.detach().cpu().numpy())), None


def efield_projection(phi, field):
    """phi [ndof] -> E [ndof, 3], the device's vector-CG1 projection
    (ESolver, efield.cpp:70-75), differentiable in phi through
    dJ/dphi = B^T M^-1 dJ/dE.  field: a puncpp_amd.FieldSolver."""
    return _EProjectionFn.apply(phi, field)


def push_deposit_collect(v0, E, runner):
    """(v0 [n,3], E [ndof,3]) -> (w [ndof], Q [64]), differentiable in v0 and
    E: w the deposit of the final positions, Q the time-averaged collected
    charge per boundary id (PushDeposit.collected_charge)."""
    return _PushDepositCollectFn.apply(v0, E, runner)


def gradient_operator(mesh):
    """Sparse G with E = G phi: at each CG1 dof the volume-weighted mean over
    the adjacent cells of -grad(phi) (phi piecewise linear).  Shape
    (3*ndof, ndof), rows ordered dof-major (E_x, E_y, E_z per dof)."""
    import scipy.sparse as sp

    loc = mesh.localizer()
    bary = loc["bary"].reshape(-1, 3, 4)
    grads = np.concatenate([bary[:, :, 1:], -bary[:, :, 1:].sum(axis=1, keepdims=True)], axis=1)  # nc,4,3
    vc = loc["vcoords"].reshape(-1, 4, 3)
    vol = np.abs(np.linalg.det(vc[:, 1:] - vc[:, :1])) / 6.0
    dofs = loc["dofs"]
    ndof = mesh.ndof
    wsum = np.zeros(ndof)
    np.add.at(wsum, dofs.ravel(), np.repeat(vol, 4))
    rows, cols, vals = [], [], []
    for a in range(4):  # the dof receiving the cell's gradient
        for b in range(4):  # the dof whose potential enters the gradient
            for d in range(3):
                rows.append(3 * dofs[:, a] + d)
                cols.append(dofs[:, b])
                vals.append(-vol * grads[:, b, d] / wsum[dofs[:, a]])
    G = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(3 * ndof, ndof)).tocsr()
    G.sum_duplicates()
    return G


def push_deposit(v0, E, runner):
    """(v0 [n,3], E [ndof,3]) -> w [ndof], differentiable in v0 and E.
    The runner must have been built for the same particles and mesh."""
    return _PushDepositFn.apply(v0, E, runner)
