"""Shared setup for the parity tests: meshes, fields, identical initial
populations on the GPU path and in the oracle, and set comparisons."""
import numpy as np

import oracle as O
import puncpp_amd as P

QE, ME = -1.0, 1.0
QI, MI = 1.0, 1836.0


def sphere_mesh(n=8, L=4.0, r=1.5):
    return P.box_mesh(n, [-L] * 3, [L] * 3, sphere_r=r)


def periodic_mesh(n=6):
    return P.box_mesh(n, [0, 0, 0], [2 * np.pi] * 3, periodic=True)


def synthetic_field(hm, seed=3, scale=0.5):
    """E at each dof: a linear field plus seeded noise (ndof x 3)."""
    rng = np.random.default_rng(seed)
    xyz = np.zeros((hm.ndof, 3))
    xyz[hm.vdof] = hm.xyz  # any image of a periodic dof
    A = rng.normal(size=(3, 3)) * 0.1
    E = scale * (xyz @ A.T + rng.normal(size=3)) + 0.05 * rng.normal(size=(hm.ndof, 3))
    return np.ascontiguousarray(E)


def loaded_pair(hm, npart=(2000, 2000), seed=0, vth=(1.0, 1.0 / np.sqrt(1836.0)),
                vd=(0.0, 0.0, 0.5), amp=0.0, kw=1.0, capacity=None):
    """Oracle population loaded by its load_particles restatement and a GPU
    population holding the identical particles (same order)."""
    o = O.OraclePop(hm, [QE, QI], [ME, MI])
    for s, (n, v) in enumerate(zip(npart, vth)):
        assert o.load(s, n, seed, amp, kw, 0, v, vd) == 0
    mesh = P.Mesh(hm)
    pop = P.Population(mesh, [(QE, ME), (QI, MI)], capacity=capacity or 4 * sum(npart) + 100000)
    st = o.get_particles()
    pop.add_located(st["x"], st["v"], st["species"], st["cell"])
    return o, mesh, pop


def canon(st):
    """Particles as a canonically ordered structured view."""
    key = np.lexsort((st["v"][:, 2], st["v"][:, 1], st["v"][:, 0], st["x"][:, 2], st["x"][:, 1],
                      st["x"][:, 0], st["cell"], st["species"]))
    return {k: v[key] for k, v in st.items()}


def near_face(o, st, tol=1e-12):
    """Number of particles within tol of a facet plane of their cell."""
    planes = o.localizer()["planes"].reshape(-1, 4, 4)
    pl = planes[st["cell"]]
    x = st["x"]
    proj = pl[:, :, 0] + pl[:, :, 1] * x[:, None, 0] + pl[:, :, 2] * x[:, None, 1] + pl[:, :, 3] * x[:, None, 2]
    return int((np.abs(proj).min(axis=1) < tol).sum())


def assert_same_sets(a, b, rel=1e-12, what=""):
    assert a["x"].shape == b["x"].shape, f"{what}: sizes {a['x'].shape[0]} vs {b['x'].shape[0]}"
    if a["x"].shape[0] == 0:
        return
    ca, cb = canon(a), canon(b)
    np.testing.assert_array_equal(ca["species"], cb["species"], err_msg=what)
    np.testing.assert_array_equal(ca["cell"], cb["cell"], err_msg=what)
    for k in ("x", "v"):
        scale = np.maximum(np.abs(cb[k]), 1e-300)
        err = np.abs(ca[k] - cb[k]) / scale
        assert err.max() <= rel, f"{what}: {k} max rel err {err.max():.3e}"


def rho_scale(o, hm, dv_inv):
    """Per-dof magnitude dv_inv_j * sum |q| lambda_j (deposit tolerance scale)."""
    st = o.get_particles()
    loc = o.localizer()
    bary = loc["bary"].reshape(-1, 3, 4)[st["cell"]]
    x = st["x"]
    l = bary[:, :, 0] + bary[:, :, 1] * x[:, None, 0] + bary[:, :, 2] * x[:, None, 1] + bary[:, :, 3] * x[:, None, 2]
    lam = np.concatenate([l, 1 - l.sum(axis=1, keepdims=True)], axis=1)
    q = np.where(st["species"] == 0, abs(QE), abs(QI))
    dofs = hm.vdof[hm.cells[st["cell"]]]
    mag = np.zeros(hm.ndof)
    np.add.at(mag, dofs.ravel(), (np.abs(lam) * q[:, None]).ravel())
    return mag * dv_inv


def first_positive_meshes():
    """Two tetrahedra T0, T1 sharing facet BCD, and the point p = (-0.05,
    0.6, 0.6) inside T1 that lies on the positive side of two facet planes of
    T0: BCD (the neighbour, proj 0.087) and ACD (an object facet, boundary
    id 2, proj 0.05).  The two vertex labelings give the object facet the
    higher local index (labeling 0) or the lower one (labeling 1): the
    first-positive walk from T0 (population.h:547-565) then ends in T1 or on
    the object, where an argmax walk would end in T1 both times and a
    last-positive walk the other way round."""
    A, B, Cc, D, E = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (-0.5, 1.0, 1.0)
    m0 = P.HostMesh(np.array([A, B, Cc, D, E]), np.array([[0, 1, 2, 3], [1, 2, 3, 4]]),
                    np.array([[0, 2, 1, 1], [1, 1, 1, 0]], dtype=np.int32))
    m1 = P.HostMesh(np.array([B, A, Cc, D, E]), np.array([[0, 1, 2, 3], [0, 2, 3, 4]]),
                    np.array([[2, 0, 1, 1], [1, 1, 1, 0]], dtype=np.int32))
    p = np.array([-0.05, 0.6, 0.6])
    return (m0, 1), (m1, -2), p


def digest_np(st, nspecies, c0=None):
    """The population digest (punc_gpu_pop_digest) of a particle state in
    numpy: per species the wrapping sum of the splitmix64-chained hash of the
    x, y, z, vx, vy, vz bit patterns and the cell id (minus c0[species])."""
    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
        return z ^ (z >> np.uint64(31))

    sp = st["species"].astype(np.int64)
    cell = st["cell"].astype(np.int64)
    if c0 is not None:
        cell = cell - np.asarray(c0, dtype=np.int64)[sp]
    with np.errstate(over="ignore"):
        x = np.ascontiguousarray(st["x"]).view(np.uint64)
        v = np.ascontiguousarray(st["v"]).view(np.uint64)
        h = mix(x[:, 0] ^ np.uint64(0x9e3779b97f4a7c15))
        for col in (x[:, 1], x[:, 2], v[:, 0], v[:, 1], v[:, 2]):
            h = mix(h ^ col)
        h = mix(h ^ (cell.astype(np.uint32).astype(np.uint64)))
        out = np.zeros(nspecies, dtype=np.uint64)
        for s in range(nspecies):
            out[s] = np.sum(h[sp == s], dtype=np.uint64)
    counts = np.bincount(sp, minlength=nspecies).astype(np.int64)
    return out, counts
