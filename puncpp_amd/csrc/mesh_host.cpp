// Host-side mesh setup: generator, localizer, patch volumes, exterior facets,
// point-location grid and Maxwellian flux tables.  See mesh_host.h.
#include "mesh_host.h"

#include <math.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

namespace punc_gpu {

namespace {

struct FacetKey {
    int32_t a, b, c;
    int64_t slot;  // cell*4 + local facet
};

// Local facet i of a tetrahedron is opposite local vertex i; its vertices are
// the other three, kept in ascending local order.
inline void facet_locals(int i, int out[3]) {
    int k = 0;
    for (int j = 0; j < 4; ++j)
        if (j != i) out[k++] = j;
}

// Facet-to-cell adjacency: adj[c*4+i] = neighbour cell, or -1 if the facet has
// one adjacent cell.  Returns false if a facet is shared by more than two cells.
bool facet_adjacency(const std::vector<int32_t>& cells, int64_t nc, std::vector<int64_t>& adj,
                     std::string& err) {
    std::vector<FacetKey> keys((size_t)nc * 4);
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t* v = &cells[(size_t)c * 4];
        for (int i = 0; i < 4; ++i) {
            int l[3];
            facet_locals(i, l);
            keys[(size_t)c * 4 + i] = FacetKey{v[l[0]], v[l[1]], v[l[2]], c * 4 + i};
        }
    }
    std::sort(keys.begin(), keys.end(), [](const FacetKey& x, const FacetKey& y) {
        if (x.a != y.a) return x.a < y.a;
        if (x.b != y.b) return x.b < y.b;
        if (x.c != y.c) return x.c < y.c;
        return x.slot < y.slot;
    });
    adj.assign((size_t)nc * 4, -1);
    size_t k = 0;
    while (k < keys.size()) {
        size_t e = k + 1;
        while (e < keys.size() && keys[e].a == keys[k].a && keys[e].b == keys[k].b &&
               keys[e].c == keys[k].c)
            ++e;
        if (e - k > 2) {
            err = "non-manifold mesh: a facet is shared by more than two cells";
            return false;
        }
        if (e - k == 2) {
            adj[keys[k].slot] = keys[k + 1].slot / 4;
            adj[keys[k + 1].slot] = keys[k].slot / 4;
        }
        k = e;
    }
    return true;
}

inline void cross3(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

// Outward unit normal of local facet i and the facet midpoint.  Convention of
// this library (the reference takes DOLFIN's Cell::normal and
// Facet::midpoint, population.h:375-382): with P0 the opposite vertex and
// P1<P2<P3 the facet vertices, n = (P2-P1) x (P3-P1) / |.|, flipped so that
// n . (P0-P1) <= 0; midpoint = ((P1+P2)+P3)/3.
void facet_normal_midpoint(const double* vc, int i, double n[3], double mid[3], double* area) {
    int l[3];
    facet_locals(i, l);
    const double* P0 = vc + 3 * i;
    const double* P1 = vc + 3 * l[0];
    const double* P2 = vc + 3 * l[1];
    const double* P3 = vc + 3 * l[2];
    double V0[3], V1[3], V2[3];
    for (int d = 0; d < 3; ++d) {
        V0[d] = P0[d] - P1[d];
        V1[d] = P2[d] - P1[d];
        V2[d] = P3[d] - P1[d];
    }
    cross3(V1, V2, n);
    double norm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < 3; ++d) n[d] = n[d] / norm;
    if (n[0] * V0[0] + n[1] * V0[1] + n[2] * V0[2] > 0.0)
        for (int d = 0; d < 3; ++d) n[d] = -n[d];
    for (int d = 0; d < 3; ++d) mid[d] = ((P1[d] + P2[d]) + P3[d]) / 3.0;
    if (area) *area = 0.5 * norm;
}

// Barycentric matrix rows 0..2 (reference population.cpp:181-240).
void barycentric_matrix(const double* vc, double* A) {
    double x1 = vc[0], y1 = vc[1], z1 = vc[2];
    double x2 = vc[3], y2 = vc[4], z2 = vc[5];
    double x3 = vc[6], y3 = vc[7], z3 = vc[8];
    double x4 = vc[9], y4 = vc[10], z4 = vc[11];
    double x13 = x1 - x3, y13 = y1 - y3, z13 = z1 - z3;
    double x14 = x1 - x4, y14 = y1 - y4, z14 = z1 - z4;
    double x24 = x2 - x4, y24 = y2 - y4, z24 = z2 - z4;
    double x31 = x3 - x1, y31 = y3 - y1, z31 = z3 - z1;
    double x32 = x3 - x2, y32 = y3 - y2, z32 = z3 - z2;
    double x34 = x3 - x4, y34 = y3 - y4, z34 = z3 - z4;
    double x42 = x4 - x2, y42 = y4 - y2, z42 = z4 - z2;
    double x43 = x4 - x3, y43 = y4 - y3, z43 = z4 - z3;
    double V01 = (x2 * (y3 * z4 - y4 * z3) + x3 * (y4 * z2 - y2 * z4) + x4 * (y2 * z3 - y3 * z2)) / 6.0;
    double V02 = (x1 * (y4 * z3 - y3 * z4) + x3 * (y1 * z4 - y4 * z1) + x4 * (y3 * z1 - y1 * z3)) / 6.0;
    double V03 = (x1 * (y2 * z4 - y4 * z2) + x2 * (y4 * z1 - y1 * z4) + x4 * (y1 * z2 - y2 * z1)) / 6.0;
    double V04 = (x1 * (y3 * z2 - y2 * z3) + x2 * (y1 * z3 - y3 * z1) + x3 * (y2 * z1 - y1 * z2)) / 6.0;
    double V = V01 + V02 + V03 + V04;
    double V6 = 6 * V;
    A[0] = V01 / V;
    A[4] = V02 / V;
    A[8] = V03 / V;
    A[1] = (y42 * z32 - y32 * z42) / V6;
    A[5] = (y31 * z43 - y34 * z13) / V6;
    A[9] = (y24 * z14 - y14 * z24) / V6;
    A[2] = (x32 * z42 - x42 * z32) / V6;
    A[6] = (x43 * z31 - x13 * z34) / V6;
    A[10] = (x14 * z24 - x24 * z14) / V6;
    A[3] = (x42 * y32 - x32 * y42) / V6;
    A[7] = (x31 * y43 - x34 * y13) / V6;
    A[11] = (x24 * y14 - x14 * y24) / V6;
}

// |det(P2-P1, P3-P1, P4-P1)| / 6.
double tet_volume(const double* vc) {
    double a[3], b[3], c[3], bc[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = vc[3 + d] - vc[d];
        b[d] = vc[6 + d] - vc[d];
        c[d] = vc[9 + d] - vc[d];
    }
    cross3(b, c, bc);
    double det = a[0] * bc[0] + a[1] * bc[1] + a[2] * bc[2];
    return fabs(det) / 6.0;
}

}  // namespace

bool meshgen_box(int n, const double lo[3], const double hi[3], double sphere_r, int periodic,
                 HostMesh& out, std::string& err) {
    if (n < 1 || n > 1024) {
        err = "meshgen_box: n must be in [1, 1024]";
        return false;
    }
    if (periodic && sphere_r > 0) {
        err = "meshgen_box: a periodic box cannot hold an object";
        return false;
    }
    const int64_t n1 = n + 1;
    const int64_t nv_full = n1 * n1 * n1;
    double h[3];
    for (int d = 0; d < 3; ++d) h[d] = (hi[d] - lo[d]) / n;
    std::vector<double> xyz((size_t)nv_full * 3);
    for (int64_t k = 0; k < n1; ++k)
        for (int64_t j = 0; j < n1; ++j)
            for (int64_t i = 0; i < n1; ++i) {
                int64_t v = i + n1 * (j + n1 * k);
                xyz[v * 3 + 0] = (i == n) ? hi[0] : lo[0] + i * h[0];
                xyz[v * 3 + 1] = (j == n) ? hi[1] : lo[1] + j * h[1];
                xyz[v * 3 + 2] = (k == n) ? hi[2] : lo[2] + k * h[2];
            }
    static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    // Cubes in Morton (Z-curve) order of their lattice indices, the six tets
    // of a cube consecutive: cells close in space get close ids, so a range
    // of cell ids (and of cell-sorted particles) is a compact region and a
    // particle crossing into a neighbouring cube usually stays in that range.
    auto spread3 = [](uint64_t v) {
        uint64_t r = 0;
        for (int b = 0; b < 21; ++b) r |= ((v >> b) & 1ull) << (3 * b);
        return r;
    };
    std::vector<std::pair<uint64_t, int64_t>> order;
    order.reserve((size_t)n * n * n);
    for (int64_t k = 0; k < n; ++k)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i)
                order.emplace_back(spread3(i) | (spread3(j) << 1) | (spread3(k) << 2), i + n * (j + n * k));
    std::sort(order.begin(), order.end());
    std::vector<int32_t> cells;
    cells.reserve((size_t)n * n * n * 24);
    const double r2 = sphere_r * sphere_r;
    for (const auto& ok : order) {
        const int64_t i = ok.second % n, j = (ok.second / n) % n, k = ok.second / ((int64_t)n * n);
        for (int p = 0; p < 6; ++p) {
            int64_t ijk[3] = {i, j, k};
            int32_t tv[4];
            tv[0] = (int32_t)(ijk[0] + n1 * (ijk[1] + n1 * ijk[2]));
            for (int s = 0; s < 2; ++s) {
                ijk[perms[p][s]] += 1;
                tv[s + 1] = (int32_t)(ijk[0] + n1 * (ijk[1] + n1 * ijk[2]));
            }
            tv[3] = (int32_t)((i + 1) + n1 * ((j + 1) + n1 * (k + 1)));
            std::sort(tv, tv + 4);
            if (sphere_r > 0) {
                double c[3] = {0, 0, 0};
                for (int q = 0; q < 4; ++q)
                    for (int d = 0; d < 3; ++d) c[d] += xyz[(size_t)tv[q] * 3 + d];
                for (int d = 0; d < 3; ++d) c[d] /= 4.0;
                if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] < r2) continue;
            }
            cells.insert(cells.end(), tv, tv + 4);
        }
    }
    const int64_t nc = (int64_t)cells.size() / 4;
    // Compact away vertices that no cell uses (the sphere interior).
    std::vector<int32_t> remap((size_t)nv_full, -1);
    for (int32_t v : cells) remap[v] = 0;
    int64_t nv = 0;
    std::vector<int64_t> lattice;  // kept vertex -> lattice index
    for (int64_t v = 0; v < nv_full; ++v)
        if (remap[v] == 0) {
            remap[v] = (int32_t)nv++;
            lattice.push_back(v);
        }
    out.nv = nv;
    out.nc = nc;
    out.xyz.resize((size_t)nv * 3);
    for (int64_t v = 0; v < nv; ++v)
        for (int d = 0; d < 3; ++d) out.xyz[v * 3 + d] = xyz[lattice[v] * 3 + d];
    out.cells.resize(cells.size());
    for (size_t q = 0; q < cells.size(); ++q) out.cells[q] = remap[cells[q]];
    // Boundary markers: facets with one adjacent cell are exterior (1) when
    // their three vertices lie on one box face, object (2) otherwise.
    std::vector<int64_t> adj;
    if (!facet_adjacency(out.cells, nc, adj, err)) return false;
    out.facet_bnd.assign((size_t)nc * 4, 0);
    for (int64_t c = 0; c < nc; ++c)
        for (int f = 0; f < 4; ++f) {
            if (adj[c * 4 + f] >= 0) continue;
            int l[3];
            facet_locals(f, l);
            bool on_box = false;
            for (int d = 0; d < 3 && !on_box; ++d) {
                for (int side = 0; side < 2 && !on_box; ++side) {
                    int64_t want = side ? n : 0;
                    bool all = true;
                    for (int q = 0; q < 3; ++q) {
                        int64_t L = lattice[out.cells[c * 4 + l[q]]];
                        int64_t idx[3] = {L % n1, (L / n1) % n1, L / (n1 * n1)};
                        all = all && idx[d] == want;
                    }
                    on_box = all;
                }
            }
            out.facet_bnd[c * 4 + f] = on_box ? 1 : 2;
        }
    out.vdof.resize((size_t)nv);
    if (periodic) {
        for (int64_t v = 0; v < nv; ++v) {
            int64_t L = lattice[v];
            int64_t i = (L % n1) % n, j = ((L / n1) % n1) % n, k = (L / (n1 * n1)) % n;
            out.vdof[v] = (int32_t)(i + n * (j + n * k));
        }
        out.ndof = (int64_t)n * n * n;
    } else {
        for (int64_t v = 0; v < nv; ++v) out.vdof[v] = (int32_t)v;
        out.ndof = nv;
    }
    return true;
}

bool build_topology(const HostMesh& m, MeshTopo& t, std::string& err) {
    const int64_t nc = m.nc, nv = m.nv;
    if (nc <= 0 || nv <= 0) {
        err = "empty mesh";
        return false;
    }
    if ((int64_t)m.cells.size() != nc * 4 || (int64_t)m.xyz.size() != nv * 3 ||
        (int64_t)m.facet_bnd.size() != nc * 4 || (int64_t)m.vdof.size() != nv) {
        err = "mesh arrays have inconsistent sizes";
        return false;
    }
    if (nc >= (int64_t)1 << 30) {
        err = "too many cells (limit 2^30)";
        return false;
    }
    for (int64_t c = 0; c < nc; ++c) {
        const int32_t* v = &m.cells[c * 4];
        for (int i = 0; i < 4; ++i)
            if (v[i] < 0 || v[i] >= nv) {
                err = "cell vertex id out of range";
                return false;
            }
        if (!(v[0] < v[1] && v[1] < v[2] && v[2] < v[3])) {
            err = "cell vertex ids must be strictly ascending (UFC ordering)";
            return false;
        }
    }
    for (int64_t v = 0; v < nv; ++v)
        if (m.vdof[v] < 0 || m.vdof[v] >= m.ndof) {
            err = "vertex dof id out of range";
            return false;
        }
    t.nc = nc;
    t.nv = nv;
    t.ndof = m.ndof;
    std::vector<int64_t> adj;
    if (!facet_adjacency(m.cells, nc, adj, err)) return false;
    t.adj.resize((size_t)nc * 4);
    t.rec.resize((size_t)nc);
    t.planes.resize((size_t)nc);
    t.vcoords.resize((size_t)nc * 12);
    t.ext.clear();
    std::vector<char> seen_bnd(64, 0);
    t.volume = 0.0;
    for (int64_t c = 0; c < nc; ++c) {
        double* vc = &t.vcoords[c * 12];
        for (int i = 0; i < 4; ++i)
            for (int d = 0; d < 3; ++d) vc[i * 3 + d] = m.xyz[(size_t)m.cells[c * 4 + i] * 3 + d];
        CellRec& r = t.rec[c];
        barycentric_matrix(vc, r.bary);
        CellPlanes& P = t.planes[c];
        for (int i = 0; i < 4; ++i) {
            r.dof[i] = m.vdof[m.cells[c * 4 + i]];
            int64_t a = adj[c * 4 + i];
            if (a < 0) {
                int32_t b = m.facet_bnd[c * 4 + i];
                if (b < 1 || b > 63) {
                    err = "boundary facet without a boundary id in [1, 63]";
                    return false;
                }
                seen_bnd[b] = 1;
                a = -(int64_t)b;
            }
            t.adj[c * 4 + i] = a;
            r.nbr[i] = (int32_t)a;
            double n[3], mid[3], area;
            facet_normal_midpoint(vc, i, n, mid, &area);
            double dot = 0;
            for (int d = 0; d < 3; ++d) dot += mid[d] * n[d];
            P.p[4 * i + 0] = -dot;
            for (int d = 0; d < 3; ++d) P.p[4 * i + 1 + d] = n[d];
            if (a == -1) {
                // Exterior facet (reference mesh.cpp:166-236): inward normal,
                // facet vertices in ascending global order, orthonormal basis.
                ExteriorFacetH ef;
                ef.area = area;
                ef.cell = (int32_t)c;
                int l[3];
                facet_locals(i, l);
                for (int q = 0; q < 3; ++q)
                    for (int d = 0; d < 3; ++d) ef.vertices[q * 3 + d] = vc[l[q] * 3 + d];
                for (int d = 0; d < 3; ++d) {
                    ef.normal[d] = -1 * n[d];
                    ef.basis[d * 3] = ef.normal[d];
                }
                double e[3], norm = 0.0;
                for (int d = 0; d < 3; ++d) {
                    e[d] = ef.vertices[d] - ef.vertices[3 + d];
                    norm += e[d] * e[d];
                }
                for (int d = 0; d < 3; ++d) {
                    e[d] /= sqrt(norm);
                    ef.basis[d * 3 + 1] = e[d];
                }
                ef.basis[2] = ef.normal[1] * e[2] - ef.normal[2] * e[1];
                ef.basis[5] = ef.normal[2] * e[0] - ef.normal[0] * e[2];
                ef.basis[8] = ef.normal[0] * e[1] - ef.normal[1] * e[0];
                t.ext.push_back(ef);
            }
        }
        t.volume += tet_volume(vc);
    }
    // Margin for the "certainly inside" shortcut of the relocation walk.
    // Notation for cell c: P_j its vertices, pi_i the affine function of the
    // stored plane of facet i (opposite P_j for j = i), lh_i the affine
    // function of stored barycentric row i, mu_j the exact barycentric
    // coordinates.  Computed values differ from exact ones by at most
    // ep_i = 8u(|c0| + |n|_1 X) and el_i = 8u(|A_i0| + |A_i|_1 X) (lambda_3:
    // the sum of the others plus 16u), u = 2^-53, X >= every |coordinate|.
    // Affine functions are fixed by their vertex values, so with
    // alpha = max |lh_i(P_j) - delta_ij|, eta_i = max_{j != i} |pi_i(P_j)| and
    // h_i = -pi_i(P_i) > 0:  if all computed lambda_i > tol then
    // S = sum |mu_j| <= 1/(1 - 8 alpha), mu_i > tol - el_i - alpha S and the
    // computed pi_i < -(tol - el_i - alpha S) h_i + S eta_i + ep_i, which is
    // negative once tol > el_i + alpha S + (S eta_i + ep_i)/h_i.  The vertex
    // values are themselves computed, so each enters with its own bound.
    {
        double X = 0.0;
        for (double v : m.xyz) X = std::max(X, fabs(v));
        X = 2.0 * X + 1.0;
        const double u = 0x1.0p-53;
        double worst = 0.0, near = 0.0;
        for (int64_t c = 0; c < nc && worst < 1e300; ++c) {
            const double* A = t.rec[c].bary;
            const double* P = &t.vcoords[c * 12];
            double el[4], lv[4][4];  // lv[i][j]: lambda_i at vertex j
            for (int r = 0; r < 3; ++r)
                el[r] = 8 * u * (fabs(A[4 * r]) + (fabs(A[4 * r + 1]) + fabs(A[4 * r + 2]) + fabs(A[4 * r + 3])) * X);
            el[3] = el[0] + el[1] + el[2] + 16 * u;
            for (int j = 0; j < 4; ++j) {
                const double* q = P + 3 * j;
                for (int r = 0; r < 3; ++r)
                    lv[r][j] = ((A[4 * r] + A[4 * r + 1] * q[0]) + A[4 * r + 2] * q[1]) + A[4 * r + 3] * q[2];
                lv[3][j] = ((1 - lv[0][j]) - lv[1][j]) - lv[2][j];
            }
            double alpha = 0.0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) alpha = std::max(alpha, fabs(lv[i][j] - (i == j ? 1.0 : 0.0)) + el[i]);
            if (!(alpha < 0.01)) {
                worst = 1e300;
                break;
            }
            const double S = 1.0 / (1.0 - 8.0 * alpha);
            for (int i = 0; i < 4; ++i) {
                const double* p = t.planes[c].p + 4 * i;
                const double ep = 8 * u * (fabs(p[0]) + (fabs(p[1]) + fabs(p[2]) + fabs(p[3])) * X);
                double h = 0.0, eta = 0.0;
                for (int j = 0; j < 4; ++j) {
                    const double* q = P + 3 * j;
                    const double v = ((p[0] + p[1] * q[0]) + p[2] * q[1]) + p[3] * q[2];
                    if (j == i)
                        h = -v - ep;
                    else
                        eta = std::max(eta, fabs(v) + ep);
                }
                worst = (h > 0) ? std::max(worst, el[i] + alpha * S + (S * eta + ep) / h) : 1e300;
                // |computed proj_i| < D: the exact stored-plane value is below
                // D + ep, and mu_i h = -pi_i(x) + sum_{j != i} mu_j pi_i(P_j),
                // so mu_i < (D + ep + S eta)/h and the computed
                // lambda_i < mu_i + el_i + alpha S.
                if (h > 0) near = std::max(near, (NEAR_FACE_DIST + ep + S * eta) / h + el[i] + alpha * S);
            }
        }
        t.inside_tol = 2 * worst;
        t.coord_bound = X;
        t.near_lam = 2 * near;
        if (!(t.inside_tol < 1e-6)) {  // degenerate mesh: always walk, always test the planes
            t.inside_tol = 0;
            t.near_lam = INFINITY;
        }
    }
    t.bnd_ids.clear();
    for (int b = 1; b < 64; ++b)
        if (seen_bnd[b]) t.bnd_ids.push_back(b);
    // Bounding box and the UniformPosition domain (min/max over all coordinates,
    // reference distributions.h:196-208).
    for (int d = 0; d < 3; ++d) {
        t.bbox_lo[d] = m.xyz[d];
        t.bbox_hi[d] = m.xyz[d];
    }
    t.coord_min = m.xyz[0];
    t.coord_max = m.xyz[0];
    for (int64_t v = 0; v < nv; ++v)
        for (int d = 0; d < 3; ++d) {
            double x = m.xyz[v * 3 + d];
            t.bbox_lo[d] = std::min(t.bbox_lo[d], x);
            t.bbox_hi[d] = std::max(t.bbox_hi[d], x);
            t.coord_min = std::min(t.coord_min, x);
            t.coord_max = std::max(t.coord_max, x);
        }
    // Point-location grid: about six cells per bucket.
    double ext[3], bvol = 1.0;
    for (int d = 0; d < 3; ++d) {
        ext[d] = t.bbox_hi[d] - t.bbox_lo[d];
        if (ext[d] <= 0) ext[d] = 1.0;
        bvol *= ext[d];
    }
    double hb = cbrt(bvol * 6.0 / (double)nc);
    int64_t nb = 1;
    for (int d = 0; d < 3; ++d) {
        int g = (int)llround(ext[d] / hb);
        g = std::max(1, std::min(g, 512));
        t.grid_n[d] = g;
        t.grid_lo[d] = t.bbox_lo[d];
        t.grid_h[d] = ext[d] / g;
        nb *= g;
    }
    std::vector<int32_t> cnt((size_t)nb + 1, 0);
    auto brange = [&](int64_t c, int lo_i[3], int hi_i[3]) {
        const double* vc = &t.vcoords[c * 12];
        for (int d = 0; d < 3; ++d) {
            double mn = vc[d], mx = vc[d];
            for (int i = 1; i < 4; ++i) {
                mn = std::min(mn, vc[i * 3 + d]);
                mx = std::max(mx, vc[i * 3 + d]);
            }
            double a = floor((mn - t.grid_lo[d]) / t.grid_h[d] - 1e-7);
            double b = floor((mx - t.grid_lo[d]) / t.grid_h[d] + 1e-7);
            lo_i[d] = (int)std::max(0.0, std::min(a, (double)t.grid_n[d] - 1));
            hi_i[d] = (int)std::max(0.0, std::min(b, (double)t.grid_n[d] - 1));
        }
    };
    for (int pass = 0; pass < 2; ++pass) {
        std::vector<int32_t> fill;
        if (pass == 1) {
            t.bucket_start.assign((size_t)nb + 1, 0);
            for (int64_t b = 0; b < nb; ++b) t.bucket_start[b + 1] = t.bucket_start[b] + cnt[b];
            t.bucket_cells.assign((size_t)t.bucket_start[nb], 0);
            fill.assign(t.bucket_start.begin(), t.bucket_start.end() - 1);
        }
        for (int64_t c = 0; c < nc; ++c) {
            int a[3], b[3];
            brange(c, a, b);
            for (int k = a[2]; k <= b[2]; ++k)
                for (int j = a[1]; j <= b[1]; ++j)
                    for (int i = a[0]; i <= b[0]; ++i) {
                        int64_t bi = i + (int64_t)t.grid_n[0] * (j + (int64_t)t.grid_n[1] * k);
                        if (pass == 0)
                            cnt[bi]++;
                        else
                            t.bucket_cells[fill[bi]++] = (int32_t)c;
                    }
        }
    }
    return true;
}

void element_volume(const HostMesh& m, const MeshTopo& t, bool voronoi, std::vector<double>& dv) {
    // Vertex -> incident cells in ascending cell order, then per vertex in
    // vertex order add each incident cell volume to the vertex's dof
    // (reference distributor.cpp:44-54).
    std::vector<int64_t> start((size_t)m.nv + 1, 0);
    for (int32_t v : m.cells) start[v + 1]++;
    for (int64_t v = 0; v < m.nv; ++v) start[v + 1] += start[v];
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    std::vector<int32_t> inc((size_t)start[m.nv]);
    for (int64_t c = 0; c < m.nc; ++c)
        for (int i = 0; i < 4; ++i) inc[pos[m.cells[c * 4 + i]]++] = (int32_t)c;
    std::vector<double> cvol((size_t)m.nc);
    for (int64_t c = 0; c < m.nc; ++c) cvol[c] = tet_volume(&t.vcoords[c * 12]);
    dv.assign((size_t)m.ndof, 0.0);
    for (int64_t v = 0; v < m.nv; ++v)
        for (int64_t k = start[v]; k < start[v + 1]; ++k) dv[m.vdof[v]] += cvol[inc[k]];
    const double g_dim = 3;
    for (int64_t i = 0; i < m.ndof; ++i) dv[i] = voronoi ? (g_dim + 1.0) / dv[i] : 1.0 / dv[i];
}

void cell_volumes(const MeshTopo& t, std::vector<double>& vol) {
    vol.resize((size_t)t.nc);
    for (int64_t c = 0; c < t.nc; ++c) vol[c] = tet_volume(&t.vcoords[c * 12]);
}

int64_t locate_host(const MeshTopo& t, const double* x) {
    return locate_in_grid(t.planes.data(), t.bucket_start.data(), t.bucket_cells.data(), t.grid_n,
                          t.grid_lo, t.grid_h, x[0], x[1], x[2]);
}

double maxwellian_flux_num(double vth, const double vd[3], const double n[3], double S) {
    double vdn = 0.0;
    for (int i = 0; i < 3; ++i) vdn = vdn + n[i] * vd[i];
    return S * (vth / (sqrt(2 * M_PI)) * exp(-0.5 * (vdn / vth) * (vdn / vth)) +
                0.5 * vdn * (1. + erf(vdn / (sqrt(2) * vth))));
}

double maxwellian_pdf(double vth, const double vd[3], const double v[3]) {
    double vth2 = vth * vth;
    double factor = (1.0 / (pow(sqrt(2. * M_PI * vth2), 3)));
    double v_sqrt = 0.0;
    for (int i = 0; i < 3; ++i) v_sqrt += (v[i] - vd[i]) * (v[i] - vd[i]);
    return factor * exp(-0.5 * v_sqrt / vth2);
}

double maxwellian_flux_max(double vth, const double vd[3], const double n[3]) {
    double vdn = 0.0;
    for (int i = 0; i < 3; ++i) vdn = vdn + n[i] * vd[i];
    double vth2 = vth * vth;
    double tmp[3];
    for (int i = 0; i < 3; ++i)
        tmp[i] = vd[i] - 0.5 * n[i] * vdn + 0.5 * n[i] * sqrt(4 * vth2 + vdn * vdn);
    double vn = 0.0;
    for (int i = 0; i < 3; ++i) vn += tmp[i] * n[i];
    return (vn > 0.0) * vn * maxwellian_pdf(vth, vd, tmp);
}

// ---------------------------------------------------------------------------
// Velocity distributions: constructors, closed-form flux tables and maxima
// (reference distributions.cpp:23-245).
// ---------------------------------------------------------------------------
static double inner3(const double* n, const double* vd) {
    double vdn = 0.0;  // std::inner_product(n, n + 3, vd, 0.0)
    for (int i = 0; i < 3; ++i) vdn = vdn + n[i] * vd[i];
    return vdn;
}

bool vdf_setup(int kind, double vth, const double vd[3], double k, double alpha, double range, VdfParams& f,
               std::string& err) {
    const int dim = 3;
    memset(&f, 0, sizeof(f));
    if (kind < VDF_MAXWELLIAN || kind > VDF_KAPPA_CAIRNS) {
        err = "unknown velocity distribution kind";
        return false;
    }
    if (!(vth > 0)) {
        err = "vth must be positive";
        return false;
    }
    if ((kind == VDF_KAPPA || kind == VDF_KAPPA_CAIRNS) && !(k > 1.5)) {
        err = "kappa must be bigger than 3/2";  // distributions.cpp:84
        return false;
    }
    f.kind = kind;
    f.vth = vth;
    for (int i = 0; i < 3; ++i) f.vd[i] = vd[i];
    f.k = k;
    f.alpha = alpha;
    f.range = range;
    f.vth2 = vth * vth;
    f.vth4 = f.vth2 * f.vth2;
    switch (kind) {
        case VDF_KAPPA:  // distributions.cpp:96-97
            f.factor = (1.0 / pow(sqrt(M_PI * (2 * k - dim) * f.vth2), dim)) *
                       (tgamma(k + 1.0) / tgamma(k + ((2.0 - dim) / 2.0)));
            break;
        case VDF_CAIRNS:  // distributions.cpp:142
            f.factor = (1.0 / (pow(sqrt(2 * M_PI * f.vth2), dim) * (1 + dim * (dim + 2) * alpha)));
            break;
        case VDF_KAPPA_CAIRNS:  // distributions.cpp:196-198
            f.factor = (1.0 / pow(sqrt(M_PI * (2. * k - dim) * f.vth2), dim)) *
                       (1.0 / (1.0 + dim * (dim + 2.0) * alpha * ((k - dim / 2.0) / (k - (dim + 2) / 2.0)))) *
                       (tgamma(k + 1.0) / tgamma(k + ((2.0 - dim) / 2.0)));
            break;
        default:  // distributions.cpp:28
            f.factor = (1.0 / (pow(sqrt(2. * M_PI * f.vth2), dim)));
    }
    if (!std::isfinite(f.factor) || !(f.factor > 0)) {
        err = "velocity distribution normalisation is not a positive finite number for these parameters";
        return false;
    }
    return true;
}

bool vdf_has_flux_num(const VdfParams& f) {
    switch (f.kind) {
        case VDF_KAPPA: {
            // distributions.cpp:86-94: std::accumulate(vd, 0) sums into an
            // int (the initial value's type), truncating at every step, so
            // drifts that truncate to zero still take the no-drift formula.
            int sum_vd = 0;
            for (int i = 0; i < 3; ++i) sum_vd = (int)(sum_vd + f.vd[i]);
            return sum_vd == 0;
        }
        case VDF_CAIRNS:
            return true;
        case VDF_KAPPA_CAIRNS:
            return false;
        default:
            return true;
    }
}

bool vdf_has_flux_max(const VdfParams& f) { return f.kind == VDF_MAXWELLIAN || f.kind == VDF_KAPPA; }

double vdf_flux_num(const VdfParams& f, const double n[3], double S) {
    const int dim = 3;
    const double vth = f.vth, k = f.k, alpha = f.alpha;
    switch (f.kind) {
        case VDF_KAPPA:  // distributions.cpp:110-118 (no drift)
            return S * ((vth / (sqrt(2 * M_PI))) * (1.0 / sqrt(k - dim / 2.0)) *
                        (tgamma(k - ((dim - 1.0) / 2.0)) / tgamma(k - dim / 2.0)));
        case VDF_CAIRNS: {  // distributions.cpp:175-185
            const double vdn = inner3(n, f.vd);
            return S * ((vth / (sqrt(2 * M_PI))) * exp(-0.5 * (vdn / vth) * (vdn / vth)) *
                            (1. + (dim + 1.) * (dim + 3.) * alpha + (vdn / vth) * (vdn / vth) * alpha) /
                            (1. + dim * (dim + 2.) * alpha) +
                        0.5 * vdn * (1. + erf(vdn / (sqrt(2) * vth))));
        }
        default:
            return maxwellian_flux_num(vth, f.vd, n, S);
    }
}

double vdf_flux_max(const VdfParams& f, const double n[3]) {
    if (f.kind == VDF_KAPPA) {  // distributions.cpp:120-128 (evaluated without the drift shift)
        const int dim = 3;
        double tmp[3];
        for (int i = 0; i < 3; ++i) tmp[i] = sqrt((2. * f.k - dim) / (2. * f.k + 1.0)) * n[i] * f.vth;
        return vdf_flux(f, tmp[0], tmp[1], tmp[2], n[0], n[1], n[2]);
    }
    return maxwellian_flux_max(f.vth, f.vd, n);
}

double vdf_max(const VdfParams& f) {
    const int dim = 3;
    const double k = f.k, alpha = f.alpha;
    double v_max[3] = {f.vd[0], f.vd[1], f.vd[2]};
    switch (f.kind) {
        case VDF_CAIRNS: {  // distributions.cpp:156-173
            if (alpha < 0.25) return f.factor;
            v_max[0] += f.vth * sqrt(2.0 + sqrt(4.0 - 1.0 / alpha));
            const double m = vdf_eval(f, v_max[0], v_max[1], v_max[2]);
            return f.factor > m ? f.factor : m;
        }
        case VDF_KAPPA_CAIRNS:  // distributions.cpp:213-230
            if (alpha >= (k - 1.) * (k + 1.) / ((2. * k - dim) * (2. * k - dim))) {
                v_max[0] += f.vth * sqrt(((2.0 * k - dim) / (k - 1.0)) +
                                         sqrt(alpha * alpha * (2 * k - dim) * (2 * k - dim) -
                                              alpha * (k - 1.0) * (k + 1.0)) /
                                             (alpha * (k - 1.0)));
                return vdf_eval(f, v_max[0], v_max[1], v_max[2]);
            }
            return f.factor;
        default:  // Maxwellian, Kappa: the value at the drift (distributions.h:244, 278)
            return f.factor;
    }
}

}  // namespace punc_gpu
