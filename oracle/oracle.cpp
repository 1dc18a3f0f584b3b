// ORACLE — test infrastructure only.
//
// A serial CPU restatement of the particle half of the PUNC++ PIC step,
// written to check the HIP path: it is imported only by tests/, by
// __graft_entry__.smoke() and by bench.py's CPU-baseline leg, never by the
// product library.  It keeps the reference's data layout and loop order
// (per-cell particle buckets of AoS particles, cell-by-cell loops, swap-with-
// back deletion) and its floating-point expression order, and is compiled
// with g++ -O2 -ffp-contract=off (no FMA), like the reference's x86-64 build.
//
// Restated reference functions (file:line in the PUNC++ snapshot):
//   Cell<3>::init_barycentric_matrix   punc/src/population.cpp:181-240
//   Cell<3>::barycentric               punc/include/punc/population.h:885-893
//   Population::init_localizer         population.h:345-389
//   Population::add_particles          population.h:411-435 (out-of-mesh points skipped)
//   Population::relocate_fast          population.h:539-569 (iterative, hop cap)
//   Population::update                 population.h:641-699
//   num_of_particles/positives/negatives  population.h:701-744
//   accel_cg1                          punc/include/punc/pusher.h:154-201
//   boris_cg1, cross                   pusher.h:282-372, 683-691
//   move, move_periodic                pusher.h:648-681
//   distribute_cg1                     punc/include/punc/distributor.h:168-210
//   element_volume                     punc/src/distributor.cpp:35-69
//   kinetic_energy                     punc/include/punc/diagnostics.h:248-265
//   particle_potential_energy_cg1      diagnostics.h:379-412
//   density_dg0, density_cg1           diagnostics.h:432-465, 492-555
//   Population::statistics             population.h:746-805
//   Population::save_file / load_file  population.h:807-880
//   inject_particles                   punc/include/punc/injector.h:157-232
//   rejection_sampler, random_facet_points, create_flux  punc/src/injector.cpp:25-171
//   load_particles                     injector.h:244-278
//   Maxwellian pdf/icdf/flux_num_particles/flux_max  punc/src/distributions.cpp:23-77
//   Kappa, Cairns, KappaCairns (pdf, max, flux number/max, flags)  distributions.cpp:79-245
//   create_flux Monte Carlo branches                  injector.cpp:127-168
//   Pdf flux operator, update_domain   punc/include/punc/distributions.h:93-102, 139-146
//   Mesh::exterior_boundaries          punc/src/mesh.cpp:166-236
//
// Deliberate, documented differences from the reference:
//   * RNG: the reference draws from an unseeded std::mt19937_64
//     (injector.h:161).  Here every draw comes from Philox4x64-10 keyed by
//     (seed, tag) with counters (step/slot, ...), the same streams the GPU uses.
//   * Point location: the reference asks DOLFIN's bounding-box tree for the
//     first colliding cell (population.cpp:24-39).  Here: the lowest-id cell
//     whose four facet planes all give proj <= 0.  Injected particles are
//     located by relocate_fast from the facet's owning cell.
//   * Facet normals/midpoints: DOLFIN's Cell::normal/Facet::midpoint are
//     replaced by n = (P2-P1)x(P3-P1)/|.| oriented away from the opposite
//     vertex, midpoint ((P1+P2)+P3)/3; cell volume |det|/6.
//   * erfc_inv (Boost) is computed by safeguarded Newton iteration on erfc.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>
#include <array>
#include <map>
#include <vector>

namespace {

constexpr int MAX_BND = 64;
constexpr int64_t HOP_CAP = 100000;
constexpr int64_t WALK_FAILED = -(int64_t(1) << 30);

struct Particle {
    double x[3];
    double v[3];
    double q;
    double m;
    int32_t species;  // bookkeeping only (absorbed counts per species)
};

struct Cell {
    int64_t id;
    int32_t verts[4];
    int32_t dofs[4];
    double vertex_coordinates[12];
    double barycentric_matrix[12];
    int64_t facet_adjacents[4];
    double facet_plane_coeffs[16];
    std::vector<Particle> particles;

    void barycentric(const double* x, double* y) const {
        const double* A = barycentric_matrix;
        y[0] = A[0] + A[1] * x[0] + A[2] * x[1] + A[3] * x[2];
        y[1] = A[4] + A[5] * x[0] + A[6] * x[1] + A[7] * x[2];
        y[2] = A[8] + A[9] * x[0] + A[10] * x[1] + A[11] * x[2];
        y[3] = 1 - y[0] - y[1] - y[2];
    }
};

struct ExtFacet {
    double area;
    double vertices[9];
    double normal[3];
    double basis[9];
    int64_t cell;
};

struct Species {
    double q, m;
    bool inject = false;
    double vth = 0, vd[3] = {0, 0, 0}, n = 0, vrange = 5.0;
    // velocity distribution: 0 Maxwellian, 1 Kappa, 2 Cairns, 3 KappaCairns
    int kind = 0;
    double k = 0, alpha = 0;
    std::vector<double> num_particles, pdf_max;  // per exterior facet
};

struct Pop {
    int64_t nv, nc, ndof;
    std::vector<double> xyz;
    std::vector<int32_t> vdof;
    std::vector<Cell> cells;
    std::vector<ExtFacet> ext;
    std::vector<Species> species;
    double cmin, cmax;
    // location grid
    int gn;
    double glo[3], gh[3];
    std::vector<std::vector<int64_t>> buckets;
};

// ---------------------------------------------------------------------------
// Philox4x64-10, 64x64->128 products from 32-bit limbs.
// ---------------------------------------------------------------------------
void mul64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
    const uint64_t a0 = a & 0xffffffffu, a1 = a >> 32, b0 = b & 0xffffffffu, b1 = b >> 32;
    const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffu) + (p10 & 0xffffffffu);
    *lo = (mid << 32) | (p00 & 0xffffffffu);
    *hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

std::array<uint64_t, 4> philox(std::array<uint64_t, 4> c, uint64_t k0, uint64_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t h0, l0, h1, l1;
        mul64(0xD2E7470EE14C6C93ull, c[0], &h0, &l0);
        mul64(0xCA5A826395121157ull, c[2], &h1, &l1);
        c = {h1 ^ c[1] ^ k0, l1, h0 ^ c[3] ^ k1, l0};
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
    return c;
}

// Uniform in (0,1): odd multiples of 2^-53.
double unif(uint64_t x) { return (double)((x >> 12) * 2 + 1) / 9007199254740992.0; }

constexpr uint64_t TAG_INJECT_COUNT = 0x494e4a43ull;
constexpr uint64_t TAG_INJECT = 0x494e4a50ull;
constexpr uint64_t TAG_LOAD = 0x4c4f4144ull;
constexpr uint64_t TAG_FLUX = 0x464c5558ull;

// erfc^-1 by Newton steps inside a shrinking bracket.  Near y = 1 the
// equation is solved as erf(x) = 1 - y (the subtraction is exact there), so
// small results keep full relative accuracy.
double erfc_inv_newton(double y) {
    if (y > 1.0) return -erfc_inv_newton(2.0 - y);  // erfc(-x) = 2 - erfc(x); 2 - y is exact
    double lo = -30.0, hi = 30.0;  // erfc decreasing: erfc(lo) = 2 > y > 0 = erfc(hi)
    double x = 0.0;
    // In the tail (y < 0.01) the Newton steps act on log(erfc(x)) - log(y),
    // which converges in a few steps where erfc itself is tiny.
    const bool central = y >= 0.5 && y <= 1.5;
    const bool tail = y < 0.01;
    const double z = 1.0 - y, ly = log(y);
    for (int it = 0; it < 200; ++it) {
        const double ec = erfc(x);
        const double f = central ? z - erf(x) : (tail ? log(ec) - ly : ec - y);
        if (f == 0.0) return x;
        if (f > 0) lo = x; else hi = x;
        double d = -2.0 / sqrt(M_PI) * exp(-x * x);
        if (tail) d /= ec;
        double xn = x - f / d;
        if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
        if (fabs(xn - x) <= 1e-17 * fabs(x) || xn == x) return xn;
        x = xn;
    }
    return x;
}

// ---------------------------------------------------------------------------
// Geometry (own conventions replacing DOLFIN's, see header)
// ---------------------------------------------------------------------------
void init_barycentric_matrix(Cell& cell) {
    const double* vc = cell.vertex_coordinates;
    double* bm = cell.barycentric_matrix;
    double x1 = vc[0]; double y1 = vc[1]; double z1 = vc[2];
    double x2 = vc[3]; double y2 = vc[4]; double z2 = vc[5];
    double x3 = vc[6]; double y3 = vc[7]; double z3 = vc[8];
    double x4 = vc[9]; double y4 = vc[10]; double z4 = vc[11];
    double x13 = x1 - x3; double y13 = y1 - y3; double z13 = z1 - z3;
    double x14 = x1 - x4; double y14 = y1 - y4; double z14 = z1 - z4;
    double x24 = x2 - x4; double y24 = y2 - y4; double z24 = z2 - z4;
    double x31 = x3 - x1; double y31 = y3 - y1; double z31 = z3 - z1;
    double x32 = x3 - x2; double y32 = y3 - y2; double z32 = z3 - z2;
    double x34 = x3 - x4; double y34 = y3 - y4; double z34 = z3 - z4;
    double x42 = x4 - x2; double y42 = y4 - y2; double z42 = z4 - z2;
    double x43 = x4 - x3; double y43 = y4 - y3; double z43 = z4 - z3;
    double V01 = (x2 * (y3 * z4 - y4 * z3) + x3 * (y4 * z2 - y2 * z4) + x4 * (y2 * z3 - y3 * z2)) / 6.0;
    double V02 = (x1 * (y4 * z3 - y3 * z4) + x3 * (y1 * z4 - y4 * z1) + x4 * (y3 * z1 - y1 * z3)) / 6.0;
    double V03 = (x1 * (y2 * z4 - y4 * z2) + x2 * (y4 * z1 - y1 * z4) + x4 * (y1 * z2 - y2 * z1)) / 6.0;
    double V04 = (x1 * (y3 * z2 - y2 * z3) + x2 * (y1 * z3 - y3 * z1) + x3 * (y2 * z1 - y1 * z2)) / 6.0;
    double V = V01 + V02 + V03 + V04;
    double V6 = 6 * V;
    bm[0] = V01 / V;
    bm[4] = V02 / V;
    bm[8] = V03 / V;
    bm[1] = (y42 * z32 - y32 * z42) / V6;
    bm[5] = (y31 * z43 - y34 * z13) / V6;
    bm[9] = (y24 * z14 - y14 * z24) / V6;
    bm[2] = (x32 * z42 - x42 * z32) / V6;
    bm[6] = (x43 * z31 - x13 * z34) / V6;
    bm[10] = (x14 * z24 - x24 * z14) / V6;
    bm[3] = (x42 * y32 - x32 * y42) / V6;
    bm[7] = (x31 * y43 - x34 * y13) / V6;
    bm[11] = (x24 * y14 - x14 * y24) / V6;
}

// Outward unit normal of local facet i (opposite local vertex i), facet
// midpoint and area.
void cell_normal(const Cell& c, int i, double n[3], double mid[3], double* area) {
    int f[3], k = 0;
    for (int j = 0; j < 4; ++j)
        if (j != i) f[k++] = j;
    const double* P0 = c.vertex_coordinates + 3 * i;
    const double* P1 = c.vertex_coordinates + 3 * f[0];
    const double* P2 = c.vertex_coordinates + 3 * f[1];
    const double* P3 = c.vertex_coordinates + 3 * f[2];
    double V0[3], V1[3], V2[3];
    for (int d = 0; d < 3; ++d) {
        V0[d] = P0[d] - P1[d];
        V1[d] = P2[d] - P1[d];
        V2[d] = P3[d] - P1[d];
    }
    n[0] = V1[1] * V2[2] - V1[2] * V2[1];
    n[1] = V1[2] * V2[0] - V1[0] * V2[2];
    n[2] = V1[0] * V2[1] - V1[1] * V2[0];
    const double norm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= norm;
    n[1] /= norm;
    n[2] /= norm;
    if (n[0] * V0[0] + n[1] * V0[1] + n[2] * V0[2] > 0.0) {
        n[0] *= -1.0;
        n[1] *= -1.0;
        n[2] *= -1.0;
    }
    for (int d = 0; d < 3; ++d) {
        double s = P1[d];
        s += P2[d];
        s += P3[d];
        mid[d] = s / 3.0;
    }
    if (area) *area = norm / 2.0;
}

double cell_volume(const Cell& c) {
    const double* v = c.vertex_coordinates;
    double a[3], b[3], e[3];
    for (int d = 0; d < 3; ++d) {
        a[d] = v[3 + d] - v[d];
        b[d] = v[6 + d] - v[d];
        e[d] = v[9 + d] - v[d];
    }
    const double det = a[0] * (b[1] * e[2] - b[2] * e[1]) + a[1] * (b[2] * e[0] - b[0] * e[2]) +
                       a[2] * (b[0] * e[1] - b[1] * e[0]);
    return fabs(det) / 6.0;
}

// ---------------------------------------------------------------------------
// Location
// ---------------------------------------------------------------------------
bool inside(const Cell& c, const double* p) {
    const double* coeffs = c.facet_plane_coeffs;
    for (int i = 0; i < 4; ++i) {
        double proj = *coeffs++;
        for (int j = 0; j < 3; ++j) proj += *coeffs++ * p[j];
        if (proj > 0) return false;
    }
    return true;
}

int64_t locate(const Pop& P, const double* p) {
    // non-finite points are outside the mesh (a NaN passes every plane test)
    for (int d = 0; d < 3; ++d)
        if (!std::isfinite(p[d])) return -1;
    int b[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor((p[d] - P.glo[d]) / P.gh[d]);
        if (f < 0) f = 0;
        if (f > P.gn - 1) f = P.gn - 1;
        b[d] = (int)f;
    }
    const auto& cand = P.buckets[b[0] + P.gn * (b[1] + P.gn * b[2])];
    for (int64_t c : cand)
        if (inside(P.cells[c], p)) return c;
    return -1;
}

int64_t locate_brute(const Pop& P, const double* p) {
    for (int d = 0; d < 3; ++d)
        if (!std::isfinite(p[d])) return -1;
    for (int64_t c = 0; c < P.nc; ++c)
        if (inside(P.cells[c], p)) return c;
    return -1;
}

// relocate_fast (population.h:539-569), iterative.
int64_t relocate_fast(const Pop& P, const double* p, int64_t cell_id) {
    for (int64_t hop = 0; hop < HOP_CAP; ++hop) {
        const Cell& c = P.cells[cell_id];
        const double* coeffs = c.facet_plane_coeffs;
        int64_t next = cell_id;
        for (int i = 0; i < 4; ++i) {
            double proj = *coeffs++;
            for (int j = 0; j < 3; ++j) proj += *coeffs++ * p[j];
            if (proj > 0) {
                next = c.facet_adjacents[i];
                break;
            }
        }
        if (next == cell_id) return cell_id;
        if (next < 0) return next;
        cell_id = next;
    }
    return WALK_FAILED;
}

void build(Pop& P, const int32_t* cells, const int32_t* facet_bnd) {
    P.cells.resize(P.nc);
    std::map<std::array<int32_t, 3>, std::vector<std::pair<int64_t, int>>> facets;
    for (int64_t c = 0; c < P.nc; ++c) {
        Cell& cell = P.cells[c];
        cell.id = c;
        for (int i = 0; i < 4; ++i) {
            cell.verts[i] = cells[c * 4 + i];
            cell.dofs[i] = P.vdof[cell.verts[i]];
            for (int d = 0; d < 3; ++d) cell.vertex_coordinates[i * 3 + d] = P.xyz[cell.verts[i] * 3 + d];
        }
        init_barycentric_matrix(cell);
        for (int i = 0; i < 4; ++i) {
            std::array<int32_t, 3> key;
            int k = 0;
            for (int j = 0; j < 4; ++j)
                if (j != i) key[k++] = cell.verts[j];
            facets[key].push_back({c, i});
        }
    }
    for (auto& kv : facets) {
        auto& lst = kv.second;
        for (auto& ci : lst) {
            Cell& cell = P.cells[ci.first];
            int64_t adj = -1;
            for (auto& o : lst)
                if (o.first != ci.first) adj = o.first;
            if (lst.size() == 1) adj = -1 * (int64_t)facet_bnd[ci.first * 4 + ci.second];
            cell.facet_adjacents[ci.second] = adj;
        }
    }
    for (int64_t c = 0; c < P.nc; ++c) {
        Cell& cell = P.cells[c];
        for (int i = 0; i < 4; ++i) {
            double n[3], mid[3], area;
            cell_normal(cell, i, n, mid, &area);
            double dot_product = 0;
            for (int j = 0; j < 3; ++j) dot_product += mid[j] * n[j];
            cell.facet_plane_coeffs[4 * i] = -dot_product;
            for (int j = 0; j < 3; ++j) cell.facet_plane_coeffs[4 * i + 1 + j] = n[j];
            if (cell.facet_adjacents[i] == -1) {
                ExtFacet ef;
                ef.area = area;
                ef.cell = c;
                int f[3], k = 0;
                for (int j = 0; j < 4; ++j)
                    if (j != i) f[k++] = j;
                for (int q = 0; q < 3; ++q)
                    for (int d = 0; d < 3; ++d) ef.vertices[q * 3 + d] = cell.vertex_coordinates[f[q] * 3 + d];
                double vertex[3], norm = 0.0;
                for (int j = 0; j < 3; ++j) {
                    ef.normal[j] = -1 * n[j];
                    ef.basis[j * 3] = ef.normal[j];
                }
                for (int d = 0; d < 3; ++d) {
                    vertex[d] = ef.vertices[d] - ef.vertices[3 + d];
                    norm += vertex[d] * vertex[d];
                }
                for (int d = 0; d < 3; ++d) {
                    vertex[d] /= sqrt(norm);
                    ef.basis[d * 3 + 1] = vertex[d];
                }
                ef.basis[2] = ef.normal[1] * vertex[2] - ef.normal[2] * vertex[1];
                ef.basis[5] = ef.normal[2] * vertex[0] - ef.normal[0] * vertex[2];
                ef.basis[8] = ef.normal[0] * vertex[1] - ef.normal[1] * vertex[0];
                P.ext.push_back(ef);
            }
        }
    }
    // Location grid: about two cells per bucket, cube buckets over the bbox.
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = P.xyz[d];
    P.cmin = P.cmax = P.xyz[0];
    for (int64_t v = 0; v < P.nv; ++v)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], P.xyz[v * 3 + d]);
            hi[d] = std::max(hi[d], P.xyz[v * 3 + d]);
            P.cmin = std::min(P.cmin, P.xyz[v * 3 + d]);
            P.cmax = std::max(P.cmax, P.xyz[v * 3 + d]);
        }
    P.gn = std::max(1, std::min(200, (int)cbrt(P.nc / 2.0)));
    for (int d = 0; d < 3; ++d) {
        P.glo[d] = lo[d];
        P.gh[d] = (hi[d] - lo[d]) / P.gn;
        if (P.gh[d] <= 0) P.gh[d] = 1.0;
    }
    P.buckets.assign((size_t)P.gn * P.gn * P.gn, {});
    for (int64_t c = 0; c < P.nc; ++c) {
        int a[3], b[3];
        for (int d = 0; d < 3; ++d) {
            double mn = 1e300, mx = -1e300;
            for (int i = 0; i < 4; ++i) {
                mn = std::min(mn, P.cells[c].vertex_coordinates[i * 3 + d]);
                mx = std::max(mx, P.cells[c].vertex_coordinates[i * 3 + d]);
            }
            double fa = floor((mn - P.glo[d]) / P.gh[d]) - 1, fb = floor((mx - P.glo[d]) / P.gh[d]) + 1;
            a[d] = (int)std::max(0.0, fa);
            b[d] = (int)std::min((double)P.gn - 1, fb);
        }
        for (int k = a[2]; k <= b[2]; ++k)
            for (int j = a[1]; j <= b[1]; ++j)
                for (int i = a[0]; i <= b[0]; ++i) P.buckets[i + P.gn * (j + P.gn * k)].push_back(c);
    }
}

}  // namespace

// ===========================================================================
// C ABI (ctypes)
// ===========================================================================
extern "C" {

void oracle_philox(const uint64_t* ctr, const uint64_t* key, uint64_t* out) {
    auto r = philox({ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

double oracle_unif(uint64_t x) { return unif(x); }
double oracle_erfc_inv(double y) { return erfc_inv_newton(y); }

void* oracle_create(int64_t nv, const double* xyz, int64_t nc, const int32_t* cells,
                    const int32_t* facet_bnd, const int32_t* vdof, int64_t ndof) {
    Pop* P = new Pop();
    P->nv = nv;
    P->nc = nc;
    P->ndof = ndof;
    P->xyz.assign(xyz, xyz + nv * 3);
    P->vdof.assign(vdof, vdof + nv);
    build(*P, cells, facet_bnd);
    return P;
}

void oracle_destroy(void* h) { delete (Pop*)h; }

void oracle_set_species(void* h, int nsp, const double* q, const double* m) {
    Pop& P = *(Pop*)h;
    P.species.assign(nsp, Species());
    for (int s = 0; s < nsp; ++s) {
        P.species[s].q = q[s];
        P.species[s].m = m[s];
    }
}

void oracle_localizer(void* h, int64_t* adj, double* planes, double* bary, double* vcoords) {
    Pop& P = *(Pop*)h;
    for (int64_t c = 0; c < P.nc; ++c) {
        const Cell& cell = P.cells[c];
        if (adj) memcpy(adj + 4 * c, cell.facet_adjacents, 4 * 8);
        if (planes) memcpy(planes + 16 * c, cell.facet_plane_coeffs, 16 * 8);
        if (bary) memcpy(bary + 12 * c, cell.barycentric_matrix, 12 * 8);
        if (vcoords) memcpy(vcoords + 12 * c, cell.vertex_coordinates, 12 * 8);
    }
}

int64_t oracle_num_exterior(void* h) { return (int64_t)((Pop*)h)->ext.size(); }

void oracle_exterior_facets(void* h, double* area, double* vertices, double* normal, double* basis,
                            int64_t* cell) {
    Pop& P = *(Pop*)h;
    for (size_t j = 0; j < P.ext.size(); ++j) {
        area[j] = P.ext[j].area;
        memcpy(vertices + 9 * j, P.ext[j].vertices, 72);
        memcpy(normal + 3 * j, P.ext[j].normal, 24);
        memcpy(basis + 9 * j, P.ext[j].basis, 72);
        cell[j] = P.ext[j].cell;
    }
}

// element_volume (distributor.cpp:35-69): vertex loop, incident cells in id order.
void oracle_element_volume(void* h, int voronoi, double* out) {
    Pop& P = *(Pop*)h;
    std::vector<std::vector<int64_t>> vcells(P.nv);
    for (int64_t c = 0; c < P.nc; ++c)
        for (int i = 0; i < 4; ++i) vcells[P.cells[c].verts[i]].push_back(c);
    std::vector<double> volumes(P.ndof, 0.0);
    for (int64_t j = 0; j < P.nv; ++j)
        for (int64_t c : vcells[j]) volumes[P.vdof[j]] += cell_volume(P.cells[c]);
    const int g_dim = 3;
    for (int64_t i = 0; i < P.ndof; ++i) volumes[i] = voronoi ? (g_dim + 1.0) / volumes[i] : 1.0 / volumes[i];
    memcpy(out, volumes.data(), P.ndof * 8);
}

int64_t oracle_locate(void* h, const double* x) { return locate(*(Pop*)h, x); }
int64_t oracle_locate_brute(void* h, const double* x) { return locate_brute(*(Pop*)h, x); }
int64_t oracle_relocate(void* h, const double* x, int64_t cell) { return relocate_fast(*(Pop*)h, x, cell); }

// add_particles (population.h:411-435), skipping points outside the mesh.
int64_t oracle_add_particles(void* h, int species, int64_t n, const double* xs, const double* vs) {
    Pop& P = *(Pop*)h;
    int64_t added = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t cell_id = locate(P, xs + 3 * i);
        if (cell_id >= 0) {
            Particle p;
            for (int j = 0; j < 3; ++j) {
                p.x[j] = xs[3 * i + j];
                p.v[j] = vs[3 * i + j];
            }
            p.q = P.species[species].q;
            p.m = P.species[species].m;
            p.species = species;
            P.cells[cell_id].particles.push_back(p);
            ++added;
        }
    }
    return added;
}

void oracle_add_located(void* h, int64_t n, const double* xs, const double* vs, const int32_t* sp,
                        const int64_t* cells) {
    Pop& P = *(Pop*)h;
    for (int64_t i = 0; i < n; ++i) {
        Particle p;
        for (int j = 0; j < 3; ++j) {
            p.x[j] = xs[3 * i + j];
            p.v[j] = vs[3 * i + j];
        }
        p.q = P.species[sp[i]].q;
        p.m = P.species[sp[i]].m;
        p.species = sp[i];
        P.cells[cells[i]].particles.push_back(p);
    }
}

int64_t oracle_size(void* h) {
    Pop& P = *(Pop*)h;
    int64_t num_particles = 0;
    for (auto& cell : P.cells) num_particles += cell.particles.size();
    return num_particles;
}

int64_t oracle_num_positives(void* h) {
    int64_t k = 0;
    for (auto& cell : ((Pop*)h)->cells)
        for (auto& p : cell.particles)
            if (p.q > 0) k++;
    return k;
}

int64_t oracle_num_negatives(void* h) {
    int64_t k = 0;
    for (auto& cell : ((Pop*)h)->cells)
        for (auto& p : cell.particles)
            if (p.q < 0) k++;
    return k;
}

void oracle_get(void* h, double* xs, double* vs, int32_t* sp, int64_t* cells) {
    Pop& P = *(Pop*)h;
    int64_t k = 0;
    for (auto& cell : P.cells)
        for (auto& p : cell.particles) {
            for (int j = 0; j < 3; ++j) {
                xs[3 * k + j] = p.x[j];
                vs[3 * k + j] = p.v[j];
            }
            sp[k] = p.species;
            cells[k] = cell.id;
            ++k;
        }
}

// accel_cg1 (pusher.h:154-201).  E: ndof*3 (component j of dof d at E[3d+j]).
double oracle_accel(void* h, const double* E, double dt) {
    Pop& P = *(Pop*)h;
    double KE = 0.0;
    double coeffs[4];
    double values[12];
    for (auto& cell : P.cells) {
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 4; ++i) values[j * 4 + i] = E[3 * cell.dofs[i] + j];
        for (auto& particle : cell.particles) {
            double m = particle.m;
            double q = particle.q;
            auto& vel = particle.v;
            auto& x = particle.x;
            cell.barycentric(x, coeffs);
            for (int j = 0; j < 3; j++) {
                double Ei = 0.0;
                for (int i = 0; i < 4; ++i) Ei += coeffs[i] * values[j * 4 + i];
                Ei *= dt * (q / m);
                KE += 0.5 * m * vel[j] * (vel[j] + Ei);
                particle.v[j] += Ei;
            }
        }
    }
    return KE;
}

// boris_cg1 (pusher.h:282-372) with cross (pusher.h:683-691).
double oracle_boris(void* h, const double* E, const double* B, double dt) {
    Pop& P = *(Pop*)h;
    double KE = 0.0;
    double t_mag2;
    double v_minus[3], v_prime[3], v_plus[3], t[3], s[3], Ei[3], coeffs[4], values[12];
    for (auto& cell : P.cells) {
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 4; ++i) values[j * 4 + i] = E[3 * cell.dofs[i] + j];
        for (auto& particle : cell.particles) {
            double m = particle.m;
            double q = particle.q;
            auto& vel = particle.v;
            auto& x = particle.x;
            cell.barycentric(x, coeffs);
            for (int j = 0; j < 3; j++) {
                Ei[j] = 0.0;
                for (int i = 0; i < 4; ++i) Ei[j] += coeffs[i] * values[j * 4 + i];
            }
            t_mag2 = 0.0;
            for (int i = 0; i < 3; ++i) {
                t[i] = tan((dt * q / (2.0 * m)) * B[i]);
                t_mag2 += t[i] * t[i];
            }
            for (int i = 0; i < 3; ++i) s[i] = 2 * t[i] / (1 + t_mag2);
            for (int i = 0; i < 3; ++i) v_minus[i] = vel[i] + 0.5 * dt * (q / m) * Ei[i];
            for (int i = 0; i < 3; i++) KE += 0.5 * m * v_minus[i] * v_minus[i];
            double r[3];
            r[0] = v_minus[1] * t[2] - v_minus[2] * t[1];
            r[1] = -v_minus[0] * t[2] + v_minus[2] * t[0];
            r[2] = v_minus[0] * t[1] - v_minus[1] * t[0];
            for (int i = 0; i < 3; ++i) v_prime[i] = v_minus[i] + r[i];
            r[0] = v_prime[1] * s[2] - v_prime[2] * s[1];
            r[1] = -v_prime[0] * s[2] + v_prime[2] * s[0];
            r[2] = v_prime[0] * s[1] - v_prime[1] * s[0];
            for (int i = 0; i < 3; ++i) v_plus[i] = v_minus[i] + r[i];
            for (int i = 0; i < 3; ++i) particle.v[i] = v_plus[i] + 0.5 * dt * (q / m) * Ei[i];
        }
    }
    return KE;
}

void oracle_move(void* h, double dt) {
    for (auto& cell : ((Pop*)h)->cells)
        for (auto& particle : cell.particles)
            for (int j = 0; j < 3; ++j) particle.x[j] += dt * particle.v[j];
}

void oracle_move_periodic(void* h, double dt, const double* Ld) {
    for (auto& cell : ((Pop*)h)->cells)
        for (auto& particle : cell.particles)
            for (int j = 0; j < 3; ++j) {
                particle.x[j] += dt * particle.v[j];
                particle.x[j] -= Ld[j] * floor(particle.x[j] / Ld[j]);
            }
}

// Population::update (population.h:641-699).  absorbed[s*64+b]: particles of
// species s that left through boundary b.  charge[b]: sum of q over them in
// the reference's loop order (Object::current before /dt).  Returns the
// number of walks that hit the hop cap.
int64_t oracle_update(void* h, int64_t* absorbed, double* charge) {
    Pop& P = *(Pop*)h;
    int64_t lost = 0;
    for (int64_t cell_id = 0; cell_id < P.nc; ++cell_id) {
        std::vector<size_t> to_delete;
        size_t num_particles = P.cells[cell_id].particles.size();
        for (size_t p_id = 0; p_id < num_particles; ++p_id) {
            auto particle = P.cells[cell_id].particles[p_id];
            int64_t new_cell_id = relocate_fast(P, particle.x, cell_id);
            if (new_cell_id != cell_id) {
                to_delete.push_back(p_id);
                if (new_cell_id >= 0) {
                    P.cells[new_cell_id].particles.push_back(particle);
                } else if (new_cell_id == WALK_FAILED) {
                    lost++;
                } else {
                    const int64_t b = -new_cell_id;
                    if (absorbed) absorbed[particle.species * MAX_BND + b] += 1;
                    if (charge) charge[b] += particle.q;
                }
            }
        }
        for (size_t it = to_delete.size(); it-- > 0;) {
            auto p_id = to_delete[it];
            P.cells[cell_id].particles[p_id] = P.cells[cell_id].particles.back();
            P.cells[cell_id].particles.pop_back();
        }
    }
    return lost;
}

// distribute_cg1 (distributor.h:168-210).
void oracle_distribute(void* h, const double* dv_inv, double* rho) {
    Pop& P = *(Pop*)h;
    std::vector<double> rho0(P.ndof, 0.0);
    double cell_coords[4];
    for (auto& cell : P.cells) {
        const int32_t* dof_id = cell.dofs;
        std::vector<double> accum(4, 0.0);
        for (auto& particle : cell.particles) {
            cell.barycentric(particle.x, cell_coords);
            for (int i = 0; i < 4; ++i) accum[i] += particle.q * cell_coords[i];
        }
        for (int i = 0; i < 4; ++i) rho0[dof_id[i]] += accum[i];
    }
    for (int64_t i = 0; i < P.ndof; ++i) rho0[i] *= dv_inv[i];
    memcpy(rho, rho0.data(), P.ndof * 8);
}

// kinetic_energy (diagnostics.h:248-265).
double oracle_kinetic_energy(void* h) {
    double KE = 0.0;
    for (auto& cell : ((Pop*)h)->cells)
        for (auto& particle : cell.particles) {
            auto m = particle.m;
            auto v = particle.v;
            for (int i = 0; i < 3; ++i) KE += 0.5 * m * v[i] * v[i];
        }
    return KE;
}

// Velocity distributions (distributions.cpp:23-245): each constructor's
// normalisation, recomputed per evaluation as a fresh object would hold it.
static double vdf_factor(const Species& s) {
    const int dim = 3;
    const double vth2 = s.vth * s.vth, k = s.k, alpha = s.alpha;
    switch (s.kind) {
        case 1:  // Kappa::Kappa, distributions.cpp:96-97
            return (1.0 / pow(sqrt(M_PI * (2 * k - dim) * vth2), dim)) *
                   (tgamma(k + 1.0) / tgamma(k + ((2.0 - dim) / 2.0)));
        case 2:  // Cairns::Cairns, distributions.cpp:142
            return (1.0 / (pow(sqrt(2 * M_PI * vth2), dim) * (1 + dim * (dim + 2) * alpha)));
        case 3:  // KappaCairns::KappaCairns, distributions.cpp:196-198
            return (1.0 / pow(sqrt(M_PI * (2. * k - dim) * vth2), dim)) *
                   (1.0 / (1.0 + dim * (dim + 2.0) * alpha * ((k - dim / 2.0) / (k - (dim + 2) / 2.0)))) *
                   (tgamma(k + 1.0) / tgamma(k + ((2.0 - dim) / 2.0)));
        default:  // Maxwellian::Maxwellian, distributions.cpp:28
            return (1.0 / (pow(sqrt(2. * M_PI * vth2), 3)));
    }
}

// Pdf::operator()(v) of each distribution.
static double vdf_value(const Species& s, const double* v) {
    const int dim = 3;
    const double vth2 = s.vth * s.vth, vth4 = vth2 * vth2, k = s.k, alpha = s.alpha;
    const double factor = vdf_factor(s);
    double v2 = 0.0;
    for (int i = 0; i < 3; ++i) v2 += (v[i] - s.vd[i]) * (v[i] - s.vd[i]);
    switch (s.kind) {
        case 1:  // distributions.cpp:99-107
            return factor * pow(1.0 + v2 / ((2 * k - dim) * vth2), -(k + 1.0));
        case 2: {  // distributions.cpp:145-154
            double v4 = v2 * v2;
            return factor * (1 + alpha * v4 / vth4) * exp(-0.5 * v2 / vth2);
        }
        case 3: {  // distributions.cpp:201-211
            double v4 = v2 * v2;
            return factor * (1.0 + alpha * v4 / vth4) * pow(1.0 + v2 / ((2 * k - dim) * vth2), -(k + 1.0));
        }
        default:  // distributions.cpp:31-39
            return factor * exp(-0.5 * v2 / vth2);
    }
}

// Flux pdf (v.n)^+ f(v) (distributions.h:93-102).
static double flux_pdf(const Species& s, const double* x, const double* n) {
    double vn = 0.0;
    for (int i = 0; i < 3; ++i) vn += x[i] * n[i];
    return (vn > 0.0) * vn * vdf_value(s, x);
}

// Pdf::max() (distributions.h:244, 278; distributions.cpp:156-173, 213-230).
static double vdf_pdf_max(const Species& s) {
    const int dim = 3;
    const double k = s.k, alpha = s.alpha, factor = vdf_factor(s);
    double v_max[3] = {s.vd[0], s.vd[1], s.vd[2]};
    if (s.kind == 2) {
        if (alpha < 0.25) return factor;
        v_max[0] += s.vth * sqrt(2.0 + sqrt(4.0 - 1.0 / alpha));
        double max = vdf_value(s, v_max);
        max = factor > max ? factor : max;
        return max;
    }
    if (s.kind == 3) {
        double max;
        if (alpha >= (k - 1.) * (k + 1.) / ((2. * k - dim) * (2. * k - dim))) {
            v_max[0] += s.vth * sqrt(((2.0 * k - dim) / (k - 1.0)) +
                                     sqrt(alpha * alpha * (2 * k - dim) * (2 * k - dim) - alpha * (k - 1.0) * (k + 1.0)) /
                                         (alpha * (k - 1.0)));
            max = vdf_value(s, v_max);
        } else {
            max = factor;
        }
        return max;
    }
    return factor;
}

// has_flux_number / has_flux_max as the constructors set them
// (distributions.cpp:23-26, 80-94, 137-140, 190-194).
static bool vdf_flux_number_flag(const Species& s) {
    if (s.kind == 1) {
        int sum_vd = 0;  // std::accumulate(_vd.begin(), _vd.end(), 0): an int accumulator
        for (int i = 0; i < 3; ++i) sum_vd = sum_vd + s.vd[i];
        return sum_vd == 0;
    }
    return s.kind != 3;
}
static bool vdf_flux_max_flag(const Species& s) { return s.kind == 0 || s.kind == 1; }

void oracle_flux(double vth, const double* vd, const double* n, double S, double* num, double* pmax) {
    Species s;
    s.vth = vth;
    for (int i = 0; i < 3; ++i) s.vd[i] = vd[i];
    double vdn = 0.0;
    for (int i = 0; i < 3; ++i) vdn = vdn + n[i] * vd[i];
    *num = S * (vth / (sqrt(2 * M_PI)) * exp(-0.5 * (vdn / vth) * (vdn / vth)) +
                0.5 * vdn * (1. + erf(vdn / (sqrt(2) * vth))));
    double vth2 = vth * vth;
    double tmp[3];
    for (int i = 0; i < 3; ++i) tmp[i] = vd[i] - 0.5 * n[i] * vdn + 0.5 * n[i] * sqrt(4 * vth2 + vdn * vdn);
    *pmax = flux_pdf(s, tmp, n);
}

// Closed-form flux number and flux-pdf maximum of the non-Maxwellian
// distributions (Kappa: distributions.cpp:109-128; Cairns: 175-185).
static double vdf_flux_num_closed(const Species& s, const double* n, double S) {
    const int dim = 3;
    const double _vth = s.vth, k = s.k, alpha = s.alpha;
    if (s.kind == 1)
        return S * ((_vth / (sqrt(2 * M_PI))) * (1.0 / sqrt(k - dim / 2.0)) *
                    (tgamma(k - ((dim - 1.0) / 2.0)) / tgamma(k - dim / 2.0)));
    double vdn = 0.0;
    for (int i = 0; i < 3; ++i) vdn = vdn + n[i] * s.vd[i];
    if (s.kind == 2)
        return S * ((_vth / (sqrt(2 * M_PI))) * exp(-0.5 * (vdn / _vth) * (vdn / _vth)) *
                        (1. + (dim + 1.) * (dim + 3.) * alpha + (vdn / _vth) * (vdn / _vth) * alpha) /
                        (1. + dim * (dim + 2.) * alpha) +
                    0.5 * vdn * (1. + erf(vdn / (sqrt(2) * _vth))));
    double num, pmax;
    oracle_flux(s.vth, s.vd, n, S, &num, &pmax);
    return num;
}
static double vdf_flux_max_closed(const Species& s, const double* n) {
    if (s.kind == 1) {
        const int dim = 3;
        double tmp[3];
        for (int i = 0; i < 3; ++i) tmp[i] = sqrt((2. * s.k - dim) / (2. * s.k + 1.0)) * n[i] * s.vth;
        return flux_pdf(s, tmp, n);
    }
    double num, pmax;
    oracle_flux(s.vth, s.vd, n, 1.0, &num, &pmax);
    return pmax;
}

// create_flux (injector.cpp:104-171) for one distribution over the exterior
// facets, every branch: closed forms where the flags say so, else n_iter
// uniform samples in the domain box per facet (Philox (seed, FLUX), counters
// (rng_sp, facet, sample, 0) for the flux number, (..., 1) for the maximum).
void oracle_vdf_flux(void* h, int kind, double vth, const double* vd, double k, double alpha, double vrange,
                     uint64_t seed, int rng_sp, int64_t n_iter, double* num_out, double* max_out) {
    Pop& P = *(Pop*)h;
    Species s;
    s.kind = kind;
    s.vth = vth;
    for (int i = 0; i < 3; ++i) s.vd[i] = vd[i];
    s.k = k;
    s.alpha = alpha;
    const int dim = 3;
    double domain[6];
    for (int i = 0; i < dim; ++i) {
        domain[i] = vd[i] - vrange * vth;
        domain[i + dim] = vd[i] + vrange * vth;
    }
    double volume = 1.0;
    for (int kk = 0; kk < dim; kk++) volume *= domain[kk + dim] - domain[kk];
    const bool cn = vdf_flux_number_flag(s), cm = vdf_flux_max_flag(s);
    for (size_t j = 0; j < P.ext.size(); ++j) {
        const ExtFacet& f = P.ext[j];
        if (cn) {
            num_out[j] = vdf_flux_num_closed(s, f.normal, f.area);
        } else {
            double sum = 0;
            double x[3];
            for (int64_t n = 0; n < n_iter; ++n) {
                auto r = philox({(uint64_t)rng_sp, (uint64_t)j, (uint64_t)n, 0}, seed, TAG_FLUX);
                for (int kk = 0; kk < dim; ++kk) x[kk] = domain[kk] + unif(r[kk]) * (domain[kk + dim] - domain[kk]);
                sum += flux_pdf(s, x, f.normal);
            }
            num_out[j] = f.area * sum * volume / n_iter;
        }
        if (cm) {
            max_out[j] = vdf_flux_max_closed(s, f.normal);
        } else {
            double max = 0, pdf_x;
            double x[3];
            for (int64_t n = 0; n < n_iter; ++n) {
                auto r = philox({(uint64_t)rng_sp, (uint64_t)j, (uint64_t)n, 1}, seed, TAG_FLUX);
                for (int kk = 0; kk < dim; ++kk) x[kk] = domain[kk] + unif(r[kk]) * (domain[kk + dim] - domain[kk]);
                pdf_x = flux_pdf(s, x, f.normal);
                max = max > pdf_x ? max : pdf_x;
            }
            max_out[j] = max * 1.01;
        }
    }
}

// Pdf value, Pdf::max() and the flags, for the pins.
double oracle_vdf_value(int kind, double vth, const double* vd, double k, double alpha, const double* v) {
    Species s;
    s.kind = kind;
    s.vth = vth;
    for (int i = 0; i < 3; ++i) s.vd[i] = vd[i];
    s.k = k;
    s.alpha = alpha;
    return vdf_value(s, v);
}
double oracle_vdf_max(int kind, double vth, const double* vd, double k, double alpha, int* flags) {
    Species s;
    s.kind = kind;
    s.vth = vth;
    for (int i = 0; i < 3; ++i) s.vd[i] = vd[i];
    s.k = k;
    s.alpha = alpha;
    flags[0] = vdf_flux_number_flag(s) ? 1 : 0;
    flags[1] = vdf_flux_max_flag(s) ? 1 : 0;
    return vdf_pdf_max(s);
}

// Species injection source with given create_flux tables.
void oracle_set_injector_vdf(void* h, int sp, int kind, double vth, const double* vd, double k, double alpha,
                             double n, double vrange, const double* num, const double* pmax) {
    Pop& P = *(Pop*)h;
    Species& s = P.species[sp];
    s.inject = true;
    s.kind = kind;
    s.vth = vth;
    for (int i = 0; i < 3; ++i) s.vd[i] = vd[i];
    s.k = k;
    s.alpha = alpha;
    s.n = n;
    s.vrange = vrange;
    s.num_particles.assign(num, num + P.ext.size());
    s.pdf_max.assign(pmax, pmax + P.ext.size());
}

// Species injection source (create_flux, injector.cpp:104-171, Maxwellian branch).
void oracle_set_injector(void* h, int sp, double vth, const double* vd, double n, double vrange) {
    Pop& P = *(Pop*)h;
    Species& s = P.species[sp];
    s.inject = true;
    s.kind = 0;
    s.vth = vth;
    for (int i = 0; i < 3; ++i) s.vd[i] = vd[i];
    s.n = n;
    s.vrange = vrange;
    s.num_particles.clear();
    s.pdf_max.clear();
    for (auto& f : P.ext) {
        double num, pmax;
        oracle_flux(vth, vd, f.normal, f.area, &num, &pmax);
        s.num_particles.push_back(num);
        s.pdf_max.push_back(pmax);
    }
}

// inject_particles (injector.h:157-232) with Philox streams.  Returns the
// number of particles added.
int64_t oracle_inject(void* h, double dt, uint64_t seed, uint64_t step) {
    Pop& P = *(Pop*)h;
    int64_t added = 0;
    const size_t num_facets = P.ext.size();
    for (size_t i = 0; i < P.species.size(); ++i) {
        Species& sp = P.species[i];
        if (!sp.inject) continue;
        std::vector<double> num = sp.num_particles;
        for (auto& num_i : num) num_i *= sp.n * dt;
        double domain[6];
        for (int d = 0; d < 3; ++d) {
            domain[d] = sp.vd[d] - sp.vrange * sp.vth;
            domain[d + 3] = sp.vd[d] + sp.vrange * sp.vth;
        }
        for (size_t j = 0; j < num_facets; ++j) {
            const ExtFacet& f = P.ext[j];
            double N_float = num[j];
            uint64_t N = static_cast<uint64_t>(N_float);
            auto rc = philox({step, (uint64_t)j, (uint64_t)i, 0}, seed, TAG_INJECT_COUNT);
            if (unif(rc[0]) < (N_float - N)) N += 1;
            const uint64_t sj = ((uint64_t)i << 32) | (uint64_t)j;
            for (uint64_t slot = 0; slot < N; ++slot) {
                auto r0 = philox({step, sj, slot, 0}, seed, TAG_INJECT);
                double xs[3], vs[3];
                double r = unif(r0[0]);
                for (int k = 0; k < 3; ++k) xs[k] = (1.0 - r) * f.vertices[k] + r * f.vertices[k + 3];
                r = sqrt(unif(r0[1]));
                for (int k = 0; k < 3; ++k) xs[k] = r * xs[k] + (1.0 - r) * f.vertices[2 * 3 + k];
                bool ok = false;
                for (uint64_t a = 1; a < 1000000ull && !ok; ++a) {
                    auto ra = philox({step, sj, slot, a}, seed, TAG_INJECT);
                    double tmp[3];
                    for (int k = 0; k < 3; ++k) tmp[k] = domain[k] + (domain[k + 3] - domain[k]) * unif(ra[k]);
                    if (unif(ra[3]) * sp.pdf_max[j] < flux_pdf(sp, tmp, f.normal)) {
                        for (int k = 0; k < 3; ++k) vs[k] = tmp[k];
                        ok = true;
                    }
                }
                if (!ok) continue;
                double rr = unif(r0[2]);
                for (int l = 0; l < 3; ++l) xs[l] = xs[l] + rr * dt * vs[l];
                int64_t c = relocate_fast(P, xs, f.cell);
                if (c >= 0) {
                    Particle p;
                    for (int l = 0; l < 3; ++l) {
                        p.x[l] = xs[l];
                        p.v[l] = vs[l];
                    }
                    p.q = sp.q;
                    p.m = sp.m;
                    p.species = (int32_t)i;
                    P.cells[c].particles.push_back(p);
                    ++added;
                }
            }
        }
    }
    return added;
}

// load_particles (injector.h:244-278) for one species: N particles, positions
// by rejection against inside(x)*(1+amp*cos(kw*x[axis])) in [cmin,cmax]^3,
// velocities vd - sqrt(2)*vth*erfc_inv(2u).  Slot k of species sp uses
// counters (k, attempt, sp, 0) and (k, 2^32-1, sp, 0), key (seed, LOAD).
// Returns the number of particles that could not be placed.
int64_t oracle_load_vdf(void* h, int sp, int64_t N, uint64_t seed, double amp, double kw, int axis, int kind,
                        double vth, const double* vd, double k, double alpha, double vrange);

int64_t oracle_load(void* h, int sp, int64_t N, uint64_t seed, double amp, double kw, int axis,
                    double vth, const double* vd) {
    return oracle_load_vdf(h, sp, N, seed, amp, kw, axis, 0, vth, vd, 0.0, 0.0, 5.0);
}

// load_particles with any velocity distribution: Maxwellian velocities by the
// inverse CDF, the others by the rejection sampler against Pdf::max() in the
// box vd +- vrange*vth (injector.h:259-272), attempts (k, a, sp, 1).
int64_t oracle_load_vdf(void* h, int sp, int64_t N, uint64_t seed, double amp, double kw, int axis, int kind,
                        double vth, const double* vd, double k, double alpha, double vrange) {
    Species vs_pdf;
    vs_pdf.kind = kind;
    vs_pdf.vth = vth;
    for (int i = 0; i < 3; ++i) vs_pdf.vd[i] = vd[i];
    vs_pdf.k = k;
    vs_pdf.alpha = alpha;
    const double vmax = vdf_pdf_max(vs_pdf);
    double vdomain[6];
    for (int i = 0; i < 3; ++i) {
        vdomain[i] = vd[i] - vrange * vth;
        vdomain[i + 3] = vd[i] + vrange * vth;
    }
    Pop& P = *(Pop*)h;
    int64_t failed = 0;
    double domain[6];
    for (int i = 0; i < 3; ++i) {
        domain[i] = P.cmin;
        domain[i + 3] = P.cmax;
    }
    const double pdf_max = 1.0 + amp;
    for (int64_t kk = 0; kk < N; ++kk) {
        double tmp[3];
        int64_t c = -1;
        bool ok = false;
        for (uint64_t a = 0; a < 1000000ull && !ok; ++a) {
            auto r = philox({(uint64_t)kk, a, (uint64_t)sp, 0}, seed, TAG_LOAD);
            for (int i = 0; i < 3; ++i) tmp[i] = domain[i] + (domain[i + 3] - domain[i]) * unif(r[i]);
            c = locate(P, tmp);
            double pdf = (c >= 0) * (1.0 + amp * cos(kw * tmp[axis]));
            if (unif(r[3]) * pdf_max < pdf) ok = true;
        }
        if (!ok) {
            failed++;
            continue;
        }
        Particle p;
        for (int j = 0; j < 3; ++j) p.x[j] = tmp[j];
        if (kind == 0) {
            auto r = philox({(uint64_t)kk, 0xFFFFFFFFull, (uint64_t)sp, 0}, seed, TAG_LOAD);
            for (int j = 0; j < 3; ++j) p.v[j] = vd[j] - sqrt(2.0) * vth * erfc_inv_newton(2 * unif(r[j]));
        } else {
            bool vok = false;
            for (uint64_t a = 0; a < 1000000ull && !vok; ++a) {
                auto r = philox({(uint64_t)kk, a, (uint64_t)sp, 1}, seed, TAG_LOAD);
                double w[3];
                for (int j = 0; j < 3; ++j) w[j] = vdomain[j] + (vdomain[j + 3] - vdomain[j]) * unif(r[j]);
                if (unif(r[3]) * vmax < vdf_value(vs_pdf, w)) {
                    for (int j = 0; j < 3; ++j) p.v[j] = w[j];
                    vok = true;
                }
            }
            if (!vok) {
                failed++;
                continue;
            }
        }
        p.q = P.species[sp].q;
        p.m = P.species[sp].m;
        p.species = sp;
        P.cells[c].particles.push_back(p);
    }
    return failed;
}

// particle_potential_energy_cg1 (diagnostics.h:379-412): PE = sum_p 0.5 q_p
// phi(x_p), phi interpolated with the barycentric coordinates.
double oracle_potential_energy(void* h, const double* phi) {
    Pop& P = *(Pop*)h;
    double phi_x, PE = 0.0;
    double coeffs[4], values[4];
    for (auto& cell : P.cells) {
        for (int i = 0; i < 4; ++i) values[i] = phi[cell.dofs[i]];
        for (auto& particle : cell.particles) {
            cell.barycentric(particle.x, coeffs);
            phi_x = 0.0;
            for (int i = 0; i < 4; ++i) phi_x += coeffs[i] * values[i];
            PE += 0.5 * particle.q * phi_x;
        }
    }
    return PE;
}

// density_cg1 (diagnostics.h:492-555): q < 0 counts as electrons, anything
// else as ions; n_j = dv_inv_j * w * sum_p psi_j(x_p).
void oracle_density_cg1(void* h, const double* dv_inv, double weight_e, double weight_i, double* ne, double* ni) {
    Pop& P = *(Pop*)h;
    std::vector<double> ne0(P.ndof, 0.0), ni0(P.ndof, 0.0);
    double cell_coords[4];
    for (auto& cell : P.cells) {
        const int32_t* dof_id = cell.dofs;
        double accum_e[4] = {0, 0, 0, 0}, accum_i[4] = {0, 0, 0, 0};
        for (auto& particle : cell.particles) {
            cell.barycentric(particle.x, cell_coords);
            if (particle.q < 0) {
                for (int i = 0; i < 4; ++i) accum_e[i] += cell_coords[i];
            } else {
                for (int i = 0; i < 4; ++i) accum_i[i] += cell_coords[i];
            }
        }
        for (int i = 0; i < 4; ++i) {
            ne0[dof_id[i]] += accum_e[i];
            ni0[dof_id[i]] += accum_i[i];
        }
    }
    for (int64_t i = 0; i < P.ndof; ++i) {
        ne0[i] *= dv_inv[i] * weight_e;
        ni0[i] *= dv_inv[i] * weight_i;
    }
    memcpy(ne, ne0.data(), P.ndof * 8);
    memcpy(ni, ni0.data(), P.ndof * 8);
}

// density_dg0 (diagnostics.h:432-465): q > 0 counts as ions, anything else as
// electrons; n_k = (sum of weights in cell k) / volume of cell k.
void oracle_density_dg0(void* h, double weight_e, double weight_i, double* ne, double* ni) {
    Pop& P = *(Pop*)h;
    for (auto& cell : P.cells) {
        double accum_e = 0.0, accum_i = 0.0;
        for (auto& particle : cell.particles) {
            if (particle.q > 0)
                accum_i += weight_i;
            else
                accum_e += weight_e;
        }
        const double vol = cell_volume(cell);
        ne[cell.id] = accum_e / vol;
        ni[cell.id] = accum_i / vol;
    }
}

// Population::statistics (population.h:746-805): running (Welford) mean and
// standard deviation of the speed of the particles with q < 0 (stats[0..1])
// and q > 0 (stats[2..3]); entries of a class with no particles are left as
// they were.
void oracle_statistics(void* h, double* stats) {
    Pop& P = *(Pop*)h;
    std::size_t m = 0, n = 0;
    double v, m_e_old = 0, m_i_old = 0;
    for (auto& cell : P.cells) {
        for (auto& particle : cell.particles) {
            v = 0;
            for (int i = 0; i < 3; ++i) v += particle.v[i] * particle.v[i];
            v = sqrt(v);
            if (particle.q < 0) {
                m++;
                if (m == 1) {
                    m_e_old = v;
                    stats[0] = v;
                    stats[1] = 0.0;
                } else {
                    stats[0] = m_e_old + (v - m_e_old) / m;
                    stats[1] += (v - m_e_old) * (v - stats[0]);
                    m_e_old = stats[0];
                }
            }
            if (particle.q > 0) {
                n++;
                if (n == 1) {
                    m_i_old = v;
                    stats[2] = v;
                    stats[3] = 0.0;
                } else {
                    stats[2] = m_i_old + (v - m_i_old) / n;
                    stats[3] += (v - m_i_old) * (v - stats[2]);
                    m_i_old = stats[2];
                }
            }
        }
    }
    if (m > 0) stats[1] = sqrt(stats[1] / (m - 1));
    if (n > 0) stats[3] = sqrt(stats[3] / (n - 1));
}

// Population::save_file (population.h:807-839): cell by cell, each particle
// as x[3], v[3], q, m - 64 raw bytes, or one text line.
int oracle_save_file(void* h, const char* fname, int binary) {
    Pop& P = *(Pop*)h;
    FILE* fout = fopen(fname, binary ? "wb" : "w");
    if (!fout) return -1;
    for (auto& cell : P.cells) {
        for (auto& particle : cell.particles) {
            if (binary) {
                double rec[8];
                for (int i = 0; i < 3; ++i) {
                    rec[i] = particle.x[i];
                    rec[3 + i] = particle.v[i];
                }
                rec[6] = particle.q;
                rec[7] = particle.m;
                fwrite(rec, sizeof(rec), 1, fout);
            } else {
                for (int i = 0; i < 3; ++i) fprintf(fout, "%.17g\t", particle.x[i]);
                for (int i = 0; i < 3; ++i) fprintf(fout, "%.17g\t", particle.v[i]);
                fprintf(fout, "%.17g\t %.17g\t", particle.q, particle.m);
                fprintf(fout, "\n");
            }
        }
    }
    fclose(fout);
    return 0;
}

// Species of a (q, m) pair: the first with both equal, else a new one.
static int species_of(Pop& P, double q, double m) {
    for (size_t s = 0; s < P.species.size(); ++s)
        if (P.species[s].q == q && P.species[s].m == m) return (int)s;
    Species sp;
    sp.q = q;
    sp.m = m;
    P.species.push_back(sp);
    return (int)P.species.size() - 1;
}

static bool add_one(Pop& P, const double* x, const double* v, double q, double m) {
    int64_t cell_id = locate(P, x);
    if (cell_id < 0) return false;  // outside the mesh: skipped
    Particle p;
    for (int j = 0; j < 3; ++j) {
        p.x[j] = x[j];
        p.v[j] = v[j];
    }
    p.q = q;
    p.m = m;
    p.species = species_of(P, q, m);
    P.cells[cell_id].particles.push_back(p);
    return true;
}

// Population::load_file (population.h:841-880): whole 64-byte records, or
// text lines whose numbers fill x, v, q, m in turn (fields a line lacks keep
// the previous line's values).  Returns the number of particles added, -1 if
// the file cannot be opened.
int64_t oracle_load_file(void* h, const char* fname, int binary) {
    Pop& P = *(Pop*)h;
    int64_t added = 0;
    if (binary) {
        FILE* fin = fopen(fname, "rb");
        if (!fin) return -1;
        double rec[8];
        while (fread(rec, sizeof(rec), 1, fin)) added += add_one(P, rec, rec + 3, rec[6], rec[7]);
        fclose(fin);
        return added;
    }
    std::ifstream in(fname);
    if (!in) return -1;
    std::string line;
    double x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    double q = 0, m = 0;
    while (std::getline(in, line)) {
        double value;
        std::stringstream ss(line);
        size_t i = 0;
        while (ss >> value) {
            if (i < 3) x[i] = value;
            else if (i < 6) v[i % 3] = value;
            else if (i == 6) q = value;
            else if (i == 7) m = value;
            ++i;
        }
        added += add_one(P, x, v, q, m);
    }
    return added;
}

}  // extern "C"
