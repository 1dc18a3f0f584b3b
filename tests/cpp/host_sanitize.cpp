// Host code under AddressSanitizer / UndefinedBehaviorSanitizer: the mesh
// set-up of the product library (puncpp_amd/csrc/mesh_host.cpp: generator,
// localizer, margins, location grid, patch volumes, flux helpers) and the
// oracle's particle path (oracle/oracle.cpp through its C ABI: load, Boris,
// move, update, inject, distribute, diagnostics, file round trip, Monte Carlo
// flux tables), on a small sphere-probe mesh.  Exits non-zero on a failed
// check; the sanitizers abort on any finding.
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "mesh_host.h"

extern "C" {
void* oracle_create(int64_t nv, const double* xyz, int64_t nc, const int32_t* cells, const int32_t* facet_bnd,
                    const int32_t* vdof, int64_t ndof);
void oracle_destroy(void* h);
void oracle_set_species(void* h, int nsp, const double* q, const double* m);
void oracle_element_volume(void* h, int voronoi, double* out);
int64_t oracle_locate(void* h, const double* x);
int64_t oracle_add_particles(void* h, int species, int64_t n, const double* xs, const double* vs);
int64_t oracle_size(void* h);
int64_t oracle_load(void* h, int sp, int64_t N, uint64_t seed, double amp, double kw, int axis, double vth,
                    const double* vd);
void oracle_set_injector(void* h, int sp, double vth, const double* vd, double n, double vrange);
double oracle_boris(void* h, const double* E, const double* B, double dt);
void oracle_move(void* h, double dt);
int64_t oracle_update(void* h, int64_t* absorbed, double* charge);
int64_t oracle_inject(void* h, double dt, uint64_t seed, uint64_t step);
void oracle_distribute(void* h, const double* dv_inv, double* rho);
double oracle_kinetic_energy(void* h);
double oracle_potential_energy(void* h, const double* phi);
void oracle_density_cg1(void* h, const double* dv_inv, double weight_e, double weight_i, double* ne, double* ni);
void oracle_density_dg0(void* h, double weight_e, double weight_i, double* ne, double* ni);
void oracle_statistics(void* h, double* stats);
int oracle_save_file(void* h, const char* fname, int binary);
int64_t oracle_load_file(void* h, const char* fname, int binary);
int64_t oracle_num_exterior(void* h);
void oracle_vdf_flux(void* h, int kind, double vth, const double* vd, double k, double alpha, double vrange,
                     uint64_t seed, int rng_sp, int64_t n_iter, double* num_out, double* max_out);
}

using namespace punc_gpu;

static int bad = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            printf("check failed: %s (line %d)\n", #c, __LINE__); \
            ++bad;                                                 \
        }                                                          \
    } while (0)

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    // mesh set-up of the product library
    HostMesh hm;
    std::string err;
    const double lo[3] = {-3.0, -3.0, -3.0}, hi[3] = {3.0, 3.0, 3.0};
    CHECK(meshgen_box(6, lo, hi, 1.0, 0, hm, err));
    MeshTopo t;
    CHECK(build_topology(hm, t, err));
    CHECK(t.inside_tol > 0.0 && t.near_lam > 0.0);
    std::vector<double> dv, vol;
    element_volume(hm, t, true, dv);
    cell_volumes(t, vol);
    double vsum = 0.0;
    for (double v : vol) vsum += v;
    CHECK(fabs(vsum - t.volume) <= 1e-9 * t.volume);
    const double pts[5][3] = {{0.5, 2.0, -2.5}, {0.0, 0.0, 0.0}, {9.0, 0.0, 0.0}, {NAN, 0.0, 0.0},
                              {-3.0, -3.0, -3.0}};
    CHECK(locate_host(t, pts[0]) >= 0);
    CHECK(locate_host(t, pts[1]) < 0);  // inside the removed sphere
    CHECK(locate_host(t, pts[2]) < 0);
    CHECK(locate_host(t, pts[3]) < 0);
    (void)locate_host(t, pts[4]);
    const double vd[3] = {0.0, 0.0, 0.5}, nrm[3] = {0.0, 0.0, 1.0};
    CHECK(maxwellian_flux_num(1.0, vd, nrm, 2.0) > 0.0);
    VdfParams f;
    CHECK(vdf_setup(3, 1.0, vd, 4.0, 0.3, 15.0, f, err));  // KappaCairns
    (void)vdf_max(f);
    for (int kind = 0; kind < 4; ++kind) {
        VdfParams g;
        if (vdf_setup(kind, 1.0, vd, 4.0, 0.3, 15.0, g, err)) {
            if (vdf_has_flux_num(g)) (void)vdf_flux_num(g, nrm, 1.0);
            if (vdf_has_flux_max(g)) (void)vdf_flux_max(g, nrm);
        }
    }
    // the oracle's particle path on the same mesh
    void* o = oracle_create(hm.nv, hm.xyz.data(), hm.nc, hm.cells.data(), hm.facet_bnd.data(), hm.vdof.data(),
                            hm.ndof);
    const double q[2] = {-1.0, 1.0}, m[2] = {1.0, 1836.0};
    oracle_set_species(o, 2, q, m);
    std::vector<double> dvo(hm.ndof);
    oracle_element_volume(o, 1, dvo.data());
    for (int64_t d = 0; d < hm.ndof; ++d) CHECK(dvo[d] == dv[d]);
    CHECK(oracle_load(o, 0, 3000, 1, 0.0, 0.0, 0, 1.0, vd) == 0);
    CHECK(oracle_load(o, 1, 3000, 1, 0.0, 0.0, 0, 1.0 / sqrt(1836.0), vd) == 0);
    double xs[9] = {0.5, 2.0, -2.5, 0.0, 0.0, 0.0, 9.0, 9.0, 9.0}, vs[9] = {0};
    CHECK(oracle_add_particles(o, 0, 3, xs, vs) == 1);  // two points outside the mesh are skipped
    for (int s = 0; s < 2; ++s) oracle_set_injector(o, s, s ? 1.0 / sqrt(1836.0) : 1.0, vd, 3000.0 / t.volume, 5.0);
    std::vector<double> E(3 * hm.ndof), rho(hm.ndof), phi(hm.ndof, 1.0), ne(hm.ndof), ni(hm.ndof);
    for (int64_t d = 0; d < 3 * hm.ndof; ++d) E[d] = 0.01 * sin(0.37 * d);
    const double B[3] = {0.0, 0.0, 1.0};
    int64_t absorbed[2 * 64];
    double charge[64];
    for (int n = 0; n < 4; ++n) {
        const int64_t before = oracle_size(o);
        oracle_distribute(o, dvo.data(), rho.data());
        (void)oracle_boris(o, E.data(), B, 0.1);
        oracle_move(o, 0.1);
        for (int k = 0; k < 2 * 64; ++k) absorbed[k] = 0;  // oracle_update accumulates into them
        for (int k = 0; k < 64; ++k) charge[k] = 0.0;
        const int64_t lost = oracle_update(o, absorbed, charge);
        CHECK(lost == 0);
        const int64_t added = oracle_inject(o, 0.1, 3, n);
        int64_t ab = 0;
        for (int k = 0; k < 2 * 64; ++k) ab += absorbed[k];
        CHECK(oracle_size(o) == before - ab + added);
    }
    CHECK(oracle_kinetic_energy(o) > 0.0);
    CHECK(isfinite(oracle_potential_energy(o, phi.data())));
    oracle_density_cg1(o, dvo.data(), 1.0, 1.0, ne.data(), ni.data());
    std::vector<double> ne0(hm.nc), ni0(hm.nc);
    oracle_density_dg0(o, 1.0, 1.0, ne0.data(), ni0.data());
    double stats[4];
    oracle_statistics(o, stats);
    const std::string fb = dir + "/pop.bin", ft = dir + "/pop.txt";
    CHECK(oracle_save_file(o, fb.c_str(), 1) == 0);
    CHECK(oracle_save_file(o, ft.c_str(), 0) == 0);
    void* o2 = oracle_create(hm.nv, hm.xyz.data(), hm.nc, hm.cells.data(), hm.facet_bnd.data(), hm.vdof.data(),
                             hm.ndof);
    CHECK(oracle_load_file(o2, fb.c_str(), 1) == oracle_size(o));
    CHECK(oracle_load_file(o2, ft.c_str(), 0) == oracle_size(o));
    const int64_t ne_ext = oracle_num_exterior(o);
    std::vector<double> num(ne_ext), pmax(ne_ext);
    oracle_vdf_flux(o, 3, 1.0, vd, 4.0, 0.3, 15.0, 0, 0, 2000, num.data(), pmax.data());
    for (int64_t j = 0; j < ne_ext; ++j) CHECK(num[j] >= 0.0 && pmax[j] > 0.0);
    oracle_destroy(o2);
    oracle_destroy(o);
    printf(bad ? "host_sanitize: %d failed\n" : "host_sanitize: ok\n", bad);
    return bad ? 1 : 0;
}
