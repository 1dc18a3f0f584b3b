// C ABI of libpunc_gpu.so: handles, device memory, and the operator entry
// points declared in include/punc_gpu.h.
#include <hip/hip_runtime.h>

#include <mutex>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <string>
#include <vector>

#include "../../include/punc_gpu.h"
#include "device.h"
#include "field.h"
#include "kernels.h"
#include "mesh_host.h"

using namespace punc_gpu;

// ---------------------------------------------------------------------------
// Error plumbing
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(const std::string& msg) {
    g_err = msg;
    return -1;
}
static int hipfail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -2;
}
#define HIPCHK(call)                                  \
    do {                                              \
        hipError_t e__ = (call);                      \
        if (e__ != hipSuccess) return hipfail(e__, #call); \
    } while (0)

extern "C" const char* punc_gpu_last_error(void) { return g_err.c_str(); }
extern "C" int punc_gpu_abi_version(void) { return 1; }

// ---------------------------------------------------------------------------
// Handles
// ---------------------------------------------------------------------------
struct punc_hostmesh_s {
    HostMesh m;
};

struct punc_mesh_s {
    int device = 0;
    HostMesh host;
    MeshTopo topo;
    CellRec* d_rec = nullptr;
    CellPlanes* d_planes = nullptr;
    int32_t* d_bstart = nullptr;
    int32_t* d_bcells = nullptr;
    // exterior facets for injection
    double* d_fverts = nullptr;
    double* d_fnormal = nullptr;
    int32_t* d_fcell = nullptr;
    DevMesh dm;
};

// Binding of a species to one instance of a batched mesh.
struct SpInstance {
    bool bound = false;
    int rng_sp = 0;
    uint64_t seed = 0;
    int64_t c0 = 0, c1 = 0, f0 = 0, f1 = 0;
};

struct CoefEntry {
    int kind;
    double dt;
    double B[3];
    SpCoef* d;
};

struct punc_pop_s {
    punc_mesh_s* mesh = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // Side stream of the step: injection generation runs there beside the
    // particle kernel (forked from and joined back to `stream` with events).
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // A re-sort's cell order is pending on the device (see DevState::perm).
    bool perm_pending = false;
    int nsp = 0;
    int max_sp = 0;
    std::vector<double> q, m;
    int cell_bits = 0;
    int64_t capacity = 0;
    int grid = 0, block = 256;
    int num_cu = 1;
    int dep_mode = 3;
    // Fused step with Maxwellian injection: the generator's blocks run in the
    // step kernel's grid (1) or as a kernel on the side stream (0).  The
    // autotuner times both.
    int gen_fused = 0;

    DevState* d_state = nullptr;
    void* bufs[2][4] = {};  // per buffer set: the (x, y), (z, vx), (vy, vz) pair arrays and cs
    int32_t* d_holes = nullptr;
    int32_t* d_slot = nullptr;  // compaction meeting slots (capacity), -1 between calls
    int32_t* d_S = nullptr;     // the re-sort's resulting order (capacity)
    unsigned long long* d_absorbed = nullptr;
    unsigned long long* d_cnt = nullptr;
    unsigned long long* d_errs = nullptr;  // [0] load failures, [1] injection failures
    double* d_E = nullptr;  // field dofs, 3 components per dof (packed as handed over)
    double* d_cellE = nullptr;  // per-cell copy of the field (16 doubles per cell)
    double4* d_walk = nullptr;  // fused step: walk list (32-byte records)
    bool field_set = false;
    double* d_rho_raw = nullptr;
    double* d_rho = nullptr;
    double* d_dv_inv = nullptr;
    bool dv_set = false;
    double* d_ke_part = nullptr;
    double* d_ke = nullptr;
    int ke_np = 0;
    std::vector<CoefEntry> coef_cache;
    // scan / sort scratch
    int64_t* d_scount = nullptr;  // cell re-sort: per (digit, tile) counts and their scan
    int64_t* d_soffset = nullptr;
    int64_t* d_sbsum = nullptr;
    int64_t* d_bsum = nullptr;
    int64_t* d_total = nullptr;  // [0] scan total, [2] injected particles kept
    int64_t bsum_len = 0;
    // captured step
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // Set when a captured step was dropped because a buffer, table or the
    // species set it captured changed (see invalidate_graph).
    bool graph_dropped = false;
    // injection
    std::vector<int> inj_on;
    std::vector<SpInstance> inst;    // per species instance binding
    int32_t* d_pair_sp = nullptr;
    int32_t* d_pair_f = nullptr;
    int64_t npairs = 0;
    std::vector<SpInject> spi;
    std::vector<double> flux_num, flux_max;  // nsp*nf (per unit density and time)
    SpInject* d_spi = nullptr;
    double* d_num = nullptr;
    double* d_pmax = nullptr;
    double inj_dt = -1.0;
    int64_t cand_cap = 0;           // candidate slots: floor(num)+1 per pair
    int64_t cand_alloc = 0;         // slots the candidate buffers hold
    int64_t* d_soff = nullptr;      // npairs: first candidate slot of each pair
    double* d_cx = nullptr;
    int32_t* d_ccs = nullptr;
    int32_t* d_kc = nullptr;        // kept per group of slots
    int32_t* d_spair = nullptr;     // cand_cap: the pair of each candidate slot
    // population digest (allocated on first use): 2 * MAX_SPECIES sums and
    // counts, then MAX_SPECIES int32 cell offsets
    unsigned long long* d_digest = nullptr;
    // diagnostics (allocated on first use)
    double* d_diag = nullptr;      // per-block partials (6 per block) + results
    double* d_phi = nullptr;       // ndof
    double* d_draw = nullptr;      // 2 * max(ndof, nc): raw density sums
    double* d_dens = nullptr;      // 2 * max(ndof, nc): last density (ne, then ni)
    double* d_dens_ema = nullptr;  // 2 * max(ndof, nc)
    double* d_rho_ema = nullptr;   // ndof
    double* d_cvol = nullptr;      // nc
    double4* d_cellphi = nullptr;  // nc: phi at each cell's four dofs
    // fused-step potential energy (phi held in the field rows' slots 12..15)
    bool pe_on = false;
    double* d_pe_part = nullptr;  // ke_np
    double* d_pe = nullptr;       // 1: the last fused step's PE (NaN when off)
    int64_t dens_len = 0;          // entries per species class of the last density (0: none)
    int64_t dens_ema_len = 0;
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
};

template <typename T>
static hipError_t dalloc(T** p, size_t count) {
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

// ---------------------------------------------------------------------------
// Synthetic mesh
// ---------------------------------------------------------------------------
extern "C" int punc_gpu_meshgen_box(int n, const double lo[3], const double hi[3], double sphere_r,
                                    int periodic, punc_hostmesh_t** out) {
    if (!out || !lo || !hi) return fail("meshgen_box: null argument");
    auto* h = new punc_hostmesh_s();
    std::string err;
    if (!meshgen_box(n, lo, hi, sphere_r, periodic, h->m, err)) {
        delete h;
        return fail(err);
    }
    *out = h;
    return 0;
}

extern "C" int punc_gpu_hostmesh_sizes(const punc_hostmesh_t* m, int64_t* nv, int64_t* nc,
                                       int64_t* ndof) {
    if (!m) return fail("hostmesh_sizes: null mesh");
    if (nv) *nv = m->m.nv;
    if (nc) *nc = m->m.nc;
    if (ndof) *ndof = m->m.ndof;
    return 0;
}

extern "C" int punc_gpu_hostmesh_copy(const punc_hostmesh_t* m, double* xyz, int32_t* cells,
                                      int32_t* facet_bnd, int32_t* vdof) {
    if (!m) return fail("hostmesh_copy: null mesh");
    if (xyz) memcpy(xyz, m->m.xyz.data(), m->m.xyz.size() * sizeof(double));
    if (cells) memcpy(cells, m->m.cells.data(), m->m.cells.size() * sizeof(int32_t));
    if (facet_bnd) memcpy(facet_bnd, m->m.facet_bnd.data(), m->m.facet_bnd.size() * sizeof(int32_t));
    if (vdof) memcpy(vdof, m->m.vdof.data(), m->m.vdof.size() * sizeof(int32_t));
    return 0;
}

extern "C" void punc_gpu_hostmesh_free(punc_hostmesh_t* m) { delete m; }

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
static void mesh_free_device(punc_mesh_s* M) {
    hipFree(M->d_rec);
    hipFree(M->d_planes);
    hipFree(M->d_bstart);
    hipFree(M->d_bcells);
    hipFree(M->d_fverts);
    hipFree(M->d_fnormal);
    hipFree(M->d_fcell);
}

extern "C" int punc_gpu_mesh_create(int64_t nv, const double* xyz, int64_t nc, const int32_t* cells,
                                    const int32_t* facet_bnd, const int32_t* vdof, int64_t ndof,
                                    punc_mesh_t** out) {
    if (!out || !xyz || !cells || !facet_bnd) return fail("mesh_create: null argument");
    if (nv <= 0 || nc <= 0) return fail("mesh_create: empty mesh");
    auto* M = new punc_mesh_s();
    M->host.nv = nv;
    M->host.nc = nc;
    M->host.xyz.assign(xyz, xyz + nv * 3);
    M->host.cells.assign(cells, cells + nc * 4);
    M->host.facet_bnd.assign(facet_bnd, facet_bnd + nc * 4);
    if (vdof) {
        M->host.vdof.assign(vdof, vdof + nv);
        M->host.ndof = ndof;
    } else {
        M->host.vdof.resize(nv);
        for (int64_t v = 0; v < nv; ++v) M->host.vdof[v] = (int32_t)v;
        M->host.ndof = nv;
    }
    std::string err;
    if (!build_topology(M->host, M->topo, err)) {
        delete M;
        return fail("mesh_create: " + err);
    }
    *out = M;
    return 0;
}

// Device copies of the mesh are made on first use by a population, so the
// host-side localizer can be built and inspected without a GPU.
static int mesh_ensure_device(punc_mesh_s* M) {
    if (M->d_rec) return 0;
    HIPCHK(hipGetDevice(&M->device));
    const MeshTopo& T = M->topo;
    const int64_t nc = T.nc;
    const size_t nf = T.ext.size();
    std::vector<double> fv(nf * 9 + 1), fn(nf * 3 + 1);
    std::vector<int32_t> fc(nf + 1);
    for (size_t j = 0; j < nf; ++j) {
        memcpy(&fv[j * 9], T.ext[j].vertices, 9 * sizeof(double));
        memcpy(&fn[j * 3], T.ext[j].normal, 3 * sizeof(double));
        fc[j] = T.ext[j].cell;
    }
#define MCHK(call)                     \
    do {                               \
        hipError_t e2 = (call);        \
        if (e2 != hipSuccess) {        \
            mesh_free_device(M);       \
            M->d_rec = nullptr;        \
            return hipfail(e2, #call); \
        }                              \
    } while (0)
    MCHK(dalloc(&M->d_rec, nc));
    MCHK(dalloc(&M->d_planes, nc));
    MCHK(dalloc(&M->d_bstart, T.bucket_start.size()));
    MCHK(dalloc(&M->d_bcells, T.bucket_cells.size()));
    MCHK(dalloc(&M->d_fverts, fv.size()));
    MCHK(dalloc(&M->d_fnormal, fn.size()));
    MCHK(dalloc(&M->d_fcell, fc.size()));
    MCHK(hipMemcpy(M->d_rec, T.rec.data(), nc * sizeof(CellRec), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->d_planes, T.planes.data(), nc * sizeof(CellPlanes), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->d_bstart, T.bucket_start.data(), T.bucket_start.size() * 4, hipMemcpyHostToDevice));
    if (!T.bucket_cells.empty())
        MCHK(hipMemcpy(M->d_bcells, T.bucket_cells.data(), T.bucket_cells.size() * 4, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->d_fverts, fv.data(), fv.size() * 8, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->d_fnormal, fn.data(), fn.size() * 8, hipMemcpyHostToDevice));
    MCHK(hipMemcpy(M->d_fcell, fc.data(), fc.size() * 4, hipMemcpyHostToDevice));
#undef MCHK
    DevMesh& D = M->dm;
    D.rec = M->d_rec;
    D.planes = M->d_planes;
    D.bstart = M->d_bstart;
    D.bcells = M->d_bcells;
    D.nc = nc;
    D.ndof = M->host.ndof;
    D.inside_tol = T.inside_tol;
    D.coord_bound = T.coord_bound;
    D.near_lam = T.near_lam;
    for (int d = 0; d < 3; ++d) {
        D.gn[d] = T.grid_n[d];
        D.glo[d] = T.grid_lo[d];
        D.gh[d] = T.grid_h[d];
    }
    return 0;
}

extern "C" void punc_gpu_mesh_destroy(punc_mesh_t* m) {
    if (!m) return;
    mesh_free_device(m);
    delete m;
}

extern "C" int punc_gpu_mesh_info(const punc_mesh_t* m, int64_t* nv, int64_t* nc, int64_t* ndof,
                                  int64_t* n_exterior, double* volume) {
    if (!m) return fail("mesh_info: null mesh");
    if (nv) *nv = m->topo.nv;
    if (nc) *nc = m->topo.nc;
    if (ndof) *ndof = m->topo.ndof;
    if (n_exterior) *n_exterior = (int64_t)m->topo.ext.size();
    if (volume) *volume = m->topo.volume;
    return 0;
}

extern "C" int punc_gpu_mesh_localizer(const punc_mesh_t* m, int64_t* adj, double* planes,
                                       double* bary, double* vcoords, int32_t* dofs) {
    if (!m) return fail("mesh_localizer: null mesh");
    const MeshTopo& T = m->topo;
    for (int64_t c = 0; c < T.nc; ++c) {
        if (adj)
            for (int i = 0; i < 4; ++i) adj[c * 4 + i] = T.adj[c * 4 + i];
        if (planes) memcpy(planes + c * 16, T.planes[c].p, 16 * sizeof(double));
        if (bary) memcpy(bary + c * 12, T.rec[c].bary, 12 * sizeof(double));
        if (vcoords) memcpy(vcoords + c * 12, &T.vcoords[c * 12], 12 * sizeof(double));
        if (dofs)
            for (int i = 0; i < 4; ++i) dofs[c * 4 + i] = T.rec[c].dof[i];
    }
    return 0;
}

extern "C" int punc_gpu_element_volume(const punc_mesh_t* m, int voronoi, double* dv_inv) {
    if (!m || !dv_inv) return fail("element_volume: null argument");
    std::vector<double> dv;
    element_volume(m->host, m->topo, voronoi != 0, dv);
    memcpy(dv_inv, dv.data(), dv.size() * sizeof(double));
    return 0;
}

extern "C" int punc_gpu_exterior_facets(const punc_mesh_t* m, double* area, double* vertices,
                                        double* normal, double* basis, int32_t* cell) {
    if (!m) return fail("exterior_facets: null mesh");
    const auto& ext = m->topo.ext;
    for (size_t j = 0; j < ext.size(); ++j) {
        if (area) area[j] = ext[j].area;
        if (vertices) memcpy(vertices + 9 * j, ext[j].vertices, 9 * sizeof(double));
        if (normal) memcpy(normal + 3 * j, ext[j].normal, 3 * sizeof(double));
        if (basis) memcpy(basis + 9 * j, ext[j].basis, 9 * sizeof(double));
        if (cell) cell[j] = ext[j].cell;
    }
    return 0;
}

extern "C" int punc_gpu_mesh_locate(const punc_mesh_t* m, int64_t n, const double* xs, int64_t* cells) {
    if (!m || (n > 0 && (!xs || !cells))) return fail("mesh_locate: null argument");
    for (int64_t i = 0; i < n; ++i) cells[i] = locate_host(m->topo, xs + 3 * i);
    return 0;
}

extern "C" int punc_gpu_maxwellian_flux(const punc_mesh_t* m, double vth, const double vd[3],
                                        double* num, double* pdf_max) {
    if (!m || !vd) return fail("maxwellian_flux: null argument");
    const auto& ext = m->topo.ext;
    for (size_t j = 0; j < ext.size(); ++j) {
        if (num) num[j] = maxwellian_flux_num(vth, vd, ext[j].normal, ext[j].area);
        if (pdf_max) pdf_max[j] = maxwellian_flux_max(vth, vd, ext[j].normal);
    }
    return 0;
}

extern "C" int punc_gpu_vdf_info(int kind, double vth, const double vd[3], double kappa, double alpha,
                                 double vdf_range, double* factor, double* pdf_max, int* has_flux_num,
                                 int* has_flux_max) {
    if (!vd) return fail("vdf_info: null argument");
    VdfParams F;
    std::string err;
    if (!vdf_setup(kind, vth, vd, kappa, alpha, vdf_range, F, err)) return fail("vdf_info: " + err);
    if (factor) *factor = F.factor;
    if (pdf_max) *pdf_max = vdf_max(F);
    if (has_flux_num) *has_flux_num = vdf_has_flux_num(F) ? 1 : 0;
    if (has_flux_max) *has_flux_max = vdf_has_flux_max(F) ? 1 : 0;
    return 0;
}

// create_flux for one species over nf facets (inward unit normals, 3 per
// facet, and areas, host arrays): closed forms where the distribution has
// them, the device Monte Carlo otherwise (on `device`, or the current device
// when it is negative).
static int vdf_flux_facets(int device, int64_t nf, const double* normal, const double* area, const VdfParams& F,
                           uint64_t seed, int rng_sp, int64_t n_iter, double* num, double* pdf_max) {
    const bool cn = vdf_has_flux_num(F), cm = vdf_has_flux_max(F);
    for (int64_t j = 0; j < nf; ++j) {
        if (cn) num[j] = vdf_flux_num(F, normal + 3 * j, area[j]);
        if (cm) pdf_max[j] = vdf_flux_max(F, normal + 3 * j);
    }
    if ((cn && cm) || nf == 0) return 0;
    if (n_iter < 1) return fail("vdf_flux: n_iter must be positive");
    if (device >= 0) HIPCHK(hipSetDevice(device));
    double *d_normal = nullptr, *d_area = nullptr, *d_num = nullptr, *d_max = nullptr;
    auto cleanup = [&]() {
        hipFree(d_normal);
        hipFree(d_area);
        hipFree(d_num);
        hipFree(d_max);
    };
    hipError_t e = dalloc(&d_normal, 3 * nf);
    if (e == hipSuccess) e = dalloc(&d_area, nf);
    if (e == hipSuccess) e = dalloc(&d_num, nf);
    if (e == hipSuccess) e = dalloc(&d_max, nf);
    if (e == hipSuccess) e = hipMemcpy(d_normal, normal, 3 * nf * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_area, area, nf * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_flux_mc(d_normal, d_area, nf, F, seed, rng_sp, n_iter, !cn, !cm, d_num, d_max, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess && !cn) e = hipMemcpy(num, d_num, nf * 8, hipMemcpyDeviceToHost);
    if (e == hipSuccess && !cm) e = hipMemcpy(pdf_max, d_max, nf * 8, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return hipfail(e, "vdf_flux");
    return 0;
}

// create_flux for one species over the mesh's exterior facets.
static int vdf_flux_tables(punc_mesh_s* M, const VdfParams& F, uint64_t seed, int rng_sp, int64_t n_iter,
                           double* num, double* pdf_max) {
    const auto& ext = M->topo.ext;
    const int64_t nf = (int64_t)ext.size();
    std::vector<double> normal((size_t)nf * 3), area((size_t)nf);
    for (int64_t j = 0; j < nf; ++j) {
        for (int k = 0; k < 3; ++k) normal[3 * j + k] = ext[j].normal[k];
        area[j] = ext[j].area;
    }
    return vdf_flux_facets(M->device, nf, normal.data(), area.data(), F, seed, rng_sp, n_iter, num, pdf_max);
}

extern "C" int punc_gpu_vdf_flux(punc_mesh_t* m, int kind, double vth, const double vd[3], double kappa,
                                 double alpha, double vdf_range, uint64_t seed, int rng_species, int64_t n_iter,
                                 double* num, double* pdf_max) {
    if (!m || !vd || !num || !pdf_max) return fail("vdf_flux: null argument");
    VdfParams F;
    std::string err;
    if (!vdf_setup(kind, vth, vd, kappa, alpha, vdf_range, F, err)) return fail("vdf_flux: " + err);
    return vdf_flux_tables(m, F, seed, rng_species, n_iter, num, pdf_max);
}

extern "C" int punc_gpu_vdf_flux_facets(int64_t nf, const double* normal, const double* area, int kind, double vth,
                                        const double vd[3], double kappa, double alpha, double vdf_range,
                                        uint64_t seed, int rng_species, int64_t n_iter, double* num,
                                        double* pdf_max) {
    if (nf < 0) return fail("vdf_flux_facets: negative facet count");
    if (nf > 0 && (!normal || !area || !num || !pdf_max)) return fail("vdf_flux_facets: null argument");
    if (!vd) return fail("vdf_flux_facets: null drift");
    VdfParams F;
    std::string err;
    if (!vdf_setup(kind, vth, vd, kappa, alpha, vdf_range, F, err)) return fail("vdf_flux_facets: " + err);
    return vdf_flux_facets(-1, nf, normal, area, F, seed, rng_species, n_iter, num, pdf_max);
}

// ---------------------------------------------------------------------------
// The flux tables without any allocation.  Callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops) stage the
// facets and the two tables through a file-scope device array of fixed size.
// Each GPU holds its own copy of the array, at its own address, from the
// first call on that GPU until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the array's
// name is not a device address in host code).  Every call on one GPU uses the
// same array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_FLUX_MAX_FACETS = 131072;
constexpr int NOALLOC_MAX_DEVICES = 64;
// per facet: 3 normal components, the area, and the two results (6 MiB)
static __device__ __align__(256) double scratch[6 * NOALLOC_FLUX_MAX_FACETS];
static double* g_scratch_dev[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The current device's scratch address, looked up once per device.
static int noalloc_scratch(double** out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail("vdf_flux_facets_noalloc: device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        double* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hipfail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail("vdf_flux_facets_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

extern "C" int punc_gpu_vdf_flux_facets_noalloc(int64_t nf, const double* normal, const double* area, int kind,
                                                double vth, const double vd[3], double kappa, double alpha,
                                                double vdf_range, uint64_t seed, int rng_species, int64_t n_iter,
                                                double* num, double* pdf_max) {
    if (nf < 0) return fail("vdf_flux_facets_noalloc: negative facet count");
    if (nf > NOALLOC_FLUX_MAX_FACETS)
        return fail("vdf_flux_facets_noalloc: " + std::to_string(nf) + " facets exceed the scratch maximum of " +
                    std::to_string(NOALLOC_FLUX_MAX_FACETS) + " (use punc_gpu_vdf_flux_facets)");
    if (nf > 0 && (!normal || !area || !num || !pdf_max)) return fail("vdf_flux_facets_noalloc: null argument");
    if (!vd) return fail("vdf_flux_facets_noalloc: null drift");
    VdfParams F;
    std::string err;
    if (!vdf_setup(kind, vth, vd, kappa, alpha, vdf_range, F, err)) return fail("vdf_flux_facets_noalloc: " + err);
    const bool cn = vdf_has_flux_num(F), cm = vdf_has_flux_max(F);
    for (int64_t j = 0; j < nf; ++j) {
        if (cn) num[j] = vdf_flux_num(F, normal + 3 * j, area[j]);
        if (cm) pdf_max[j] = vdf_flux_max(F, normal + 3 * j);
    }
    if ((cn && cm) || nf == 0) return 0;
    if (n_iter < 1) return fail("vdf_flux_facets_noalloc: n_iter must be positive");
    double* base = nullptr;
    int r = noalloc_scratch(&base);
    if (r) return r;
    double* d_normal = base;
    double* d_area = base + 3 * nf;
    double* d_num = base + 4 * nf;
    double* d_max = base + 5 * nf;
    // every entry the kernel reads is written here; the results are written
    // by the kernel for every facet, so nothing needs clearing
    HIPCHK(hipMemcpy(d_normal, normal, 3 * nf * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_area, area, nf * 8, hipMemcpyHostToDevice));
    HIPCHK(launch_flux_mc(d_normal, d_area, nf, F, seed, rng_species, n_iter, !cn, !cm, d_num, d_max, nullptr));
    HIPCHK(hipStreamSynchronize(nullptr));
    if (!cn) HIPCHK(hipMemcpy(num, d_num, nf * 8, hipMemcpyDeviceToHost));
    if (!cm) HIPCHK(hipMemcpy(pdf_max, d_max, nf * 8, hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// Population
// ---------------------------------------------------------------------------
static void graph_free(punc_pop_s* P);

// A captured step holds raw device pointers (species coefficients, injection
// tables, candidate buffers) and the species count of its capture: whatever
// frees or rebuilds one of them, or changes what the replay should compute,
// drops the graph first, so that a later punc_gpu_graph_launch fails instead
// of reading freed memory or replaying stale parameters.
static void invalidate_graph(punc_pop_s* P) {
    if (P->graph_exec) P->graph_dropped = true;
    graph_free(P);
}

static void pop_free(punc_pop_s* P) {
    graph_free(P);
    hipFree(P->d_digest);
    hipFree(P->d_pair_sp);
    hipFree(P->d_pair_f);
    if (P->d_state) hipFree(P->d_state);
    for (int b = 0; b < 2; ++b)
        for (int k = 0; k < 4; ++k) hipFree(P->bufs[b][k]);
    hipFree(P->d_holes);
    hipFree(P->d_slot);
    hipFree(P->d_S);
    hipFree(P->d_absorbed);
    hipFree(P->d_cnt);
    hipFree(P->d_errs);
    hipFree(P->d_E);
    hipFree(P->d_cellE);
    hipFree(P->d_walk);
    hipFree(P->d_rho_raw);
    hipFree(P->d_rho);
    hipFree(P->d_dv_inv);
    hipFree(P->d_ke_part);
    hipFree(P->d_ke);
    hipFree(P->d_diag);
    hipFree(P->d_phi);
    hipFree(P->d_draw);
    hipFree(P->d_dens);
    hipFree(P->d_dens_ema);
    hipFree(P->d_rho_ema);
    hipFree(P->d_cvol);
    hipFree(P->d_cellphi);
    hipFree(P->d_pe_part);
    hipFree(P->d_pe);
    for (auto& c : P->coef_cache) hipFree(c.d);
    hipFree(P->d_scount);
    hipFree(P->d_soffset);
    hipFree(P->d_sbsum);
    hipFree(P->d_bsum);
    hipFree(P->d_total);
    hipFree(P->d_spi);
    hipFree(P->d_num);
    hipFree(P->d_pmax);
    hipFree(P->d_soff);
    hipFree(P->d_cx);
    hipFree(P->d_ccs);
    hipFree(P->d_kc);
    hipFree(P->d_spair);
    for (auto& e : P->ev) {
        hipEventDestroy(e.first);
        hipEventDestroy(e.second);
    }
    if (P->ev_fork) hipEventDestroy(P->ev_fork);
    if (P->ev_join) hipEventDestroy(P->ev_join);
    if (P->side) hipStreamDestroy(P->side);
    if (P->own_stream && P->stream) hipStreamDestroy(P->stream);
}

extern "C" int punc_gpu_pop_create(punc_mesh_t* m, int nspecies, const double* q, const double* mass,
                                   int64_t capacity, void* stream, punc_pop_t** out) {
    if (!m || !out || (nspecies > 0 && (!q || !mass))) return fail("pop_create: null argument");
    {
        int r0 = mesh_ensure_device(m);
        if (r0) return r0;
    }
    if (nspecies < 0 || nspecies > MAX_SPECIES) return fail("pop_create: nspecies must be in [0, 128]");
    if (capacity < 1 || capacity >= ((int64_t)1 << 31)) return fail("pop_create: capacity must be in [1, 2^31)");
    int cell_bits = 1;
    while (((int64_t)1 << cell_bits) < m->topo.nc) ++cell_bits;
    // The packed cell field keeps 31 - cell_bits bits for the species index.
    const int max_sp = (int)std::min<int64_t>(MAX_SPECIES, (int64_t)1 << std::max(0, 31 - cell_bits));
    if (nspecies > max_sp) return fail("pop_create: too many cells x species for the packed cell field");
    auto* P = new punc_pop_s();
    P->mesh = m;
    P->device = m->device;
    P->nsp = nspecies;
    P->max_sp = max_sp;
    P->q.assign(q, q + nspecies);
    P->m.assign(mass, mass + nspecies);
    P->cell_bits = cell_bits;
    P->capacity = capacity;
#define PCHK(call)                     \
    do {                               \
        hipError_t e2 = (call);        \
        if (e2 != hipSuccess) {        \
            pop_free(P);               \
            delete P;                  \
            return hipfail(e2, #call); \
        }                              \
    } while (0)
    PCHK(hipSetDevice(P->device));
    if (stream) {
        P->stream = (hipStream_t)stream;
    } else {
        PCHK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
        P->own_stream = true;
    }
    PCHK(hipStreamCreateWithFlags(&P->side, hipStreamNonBlocking));
    PCHK(hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming));
    PCHK(hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming));
    hipDeviceProp_t prop;
    PCHK(hipGetDeviceProperties(&prop, P->device));
    P->ke_np = 65536;
    // Default launch: 128 blocks per CU (short chunks, many blocks in
    // flight), deposit mode 3; the autotuner can pick another shape.
    P->grid = std::min(prop.multiProcessorCount * 128, P->ke_np);
    P->num_cu = prop.multiProcessorCount;
    for (int b = 0; b < 2; ++b) {
        for (int k = 0; k < 3; ++k) PCHK(dalloc((double**)&P->bufs[b][k], 2 * capacity));
        PCHK(dalloc((int32_t**)&P->bufs[b][3], capacity));
    }
    PCHK(dalloc(&P->d_state, 1));
    PCHK(dalloc(&P->d_holes, capacity));
    PCHK(dalloc(&P->d_walk, capacity));
    PCHK(dalloc(&P->d_slot, capacity));
    PCHK(dalloc(&P->d_S, capacity));
    PCHK(dalloc(&P->d_absorbed, (size_t)MAX_SPECIES * MAX_BND));
    PCHK(dalloc(&P->d_cnt, MAX_SPECIES));
    PCHK(dalloc(&P->d_errs, 4));
    const int64_t ndof = m->topo.ndof;
    PCHK(dalloc(&P->d_E, 3 * ndof));
    PCHK(dalloc(&P->d_rho_raw, ndof));
    PCHK(dalloc(&P->d_rho, ndof));
    PCHK(dalloc(&P->d_dv_inv, ndof));
    PCHK(dalloc(&P->d_ke_part, P->ke_np));
    PCHK(dalloc(&P->d_pe_part, P->ke_np));
    PCHK(dalloc(&P->d_pe, 1));
    PCHK(dalloc(&P->d_ke, 4));
    const int64_t nc = m->topo.nc;
    PCHK(dalloc(&P->d_cellE, nc * 16));
    PCHK(hipMemset(P->d_cellE, 0, nc * 16 * sizeof(double)));
    const int64_t nsc = sort_count_entries(capacity);
    PCHK(dalloc(&P->d_scount, nsc));
    PCHK(dalloc(&P->d_soffset, nsc));
    PCHK(dalloc(&P->d_sbsum, nsc / 1024 + 2));
    P->bsum_len = std::max<int64_t>(nc / 1024 + 2, 4096);
    PCHK(dalloc(&P->d_bsum, P->bsum_len));
    PCHK(dalloc(&P->d_total, 8));
    DevState s;
    memset(&s, 0, sizeof(s));
    PopArrays* sets[2] = {&s.cur, &s.alt};
    for (int b = 0; b < 2; ++b) {
        sets[b]->xy = (double*)P->bufs[b][0];
        sets[b]->zvx = (double*)P->bufs[b][1];
        sets[b]->vyz = (double*)P->bufs[b][2];
        sets[b]->cs = (int32_t*)P->bufs[b][3];
    }
    s.capacity = capacity;
    PCHK(hipMemcpy(P->d_state, &s, sizeof(s), hipMemcpyHostToDevice));
    PCHK(hipMemset(P->d_absorbed, 0, (size_t)MAX_SPECIES * MAX_BND * 8));
    PCHK(hipMemset(P->d_errs, 0, 4 * 8));
    PCHK(hipMemset(P->d_E, 0, 3 * ndof * sizeof(double)));
    PCHK(hipMemset(P->d_rho_raw, 0, ndof * 8));  // zero outside a deposit (the scale re-zeroes it)
    PCHK(hipMemset(P->d_slot, 0xff, (size_t)capacity * 4));  // compaction slots: -1 between calls
    PCHK(hipMemset(P->d_rho, 0, ndof * 8));
    PCHK(hipMemset(P->d_ke, 0, 4 * 8));
    PCHK(hipMemset(P->d_total, 0, 8 * 8));
    // the fills above run on the null stream, which does not order the
    // population's non-blocking stream behind them: finish them here
    PCHK(hipDeviceSynchronize());
    P->inj_on.assign(nspecies, 0);
    P->inst.assign(nspecies, SpInstance{});
    P->spi.assign(nspecies, SpInject{});
    const size_t nf = m->topo.ext.size();
    P->flux_num.assign((size_t)nspecies * nf, 0.0);
    P->flux_max.assign((size_t)nspecies * nf, 0.0);
#undef PCHK
    *out = P;
    return 0;
}

extern "C" void punc_gpu_pop_destroy(punc_pop_t* p) {
    if (!p) return;
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    pop_free(p);
    delete p;
}

extern "C" int punc_gpu_pop_add_species(punc_pop_t* p, double q, double mass, int* index) {
    if (!p || !index) return fail("pop_add_species: null argument");
    if (!(mass != 0.0)) return fail("pop_add_species: mass must be nonzero");
    for (int s = 0; s < p->nsp; ++s)
        if (p->q[s] == q && p->m[s] == mass) {
            *index = s;
            return 0;
        }
    if (p->nsp >= p->max_sp) return fail("pop_add_species: species limit of this population reached");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamSynchronize(p->stream));
    invalidate_graph(p);
    p->q.push_back(q);
    p->m.push_back(mass);
    p->nsp += 1;
    for (auto& c : p->coef_cache) hipFree(c.d);
    p->coef_cache.clear();
    const size_t nf = p->mesh->topo.ext.size();
    p->inj_on.push_back(0);
    p->inst.push_back(SpInstance{});
    p->spi.push_back(SpInject{});
    p->flux_num.resize((size_t)p->nsp * nf, 0.0);
    p->flux_max.resize((size_t)p->nsp * nf, 0.0);
    p->inj_dt = -1.0;
    *index = p->nsp - 1;
    return 0;
}

extern "C" int punc_gpu_pop_species(punc_pop_t* p, int* nspecies, double* q, double* mass) {
    if (!p || !nspecies) return fail("pop_species: null argument");
    *nspecies = p->nsp;
    for (int s = 0; s < p->nsp; ++s) {
        if (q) q[s] = p->q[s];
        if (mass) mass[s] = p->m[s];
    }
    return 0;
}

static int read_state(punc_pop_s* P, DevState& s) {
    HIPCHK(hipMemcpyAsync(&s, P->d_state, sizeof(DevState), hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return 0;
}

static int check_errors(punc_pop_s* P, const char* what) {
    DevState s;
    int r = read_state(P, s);
    if (r) return r;
    unsigned long long errs[4];
    HIPCHK(hipMemcpy(errs, P->d_errs, sizeof(errs), hipMemcpyDeviceToHost));
    if (s.n_lost > 0)
        return fail(std::string(what) + ": " + std::to_string(s.n_lost) +
                    " particle(s) exceeded the relocation hop cap and were removed");
    if (s.overflow > 0)
        return fail(std::string(what) + ": population capacity exceeded (" +
                    std::to_string(s.overflow) + " particle(s) dropped)");
    if (errs[0] || errs[1])
        return fail(std::string(what) + ": rejection sampler did not converge");
    if (errs[2]) return fail(std::string(what) + ": file records whose (q, m) matched no species");
    return 0;
}

// Operations other than the fused step see the particles in the array
// order, so a pending cell order is applied first.
static int ensure_order(punc_pop_s* P) {
    if (!P->perm_pending) return 0;
    HIPCHK(launch_apply_order(P->d_state, P->capacity, P->stream));
    P->perm_pending = false;
    return 0;
}

static int append_host(punc_pop_s* P, int64_t n, const double* xs, const double* vs,
                       const int32_t* cs_packed) {
    int r0 = ensure_order(P);
    if (r0) return r0;
    DevState s;
    int r = read_state(P, s);
    if (r) return r;
    if (s.n + n > P->capacity) return fail("population capacity exceeded");
    // the pair arrays: value q (0..5: x, y, z, vx, vy, vz) of particle i is
    // element 2i + q % 2 of pair array q / 2
    std::vector<double> col((size_t)(2 * n));
    double* dst[3] = {s.cur.xy, s.cur.zvx, s.cur.vyz};
    for (int a = 0; a < 3; ++a) {
        for (int h = 0; h < 2; ++h) {
            const int q = 2 * a + h;
            const double* src = q < 3 ? xs : vs;
            for (int64_t i = 0; i < n; ++i) col[2 * i + h] = src[i * 3 + (q % 3)];
        }
        HIPCHK(hipMemcpy(dst[a] + 2 * s.n, col.data(), n * 16, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(s.cur.cs + s.n, cs_packed, n * 4, hipMemcpyHostToDevice));
    int64_t nn = s.n + n;
    HIPCHK(hipMemcpy((char*)P->d_state + offsetof(DevState, n), &nn, 8, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int punc_gpu_pop_add(punc_pop_t* p, int species, int64_t n, const double* xs,
                                const double* vs, int64_t* n_added) {
    if (!p || (n > 0 && (!xs || !vs))) return fail("pop_add: null argument");
    if (species < 0 || species >= p->nsp) return fail("pop_add: species out of range");
    HIPCHK(hipSetDevice(p->device));
    std::vector<double> kx, kv;
    std::vector<int32_t> kc;
    kx.reserve(n * 3);
    kv.reserve(n * 3);
    kc.reserve(n);
    const SpInstance& I = p->inst[species];
    const MeshTopo& T = p->mesh->topo;
    const int32_t c0 = I.bound ? (int32_t)I.c0 : 0, c1 = I.bound ? (int32_t)I.c1 : (int32_t)T.nc;
    for (int64_t i = 0; i < n; ++i) {
        const double* x = xs + 3 * i;
        const int64_t c = locate_in_grid(T.planes.data(), T.bucket_start.data(), T.bucket_cells.data(),
                                         T.grid_n, T.grid_lo, T.grid_h, x[0], x[1], x[2], c0, c1);
        if (c < 0) continue;
        kx.insert(kx.end(), xs + 3 * i, xs + 3 * i + 3);
        kv.insert(kv.end(), vs + 3 * i, vs + 3 * i + 3);
        kc.push_back(pack_cs(species, (int32_t)c, p->cell_bits));
    }
    const int64_t k = (int64_t)kc.size();
    if (k > 0) {
        int r = append_host(p, k, kx.data(), kv.data(), kc.data());
        if (r) return r;
    }
    if (n_added) *n_added = k;
    return 0;
}

extern "C" int punc_gpu_pop_add_located(punc_pop_t* p, int64_t n, const double* xs, const double* vs,
                                        const int32_t* species, const int64_t* cells) {
    if (!p || (n > 0 && (!xs || !vs || !species || !cells))) return fail("pop_add_located: null argument");
    HIPCHK(hipSetDevice(p->device));
    std::vector<int32_t> cs((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        if (species[i] < 0 || species[i] >= p->nsp) return fail("pop_add_located: species out of range");
        if (cells[i] < 0 || cells[i] >= p->mesh->topo.nc) return fail("pop_add_located: cell out of range");
        cs[i] = pack_cs(species[i], (int32_t)cells[i], p->cell_bits);
    }
    if (n == 0) return 0;
    return append_host(p, n, xs, vs, cs.data());
}

static int pop_load_common(punc_pop_t* p, int species, int64_t N, uint64_t seed, double amp, double kw, int axis,
                           const VdfParams& F) {
    if (species < 0 || species >= p->nsp) return fail("pop_load: species out of range");
    if (axis < 0 || axis > 2) return fail("pop_load: axis must be 0, 1 or 2");
    if (N < 0) return fail("pop_load: negative count");
    HIPCHK(hipSetDevice(p->device));
    {
        int r0 = ensure_order(p);
        if (r0) return r0;
    }
    DevState s;
    int r = read_state(p, s);
    if (r) return r;
    if (s.n + N > p->capacity) return fail("pop_load: population capacity exceeded");
    const MeshTopo& T = p->mesh->topo;
    const SpInstance& I = p->inst[species];
    double cmin = T.coord_min, cmax = T.coord_max;
    int32_t c0 = 0, c1 = (int32_t)T.nc;
    int rng_sp = species;
    if (I.bound) {
        // The instance's own UniformPosition domain (distributions.h:196-208).
        c0 = (int32_t)I.c0;
        c1 = (int32_t)I.c1;
        cmin = T.vcoords[c0 * 12];
        cmax = cmin;
        for (int64_t k = c0 * 12; k < (int64_t)c1 * 12; ++k) {
            cmin = std::min(cmin, T.vcoords[k]);
            cmax = std::max(cmax, T.vcoords[k]);
        }
        rng_sp = I.rng_sp;
        seed = I.seed;
    }
    HIPCHK(launch_load(p->mesh->dm, p->d_state, s.n, N, species, rng_sp, c0, c1, p->cell_bits, seed, cmin,
                       cmax, amp, kw, axis, F, vdf_max(F), p->d_errs, p->stream));
    int64_t nn = s.n + N;
    HIPCHK(hipMemcpyAsync((char*)p->d_state + offsetof(DevState, n), &nn, 8, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return check_errors(p, "pop_load");
}

extern "C" int punc_gpu_pop_load(punc_pop_t* p, int species, int64_t N, uint64_t seed, double amp,
                                 double kw, int axis, double vth, const double vd[3]) {
    if (!p || !vd) return fail("pop_load: null argument");
    VdfParams F;
    std::string err;
    if (!vdf_setup(VDF_MAXWELLIAN, vth, vd, 0.0, 0.0, 5.0, F, err)) return fail("pop_load: " + err);
    return pop_load_common(p, species, N, seed, amp, kw, axis, F);
}

extern "C" int punc_gpu_pop_load_vdf(punc_pop_t* p, int species, int64_t N, uint64_t seed, double amp, double kw,
                                     int axis, int kind, double vth, const double vd[3], double kappa, double alpha,
                                     double vdf_range) {
    if (!p || !vd) return fail("pop_load_vdf: null argument");
    VdfParams F;
    std::string err;
    if (!vdf_setup(kind, vth, vd, kappa, alpha, vdf_range, F, err)) return fail("pop_load_vdf: " + err);
    return pop_load_common(p, species, N, seed, amp, kw, axis, F);
}

extern "C" int punc_gpu_pop_size(punc_pop_t* p, int64_t* n) {
    if (!p || !n) return fail("pop_size: null argument");
    HIPCHK(hipSetDevice(p->device));
    DevState s;
    int r = read_state(p, s);
    if (r) return r;
    *n = s.n;
    return 0;
}

extern "C" int punc_gpu_pop_counts(punc_pop_t* p, int64_t* per_species) {
    if (!p || !per_species) return fail("pop_counts: null argument");
    HIPCHK(hipSetDevice(p->device));
    {
        int r0 = ensure_order(p);
        if (r0) return r0;
    }
    HIPCHK(launch_fill_u64(p->d_cnt, p->nsp, p->stream));
    HIPCHK(launch_count_species(p->d_state, p->cell_bits, p->nsp, p->d_cnt, p->grid, p->stream));
    HIPCHK(hipMemcpyAsync(per_species, p->d_cnt, p->nsp * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int punc_gpu_pop_get(punc_pop_t* p, double* xs, double* vs, int32_t* species, int64_t* cells) {
    if (!p) return fail("pop_get: null population");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    DevState s;
    r = read_state(p, s);
    if (r) return r;
    const int64_t n = s.n;
    if (n == 0) return 0;
    std::vector<double> col((size_t)(2 * n));
    const double* src[3] = {s.cur.xy, s.cur.zvx, s.cur.vyz};
    for (int a = 0; a < 3; ++a) {
        // pair array a holds values 2a and 2a + 1 of (x, y, z, vx, vy, vz)
        const bool need = a == 0 ? xs != nullptr : a == 2 ? vs != nullptr : (xs != nullptr || vs != nullptr);
        if (!need) continue;
        HIPCHK(hipMemcpy(col.data(), src[a], n * 16, hipMemcpyDeviceToHost));
        for (int h = 0; h < 2; ++h) {
            const int q = 2 * a + h;
            double* dst = q < 3 ? xs : vs;
            if (!dst) continue;
            for (int64_t i = 0; i < n; ++i) dst[i * 3 + (q % 3)] = col[2 * i + h];
        }
    }
    if (species || cells) {
        std::vector<int32_t> cs((size_t)n);
        HIPCHK(hipMemcpy(cs.data(), s.cur.cs, n * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            if (species) species[i] = cs_species(cs[i], p->cell_bits);
            if (cells) cells[i] = cs_cell(cs[i], p->cell_bits);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Operators
// ---------------------------------------------------------------------------
// Per-species coefficients for (kind, dt, B): kind 1 accel, 2 boris, 0 none.
static int get_coef(punc_pop_s* P, int kind, double dt, const double* B, const SpCoef** out) {
    double Bv[3] = {0, 0, 0};
    if (B)
        for (int i = 0; i < 3; ++i) Bv[i] = B[i];
    for (auto& c : P->coef_cache)
        if (c.kind == kind && c.dt == dt && c.B[0] == Bv[0] && c.B[1] == Bv[1] && c.B[2] == Bv[2]) {
            *out = c.d;
            return 0;
        }
    std::vector<SpCoef> h((size_t)P->nsp);
    for (int s = 0; s < P->nsp; ++s) {
        SpCoef& K = h[s];
        memset(&K, 0, sizeof(K));
        const double q = P->q[s], m = P->m[s];
        K.q = q;
        K.m = m;
        K.qm = q / m;
        K.half_m = 0.5 * m;
        if (kind == 1) {
            K.a = dt * (q / m);  // pusher.h:194
        } else if (kind == 2) {
            K.a = 0.5 * dt * (q / m);  // pusher.h:345,367
            double t_mag2 = 0.0;
            for (int i = 0; i < 3; ++i) {
                K.t[i] = tan((dt * q / (2.0 * m)) * Bv[i]);  // pusher.h:334
                t_mag2 += K.t[i] * K.t[i];
            }
            for (int i = 0; i < 3; ++i) K.s[i] = 2 * K.t[i] / (1 + t_mag2);  // pusher.h:340
        }
    }
    if (P->coef_cache.size() >= 8) {
        hipStreamSynchronize(P->stream);
        invalidate_graph(P);
        for (auto& c : P->coef_cache) hipFree(c.d);
        P->coef_cache.clear();
    }
    CoefEntry e;
    e.kind = kind;
    e.dt = dt;
    for (int i = 0; i < 3; ++i) e.B[i] = Bv[i];
    HIPCHK(dalloc(&e.d, P->nsp));
    HIPCHK(hipMemcpy(e.d, h.data(), P->nsp * sizeof(SpCoef), hipMemcpyHostToDevice));
    P->coef_cache.push_back(e);
    *out = e.d;
    return 0;
}

// Removal's grid: its loops stride over the holes and the tail, a few
// thousand to a few hundred thousand per step, so four blocks per compute
// unit cover a step in one or two passes (the particle grid, up to 32768
// blocks, mostly launched blocks that had nothing to do).
static int compact_grid(const punc_pop_s* P) { return std::min(P->grid, 4 * P->num_cu); }

static ParticleLaunch base_launch(punc_pop_s* P) {
    ParticleLaunch L;
    memset(&L, 0, sizeof(L));
    L.mesh = P->mesh->dm;
    L.state = P->d_state;
    L.nsp = P->nsp;
    L.cellE = P->d_cellE;
    L.walk = P->d_walk;
    L.cell_bits = P->cell_bits;
    L.rho_raw = P->d_rho_raw;
    L.ke_part = P->d_ke_part;
    L.holes = P->d_holes;
    L.absorbed = P->d_absorbed;
    L.grid = P->grid;
    L.block = P->block;
    L.dep_mode = P->dep_mode;
    return L;
}

// passes: launch the fused step's two passes instead of one particle kernel.
static hipError_t timed_launch(punc_pop_s* P, const ParticleLaunch& L, bool passes = false) {
    auto go = [&]() { return passes ? launch_step_passes(L, P->stream) : launch_particles(L, P->stream); };
    if (!P->timing) return go();
    if (P->ev_used == P->ev.size()) {
        std::pair<hipEvent_t, hipEvent_t> pr;
        hipError_t e = hipEventCreate(&pr.first);
        if (e != hipSuccess) return e;
        e = hipEventCreate(&pr.second);
        if (e != hipSuccess) return e;
        P->ev.push_back(pr);
    }
    auto& pr = P->ev[P->ev_used++];
    hipError_t e = hipEventRecord(pr.first, P->stream);
    if (e != hipSuccess) return e;
    e = go();
    if (e != hipSuccess) return e;
    return hipEventRecord(pr.second, P->stream);
}

static int push_common(punc_pop_s* P, int kind, double dt, const double* B, double* KE) {
    if (!P->field_set) return fail("push: no electric field set (punc_gpu_set_field)");
    int r = ensure_order(P);
    if (r) return r;
    const SpCoef* coef;
    r = get_coef(P, kind, dt, B, &coef);
    if (r) return r;
    ParticleLaunch L = base_launch(P);
    L.coef = coef;
    L.push = kind;
    HIPCHK(timed_launch(P, L));
    HIPCHK(launch_sum_partials(P->d_ke_part, L.grid, P->d_ke, P->stream));
    if (KE) {
        HIPCHK(hipMemcpyAsync(KE, P->d_ke, 8, hipMemcpyDeviceToHost, P->stream));
        HIPCHK(hipStreamSynchronize(P->stream));
    }
    return 0;
}

// Page-locked host memory for the arrays that cross the host boundary each
// step (field dofs in, charge dofs out): copies from and to it are DMA
// transfers at the link's rate, with no staging through pageable memory.
extern "C" int punc_gpu_host_alloc(int64_t bytes, void** ptr) {
    if (!ptr || bytes < 0) return fail("host_alloc: null pointer or negative size");
    *ptr = nullptr;
    HIPCHK(hipHostMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault));
    return 0;
}

extern "C" int punc_gpu_host_free(void* ptr) {
    if (ptr) HIPCHK(hipHostFree(ptr));
    return 0;
}

extern "C" int punc_gpu_set_field(punc_pop_t* p, const double* E) {
    if (!p || !E) return fail("set_field: null argument");
    HIPCHK(hipSetDevice(p->device));
    const int64_t ndof = p->mesh->topo.ndof;
    // One contiguous copy: a pitched 2D copy from pageable memory runs row by row.
    HIPCHK(hipMemcpyAsync(p->d_E, E, 3 * ndof * sizeof(double), hipMemcpyHostToDevice, p->stream));
    HIPCHK(launch_cell_field(p->mesh->dm, p->d_E, p->d_cellE, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    p->field_set = true;
    return 0;
}

extern "C" int punc_gpu_accel(punc_pop_t* p, double dt, double* KE) {
    if (!p) return fail("accel: null population");
    HIPCHK(hipSetDevice(p->device));
    return push_common(p, 1, dt, nullptr, KE);
}

extern "C" int punc_gpu_boris(punc_pop_t* p, const double B[3], double dt, double* KE) {
    if (!p || !B) return fail("boris: null argument");
    HIPCHK(hipSetDevice(p->device));
    return push_common(p, 2, dt, B, KE);
}

static int move_common(punc_pop_s* P, double dt, const double* Ld) {
    int r = ensure_order(P);
    if (r) return r;
    const SpCoef* coef;
    r = get_coef(P, 0, 0.0, nullptr, &coef);
    if (r) return r;
    ParticleLaunch L = base_launch(P);
    L.coef = coef;
    L.dt = dt;
    L.move = Ld ? 2 : 1;
    if (Ld)
        for (int i = 0; i < 3; ++i) L.Ld[i] = Ld[i];
    HIPCHK(timed_launch(P, L));
    HIPCHK(hipStreamSynchronize(P->stream));
    return 0;
}

extern "C" int punc_gpu_move(punc_pop_t* p, double dt) {
    if (!p) return fail("move: null population");
    HIPCHK(hipSetDevice(p->device));
    return move_common(p, dt, nullptr);
}

extern "C" int punc_gpu_move_periodic(punc_pop_t* p, double dt, const double Ld[3]) {
    if (!p || !Ld) return fail("move_periodic: null argument");
    HIPCHK(hipSetDevice(p->device));
    return move_common(p, dt, Ld);
}

static int read_absorbed(punc_pop_s* P, int64_t* absorbed) {
    if (absorbed)
        HIPCHK(hipMemcpyAsync(absorbed, P->d_absorbed, (size_t)P->nsp * MAX_BND * 8,
                              hipMemcpyDeviceToHost, P->stream));
    HIPCHK(launch_fill_u64(P->d_absorbed, (int64_t)P->nsp * MAX_BND, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return 0;
}

extern "C" int punc_gpu_update(punc_pop_t* p, int64_t* absorbed) {
    if (!p) return fail("update: null population");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    const SpCoef* coef;
    r = get_coef(p, 0, 0.0, nullptr, &coef);
    if (r) return r;
    HIPCHK(launch_fill_u64(p->d_absorbed, (int64_t)p->nsp * MAX_BND, p->stream));
    ParticleLaunch L = base_launch(p);
    L.coef = coef;
    L.locate = true;
    HIPCHK(timed_launch(p, L));
    HIPCHK(launch_compact(p->d_state, p->d_holes, p->d_slot, compact_grid(p), false, p->stream));
    r = read_absorbed(p, absorbed);
    if (r) return r;
    return check_errors(p, "update");
}

static int upload_dv(punc_pop_s* P, const double* dv_inv) {
    if (dv_inv) {
        HIPCHK(hipMemcpyAsync(P->d_dv_inv, dv_inv, P->mesh->topo.ndof * 8, hipMemcpyHostToDevice, P->stream));
        P->dv_set = true;
    }
    if (!P->dv_set) {
        std::vector<double> dv;
        element_volume(P->mesh->host, P->mesh->topo, true, dv);
        HIPCHK(hipMemcpy(P->d_dv_inv, dv.data(), dv.size() * 8, hipMemcpyHostToDevice));
        P->dv_set = true;
    }
    return 0;
}

extern "C" int punc_gpu_distribute(punc_pop_t* p, const double* dv_inv, double* rho) {
    if (!p) return fail("distribute: null population");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    r = upload_dv(p, dv_inv);
    if (r) return r;
    const SpCoef* coef;
    r = get_coef(p, 0, 0.0, nullptr, &coef);
    if (r) return r;
    const int64_t ndof = p->mesh->topo.ndof;
    HIPCHK(launch_fill_f64(p->d_rho_raw, ndof, 0.0, p->stream));
    ParticleLaunch L = base_launch(p);
    L.coef = coef;
    L.deposit = true;
    HIPCHK(timed_launch(p, L));
    HIPCHK(launch_scale_rho(p->d_rho_raw, p->d_dv_inv, p->d_rho, ndof, p->stream));
    if (rho) HIPCHK(hipMemcpyAsync(rho, p->d_rho, ndof * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int punc_gpu_kinetic_energy(punc_pop_t* p, double* KE) {
    if (!p || !KE) return fail("kinetic_energy: null argument");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    const SpCoef* coef;
    r = get_coef(p, 0, 0.0, nullptr, &coef);
    if (r) return r;
    HIPCHK(launch_kinetic(p->d_state, coef, p->cell_bits, p->d_ke_part, p->grid, p->stream));
    HIPCHK(launch_sum_partials(p->d_ke_part, p->grid, p->d_ke + 1, p->stream));
    HIPCHK(hipMemcpyAsync(KE, p->d_ke + 1, 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Field solve (PoissonSolver, poisson.cpp:260-376; ESolver, efield.cpp:41-75)
// ---------------------------------------------------------------------------
struct punc_field_s {
    punc_mesh_s* mesh = nullptr;
    FieldSolver* F = nullptr;
    double* d_in = nullptr;  // host-given rho, phi (ndof) or dJ/dE (3 ndof) + dJ/dphi (ndof) staged on the device
};

extern "C" int punc_gpu_field_create(punc_mesh_t* m, double eps0, int dirichlet_exterior, int n_objects,
                                     int remove_null_space, punc_field_t** out) {
    if (!m || !out) return fail("field_create: null argument");
    HIPCHK(hipSetDevice(m->device));
    std::string err;
    FieldSolver* F = field_create(m->host, m->topo, eps0, dirichlet_exterior != 0, n_objects, remove_null_space != 0,
                                  m->device, nullptr, err);
    if (!F) return fail("field_create: " + err);
    punc_field_s* h = new punc_field_s();
    h->mesh = m;
    h->F = F;
    hipError_t e = dalloc(&h->d_in, 4 * (size_t)m->topo.ndof);
    if (e != hipSuccess) {
        field_destroy(F);
        delete h;
        return hipfail(e, "field_create");
    }
    *out = h;
    return 0;
}

extern "C" void punc_gpu_field_destroy(punc_field_t* f) {
    if (!f) return;
    field_destroy(f->F);
    hipFree(f->d_in);
    delete f;
}

extern "C" int punc_gpu_field_set_max_iterations(punc_field_t* f, int maxit) {
    if (!f || maxit < 1) return fail("field_set_max_iterations: bad argument");
    field_set_max_iterations(f->F, maxit);
    return 0;
}

extern "C" int punc_gpu_field_set_warm_start(punc_field_t* f, int warm) {
    if (!f) return fail("field_set_warm_start: null field");
    field_set_warm_start(f->F, warm != 0);
    return 0;
}

extern "C" int punc_gpu_field_set_preconditioner(punc_field_t* f, int kind) {
    if (!f) return fail("field_set_preconditioner: null field");
    std::string err;
    if (!field_set_preconditioner(f->F, kind, err)) return fail("field_set_preconditioner: " + err);
    return 0;
}

extern "C" int punc_gpu_field_get_preconditioner(punc_field_t* f, int* kind, int* levels, int64_t* coarse) {
    if (!f) return fail("field_get_preconditioner: null field");
    const int k = field_preconditioner(f->F, levels, coarse);
    if (kind) *kind = k;
    return 0;
}

extern "C" int punc_gpu_poisson_solve(punc_field_t* f, const double* rho, const double* exterior_values,
                                      const double* object_potentials, double* phi, int* iterations) {
    if (!f || !rho || !phi) return fail("poisson_solve: null argument");
    const int64_t n = f->mesh->topo.ndof;
    HIPCHK(hipMemcpy(f->d_in, rho, n * 8, hipMemcpyHostToDevice));
    std::string err;
    if (!field_poisson(f->F, f->d_in, exterior_values, object_potentials, iterations, err))
        return fail("poisson_solve: " + err);
    HIPCHK(hipMemcpy(phi, field_phi(f->F), n * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int punc_gpu_object_charges(punc_field_t* f, const double* phi, double* charges) {
    if (!f || !phi || !charges) return fail("object_charges: null argument");
    const int64_t n = f->mesh->topo.ndof;
    HIPCHK(hipMemcpy(f->d_in, phi, n * 8, hipMemcpyHostToDevice));
    std::string err;
    if (!field_object_charges(f->F, f->d_in, charges, err)) return fail("object_charges: " + err);
    return 0;
}

extern "C" int punc_gpu_circuit_set(punc_field_t* f, double dt, int n_vsources, const int* v_a, const int* v_b,
                                    const double* v_value, int n_isources, const int* i_a, const int* i_b,
                                    const double* i_value, double* inv_capacitance) {
    if (!f || n_vsources < 0 || n_isources < 0) return fail("circuit_set: bad argument");
    if ((n_vsources > 0 && (!v_a || !v_b || !v_value)) || (n_isources > 0 && (!i_a || !i_b || !i_value)))
        return fail("circuit_set: null source arrays");
    std::vector<CircuitSource> vs, is;
    for (int k = 0; k < n_vsources; ++k) vs.push_back({v_a[k], v_b[k], v_value[k]});
    for (int k = 0; k < n_isources; ++k) is.push_back({i_a[k], i_b[k], i_value[k]});
    std::string err;
    if (!field_set_circuit(f->F, dt, vs, is, inv_capacitance, err)) return fail("circuit_set: " + err);
    return 0;
}

extern "C" int punc_gpu_poisson_solve_circuit(punc_field_t* f, const double* rho, double* charges,
                                              double* potentials, double* phi, int* iterations) {
    if (!f || !rho || !charges || !potentials) return fail("poisson_solve_circuit: null argument");
    const int64_t n = f->mesh->topo.ndof;
    HIPCHK(hipMemcpy(f->d_in, rho, n * 8, hipMemcpyHostToDevice));
    std::string err;
    if (!field_poisson_circuit(f->F, f->d_in, charges, potentials, iterations, err))
        return fail("poisson_solve_circuit: " + err);
    if (phi) HIPCHK(hipMemcpy(phi, field_phi(f->F), n * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int punc_gpu_efield_solve(punc_field_t* f, const double* phi, double* E, int* iterations) {
    if (!f || !phi || !E) return fail("efield_solve: null argument");
    const int64_t n = f->mesh->topo.ndof;
    HIPCHK(hipMemcpy(f->d_in, phi, n * 8, hipMemcpyHostToDevice));
    std::string err;
    if (!field_efield(f->F, f->d_in, iterations, err)) return fail("efield_solve: " + err);
    HIPCHK(hipMemcpy(E, field_E(f->F), n * 24, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int punc_gpu_efield_adjoint(punc_field_t* f, const double* gE, double* gphi, int* iterations) {
    if (!f || !gE || !gphi) return fail("efield_adjoint: null argument");
    HIPCHK(hipSetDevice(f->mesh->device));
    const int64_t n = f->mesh->topo.ndof;
    HIPCHK(hipMemcpy(f->d_in, gE, n * 24, hipMemcpyHostToDevice));
    std::string err;
    if (!field_efield_adjoint(f->F, f->d_in, f->d_in + 3 * n, iterations, err)) return fail("efield_adjoint: " + err);
    HIPCHK(hipMemcpy(gphi, f->d_in + 3 * n, n * 8, hipMemcpyDeviceToHost));
    return 0;
}

static int field_solve_pop_common(punc_field_t* f, punc_pop_t* p, const double* exterior_values,
                                  const double* object_potentials, double* charges, double* potentials, double* phi,
                                  double* E, int* iterations) {
    if (!f || !p) return fail("field_solve_pop: null argument");
    if (p->mesh != f->mesh) return fail("field_solve_pop: the population lives on another mesh");
    HIPCHK(hipSetDevice(p->device));
    const int64_t n = f->mesh->topo.ndof;
    // the population's deposited charge density (its stream's work first)
    HIPCHK(hipStreamSynchronize(p->stream));
    std::string err;
    int it_phi = 0, it_e = 0;
    const bool ok = charges ? field_poisson_circuit(f->F, p->d_rho, charges, potentials, &it_phi, err)
                            : field_poisson(f->F, p->d_rho, exterior_values, object_potentials, &it_phi, err);
    if (!ok) return fail("field_solve_pop: " + err);
    if (!field_efield(f->F, field_phi(f->F), &it_e, err)) return fail("field_solve_pop: " + err);
    // the field the next push reads: packed dofs and the per-cell copy, on
    // the population's stream (a device-to-device hipMemcpy may return before
    // the copy is done, and the null stream does not order the population's
    // non-blocking stream behind it)
    HIPCHK(hipMemcpyAsync(p->d_E, field_E(f->F), n * 24, hipMemcpyDeviceToDevice, p->stream));
    HIPCHK(launch_cell_field(p->mesh->dm, p->d_E, p->d_cellE, p->stream));
    // the potential too: the next fused step reports the potential energy
    HIPCHK(launch_cell_potential(p->mesh->dm, field_phi(f->F), p->d_cellE, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    p->field_set = true;
    if (!p->pe_on) invalidate_graph(p);
    p->pe_on = true;
    if (phi) HIPCHK(hipMemcpy(phi, field_phi(f->F), n * 8, hipMemcpyDeviceToHost));
    if (E) HIPCHK(hipMemcpy(E, field_E(f->F), n * 24, hipMemcpyDeviceToHost));
    if (iterations) {
        iterations[0] = it_phi;
        iterations[1] = it_e;
    }
    return 0;
}

extern "C" int punc_gpu_field_solve_pop(punc_field_t* f, punc_pop_t* p, const double* exterior_values,
                                        const double* object_potentials, double* phi, double* E, int* iterations) {
    return field_solve_pop_common(f, p, exterior_values, object_potentials, nullptr, nullptr, phi, E, iterations);
}

extern "C" int punc_gpu_field_solve_pop_circuit(punc_field_t* f, punc_pop_t* p, double* charges, double* potentials,
                                                double* phi, double* E, int* iterations) {
    if (!charges || !potentials) return fail("field_solve_pop_circuit: null argument");
    return field_solve_pop_common(f, p, nullptr, nullptr, charges, potentials, phi, E, iterations);
}

// ---------------------------------------------------------------------------
// Population files (Population::save_file / load_file, population.h:807-880)
// ---------------------------------------------------------------------------
constexpr int64_t FILE_CHUNK = int64_t(1) << 22;  // records per device round trip (256 MB)

// One text line per record: "%.17g\t" for x and v, "%.17g\t %.17g\t" for
// q and m, then "\n" (population.h:822-835).  Formatted by up to 16 host
// threads, written in record order.
static bool write_text_records(FILE* f, const double* rec, int64_t cnt) {
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>({16, (int64_t)std::thread::hardware_concurrency(),
                                                                (cnt + 4095) / 4096}));
    std::vector<std::string> out((size_t)nt);
    auto work = [&](int t) {
        const int64_t a = cnt * t / nt, b = cnt * (t + 1) / nt;
        std::string& o = out[t];
        o.reserve((size_t)(b - a) * 200);
        char buf[64];
        for (int64_t k = a; k < b; ++k) {
            const double* r = rec + 8 * k;
            for (int i = 0; i < 6; ++i) {
                const int len = snprintf(buf, sizeof(buf), "%.17g\t", r[i]);
                o.append(buf, (size_t)len);
            }
            const int len = snprintf(buf, sizeof(buf), "%.17g\t %.17g\t", r[6], r[7]);
            o.append(buf, (size_t)len);
            o.push_back('\n');
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    for (auto& o : out)
        if (fwrite(o.data(), 1, o.size(), f) != o.size()) return false;
    return true;
}

extern "C" int punc_gpu_pop_save_file(punc_pop_t* p, const char* fname, int binary) {
    if (!p || !fname) return fail("pop_save_file: null argument");
    HIPCHK(hipSetDevice(p->device));
    // cell order, as the reference writes cell by cell (a re-sort, applied now)
    int r = punc_gpu_sort(p);
    if (r) return r;
    r = ensure_order(p);
    if (r) return r;
    const SpCoef* coef;
    r = get_coef(p, 0, 0.0, nullptr, &coef);
    if (r) return r;
    DevState s;
    r = read_state(p, s);
    if (r) return r;
    FILE* f = fopen(fname, binary ? "wb" : "w");
    if (!f) return fail(std::string("pop_save_file: cannot open ") + fname);
    const int64_t chunk = std::min<int64_t>(FILE_CHUNK, std::max<int64_t>(s.n, 1));
    double* d_rec = nullptr;
    hipError_t e = dalloc(&d_rec, (size_t)chunk * 8);
    std::vector<double> h;
    bool io_ok = true;
    if (e == hipSuccess) h.resize((size_t)chunk * 8);
    for (int64_t first = 0; e == hipSuccess && io_ok && first < s.n; first += chunk) {
        const int64_t cnt = std::min(chunk, s.n - first);
        e = launch_pack_records(p->d_state, coef, p->cell_bits, first, cnt, d_rec, p->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d_rec, cnt * 64, hipMemcpyDeviceToHost, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) break;
        io_ok = binary ? fwrite(h.data(), 64, (size_t)cnt, f) == (size_t)cnt : write_text_records(f, h.data(), cnt);
    }
    hipFree(d_rec);
    io_ok = (fclose(f) == 0) && io_ok;
    if (e != hipSuccess) return hipfail(e, "pop_save_file");
    if (!io_ok) return fail(std::string("pop_save_file: write error on ") + fname);
    return 0;
}

// Text records as the reference reads them (population.h:855-878): the
// numbers of each line fill x, v, q, m in turn; fields a line does not give
// keep the previous line's values (zeros before the first), numbers after
// the eighth are ignored, and every line - an empty one too - is a particle.
static void parse_text_line(const char* s, const char* end, double cur[8]) {
    int i = 0;
    while (s < end) {
        while (s < end && (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\v' || *s == '\f')) ++s;
        if (s >= end) break;
        char* stop = nullptr;
        const double v = strtod(s, &stop);
        if (stop == s) break;  // extraction fails: the rest of the line is ignored
        if (i < 8) cur[i] = v;
        ++i;
        s = stop;
    }
}

extern "C" int punc_gpu_pop_load_file(punc_pop_t* p, const char* fname, int binary, int64_t* n_added) {
    if (!p || !fname) return fail("pop_load_file: null argument");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    FILE* f = fopen(fname, binary ? "rb" : "r");
    if (!f) return fail(std::string("pop_load_file: cannot open ") + fname);
    // A failed load leaves the population as it was: the count, the species
    // list and the overflow counter from here are restored on every error.
    DevState s0;
    r = read_state(p, s0);
    if (r) {
        fclose(f);
        return r;
    }
    const int nsp0 = p->nsp;
    auto rollback = [&]() {
        if (p->nsp > nsp0) {
            invalidate_graph(p);
            const size_t nf = p->mesh->topo.ext.size();
            p->q.resize(nsp0);
            p->m.resize(nsp0);
            p->inj_on.resize(nsp0);
            p->inst.resize(nsp0);
            p->spi.resize(nsp0);
            p->flux_num.resize((size_t)nsp0 * nf);
            p->flux_max.resize((size_t)nsp0 * nf);
            for (auto& c : p->coef_cache) hipFree(c.d);
            p->coef_cache.clear();
            p->inj_dt = -1.0;
            p->nsp = nsp0;
        }
        const int64_t keep[2] = {s0.n, s0.overflow};
        hipStreamSynchronize(p->stream);
        hipMemcpy((char*)p->d_state + offsetof(DevState, n), &keep[0], 8, hipMemcpyHostToDevice);
        hipMemcpy((char*)p->d_state + offsetof(DevState, overflow), &keep[1], 8, hipMemcpyHostToDevice);
    };
    const int64_t chunk = FILE_CHUNK;
    double *d_rec = nullptr, *d_cx = nullptr, *d_sp = nullptr;
    int32_t* d_ccs = nullptr;
    int64_t *d_keep = nullptr, *d_pos = nullptr, *d_bsum = nullptr, *d_tot = nullptr;
    char* lbuf = nullptr;  // getline's line buffer
    size_t lcap = 0;
    auto cleanup = [&]() {
        fclose(f);
        free(lbuf);
        hipFree(d_rec);
        hipFree(d_cx);
        hipFree(d_sp);
        hipFree(d_ccs);
        hipFree(d_keep);
        hipFree(d_pos);
        hipFree(d_bsum);
        hipFree(d_tot);
    };
    hipError_t e = dalloc(&d_rec, (size_t)chunk * 8);
    if (e == hipSuccess) e = dalloc(&d_cx, (size_t)chunk * 6);
    if (e == hipSuccess) e = dalloc(&d_sp, 2 * (size_t)p->max_sp);
    if (e == hipSuccess) e = dalloc(&d_ccs, (size_t)chunk);
    if (e == hipSuccess) e = dalloc(&d_keep, (size_t)chunk);
    if (e == hipSuccess) e = dalloc(&d_pos, (size_t)chunk);
    if (e == hipSuccess) e = dalloc(&d_bsum, (size_t)(chunk / 1024 + 2));
    if (e == hipSuccess) e = dalloc(&d_tot, 1);
    if (e != hipSuccess) {
        cleanup();
        return hipfail(e, "pop_load_file");
    }
    std::vector<double> h((size_t)chunk * 8);
    double cur[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t added = 0;
    bool eof = false;
    while (!eof) {
        int64_t cnt = 0;
        if (binary) {
            cnt = (int64_t)fread(h.data(), 64, (size_t)chunk, f);  // whole records only
            eof = cnt < chunk;
        } else {
            while (cnt < chunk) {
                const ssize_t len = getline(&lbuf, &lcap, f);
                if (len < 0) {  // std::getline: no line left
                    eof = true;
                    break;
                }
                const char* end = lbuf + len;
                if (len > 0 && end[-1] == '\n') --end;
                parse_text_line(lbuf, end, cur);
                memcpy(&h[8 * cnt], cur, 64);
                ++cnt;
            }
        }
        if (cnt == 0) break;
        // register every (q, m) pair of the chunk as a species (first seen, first numbered)
        for (int64_t k = 0; k < cnt; ++k) {
            const double q = h[8 * k + 6], m = h[8 * k + 7];
            bool known = false;
            for (int sp = 0; sp < p->nsp && !known; ++sp) known = p->q[sp] == q && p->m[sp] == m;
            if (!known) {
                int idx;
                r = punc_gpu_pop_add_species(p, q, m, &idx);
                if (r) {
                    cleanup();
                    rollback();
                    return r;
                }
            }
        }
        std::vector<double> sp((size_t)2 * p->nsp);
        for (int j = 0; j < p->nsp; ++j) {
            sp[j] = p->q[j];
            sp[p->nsp + j] = p->m[j];
        }
        e = hipMemcpyAsync(d_sp, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, p->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(d_rec, h.data(), cnt * 64, hipMemcpyHostToDevice, p->stream);
        if (e == hipSuccess)
            e = launch_unpack_append(p->mesh->dm, p->d_state, d_rec, cnt, d_sp, d_sp + p->nsp, p->nsp, p->cell_bits,
                                     d_cx, d_ccs, d_keep, d_pos, d_bsum, d_tot, p->d_errs + 2, p->stream);
        int64_t kept = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&kept, d_tot, 8, hipMemcpyDeviceToHost, p->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
        if (e != hipSuccess) {
            cleanup();
            rollback();
            return hipfail(e, "pop_load_file");
        }
        added += kept;
    }
    cleanup();
    DevState s1;
    r = read_state(p, s1);
    if (r) return r;
    if (s1.overflow > s0.overflow) {
        const int64_t dropped = s1.overflow - s0.overflow;
        rollback();
        if (n_added) *n_added = 0;
        return fail("pop_load_file: population capacity exceeded (" + std::to_string(dropped) +
                    " record(s) did not fit); nothing was loaded");
    }
    if (n_added) *n_added = added;
    return check_errors(p, "pop_load_file");
}

// ---------------------------------------------------------------------------
// Diagnostics (reference diagnostics.h, Population::statistics)
// ---------------------------------------------------------------------------
static int diag_grid(const punc_pop_s* P) { return std::max(1, std::min(P->grid, P->num_cu * 16)); }

static int diag_alloc(punc_pop_s* P) {
    if (P->d_diag) return 0;
    const int64_t ndof = P->mesh->topo.ndof, nc = P->mesh->topo.nc;
    const int64_t len = std::max(ndof, nc);
    HIPCHK(dalloc(&P->d_diag, 6 * (size_t)P->ke_np + 8));
    HIPCHK(dalloc(&P->d_phi, ndof));
    HIPCHK(dalloc(&P->d_draw, 2 * len));
    HIPCHK(dalloc(&P->d_dens, 2 * len));
    HIPCHK(dalloc(&P->d_dens_ema, 2 * len));
    HIPCHK(dalloc(&P->d_rho_ema, ndof));
    std::vector<double> vol;
    cell_volumes(P->mesh->topo, vol);
    HIPCHK(dalloc(&P->d_cvol, nc));
    HIPCHK(dalloc(&P->d_cellphi, nc));
    HIPCHK(hipMemcpy(P->d_cvol, vol.data(), nc * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(P->d_rho_ema, 0, ndof * 8));
    HIPCHK(hipMemset(P->d_dens_ema, 0, 2 * len * 8));
    return 0;
}

static int diag_begin(punc_pop_s* P, const SpCoef** coef) {
    HIPCHK(hipSetDevice(P->device));
    int r = ensure_order(P);
    if (r) return r;
    r = diag_alloc(P);
    if (r) return r;
    return get_coef(P, 0, 0.0, nullptr, coef);
}

extern "C" int punc_gpu_potential_energy(punc_pop_t* p, const double* phi, double* PE) {
    if (!p || !phi || !PE) return fail("potential_energy: null argument");
    const SpCoef* coef;
    int r = diag_begin(p, &coef);
    if (r) return r;
    const int g = p->grid;  // the fused step's launch grid: the same pipeline (at most ke_np partials)
    HIPCHK(hipMemcpyAsync(p->d_phi, phi, p->mesh->topo.ndof * 8, hipMemcpyHostToDevice, p->stream));
    HIPCHK(launch_potential_energy(p->mesh->dm, p->d_state, coef, p->nsp, p->cell_bits, p->d_phi, p->d_cellphi,
                                   p->d_diag, g, p->stream));
    HIPCHK(launch_sum_partials(p->d_diag, g, p->d_diag + 6 * (size_t)p->ke_np, p->stream));
    HIPCHK(hipMemcpyAsync(PE, p->d_diag + 6 * (size_t)p->ke_np, 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

static int density_common(punc_pop_s* P, bool dg0, const double* dv_inv, double we, double wi, double* ne,
                          double* ni) {
    const SpCoef* coef;
    int r = diag_begin(P, &coef);
    if (r) return r;
    const int64_t len = dg0 ? P->mesh->topo.nc : P->mesh->topo.ndof;
    const double* scale = P->d_cvol;
    if (!dg0) {
        r = upload_dv(P, dv_inv);
        if (r) return r;
        scale = P->d_dv_inv;
    }
    HIPCHK(launch_fill_f64(P->d_draw, 2 * len, 0.0, P->stream));
    // CG1 runs at the fused step's launch grid (the same pipeline and deposit);
    // DG0 at fewer, longer chunks: every block flushes its table once, and the
    // flush's global atomics are what that pass costs.
    HIPCHK(launch_density(P->mesh->dm, P->d_state, coef, P->nsp, P->cell_bits, dg0, P->d_draw,
                          dg0 ? diag_grid(P) : P->grid,
                          P->stream));
    HIPCHK(launch_density_scale(dg0, P->d_draw, scale, we, wi, P->d_dens, P->d_dens + len, len, P->stream));
    P->dens_len = len;
    if (ne) HIPCHK(hipMemcpyAsync(ne, P->d_dens, len * 8, hipMemcpyDeviceToHost, P->stream));
    if (ni) HIPCHK(hipMemcpyAsync(ni, P->d_dens + len, len * 8, hipMemcpyDeviceToHost, P->stream));
    HIPCHK(hipStreamSynchronize(P->stream));
    return 0;
}

extern "C" int punc_gpu_density_cg1(punc_pop_t* p, const double* dv_inv, double weight_e, double weight_i,
                                    double* ne, double* ni) {
    if (!p) return fail("density_cg1: null population");
    return density_common(p, false, dv_inv, weight_e, weight_i, ne, ni);
}

extern "C" int punc_gpu_density_dg0(punc_pop_t* p, double weight_e, double weight_i, double* ne, double* ni) {
    if (!p) return fail("density_dg0: null population");
    return density_common(p, true, nullptr, weight_e, weight_i, ne, ni);
}

extern "C" int punc_gpu_statistics(punc_pop_t* p, double* stats) {
    if (!p || !stats) return fail("statistics: null argument");
    const SpCoef* coef;
    int r = diag_begin(p, &coef);
    if (r) return r;
    const int g = diag_grid(p);
    double* out = p->d_diag + 6 * (size_t)p->ke_np;
    HIPCHK(launch_statistics(p->d_state, coef, p->nsp, p->cell_bits, p->d_diag, out, g, p->stream));
    double h[6];
    HIPCHK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    // population.h:802-803: sample standard deviation from M2; a class with
    // no particles leaves its two entries as they were.
    for (int k = 0; k < 2; ++k) {
        if (h[3 * k] > 0) {
            stats[2 * k] = h[3 * k + 1];
            stats[2 * k + 1] = sqrt(h[3 * k + 2] / (h[3 * k] - 1));
        }
    }
    return 0;
}

// ema (diagnostics.h:573-588) on device-resident state: w = 1 - exp(-dt/tau).
extern "C" int punc_gpu_rho_ema(punc_pop_t* p, double dt, double tau, int reset, double* g) {
    if (!p) return fail("rho_ema: null population");
    HIPCHK(hipSetDevice(p->device));
    int r = diag_alloc(p);
    if (r) return r;
    const int64_t ndof = p->mesh->topo.ndof;
    const double w = 1.0 - exp(-dt / tau);
    if (reset) HIPCHK(launch_fill_f64(p->d_rho_ema, ndof, 0.0, p->stream));
    HIPCHK(launch_ema(p->d_rho, p->d_rho_ema, w, ndof, p->stream));
    if (g) HIPCHK(hipMemcpyAsync(g, p->d_rho_ema, ndof * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int punc_gpu_ema(int64_t n, const double* f, double* g, double dt, double tau) {
    if (n < 0) return fail("ema: negative length");
    if (n == 0) return 0;
    if (!f || !g) return fail("ema: null argument");
    const double w = 1.0 - exp(-dt / tau);
    double* d = nullptr;
    HIPCHK(dalloc(&d, 2 * (size_t)n));
    hipError_t e = hipMemcpy(d, f, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + n, g, n * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_ema(d, d + n, w, n, nullptr);
    if (e == hipSuccess) e = hipMemcpy(g, d + n, n * 8, hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess) return hipfail(e, "ema");
    return 0;
}

extern "C" int punc_gpu_density_ema(punc_pop_t* p, double dt, double tau, int reset, double* ne_g,
                                    double* ni_g) {
    if (!p) return fail("density_ema: null population");
    if (p->dens_len == 0) return fail("density_ema: no density computed yet (punc_gpu_density_cg1/dg0)");
    HIPCHK(hipSetDevice(p->device));
    const int64_t len = p->dens_len;
    if (!reset && p->dens_ema_len != len)
        return fail("density_ema: the averaged density has another length (pass reset = 1)");
    const double w = 1.0 - exp(-dt / tau);
    if (reset) HIPCHK(launch_fill_f64(p->d_dens_ema, 2 * len, 0.0, p->stream));
    p->dens_ema_len = len;
    HIPCHK(launch_ema(p->d_dens, p->d_dens_ema, w, 2 * len, p->stream));
    if (ne_g) HIPCHK(hipMemcpyAsync(ne_g, p->d_dens_ema, len * 8, hipMemcpyDeviceToHost, p->stream));
    if (ni_g) HIPCHK(hipMemcpyAsync(ni_g, p->d_dens_ema + len, len * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int punc_gpu_sort(punc_pop_t* p) {
    if (!p) return fail("sort: null population");
    HIPCHK(hipSetDevice(p->device));
    SortLaunch L;
    L.state = p->d_state;
    L.cell_bits = p->cell_bits;
    L.capacity = p->capacity;
    // The (cell, index) pairs of the passes live in the walk-list buffer
    // (32 bytes per slot), which is only used inside a step.  The resulting
    // order is read by the next fused step, which writes the walk list: it
    // goes to the compaction scratch, used only after that step's particle
    // kernel has finished.
    uint32_t* w = (uint32_t*)p->d_walk;
    L.kv0 = (uint64_t*)w;
    L.kv1 = (uint64_t*)(w + 2 * p->capacity);
    L.perm = (uint32_t*)p->d_S;
    L.counts = p->d_scount;
    L.offsets = p->d_soffset;
    L.bsum = p->d_sbsum;
    L.total = p->d_total;
    HIPCHK(launch_sort(L, p->stream));
    p->perm_pending = true;
    return 0;
}

// ---------------------------------------------------------------------------
// Injection
// ---------------------------------------------------------------------------
extern "C" int punc_gpu_injector_set(punc_pop_t* p, int species, double vth, const double vd[3],
                                     double density, double vrange) {
    // Maxwellian: closed-form flux number and maximum (distributions.cpp:52-72)
    return punc_gpu_injector_set_vdf(p, species, VDF_MAXWELLIAN, vth, vd, 0.0, 0.0, density, vrange, nullptr,
                                     nullptr);
}

extern "C" int punc_gpu_injector_set_vdf(punc_pop_t* p, int species, int kind, double vth, const double vd[3],
                                         double kappa, double alpha, double density, double vdf_range,
                                         const double* num, const double* pdf_max) {
    if (!p || !vd) return fail("injector_set_vdf: null argument");
    if (species < 0 || species >= p->nsp) return fail("injector_set_vdf: species out of range");
    VdfParams F;
    std::string err;
    if (!vdf_setup(kind, vth, vd, kappa, alpha, vdf_range, F, err)) return fail("injector_set_vdf: " + err);
    const size_t nf = p->mesh->topo.ext.size();
    std::vector<double> tn(nf), tm(nf);
    if (num && pdf_max) {
        memcpy(tn.data(), num, nf * 8);
        memcpy(tm.data(), pdf_max, nf * 8);
    } else {
        // create_flux with the default stream: seed 0, n_iter 500000
        int r = vdf_flux_tables(p->mesh, F, 0, species, 500000, tn.data(), tm.data());
        if (r) return r;
    }
    SpInject& S = p->spi[species];
    SpInject T;
    memset(&T, 0, sizeof(T));
    T.vth = vth;
    for (int i = 0; i < 3; ++i) T.vd[i] = vd[i];
    T.vth2 = F.vth2;
    T.factor = F.factor;
    T.vrange = vdf_range;
    T.q = p->q[species];
    T.m = p->m[species];
    T.density = density;
    T.kind = kind;
    T.k = F.k;
    T.alpha = F.alpha;
    // Unchanged parameters and tables keep the built pair table (a time loop
    // that re-registers its species every step pays nothing).
    const bool same = p->inj_on[species] && memcmp(&S, &T, offsetof(SpInject, seed)) == 0 && S.kind == T.kind &&
                      S.k == T.k && S.alpha == T.alpha &&
                      memcmp(&p->flux_num[species * nf], tn.data(), nf * 8) == 0 &&
                      memcmp(&p->flux_max[species * nf], tm.data(), nf * 8) == 0;
    if (same) return 0;
    S = T;
    p->inj_on[species] = 1;
    memcpy(&p->flux_num[species * nf], tn.data(), nf * 8);
    memcpy(&p->flux_max[species * nf], tm.data(), nf * 8);
    p->inj_dt = -1.0;  // rebuild the pair table
    invalidate_graph(p);
    return 0;
}

extern "C" int punc_gpu_species_instance(punc_pop_t* p, int species, int rng_species, uint64_t seed,
                                         int64_t cell_begin, int64_t cell_end, int64_t facet_begin,
                                         int64_t facet_end) {
    if (!p) return fail("species_instance: null population");
    if (species < 0 || species >= p->nsp) return fail("species_instance: species out of range");
    const int64_t nf = (int64_t)p->mesh->topo.ext.size(), nc = p->mesh->topo.nc;
    if (facet_begin < 0 || facet_end > nf || facet_begin > facet_end)
        return fail("species_instance: bad exterior facet range");
    if (cell_begin < 0 || cell_end > nc || cell_begin >= cell_end) return fail("species_instance: bad cell range");
    if (rng_species < 0) return fail("species_instance: negative stream species id");
    SpInstance& I = p->inst[species];
    I.bound = true;
    I.rng_sp = rng_species;
    I.seed = seed;
    I.c0 = cell_begin;
    I.c1 = cell_end;
    I.f0 = facet_begin;
    I.f1 = facet_end;
    p->inj_dt = -1.0;
    invalidate_graph(p);
    return 0;
}

// Builds the (species, facet) pair table for dt: species-major, facets
// ascending, one entry per facet of each injecting species.
static int inject_prepare(punc_pop_s* P, double dt) {
    if (P->inj_dt == dt && P->d_num) return 0;
    const int64_t nf = (int64_t)P->mesh->topo.ext.size();
    std::vector<int32_t> psp, pf;
    std::vector<double> num, pmax;
    std::vector<int64_t> soff;  // first candidate slot of each pair
    std::vector<int32_t> spair; // pair of each candidate slot
    int64_t cap = 0;
    for (int s = 0; s < P->nsp; ++s) {
        if (!P->inj_on[s]) continue;
        const SpInstance& I = P->inst[s];
        const int64_t fb = I.bound ? I.f0 : 0, fe = I.bound ? I.f1 : nf;
        SpInject& S = P->spi[s];
        S.use_seed = I.bound ? 1 : 0;
        S.seed = I.seed;
        S.rng_sp = I.bound ? I.rng_sp : s;
        S.facet_base = fb;
        const double nd = P->spi[s].density * dt;
        for (int64_t j = fb; j < fe; ++j) {
            double v = P->flux_num[s * nf + j];
            v *= nd;  // injector.h:174: num_i *= n*dt
            psp.push_back(s);
            pf.push_back((int32_t)j);
            num.push_back(v);
            pmax.push_back(P->flux_max[s * nf + j]);
            soff.push_back(cap);
            const int64_t mx = (int64_t)v + 1;  // the most candidates a step can draw: floor(num) + 1
            spair.insert(spair.end(), (size_t)mx, (int32_t)(psp.size() - 1));
            cap += mx;
        }
    }
    const int64_t np = (int64_t)psp.size();
    cap = std::max<int64_t>(cap, 1);
    HIPCHK(hipStreamSynchronize(P->stream));
    invalidate_graph(P);  // the tables below are freed and rebuilt
    hipFree(P->d_num);
    hipFree(P->d_pmax);
    hipFree(P->d_spi);
    hipFree(P->d_pair_sp);
    hipFree(P->d_pair_f);
    hipFree(P->d_soff);
    P->d_num = P->d_pmax = nullptr;
    P->d_soff = nullptr;
    HIPCHK(dalloc(&P->d_num, np));
    HIPCHK(dalloc(&P->d_pmax, np));
    HIPCHK(dalloc(&P->d_pair_sp, np));
    HIPCHK(dalloc(&P->d_pair_f, np));
    HIPCHK(dalloc(&P->d_soff, np));
    HIPCHK(dalloc(&P->d_spi, P->nsp));
    if (np) {
        HIPCHK(hipMemcpy(P->d_soff, soff.data(), np * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P->d_num, num.data(), np * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P->d_pmax, pmax.data(), np * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P->d_pair_sp, psp.data(), np * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(P->d_pair_f, pf.data(), np * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(P->d_spi, P->spi.data(), P->nsp * sizeof(SpInject), hipMemcpyHostToDevice));
    P->npairs = np;
    // The slot layout is fixed by the tables: cand_cap is exactly the sum,
    // the buffers grow when it does.
    if (cap > P->cand_alloc) {
        hipFree(P->d_cx);
        hipFree(P->d_ccs);
        hipFree(P->d_kc);
        hipFree(P->d_spair);
        P->d_cx = nullptr;
        P->d_ccs = P->d_kc = P->d_spair = nullptr;
        P->cand_alloc = cap;
        HIPCHK(dalloc(&P->d_cx, (size_t)cap * 6));
        HIPCHK(dalloc(&P->d_ccs, cap));
        HIPCHK(dalloc(&P->d_kc, inject_group_count(cap)));
        HIPCHK(dalloc(&P->d_spair, cap));
    }
    if (!spair.empty()) HIPCHK(hipMemcpy(P->d_spair, spair.data(), spair.size() * 4, hipMemcpyHostToDevice));
    P->cand_cap = cap;
    P->inj_dt = dt;
    return 0;
}

// Enqueues one injection (no host synchronization); the step index is read
// from the device state.
// Every species that injects draws from a Maxwellian (the generator that
// fits four wavefronts per SIMD, and can share the fused step's grid).
static bool all_maxwellian(const punc_pop_s* P) {
    for (int sp = 0; sp < P->nsp; ++sp)
        if (P->inj_on[sp] && P->spi[sp].kind != VDF_MAXWELLIAN) return false;
    return true;
}

static int inject_launch_args(punc_pop_s* P, double dt, uint64_t seed, double* rho_raw, InjectLaunch& L) {
    const SpCoef* coef;
    int r = get_coef(P, 0, 0.0, nullptr, &coef);
    if (r) return r;
    memset(&L, 0, sizeof(L));
    L.mesh = P->mesh->dm;
    L.inj.fverts = P->mesh->d_fverts;
    L.inj.fnormal = P->mesh->d_fnormal;
    L.inj.fcell = P->mesh->d_fcell;
    L.inj.pair_sp = P->d_pair_sp;
    L.inj.pair_f = P->d_pair_f;
    L.inj.num = P->d_num;
    L.inj.pmax = P->d_pmax;
    L.inj.npairs = P->npairs;
    L.state = P->d_state;
    L.spi = P->d_spi;
    L.coef = coef;
    L.seed = seed;
    L.dt = dt;
    L.cell_bits = P->cell_bits;
    L.cand_cap = P->cand_cap;
    L.soff = P->d_soff;
    L.spair = P->d_spair;
    L.nkept = P->d_total + 2;
    L.cx = P->d_cx;
    L.ccs = P->d_ccs;
    L.kc = P->d_kc;
    L.rho_raw = rho_raw;
    L.errs = P->d_errs + 1;
    L.all_maxwellian = all_maxwellian(P);
    return 0;
}

static int inject_enqueue(punc_pop_s* P, double dt, uint64_t seed, double* rho_raw) {
    if (P->npairs == 0) return 0;
    InjectLaunch L;
    memset(&L, 0, sizeof(L));
    int r = inject_launch_args(P, dt, seed, rho_raw, L);
    if (r) return r;
    HIPCHK(launch_inject(L, P->stream));
    return 0;
}

extern "C" int punc_gpu_inject(punc_pop_t* p, double dt, uint64_t seed, uint64_t step, int64_t* n_injected) {
    if (!p) return fail("inject: null population");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    r = inject_prepare(p, dt);
    if (r) return r;
    HIPCHK(launch_fill_u64((unsigned long long*)(p->d_total + 1), 2, p->stream));
    HIPCHK(launch_set_step(p->d_state, step, p->stream));
    r = inject_enqueue(p, dt, seed, nullptr);
    if (r) return r;
    HIPCHK(launch_set_step(p->d_state, step, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    r = check_errors(p, "inject");
    if (r) return r;
    if (n_injected) {
        if (p->npairs == 0) {
            *n_injected = 0;
        } else {
            DevState s;
            r = read_state(p, s);
            if (r) return r;
            *n_injected = s.inj_kept;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Fused step and its graph
// ---------------------------------------------------------------------------
struct StepArgs {
    int pusher;
    double push_dt;
    double B[3];
    double dt;
    bool periodic;
    double Ld[3];
    bool inject;
    uint64_t seed;
};

static int step_check(punc_pop_s* P, int pusher, const double* B) {
    if (pusher != 1 && pusher != 2) return fail("step: pusher must be 1 (accel_cg1) or 2 (boris_cg1)");
    if (pusher == 2 && !B) return fail("step: boris needs B");
    if (!P->field_set) return fail("step: no electric field set (punc_gpu_set_field)");
    return 0;
}

// Everything that allocates or copies from the host, done before enqueueing.
static int step_prepare(punc_pop_s* P, const StepArgs& a, const SpCoef** coef) {
    int r = upload_dv(P, nullptr);
    if (r) return r;
    r = get_coef(P, a.pusher, a.push_dt, a.pusher == 2 ? a.B : nullptr, coef);
    if (r) return r;
    const SpCoef* c0;
    r = get_coef(P, 0, 0.0, nullptr, &c0);
    if (r) return r;
    if (a.inject) {
        r = inject_prepare(P, a.dt);
        if (r) return r;
    }
    return 0;
}

// Kernel launches only: safe to capture into a graph.
static int step_enqueue(punc_pop_s* P, const StepArgs& a, const SpCoef* coef, bool advance_step) {
    const int64_t ndof = P->mesh->topo.ndof;
    // Injection generation depends only on the mesh, the injection tables and
    // the step counter: it runs on the side stream beside the particle kernel
    // and is joined back before its particles are appended.
    InjectLaunch IL;
    memset(&IL, 0, sizeof(IL));
    const bool inj = a.inject && P->npairs > 0;
    bool fused_gen = false;
    if (inj) {
        int r = inject_launch_args(P, a.dt, a.seed, P->d_rho_raw, IL);
        if (r) return r;
        fused_gen = IL.all_maxwellian && P->gen_fused;
        if (!fused_gen) HIPCHK(hipEventRecord(P->ev_fork, P->stream));
    }
    // rho_raw is zero here: every deposit ends with a scale that re-zeroes it.
    ParticleLaunch L = base_launch(P);
    L.coef = coef;
    L.push = a.pusher;
    L.dt = a.dt;
    L.move = a.periodic ? 2 : 1;
    for (int i = 0; i < 3; ++i) L.Ld[i] = a.Ld[i];
    L.locate = true;
    L.deposit = true;
    if (P->pe_on) L.pe_part = P->d_pe_part;
    if (fused_gen) L.gen = &IL;
    // The fused kernel is enqueued before the injection branch: enqueued
    // first, the injection's blocks take compute units from the particle
    // kernel and the step is slower (measured).
    HIPCHK(timed_launch(P, L, true));
    // Injection generation depends only on the mesh, the injection tables and
    // the step counter: it runs on the side stream beside the particle kernel
    // and is joined back before its particles are appended.
    if (inj && !fused_gen) {
        HIPCHK(hipStreamWaitEvent(P->side, P->ev_fork, 0));
        HIPCHK(launch_inject_gen(IL, P->side));
        HIPCHK(hipEventRecord(P->ev_join, P->side));
    }
    // Three kernels close the step: removal, append of the injected
    // particles, and scale + bookkeeping (buffer swap, KE/PE sums, count,
    // step counter) - no single-block kernels in between.
    HIPCHK(launch_compact(P->d_state, P->d_holes, P->d_slot, compact_grid(P), true, P->stream));
    if (inj) {
        if (!fused_gen) HIPCHK(hipStreamWaitEvent(P->stream, P->ev_join, 0));
        HIPCHK(launch_inject_append(IL, P->stream, false));
    }
    StepClose C;
    C.ke_part = P->d_ke_part;
    C.pe_part = P->pe_on ? P->d_pe_part : nullptr;
    C.np = L.grid;
    C.ke_out = P->d_ke;
    C.pe_out = P->d_pe;
    C.nkept = inj ? IL.nkept : nullptr;
    C.advance = advance_step ? 1 : 0;
    HIPCHK(launch_scale_finish(P->d_rho_raw, P->d_dv_inv, P->d_rho, ndof, P->d_state, C, P->stream));
    return 0;
}

static StepArgs make_args(int pusher, double push_dt, const double* B, double dt, const double* Ld,
                          int inject, uint64_t seed) {
    StepArgs a;
    memset(&a, 0, sizeof(a));
    a.pusher = pusher;
    a.push_dt = push_dt;
    if (B)
        for (int i = 0; i < 3; ++i) a.B[i] = B[i];
    a.dt = dt;
    a.periodic = Ld != nullptr;
    if (Ld)
        for (int i = 0; i < 3; ++i) a.Ld[i] = Ld[i];
    a.inject = inject != 0;
    a.seed = seed;
    return a;
}

extern "C" int punc_gpu_step(punc_pop_t* p, int pusher, double push_dt, const double B[3], double dt,
                             const double* Ld, int inject, uint64_t seed, uint64_t step) {
    if (!p) return fail("step: null population");
    int r = step_check(p, pusher, B);
    if (r) return r;
    HIPCHK(hipSetDevice(p->device));
    StepArgs a = make_args(pusher, push_dt, B, dt, Ld, inject, seed);
    const SpCoef* coef;
    r = step_prepare(p, a, &coef);
    if (r) return r;
    HIPCHK(launch_set_step(p->d_state, step, p->stream));
    r = step_enqueue(p, a, coef, false);
    if (r) return r;
    p->perm_pending = false;  // the fused step applied it
    return 0;
}

static void graph_free(punc_pop_s* P) {
    if (P->graph_exec) hipGraphExecDestroy(P->graph_exec);
    if (P->graph) hipGraphDestroy(P->graph);
    P->graph_exec = nullptr;
    P->graph = nullptr;
}

extern "C" int punc_gpu_step_graph(punc_pop_t* p, int pusher, double push_dt, const double B[3], double dt,
                                   const double* Ld, int inject, uint64_t seed, uint64_t first_step) {
    if (!p) return fail("step_graph: null population");
    int r = step_check(p, pusher, B);
    if (r) return r;
    HIPCHK(hipSetDevice(p->device));
    r = ensure_order(p);
    if (r) return r;
    StepArgs a = make_args(pusher, push_dt, B, dt, Ld, inject, seed);
    const SpCoef* coef;
    r = step_prepare(p, a, &coef);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(p->stream));
    graph_free(p);
    p->graph_dropped = false;
    HIPCHK(launch_set_step(p->d_state, first_step, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    const bool was_timing = p->timing;
    p->timing = false;  // no event records inside the graph
    HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    r = step_enqueue(p, a, coef, true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->stream, &g);
    p->timing = was_timing;
    if (r) {
        if (g) hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return hipfail(e, "hipStreamEndCapture");
    p->graph = g;
    e = hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        graph_free(p);
        return hipfail(e, "hipGraphInstantiate");
    }
    return 0;
}

extern "C" int punc_gpu_graph_launch(punc_pop_t* p, int count) {
    if (!p) return fail("graph_launch: null population");
    if (!p->graph_exec)
        return fail(p->graph_dropped ? "graph_launch: the captured step was dropped because the population's species, "
                                       "injection tables or coefficients changed since the capture; capture again "
                                       "(punc_gpu_step_graph)"
                                     : "graph_launch: no captured step (punc_gpu_step_graph)");
    HIPCHK(hipSetDevice(p->device));
    for (int k = 0; k < count; ++k) {
        HIPCHK(hipGraphLaunch(p->graph_exec, p->stream));
        p->perm_pending = false;  // the captured fused step applied it
    }
    return 0;
}

extern "C" int punc_gpu_step_results(punc_pop_t* p, double* KE, int64_t* absorbed, int64_t* n_injected) {
    if (!p) return fail("step_results: null population");
    HIPCHK(hipSetDevice(p->device));
    if (KE) HIPCHK(hipMemcpyAsync(KE, p->d_ke, 8, hipMemcpyDeviceToHost, p->stream));
    int r = read_absorbed(p, absorbed);
    if (r) return r;
    r = check_errors(p, "step");
    if (r) return r;
    if (n_injected) {
        DevState s;
        r = read_state(p, s);
        if (r) return r;
        *n_injected = s.inj_kept;
    }
    return 0;
}

extern "C" int punc_gpu_pop_digest(punc_pop_t* p, uint64_t* digest, int64_t* counts) {
    if (!p || !digest) return fail("pop_digest: null argument");
    HIPCHK(hipSetDevice(p->device));
    int r = ensure_order(p);
    if (r) return r;
    if (!p->d_digest) HIPCHK(dalloc(&p->d_digest, 2 * MAX_SPECIES + MAX_SPECIES / 2));
    const int nsp = p->nsp;
    if (nsp == 0) return 0;
    std::vector<int32_t> c0((size_t)nsp);
    for (int s = 0; s < nsp; ++s) c0[s] = p->inst[s].bound ? (int32_t)p->inst[s].c0 : 0;
    int32_t* d_c0 = (int32_t*)(p->d_digest + 2 * MAX_SPECIES);
    HIPCHK(hipMemcpyAsync(d_c0, c0.data(), nsp * 4, hipMemcpyHostToDevice, p->stream));
    HIPCHK(launch_fill_u64(p->d_digest, 2 * nsp, p->stream));
    HIPCHK(launch_pop_digest(p->d_state, p->cell_bits, d_c0, nsp, p->d_digest, p->num_cu * 8, p->stream));
    std::vector<unsigned long long> h((size_t)2 * nsp);
    HIPCHK(hipMemcpyAsync(h.data(), p->d_digest, 2 * nsp * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    for (int s = 0; s < nsp; ++s) {
        digest[s] = h[s];
        if (counts) counts[s] = (int64_t)h[nsp + s];
    }
    return 0;
}

extern "C" int punc_gpu_near_face(punc_pop_t* p, int64_t* n_near) {
    if (!p || !n_near) return fail("near_face: null argument");
    HIPCHK(hipSetDevice(p->device));
    DevState s;
    int r = read_state(p, s);
    if (r) return r;
    *n_near = s.n_near;
    HIPCHK(launch_fill_u64((unsigned long long*)&p->d_state->n_near, 1, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

static int diag_alloc(punc_pop_s* P);

extern "C" int punc_gpu_set_potential(punc_pop_t* p, const double* phi) {
    if (!p) return fail("set_potential: null population");
    HIPCHK(hipSetDevice(p->device));
    int r = diag_alloc(p);
    if (r) return r;
    if (phi) HIPCHK(hipMemcpyAsync(p->d_phi, phi, p->mesh->topo.ndof * 8, hipMemcpyHostToDevice, p->stream));
    HIPCHK(launch_cell_potential(p->mesh->dm, phi ? p->d_phi : nullptr, p->d_cellE, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    if (p->pe_on != (phi != nullptr)) invalidate_graph(p);  // a captured step holds the other kernel
    p->pe_on = phi != nullptr;
    return 0;
}

extern "C" int punc_gpu_step_potential_energy(punc_pop_t* p, double* PE) {
    if (!p || !PE) return fail("step_potential_energy: null argument");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipMemcpyAsync(PE, p->d_pe, 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int punc_gpu_get_rho(punc_pop_t* p, double* rho) {
    if (!p || !rho) return fail("get_rho: null argument");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipMemcpyAsync(rho, p->d_rho, p->mesh->topo.ndof * 8, hipMemcpyDeviceToHost, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

// ---------------------------------------------------------------------------
// Launch control and timing
// ---------------------------------------------------------------------------
extern "C" int punc_gpu_set_launch(punc_pop_t* p, int grid, int block, int deposit_mode) {
    if (!p) return fail("set_launch: null population");
    if (grid < 1 || grid > p->ke_np) return fail("set_launch: grid out of range");
    if (block != 64 && block != 128 && block != 256) return fail("set_launch: block must be 64, 128 or 256");
    if (deposit_mode < 0 || deposit_mode > 4) return fail("set_launch: deposit_mode must be 0 to 4");
    graph_free(p);  // a captured graph holds the old launch shape
    p->grid = grid;
    p->block = block;
    p->dep_mode = deposit_mode;
    return 0;
}

extern "C" int punc_gpu_get_launch(punc_pop_t* p, int* grid, int* block, int* deposit_mode) {
    if (!p) return fail("get_launch: null population");
    if (grid) *grid = p->grid;
    if (block) *block = p->block;
    if (deposit_mode) *deposit_mode = p->dep_mode;
    return 0;
}

extern "C" int punc_gpu_set_inject_in_step(punc_pop_t* p, int on) {
    if (!p) return fail("set_inject_in_step: null population");
    if ((on != 0) != (p->gen_fused != 0)) invalidate_graph(p);  // a captured step holds the other form
    p->gen_fused = on != 0;
    return 0;
}

extern "C" int punc_gpu_get_inject_in_step(punc_pop_t* p, int* on) {
    if (!p || !on) return fail("get_inject_in_step: null argument");
    *on = p->gen_fused;
    return 0;
}

extern "C" int punc_gpu_timing(punc_pop_t* p, int enable) {
    if (!p) return fail("timing: null population");
    p->timing = enable != 0;
    p->ev_used = 0;
    p->timed_ms = 0.0;
    p->timed_launches = 0;
    return 0;
}

extern "C" int punc_gpu_timing_read(punc_pop_t* p, double* ms_total, int64_t* launches) {
    if (!p) return fail("timing_read: null population");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamSynchronize(p->stream));
    for (size_t k = 0; k < p->ev_used; ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[k].first, p->ev[k].second));
        p->timed_ms += ms;
        p->timed_launches += 1;
    }
    p->ev_used = 0;
    if (ms_total) *ms_total = p->timed_ms;
    if (launches) *launches = p->timed_launches;
    return 0;
}

extern "C" int punc_gpu_synchronize(punc_pop_t* p) {
    if (!p) return fail("synchronize: null population");
    HIPCHK(hipSetDevice(p->device));
    HIPCHK(hipStreamSynchronize(p->stream));
    return 0;
}

extern "C" int punc_gpu_copy_probe(int64_t nbytes, int reps, double* ms_per_copy) {
    if (nbytes < 16 || reps < 1 || !ms_per_copy) return fail("copy_probe: bad argument");
    const int64_t n16 = nbytes / 16;
    void *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc(&a, n16 * 16));
    hipError_t e = hipMalloc(&b, n16 * 16);
    if (e != hipSuccess) {
        hipFree(a);
        return hipfail(e, "hipMalloc");
    }
    hipStream_t st;
    hipEvent_t e0, e1;
    hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t prop;
    hipGetDeviceProperties(&prop, dev);
    const int grid = prop.multiProcessorCount * 8;
    hipMemsetAsync(a, 0, n16 * 16, st);
    e = launch_copy16(a, b, n16, grid, st);  // warm-up
    hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy16(a, b, n16, grid, st);
    hipEventRecord(e1, st);
    hipStreamSynchronize(st);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipStreamDestroy(st);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return hipfail(e, "copy kernel");
    *ms_per_copy = ms / reps;
    return 0;
}

// ---------------------------------------------------------------------------
// Autotuner: times the fused step for each launch shape on a snapshot of the
// population and restores the snapshot afterwards, so tuning leaves the
// simulation state untouched.
// ---------------------------------------------------------------------------
extern "C" int punc_gpu_autotune(punc_pop_t* p, int pusher, double push_dt, const double B[3], double dt,
                                 const double* Ld, int inject, uint64_t seed, int reps, int* best_grid,
                                 int* best_block, int* best_mode, double* best_ms) {
    if (!p) return fail("autotune: null population");
    if (reps < 1) return fail("autotune: reps must be >= 1");
    int r = step_check(p, pusher, B);
    if (r) return r;
    HIPCHK(hipSetDevice(p->device));
    r = ensure_order(p);
    if (r) return r;
    HIPCHK(hipSetDevice(p->device));
    StepArgs a = make_args(pusher, push_dt, B, dt, Ld, inject, seed);
    const SpCoef* coef;
    r = step_prepare(p, a, &coef);
    if (r) return r;
    DevState s0;
    r = read_state(p, s0);
    if (r) return r;
    const int64_t n = s0.n;
    const size_t nab = (size_t)MAX_SPECIES * MAX_BND;
    // snapshot: particle arrays, device state, absorbed counters
    std::vector<void*> snap(4, nullptr);
    void* snap_ab = nullptr;
    auto release = [&]() {
        for (void* q : snap) hipFree(q);
        hipFree(snap_ab);
    };
    // bytes per particle of each array: the three pair arrays, then cs
    auto snap_bytes = [&](int k) { return (size_t)std::max<int64_t>(n, 1) * (k < 3 ? 16 : 4); };
    const void* src[4] = {s0.cur.xy, s0.cur.zvx, s0.cur.vyz, s0.cur.cs};
    for (int k = 0; k < 4; ++k) {
        const size_t bytes = snap_bytes(k);
        hipError_t e = hipMalloc(&snap[k], bytes);
        if (e == hipSuccess) e = hipMemcpyAsync(snap[k], src[k], bytes, hipMemcpyDeviceToDevice, p->stream);
        if (e != hipSuccess) {
            release();
            return hipfail(e, "autotune snapshot");
        }
    }
    const int64_t ndof = p->mesh->topo.ndof;
    hipError_t e = hipMalloc(&snap_ab, nab * 8 + ndof * 8 + 8);
    if (e == hipSuccess) e = hipMemcpyAsync(snap_ab, p->d_absorbed, nab * 8, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync((char*)snap_ab + nab * 8, p->d_rho, ndof * 8, hipMemcpyDeviceToDevice, p->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync((char*)snap_ab + nab * 8 + ndof * 8, p->d_ke, 8, hipMemcpyDeviceToDevice, p->stream);
    if (e != hipSuccess) {
        release();
        return hipfail(e, "autotune snapshot");
    }
    hipDeviceProp_t prop;
    hipGetDeviceProperties(&prop, p->device);
    const int cus = prop.multiProcessorCount;
    const int g0 = p->grid, b0 = p->block, m0 = p->dep_mode;
    int bg = g0, bb = b0, bm = m0;
    double bms = 1e300;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    const int grids[] = {4, 8, 16, 32, 64, 128};
    const int blocks[] = {128, 256};
    const int modes[] = {0, 1, 2, 3, 4};
    int rc = 0;
    // restore the snapshot, one untimed step, then reps timed steps
    auto time_steps = [&](double* per) {
        for (int k = 0; k < 4 && !rc; ++k) {
            void* dst[4] = {s0.cur.xy, s0.cur.zvx, s0.cur.vyz, s0.cur.cs};
            if (hipMemcpyAsync(dst[k], snap[k], snap_bytes(k), hipMemcpyDeviceToDevice, p->stream) != hipSuccess)
                rc = fail("autotune: restore failed");
        }
        if (!rc && hipMemcpyAsync(p->d_state, &s0, sizeof(DevState), hipMemcpyHostToDevice, p->stream) != hipSuccess)
            rc = fail("autotune: restore failed");
        if (!rc) rc = step_enqueue(p, a, coef, true);
        if (!rc && hipEventRecord(e0, p->stream) != hipSuccess) rc = fail("autotune: event");
        for (int it = 0; it < reps && !rc; ++it) rc = step_enqueue(p, a, coef, true);
        if (!rc && hipEventRecord(e1, p->stream) != hipSuccess) rc = fail("autotune: event");
        if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = fail("autotune: step failed");
        float ms = 0.f;
        if (!rc) hipEventElapsedTime(&ms, e0, e1);
        *per = ms / reps;
    };
    const int gf0 = p->gen_fused;
    p->gen_fused = 0;
    for (int gm : grids) {
        for (int bl : blocks) {
            for (int md : modes) {
                const int grid = std::min(cus * gm, p->ke_np);
                p->grid = grid;
                p->block = bl;
                p->dep_mode = md;
                double per = 0.0;
                time_steps(&per);
                if (rc) break;
                if (per < bms) {
                    bms = per;
                    bg = grid;
                    bb = bl;
                    bm = md;
                }
            }
            if (rc) break;
        }
        if (rc) break;
    }
    // Where the Maxwellian injection generator runs (a second stream or the
    // step kernel's own grid): timed at the best launch shape.
    int bgf = 0;
    if (!rc && inject && p->npairs > 0 && all_maxwellian(p)) {
        p->grid = bg;
        p->block = bb;
        p->dep_mode = bm;
        p->gen_fused = 1;
        double per = 0.0;
        time_steps(&per);
        if (!rc && per < bms) {
            bms = per;
            bgf = 1;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    // restore the state the caller handed in
    void* dst[4] = {s0.cur.xy, s0.cur.zvx, s0.cur.vyz, s0.cur.cs};
    for (int k = 0; k < 4; ++k) hipMemcpyAsync(dst[k], snap[k], snap_bytes(k), hipMemcpyDeviceToDevice, p->stream);
    hipMemcpyAsync(p->d_state, &s0, sizeof(DevState), hipMemcpyHostToDevice, p->stream);
    hipMemcpyAsync(p->d_absorbed, snap_ab, nab * 8, hipMemcpyDeviceToDevice, p->stream);
    hipMemcpyAsync(p->d_rho, (char*)snap_ab + nab * 8, ndof * 8, hipMemcpyDeviceToDevice, p->stream);
    hipMemcpyAsync(p->d_ke, (char*)snap_ab + nab * 8 + ndof * 8, 8, hipMemcpyDeviceToDevice, p->stream);
    hipStreamSynchronize(p->stream);
    release();
    graph_free(p);
    if (rc) {
        p->grid = g0;
        p->block = b0;
        p->dep_mode = m0;
        p->gen_fused = gf0;
        return rc;
    }
    p->grid = bg;
    p->block = bb;
    p->dep_mode = bm;
    p->gen_fused = bgf;
    if (best_grid) *best_grid = bg;
    if (best_block) *best_block = bb;
    if (best_mode) *best_mode = bm;
    if (best_ms) *best_ms = bms;
    return 0;
}

// Device view of a mesh for the other translation units of the library.
int punc_internal_mesh_view(punc_mesh_t* m, DevMesh* out, int64_t* ndof, std::string& err) {
    if (!m) {
        err = "null mesh";
        return -1;
    }
    if (mesh_ensure_device(m)) {
        err = g_err;
        return -1;
    }
    *out = m->dm;
    *ndof = m->topo.ndof;
    return 0;
}

// Host value of the mesh's "certainly inside" margin (0: shortcut off).
extern "C" int punc_internal_inside_tol(const punc_mesh_t* m, double* tol) {
    if (!m || !tol) return -1;
    *tol = m->topo.inside_tol;
    return 0;
}
