// Host-side mesh setup for the PUNC particle path: synthetic mesh generation,
// localizer construction (facet neighbours, facet planes, barycentric
// matrices), patch volumes, exterior facets and the point-location grid.
// This is setup work done once per mesh; the per-step work is on the GPU.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "punc_types.h"

namespace punc_gpu {

// Raw tetrahedral mesh as a caller (or the generator) supplies it.
//  xyz        : nv*3 vertex coordinates
//  cells      : nc*4 vertex ids per cell, ascending (UFC ordering)
//  facet_bnd  : nc*4 boundary id of local facet i (opposite local vertex i);
//               read only for facets with a single adjacent cell: 1 = exterior,
//               2.. = objects (reference mesh.h:61-71 numbering)
//  vdof       : nv   CG1 dof id of each vertex (identity unless periodic)
struct HostMesh {
    int64_t nv = 0, nc = 0, ndof = 0;
    std::vector<double> xyz;
    std::vector<int32_t> cells;
    std::vector<int32_t> facet_bnd;
    std::vector<int32_t> vdof;
};

// Structured box [lo, hi]^3 of n^3 cubes, each split into 6 Kuhn tetrahedra
// sharing the cube's main diagonal (conforming).  Cells whose centroid lies
// within sphere_r of the origin are removed and their faces become object
// boundary 2.  periodic != 0 identifies opposite vertices in the dof map.
bool meshgen_box(int n, const double lo[3], const double hi[3], double sphere_r,
                 int periodic, HostMesh& out, std::string& err);

struct ExteriorFacetH {
    double area;
    double vertices[9];  // ascending global vertex order
    double normal[3];    // inward unit normal
    double basis[9];
    int32_t cell;        // the one adjacent cell
};

// Localizer and derived data for one mesh.
struct MeshTopo {
    int64_t nv = 0, nc = 0, ndof = 0;
    std::vector<CellRec> rec;         // packed barycentric rows, dofs, neighbours
    std::vector<CellPlanes> planes;   // packed facet planes
    std::vector<int64_t> adj;         // nc*4 facet neighbours as in the reference (signed long)
    std::vector<double> vcoords;      // nc*12 vertex coordinates per cell
    std::vector<ExteriorFacetH> ext;  // exterior facets (bnd id 1)
    std::vector<int32_t> bnd_ids;     // distinct boundary ids present (>= 1)
    double bbox_lo[3], bbox_hi[3];
    double coord_min = 0, coord_max = 0;  // min/max over all coordinates (UniformPosition domain)
    // Barycentric margin above which a point is certainly inside a cell for
    // the plane test: if all four computed lambda_i exceed it, every computed
    // facet projection is negative (rounding bounds of both, see
    // build_topology).  0 disables the shortcut.
    double inside_tol = 0;
    double coord_bound = 0;  // |coordinate| bound the margin was derived for
    // Every point whose computed facet projection |proj_i| is below
    // NEAR_FACE_DIST (within 1e-12 of a facet plane) has a computed
    // barycentric coordinate lambda_i below near_lam (rounding bounds as for
    // inside_tol); points above it in all four are not near a facet.
    // Infinity when the margin is disabled.
    double near_lam = 0;
    double volume = 0;                // total mesh volume
    // Uniform bucket grid for point location.
    int grid_n[3] = {1, 1, 1};
    double grid_lo[3], grid_h[3];
    std::vector<int32_t> bucket_start, bucket_cells;
};

bool build_topology(const HostMesh& m, MeshTopo& t, std::string& err);

// CG1 patch volumes per dof; voronoi: (D+1)/sum, else 1/sum
// (reference distributor.cpp:35-69).
void element_volume(const HostMesh& m, const MeshTopo& t, bool voronoi, std::vector<double>& dv_inv);

// Host point location: lowest-id cell whose four facet planes all give
// proj <= 0; -1 if none.  Replaces DOLFIN's BoundingBoxTree first collision
// (reference population.cpp:24-39); the two can differ only for points on faces.
int64_t locate_host(const MeshTopo& t, const double* x);

// Volume of every cell, |det(P2-P1, P3-P1, P4-P1)| / 6 (DOLFIN Cell::volume as
// used by reference diagnostics.h:461-462).
void cell_volumes(const MeshTopo& t, std::vector<double>& vol);

// Maxwellian flux helpers (reference distributions.cpp:52-72).
double maxwellian_flux_num(double vth, const double vd[3], const double n[3], double S);
double maxwellian_flux_max(double vth, const double vd[3], const double n[3]);
double maxwellian_pdf(double vth, const double vd[3], const double v[3]);

// Velocity distributions (reference distributions.cpp:23-245): parameters
// with the constructor's normalisation factor, the create_flux flags
// (has_flux_number, has_flux_max), the closed-form flux number and flux-pdf
// maximum for a facet of unit normal n and area S, and the pdf maximum used
// by the loading sampler.
bool vdf_setup(int kind, double vth, const double vd[3], double k, double alpha, double range, VdfParams& f,
               std::string& err);
bool vdf_has_flux_num(const VdfParams& f);
bool vdf_has_flux_max(const VdfParams& f);
double vdf_flux_num(const VdfParams& f, const double n[3], double S);
double vdf_flux_max(const VdfParams& f, const double n[3]);
double vdf_max(const VdfParams& f);

}  // namespace punc_gpu
