// Drop-in for punc/include/punc/mesh.h (PUNC++) on the MI355X particle path.
//
// punc::Mesh holds a tetrahedral mesh and its device-side localizer.  The
// reference loads a DOLFIN mesh from file (mesh.cpp:41-92); DOLFIN is not part
// of this library, so a Mesh is built from plain arrays (vertex coordinates,
// ascending cell vertex ids, per-(cell, local facet) boundary ids, vertex ->
// CG1 dof map) or by the structured box generator.  Boundary ids follow the
// reference (mesh.h:61-71): 0 interior, 1 exterior, 2.. objects.
#ifndef PUNC_GPU_MESH_H
#define PUNC_GPU_MESH_H

#include <punc_gpu.h>

#include <array>
#include <cstdlib>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace punc {

//! Throws std::runtime_error with the library's message on a nonzero status.
inline void check(int rc) {
    if (rc != 0) throw std::runtime_error(punc_gpu_last_error());
}

//! Exterior facet (reference mesh.h:38-44) plus its owning cell.
struct ExteriorFacet {
    double area;                   ///< Area of the facet
    std::vector<double> vertices;  ///< 3x3 vertex coordinates, ascending vertex ids
    std::vector<double> normal;    ///< Inward unit normal
    std::vector<double> basis;     ///< Orthonormal basis, first column = normal
    int cell;                      ///< The adjacent cell
};

class Mesh {
public:
    using size_t = std::size_t;

    size_t dim = 3;          ///< Geometric dimension (3 only)
    size_t ext_bnd_id = 1;   ///< Id of the exterior boundary
    size_t num_objects = 0;  ///< Number of objects
    std::vector<ExteriorFacet> exterior_facets;

    std::vector<double> xyz;          ///< nv*3
    std::vector<int32_t> cells;       ///< nc*4
    std::vector<int32_t> facet_bnd;   ///< nc*4
    std::vector<int32_t> vdof;        ///< nv
    int64_t ndof = 0;

    Mesh(std::vector<double> xyz_, std::vector<int32_t> cells_, std::vector<int32_t> facet_bnd_,
         std::vector<int32_t> vdof_ = {}, int64_t ndof_ = -1)
        : xyz(std::move(xyz_)), cells(std::move(cells_)), facet_bnd(std::move(facet_bnd_)),
          vdof(std::move(vdof_)), ndof(ndof_) {
        const int64_t nv = (int64_t)xyz.size() / 3;
        if (vdof.empty()) {
            vdof.resize(nv);
            for (int64_t v = 0; v < nv; ++v) vdof[v] = (int32_t)v;
            ndof = nv;
        }
        punc_mesh_t* h = nullptr;
        check(punc_gpu_mesh_create(nv, xyz.data(), (int64_t)cells.size() / 4, cells.data(),
                                   facet_bnd.data(), vdof.data(), ndof, &h));
        handle_.reset(h, punc_gpu_mesh_destroy);
        init_info();
    }

    //! Structured box of n^3 cubes x 6 tetrahedra, optional spherical object.
    static Mesh box(int n, std::array<double, 3> lo, std::array<double, 3> hi, double sphere_r = 0.0,
                    bool periodic = false) {
        punc_hostmesh_t* hm = nullptr;
        check(punc_gpu_meshgen_box(n, lo.data(), hi.data(), sphere_r, periodic ? 1 : 0, &hm));
        int64_t nv, nc, nd;
        punc_gpu_hostmesh_sizes(hm, &nv, &nc, &nd);
        std::vector<double> x(nv * 3);
        std::vector<int32_t> c(nc * 4), b(nc * 4), d(nv);
        punc_gpu_hostmesh_copy(hm, x.data(), c.data(), b.data(), d.data());
        punc_gpu_hostmesh_free(hm);
        return Mesh(std::move(x), std::move(c), std::move(b), std::move(d), nd);
    }

    int64_t num_vertices() const { return (int64_t)xyz.size() / 3; }
    int64_t num_cells() const { return (int64_t)cells.size() / 4; }

    //! Mesh::volume (mesh.cpp:147-164).
    double volume() const { return volume_; }

    //! Mesh::domain_size (mesh.cpp:124-145).
    std::vector<double> domain_size() const {
        std::vector<double> Ld(3, 0.0);
        for (size_t i = 0; i < xyz.size(); ++i)
            if (xyz[i] >= Ld[i % 3]) Ld[i % 3] = xyz[i];
        return Ld;
    }

    punc_mesh_t* handle() const { return handle_.get(); }

private:
    std::shared_ptr<punc_mesh_t> handle_;
    double volume_ = 0.0;

    void init_info() {
        int64_t nv, nc, nd, ne;
        check(punc_gpu_mesh_info(handle(), &nv, &nc, &nd, &ne, &volume_));
        std::vector<double> area(ne), verts(ne * 9), normal(ne * 3), basis(ne * 9);
        std::vector<int32_t> cell(ne);
        check(punc_gpu_exterior_facets(handle(), area.data(), verts.data(), normal.data(), basis.data(),
                                       cell.data()));
        exterior_facets.clear();
        for (int64_t j = 0; j < ne; ++j)
            exterior_facets.push_back(ExteriorFacet{area[j],
                                                    {verts.begin() + 9 * j, verts.begin() + 9 * j + 9},
                                                    {normal.begin() + 3 * j, normal.begin() + 3 * j + 3},
                                                    {basis.begin() + 9 * j, basis.begin() + 9 * j + 9},
                                                    cell[j]});
        std::vector<int64_t> adj(nc * 4);
        check(punc_gpu_mesh_localizer(handle(), adj.data(), nullptr, nullptr, nullptr, nullptr));
        std::vector<char> seen(64, 0);
        for (int64_t a : adj)
            if (a < -1 && a > -64) seen[-a] = 1;
        num_objects = 0;
        for (char s : seen) num_objects += s;
    }
};

//! The function spaces the particle path uses (the stand-in for
//! df::FunctionSpace): CG1 or DG0, scalar or 3-vector, on a Mesh.  A space
//! fixes the length and layout of a Function's values.
class FunctionSpace {
public:
    FunctionSpace(const Mesh& mesh, int value_dim, bool dg0) : mesh_(&mesh), value_dim_(value_dim), dg0_(dg0) {}
    const Mesh& mesh() const { return *mesh_; }
    int value_dim() const { return value_dim_; }
    bool dg0() const { return dg0_; }
    //! Number of values (df::FunctionSpace::dim).
    int64_t dim() const { return (dg0_ ? mesh_->num_cells() : mesh_->ndof) * value_dim_; }

private:
    const Mesh* mesh_;
    int value_dim_;
    bool dg0_;
};

//! CG1_space, CG1_vector_space, DG0_space, DG0_vector_space (reference
//! poisson.h:146-174).  A periodic mesh carries its constraint in the
//! vertex -> dof map, so CG1_space takes no constraint argument.
inline FunctionSpace CG1_space(const Mesh& mesh) { return FunctionSpace(mesh, 1, false); }
inline FunctionSpace CG1_vector_space(const Mesh& mesh) { return FunctionSpace(mesh, 3, false); }
inline FunctionSpace DG0_space(const Mesh& mesh) { return FunctionSpace(mesh, 1, true); }
inline FunctionSpace DG0_vector_space(const Mesh& mesh) { return FunctionSpace(mesh, 3, true); }

//! Allocator of page-locked host memory (punc_gpu_host_alloc), so that a
//! Function's values go to and from the device as one DMA transfer.  Where
//! page-locked memory cannot be had (no device), ordinary heap memory is used;
//! a 64-byte header records which, for deallocate.
template <class T>
struct PinnedAllocator {
    using value_type = T;
    PinnedAllocator() = default;
    template <class U>
    PinnedAllocator(const PinnedAllocator<U>&) {}
    T* allocate(std::size_t n) {
        const std::size_t bytes = n * sizeof(T) + 64;
        void* p = nullptr;
        bool pinned = punc_gpu_host_alloc((int64_t)bytes, &p) == 0 && p;
        if (!pinned) p = std::malloc(bytes);
        if (!p) throw std::bad_alloc();
        *static_cast<int*>(p) = pinned ? 1 : 0;
        return reinterpret_cast<T*>(static_cast<char*>(p) + 64);
    }
    void deallocate(T* q, std::size_t) {
        if (!q) return;
        char* p = reinterpret_cast<char*>(q) - 64;
        if (*reinterpret_cast<int*>(p))
            punc_gpu_host_free(p);
        else
            std::free(p);
    }
    template <class U>
    bool operator==(const PinnedAllocator<U>&) const { return true; }
    template <class U>
    bool operator!=(const PinnedAllocator<U>&) const { return false; }
};

//! A CG1 function on a mesh: ndof values per component (the stand-in for
//! df::Function on the particle path; E has 3 components, rho has 1).
class Function {
public:
    //! dg0: one value per cell and component (cell-major) instead of per dof.
    Function(const Mesh& mesh, int value_dim, bool dg0 = false)
        : value_dim(value_dim), dg0(dg0), values((dg0 ? mesh.num_cells() : mesh.ndof) * value_dim, 0.0) {}
    //! A zero function in V, as df::Function(std::make_shared<const df::FunctionSpace>(V)).
    explicit Function(const FunctionSpace& V) : Function(V.mesh(), V.value_dim(), V.dg0()) {}
    explicit Function(const std::shared_ptr<const FunctionSpace>& V) : Function(*V) {}
    int value_dim;
    bool dg0;
    std::vector<double, PinnedAllocator<double>> values;  ///< dof-major: values[dof*value_dim + component]
};

}  // namespace punc

#endif
