// Shared host/device definitions for the PUNC particle path on MI355X (gfx950).
//
// Everything here is compiled twice: by hipcc for the device kernels and by the
// host pass for the setup code in mesh.cpp.  All floating-point expressions are
// written in the evaluation order of the PUNC++ particle code and the library is
// built with -ffp-contract=off, so host and device produce identical bits for
// +, -, *, / and sqrt.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PUNC_HD __host__ __device__ inline
#else
#define PUNC_HD inline
#endif

namespace punc_gpu {

// ----------------------------------------------------------------------------
// Packed per-cell mesh data.
//
// CellRec (128 B, one cache line): the barycentric matrix rows 0..2 of
// Cell<3>::init_barycentric_matrix (reference population.cpp:181-240, row 3 is
// implicit: lambda3 = 1 - l0 - l1 - l2, population.h:885-893), the four CG1 dof
// ids of the cell's vertices (local vertex order), and the four facet
// neighbours (cell id, or -bnd_id on a boundary facet; population.h:363-373).
//
// CellPlanes (128 B): the facet planes (c0, nx, ny, nz) per local facet,
// c0 = -(midpoint . n) with n the outward unit normal (population.h:375-382).
// Local facet i is opposite local vertex i.
// ----------------------------------------------------------------------------
struct alignas(128) CellRec {
    double bary[12];
    int32_t dof[4];
    int32_t nbr[4];
};
struct alignas(128) CellPlanes {
    double p[16];
};
static_assert(sizeof(CellRec) == 128, "CellRec must be one 128-byte line");
static_assert(sizeof(CellPlanes) == 128, "CellPlanes must be one 128-byte line");

// Particle cell field: bits [0, cell_bits) hold the cell id, the bits above it
// hold the species index; a negative value marks a particle that left through
// boundary -value during the last relocation.
PUNC_HD int32_t pack_cs(int32_t species, int32_t cell, int cell_bits) {
    return (species << cell_bits) | cell;
}
PUNC_HD int32_t cs_cell(int32_t cs, int cell_bits) { return cs & ((1 << cell_bits) - 1); }
PUNC_HD int32_t cs_species(int32_t cs, int cell_bits) { return cs >> cell_bits; }

// Relocation walk outcome when the hop cap is exceeded (never a valid -bnd id).
constexpr int32_t WALK_FAILED = -(1 << 30);
constexpr int WALK_MAX_HOPS = 100000;
// A particle whose position gives |proj| below this for some facet plane of
// its cell is "near a face": counted and reported (a walk decision there
// rests on rounding, and the reference's own location may differ).
constexpr double NEAR_FACE_DIST = 1e-12;

// ----------------------------------------------------------------------------
// Counter-based RNG: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11).  Replaces the reference's unseeded
// std::mt19937_64 (injector.h:161) so that every draw is addressable by
// (seed, step, species, facet, slot, attempt) and identical on CPU and GPU.
// ----------------------------------------------------------------------------
struct U64x4 { uint64_t v[4]; };

PUNC_HD void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    // Four 32x32->64 products (v_mad_u64_u32, each adding a carry word that
    // cannot overflow it); the low word comes from the same partial sums.
    // a * b and __umul64hi(a, b) separately cost two more 32-bit multiplies,
    // which dominate the injection sampler's Philox rounds.
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t t = (uint64_t)a0 * b0;
    const uint64_t u = (uint64_t)a1 * b0 + (t >> 32);
    const uint64_t v = (uint64_t)a0 * b1 + (uint32_t)u;
    *hi = (uint64_t)a1 * b1 + (u >> 32) + (v >> 32);
    *lo = (v << 32) | (uint32_t)t;
#else
    unsigned __int128 p = (unsigned __int128)a * b;
    *lo = (uint64_t)p;
    *hi = (uint64_t)(p >> 64);
#endif
}

PUNC_HD U64x4 philox4x64_10(U64x4 c, uint64_t k0, uint64_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ull, c.v[0], &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ull, c.v[2], &hi1, &lo1);
        U64x4 n;
        n.v[0] = hi1 ^ c.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ c.v[3] ^ k1;
        n.v[3] = lo0;
        c = n;
        k0 += 0x9E3779B97F4A7C15ull;
        k1 += 0xBB67AE8584CAA73Bull;
    }
    return c;
}

// Uniform double strictly inside (0, 1): ((x >> 12) + 1/2) * 2^-52.
PUNC_HD double u01(uint64_t x) {
    return ((double)(x >> 12) * 2.0 + 1.0) * 0x1.0p-53;
}

// Stream tags (second key word) keep the draws of different phases disjoint.
constexpr uint64_t TAG_INJECT_COUNT = 0x494e4a43ull;  // per (species, facet) count rounding
constexpr uint64_t TAG_INJECT = 0x494e4a50ull;        // per injected candidate
constexpr uint64_t TAG_LOAD = 0x4c4f4144ull;          // initial loading
constexpr uint64_t TAG_FLUX = 0x464c5558ull;          // flux-table Monte Carlo (create_flux)

// ----------------------------------------------------------------------------
// Velocity distributions (reference distributions.cpp:23-245): the pdf value
// at v, each expression in the reference's evaluation order.  factor, vth2
// and vth4 come from vdf_setup (host, once per species).
// ----------------------------------------------------------------------------
enum VdfKind : int32_t { VDF_MAXWELLIAN = 0, VDF_KAPPA = 1, VDF_CAIRNS = 2, VDF_KAPPA_CAIRNS = 3 };

struct VdfParams {
    int32_t kind;
    int32_t pad;
    double vth, vth2, vth4, factor;
    double k;      // spectral index kappa (Kappa, KappaCairns)
    double alpha;  // Cairns alpha (Cairns, KappaCairns)
    double vd[3];
    double range;  // sampling box vd +- range * vth
};

PUNC_HD double vdf_eval(const VdfParams& f, double w0, double w1, double w2) {
    const int dim = 3;
    double v2 = 0.0;
    v2 += (w0 - f.vd[0]) * (w0 - f.vd[0]);
    v2 += (w1 - f.vd[1]) * (w1 - f.vd[1]);
    v2 += (w2 - f.vd[2]) * (w2 - f.vd[2]);
    switch (f.kind) {
        case VDF_KAPPA:  // distributions.cpp:99-107
            return f.factor * pow(1.0 + v2 / ((2 * f.k - dim) * f.vth2), -(f.k + 1.0));
        case VDF_CAIRNS: {  // distributions.cpp:145-154
            const double v4 = v2 * v2;
            return f.factor * (1 + f.alpha * v4 / f.vth4) * exp(-0.5 * v2 / f.vth2);
        }
        case VDF_KAPPA_CAIRNS: {  // distributions.cpp:201-211
            const double v4 = v2 * v2;
            return f.factor * (1.0 + f.alpha * v4 / f.vth4) * pow(1.0 + v2 / ((2 * f.k - dim) * f.vth2), -(f.k + 1.0));
        }
        default:  // Maxwellian, distributions.cpp:31-39
            return f.factor * exp(-0.5 * v2 / f.vth2);
    }
}

// The flux pdf of a Maxwellian only: the same expressions, in the same
// order, as vdf_flux with f.kind == VDF_MAXWELLIAN (bit-identical), without
// the other distributions' code.
PUNC_HD double vdf_flux_maxwellian(double factor, double vth2, const double vd[3], double w0, double w1, double w2,
                                   double n0, double n1, double n2) {
    double vn = 0.0;
    vn += w0 * n0;
    vn += w1 * n1;
    vn += w2 * n2;
    double v2 = 0.0;
    v2 += (w0 - vd[0]) * (w0 - vd[0]);
    v2 += (w1 - vd[1]) * (w1 - vd[1]);
    v2 += (w2 - vd[2]) * (w2 - vd[2]);
    return (vn > 0.0) * vn * (factor * exp(-0.5 * v2 / vth2));
}

// Flux pdf (v.n)^+ f(v) (distributions.h:93-102).
PUNC_HD double vdf_flux(const VdfParams& f, double w0, double w1, double w2, double n0, double n1, double n2) {
    double vn = 0.0;
    vn += w0 * n0;
    vn += w1 * n1;
    vn += w2 * n2;
    return (vn > 0.0) * vn * vdf_eval(f, w0, w1, w2);
}

// ----------------------------------------------------------------------------
// Inverse complementary error function.  Initial guess from Abramowitz &
// Stegun 26.2.23 (normal quantile, |err| < 4.5e-4), then three Halley steps on
// erfc(x) = y, which converge cubically to full double precision.  Boost's
// erfc_inv (distributions.cpp:47) is not available here.
// ----------------------------------------------------------------------------
PUNC_HD double erfc_inv_approx(double y) {
    // erfc(-x) = 2 - erfc(x): map y > 1 to 2 - y (an exact subtraction).
    const double sgn = y > 1.0 ? -1.0 : 1.0;
    if (y > 1.0) y = 2.0 - y;
    // p = y/2 is the lower-tail probability of a standard normal at -sqrt(2) x.
    double p = 0.5 * y;
    double pp = p < 0.5 ? p : 1.0 - p;
    double t = sqrt(-2.0 * log(pp));
    double z = t - (2.515517 + 0.802853 * t + 0.010328 * t * t) /
                       (1.0 + 1.432788 * t + 0.189269 * t * t + 0.001308 * t * t * t);
    // z approximates the upper quantile for pp; erfc(x)=y means x = z/sqrt(2) when p<0.5.
    double x = (p < 0.5 ? z : -z) * 0.70710678118654752440;
    const double two_over_sqrtpi = 1.12837916709551257390;
    // Near y = 1 solve erf(x) = 1 - y instead (exact subtraction, full
    // relative accuracy for small x).
    const bool central = y >= 0.5 && y <= 1.5;
    const double zc = 1.0 - y;
    for (int it = 0; it < 3; ++it) {
        double f = central ? zc - erf(x) : erfc(x) - y;
        double fp = -two_over_sqrtpi * exp(-x * x);
        double d = f / fp;
        x = x - d / (1.0 + x * d);
    }
    return sgn * x;
}

}  // namespace punc_gpu

namespace punc_gpu {

// Facet-plane projection in the reference's evaluation order
// (population.h:547-552): proj = ((c0 + c1*x) + c2*y) + c3*z.
PUNC_HD double plane_proj(const double* p, double x, double y, double z) {
    return ((p[0] + p[1] * x) + p[2] * y) + p[3] * z;
}

// Point location through a uniform bucket grid: the lowest-id candidate cell
// whose four facet planes all give proj <= 0, or -1.  Bucket lists are in
// ascending cell order and hold every cell whose bounding box meets the bucket,
// so the result equals a brute-force scan over all cells in id order.
PUNC_HD int32_t locate_in_grid(const CellPlanes* planes, const int32_t* bucket_start,
                               const int32_t* bucket_cells, const int* gn, const double* glo,
                               const double* gh, double x, double y, double z,
                               int32_t cell_lo = 0, int32_t cell_hi = 0x7fffffff) {
    // A non-finite coordinate is outside every cell (a NaN would pass every
    // "proj > 0" test and land in the first candidate).
    if (!(fabs(x) <= 1.79e308 && fabs(y) <= 1.79e308 && fabs(z) <= 1.79e308)) return -1;
    double q[3] = {x, y, z};
    int b[3];
    for (int d = 0; d < 3; ++d) {
        double f = floor((q[d] - glo[d]) / gh[d]);
        int i = (f < 0.0) ? 0 : (f >= (double)gn[d] ? gn[d] - 1 : (int)f);
        b[d] = i;
    }
    int64_t bi = (int64_t)b[0] + (int64_t)gn[0] * ((int64_t)b[1] + (int64_t)gn[1] * b[2]);
    for (int32_t k = bucket_start[bi]; k < bucket_start[bi + 1]; ++k) {
        int32_t c = bucket_cells[k];
        if (c < cell_lo || c >= cell_hi) continue;
        const double* p = planes[c].p;
        bool inside = true;
        for (int f = 0; f < 4; ++f) inside = inside && !(plane_proj(p + 4 * f, x, y, z) > 0.0);
        if (inside) return c;
    }
    return -1;
}

}  // namespace punc_gpu
