// Scene hook of the GPU path: the seeded particle set-ups that the
// reference builds in Standard_Tests::Initialize_Particles
// (mpm_3d/Standard_Tests/Standard_Tests.h:47-154,
// mpm_2d/Standard_Tests/Standard_Tests.h:47-323), plus the snow scenes the
// benchmark configurations name.
//
// Sampling: stratified jitter (shapes without a fixed count).  Every cell of the shape's bounding box is
// split into ppc sub-cells (2 per axis), each sub-cell gets one uniform
// sample from a counter-based generator keyed by (seed, cell, sub-cell),
// and samples inside the shape are kept.  Particle mass is
// density * dx^dim / ppc.  The reference's own Standard_Tests scenes draw a
// fixed number of uniform samples inside a sphere or box
// (Random_Numbers::Get_Vector_In_Sphere / Get_Uniform_Vector) with mass
// mass_density * shape volume / count; those are reproduced by rejection
// sampling from the same counter-based generator (Nova's generator is not
// available, so the positions are new, the statistics are the same).  Each
// shape names its entry in the scene's material table.
//
// Sized scenes (target_n > 0): the preset's shapes are scaled toward the
// volume that holds target_n particles at 8 per cell (clamped so the shapes
// stay inside the domain), and exactly target_n uniform samples are drawn
// inside them, split over the shapes by volume -- the same fixed-count
// sampling as test 18.  Batched and served scenes draw grid size and
// particle count independently, so their particles-per-cell varies.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <cfloat>

#include "../../include/mpm_gpu.h"
#include "mpm_internal.h"

namespace {

inline uint64_t mix64(uint64_t z) {
    // SplitMix64 finaliser (Steele, Lea, Flood 2014).
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t key3(uint64_t seed, uint64_t a, uint64_t b) {
    return mix64(mix64(mix64(seed) ^ a) ^ b);
}
// 24 random bits -> float in [0,1)
inline float unit_float(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }

enum ShapeKind { SPHERE = 0, OBOX = 1 };
struct Shape {
    ShapeKind kind;
    double c[3];
    double r;            // sphere radius
    double axis[3][3];   // box axes (unit)
    double half[3];      // box half extents
    float v[3];          // initial velocity
    int count = 0;       // >0: this many uniform samples (reference Standard_Tests style)
    double mass_density = 0.0;  // with count: particle mass = mass_density * volume / count
    int material = 0;    // entry of the scene's material table
    bool inside(const double* p, int dim) const {
        if (kind == SPHERE) {
            double s = 0;
            for (int a = 0; a < dim; ++a) s += (p[a] - c[a]) * (p[a] - c[a]);
            return s <= r * r;
        }
        for (int k = 0; k < dim; ++k) {
            double s = 0;
            for (int a = 0; a < dim; ++a) s += (p[a] - c[a]) * axis[k][a];
            if (std::fabs(s) > half[k]) return false;
        }
        return true;
    }
    void bounds(int dim, double* lo, double* hi) const {
        for (int a = 0; a < dim; ++a) {
            double e;
            if (kind == SPHERE) e = r;
            else {
                e = 0;
                for (int k = 0; k < dim; ++k) e += std::fabs(axis[k][a]) * half[k];
            }
            lo[a] = c[a] - e;
            hi[a] = c[a] + e;
        }
    }
};

struct Spec {
    mpm_scene_desc desc;
    std::vector<Shape> shapes;
    double density = 400.0;
    int64_t target_n = 0;    // >0: exactly this many uniform samples over all shapes
    bool fixed() const {
        for (const Shape& s : shapes)
            if (s.count > 0) return true;
        return false;
    }
};

double shape_volume(const Shape& s, int dim) {
    if (s.kind == SPHERE) return dim == 3 ? 4.0 / 3.0 * M_PI * s.r * s.r * s.r : M_PI * s.r * s.r;
    double v = 1.0;
    for (int k = 0; k < dim; ++k) v *= 2.0 * s.half[k];
    return v;
}

// Volume that holds n particles at 8 per cell of a counts^3 grid.
double volume_for(int64_t n, int counts) { return (double)n / (8.0 * (double)counts * counts * counts); }

double clampd(double x, double lo, double hi) { return std::max(lo, std::min(hi, x)); }
// x in [lo, hi] at fraction u; the midpoint if the interval is empty.
double place(double lo, double hi, double u) { return hi >= lo ? lo + (hi - lo) * u : 0.5 * (lo + hi); }

void snow_material(mpm_material& m, double E, double nu, double theta_c, double theta_s, double xi) {
    m.mu0 = (float)(E / (2.0 * (1.0 + nu)));
    m.lambda0 = (float)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    m.compression_yield = (float)(1.0 - theta_c);
    m.stretching_yield = (float)(1.0 + theta_s);
    m.hardening_factor = (float)xi;
    m.hardening_max = FLT_MAX;
    m.plastic = 1;
    m.bulk_modulus = 1e7f;
    m.gamma = 7.0f;
    m.density = 1.0f;
}

void base_desc(mpm_scene_desc& d, int dim, int counts) {
    std::memset(&d, 0, sizeof(d));
    d.dim = dim;
    d.counts = counts;
    d.gravity[1] = -9.8f;
    d.flip = 0.9f;
    d.cfl = 0.1f;
    d.min_dt = 1e-6f;
    d.max_dt = 5e-3f;
    d.frame_rate = 24.0f;
    d.cg_tolerance = 1e-6f;
    d.cg_max_iterations = 50;
    d.cg_min_iterations = 0;
    d.crop_ghost_x = 1;
    d.project_mode = 1;
    // rho=400, E=1.4e5, nu=0.2, theta_c=2.5e-2, theta_s=7.5e-3, xi=10 (Stomakhin et al. 2013)
    snow_material(d.material, 1.4e5, 0.2, 2.5e-2, 7.5e-3, 10.0);
}

// Upper bound of dt: the reference's 5e-3 (MPM_Driver.cpp:59), tightened to
// the elastic CFL dx / c_p with c_p = sqrt((lambda0 + 2 mu0) / rho) of the
// stiffest material (an EOS fluid at its initial density: K gamma instead of
// lambda0 + 2 mu0).  The reference's CFL (cfl dx / max|V|) ignores the wave
// speed and relies on a CG run to 1e-7; with at most 50 unpreconditioned
// iterations a stiff scene at dt = 5e-3 leaves residuals of order 10^2 on
// light nodes and blows up.
void elastic_dt_bound(mpm_scene_desc& d, double rho) {
    const int nmat = std::max(1, (int)d.nmaterials);
    for (int k = 0; k < nmat; ++k) {
        const mpm_material& m = k == 0 ? d.material : d.extra_materials[k - 1];
        const double modulus = m.eos ? (double)m.bulk_modulus * m.gamma * std::pow((double)m.density, m.gamma - 1.0)
                                     : m.lambda0 + 2.0 * m.mu0;
        const double dt = (1.0 / d.counts) / std::sqrt(modulus / rho);
        if (dt < d.max_dt) d.max_dt = (float)dt;
    }
}

// A material of the reference's Standard_Tests:
// Compute_Lame_Parameters(E, nu) and the plasticity fields
// (MPM_Constitutive_Model.h:105, MPM_Constitutive_Model.cpp:22-27).
mpm_material ref_material(double E, double nu, int plastic, double sy, double cy, double xi) {
    mpm_material m{};
    m.mu0 = (float)(E / (2.0 * (1.0 + nu)));
    m.lambda0 = (float)(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)));
    m.stretching_yield = (float)sy;
    m.compression_yield = (float)cy;
    m.hardening_factor = (float)xi;
    m.hardening_max = FLT_MAX;
    m.plastic = plastic;
    m.bulk_modulus = 1e7f;  // MPM_Particle::Initialize (MPM_Particle.h:51-56)
    m.gamma = 7.0f;
    m.density = 1.0f;
    return m;
}

// Scene description of the reference's Standard_Tests scenes: gravity
// (0,-2,0), CG tolerance 1e-7 and cap 10000 (MPM_Example.cpp:77-82).
void ref_desc(mpm_scene_desc& d, int dim, int counts) {
    base_desc(d, dim, counts > 0 ? counts : (dim == 3 ? 32 : 64));
    d.gravity[1] = -2.0f;
    d.cg_tolerance = 1e-7f;
    d.cg_max_iterations = 10000;
}

Shape box2(double x0, double y0, double x1, double y1) {
    Shape s{};
    s.kind = OBOX;
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) s.axis[k][a] = (a == k) ? 1.0 : 0.0;
    s.c[0] = 0.5 * (x0 + x1);
    s.c[1] = 0.5 * (y0 + y1);
    s.half[0] = 0.5 * (x1 - x0);
    s.half[1] = 0.5 * (y1 - y0);
    return s;
}

void add_plane(mpm_scene_desc& d, double nx, double ny, double nz, double sx, double sy, double sz) {
    mpm_plane& p = d.planes[d.nplanes++];
    double l = std::sqrt(nx * nx + ny * ny + nz * nz);
    p.normal[0] = (float)(nx / l);
    p.normal[1] = (float)(ny / l);
    p.normal[2] = (float)(nz / l);
    p.surface[0] = (float)sx;
    p.surface[1] = (float)sy;
    p.surface[2] = (float)sz;
    p.mu = 0.0f;
    p.sticky = 0;
}

Shape sphere(double x, double y, double z, double r, float vx, float vy, float vz) {
    Shape s{};
    s.kind = SPHERE;
    s.c[0] = x; s.c[1] = y; s.c[2] = z;
    s.r = r;
    s.v[0] = vx; s.v[1] = vy; s.v[2] = vz;
    return s;
}

bool build_spec(const std::string& name, int counts, uint64_t seed, int64_t target_n, Spec& sp) {
    mpm_scene_desc& d = sp.desc;
    sp.target_n = target_n;
    if (name == "disk2d") {
        // 2D snow disk r=0.1 at (0.5,0.7), V=(0,-2), ppc 4, ground y=0.05.
        base_desc(d, 2, counts > 0 ? counts : 64);
        add_plane(d, 0, 1, 0, 0, 0.05, 0);
        sp.shapes.push_back(sphere(0.5, 0.7, 0, 0.1, 0.f, -2.f, 0.f));
    } else if (name == "two_balls") {
        // Two snowballs r=0.12 at x=0.3 and x=0.7 moving head-on at 3 m/s each, ground y=0.05.
        base_desc(d, 3, counts > 0 ? counts : 128);
        add_plane(d, 0, 1, 0, 0, 0.05, 0);
        double r = 0.12;
        if (target_n > 0) r = clampd(std::cbrt(3.0 * volume_for(target_n, d.counts) / (8.0 * M_PI)), 0.03, 0.18);
        sp.shapes.push_back(sphere(0.3, 0.5, 0.5, r, 3.f, 0.f, 0.f));
        sp.shapes.push_back(sphere(0.7, 0.5, 0.5, r, -3.f, 0.f, 0.f));
    } else if (name == "avalanche") {
        // Snow slab 0.6 x 0.1 x 0.125 resting on a 30-degree incline; ground y=0.05 and
        // a wall at x=0.05 catch the slide.
        base_desc(d, 3, counts > 0 ? counts : 256);
        const double s30 = 0.5, c30 = std::sqrt(3.0) / 2.0;
        const double p0[3] = {0.1, 0.1, 0.5};
        add_plane(d, -s30, c30, 0, p0[0], p0[1], p0[2]);
        add_plane(d, 0, 1, 0, 0, 0.05, 0);
        add_plane(d, 1, 0, 0, 0.05, 0, 0);
        Shape s{};
        s.kind = OBOX;
        const double t[3] = {c30, s30, 0}, n[3] = {-s30, c30, 0}, z[3] = {0, 0, 1};
        // sized: the slab scaled by f, its lower end 0.1 up the incline from p0
        const double f = target_n > 0 ? clampd(std::cbrt(volume_for(target_n, d.counts) / 0.0075), 0.3, 1.3) : 1.0;
        const double along = target_n > 0 ? 0.1 + 0.3 * f : 0.6;
        for (int a = 0; a < 3; ++a) {
            s.c[a] = p0[a] + along * t[a] + 0.05 * f * n[a];
            s.axis[0][a] = t[a];
            s.axis[1][a] = n[a];
            s.axis[2][a] = z[a];
        }
        s.half[0] = 0.3 * f; s.half[1] = 0.05 * f; s.half[2] = 0.0625 * f;
        sp.shapes.push_back(s);
    } else if (name == "ref3d_test18" || name == "ref2d_test18") {
        // Reference test 18: two balls r=0.1 at x=0.8 / 0.2 with V=(-2,0.1) / (2,0.1),
        // E=30, nu=0.4, yields 1.01 / 0.95, xi=10, mass density 2, ground y=0.1
        // (mpm_3d/Standard_Tests/Standard_Tests.h:51-112,
        // mpm_2d/Standard_Tests/Standard_Tests.h:190-246).
        const int dim = name == "ref3d_test18" ? 3 : 2;
        ref_desc(d, dim, counts);
        d.material = ref_material(30.0, 0.4, 1, 1.01, 0.95, 10.0);
        add_plane(d, 0, 1, 0, 0, 0.1, 0);
        sp.shapes.push_back(sphere(0.8, 0.5, 0.5, 0.1, -2.f, 0.1f, 0.f));
        sp.shapes.push_back(sphere(0.2, 0.5, 0.5, 0.1, 2.f, 0.1f, 0.f));
        for (Shape& s : sp.shapes) {
            s.count = dim == 3 ? 20000 : 2000;
            s.mass_density = 2.0;
        }
    } else if (name == "ref2d_test1") {
        // Reference test 1: one particle in a ball r=0.1 at (0.2,0.5), V=(0.5,0), mass
        // density 2 over the ball, E=40, nu=0.4, yields 1.01 / 0.95, xi=10, no planes
        // (mpm_2d/Standard_Tests/Standard_Tests.h:55-81).
        ref_desc(d, 2, counts);
        d.material = ref_material(40.0, 0.4, 1, 1.01, 0.95, 10.0);
        Shape s = sphere(0.2, 0.5, 0, 0.1, 0.5f, 0.f, 0.f);
        s.count = 1;
        s.mass_density = 2.0;
        sp.shapes.push_back(s);
    } else if (name == "ref2d_test15") {
        // Reference test 15: 2000 particles in the box [0.4,0.6]^2, V=(-2,0.1), mass
        // density 2, E=40, nu=0.2, yields 1.01 / 0.95, xi=10, no planes
        // (mpm_2d/Standard_Tests/Standard_Tests.h:86-109).
        ref_desc(d, 2, counts);
        d.material = ref_material(40.0, 0.2, 1, 1.01, 0.95, 10.0);
        Shape s = box2(0.4, 0.4, 0.6, 0.6);
        s.v[0] = -2.f;
        s.v[1] = 0.1f;
        s.count = 2000;
        s.mass_density = 2.0;
        sp.shapes.push_back(s);
    } else if (name == "ref2d_test16") {
        // Reference test 16, snowball hitting the ground: a ball r=0.1 at (0.5,0.5) with
        // V=(0,-1.5), 2000 particles, mass density 2, E=50, nu=0.45, yields 1.005 / 0.975,
        // xi=10; below it a layer [0,1]x[0.1,0.125] of 1000 particles at rest, mass density
        // 1, E=5, nu=0.4, yields 1.001 / 0.975, xi=5 (the second material); no planes
        // (mpm_2d/Standard_Tests/Standard_Tests.h:111-155).
        ref_desc(d, 2, counts);
        d.material = ref_material(50.0, 0.45, 1, 1.005, 0.975, 10.0);
        d.nmaterials = 2;
        d.extra_materials[0] = ref_material(5.0, 0.4, 1, 1.001, 0.975, 5.0);
        Shape ball = sphere(0.5, 0.5, 0, 0.1, 0.f, -1.5f, 0.f);
        ball.count = 2000;
        ball.mass_density = 2.0;
        Shape layer = box2(0.0, 0.1, 1.0, 0.125);
        layer.count = 1000;
        layer.mass_density = 1.0;
        layer.material = 1;
        sp.shapes.push_back(ball);
        sp.shapes.push_back(layer);
    } else if (name == "ref2d_test17") {
        // Reference test 17, snowball hitting the wall: a ball r=0.1 at (0.5,0.5) with
        // V=(-2,0.1), 2000 particles, mass density 2, E=30, nu=0.4, yields 1.01 / 0.95,
        // xi=10; a wall x=0.1 and the ground y=0.1
        // (mpm_2d/Standard_Tests/Standard_Tests.h:157-188).
        ref_desc(d, 2, counts);
        d.material = ref_material(30.0, 0.4, 1, 1.01, 0.95, 10.0);
        add_plane(d, 1, 0, 0, 0.1, 0, 0);
        add_plane(d, 0, 1, 0, 0, 0.1, 0);
        Shape s = sphere(0.5, 0.5, 0, 0.1, -2.f, 0.1f, 0.f);
        s.count = 2000;
        s.mass_density = 2.0;
        sp.shapes.push_back(s);
    } else if (name == "ref2d_test20") {
        // Reference test 20, fluid source: no particles at the start, a box of ground, ceiling
        // and side walls, and a source region [0.45,0.55]^2 emitting EOS fluid particles (K=1,
        // gamma=7, density 1; mass 2 * 5e-4 each, V=(0,-1)) through MPM_Example::Add_Fluid_Source
        // (mpm_2d/Standard_Tests/Standard_Tests.h:285-300, MPM_Example.cpp:977-1001).  The source
        // is the host's (mpm_add_particles); the preset holds the walls and the fluid material.
        ref_desc(d, 2, counts);
        d.material = ref_material(40.0, 0.2, 0, 1.0, 1.0, 0.0);
        d.material.eos = 1;
        d.material.bulk_modulus = 1.0f;
        add_plane(d, 0, 1, 0, 0, 0.1, 0);
        add_plane(d, 0, -1, 0, 0, 0.9, 0);
        add_plane(d, 1, 0, 0, 0.1, 0, 0);
        add_plane(d, -1, 0, 0, 0.9, 0, 0);
    } else if (name == "ref2d_test19" || name == "ref3d_test19") {
        // Reference test 19, fluid: a ball r=0.1 at the centre with V=(0,-1), mass density
        // 2, every particle an EOS fluid particle with K=1, gamma=7, density 1 (E and nu are
        // set but unused); 2D: 2000 particles in a box of ground, ceiling and side walls
        // (mpm_2d/Standard_Tests/Standard_Tests.h:248-283); 3D: 20000 particles, the
        // ground and four side walls (mpm_3d/Standard_Tests/Standard_Tests.h:113-150).
        const int dim = name == "ref3d_test19" ? 3 : 2;
        ref_desc(d, dim, counts);
        d.material = dim == 3 ? ref_material(30.0, 0.4, 0, 1.0, 1.0, 0.0) : ref_material(40.0, 0.2, 0, 1.0, 1.0, 0.0);
        d.material.eos = 1;
        d.material.bulk_modulus = 1.0f;
        add_plane(d, 0, 1, 0, 0, 0.1, 0);
        if (dim == 2) {
            add_plane(d, 0, -1, 0, 0, 0.9, 0);
            add_plane(d, 1, 0, 0, 0.1, 0, 0);
            add_plane(d, -1, 0, 0, 0.9, 0, 0);
        } else {
            add_plane(d, 1, 0, 0, 0.1, 0, 0);
            add_plane(d, -1, 0, 0, 0.9, 0, 0);
            add_plane(d, 0, 0, 1, 0, 0, 0.1);
            add_plane(d, 0, 0, -1, 0, 0, 0.9);
        }
        Shape s = sphere(0.5, 0.5, dim == 3 ? 0.5 : 0.0, 0.1, 0.f, -1.f, 0.f);
        s.count = dim == 3 ? 20000 : 2000;
        s.mass_density = 2.0;
        sp.shapes.push_back(s);
    } else if (name == "blob" || name == "slab") {
        // Seeded random snow blob (sphere) or slab (box) for batched and served scenes.
        base_desc(d, 3, counts > 0 ? counts : 64);
        add_plane(d, 0, 1, 0, 0, 0.05, 0);
        const double u0 = unit_float(key3(seed, 11, 0)), u1 = unit_float(key3(seed, 11, 1));
        const double u2 = unit_float(key3(seed, 11, 2)), u3 = unit_float(key3(seed, 11, 3));
        if (name == "blob") {
            if (target_n > 0) {
                const double r = clampd(std::cbrt(3.0 * volume_for(target_n, d.counts) / (4.0 * M_PI)), 0.03, 0.3);
                sp.shapes.push_back(sphere(place(0.06 + r, 0.94 - r, u1), place(0.07 + r, 0.94 - r, u2),
                                           place(0.06 + r, 0.94 - r, u3), r,
                                           (float)(2.0 * (u1 - 0.5)), (float)(-1.0 - u2), 0.f));
            } else {
                const double r = 0.08 + 0.10 * u0;
                sp.shapes.push_back(sphere(0.3 + 0.4 * u1, 0.35 + 0.3 * u2, 0.3 + 0.4 * u3, r,
                                           (float)(2.0 * (u1 - 0.5)), (float)(-1.0 - u2), 0.f));
            }
        } else {
            Shape s{};
            s.kind = OBOX;
            for (int a = 0; a < 3; ++a)
                for (int k = 0; k < 3; ++k) s.axis[k][a] = (a == k) ? 1.0 : 0.0;
            s.half[0] = 0.12 + 0.15 * u0;
            s.half[1] = 0.03 + 0.04 * u1;
            s.half[2] = 0.08 + 0.1 * u2;
            if (target_n > 0) {
                const double f = std::cbrt(volume_for(target_n, d.counts) / shape_volume(s, 3));
                s.half[0] = clampd(f * s.half[0], 0.01, 0.44);
                s.half[1] = clampd(f * s.half[1], 0.01, 0.3);
                s.half[2] = clampd(f * s.half[2], 0.01, 0.44);
            }
            s.c[0] = 0.5; s.c[1] = std::min(0.06 + s.half[1] + 0.2 * u3, 0.94 - s.half[1]); s.c[2] = 0.5;
            s.v[0] = 0.f; s.v[1] = -0.5f; s.v[2] = 0.f;
            sp.shapes.push_back(s);
        }
    } else {
        return false;
    }
    double rho = sp.shapes.empty() ? 2.0 : sp.density;
    if (sp.fixed()) {
        rho = 1e30;
        for (const Shape& s : sp.shapes) rho = std::min(rho, s.mass_density);
    }
    elastic_dt_bound(d, rho);
    return true;
}

// Generates particles; if x == nullptr only counts.
int64_t generate(const Spec& sp, uint64_t seed, int64_t cap, float* x, float* v, float* mass,
                 int32_t* material = nullptr) {
    const mpm_scene_desc& d = sp.desc;
    const int dim = d.dim;
    const int n = d.counts;
    const double dx = 1.0 / n;
    int64_t count = 0;
    auto emit = [&](const double* p, const Shape& s, float m) {
        if (x && count < cap) {
            for (int a = 0; a < dim; ++a) {
                x[count * dim + a] = (float)p[a];
                v[count * dim + a] = s.v[a];
            }
            mass[count] = m;
        }
        if (material && count < cap) material[count] = s.material;
        ++count;
    };
    if (sp.target_n > 0) {
        // exactly target_n uniform samples, split over the shapes by volume
        double vtot = 0.0;
        for (const Shape& s : sp.shapes) vtot += shape_volume(s, dim);
        int64_t given = 0;
        for (size_t si = 0; si < sp.shapes.size(); ++si) {
            const Shape& s = sp.shapes[si];
            const double vs = shape_volume(s, dim);
            const int64_t ns = si + 1 == sp.shapes.size() ? sp.target_n - given
                                                          : (int64_t)std::llround((double)sp.target_n * vs / vtot);
            given += ns;
            if (ns <= 0) continue;
            const float m = (float)(sp.density * vs / (double)ns);
            double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            s.bounds(dim, lo, hi);
            uint64_t attempt = 0;
            for (int64_t i = 0; i < ns;) {
                double p[3] = {0, 0, 0};
                for (int a = 0; a < dim; ++a)
                    p[a] = lo[a] + (hi[a] - lo[a]) * unit_float(key3(seed ^ (si * 0x1000193ull), attempt, (uint64_t)a + 7));
                ++attempt;
                if (!s.inside(p, dim)) continue;
                emit(p, s, m);
                ++i;
            }
        }
        return count;
    }
    if (sp.fixed()) {
        for (size_t si = 0; si < sp.shapes.size(); ++si) {
            const Shape& s = sp.shapes[si];
            const float m = (float)(s.mass_density * shape_volume(s, dim) / s.count);
            double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
            s.bounds(dim, lo, hi);
            uint64_t attempt = 0;
            for (int i = 0; i < s.count;) {
                double p[3] = {0, 0, 0};
                for (int a = 0; a < dim; ++a) {
                    const double u = unit_float(key3(seed ^ (si * 0x1000193ull), attempt, a));
                    p[a] = s.kind == SPHERE ? s.c[a] + s.r * (2.0 * u - 1.0) : lo[a] + (hi[a] - lo[a]) * u;
                }
                ++attempt;
                if (!s.inside(p, dim)) continue;
                emit(p, s, m);
                ++i;
            }
        }
        return count;
    }
    const int sub = 2;  // 2 sub-cells per axis: ppc 4 (2D) / 8 (3D)
    const int ppc = dim == 3 ? 8 : 4;
    const float m = (float)(sp.density * std::pow(dx, dim) / ppc);
    for (size_t si = 0; si < sp.shapes.size(); ++si) {
        const Shape& s = sp.shapes[si];
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        s.bounds(dim, lo, hi);
        int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) {
            clo[a] = std::max(1, (int)std::floor(lo[a] / dx) - 1);
            chi[a] = std::min(n - 2, (int)std::floor(hi[a] / dx) + 1);
        }
        const int zlo = dim == 3 ? clo[2] : 0, zhi = dim == 3 ? chi[2] : 0;
        for (int i = clo[0]; i <= chi[0]; ++i)
            for (int j = clo[1]; j <= chi[1]; ++j)
                for (int k = zlo; k <= zhi; ++k) {
                    const uint64_t cell = ((uint64_t)i * n + j) * n + k;
                    const int nsub = dim == 3 ? 8 : 4;
                    for (int q = 0; q < nsub; ++q) {
                        const int qi[3] = {q & 1, (q >> 1) & 1, (q >> 2) & 1};
                        const int ci[3] = {i, j, k};
                        double p[3] = {0, 0, 0};
                        for (int a = 0; a < dim; ++a) {
                            const double u = unit_float(key3(seed + 0x51ull * si, cell, (uint64_t)(q * 4 + a)));
                            p[a] = (ci[a] + (qi[a] + u) / sub) * dx;
                        }
                        if (!s.inside(p, dim)) continue;
                        emit(p, s, m);
                    }
                }
    }
    return count;
}

}  // namespace

extern "C" int mpm_scene_preset_sized(const char* name, int32_t counts, uint64_t seed, int64_t target_n,
                                      mpm_scene_desc* desc, int64_t* n_particles) {
    if (!name || !desc) return mpm_set_error(MPM_E_INVALID, "mpm_scene_preset: null argument");
    if (target_n < 0 || target_n > ((int64_t)1 << 30)) return mpm_set_error(MPM_E_INVALID, "mpm_scene_preset: bad target_n");
    Spec sp;
    if (!build_spec(name, counts, seed, target_n, sp))
        return mpm_set_error(MPM_E_INVALID, std::string("unknown scene ") + name);
    if (target_n > 0 && sp.fixed())
        return mpm_set_error(MPM_E_INVALID, std::string("scene ") + name + " has a fixed particle count");
    *desc = sp.desc;
    if (n_particles) *n_particles = target_n > 0 ? target_n : generate(sp, seed, 0, nullptr, nullptr, nullptr);
    return MPM_OK;
}

extern "C" int mpm_scene_generate_sized(const char* name, int32_t counts, uint64_t seed, int64_t target_n,
                                        int64_t cap, float* x, float* v, float* mass, int64_t* n_out) {
    if (!name) return mpm_set_error(MPM_E_INVALID, "mpm_scene_generate: null argument");
    mpm_scene_desc d;
    int64_t n = 0;
    int rc = mpm_scene_preset_sized(name, counts, seed, target_n, &d, &n);
    if (rc) return rc;
    if (n_out) *n_out = n;
    if (n == 0) return MPM_OK;  // a scene without particles (the fluid-source test)
    if (!x || !v || !mass) return mpm_set_error(MPM_E_INVALID, "mpm_scene_generate: null argument");
    if (n > cap) return mpm_set_error(MPM_E_CAPACITY, "mpm_scene_generate: capacity too small");
    Spec sp;
    build_spec(name, counts, seed, target_n, sp);
    generate(sp, seed, cap, x, v, mass);
    return MPM_OK;
}

extern "C" int mpm_scene_preset(const char* name, int32_t counts, uint64_t seed, mpm_scene_desc* desc,
                                int64_t* n_particles) {
    return mpm_scene_preset_sized(name, counts, seed, 0, desc, n_particles);
}

extern "C" int mpm_scene_generate(const char* name, int32_t counts, uint64_t seed, int64_t cap, float* x, float* v,
                                  float* mass, int64_t* n_out) {
    return mpm_scene_generate_sized(name, counts, seed, 0, cap, x, v, mass, n_out);
}

extern "C" int mpm_scene_materials(const char* name, int32_t counts, uint64_t seed, int64_t target_n, int64_t cap,
                                   int32_t* material, int64_t* n_out) {
    if (!name) return mpm_set_error(MPM_E_INVALID, "mpm_scene_materials: null argument");
    mpm_scene_desc d;
    int64_t n = 0;
    int rc = mpm_scene_preset_sized(name, counts, seed, target_n, &d, &n);
    if (rc) return rc;
    if (n_out) *n_out = n;
    if (n == 0) return MPM_OK;
    if (!material) return mpm_set_error(MPM_E_INVALID, "mpm_scene_materials: null argument");
    if (n > cap) return mpm_set_error(MPM_E_CAPACITY, "mpm_scene_materials: capacity too small");
    Spec sp;
    build_spec(name, counts, seed, target_n, sp);
    std::vector<float> x((size_t)n * d.dim), v((size_t)n * d.dim), m((size_t)n);
    generate(sp, seed, cap, x.data(), v.data(), m.data(), material);
    return MPM_OK;
}
