// MPM_Example on the GPU path: the reference's simulation-stage surface
// (MPM_Example.h:229-336) kept method for method, with each stage mapped to
// the libmpm_gpu C ABI (include/mpm_gpu.h).  A scene plugin subclasses it and
// fills `particles` and `barriers` in Initialize_Particles, as the
// reference's Standard_Tests do (mpm_3d/Standard_Tests/Standard_Tests.h:47-154).
//
// Two modes:
//   fused (default)  the stage calls of MPM_Driver::Advance_Step
//                    (MPM_Driver.cpp:103-116) are collected and run as one
//                    mpm_substep when Update_Particle_Velocities_And_Positions
//                    is reached -- one launch sequence per substep.
//   staged           every stage method runs its own mpm_run_stage, for
//                    stage-by-stage parity checks.
#pragma once

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <sys/stat.h>
#include <cerrno>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpm_gpu.h"

namespace mpm_host {

inline void check(int rc, const char* what) {
    if (rc != MPM_OK) throw std::runtime_error(std::string(what) + ": " + mpm_last_error());
}

template <class T, int d>
struct Vec {
    T v[d];
    Vec() { for (int i = 0; i < d; ++i) v[i] = T(0); }
    T& operator()(int i) { return v[i]; }
    const T& operator()(int i) const { return v[i]; }
};

// MPM_Particle (MPM_Particle.h:14-60) restricted to the state the substep
// carries: position, velocity, mass, rest volume, elastic F and det(Fp), and
// for an EOS fluid particle its density.  The constitutive parameters and the
// eos flag, bulk modulus and gamma live in the example's material table;
// `material` names the particle's entry.
template <class T, int d>
struct MPM_Particle {
    bool valid = true;
    Vec<T, d> X, V;
    T mass = 0, volume = 0;
    T Fe[d * d];
    T Jp = 1;
    int material = 0;
    T density = 1;
    MPM_Particle() { for (int i = 0; i < d * d; ++i) Fe[i] = (i % (d + 1) == 0) ? T(1) : T(0); }
};

// MPM_Plane_Barrier (MPM_Plane_Barrier.h:13-32)
template <class T, int d>
struct MPM_Plane_Barrier {
    T mu;
    Vec<T, d> normal, surface;
    bool sticky;
    MPM_Plane_Barrier(T mu_input, const Vec<T, d>& n, const Vec<T, d>& s, bool sticky_input = false)
        : mu(mu_input), normal(n), surface(s), sticky(sticky_input) {}
};

template <class T, int d>
class MPM_Example {
  public:
    using TV = Vec<T, d>;
    using T_Particle = MPM_Particle<T, d>;
    using T_Barrier = MPM_Plane_Barrier<T, d>;

    // options (MPM_Example.cpp:856-935 and the Nova base options)
    int test_number = 1;
    int threads = 1;       // accepted for CLI compatibility; the GPU ignores it
    int levels = 1;
    T cfl = (T)0.1;
    int counts = 32;
    int first_frame = 0, last_frame = 1, current_frame = 0, output_number = 0;
    T frame_rate = 24;
    bool restart = false;
    int restart_frame = 0;
    bool staged = false;
    std::string output_directory = "output";
    std::string frame_title;

    // scene state filled by Initialize_Particles
    std::vector<T_Particle> particles;
    std::vector<T_Barrier> barriers;
    TV gravity;
    T flip = (T)0.9;
    T solver_tolerance = (T)1e-7;
    int solver_iterations = 10000;
    mpm_material material{};                  // table entry 0
    std::vector<mpm_material> extra_materials;  // entries 1.. (at most MPM_MAX_MATERIALS - 1)
    T min_dt = (T)1e-6, max_dt = (T)5e-3;
    int project_mode = 1, crop_ghost_x = 1;

    // fluid source (MPM_Example.h fluid_source; Add_Fluid_Source,
    // MPM_Example.cpp:977-1027).  The reference's call is commented out
    // (MPM_Driver.cpp:108); fluid_source_on turns it on.
    bool fluid_source_on = false, has_fluid_source = false;
    TV fluid_source_min, fluid_source_max;
    int fluid_material = 0;              // entry of the material table (an EOS material)
    int64_t fluid_capacity = 1 << 20;    // particle room kept for emitted particles
    uint64_t fluid_seed = 0, fluid_draws = 0;

    MPM_Example() {
        gravity(1) = (T)-2;  // MPM_Example.cpp:81
        material.hardening_max = FLT_MAX;
    }
    virtual ~MPM_Example() { if (handle) mpm_destroy(handle); }

    virtual void Initialize_Particles(int test_case) = 0;

    // ---- lifecycle --------------------------------------------------------
    // MPM_Example::Initialize (MPM_Example.cpp:152-161): scene hook, then the
    // particles and the scene description go to the device.
    void Initialize() {
        Initialize_Particles(test_number);
        Populate_Simulated_Particles();
        mpm_scene_desc desc = Scene_Description();
        const int64_t n = (int64_t)particles.size();
        const int64_t room = Emitting() ? fluid_capacity : 0;
        if (handle) mpm_destroy(handle);
        check(mpm_create(&desc, std::max<int64_t>(n + room, 1), &handle), "mpm_create");
        Upload();
        waiting = true;
        initialized = false;
    }
    bool Emitting() const { return fluid_source_on && has_fluid_source; }
    // MPM_Example::Add_Fluid_Source (MPM_Example.cpp:977-1027): ten EOS
    // particles, uniform in the source box, V = (0,-1), mass density 2 over
    // 5e-4 each in 2D and a tenth of the box in 3D, volume from the grid
    // (waiting).  Positions come from a counter-based generator (the
    // reference's Random_Numbers is Nova's).
    void Add_Fluid_Source() {
        if (!Emitting()) return;
        if (!initialized) check(mpm_initialize(handle), "mpm_initialize");
        initialized = true;
        waiting = false;
        constexpr int number_fluid_particles = 10;
        T measure = 1;
        for (int a = 0; a < d; ++a) measure *= fluid_source_max(a) - fluid_source_min(a);
        const T area_per_particle = d == 2 ? (T)5e-4 : measure / number_fluid_particles;
        const T mass_density = 2;
        std::vector<float> x(number_fluid_particles * d), v(number_fluid_particles * d, 0.0f),
            m(number_fluid_particles, (float)(mass_density * area_per_particle));
        std::vector<int32_t> mat(number_fluid_particles, fluid_material);
        for (int i = 0; i < number_fluid_particles; ++i) {
            for (int a = 0; a < d; ++a) {
                const double u = Unit_Sample(fluid_seed, fluid_draws, (uint64_t)a);
                x[i * d + a] = (float)(fluid_source_min(a) + (fluid_source_max(a) - fluid_source_min(a)) * u);
            }
            v[i * d + 1] = -1.0f;
            ++fluid_draws;
        }
        check(mpm_add_particles(handle, 0, number_fluid_particles, x.data(), v.data(), m.data(), mat.data()),
              "mpm_add_particles");
        emitted_since_download += number_fluid_particles;
        host_stale = true;
    }
    mpm_scene_desc Scene_Description() const {
        mpm_scene_desc desc{};
        desc.dim = d;
        desc.counts = counts;
        for (int a = 0; a < d; ++a) desc.gravity[a] = (float)gravity(a);
        desc.flip = (float)flip;
        desc.cfl = (float)cfl;
        desc.min_dt = (float)min_dt;
        desc.max_dt = (float)max_dt;
        desc.frame_rate = (float)frame_rate;
        desc.cg_tolerance = (float)solver_tolerance;
        desc.cg_max_iterations = solver_iterations;
        desc.cg_min_iterations = 0;
        desc.nplanes = (int)barriers.size();
        for (size_t i = 0; i < barriers.size() && i < MPM_MAX_PLANES; ++i) {
            for (int a = 0; a < d; ++a) {
                desc.planes[i].normal[a] = (float)barriers[i].normal(a);
                desc.planes[i].surface[a] = (float)barriers[i].surface(a);
            }
            desc.planes[i].mu = (float)barriers[i].mu;
            desc.planes[i].sticky = barriers[i].sticky ? 1 : 0;
        }
        desc.material = material;
        desc.nmaterials = 1 + (int)extra_materials.size();
        for (size_t k = 0; k < extra_materials.size() && k + 1 < MPM_MAX_MATERIALS; ++k)
            desc.extra_materials[k] = extra_materials[k];
        desc.crop_ghost_x = crop_ghost_x;
        desc.project_mode = project_mode;
        return desc;
    }
    // Populate_Simulated_Particles (MPM_Example.cpp:347-355): the device keeps
    // only live particles; this drops removed ones from the host copy.
    void Populate_Simulated_Particles() {
        std::vector<T_Particle> keep;
        keep.reserve(particles.size());
        for (auto& p : particles) if (p.valid) keep.push_back(p);
        particles.swap(keep);
    }

    // ---- stages (MPM_Driver.cpp:103-116) ------------------------------------
    // The device rebuilds its block-sparse grid every substep during binning,
    // so the reference's allocation and clearing stages have nothing to do.
    void Initialize_SPGrid() {}
    void Reset_Grid_Based_Variables() {}
    void Update_Particle_Weights() {}
    void Group_Particles() {}
    // Rasterize_Voxels (MPM_Example.cpp:291-313): binning builds the active
    // blocks.  In fused mode the first call (MPM_Driver::Initialize) runs the
    // whole initial grid build, P2G and volume estimate at once.
    void Rasterize_Voxels() {
        if (staged) {
            Stage(MPM_STAGE_BIN, 0);
        } else if (!initialized) {
            check(mpm_initialize(handle), "mpm_initialize");
            initialized = true;
            waiting = false;
            host_stale = true;  // volumes were set on the device
        }
    }
    void Rasterize() { if (staged) Stage(MPM_STAGE_P2G, 0); }
    void Process_Waiting_Particles() {
        if (staged && waiting) {
            Stage(MPM_STAGE_WAITING, 0);
            waiting = false;
            initialized = true;
            host_stale = true;
        }
    }
    void Update_Constitutive_Model_State() { if (staged) Stage(MPM_STAGE_CONSTITUTIVE, 0); }
    // Update_Particle_Velocities_And_Positions (MPM_Example.cpp:623-666),
    // including Apply_Force (:695-716).
    void Update_Particle_Velocities_And_Positions(const T dt) {
        if (staged) {
            Stage(MPM_STAGE_EXPLICIT_FORCE, dt);
            Stage(MPM_STAGE_COLLISION, dt);
            Stage(MPM_STAGE_CG, dt);
            Stage(MPM_STAGE_G2P, dt);
        } else {
            check(mpm_substep(handle, (float)dt, nullptr), "mpm_substep");
        }
        host_stale = true;
    }
    // Max_Particle_Velocity (MPM_Example.cpp:375-387), reduced on the device
    T Max_Particle_Velocity() const {
        float v = 0;
        check(mpm_max_particle_velocity(handle, &v), "mpm_max_particle_velocity");
        return (T)v;
    }
    T Time_At_Frame(int frame) const { return (T)frame / frame_rate; }

    // ---- output (Read_Write/Read_Write_MPM_Particles.h:14-47) -----------------
    // <dir>/<frame>/particles: size_t count, then X, V, mass, volume of every
    // valid particle in the reference's layout; <dir>/<frame>/state adds F and
    // Jp so a restart is lossless (the reference's restart drops them).
    void Write_Output_Files(int frame) {
        Download();
        const std::string dir = output_directory + "/" + std::to_string(frame);
        Make_Directory(output_directory);
        Make_Directory(dir);
        {
            std::ofstream f(dir + "/frame_title");
            f << frame_title;
        }
        std::ofstream f(dir + "/particles", std::ios::binary);
        const size_t count = particles.size();
        f.write((const char*)&count, sizeof(count));
        for (auto& p : particles) {
            f.write((const char*)p.X.v, sizeof(T) * d);
            f.write((const char*)p.V.v, sizeof(T) * d);
            f.write((const char*)&p.mass, sizeof(T));
            f.write((const char*)&p.volume, sizeof(T));
        }
        std::ofstream g(dir + "/state", std::ios::binary);
        g.write((const char*)&count, sizeof(count));
        for (auto& p : particles) {
            g.write((const char*)p.Fe, sizeof(T) * d * d);
            g.write((const char*)&p.Jp, sizeof(T));
            g.write((const char*)&p.material, sizeof(int));
            g.write((const char*)&p.density, sizeof(T));
        }
        Write_Hierarchy(frame);
    }
    // Grid hierarchy for the viewer (MPM_Example.cpp:1061-1062, read back by
    // plugin_Non_Ficks_Diffusion/MPM_Renderable.h:68-78 and 125-168): one
    // level of 4^3 (2D: 8^2) blocks of the last binning.
    //   common/hierarchy.struct3d|2d: int levels, unsigned elements per block,
    //     unsigned block size per axis;
    //   common/fine_grid: cell counts (int per axis), domain min and max (T per
    //     axis);
    //   <frame>/levels: the level count as text;
    //   <frame>/block_offsets: per level an unsigned block count, then each
    //     block's first cell index (int per axis, cells counted from 1);
    //   <frame>/flags: per block its unsigned flags, last axis fastest.
    void Write_Hierarchy(int frame) {
        const std::string common = output_directory + "/common";
        Make_Directory(common);
        {
            std::ofstream f(common + (d == 3 ? "/hierarchy.struct3d" : "/hierarchy.struct2d"), std::ios::binary);
            const int levels = 1;
            const unsigned B = d == 3 ? 4u : 8u, elements = 64u;  // 4^3 or 8^2
            f.write((const char*)&levels, sizeof(levels));
            f.write((const char*)&elements, sizeof(elements));
            for (int a = 0; a < d; ++a) f.write((const char*)&B, sizeof(B));
        }
        {
            std::ofstream f(common + "/fine_grid", std::ios::binary);
            for (int a = 0; a < d; ++a) f.write((const char*)&counts, sizeof(counts));
            const T lo = 0, hi = 1;
            for (int a = 0; a < d; ++a) f.write((const char*)&lo, sizeof(lo));
            for (int a = 0; a < d; ++a) f.write((const char*)&hi, sizeof(hi));
        }
        const std::string dir = output_directory + "/" + std::to_string(frame);
        {
            std::ofstream f(dir + "/levels");
            f << 1;
        }
        int32_t nblocks = 0;
        check(mpm_read_block_flags(handle, 0, nullptr, nullptr, &nblocks), "mpm_read_block_flags");
        std::vector<int32_t> origin((size_t)std::max(nblocks, 1) * d);
        std::vector<uint32_t> flags((size_t)std::max(nblocks, 1) * 64);
        check(mpm_read_block_flags(handle, nblocks, origin.data(), flags.data(), &nblocks), "mpm_read_block_flags");
        std::ofstream fo(dir + "/block_offsets", std::ios::binary);
        const unsigned nb = (unsigned)nblocks;
        fo.write((const char*)&nb, sizeof(nb));
        for (int b = 0; b < nblocks; ++b)
            for (int a = 0; a < d; ++a) {
                const int32_t index = origin[(size_t)b * d + a] + 1;
                fo.write((const char*)&index, sizeof(index));
            }
        std::ofstream ff(dir + "/flags", std::ios::binary);
        ff.write((const char*)flags.data(), (std::streamsize)((size_t)nblocks * 64 * sizeof(uint32_t)));
    }
    // Read_Output_Files (MPM_Example.cpp:1075-1079); restores F, Jp, the
    // material entry and the density too when the state file exists.
    void Read_Output_Files(int frame) {
        const std::string dir = output_directory + "/" + std::to_string(frame);
        std::ifstream f(dir + "/particles", std::ios::binary);
        if (!f) throw std::runtime_error("cannot read " + dir + "/particles");
        size_t count = 0;
        f.read((char*)&count, sizeof(count));
        particles.assign(count, T_Particle());
        for (auto& p : particles) {
            f.read((char*)p.X.v, sizeof(T) * d);
            f.read((char*)p.V.v, sizeof(T) * d);
            f.read((char*)&p.mass, sizeof(T));
            f.read((char*)&p.volume, sizeof(T));
        }
        std::ifstream g(dir + "/state", std::ios::binary);
        if (g) {
            size_t c2 = 0;
            g.read((char*)&c2, sizeof(c2));
            if (c2 == count)
                for (auto& p : particles) {
                    g.read((char*)p.Fe, sizeof(T) * d * d);
                    g.read((char*)&p.Jp, sizeof(T));
                    g.read((char*)&p.material, sizeof(int));
                    g.read((char*)&p.density, sizeof(T));
                }
        }
        mpm_scene_desc desc = Scene_Description();
        if (handle) mpm_destroy(handle);
        check(mpm_create(&desc, (int64_t)std::max<size_t>(count, 1), &handle), "mpm_create");
        Upload();
        waiting = false;
        // one grid build so the first substep starts from a binned state
        check(mpm_initialize(handle), "mpm_initialize");
        initialized = true;
    }

    mpm_handle Handle() const { return handle; }

  protected:
    mpm_handle handle = nullptr;
    bool waiting = false, initialized = false, host_stale = false;
    int64_t emitted_since_download = 0;

    // SplitMix64 (Steele, Lea, Flood 2014) over (seed, draw, axis): 24 bits in [0,1)
    static double Unit_Sample(uint64_t seed, uint64_t draw, uint64_t axis) {
        auto mix = [](uint64_t z) {
            z += 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        return (double)(mix(mix(mix(seed ^ 0xF1u) ^ draw) ^ axis) >> 40) / 16777216.0;
    }

    void Stage(int stage, T dt) { check(mpm_run_stage(handle, stage, (float)dt), "mpm_run_stage"); }

    bool Is_EOS(int entry) const {
        return entry == 0 ? material.eos != 0 : extra_materials.at((size_t)entry - 1).eos != 0;
    }
    // The device keeps an EOS particle's density in its Jp slot.
    void Upload() {
        const size_t n = particles.size();
        std::vector<float> x(n * d), v(n * d), F(n * d * d), Jp(n), m(n), vol(n);
        std::vector<int32_t> mat(n);
        bool have_volume = n > 0;
        for (size_t i = 0; i < n; ++i) {
            const T_Particle& p = particles[i];
            for (int a = 0; a < d; ++a) { x[i * d + a] = (float)p.X(a); v[i * d + a] = (float)p.V(a); }
            for (int k = 0; k < d * d; ++k) F[i * d * d + k] = (float)p.Fe[k];
            mat[i] = p.material;
            Jp[i] = (float)(Is_EOS(p.material) ? p.density : p.Jp);
            m[i] = (float)p.mass;
            vol[i] = (float)p.volume;
            have_volume = have_volume && p.volume > 0;
        }
        check(mpm_upload_particles_ex(handle, (int64_t)n, x.data(), v.data(), F.data(), Jp.data(), m.data(),
                                      have_volume ? vol.data() : nullptr, mat.data()),
              "mpm_upload_particles");
        host_stale = false;
    }
    void Download() {
        if (!host_stale) return;
        const size_t cap = particles.size() + (size_t)emitted_since_download;
        std::vector<float> x(cap * d), v(cap * d), F(cap * d * d), Jp(cap), m(cap), vol(cap);
        std::vector<int32_t> mat(cap);
        int64_t n = 0;
        check(mpm_download_particles_ex(handle, (int64_t)cap, x.data(), v.data(), F.data(), Jp.data(), m.data(),
                                        vol.data(), nullptr, mat.data(), &n),
              "mpm_download_particles");
        particles.assign((size_t)n, T_Particle());
        for (size_t i = 0; i < (size_t)n; ++i) {
            T_Particle& p = particles[i];
            for (int a = 0; a < d; ++a) { p.X(a) = x[i * d + a]; p.V(a) = v[i * d + a]; }
            for (int k = 0; k < d * d; ++k) p.Fe[k] = F[i * d * d + k];
            p.material = mat[i];
            if (Is_EOS(p.material)) p.density = Jp[i];
            else p.Jp = Jp[i];
            p.mass = m[i];
            p.volume = vol[i];
        }
        emitted_since_download = 0;
        host_stale = false;
    }
    static void Make_Directory(const std::string& path) {
        if (mkdir(path.c_str(), 0755) != 0 && errno != EEXIST)
            throw std::runtime_error("cannot create directory " + path);
    }
};

}  // namespace mpm_host
