// Standard_Tests on the GPU path: the reference's scene plugin
// (mpm_3d/Standard_Tests/Standard_Tests.h:16-154,
// mpm_2d/Standard_Tests/Standard_Tests.h) as an MPM_Example subclass whose
// Initialize_Particles fills `particles` and `barriers`.  The particle sets
// come from the library's scene generator (mpm_scene_preset /
// mpm_scene_generate), so the executable, the Python tests and the
// benchmark run the same scenes.
//
// test 1  (2D)   one plastic particle moving right
// test 15 (2D)   a plastic square thrown left
// test 16 (2D)   a snowball falling on a softer snow layer (two materials)
// test 17 (2D)   a snowball thrown at a wall above the ground
// test 18        two plastic balls (E 30, nu 0.4) thrown at each other,
//                ground y = 0.1 (3D and 2D test 18 of the reference)
// test 19        an EOS fluid ball falling into a box (3D and 2D)
// test 20 (2D)   a fluid source: no particles, the box of test 19 and a
//                source region [0.45,0.55]^2 (emits with -fluid_source)
// test 100       the benchmark scenes: 3D two snowballs head-on,
//                2D snow disk dropped on the ground
// test 101       3D snow slab on a 30-degree incline (avalanche)
#pragma once

#include <string>

#include "MPM_Example.h"

namespace mpm_host {

template <class T, int d>
class Standard_Tests : public MPM_Example<T, d> {
  public:
    using Base = MPM_Example<T, d>;
    using typename Base::TV;
    using typename Base::T_Particle;
    using typename Base::T_Barrier;
    uint64_t seed = 0;

    static std::string Preset_Name(int test_case) {
        if (test_case == 18) return d == 3 ? "ref3d_test18" : "ref2d_test18";
        if (test_case == 19) return d == 3 ? "ref3d_test19" : "ref2d_test19";
        if (test_case == 20 && d == 2) return "ref2d_test20";
        if (d == 2 && (test_case == 1 || test_case == 15 || test_case == 16 || test_case == 17))
            return "ref2d_test" + std::to_string(test_case);
        if (test_case == 100) return d == 3 ? "two_balls" : "disk2d";
        if (test_case == 101 && d == 3) return "avalanche";
        return "";
    }

    void Parse_Options() {
        this->output_directory = std::string(d == 3 ? "mpm_3d_" : "mpm_2d_") + std::to_string(this->test_number);
    }

    void Initialize_Particles(int test_case) override {
        const std::string name = Preset_Name(test_case);
        if (name.empty()) throw std::runtime_error("unknown test number " + std::to_string(test_case));
        mpm_scene_desc desc{};
        int64_t n = 0;
        check(mpm_scene_preset(name.c_str(), this->counts, seed, &desc, &n), "mpm_scene_preset");
        if (desc.dim != d) throw std::runtime_error("test " + std::to_string(test_case) + " is not a " +
                                                    std::to_string(d) + "D scene");
        std::vector<float> x((size_t)n * d), v((size_t)n * d), m((size_t)n);
        int64_t got = 0;
        check(mpm_scene_generate(name.c_str(), this->counts, seed, n, x.data(), v.data(), m.data(), &got),
              "mpm_scene_generate");
        std::vector<int32_t> mat((size_t)n);
        check(mpm_scene_materials(name.c_str(), this->counts, seed, 0, n, mat.data(), &got), "mpm_scene_materials");
        this->material = desc.material;
        this->extra_materials.assign(desc.extra_materials,
                                     desc.extra_materials + std::max(0, (int)desc.nmaterials - 1));
        this->particles.assign((size_t)got, T_Particle());
        for (size_t i = 0; i < (size_t)got; ++i) {
            T_Particle& p = this->particles[i];
            for (int a = 0; a < d; ++a) {
                p.X(a) = x[i * d + a];
                p.V(a) = v[i * d + a];
            }
            p.mass = m[i];
            p.material = mat[i];
            const mpm_material& pm = mat[i] == 0 ? desc.material : desc.extra_materials[mat[i] - 1];
            p.density = pm.density;
        }
        this->barriers.clear();
        for (int k = 0; k < desc.nplanes; ++k) {
            TV n_, s_;
            for (int a = 0; a < d; ++a) { n_(a) = desc.planes[k].normal[a]; s_(a) = desc.planes[k].surface[a]; }
            this->barriers.push_back(T_Barrier(desc.planes[k].mu, n_, s_, desc.planes[k].sticky != 0));
        }
        for (int a = 0; a < d; ++a) this->gravity(a) = desc.gravity[a];
        this->flip = desc.flip;
        this->solver_tolerance = desc.cg_tolerance;
        this->solver_iterations = desc.cg_max_iterations;
        this->min_dt = desc.min_dt;
        this->max_dt = desc.max_dt;
        this->frame_rate = desc.frame_rate;
        this->project_mode = desc.project_mode;
        this->crop_ghost_x = desc.crop_ghost_x;
        this->has_fluid_source = test_case == 20;
        if (this->has_fluid_source) {
            // mpm_2d/Standard_Tests/Standard_Tests.h:290-291
            for (int a = 0; a < d; ++a) {
                this->fluid_source_min(a) = (T)0.45;
                this->fluid_source_max(a) = (T)0.55;
            }
            this->fluid_material = 0;
            this->fluid_seed = seed;
        }
    }
};

}  // namespace mpm_host
