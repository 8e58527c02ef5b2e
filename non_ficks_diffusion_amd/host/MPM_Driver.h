// MPM_Driver on the GPU path: the reference's frame loop (MPM_Driver.h:13-36,
// MPM_Driver.cpp:18-116) unchanged in structure; it drives an MPM_Example
// whose stages run on the device.
#pragma once

#include <algorithm>
#include <cstdio>
#include <string>

#include "MPM_Example.h"

namespace mpm_host {

template <class T, int d>
class MPM_Driver {
  public:
    MPM_Example<T, d>& example;
    T time = 0;
    int substeps_total = 0;
    bool verbose = false;

    explicit MPM_Driver(MPM_Example<T, d>& e) : example(e) {}

    // MPM_Driver::Initialize (MPM_Driver.cpp:18-32)
    void Initialize() {
        time = (T)0;
        if (!example.restart) {
            example.Initialize();
            example.Reset_Grid_Based_Variables();
            example.Update_Particle_Weights();
            example.Group_Particles();
            example.Rasterize_Voxels();
            example.Rasterize();
            example.Process_Waiting_Particles();
        } else {
            example.Read_Output_Files(example.restart_frame);
            example.current_frame = example.restart_frame;
            example.output_number = example.restart_frame;
            time = example.Time_At_Frame(example.restart_frame);
        }
    }
    // MPM_Driver::Execute_Main_Program (MPM_Driver.cpp:45-51)
    void Execute_Main_Program() {
        Initialize();
        Simulate_To_Frame(example.last_frame);
    }
    // MPM_Driver::Advance_To_Target_Time (MPM_Driver.cpp:55-80)
    void Advance_To_Target_Time(const T target_time) {
        const T min_dt = example.min_dt, max_dt = example.max_dt, cfl = example.cfl;
        const T dx_min = (T)1 / (T)example.counts;
        bool done = false;
        example.Populate_Simulated_Particles();
        for (int substep = 1; !done; ++substep) {
            const T max_v = example.Max_Particle_Velocity();
            T dt = std::max(min_dt, std::min(max_dt, cfl * dx_min / std::max(max_v, (T)1e-2)));
            if (target_time - time < dt * (T)1.001) {
                dt = target_time - time;
                done = true;
            } else if (target_time - time < (T)2 * dt) {
                dt = (target_time - time) * (T).5;
            }
            Advance_Step(dt);
            if (verbose) std::printf("substep %d dt %g\n", substep, (double)dt);
            time += dt;
            ++substeps_total;
        }
    }
    // MPM_Driver::Simulate_To_Frame (MPM_Driver.cpp:84-99)
    void Simulate_To_Frame(const int target_frame) {
        example.frame_title = "Frame " + std::to_string(example.current_frame);
        if (!example.restart) example.Write_Output_Files(example.current_frame);
        while (example.current_frame < target_frame) {
            ++example.current_frame;
            Advance_To_Target_Time(example.Time_At_Frame(example.current_frame));
            example.frame_title = "Frame " + std::to_string(example.current_frame);
            example.Write_Output_Files(++example.output_number);
            if (verbose) std::printf("TIME = %g\n", (double)time);
        }
    }
    // MPM_Driver::Advance_Step (MPM_Driver.cpp:103-116), call for call
    void Advance_Step(const T dt) {
        example.Initialize_SPGrid();
        example.Reset_Grid_Based_Variables();
        example.Add_Fluid_Source();  // commented out in the reference; runs with -fluid_source
        example.Update_Particle_Weights();
        example.Group_Particles();
        example.Rasterize_Voxels();
        example.Rasterize();
        example.Process_Waiting_Particles();
        example.Update_Constitutive_Model_State();
        example.Update_Particle_Velocities_And_Positions(dt);
    }
};

}  // namespace mpm_host
