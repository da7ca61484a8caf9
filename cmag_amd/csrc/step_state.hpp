// step_state.hpp - the per-system time state that PeristalsisSystem::Update
// keeps on the host (PeristalsisSystem.cpp:330-343), as a record that both the
// device (one thread per system, in front of each step) and the host (a mirror
// for queries) advance with the same function, so a step needs no host
// decision and a whole step, boundary configuration included, fits in a HIP
// graph.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cmag_params.h"

namespace cmag {

struct StepState {
    float elapsed;
    float time_shift, time_relax, wave_height;
    float wave_height_inc;   // delta_time * 10^-3, computed on the host as Update() does
    float t_w;               // wave time of the current step
    float configure_height;  // wall height of this step's boundary configuration
    int32_t is_config;       // boundary-configuration phase (after this step's switch)
    int32_t configure_now;   // configure the walls in front of this step's hash
    int32_t pad[3];
};
static_assert(sizeof(StepState) == 48, "StepState layout");

// a fresh system's state (PeristalsisSystem::Reset, PeristalsisSystem.cpp:213-237)
inline StepState initial_step_state(const cmag_params& P, float elapsed) {
    StepState s{};
    s.elapsed = elapsed;
    s.time_relax = 1000 * P.delta_time;
    s.wave_height_inc = P.delta_time * powf(10.0f, -3.0f);
    s.is_config = P.system == CMAG_SYS_PERISTALSIS && P.is_boundary_configuration ? 1 : 0;
    return s;
}

// one Update()'s state machine: Peristalsis shifts the wave clock and raises
// the walls by wave_height_inc until they reach the amplitude, then relaxes
// for 1000 steps and leaves the configuration phase; every system advances
// its clock by delta_time
__host__ __device__ inline void advance_step_state(const cmag_params& P, StepState& s) {
#pragma clang fp contract(off)
    s.configure_now = 0;
    s.t_w = 0.0f;
    if (P.system == CMAG_SYS_PERISTALSIS) {
        if (s.is_config) {
            s.time_shift += P.delta_time;
            if (s.wave_height < P.amplitude) {
                s.configure_now = 1;
                s.configure_height = s.wave_height;
                s.wave_height += s.wave_height_inc;
            } else {
                s.time_relax -= P.delta_time;
                if (s.time_relax < 0.0f) s.is_config = !s.is_config;
            }
        }
        s.t_w = s.is_config ? 0.0f : s.elapsed - s.time_shift;
    }
    s.elapsed += P.delta_time;
}

}  // namespace cmag
