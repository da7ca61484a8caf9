// cmag/systems.hpp - the reference's C++ System classes over libcmag_hip.so.
//
// Header-only: each class owns one cmag_system handle of the C ABI
// (cmag_hip.h).  Constructor arguments, method names, enums and getters follow
//   DamBreakSystem        DamBreak.Core/fluidSystem.h:6-113   (FluidSystem alias)
//   PoiseuilleFlowSystem  Poiseuille.Core/poiseuilleFlowSystem.h:6-112
//   PeristalsisSystem     Peristalsis.Core/PeristalsisSystem.h:6-132
// so callers switch by changing the include and linking libcmag_hip.so.
// Differences: errors throw std::runtime_error instead of exiting the
// process; update(float dt) and getArray(...) are added; bUseOpenGL is
// accepted but no GL buffers are created (getCurrentReadBuffer and
// getColorBuffer return 0, positions stay in the device buffer
// getCudaPosVBO() returns).  Device pointers are borrowed and stay valid
// until the object is destroyed.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../cmag_hip.h"

namespace cmag_cpp {

inline void check(int rc, const char* what) {
    if (rc != CMAG_OK) throw std::runtime_error(std::string(what) + ": " + cmag_last_error());
}

// shared state and the getters all three reference headers carry
class SystemHandle {
public:
    SystemHandle(const SystemHandle&) = delete;
    SystemHandle& operator=(const SystemHandle&) = delete;
    ~SystemHandle() { cmag_system_destroy(s_); }

    int getNumParticles() const { return (int)cmag_system_num_particles(s_); }
    float getHalfWorldXSize() const { return p().grid_size[0] * p().particle_radius; }
    float getHalfWorldYSize() const { return p().grid_size[1] * p().particle_radius; }
    float getHalfWorldZSize() const { return p().grid_size[2] * p().particle_radius; }
    unsigned int getCurrentReadBuffer() const { return 0u; }
    unsigned int getColorBuffer() const { return 0u; }
    void* getCudaPosVBO() const { return ptr(CMAG_ARRAY_POSITION); }
    void* getCudaVelVBO() const { return ptr(CMAG_ARRAY_VELOCITY); }
    // the colour buffer the reference keeps for its GL colour VBO (float4 RGBA
    // per particle), made on first call (cmag_system_color_buffer)
    void* getCudaColorVBO() const {
        void* c = nullptr;
        check(cmag_system_color_buffer(s_, &c), "getCudaColorVBO");
        return c;
    }
    void* getCudaHash() const { return ptr(CMAG_ARRAY_HASH); }
    void* getCudaIndex() const { return ptr(CMAG_ARRAY_INDEX); }
    void* getCudaSortedPosition() const { return ptr(CMAG_ARRAY_SORTEDPOS); }
    uint3 getGridSize() const { return make_uint3(p().grid_size[0], p().grid_size[1], p().grid_size[2]); }
    float3 getWorldOrigin() const {
        return make_float3(p().world_origin[0], p().world_origin[1], p().world_origin[2]);
    }
    float3 getCellSize() const { return make_float3(p().cell_size[0], p().cell_size[1], p().cell_size[2]); }

    // neighbour-walk order (DamBreak: sub-cell order by default; off = the
    // reference summation order)
    void setSubcellWalk(bool on) { check(cmag_system_set_subcell_walk(s_, on ? 1 : 0), "setSubcellWalk"); }
    bool getSubcellWalk() const { return cmag_system_subcell_walk(s_) != 0; }

    // access to the C handle (streams, graphs, stage timing)
    cmag_system* handle() const { return s_; }
    cmag_params params() const { return p(); }

protected:
    explicit SystemHandle(cmag_system* s) : s_(s) {}
    void* ptr(int a) const { return cmag_system_device_ptr(s_, a); }
    cmag_params p() const {
        cmag_params q;
        check(cmag_system_get_params(s_, &q), "get_params");
        return q;
    }
    void step() { check(cmag_system_update(s_), "update"); }
    void step(float dt) { check(cmag_system_update_dt(s_, dt), "update"); }
    void set(int a, const float* d, int start, int count) {
        check(cmag_system_set_array(s_, a, d, start, count), "setArray");
    }
    void get(int a, float* d, int start, int count) {
        check(cmag_system_get_array(s_, a, d, start, count), "getArray");
    }
    cmag_system* s_ = nullptr;

    static cmag_system* make_dambreak(uint3 f, int bo, uint3 g, float r) {
        cmag_system* s = nullptr;
        check(cmag_dambreak_create(f.x, f.y, f.z, bo, g.x, g.y, g.z, r, &s), "DamBreakSystem");
        return s;
    }
    static cmag_system* make_poiseuille(uint3 f, int bo, uint3 g, float r) {
        cmag_system* s = nullptr;
        check(cmag_poiseuille_create(f.x, f.y, f.z, bo, g.x, g.y, g.z, r, &s), "PoiseuilleFlowSystem");
        return s;
    }
    static cmag_system* make_peristalsis(float dt, uint3 f, float a, float c, float cs, float3 gr, int bo,
                                         uint3 g, float r) {
        cmag_system* s = nullptr;
        check(cmag_peristalsis_create(dt, f.x, f.y, f.z, a, c, cs, gr.x, gr.y, gr.z, bo, g.x, g.y, g.z, r, &s),
              "PeristalsisSystem");
        return s;
    }
};

}  // namespace cmag_cpp

// DamBreak.Core/fluidSystem.h:6-113
class DamBreakSystem : public cmag_cpp::SystemHandle {
public:
    DamBreakSystem(uint3 fluidParticlesSize, int boundaryOffset, uint3 gridSize, float particleRadius,
                   bool bUseOpenGL = false)
        : SystemHandle(make_dambreak(fluidParticlesSize, boundaryOffset, gridSize, particleRadius)) {
        (void)bUseOpenGL;
    }

    enum ParticleArray { POSITION, VELOCITY, MEASURES, ACCELERATION, VELOCITYLEAPFROG };

    void update() { step(); }
    void update(float dt) { step(dt); }
    void reset() { cmag_cpp::check(cmag_system_reset(s_), "reset"); }
    void setArray(ParticleArray array, const float* data, int start, int count) { set(id(array), data, start, count); }
    void getArray(ParticleArray array, float* data, int start, int count) { get(id(array), data, start, count); }
    float getElapsedTime() const { return cmag_system_elapsed_time(s_); }
    void* getCudaMeasures() const { return ptr(CMAG_ARRAY_MEASURES); }
    void* getCudaAcceleration() const { return ptr(CMAG_ARRAY_ACCELERATION); }
    void changeRightBoundary() { cmag_cpp::check(cmag_system_change_right_boundary(s_), "changeRightBoundary"); }
    void removeRightBoundary() { cmag_cpp::check(cmag_system_remove_right_boundary(s_), "removeRightBoundary"); }
    float getParticleRadius() const { return p().particle_radius; }
    float3 getGravity() const { return make_float3(p().gravity[0], p().gravity[1], p().gravity[2]); }

private:
    static int id(ParticleArray a) {
        static const int ids[] = {CMAG_ARRAY_POSITION, CMAG_ARRAY_VELOCITY, CMAG_ARRAY_MEASURES,
                                  CMAG_ARRAY_ACCELERATION, CMAG_ARRAY_VELOCITYLEAPFROG};
        return ids[a];
    }
};

// the north-star name of the DamBreak system
using FluidSystem = DamBreakSystem;

// Poiseuille.Core/poiseuilleFlowSystem.h:6-112
class PoiseuilleFlowSystem : public cmag_cpp::SystemHandle {
public:
    PoiseuilleFlowSystem(uint3 fluidParticlesSize, int boundaryOffset, uint3 gridSize, float particleRadius,
                         bool bUseOpenGL = false)
        : SystemHandle(make_poiseuille(fluidParticlesSize, boundaryOffset, gridSize, particleRadius)) {
        (void)bUseOpenGL;
    }

    enum ParticleArray { POSITION, VELOCITY, MEASURES, ACCELERATION, VELOCITYLEAPFROG };

    void update() { step(); }
    void update(float dt) { step(dt); }
    void reset() { cmag_cpp::check(cmag_system_reset(s_), "reset"); }
    void setArray(ParticleArray array, const float* data, int start, int count) { set(id(array), data, start, count); }
    void getArray(ParticleArray array, float* data, int start, int count) { get(id(array), data, start, count); }
    float getElapsedTime() const { return cmag_system_elapsed_time(s_); }
    void* getCudaMeasures() const { return ptr(CMAG_ARRAY_MEASURES); }
    void* getCudaAcceleration() const { return ptr(CMAG_ARRAY_ACCELERATION); }
    void* getLeapFrogVelocity() const { return ptr(CMAG_ARRAY_VELOCITYLEAPFROG); }
    void changeGravity() { cmag_cpp::check(cmag_system_change_gravity(s_), "changeGravity"); }
    float getParticleRadius() const { return p().particle_radius; }

private:
    static int id(ParticleArray a) {
        static const int ids[] = {CMAG_ARRAY_POSITION, CMAG_ARRAY_VELOCITY, CMAG_ARRAY_MEASURES,
                                  CMAG_ARRAY_ACCELERATION, CMAG_ARRAY_VELOCITYLEAPFROG};
        return ids[a];
    }
};

// Peristalsis.Core/PeristalsisSystem.h:6-132
class PeristalsisSystem : public cmag_cpp::SystemHandle {
public:
    PeristalsisSystem(float deltaTime, uint3 fluid_size, float amplitude, float wave_speed, float soundspeed,
                      float3 gravity, int boundaryOffset, uint3 gridSize, float radius, bool bUseOpenGL = false)
        : SystemHandle(make_peristalsis(deltaTime, fluid_size, amplitude, wave_speed, soundspeed, gravity,
                                        boundaryOffset, gridSize, radius)) {
        (void)bUseOpenGL;
    }

    enum ParticleArray {
        POSITION, PREDICTEDPOSITION, VELOCITY, MEASURES, VISCOUSFORCE, PRESSUREFORCE, VELOCITYLEAPFROG
    };

    void Update() { step(); }
    void Update(float dt) { step(dt); }
    void Reset() { cmag_cpp::check(cmag_system_reset(s_), "Reset"); }
    void setArray(ParticleArray array, const float* data, int start, int count) { set(id(array), data, start, count); }
    void getArray(ParticleArray array, float* data, int start, int count) { get(id(array), data, start, count); }
    float GetElapsedTime() const { return cmag_system_elapsed_time(s_); }
    void* getSortedVelocity() const { return ptr(CMAG_ARRAY_SORTEDVEL); }
    void* getMeasures() const { return ptr(CMAG_ARRAY_MEASURES); }
    void* viscous_force() const { return ptr(CMAG_ARRAY_ACCELERATION); }
    void* pressure_force() const { return ptr(CMAG_ARRAY_PRESSUREFORCE); }
    void* getPredictedPos() const { return ptr(CMAG_ARRAY_PREDICTEDPOSITION); }
    void* getLeapFrogVelocity() const { return ptr(CMAG_ARRAY_VELOCITYLEAPFROG); }
    float getradius() const { return p().particle_radius; }
    // PeristalsisSystem::Coloring into the colour buffer of getCudaColorVBO()
    void Coloring() { cmag_cpp::check(cmag_system_coloring(s_), "Coloring"); }

private:
    static int id(ParticleArray a) {
        static const int ids[] = {CMAG_ARRAY_POSITION, CMAG_ARRAY_PREDICTEDPOSITION, CMAG_ARRAY_VELOCITY,
                                  CMAG_ARRAY_MEASURES, CMAG_ARRAY_ACCELERATION, CMAG_ARRAY_PRESSUREFORCE,
                                  CMAG_ARRAY_VELOCITYLEAPFROG};
        return ids[a];
    }
};
