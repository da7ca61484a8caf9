// cmag/render.hpp - ParticleRenderer of the reference demos
// (DamBreak.Demo/render_particles.h) without OpenGL: the same setters, and
// display() draws the point sprites into an RGBA8 image in device memory
// (cmag_renderer_* of cmag_hip.h) instead of a GL window.
//
// Differences a caller sees: setVertexBuffer and setColorBuffer take device
// pointers (a System's getCudaPosVBO() and getCudaColorVBO()) rather than GL
// buffer names; the camera the demo sets with glTranslatef / glRotatef is
// setCamera; image() and read() give the frame.
#pragma once

#include <cstdint>
#include <vector>

#include "cmag/systems.hpp"

class ParticleRenderer {
public:
    ParticleRenderer() { cmag_render_view_default(&view_); }
    ~ParticleRenderer() { cmag_renderer_destroy(r_); }
    ParticleRenderer(const ParticleRenderer&) = delete;
    ParticleRenderer& operator=(const ParticleRenderer&) = delete;

    void setPositions(const float* dev_pos4, int numParticles) { setVertexBuffer(dev_pos4, numParticles); }
    void setVertexBuffer(const void* dev_pos4, int numParticles) {
        pos_ = static_cast<const float*>(dev_pos4);
        n_ = numParticles;
    }
    void setColorBuffer(const void* dev_color4) { color_ = static_cast<const float*>(dev_color4); }
    // the reference stores pointSize without using it (its shader sizes the
    // sprites); the largest sprite side is setMaxPointSize here
    void setPointSize(float size) { point_size_ = size; }
    void setMaxPointSize(float pixels) { view_.max_point_size = pixels; }
    void setradius(float r) { view_.point_radius = r; }
    void setFOV(float fov) { view_.fov_deg = fov; }
    void setWindowSize(int w, int h) {
        if (r_ && (w != w_ || h != h_)) {
            cmag_renderer_destroy(r_);
            r_ = nullptr;
        }
        w_ = w;
        h_ = h;
    }
    void setCamera(float tx, float ty, float tz, float rot_x_deg, float rot_y_deg) {
        view_.translate[0] = tx;
        view_.translate[1] = ty;
        view_.translate[2] = tz;
        view_.rotate[0] = rot_x_deg;
        view_.rotate[1] = rot_y_deg;
    }
    void setClearColor(float r, float g, float b, float a) {
        view_.clear[0] = r;
        view_.clear[1] = g;
        view_.clear[2] = b;
        view_.clear[3] = a;
    }

    // one frame on stream (default: the null stream)
    void display(cmag_stream_t stream = nullptr) {
        if (!r_) cmag_cpp::check(cmag_renderer_create((uint32_t)w_, (uint32_t)h_, &r_), "ParticleRenderer");
        cmag_cpp::check(cmag_renderer_display(r_, &view_, pos_, color_, (uint32_t)n_, stream), "display");
    }
    const void* image() const { return r_ ? cmag_renderer_image(r_) : nullptr; }
    // the last frame as width * height RGBA8 pixels, top row first
    std::vector<uint8_t> read(cmag_stream_t stream = nullptr) {
        std::vector<uint8_t> out((size_t)w_ * (size_t)h_ * 4);
        if (r_) cmag_cpp::check(cmag_renderer_read(r_, out.data(), stream), "read");
        return out;
    }
    int width() const { return w_; }
    int height() const { return h_; }

private:
    cmag_render_view view_{};
    cmag_renderer* r_ = nullptr;
    const float* pos_ = nullptr;
    const float* color_ = nullptr;
    int n_ = 0;
    float point_size_ = 1.0f;
    int w_ = 1280, h_ = 1024;  // the demo's window (DamBreak.Demo/main.cpp:27)
};
