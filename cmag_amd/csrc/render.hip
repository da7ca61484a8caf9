// render.hip - the demos' particle colours and point-sprite renderer without
// OpenGL (gfx950).
//
// Colours: the float4 RGBA buffer the reference fills for its GL colour VBO,
//   DamBreakSystem::_initialize      fluidSystem.cpp:160-181 (colorRamp :97-114)
//   PoiseuilleFlowSystem::_initialize poiseuilleFlowSystem.cpp:148-169
//   PeristalsisSystem::Coloring       PeristalsisSystem.cpp:409-442
// Renderer: what DamBreak.Demo/render_particles.cpp:84-96 draws with the
// sphere pixel shader of shaders.cpp (point sprites of pointRadius *
// pointScale / eye distance pixels, each pixel inside the sprite's disc lit by
// max(0, L . N), depth test GL_LESS), drawn into an RGBA8 image in device
// memory.  Three launches per frame: clear the depth-and-winner buffer, splat
// every particle's disc with a 64-bit atomic minimum of (depth, particle)
// per pixel, resolve each pixel from its winner.  A point whose centre lies
// outside the view volume is dropped whole.
#include <algorithm>
#include <cmath>
#include <new>

#include "common.hpp"

namespace cmag {

// colorRamp, fluidSystem.cpp:97-114: seven colours, t in [0, 1)
__device__ __forceinline__ float4 color_ramp(float t) {
#pragma clang fp contract(off)
    const float c[7][3] = {{1.0f, 0.0f, 0.0f}, {1.0f, 0.5f, 0.0f}, {1.0f, 1.0f, 0.0f}, {0.0f, 1.0f, 0.0f},
                           {0.0f, 1.0f, 1.0f}, {0.0f, 0.0f, 1.0f}, {1.0f, 0.0f, 1.0f}};
    t = t * 6.0f;
    const int i = (int)t;
    const float u = t - floorf(t);
    float r[3];
    for (int k = 0; k < 3; ++k) r[k] = c[i][k] + u * (c[i + 1][k] - c[i][k]);
    return make_float4(r[0], r[1], r[2], 1.0f);
}

enum ColorRule { RULE_DAMBREAK = 0, RULE_POISEUILLE = 1, RULE_PERISTALSIS = 2 };

__global__ __launch_bounds__(256) void fill_colors_kernel(int rule, const float4* __restrict__ pos,
                                                          float4* __restrict__ color, uint32_t n, uint32_t fluid,
                                                          uint32_t type_one, uint32_t gx, float x_mark) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 c;
    if (rule == RULE_PERISTALSIS) {
        // walls black; the fluid left of x_mark red, the rest white (the
        // reference writes 255, which the framebuffer clamps to 1)
        const float4 p = pos[i];
        if (p.w != 0.0f) c = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        else if (p.x < x_mark) c = make_float4(255.0f, 0.0f, 0.0f, 1.0f);
        else c = make_float4(255.0f, 255.0f, 255.0f, 1.0f);
    } else {
        float t = 0.7f;
        if (rule == RULE_DAMBREAK) {
            if (i < type_one) t = 0.2f;
            if (i < fluid) t = 0.5f;
        } else {
            if (i < fluid) t = 0.5f;
            if (i % gx == 0u && i < fluid) t = 0.2f;
        }
        c = color_ramp(t);
    }
    color[i] = c;
}

int launch_fill_colors(const cmag_params& P, const float* pos4, float* color4, uint32_t n, hipStream_t s) {
    if (n == 0) return CMAG_OK;
    const uint32_t fluid = P.fluid_size[0] * P.fluid_size[1] * P.fluid_size[2];
    int rule = RULE_DAMBREAK;
    uint32_t type_one = fluid;
    if (P.system == CMAG_SYS_DAMBREAK) {
        // fluid + (gridSize.x - boundaryOffset + 1) + 2 (gridSize.y - boundaryOffset)
        const int64_t extra = ((int64_t)P.grid_size[0] - P.boundary_offset + 1) +
                              2 * ((int64_t)P.grid_size[1] - P.boundary_offset);
        type_one = (uint32_t)std::max<int64_t>(0, (int64_t)fluid + extra);
    } else if (P.system == CMAG_SYS_POISEUILLE) {
        rule = RULE_POISEUILLE;
    } else {
        rule = RULE_PERISTALSIS;
    }
    const float x_mark = P.world_origin[0] + 4.0f * P.particle_radius;
    hipLaunchKernelGGL(fill_colors_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, rule,
                       reinterpret_cast<const float4*>(pos4), reinterpret_cast<float4*>(color4), n, fluid, type_one,
                       P.grid_size[0] ? P.grid_size[0] : 1u, x_mark);
    return check_launch("fill_colors_kernel");
}

// ---- renderer ---------------------------------------------------------------
// Eye space = M p (M = T(translate) Rx(rotate[0]) Ry(rotate[1]), as the demo's
// glTranslatef / glRotatef sequence builds it); clip space = gluPerspective.
struct View {
    float m[12];        // 3 x 4 row-major modelview
    float fx, fy;       // f / aspect, f (f = 1 / tan(fovy / 2))
    float za, zb;       // clip z = za z_e + zb
    float point_scale;  // window_h / tan(fov / 2)
    float radius;       // pointRadius (world units)
    float max_size;     // largest sprite side in pixels
    int w, h;
};

struct Sprite {
    float xw, yw, zw, size;  // window centre (origin bottom-left), depth in [0, 1), side in pixels
    bool visible;
};

__device__ __forceinline__ Sprite project(const View& v, const float4 p) {
#pragma clang fp contract(off)
    Sprite s;
    const float xe = v.m[0] * p.x + v.m[1] * p.y + v.m[2] * p.z + v.m[3];
    const float ye = v.m[4] * p.x + v.m[5] * p.y + v.m[6] * p.z + v.m[7];
    const float ze = v.m[8] * p.x + v.m[9] * p.y + v.m[10] * p.z + v.m[11];
    const float cw = -ze;
    const float cx = v.fx * xe, cy = v.fy * ye, cz = v.za * ze + v.zb;
    s.visible = cw > 0.0f && fabsf(cx) <= cw && fabsf(cy) <= cw && fabsf(cz) <= cw;
    const float inv_w = 1.0f / cw;
    s.xw = (cx * inv_w + 1.0f) * 0.5f * (float)v.w;
    s.yw = (cy * inv_w + 1.0f) * 0.5f * (float)v.h;
    s.zw = (cz * inv_w + 1.0f) * 0.5f;
    const float dist = sqrtf(xe * xe + ye * ye + ze * ze);
    s.size = fminf(fmaxf(v.radius * (v.point_scale / dist), 1.0f), v.max_size);
    s.visible = s.visible && s.zw < 1.0f;
    return s;
}

// the sprite's normal at the centre of pixel (px, py): N.xy from the point
// coordinate (s, t) with t from the top, N.xy = (2s - 1, 1 - 2t); inside the
// disc when |N.xy| <= 1
__device__ __forceinline__ float2 sprite_xy(const Sprite& s, int px, int py) {
#pragma clang fp contract(off)
    const float inv = 1.0f / s.size;
    return make_float2(2.0f * (((float)px + 0.5f - s.xw) * inv), 2.0f * (((float)py + 0.5f - s.yw) * inv));
}

__global__ __launch_bounds__(256) void clear_depth_kernel(unsigned long long* __restrict__ zbuf, uint32_t npix) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npix) zbuf[p] = ~0ull;
}

__global__ __launch_bounds__(256) void splat_kernel(const View v, const float4* __restrict__ pos, uint32_t n,
                                                    unsigned long long* __restrict__ zbuf) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Sprite s = project(v, pos[i]);
    if (!s.visible) return;
    // pixels whose centres lie in the sprite's square [x - size/2, x + size/2)
    const float half = 0.5f * s.size;
    const int x0 = max(0, (int)ceilf(s.xw - half - 0.5f)), x1 = min(v.w - 1, (int)ceilf(s.xw + half - 0.5f) - 1);
    const int y0 = max(0, (int)ceilf(s.yw - half - 0.5f)), y1 = min(v.h - 1, (int)ceilf(s.yw + half - 0.5f) - 1);
    const unsigned long long key = ((unsigned long long)__float_as_uint(s.zw) << 32) | i;
    for (int py = y0; py <= y1; ++py)
        for (int px = x0; px <= x1; ++px) {
            const float2 nxy = sprite_xy(s, px, py);
            if (nxy.x * nxy.x + nxy.y * nxy.y > 1.0f) continue;
            atomicMin(&zbuf[(size_t)py * (size_t)v.w + (size_t)px], key);
        }
}

// row r of the image is window row h - 1 - r (the image is stored top row first)
__global__ __launch_bounds__(256) void resolve_kernel(const View v, const float4* __restrict__ pos,
                                                      const float4* __restrict__ color,
                                                      const unsigned long long* __restrict__ zbuf, float4 clear,
                                                      uchar4* __restrict__ image) {
#pragma clang fp contract(off)
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)(v.w * v.h)) return;
    const int px = (int)(p % (uint32_t)v.w), py = (int)(p / (uint32_t)v.w);
    const unsigned long long key = zbuf[p];
    float4 c = clear;
    if (key != ~0ull) {
        const uint32_t i = (uint32_t)(key & 0xffffffffull);
        const Sprite s = project(v, pos[i]);
        const float2 nxy = sprite_xy(s, px, py);
        const float mag = nxy.x * nxy.x + nxy.y * nxy.y;
        const float nz = sqrtf(fmaxf(1.0f - mag, 0.0f));
        const float diffuse = fmaxf(0.0f, 0.577f * nxy.x + 0.577f * nxy.y + 0.577f * nz);
        const float4 col = color ? color[i] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        c = make_float4(col.x * diffuse, col.y * diffuse, col.z * diffuse, col.w * diffuse);
    }
    auto u8 = [](float x) { return (unsigned char)(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f + 0.5f); };
    image[(size_t)(v.h - 1 - py) * (size_t)v.w + (size_t)px] = make_uchar4(u8(c.x), u8(c.y), u8(c.z), u8(c.w));
}

}  // namespace cmag

using namespace cmag;

struct cmag_renderer {
    uint32_t w = 0, h = 0;
    unsigned long long* zbuf = nullptr;
    uchar4* image = nullptr;
    ~cmag_renderer() {
        if (zbuf) (void)hipFree(zbuf);
        if (image) (void)hipFree(image);
    }
};

namespace {

// the demo's view: M = T(t) Rx(a) Ry(b); gluPerspective(fovy, w / h, near, far)
View make_view(const cmag_render_view& rv, uint32_t w, uint32_t h) {
    const double pi = 3.14159265358979323846;
    const double a = rv.rotate[0] * pi / 180.0, b = rv.rotate[1] * pi / 180.0;
    const double ca = std::cos(a), sa = std::sin(a), cb = std::cos(b), sb = std::sin(b);
    // Rx(a) Ry(b): rows of the rotation
    const double R[3][3] = {{cb, 0.0, sb}, {sa * sb, ca, -sa * cb}, {-ca * sb, sa, ca * cb}};
    View v{};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) v.m[4 * r + c] = (float)R[r][c];
        v.m[4 * r + 3] = rv.translate[r];
    }
    const double f = 1.0 / std::tan(0.5 * rv.fov_deg * pi / 180.0);
    const double aspect = (double)w / (double)h;
    v.fx = (float)(f / aspect);
    v.fy = (float)f;
    const double zn = rv.znear, zf = rv.zfar;
    v.za = (float)((zf + zn) / (zn - zf));
    v.zb = (float)(2.0 * zf * zn / (zn - zf));
    v.point_scale = (float)((double)h / std::tan(0.5 * rv.fov_deg * pi / 180.0));
    v.radius = rv.point_radius;
    v.max_size = rv.max_point_size > 0.0f ? std::min(std::max(rv.max_point_size, 1.0f), 256.0f) : 64.0f;
    v.w = (int)w;
    v.h = (int)h;
    return v;
}

}  // namespace

extern "C" {

int cmag_fill_colors(const cmag_params* p, const float* pos4, float* color4, uint32_t n, cmag_stream_t stream) {
    if (!p || (n && (!color4 || !pos4))) return fail(CMAG_E_INVALID, "cmag_fill_colors: null pointer");
    return launch_fill_colors(*p, pos4, color4, n, as_stream(stream));
}

void cmag_render_view_default(cmag_render_view* v) {
    if (!v) return;
    *v = cmag_render_view{};
    v->translate[0] = 0.5f;
    v->translate[1] = 0.0f;
    v->translate[2] = -2.5f;
    v->fov_deg = 60.0f;
    v->znear = 0.001f;
    v->zfar = 100.0f;
    v->point_radius = 0.125f * 0.5f;
    v->max_point_size = 64.0f;
    for (int k = 0; k < 4; ++k) v->clear[k] = 1.0f;
}

int cmag_renderer_create(uint32_t width, uint32_t height, cmag_renderer** out) {
    if (!out || width == 0 || height == 0 || (uint64_t)width * height > (1ull << 28))
        return fail(CMAG_E_INVALID, "cmag_renderer_create: image size outside [1, 2^28] pixels");
    *out = nullptr;
    cmag_renderer* r = new (std::nothrow) cmag_renderer;
    if (!r) return fail(CMAG_E_INVALID, "cmag_renderer_create: out of host memory");
    r->w = width;
    r->h = height;
    const size_t npix = (size_t)width * height;
    hipError_t e = hipMalloc((void**)&r->zbuf, sizeof(unsigned long long) * npix);
    if (e == hipSuccess) e = hipMalloc((void**)&r->image, sizeof(uchar4) * npix);
    if (e != hipSuccess) {
        delete r;
        return hip_fail(e, "hipMalloc (renderer buffers)");
    }
    *out = r;
    return CMAG_OK;
}

int cmag_renderer_destroy(cmag_renderer* r) {
    delete r;
    return CMAG_OK;
}

int cmag_renderer_display(cmag_renderer* r, const cmag_render_view* view, const float* pos4, const float* color4,
                          uint32_t n, cmag_stream_t stream) {
    if (!r || !view || (n && !pos4)) return fail(CMAG_E_INVALID, "cmag_renderer_display: null pointer");
    if (!(view->point_radius > 0.0f) || !(view->znear > 0.0f) || !(view->zfar > view->znear) ||
        !(view->fov_deg > 0.0f && view->fov_deg < 180.0f))
        return fail(CMAG_E_INVALID, "cmag_renderer_display: view needs radius > 0, 0 < near < far, 0 < fov < 180");
    hipStream_t s = as_stream(stream);
    const View v = make_view(*view, r->w, r->h);
    const uint32_t npix = r->w * r->h;
    hipLaunchKernelGGL(clear_depth_kernel, dim3(div_up(npix, 256)), dim3(256), 0, s, r->zbuf, npix);
    CMAG_TRY(check_launch("clear_depth_kernel"));
    if (n) {
        hipLaunchKernelGGL(splat_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, v,
                           reinterpret_cast<const float4*>(pos4), n, r->zbuf);
        CMAG_TRY(check_launch("splat_kernel"));
    }
    const float4 clear = make_float4(view->clear[0], view->clear[1], view->clear[2], view->clear[3]);
    hipLaunchKernelGGL(resolve_kernel, dim3(div_up(npix, 256)), dim3(256), 0, s, v,
                       reinterpret_cast<const float4*>(pos4), reinterpret_cast<const float4*>(color4), r->zbuf,
                       clear, r->image);
    return check_launch("resolve_kernel");
}

void* cmag_renderer_image(cmag_renderer* r) { return r ? (void*)r->image : nullptr; }

int cmag_renderer_read(cmag_renderer* r, uint8_t* rgba, cmag_stream_t stream) {
    if (!r || !rgba) return fail(CMAG_E_INVALID, "cmag_renderer_read: null pointer");
    hipStream_t s = as_stream(stream);
    CMAG_HIP_TRY(hipMemcpyAsync(rgba, r->image, sizeof(uchar4) * (size_t)r->w * r->h, hipMemcpyDeviceToHost, s));
    CMAG_HIP_TRY(hipStreamSynchronize(s));
    return CMAG_OK;
}

}  // extern "C"
