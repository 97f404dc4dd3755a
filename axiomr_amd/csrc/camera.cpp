// camera.cpp - the reference's fly camera (src/camera.cpp, include/camera.hpp)
// as a caller-owned state record.  The reference polls a Window for the mouse
// and the W/S/A/D keys; here the same per-frame facts arrive as an explicit
// axr_camera_input, so a recorded input sequence replays to the same camera
// path every time and a whole path can be produced in one call for a batched
// draw.  Host only: a camera step is a few hundred flops.
#include <cmath>
#include <cstring>
#include "../../include/axiomr.h"
#include "axr_math.hpp"

using namespace axr;

namespace {

constexpr float kPitchLimitDeg = 89.0f;                 // camera.cpp:10
constexpr float kMinDelta = 0.0001f;                    // camera.cpp:11
constexpr float kMaxDelta = 0.5f;                       // camera.cpp:12
constexpr float kMaxRotation = 0.785398163397448309616f;  // radians(45), camera.cpp:131
constexpr float kRadPerDeg = 0.01745329251994329576923690768489f;
constexpr float kDegPerRad = 57.295779513082320876798154814105f;
constexpr float kNear = 0.1f, kFar = 100.0f;            // camera.hpp:49-50
constexpr float kSensitivity = 4.0f;                    // camera.hpp:53

float clamp_fov(float f) { return f < 1.0f ? 1.0f : (179.0f < f ? 179.0f : f); }   // camera.cpp:14-16
float clamp_aspect(float a) { return a < 0.1f ? 0.1f : a; }                        // camera.cpp:18-20

vec3 get3(const float* p) { return make3(p[0], p[1], p[2]); }
void put3(float* d, vec3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
quat get_q(const axr_camera* c) {
    return quat{c->orientation[0], c->orientation[1], c->orientation[2], c->orientation[3]};
}
void put_q(axr_camera* c, quat q) {
    c->orientation[0] = q.w; c->orientation[1] = q.x; c->orientation[2] = q.y; c->orientation[3] = q.z;
}

// reference camera.cpp:34-53: the rotation turning -z onto the view
// direction, built from an angle and an axis (the constructor's own form;
// setTarget uses the shortest-arc form instead).
quat look_orientation(vec3 pos, vec3 target) {
    vec3 fwd = make3(0.0f, 0.0f, -1.0f);
    vec3 dir = normalize(target - pos);
    float d = dot(fwd, dir);
    quat q;
    if (std::fabs(d + 1.0f) < 0.000001f) {
        q = angle_axis(3.14159265358979323846f, make3(0.0f, 1.0f, 0.0f));
    } else {
        vec3 axis = cross(fwd, dir);
        if (dot(axis, axis) < 0.000001f) q = quat{1.0f, 0.0f, 0.0f, 0.0f};
        else q = angle_axis(std::acos(d), normalize(axis));
    }
    return qnormalize(q);
}

// reference camera.cpp:57-62
void update_vectors(axr_camera* c) {
    quat q = get_q(c);
    put3(c->forward, normalize(qrotate(q, make3(0.0f, 0.0f, -1.0f))));
    put3(c->right, normalize(qrotate(q, make3(1.0f, 0.0f, 0.0f))));
    put3(c->up, normalize(qrotate(q, make3(0.0f, 1.0f, 0.0f))));
    c->view_dirty = 1;
}

mat4 load4(const float* m) { mat4 r; std::memcpy(&r, m, 64); return r; }

// reference camera.cpp:64-72
void update_view(axr_camera* c) {
    if (!c->view_dirty) return;
    mat4 view = mat4_cast(qinverse(get_q(c))) * translate(identity4(), -get3(c->position));
    mat4 vp = load4(c->proj) * view;
    std::memcpy(c->view, &view, 64);
    std::memcpy(c->view_proj, &vp, 64);
    c->view_dirty = 0;
}

// reference camera.cpp:74-80
void update_projection(axr_camera* c) {
    if (!c->proj_dirty) return;
    mat4 proj = perspective(c->fov * kRadPerDeg, c->aspect, c->znear, c->zfar);
    mat4 vp = proj * load4(c->view);
    std::memcpy(c->proj, &proj, 64);
    std::memcpy(c->view_proj, &vp, 64);
    c->proj_dirty = 0;
}

// reference camera.cpp:129-160
void apply_rotation(axr_camera* c, float dyaw, float dpitch) {
    dyaw = sclamp(dyaw, -kMaxRotation, kMaxRotation);
    dpitch = sclamp(dpitch, -kMaxRotation, kMaxRotation);
    quat q = get_q(c);
    quat yaw = angle_axis(dyaw, make3(0.0f, 1.0f, 0.0f));
    quat pitch = angle_axis(dpitch, get3(c->right));
    quat next = qmul(qmul(yaw, pitch), q);
    vec3 up = normalize(qrotate(next, make3(0.0f, 1.0f, 0.0f)));
    float up_dot = dot(up, make3(0.0f, 1.0f, 0.0f));
    if (up_dot < 0.0f) {
        q = qnormalize(qmul(yaw, q));               // pitching would turn the camera over
    } else {
        vec3 fwd = normalize(qrotate(next, make3(0.0f, 0.0f, -1.0f)));
        float pitch_deg = std::asin(sclamp(fwd.y, -1.0f, 1.0f)) * kDegPerRad;
        q = std::fabs(pitch_deg) <= kPitchLimitDeg ? qnormalize(next) : qnormalize(qmul(yaw, q));
    }
    put_q(c, q);
    update_vectors(c);
}

// reference camera.cpp:100-127 (mouse button 1 is the right button)
void handle_input(axr_camera* c, const axr_camera_input& in) {
    if (in.right_button) {
        c->delta_yaw = -(in.mouse_x - c->last_mouse[0]) * c->mouse_sensitivity;
        c->delta_pitch = (in.mouse_y - c->last_mouse[1]) * c->mouse_sensitivity;
        c->view_dirty = 1;
    } else {
        c->delta_yaw = 0.0f;
        c->delta_pitch = 0.0f;
    }
    c->last_mouse[0] = in.mouse_x;
    c->last_mouse[1] = in.mouse_y;
    vec3 v = make3(0.0f, 0.0f, 0.0f);
    vec3 f = get3(c->forward), r = get3(c->right);
    if (in.keys & AXR_KEY_W) v = v + f;
    if (in.keys & AXR_KEY_S) v = v - f;
    if (in.keys & AXR_KEY_A) v = v - r;
    if (in.keys & AXR_KEY_D) v = v + r;
    put3(c->velocity, v);
}

// reference camera.cpp:83-98.  Rotation runs only while BOTH mouse deltas
// are non-zero, exactly as the reference tests them (before the deltas are
// scaled, clamped to +-0.5 rad and zeroed below 1e-4 rad).
void update(axr_camera* c, float dt) {
    if (c->delta_pitch != 0.0f && c->delta_yaw != 0.0f) {
        float dyaw = (c->delta_yaw * dt) * kRadPerDeg;
        float dpitch = (c->delta_pitch * dt) * kRadPerDeg;
        dyaw = sclamp(dyaw, -kMaxDelta, kMaxDelta);
        dpitch = sclamp(dpitch, -kMaxDelta, kMaxDelta);
        if (std::fabs(dyaw) < kMinDelta) dyaw = 0.0f;
        if (std::fabs(dpitch) < kMinDelta) dpitch = 0.0f;
        apply_rotation(c, dyaw, dpitch);
    }
    vec3 v = get3(c->velocity);
    if (dot(v, v) > 0.0f) {
        put3(c->position, get3(c->position) + v * dt);
        c->view_dirty = 1;
    }
    update_view(c);
}

bool finite3(const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

}  // namespace

extern "C" {

int axr_camera_init(axr_camera* cam, const float position[3], const float target[3], float fov_deg, float aspect) {
    if (!cam || !position || !target || !finite3(position) || !finite3(target)) return AXR_ERR_INVALID;
    std::memset(cam, 0, sizeof(*cam));
    std::memcpy(cam->position, position, 12);
    cam->fov = clamp_fov(fov_deg);
    cam->aspect = clamp_aspect(aspect);
    cam->znear = kNear;
    cam->zfar = kFar;
    cam->mouse_sensitivity = kSensitivity;
    cam->viewport[2] = 800;                                  // camera.hpp:60-63
    cam->viewport[3] = 600;
    mat4 id = identity4();
    std::memcpy(cam->view, &id, 64);
    std::memcpy(cam->proj, &id, 64);
    std::memcpy(cam->view_proj, &id, 64);
    cam->proj_dirty = 1;
    put_q(cam, look_orientation(get3(position), get3(target)));
    update_vectors(cam);
    update_projection(cam);          // the view stays identity until the first update
    return AXR_OK;
}

int axr_camera_handle_input(axr_camera* cam, const axr_camera_input* in) {
    if (!cam || !in) return AXR_ERR_INVALID;
    handle_input(cam, *in);
    return AXR_OK;
}

int axr_camera_update(axr_camera* cam, float dt) {
    if (!cam) return AXR_ERR_INVALID;
    update(cam, dt);
    return AXR_OK;
}

int axr_camera_set_position(axr_camera* cam, const float position[3]) {
    if (!cam || !position) return AXR_ERR_INVALID;
    std::memcpy(cam->position, position, 12);
    cam->view_dirty = 1;
    return AXR_OK;
}

int axr_camera_set_target(axr_camera* cam, const float target[3]) {
    if (!cam || !target) return AXR_ERR_INVALID;
    vec3 dir = normalize(get3(target) - get3(cam->position));
    put_q(cam, rotation_between(make3(0.0f, 0.0f, -1.0f), dir));
    update_vectors(cam);
    return AXR_OK;
}

int axr_camera_set_fov(axr_camera* cam, float fov_deg) {
    if (!cam) return AXR_ERR_INVALID;
    cam->fov = clamp_fov(fov_deg);
    cam->proj_dirty = 1;
    return AXR_OK;
}

int axr_camera_set_aspect_ratio(axr_camera* cam, float aspect) {
    if (!cam) return AXR_ERR_INVALID;
    cam->aspect = clamp_aspect(aspect);
    cam->proj_dirty = 1;
    return AXR_OK;
}

int axr_camera_set_viewport(axr_camera* cam, int x, int y, int width, int height) {
    if (!cam) return AXR_ERR_INVALID;
    cam->viewport[0] = x; cam->viewport[1] = y; cam->viewport[2] = width; cam->viewport[3] = height;
    return AXR_OK;
}

int axr_camera_viewport_matrix(const axr_camera* cam, float* viewport16) {
    if (!cam || !viewport16) return AXR_ERR_INVALID;
    // reference camera.cpp:191-205
    float hw = cam->viewport[2] * 0.5f, hh = cam->viewport[3] * 0.5f;
    mat4 m = identity4();
    m.c[0].x = hw;
    m.c[1].y = -hh;
    m.c[2].z = 0.5f;
    m.c[3].x = cam->viewport[0] + hw;
    m.c[3].y = cam->viewport[1] + hh;
    m.c[3].z = 0.5f;
    std::memcpy(viewport16, &m, 64);
    return AXR_OK;
}

int axr_camera_path(axr_camera* cam, const axr_camera_input* inputs, const float* dts, uint64_t n,
                    float* view_proj16n, float* positions3n) {
    if (!cam || (n && (!inputs || !dts))) return AXR_ERR_INVALID;
    for (uint64_t i = 0; i < n; ++i) {
        handle_input(cam, inputs[i]);        // reference src/renderer.cpp:122-123
        update(cam, dts[i]);
        if (view_proj16n) std::memcpy(view_proj16n + 16 * i, cam->view_proj, 64);
        if (positions3n) std::memcpy(positions3n + 3 * i, cam->position, 12);
    }
    return AXR_OK;
}

int axr_camera_matrices(const float pos[3], const float target[3], float fov_deg, float aspect, float znear,
                        float zfar, int vp_w, int vp_h, float* view16, float* proj16, float* view_proj16,
                        float* viewport16) {
    // A camera constructed at pos looking at target, with the given planes
    // and viewport, after its first update(): the matrices a draw reads.
    axr_camera cam;
    int rc = axr_camera_init(&cam, pos, target, fov_deg, aspect);
    if (rc != AXR_OK) return rc;
    cam.znear = znear;
    cam.zfar = zfar;
    cam.proj_dirty = 1;
    update_projection(&cam);
    axr_camera_set_viewport(&cam, 0, 0, vp_w, vp_h);
    update(&cam, 0.0f);
    if (view16) std::memcpy(view16, cam.view, 64);
    if (proj16) std::memcpy(proj16, cam.proj, 64);
    if (view_proj16) std::memcpy(view_proj16, cam.view_proj, 64);
    if (viewport16) axr_camera_viewport_matrix(&cam, viewport16);
    return AXR_OK;
}

}  // extern "C"
