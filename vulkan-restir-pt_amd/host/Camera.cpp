// Host camera: src/Camera.cpp restated without glm.
#include "Camera.h"
#include <cmath>
#include <cstring>

namespace {
const float kDeg = 0.017453292519943295f;
Vec3 add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
Vec3 sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
Vec3 scale(Vec3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
Vec3 normalize(Vec3 a) { return scale(a, 1.0f / std::sqrt(dot(a, a))); }
float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

Mat4 identity() { Mat4 r{}; r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f; return r; }
Mat4 mul(const Mat4& a, const Mat4& b) {
    Mat4 r{};
    for (int c = 0; c < 4; c++)
        for (int row = 0; row < 4; row++) {
            float s = 0.0f;
            for (int k = 0; k < 4; k++) s += a.m[k * 4 + row] * b.m[c * 4 + k];
            r.m[c * 4 + row] = s;
        }
    return r;
}
// glm::rotate(mat4(1), angle, axis)
Mat4 rotation(float angle, Vec3 axis) {
    float c = std::cos(angle), s = std::sin(angle);
    Vec3 a = normalize(axis);
    Vec3 t = scale(a, 1.0f - c);
    Mat4 r = identity();
    r.m[0] = c + t.x * a.x;        r.m[1] = t.x * a.y + s * a.z;  r.m[2] = t.x * a.z - s * a.y;
    r.m[4] = t.y * a.x - s * a.z;  r.m[5] = c + t.y * a.y;        r.m[6] = t.y * a.z + s * a.x;
    r.m[8] = t.z * a.x + s * a.y;  r.m[9] = t.z * a.y - s * a.x;  r.m[10] = c + t.z * a.z;
    return r;
}
// glm::lookAt (right handed)
Mat4 lookAtRH(Vec3 eye, Vec3 center, Vec3 up) {
    Vec3 f = normalize(sub(center, eye));
    Vec3 s = normalize(cross(f, up));
    Vec3 u = cross(s, f);
    Mat4 r = identity();
    r.m[0] = s.x; r.m[4] = s.y; r.m[8] = s.z;
    r.m[1] = u.x; r.m[5] = u.y; r.m[9] = u.z;
    r.m[2] = -f.x; r.m[6] = -f.y; r.m[10] = -f.z;
    r.m[12] = -dot(s, eye); r.m[13] = -dot(u, eye); r.m[14] = dot(f, eye);
    return r;
}
// glm::perspective (right handed, depth -1..1)
Mat4 perspectiveRH(float fovy, float aspect, float zNear, float zFar) {
    float t = std::tan(fovy / 2.0f);
    Mat4 r{};
    r.m[0] = 1.0f / (aspect * t);
    r.m[5] = 1.0f / t;
    r.m[10] = -(zFar + zNear) / (zFar - zNear);
    r.m[11] = -1.0f;
    r.m[14] = -(2.0f * zFar * zNear) / (zFar - zNear);
    return r;
}
}  // namespace

Camera::Camera(Vec3 pos, Vec3 angle) : mPos(pos), mAngle(angle) { update(); }

void Camera::move(Vec3 v) { mPos = add(mPos, v); update(); }
void Camera::rotate(Vec3 a) {
    mAngle = add(mAngle, a);
    mAngle.y = clampf(mAngle.y, -89.0f, 89.0f);
    update();
}
void Camera::setFOV(float fov) { mFOV = clampf(fov, 0.1f, 90.0f); update(); }
void Camera::lookAt(Vec3 focus) { setDir(sub(focus, mPos)); }
void Camera::setDir(Vec3 dir) {
    dir = normalize(dir);
    float lenXY = std::sqrt(dir.x * dir.x + dir.y * dir.y);
    mAngle.y = std::asin(dir.z) / kDeg;
    mAngle.x = std::asin(dir.y / lenXY) / kDeg - 90.0f;
    if (dir.x < 0) mAngle.x += 360.0f;
    update();
}
void Camera::setPos(Vec3 p) { mPos = p; update(); }
void Camera::setAngle(Vec3 a) { mAngle = a; update(); }
void Camera::setFilmSize(uint32_t w, uint32_t h) { mFilmSize[0] = w; mFilmSize[1] = h; update(); }
void Camera::setPlanes(float n, float f) { mNear = n; mFar = f; update(); }

void Camera::nextFrame(uint32_t seed) {
    mSeed = seed;
    mLastProjView = mProjView;
    mFrameIndex++;
}
void Camera::setClearFlag() { mFrameIndex = 0x80000000u; }

// src/Camera.cpp:78-98
void Camera::update() {
    float x = std::sin(mAngle.x * kDeg) * std::cos(mAngle.y * kDeg);
    float y = std::cos(mAngle.x * kDeg) * std::cos(mAngle.y * kDeg);
    float z = std::sin(mAngle.y * kDeg);
    mFront = normalize({x, y, z});
    mRight = normalize(cross(mFront, {0.0f, 0.0f, 1.0f}));
    Mat4 rot = rotation(mAngle.z, mFront);
    Vec3 r = mRight;
    Vec3 rr = {rot.m[0] * r.x + rot.m[4] * r.y + rot.m[8] * r.z + rot.m[12],
               rot.m[1] * r.x + rot.m[5] * r.y + rot.m[9] * r.z + rot.m[13],
               rot.m[2] * r.x + rot.m[6] * r.y + rot.m[10] * r.z + rot.m[14]};
    mRight = normalize(rr);
    mUp = normalize(cross(mRight, mFront));
    Vec3 at = add(mPos, {x, y, z});
    mViewMatrix = lookAtRH(mPos, at, mUp);
    float aspect = (float)mFilmSize[0] / (float)mFilmSize[1];
    mProjMatrix = perspectiveRH(mFOV * kDeg, aspect, mNear, mFar);
    mProjMatrix.m[5] *= -1.0f;
    mProjView = mul(mProjMatrix, mViewMatrix);
    mFrameIndex = 0;
}

void Camera::toDevice(rt_camera& o) const {
    std::memset(&o, 0, sizeof(o));
    std::memcpy(o.view, mViewMatrix.m, 64);
    std::memcpy(o.proj, mProjMatrix.m, 64);
    std::memcpy(o.projView, mProjView.m, 64);
    std::memcpy(o.lastProjView, mLastProjView.m, 64);
    o.pos[0] = mPos.x; o.pos[1] = mPos.y; o.pos[2] = mPos.z; o.FOV = mFOV;
    o.angle[0] = mAngle.x; o.angle[1] = mAngle.y; o.angle[2] = mAngle.z; o.nearZ = mNear;
    o.front[0] = mFront.x; o.front[1] = mFront.y; o.front[2] = mFront.z; o.farZ = mFar;
    o.right[0] = mRight.x; o.right[1] = mRight.y; o.right[2] = mRight.z; o.lensRadius = mLensRadius;
    o.up[0] = mUp.x; o.up[1] = mUp.y; o.up[2] = mUp.z; o.focalDist = mFocalDist;
    o.filmSize[0] = mFilmSize[0]; o.filmSize[1] = mFilmSize[1];
    o.frameIndex = mFrameIndex;
    o.seed = mSeed;
}
