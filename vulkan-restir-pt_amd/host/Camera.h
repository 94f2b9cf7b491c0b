// Camera of the host layer: same interface and conventions as the reference's
// Camera (src/Camera.h:8-69, src/Camera.cpp): z-up, angles in degrees
// (yaw, pitch, roll-in-radians), vertical FOV in degrees, lookAt/perspective
// matrices with the Vulkan y flip. toDevice() produces the 352-B uniform of
// layouts.glsl:27-51.
#pragma once
#include "restirpt.h"
#include <cstdint>

struct Vec3 { float x = 0, y = 0, z = 0; };
struct Mat4 { float m[16]; };  // column-major

class Camera {
public:
    Camera(Vec3 pos = {0, 0, 0}, Vec3 angle = {90.0f, 0.0f, 0.0f});

    void move(Vec3 v);
    void rotate(Vec3 angle);
    void setFOV(float fov);
    void lookAt(Vec3 focus);
    void setDir(Vec3 dir);
    void setPos(Vec3 pos);
    void setAngle(Vec3 angle);
    void setFilmSize(uint32_t w, uint32_t h);
    void setPlanes(float nearZ, float farZ);
    // setLensRadius / setFocalDist (src/Camera.h:35-38)
    void setLens(float lensRadius, float focalDist) { mLensRadius = lensRadius; mFocalDist = focalDist; }

    Vec3 pos() const { return mPos; }
    Vec3 angle() const { return mAngle; }
    float FOV() const { return mFOV; }
    const Mat4& projView() const { return mProjView; }

    // src/Camera.cpp:68-76: new per-frame seed, remember projView, count frames
    void nextFrame(uint32_t seed);
    void setClearFlag();
    void update();

    void toDevice(rt_camera& out) const;

private:
    Mat4 mViewMatrix{}, mProjMatrix{}, mProjView{}, mLastProjView{};
    Vec3 mPos{};
    float mFOV = 45.0f;
    Vec3 mAngle{};
    float mNear = 1e-3f;
    Vec3 mFront{};
    float mFar = 1e3f;
    Vec3 mRight{};
    float mLensRadius = 0.0f;
    Vec3 mUp{0, 0, 1};
    float mFocalDist = 1.0f;
    uint32_t mFilmSize[2] = {1, 1};
    uint32_t mFrameIndex = 0;
    uint32_t mSeed = 0;
};
