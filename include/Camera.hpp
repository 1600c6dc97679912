/*
 * Camera.hpp -- pinhole camera (Camera.hpp of the reference).
 * Ray directions are generated on the GPU (vt_camera_rays); every move
 * keeps the reference semantics (Camera.cu:85-156), including that
 * MoveUpward moves along `right` and StrafeRight along `up`.
 * Additions: setPose / lookAt (explicit orthonormal poses for camera paths)
 * and state() (the pose as the C ABI's vt_camera).
 */
#ifndef VT_CAMERA_HPP
#define VT_CAMERA_HPP

#include "Screen.hpp"

class Camera
{
    private:
        void UpdateRays();
        vt_context *ctx;
        bool ownCtx;
        bool raysDirty;
        Screen *screen;
        int numRayPixel;
        int numRays;
        float3 *rayDirections;
        float3 look;
        float3 up;
        float3 right;
        float3 position;
        float RotatedX;
        float RotatedY;
        float RotatedZ;
        void init(int nRP, int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w);
    public:
        /* Position (0,0,0), looking at (0,0,-1), up (0,1,0) (Camera.cu:12-29). */
        Camera(int nRP, int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w);
        /* Same, on an existing context (shared with an Octree/lruCache). */
        Camera(vt_context *context, int nRP, int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w);
        ~Camera();

        void Move(float3 Direction);
        void RotateX(float Angle);
        void RotateY(float Angle);
        void RotateZ(float Angle);
        void MoveForward(float Distance);
        void MoveUpward(float Distance);
        void StrafeRight(float Distance);
        int get_H();
        int get_W();
        float get_h();
        float get_w();
        float get_Distance();
        float get_fovH();
        float get_fovW();
        int get_numRayPixel();
        int get_numRays();
        float3 *get_rayDirections();   /* device array of W*H float3 */
        float3 get_look();
        float3 get_up();
        float3 get_right();
        float3 get_position();
        float get_RotatedX();
        float get_RotatedY();
        float get_RotatedZ();
        void set_H(int pH);
        void set_W(int pW);
        void set_Distance(float pd);
        void set_fovH(float pfh);
        void set_fovW(float pfw);
        void set_RayPerPixel(int rpp);

        void setPose(float3 pos, float3 lookDir, float3 upDir, float3 rightDir);
        void lookAt(float3 eye, float3 target, float3 upHint);
        vt_camera state();
        vt_context *context();
};

#endif
