/*
 * visualTur.hpp -- the frame renderer (visualTur.hpp of the reference).
 * updateVisibleCubes renders with the fused GPU pass loop; the reference's
 * staged loop over Octree / lruCache / rayCaster is kept as
 * updateVisibleCubesStaged and produces the same pixels.
 */
#ifndef VT_VISUALTUR_HPP
#define VT_VISUALTUR_HPP

#include "Octree.hpp"
#include "lruCache.hpp"
#include "rayCaster.hpp"

typedef struct
{
    /* Camera settings */
    int W;
    int H;
    float distance;
    float fov_H;
    float fov_W;
    int numRayPx;
    /* Cube Cache settings */
    int maxElementsCache;
    int maxElementsCache_CPU;
    int3 dimCubeCache;
    int cubeInc;
    int levelCubes;
    /* hdf5 settings */
    char *hdf5File;
    char *dataset_name;
    /* octree settings */
    char *octreeFile;
    int octreeLevel;
} visualTurParams_t;

class visualTur
{
    private:
        vt_context *ctx;
        Camera *camera;
        visibleCube_t *visibleCubesCPU;
        visibleCube_t *visibleCubesGPU;
        lruCache *cache;
        int cubeLevel;
        Octree *octree;
        float iso;
        int octreeLevel;
        rayCaster *raycaster;
        bool framePainted;
        vt_frame_stats lastStats;
        void resetVisibleCubes();
        double timingO;
        double timingC;
        double timingR;
    public:
        visualTur(visualTurParams_t initParams);
        ~visualTur();

        void changeScreen(int pW, int pH, float pfovW, float pfovH, float pDistance);
        void changeNumRays(int pnR);
        void changeCacheParameters(int nE, int3 cDim, int cInc);

        void camera_Move(float3 Direction);
        void camera_RotateX(float Angle);
        void camera_RotateY(float Angle);
        void camera_RotateZ(float Angle);
        void camera_MoveForward(float Distance);
        void camera_MoveUpward(float Distance);
        void camera_StrafeRight(float Distance);

        /* Renders the frame into the device buffer pixelBuffer (W*H*4
         * floats).  As in the reference, a second call without a camera or
         * screen change does nothing (visualTur.cu:167-200). */
        void updateVisibleCubes(float *pixelBuffer);
        /* The reference's pass loop driven through Octree, lruCache and
         * rayCaster (visualTur.cu:175-220). */
        void updateVisibleCubesStaged(float *pixelBuffer);
        /* Always renders; optional per-ray hit records (device). */
        void renderFrame(float *pixelBuffer, vt_hit *hits);
        vt_frame_stats lastFrameStats();
        Camera *getCamera();
};

#endif
