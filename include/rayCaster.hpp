/* rayCaster.hpp -- trilinear march + shading (rayCaster.hpp of the reference). */
#ifndef VT_RAYCASTER_HPP
#define VT_RAYCASTER_HPP

#include "Camera.hpp"
#include "config.hpp"

class rayCaster
{
    private:
        float iso;
        float3 lightPosition;
        float step;
    public:
        rayCaster(float isosurface, float3 lposition);
        ~rayCaster();
        /* March the CACHED rays of the device array `cube` (rayCaster.cu:375-384). */
        void render(Camera *camera, int levelO, int levelC, int nLevel, visibleCube_t *cube, int3 cubeDim,
                    int3 cubeInc, float *buffer);
        /* One brick of (cubeDim+2*cubeInc)^3 device floats (rayCaster.cu:385-394). */
        void renderCube(Camera *camera, float *cube, int3 minBox, int3 cubeDim, int3 cubeInc, float *buffer);
};

#endif
