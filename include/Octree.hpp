/*
 * Octree.hpp -- octree of cells crossing the isovalue (Octree.hpp of the
 * reference).  Reads the reference binary file and uploads levels
 * 0..maxLevel as the GPU membership tables; traversal state lives on the
 * device, one node code per ray instead of a 32-entry stack.
 */
#ifndef VT_OCTREE_HPP
#define VT_OCTREE_HPP

#include "Camera.hpp"
#include "config.hpp"

class Octree
{
    private:
        Camera *camera;
        vt_context *ctx;
        float isosurface;
        int3 realDim;
        int dimension;
        int nLevels;
        int maxLevel;
    public:
        /* Throws std::runtime_error on a missing file or bad magic
         * (the reference throws std::exception, Octree.cu:781-785). */
        Octree(const char *file_name, Camera *p_camera, int p_maxLevel);
        ~Octree();
        int getnLevels();
        float getIsosurface();
        void resetState();
        /* Next node of level finalLevel for every NOCUBE ray; the device
         * array is copied to visibleCPU afterwards (Octree.cu:883-901). */
        bool getBoxIntersected(int finalLevel, visibleCube_t *visibleGPU, visibleCube_t *visibleCPU);
};

#endif
