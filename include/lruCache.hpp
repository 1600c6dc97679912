/*
 * lruCache.hpp -- brick cache (lruCache.hpp of the reference).
 * A device pool of bricks with a page table and clock eviction, backed by
 * a pinned host tier filled from the HDF5 file.  Bricks a pass already uses
 * are pinned while the pass loads the missing ones.
 */
#ifndef VT_LRUCACHE_HPP
#define VT_LRUCACHE_HPP

#include <string>

#include "config.hpp"

class lruCache
{
    private:
        vt_context *ctx;
        bool ownCtx;
        int numElements;
        int numElementsCPU;
        int3 cubeDim;
        int3 cubeInc;
        int levelCube;
        int levelOctree;
        std::string file;
        std::string dataset;
        void create();
    public:
        lruCache(char *file_name, char *dataset_name, int maxElements, int3 cDim, int cI, int levelC,
                 int nLevelsO, int maxElementsCPU);
        /* Same, on an existing context (the one the frame renders in). */
        lruCache(vt_context *context, char *file_name, char *dataset_name, int maxElements, int3 cDim, int cI,
                 int levelC, int nLevelsO, int maxElementsCPU);
        ~lruCache();
        int get_numElements();
        int3 get_cubeDim();
        int get_cubeInc();
        void changeDimensionCube(int maxElements, int3 cDim, int cI);
        void updateCache(visibleCube_t *visibleCubes, int num, int nLevels);
        void printStatistics();
        vt_cache_stats statistics();
};

#endif
