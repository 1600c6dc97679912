/*
 * config.hpp -- shared types of the visualTur C++ surface (config.hpp:5-19
 * of the reference): node codes, per-ray states and visibleCube_t.
 * Vector types are HIP's (int3, float3, make_int3, make_float3).
 */
#ifndef VT_CONFIG_HPP
#define VT_CONFIG_HPP

#include <hip/hip_vector_types.h>

#include "visualtur_amd.h"

/* identifier type of an octree node (Morton code with a sentinel bit) */
typedef unsigned long long index_node_t;

#define CUBE     (unsigned char)8
#define PAINTED  (unsigned char)4
#define CACHED   (unsigned char)2
#define NOCACHED (unsigned char)1
#define NOCUBE   (unsigned char)0

/* same 32-byte layout as the reference struct {id, data, state, cubeID} */
typedef vt_visible_cube visibleCube_t;

#endif
