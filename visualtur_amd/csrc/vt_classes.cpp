// vt_classes.cpp -- the reference's C++ class surface on top of the C ABI.
//
// visualTur, Camera, Screen, Octree, lruCache, rayCaster and FileManager
// keep the reference signatures (inc/*.hpp) so that callers of the
// reference compile against this library.  Every GPU operation goes
// through include/visualtur_amd.h; a non-zero status becomes a
// std::runtime_error carrying vt_last_error().
#include <math.h>
#include <stdio.h>

#include <stdexcept>
#include <string>

#include "Camera.hpp"
#include "FileManager.hpp"
#include "Octree.hpp"
#include "Screen.hpp"
#include "lruCache.hpp"
#include "rayCaster.hpp"
#include "visualTur.hpp"

namespace {

void check(int rc, const char* what)
{
    if (rc != VT_OK) {
        char buf[512];
        vt_last_error(nullptr, buf, sizeof buf);
        throw std::runtime_error(std::string(what) + ": " + buf);
    }
}

// float3 helpers (component-wise, in the operand order the reference uses)
float3 f3(float x, float y, float z)
{
    float3 v;
    v.x = x; v.y = y; v.z = z;
    return v;
}
float3 add(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
float3 sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
float3 scale(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
float3 cross(float3 a, float3 b)
{
    return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
float3 unit(float3 a)
{
    double l = sqrt((double)a.x * a.x + (double)a.y * a.y + (double)a.z * a.z);
    return f3((float)(a.x / l), (float)(a.y / l), (float)(a.z / l));
}

vt_context* new_context()
{
    vt_context* c = nullptr;
    check(vt_context_create(-1, &c), "vt_context_create");
    return c;
}

}  // namespace

// ================================================================== Screen
// Screen.cpp:11-101

Screen::Screen(int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w)
    : H(p_H), W(p_W), distance(p_d), fov_H(p_fov_h), fov_W(p_fov_w)
{
    createScreen();
}

Screen::~Screen() {}

void Screen::createScreen()
{
    check(vt_screen_size(distance, fov_H, fov_W, &h, &w), "Screen");
}

int Screen::get_H() { return H; }
int Screen::get_W() { return W; }
float Screen::get_h() { return h; }
float Screen::get_w() { return w; }
float Screen::get_Distance() { return distance; }
float Screen::get_fovH() { return fov_H; }
float Screen::get_fovW() { return fov_W; }
void Screen::set_H(int pH) { H = pH; }
void Screen::set_W(int pW) { W = pW; }
void Screen::set_Distance(float pd) { distance = pd; createScreen(); }
void Screen::set_fovH(float pfh) { fov_H = pfh; createScreen(); }
void Screen::set_fovW(float pfw) { fov_W = pfw; createScreen(); }

// ================================================================== Camera
// Camera.cu:12-244

Camera::Camera(int nRP, int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w)
    : ctx(new_context()), ownCtx(true), raysDirty(true), screen(nullptr), rayDirections(nullptr)
{
    init(nRP, p_H, p_W, p_d, p_fov_h, p_fov_w);
}

Camera::Camera(vt_context* context, int nRP, int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w)
    : ctx(context), ownCtx(false), raysDirty(true), screen(nullptr), rayDirections(nullptr)
{
    init(nRP, p_H, p_W, p_d, p_fov_h, p_fov_w);
}

void Camera::init(int nRP, int p_H, int p_W, float p_d, float p_fov_h, float p_fov_w)
{
    screen = new Screen(p_H, p_W, p_d, p_fov_h, p_fov_w);
    numRayPixel = nRP;
    look = f3(0.0f, 0.0f, -1.0f);
    up = f3(0.0f, 1.0f, 0.0f);
    position = f3(0.0f, 0.0f, 0.0f);
    right = cross(look, up);
    RotatedX = RotatedY = RotatedZ = 0.0f;
    numRays = screen->get_H() * screen->get_W() * numRayPixel * numRayPixel;
    check(vt_malloc(ctx, (size_t)numRays * sizeof(float3), (void**)&rayDirections), "Camera rays");
    UpdateRays();
}

Camera::~Camera()
{
    if (rayDirections) vt_free(ctx, rayDirections);
    delete screen;
    if (ownCtx) vt_context_destroy(ctx);
}

void Camera::UpdateRays()
{
    vt_camera c = state();
    check(vt_camera_rays(ctx, &c, (float*)rayDirections), "Camera::UpdateRays");
    raysDirty = false;
}

void Camera::Move(float3 Direction)
{
    position = add(position, Direction);
    UpdateRays();
}

void Camera::RotateX(float Angle)
{
    float s = (float)sin(Angle * (M_PI / 180.0));
    float c = (float)cos(Angle * (M_PI / 180.0));
    RotatedX += Angle;
    look = add(scale(look, c), scale(up, s));   // the reference discards normalize(look)
    up = scale(cross(look, right), -1.0f);
    UpdateRays();
}

void Camera::RotateY(float Angle)
{
    float s = (float)sin(Angle * (M_PI / 180.0));
    float c = (float)cos(Angle * (M_PI / 180.0));
    RotatedY += Angle;
    look = sub(scale(look, c), scale(right, s));
    right = cross(look, up);
    UpdateRays();
}

void Camera::RotateZ(float Angle)
{
    float s = (float)sin(Angle * (M_PI / 180.0));
    float c = (float)cos(Angle * (M_PI / 180.0));
    RotatedZ += Angle;
    right = add(scale(right, c), scale(up, s));
    up = scale(cross(look, right), -1.0f);
    UpdateRays();
}

void Camera::MoveForward(float Distance)
{
    position = add(position, scale(look, -Distance));
    UpdateRays();
}

// Camera.cu:146-156: MoveUpward moves along `right`, StrafeRight along `up`
void Camera::MoveUpward(float Distance)
{
    position = add(position, scale(right, Distance));
    UpdateRays();
}

void Camera::StrafeRight(float Distance)
{
    position = add(position, scale(up, Distance));
    UpdateRays();
}

int Camera::get_H() { return screen->get_H(); }
int Camera::get_W() { return screen->get_W(); }
float Camera::get_h() { return screen->get_h(); }
float Camera::get_w() { return screen->get_w(); }
float Camera::get_Distance() { return screen->get_Distance(); }
float Camera::get_fovH() { return screen->get_fovH(); }
float Camera::get_fovW() { return screen->get_fovW(); }
int Camera::get_numRayPixel() { return numRayPixel; }
int Camera::get_numRays() { return numRays; }
float3* Camera::get_rayDirections()
{
    if (raysDirty) UpdateRays();
    return rayDirections;
}
float3 Camera::get_look() { return look; }
float3 Camera::get_up() { return up; }
float3 Camera::get_right() { return right; }
float3 Camera::get_position() { return position; }
float Camera::get_RotatedX() { return RotatedX; }
float Camera::get_RotatedY() { return RotatedY; }
float Camera::get_RotatedZ() { return RotatedZ; }

static void realloc_rays(vt_context* ctx, float3** rays, int numRays)
{
    if (*rays) vt_free(ctx, *rays);
    *rays = nullptr;
    check(vt_malloc(ctx, (size_t)numRays * sizeof(float3), (void**)rays), "Camera rays");
}

void Camera::set_H(int pH)
{
    screen->set_H(pH);
    numRays = screen->get_H() * screen->get_W() * numRayPixel * numRayPixel;
    realloc_rays(ctx, &rayDirections, numRays);
    UpdateRays();
}

void Camera::set_W(int pW)
{
    screen->set_W(pW);
    numRays = screen->get_H() * screen->get_W() * numRayPixel * numRayPixel;
    realloc_rays(ctx, &rayDirections, numRays);
    UpdateRays();
}

void Camera::set_Distance(float pd)
{
    screen->set_Distance(pd);
    UpdateRays();
}

void Camera::set_fovH(float pfh)
{
    screen->set_fovH(pfh);
    UpdateRays();
}

void Camera::set_fovW(float pfw)
{
    screen->set_fovW(pfw);
    UpdateRays();
}

void Camera::set_RayPerPixel(int rpp)
{
    numRayPixel = rpp;
    numRays = screen->get_H() * screen->get_W() * numRayPixel * numRayPixel;
    realloc_rays(ctx, &rayDirections, numRays);
    UpdateRays();
}

void Camera::setPose(float3 pos, float3 lookDir, float3 upDir, float3 rightDir)
{
    position = pos;
    look = lookDir;
    up = upDir;
    right = rightDir;
    UpdateRays();
}

void Camera::lookAt(float3 eye, float3 target, float3 upHint)
{
    float3 l = unit(sub(target, eye));
    float3 r = unit(cross(l, upHint));
    setPose(eye, l, cross(r, l), r);
}

vt_camera Camera::state()
{
    vt_camera c;
    c.pos[0] = position.x; c.pos[1] = position.y; c.pos[2] = position.z;
    c.look[0] = look.x; c.look[1] = look.y; c.look[2] = look.z;
    c.up[0] = up.x; c.up[1] = up.y; c.up[2] = up.z;
    c.right[0] = right.x; c.right[1] = right.y; c.right[2] = right.z;
    c.h = screen->get_h();
    c.w = screen->get_w();
    c.distance = screen->get_Distance();
    c.W = screen->get_W();
    c.H = screen->get_H();
    return c;
}

vt_context* Camera::context() { return ctx; }

// ================================================================== Octree
// Octree.cu:769-911

Octree::Octree(const char* file_name, Camera* p_camera, int p_maxLevel)
    : camera(p_camera), ctx(p_camera->context()), maxLevel(p_maxLevel)
{
    int32_t rd[3];
    int32_t n = 0;
    check(vt_octree_load_file(ctx, file_name, maxLevel, &isosurface, &n, rd), "Octree");
    nLevels = n;
    dimension = 1 << n;
    realDim = make_int3(rd[0], rd[1], rd[2]);
    resetState();
}

Octree::~Octree() {}

float Octree::getIsosurface() { return isosurface; }
int Octree::getnLevels() { return nLevels; }

void Octree::resetState()
{
    check(vt_stage_reset(ctx, camera->get_numRays()), "Octree::resetState");
}

bool Octree::getBoxIntersected(int finalLevel, visibleCube_t* visibleGPU, visibleCube_t* visibleCPU)
{
    vt_camera c = camera->state();
    check(vt_stage_traverse(ctx, &c, finalLevel, visibleGPU), "Octree::getBoxIntersected");
    check(vt_memcpy_d2h(ctx, visibleCPU, visibleGPU, (size_t)camera->get_numRays() * sizeof(visibleCube_t)),
          "Octree::getBoxIntersected copy");
    return true;
}

// ================================================================ lruCache
// lruCache.cu:101-322

lruCache::lruCache(char* file_name, char* dataset_name, int maxElements, int3 cDim, int cI, int levelC,
                   int nLevelsO, int maxElementsCPU)
    : ctx(new_context()), ownCtx(true), numElements(maxElements), numElementsCPU(maxElementsCPU), cubeDim(cDim),
      cubeInc(make_int3(cI, cI, cI)), levelCube(levelC), levelOctree(nLevelsO), file(file_name),
      dataset(dataset_name)
{
    create();
}

lruCache::lruCache(vt_context* context, char* file_name, char* dataset_name, int maxElements, int3 cDim, int cI,
                   int levelC, int nLevelsO, int maxElementsCPU)
    : ctx(context), ownCtx(false), numElements(maxElements), numElementsCPU(maxElementsCPU), cubeDim(cDim),
      cubeInc(make_int3(cI, cI, cI)), levelCube(levelC), levelOctree(nLevelsO), file(file_name),
      dataset(dataset_name)
{
    create();
}

void lruCache::create()
{
    if (cubeDim.x != cubeDim.y || cubeDim.x != cubeDim.z)
        throw std::runtime_error("lruCache: bricks must be cubic (posToIndex strides by dim.x)");
    check(vt_cache_create(ctx, levelCube, cubeDim.x, cubeInc.x, numElements, numElementsCPU), "lruCache");
    check(vt_cache_attach_hdf5(ctx, file.c_str(), dataset.c_str()), "lruCache file");
}

lruCache::~lruCache()
{
    vt_cache_destroy(ctx);
    if (ownCtx) vt_context_destroy(ctx);
}

int lruCache::get_numElements() { return numElements; }
int3 lruCache::get_cubeDim() { return cubeDim; }
int lruCache::get_cubeInc() { return cubeInc.x; }

void lruCache::changeDimensionCube(int maxElements, int3 cDim, int cI)
{
    check(vt_cache_destroy(ctx), "lruCache::changeDimensionCube");
    numElements = maxElements;
    cubeDim = cDim;
    cubeInc = make_int3(cI, cI, cI);
    create();
}

void lruCache::updateCache(visibleCube_t* visibleCubes, int num, int nLevels)
{
    check(vt_stage_cache(ctx, visibleCubes, num, levelOctree, nLevels), "lruCache::updateCache");
}

vt_cache_stats lruCache::statistics()
{
    vt_cache_stats s;
    check(vt_cache_get_stats(ctx, &s), "lruCache::statistics");
    return s;
}

void lruCache::printStatistics()
{
    vt_cache_stats s = statistics();
    printf("GPU slots %lld resident %lld missed %lld uploads %lld evictions %lld | CPU slots %lld resident %lld "
           "file reads %lld\n",
           (long long)s.gpu_slots, (long long)s.gpu_resident, (long long)s.lookups_missed, (long long)s.uploads,
           (long long)s.evictions, (long long)s.cpu_slots, (long long)s.cpu_resident, (long long)s.file_reads);
}

// =============================================================== rayCaster
// rayCaster.cu:365-394

rayCaster::rayCaster(float isosurface, float3 lposition) : iso(isosurface), lightPosition(lposition), step(0.5f) {}

rayCaster::~rayCaster() {}

void rayCaster::render(Camera* camera, int levelO, int levelC, int nLevel, visibleCube_t* cube, int3 cubeDim,
                       int3 cubeInc, float* buffer)
{
    vt_camera c = camera->state();
    vt_render_params p;
    p.levelO = levelO;
    p.levelC = levelC;
    p.iso = iso;
    p.light[0] = lightPosition.x; p.light[1] = lightPosition.y; p.light[2] = lightPosition.z;
    p.color_scale = 256.0f;
    check(vt_stage_raycast(camera->context(), &c, &p, nLevel, cubeDim.x, cubeInc.x, cube, buffer), "rayCaster::render");
}

void rayCaster::renderCube(Camera* camera, float* cube, int3 minBox, int3 cubeDim, int3 cubeInc, float* buffer)
{
    vt_camera c = camera->state();
    float light[3] = {lightPosition.x, lightPosition.y, lightPosition.z};
    int32_t mb[3] = {minBox.x, minBox.y, minBox.z};
    int32_t cd[3] = {cubeDim.x, cubeDim.y, cubeDim.z};
    check(vt_render_cube(camera->context(), &c, iso, light, cube, mb, cd, cubeInc.x, buffer), "rayCaster::renderCube");
}

// ============================================================= FileManager
// FileManager.cu:11-243

FileManager::FileManager(const char* file_name, const char* dataset_name) : file(nullptr)
{
    int64_t d[3];
    check(vt_file_open(file_name, dataset_name, &file, d), "FileManager");
    for (int k = 0; k < 3; ++k) dims[k] = (unsigned long long)d[k];
}

FileManager::~FileManager() { vt_file_close(file); }

int FileManager::readHDF5(float* volume)
{
    int32_t s[3] = {0, 0, 0};
    int32_t e[3] = {(int32_t)dims[0], (int32_t)dims[1], (int32_t)dims[2]};
    return vt_file_read_box(file, s, e, volume) == VT_OK ? 0 : -1;
}

void FileManager::readHDF5_Voxel_Array(int3 s, int3 e, float* data)
{
    int32_t s3[3] = {s.x, s.y, s.z}, e3[3] = {e.x, e.y, e.z};
    check(vt_file_read_box(file, s3, e3, data), "FileManager::readHDF5_Voxel_Array");
}

int FileManager::CreateFile(float* volume, char* newname, char* dataset_name, int x, int y, int z)
{
    return vt_file_create(newname, dataset_name, volume, x, y, z, 0) == VT_OK ? 0 : -1;
}

// =============================================================== visualTur
// visualTur.cu:10-228

visualTur::visualTur(visualTurParams_t p)
    : ctx(new_context()), camera(nullptr), visibleCubesCPU(nullptr), visibleCubesGPU(nullptr), cache(nullptr),
      octree(nullptr), raycaster(nullptr), framePainted(false), timingO(0.0), timingC(0.0), timingR(0.0)
{
    lastStats = vt_frame_stats();
    camera = new Camera(ctx, p.numRayPx, p.H, p.W, p.distance, p.fov_H, p.fov_W);
    visibleCubesCPU = new visibleCube_t[camera->get_numRays()];
    check(vt_malloc(ctx, (size_t)camera->get_numRays() * sizeof(visibleCube_t), (void**)&visibleCubesGPU),
          "visualTur");
    resetVisibleCubes();
    octreeLevel = p.octreeLevel;
    octree = new Octree(p.octreeFile, camera, octreeLevel);
    cache = new lruCache(ctx, p.hdf5File, p.dataset_name, p.maxElementsCache, p.dimCubeCache, p.cubeInc,
                         p.levelCubes, p.octreeLevel, p.maxElementsCache_CPU);
    cubeLevel = p.levelCubes;
    iso = octree->getIsosurface();
    raycaster = new rayCaster(iso, make_float3(0.0f, 512.0f, 0.0f));
}

visualTur::~visualTur()
{
    delete raycaster;
    delete cache;
    delete octree;
    delete[] visibleCubesCPU;
    if (visibleCubesGPU) vt_free(ctx, visibleCubesGPU);
    delete camera;
    vt_context_destroy(ctx);
}

void visualTur::resetVisibleCubes()
{
    check(vt_memset(ctx, visibleCubesGPU, 0, (size_t)camera->get_numRays() * sizeof(visibleCube_t)),
          "visualTur::resetVisibleCubes");
    framePainted = false;
}

static void realloc_cubes(vt_context* ctx, Camera* camera, visibleCube_t** cpu, visibleCube_t** gpu)
{
    delete[] *cpu;
    if (*gpu) vt_free(ctx, *gpu);
    *gpu = nullptr;
    *cpu = new visibleCube_t[camera->get_numRays()];
    check(vt_malloc(ctx, (size_t)camera->get_numRays() * sizeof(visibleCube_t), (void**)gpu), "visualTur");
}

void visualTur::changeScreen(int pW, int pH, float pfovW, float pfovH, float pDistance)
{
    camera->set_W(pW);
    camera->set_H(pH);
    camera->set_fovW(pfovW);
    camera->set_fovH(pfovH);
    camera->set_Distance(pDistance);
    realloc_cubes(ctx, camera, &visibleCubesCPU, &visibleCubesGPU);
    resetVisibleCubes();
    octree->resetState();
}

void visualTur::changeNumRays(int pnR)
{
    camera->set_RayPerPixel(pnR);
    realloc_cubes(ctx, camera, &visibleCubesCPU, &visibleCubesGPU);
    resetVisibleCubes();
    octree->resetState();
}

void visualTur::changeCacheParameters(int nE, int3 cDim, int cInc)
{
    cache->changeDimensionCube(nE, cDim, cInc);
    realloc_cubes(ctx, camera, &visibleCubesCPU, &visibleCubesGPU);
    resetVisibleCubes();
    octree->resetState();
}

#define CAMERA_MOVE(name, type, call)      \
    void visualTur::name(type v)           \
    {                                      \
        camera->call(v);                   \
        resetVisibleCubes();               \
        octree->resetState();              \
    }
CAMERA_MOVE(camera_Move, float3, Move)
CAMERA_MOVE(camera_RotateX, float, RotateX)
CAMERA_MOVE(camera_RotateY, float, RotateY)
CAMERA_MOVE(camera_RotateZ, float, RotateZ)
CAMERA_MOVE(camera_MoveForward, float, MoveForward)
CAMERA_MOVE(camera_MoveUpward, float, MoveUpward)
CAMERA_MOVE(camera_StrafeRight, float, StrafeRight)
#undef CAMERA_MOVE

void visualTur::renderFrame(float* pixelBuffer, vt_hit* hits)
{
    vt_camera c = camera->state();
    vt_render_params p;
    p.levelO = octreeLevel;
    p.levelC = cubeLevel;
    p.iso = iso;
    p.light[0] = 0.0f; p.light[1] = 512.0f; p.light[2] = 0.0f;
    p.color_scale = 256.0f;
    check(vt_render_frame(ctx, &c, &p, pixelBuffer, hits, &lastStats), "visualTur::updateVisibleCubes");
    timingR += lastStats.ms_total * 1e-3;
}

void visualTur::updateVisibleCubes(float* pixelBuffer)
{
    if (framePainted) return;   // every ray is already PAINTED
    renderFrame(pixelBuffer, nullptr);
    framePainted = true;
}

void visualTur::updateVisibleCubesStaged(float* pixelBuffer)
{
    const int num = camera->get_numRays();
    int3 inc = make_int3(cache->get_cubeInc(), cache->get_cubeInc(), cache->get_cubeInc());
    for (;;) {
        octree->getBoxIntersected(octreeLevel, visibleCubesGPU, visibleCubesCPU);
        cache->updateCache(visibleCubesCPU, num, octree->getnLevels());
        int painted = 0;
        for (int i = 0; i < num; ++i)
            if (visibleCubesCPU[i].state == PAINTED) painted++;
        if (painted == num) break;
        check(vt_memcpy_h2d(ctx, visibleCubesGPU, visibleCubesCPU, (size_t)num * sizeof(visibleCube_t)),
              "visualTur copy");
        raycaster->render(camera, octreeLevel, cubeLevel, octree->getnLevels(), visibleCubesGPU,
                          cache->get_cubeDim(), inc, pixelBuffer);
    }
    framePainted = true;
}

vt_frame_stats visualTur::lastFrameStats() { return lastStats; }
Camera* visualTur::getCamera() { return camera; }
