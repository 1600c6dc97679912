// Drives the C++ class surface the way the reference's drivers do
// (testFileManager.cu:39-104, testVisualTur.cu:25-109, testRayCaster.cu:11-69)
// and writes the raw outputs for the Python test to compare with the oracle.
//
// usage: drive_classes <volume.h5> <dataset> <octree file> <out dir>
//                      <W> <H> <levelCubes> <octreeLevel> <cubeInc> <gpu slots> <cpu slots>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "FileManager.hpp"
#include "visualTur.hpp"

static int failures = 0;
#define EXPECT(c, msg)                                           \
    do {                                                         \
        if (!(c)) {                                              \
            fprintf(stderr, "FAILED: %s (%s:%d)\n", msg, __FILE__, __LINE__); \
            failures++;                                          \
        }                                                        \
    } while (0)

static void dump(const std::string& path, const void* p, size_t n)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f || fwrite(p, 1, n, f) != n) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    fclose(f);
}

static void dump_camera(const std::string& path, Camera* c)
{
    vt_camera s = c->state();
    dump(path, &s, sizeof s);
}

// testFileManager.cu:8-36: outside the dataset zero, inside the shifted base read
static void check_filemanager(const char* file, const char* dataset)
{
    FileManager fm(file, dataset);
    const int N = 10;
    std::vector<float> base(N * N * N), box(N * N * N);
    fm.readHDF5_Voxel_Array(make_int3(0, 0, 0), make_int3(N, N, N), base.data());
    int offs[4][3] = {{-3, 0, 0}, {0, -4, 0}, {-2, -2, -5}, {-9, -9, -9}};
    for (auto& o : offs) {
        fm.readHDF5_Voxel_Array(make_int3(o[0], o[1], o[2]), make_int3(o[0] + N, o[1] + N, o[2] + N), box.data());
        for (int x = 0; x < N; ++x)
            for (int y = 0; y < N; ++y)
                for (int z = 0; z < N; ++z) {
                    int gx = x + o[0], gy = y + o[1], gz = z + o[2];
                    float v = box[(x * N + y) * N + z];
                    if (gx < 0 || gy < 0 || gz < 0) EXPECT(v == 0.0f, "zero outside the dataset");
                    else EXPECT(v == base[(gx * N + gy) * N + gz], "shifted read equals base read");
                }
    }
    std::vector<float> all((size_t)fm.dims[0] * fm.dims[1] * fm.dims[2]);
    EXPECT(fm.readHDF5(all.data()) == 0, "readHDF5");
    EXPECT(all[(size_t)(3 * fm.dims[1] + 4) * fm.dims[2] + 5] == base[(3 * N + 4) * N + 5], "full read");
    bool threw = false;
    try { FileManager bad("/nonexistent/file.h5", dataset); } catch (const std::runtime_error&) { threw = true; }
    EXPECT(threw, "missing file throws");
}

int main(int argc, char** argv)
{
    if (argc != 12) {
        fprintf(stderr, "usage: drive_classes h5 dataset octree outdir W H levelCubes octreeLevel cubeInc gpuSlots cpuSlots\n");
        return 2;
    }
    const char* h5 = argv[1];
    const char* ds = argv[2];
    const char* oct = argv[3];
    std::string out = argv[4];
    int W = atoi(argv[5]), H = atoi(argv[6]);
    int levelC = atoi(argv[7]), levelO = atoi(argv[8]), inc = atoi(argv[9]);
    int gslots = atoi(argv[10]), cslots = atoi(argv[11]);
    try {
        check_filemanager(h5, ds);

        FileManager fm(h5, ds);
        int D = 1;
        while (D < (int)fm.dims[0] || D < (int)fm.dims[1] || D < (int)fm.dims[2]) D <<= 1;
        int n = 0;
        while ((1 << n) < D) n++;
        int cube = 1 << (n - levelC);

        visualTurParams_t params;
        params.W = W;
        params.H = H;
        params.fov_H = 22.5f;
        params.fov_W = 22.5f;
        params.distance = 50.0f;
        params.numRayPx = 1;
        params.maxElementsCache = gslots;
        params.maxElementsCache_CPU = cslots;
        params.dimCubeCache = make_int3(cube, cube, cube);
        params.cubeInc = inc;
        params.levelCubes = levelC;
        params.octreeLevel = levelO;
        params.hdf5File = (char*)h5;
        params.dataset_name = (char*)ds;
        params.octreeFile = (char*)oct;
        visualTur* vt = new visualTur(params);

        float* screenG = nullptr;
        vt_context* ctx = vt->getCamera()->context();
        size_t bytes = (size_t)W * H * 4 * sizeof(float);
        if (vt_malloc(ctx, bytes, (void**)&screenG) != VT_OK) return 3;
        std::vector<float> a((size_t)W * H * 4), b((size_t)W * H * 4);

        // testVisualTur.cu:48-50: place the camera, then frames with strafes
        vt->camera_Move(make_float3(0.5f * D, 0.5f * D, 0.5f * D + 2.5f * D));
        vt->camera_MoveForward(1.0f);
        for (int m = 0; m < 3; ++m) {
            vt_memset(ctx, screenG, 0, bytes);
            vt->updateVisibleCubes(screenG);
            vt_memcpy_d2h(ctx, a.data(), screenG, bytes);
            // the reference's second call without a camera move paints nothing
            vt_memset(ctx, screenG, 0, bytes);
            vt->updateVisibleCubes(screenG);
            vt_memcpy_d2h(ctx, b.data(), screenG, bytes);
            bool untouched = true;
            for (float v : b) untouched = untouched && v == 0.0f;
            EXPECT(untouched, "second updateVisibleCubes is a no-op");
            dump(out + "/fused_" + std::to_string(m) + ".bin", a.data(), bytes);
            dump_camera(out + "/camera_" + std::to_string(m) + ".bin", vt->getCamera());
            // the reference's staged loop over Octree / lruCache / rayCaster
            vt->camera_Move(make_float3(0.0f, 0.0f, 0.0f));   // resets the per-ray state
            vt_memset(ctx, screenG, 0, bytes);
            vt->updateVisibleCubesStaged(screenG);
            vt_memcpy_d2h(ctx, b.data(), screenG, bytes);
            EXPECT(memcmp(a.data(), b.data(), bytes) == 0, "staged pass loop equals the fused frame");
            dump(out + "/staged_" + std::to_string(m) + ".bin", b.data(), bytes);
            vt->camera_StrafeRight(0.5f * D / 16.0f);
            vt->camera_RotateY(7.0f);
        }
        vt_free(ctx, screenG);

        // testRayCaster.cu:21-45: one brick read with its halo, rendered alone
        {
            Camera cam(1, H, W, 50.0f, 22.5f, 22.5f);
            cam.Move(make_float3(0.5f * D, 0.5f * D, 0.5f * D + 2.5f * D));
            int3 mb = make_int3(0, 0, 0);
            int S = D + 2 * inc;
            std::vector<float> brick((size_t)S * S * S);
            fm.readHDF5_Voxel_Array(make_int3(-inc, -inc, -inc), make_int3(D + inc, D + inc, D + inc), brick.data());
            float* d_brick = nullptr;
            float* d_screen = nullptr;
            vt_malloc(cam.context(), brick.size() * 4, (void**)&d_brick);
            vt_malloc(cam.context(), bytes, (void**)&d_screen);
            vt_memcpy_h2d(cam.context(), d_brick, brick.data(), brick.size() * 4);
            Octree octree(oct, &cam, 0);
            rayCaster rc(octree.getIsosurface(), make_float3(0.0f, 512.0f, 0.0f));
            rc.renderCube(&cam, d_brick, mb, make_int3(D, D, D), make_int3(inc, inc, inc), d_screen);
            vt_memcpy_d2h(cam.context(), a.data(), d_screen, bytes);
            dump(out + "/cube.bin", a.data(), bytes);
            dump_camera(out + "/camera_cube.bin", &cam);
            vt_free(cam.context(), d_brick);
            vt_free(cam.context(), d_screen);
        }
        delete vt;
        bool threw = false;
        try {
            params.octreeFile = (char*)"/nonexistent/octree";
            visualTur bad(params);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        EXPECT(threw, "missing octree file throws");
    } catch (const std::exception& e) {
        fprintf(stderr, "exception: %s\n", e.what());
        return 4;
    }
    printf("drive_classes: %d failures\n", failures);
    return failures ? 1 : 0;
}
