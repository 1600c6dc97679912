// Headless counterpart of the reference's viewer (drawPixel.cu): the same
// prompts (display size and camera position, drawPixel.cu:190-203), the same
// visualTur parameters (drawPixel.cu:214-236) and the same keys moving the
// camera (KeyDown, drawPixel.cu:127-176), each followed by a redraw
// (display, drawPixel.cu:108-124).  There is no window on a compute node, so
// every drawn frame goes to a PPM file in the output directory (laid out by
// vt_image_write_ppm as the reference's image writer lays out its bitmap),
// optionally with the raw float4 pixels and the camera state, and the frame
// rate goes to stderr instead of the window title (computeFPS,
// drawPixel.cu:26-44).
//
// usage: vt_view <hdf5_file> <dataset_name> <octree_file> [device] [out_dir] [raw]
//   stdin: Width, Height, then the camera position X, Y, Z (integers), then
//   keys, whitespace ignored:
//     a / d  rotate about y by +5 / -5 degrees    w / s  move forward by -5 / +5
//     x / y  rotate about x by +5 / -5 degrees    c / v  strafe right by -5 / +5
//     f / r  move upward by -5 / +5               m / n  rotate about z by -5 / +5
//     q      quit (so does the end of the input)
//   Any other key redraws without a move, as display() does after a key
//   KeyDown does not bind; a redraw without a move paints nothing new.
//   Frame k is out_dir/frame_<k>.ppm (k from 0, the first frame drawn before
//   any key); with "raw", also frame_<k>.rgba (W*H float4) and
//   frame_<k>.camera (the vt_camera state).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <chrono>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "FileManager.hpp"
#include "visualTur.hpp"

namespace {

// getnLevelFile (drawPixel.cu:46-106): levels of the cube that holds the
// dataset's largest dimension
int levels_of(const char* file, const char* dataset)
{
    FileManager fm(file, dataset);
    int dimension = (int)fm.dims[0];
    if ((int)fm.dims[1] > dimension) dimension = (int)fm.dims[1];
    if ((int)fm.dims[2] > dimension) dimension = (int)fm.dims[2];
    int n = 0;
    while ((1 << n) < dimension) n++;
    return n;
}

bool write_file(const std::string& path, const void* p, size_t n)
{
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = fwrite(p, 1, n, f) == n;
    return fclose(f) == 0 && ok;
}

// KeyDown (drawPixel.cu:127-176); false for a key it does not bind
bool apply_key(visualTur* v, char key)
{
    switch (key) {
    case 'a': v->camera_RotateY(5.0f); return true;
    case 'd': v->camera_RotateY(-5.0f); return true;
    case 'w': v->camera_MoveForward(-5.0f); return true;
    case 's': v->camera_MoveForward(5.0f); return true;
    case 'x': v->camera_RotateX(5.0f); return true;
    case 'y': v->camera_RotateX(-5.0f); return true;
    case 'c': v->camera_StrafeRight(-5.0f); return true;
    case 'v': v->camera_StrafeRight(5.0f); return true;
    case 'f': v->camera_MoveUpward(-5.0f); return true;
    case 'r': v->camera_MoveUpward(5.0f); return true;
    case 'm': v->camera_RotateZ(-5.0f); return true;
    case 'n': v->camera_RotateZ(5.0f); return true;
    default: return false;
    }
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: vt_view hdf5_file dataset_name octree_file [device] [out_dir] [raw]\n");
        return 2;
    }
    const int device = argc > 4 ? atoi(argv[4]) : 0;
    const std::string out = argc > 5 ? argv[5] : ".";
    const bool raw = argc > 6 && std::string(argv[6]) == "raw";

    int W = 0, H = 0, x = 0, y = 0, z = 0;
    std::cout << "Display Resolution:" << std::endl << "Width: ";
    std::cin >> W;
    std::cout << "Height: ";
    std::cin >> H;
    std::cout << "Camera position (X,Y,Z):" << std::endl << "X: ";
    std::cin >> x;
    std::cout << "Y: ";
    std::cin >> y;
    std::cout << "Z: ";
    std::cin >> z;
    std::cout << std::endl;
    if (!std::cin || W <= 0 || H <= 0) {
        fprintf(stderr, "vt_view: expected a positive width and height and a camera position on stdin\n");
        return 2;
    }

    visualTur* view = nullptr;
    float* screenG = nullptr;
    int rc = 0;
    try {
        const int nLevel = levels_of(argv[1], argv[2]);
        size_t free_bytes = 0, total_bytes = 0;
        if (vt_device_memory(device, &free_bytes, &total_bytes) != VT_OK)
            throw std::runtime_error("cannot query device " + std::to_string(device));

        visualTurParams_t params;
        params.W = W;
        params.H = H;
        params.fov_H = 35.0f;
        params.fov_W = 35.0f;
        params.distance = 50.0f;
        params.numRayPx = 1;
        params.maxElementsCache = (int)(3 * (free_bytes / (38 * 38 * 38 * 4)) / 4);
        params.maxElementsCache_CPU = 5000;
        params.dimCubeCache = make_int3(32, 32, 32);
        params.cubeInc = 2;
        params.levelCubes = nLevel - 5;
        params.octreeLevel = (nLevel - 5) + 3;
        params.hdf5File = argv[1];
        params.dataset_name = argv[2];
        params.octreeFile = argv[3];
        view = new visualTur(params);
        view->camera_Move(make_float3((float)x, (float)y, (float)z));

        vt_context* ctx = view->getCamera()->context();
        const size_t bytes = (size_t)W * H * 4 * sizeof(float);
        if (vt_malloc(ctx, bytes, (void**)&screenG) != VT_OK) throw std::runtime_error("screen allocation failed");
        if (vt_memset(ctx, screenG, 0, bytes) != VT_OK) throw std::runtime_error("screen clear failed");
        std::vector<float> screenC((size_t)W * H * 4);

        int frame = 0;
        double avg_ms = 0.0;
        auto display = [&]() {
            const auto t0 = std::chrono::steady_clock::now();
            view->updateVisibleCubes(screenG);
            if (vt_memcpy_d2h(ctx, screenC.data(), screenG, bytes) != VT_OK)
                throw std::runtime_error("screen read-back failed");
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            avg_ms = frame == 0 ? ms : 0.9 * avg_ms + 0.1 * ms;
            char name[64];
            snprintf(name, sizeof name, "/frame_%04d", frame);
            const std::string base = out + name;
            if (vt_image_write_ppm(screenC.data(), W, H, (base + ".ppm").c_str()) != VT_OK)
                throw std::runtime_error("cannot write " + base + ".ppm");
            if (raw) {
                const vt_camera cam = view->getCamera()->state();
                if (!write_file(base + ".rgba", screenC.data(), bytes) ||
                    !write_file(base + ".camera", &cam, sizeof cam))
                    throw std::runtime_error("cannot write the raw frame " + base);
            }
            fprintf(stderr, "frame %d: %.3f ms, %3.1f fps\n", frame, ms, avg_ms > 0.0 ? 1000.0 / avg_ms : 0.0);
            frame++;
        };

        display();   // the first frame, drawn when the window would appear
        char key;
        while (std::cin.get(key)) {
            if (key == ' ' || key == '\n' || key == '\t' || key == '\r') continue;
            if (key == 'q') break;
            apply_key(view, key);
            display();
        }
        printf("vt_view: %d frames\n", frame);
    } catch (const std::exception& e) {
        fprintf(stderr, "vt_view: %s\n", e.what());
        rc = 4;
    }
    if (screenG) vt_free(view->getCamera()->context(), screenG);
    delete view;
    return rc;
}
