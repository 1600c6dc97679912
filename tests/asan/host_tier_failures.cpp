// Host-only harness of the pinned host tier's failure paths (HostTier in
// vt_context.hpp), built with AddressSanitizer by tests/test_host_tier_asan.py
// and run on the CPU: it makes no HIP call (the tier's slots are ordinary
// memory here).  A failing brick source must leave the tier's tables
// consistent: a source that cannot be opened takes no slot, a failed read
// frees its slot, and the clock never indexes an empty slot's brick.
#include <cstdio>
#include <vector>

#include "vt_context.hpp"

namespace vt_rt {
thread_local std::string g_err;
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
}  // namespace vt_rt

static int g_bad = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            g_bad++;                                                    \
        }                                                               \
    } while (0)

// slot_of and brick_of are inverse maps over the used slots
static void check_tables(const HostTier& h)
{
    for (int64_t s = 0; s < h.capacity; ++s) {
        const int64_t b = h.brick_of[(size_t)s];
        if (b >= 0) CHECK(h.slot_of[(size_t)b] == (int32_t)s);
    }
    for (int64_t b = 0; b < h.nbricks; ++b) {
        const int32_t s = h.slot_of[(size_t)b];
        if (s >= 0) CHECK(s < h.capacity && h.brick_of[(size_t)s] == b);
    }
}

static void setup(HostTier& h, std::vector<float>& pool, const std::string& file, int64_t capacity)
{
    h.level_c = 2;   // 4^3 = 64 bricks of 8^3 cells in a 32^3 volume
    h.cube = 8;
    h.inc = 1;
    h.S = h.cube + 2 * h.inc;
    h.slot_elems = (size_t)h.S * h.S * h.S;
    h.nbricks = 64;
    h.capacity = capacity;
    pool.assign((size_t)capacity * h.slot_elems, -1.0f);
    h.pool = pool.data();
    h.pool_dev = pool.data();
    h.file = file;
    h.dataset = "value";
    h.slot_of.assign((size_t)h.nbricks, -1);
    h.brick_of.assign((size_t)capacity, -1);
    h.ref.assign((size_t)capacity, 0);
    h.pin.assign((size_t)capacity, 0);
}

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    const std::string good = argv[1], missing = argv[2];
    // a 32^3 volume: voxel (x, y, z) = x * 1024 + y * 32 + z
    std::vector<float> vol(32 * 32 * 32);
    for (size_t i = 0; i < vol.size(); ++i) vol[i] = (float)i;
    std::string err;
    CHECK(vt::h5_create_volume(good, "value", vol.data(), 32, 32, 32, 8, err));

    {   // on demand (5 slots for 64 bricks)
        HostTier h;
        std::vector<float> pool;
        setup(h, pool, missing, 5);
        CHECK(h.start(5, nullptr) == VT_OK && !h.full_mode);
        const float* src = nullptr;
        int64_t misses = 0;
        // a source that cannot be opened takes no slot
        CHECK(h.get(3, 1, &src, &misses) == VT_E_IO);
        CHECK(h.used == 0 && h.slot_of[3] == -1);
        check_tables(h);
        // the source works: the clock cycles every brick through the slots
        h.file = good;
        uint64_t epoch = 2;
        for (int round = 0; round < 20; ++round)
            for (int64_t b = 0; b < 64; b += 1 + (round % 3)) {
                CHECK(h.get(b, epoch++, &src, &misses) == VT_OK);
                int x, y, z;
                node_min((uint64_t)b + 64, 2, 5, x, y, z);
                // the slot's interior corner holds voxel (x, y, z)
                const float* p = src + ((size_t)h.inc * h.S + h.inc) * h.S + h.inc;
                CHECK(*p == (float)(x * 1024 + y * 32 + z));
                check_tables(h);
            }
        // a read that fails frees its slot; the next good read reuses it
        h.sync_reader.close();
        h.dataset = "missing_dataset";
        int64_t absent = -1;
        for (int64_t b = 0; b < 64 && absent < 0; ++b)
            if (h.slot_of[(size_t)b] < 0) absent = b;
        CHECK(absent >= 0);
        CHECK(h.get(absent, epoch++, &src, &misses) == VT_E_IO);
        CHECK(h.slot_of[(size_t)absent] == -1);
        check_tables(h);
        h.dataset = "value";
        for (int64_t b = 63; b >= 0; --b) {
            CHECK(h.get(b, epoch++, &src, &misses) == VT_OK);
            check_tables(h);
        }
        h.pool = nullptr;   // ordinary memory here: the destructor frees no pinned pool
    }
    {   // filled ahead by the reader thread from a file that cannot be opened
        HostTier h;
        std::vector<float> pool;
        setup(h, pool, missing, 64);
        CHECK(h.start(5, nullptr) == VT_OK && h.full_mode);
        const float* src = nullptr;
        int64_t misses = 0;
        CHECK(h.get(10, 1, &src, &misses) == VT_E_IO);
        if (h.reader.joinable()) h.reader.join();
        h.pool = nullptr;
    }
    {   // filled ahead from the good file: every brick readable
        HostTier h;
        std::vector<float> pool;
        setup(h, pool, good, 64);
        CHECK(h.start(5, nullptr) == VT_OK && h.full_mode);
        const float* src = nullptr;
        int64_t misses = 0;
        for (int64_t b = 0; b < 64; ++b) CHECK(h.get(b, 1, &src, &misses) == VT_OK);
        if (h.reader.joinable()) h.reader.join();
        h.pool = nullptr;
    }
    std::printf("host tier failure paths: %d failed checks\n", g_bad);
    return g_bad == 0 ? 0 : 1;
}
