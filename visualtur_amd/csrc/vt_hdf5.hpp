// vt_hdf5.hpp -- HDF5 volume access (FileManager semantics).
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

namespace vt {

// One open float32 3-D dataset, C order [X][Y][Z].  Not shared between
// threads: each reader owns its own H5File (the file dataspace selection is
// per handle, unlike the reference's shared `spaceid`, FileManager.cu:125).
class H5Volume {
public:
    H5Volume() = default;
    ~H5Volume();
    H5Volume(const H5Volume&) = delete;
    H5Volume& operator=(const H5Volume&) = delete;

    // FileManager.cu:11-37, but read-only and returning an error instead of
    // exit(0).  chunk_cache_bytes sizes the raw-data chunk cache.
    bool open(const std::string& file, const std::string& dataset, size_t chunk_cache_bytes,
              std::string& err);
    void close();
    bool is_open() const { return dset_ >= 0; }
    const int64_t* dims() const { return dims_; }

    // FileManager.cu:86-167: fill the box [s,e) (C order, extent |e-s|) with
    // zeros, then read the part inside the dataset.
    bool read_box(const int s[3], const int e[3], float* out, std::string& err);

private:
    int64_t file_ = -1, dset_ = -1, space_ = -1;
    int64_t dims_[3] = {0, 0, 0};
};

// FileManager::CreateFile (FileManager.cu:169-243) with an optional chunk
// shape (0 = contiguous layout as in the reference).
bool h5_create_volume(const std::string& file, const std::string& dataset, const float* data,
                      int64_t X, int64_t Y, int64_t Z, int chunk, std::string& err);

// Streaming writer for volumes larger than one host buffer: create the
// dataset, then write x-slabs [x0, x0+nx).
class H5Writer {
public:
    ~H5Writer();
    bool create(const std::string& file, const std::string& dataset, int64_t X, int64_t Y,
                int64_t Z, int chunk, std::string& err);
    bool write_slab(int64_t x0, int64_t nx, const float* data, std::string& err);
    bool close(std::string& err);

private:
    int64_t file_ = -1, dset_ = -1, space_ = -1;
    int64_t dims_[3] = {0, 0, 0};
};

// A brick-major volume file: dataset `name` of shape (K, S, S, S) float32,
// record k the brick of Morton code codes[k] at levelC (its cube plus an
// inc-voxel halo, zero outside the volume: the FileManager box read of
// [min - inc, min - inc + S)), one record per chunk, so that a brick is one
// contiguous read; dataset `name`_codes holds the K codes (ascending), and
// attributes on `name` hold levelC, cubeDim, cubeInc and the volume dims.
class H5BrickWriter {
public:
    ~H5BrickWriter();
    bool create(const std::string& file, const std::string& name, const uint64_t* codes, int64_t K, int level_c,
                int cube, int inc, const int64_t dims[3], std::string& err);
    bool write(int64_t k0, int64_t n, const float* data, std::string& err);   // records [k0, k0 + n)
    bool close(std::string& err);

private:
    int64_t file_ = -1, dset_ = -1, space_ = -1;
    int64_t K_ = 0;
    int S_ = 0;
};

class H5BrickReader {
public:
    ~H5BrickReader();
    bool open(const std::string& file, const std::string& name, std::string& err);
    bool read(int64_t k, float* out, std::string& err);   // record k, S^3 floats
    void close();
    bool is_open() const { return dset_ >= 0; }
    int level_c = 0, cube = 0, inc = 0, S = 0;
    int64_t dims[3] = {0, 0, 0};
    std::vector<uint64_t> codes;

private:
    int64_t file_ = -1, dset_ = -1, space_ = -1;
};

}  // namespace vt
