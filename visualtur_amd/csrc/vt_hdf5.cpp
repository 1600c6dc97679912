// vt_hdf5.cpp -- HDF5 volume access (FileManager semantics).
#include "vt_hdf5.hpp"

#include <hdf5.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>

namespace vt {

namespace {
struct ErrSilencer {   // keep libhdf5 from printing its own error stack
    H5E_auto2_t fn = nullptr;
    void* data = nullptr;
    ErrSilencer() { H5Eget_auto2(H5E_DEFAULT, &fn, &data); H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
    ~ErrSilencer() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};
}  // namespace

H5Volume::~H5Volume() { close(); }

void H5Volume::close()
{
    if (space_ >= 0) H5Sclose((hid_t)space_);
    if (dset_ >= 0) H5Dclose((hid_t)dset_);
    if (file_ >= 0) H5Fclose((hid_t)file_);
    space_ = dset_ = file_ = -1;
}

bool H5Volume::open(const std::string& file, const std::string& dataset, size_t chunk_cache_bytes,
                    std::string& err)
{
    ErrSilencer quiet;
    close();
    hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { err = "unable to open the requested file: " + file; return false; }
    file_ = f;
    hid_t dapl = H5Pcreate(H5P_DATASET_ACCESS);
    if (chunk_cache_bytes > 0) {
        // slots: a prime well above the number of cached chunks
        size_t nslots = std::max<size_t>(521, chunk_cache_bytes / 4096);
        nslots |= 1;
        H5Pset_chunk_cache(dapl, nslots, chunk_cache_bytes, 1.0);
    }
    hid_t d = H5Dopen2(f, dataset.c_str(), dapl);
    H5Pclose(dapl);
    if (d < 0) { err = "unable to open the requested data set: " + dataset; close(); return false; }
    dset_ = d;
    hid_t sp = H5Dget_space(d);
    if (sp < 0) { err = "unable to open the requested data space"; close(); return false; }
    space_ = sp;
    int nd = H5Sget_simple_extent_ndims(sp);
    if (nd != 3) { err = "dataset is not 3-D"; close(); return false; }
    hsize_t dd[3];
    H5Sget_simple_extent_dims(sp, dd, nullptr);
    for (int k = 0; k < 3; ++k) dims_[k] = (int64_t)dd[k];
    return true;
}

bool H5Volume::read_box(const int s[3], const int e[3], float* out, std::string& err)
{
    if (dset_ < 0) { err = "file not open"; return false; }
    hsize_t dim[3];
    for (int k = 0; k < 3; ++k) dim[k] = (hsize_t)std::abs(e[k] - s[k]);
    memset(out, 0, (size_t)(dim[0] * dim[1] * dim[2]) * sizeof(float));
    // completely outside (FileManager.cu:94): zeros, not an error
    for (int k = 0; k < 3; ++k)
        if (s[k] >= dims_[k] || e[k] < 0) return true;
    hsize_t off_out[3], off_in[3], cnt[3];
    for (int k = 0; k < 3; ++k) {
        off_out[k] = s[k] < 0 ? (hsize_t)(-s[k]) : 0;
        off_in[k] = s[k] < 0 ? 0 : (hsize_t)s[k];
        int64_t hi = e[k] > dims_[k] ? dims_[k] : e[k];
        if (hi <= (int64_t)off_in[k]) return true;   // empty intersection
        cnt[k] = (hsize_t)(hi - (int64_t)off_in[k]);
    }
    ErrSilencer quiet;
    hid_t fsp = H5Scopy((hid_t)space_);
    if (H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off_in, nullptr, cnt, nullptr) < 0) {
        H5Sclose(fsp); err = "defining hyperslab in the dataset"; return false;
    }
    hid_t msp = H5Screate_simple(3, dim, nullptr);
    if (H5Sselect_hyperslab(msp, H5S_SELECT_SET, off_out, nullptr, cnt, nullptr) < 0) {
        H5Sclose(msp); H5Sclose(fsp); err = "defining the memory hyperslab"; return false;
    }
    herr_t st = H5Dread((hid_t)dset_, H5T_IEEE_F32LE, msp, fsp, H5P_DEFAULT, out);
    H5Sclose(msp);
    H5Sclose(fsp);
    if (st < 0) { err = "reading data from hyperslab in the file"; return false; }
    return true;
}

bool h5_create_volume(const std::string& file, const std::string& dataset, const float* data,
                      int64_t X, int64_t Y, int64_t Z, int chunk, std::string& err)
{
    H5Writer w;
    if (!w.create(file, dataset, X, Y, Z, chunk, err)) return false;
    if (!w.write_slab(0, X, data, err)) return false;
    return w.close(err);
}

H5Writer::~H5Writer()
{
    std::string e;
    close(e);
}

bool H5Writer::create(const std::string& file, const std::string& dataset, int64_t X, int64_t Y,
                      int64_t Z, int chunk, std::string& err)
{
    ErrSilencer quiet;
    hid_t f = H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) { err = "unable to create the HDF5 file: " + file; return false; }
    file_ = f;
    hsize_t d[3] = {(hsize_t)X, (hsize_t)Y, (hsize_t)Z};
    dims_[0] = X; dims_[1] = Y; dims_[2] = Z;
    hid_t sp = H5Screate_simple(3, d, nullptr);
    space_ = sp;
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    if (chunk > 0) {
        hsize_t c[3] = {(hsize_t)std::min<int64_t>(chunk, X), (hsize_t)std::min<int64_t>(chunk, Y),
                        (hsize_t)std::min<int64_t>(chunk, Z)};
        H5Pset_chunk(dcpl, 3, c);
    }
    hid_t ds = H5Dcreate2(f, dataset.c_str(), H5T_IEEE_F32LE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Pclose(dcpl);
    if (ds < 0) { err = "unable to create the data set"; return false; }
    dset_ = ds;
    return true;
}

bool H5Writer::write_slab(int64_t x0, int64_t nx, const float* data, std::string& err)
{
    if (dset_ < 0) { err = "writer not open"; return false; }
    if (x0 < 0 || nx <= 0 || x0 + nx > dims_[0]) { err = "slab outside the dataset"; return false; }
    ErrSilencer quiet;
    hsize_t off[3] = {(hsize_t)x0, 0, 0};
    hsize_t cnt[3] = {(hsize_t)nx, (hsize_t)dims_[1], (hsize_t)dims_[2]};
    hid_t fsp = H5Scopy((hid_t)space_);
    H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
    hid_t msp = H5Screate_simple(3, cnt, nullptr);
    herr_t st = H5Dwrite((hid_t)dset_, H5T_IEEE_F32LE, msp, fsp, H5P_DEFAULT, data);
    H5Sclose(msp);
    H5Sclose(fsp);
    if (st < 0) { err = "unable to write data set"; return false; }
    return true;
}

bool H5Writer::close(std::string& err)
{
    bool ok = true;
    if (dset_ >= 0 && H5Dclose((hid_t)dset_) < 0) { err = "unable to close the data set"; ok = false; }
    if (space_ >= 0) H5Sclose((hid_t)space_);
    if (file_ >= 0 && H5Fclose((hid_t)file_) < 0) { err = "unable to close the file"; ok = false; }
    dset_ = space_ = file_ = -1;
    return ok;
}

// ------------------------------------------------------ brick-major files

namespace {

bool put_attr_i64(hid_t obj, const char* name, const int64_t* v, hsize_t n)
{
    hid_t sp = H5Screate_simple(1, &n, nullptr);
    hid_t a = H5Acreate2(obj, name, H5T_STD_I64LE, sp, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = a >= 0 && H5Awrite(a, H5T_NATIVE_INT64, v) >= 0;
    if (a >= 0) H5Aclose(a);
    H5Sclose(sp);
    return ok;
}

bool get_attr_i64(hid_t obj, const char* name, int64_t* v, hsize_t n)
{
    if (H5Aexists(obj, name) <= 0) return false;
    hid_t a = H5Aopen(obj, name, H5P_DEFAULT);
    if (a < 0) return false;
    hid_t sp = H5Aget_space(a);
    const bool ok = H5Sget_simple_extent_npoints(sp) == (hssize_t)n && H5Aread(a, H5T_NATIVE_INT64, v) >= 0;
    H5Sclose(sp);
    H5Aclose(a);
    return ok;
}

}  // namespace

H5BrickWriter::~H5BrickWriter()
{
    std::string e;
    close(e);
}

bool H5BrickWriter::create(const std::string& file, const std::string& name, const uint64_t* codes, int64_t K,
                           int level_c, int cube, int inc, const int64_t dims[3], std::string& err)
{
    ErrSilencer quiet;
    if (K <= 0) { err = "no bricks to write"; return false; }
    hid_t f = H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) { err = "unable to create the HDF5 file: " + file; return false; }
    file_ = f;
    S_ = cube + 2 * inc;
    K_ = K;
    hsize_t d[4] = {(hsize_t)K, (hsize_t)S_, (hsize_t)S_, (hsize_t)S_};
    hid_t sp = H5Screate_simple(4, d, nullptr);
    space_ = sp;
    hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
    hsize_t c[4] = {1, (hsize_t)S_, (hsize_t)S_, (hsize_t)S_};
    H5Pset_chunk(dcpl, 4, c);
    hid_t ds = H5Dcreate2(f, name.c_str(), H5T_IEEE_F32LE, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    H5Pclose(dcpl);
    if (ds < 0) { err = "unable to create the brick data set"; return false; }
    dset_ = ds;
    const int64_t meta[3] = {level_c, cube, inc};
    if (!put_attr_i64(ds, "levelC_cubeDim_cubeInc", meta, 3) || !put_attr_i64(ds, "dims", dims, 3)) {
        err = "unable to write the brick attributes";
        return false;
    }
    hsize_t kk = (hsize_t)K;
    hid_t csp = H5Screate_simple(1, &kk, nullptr);
    hid_t cds = H5Dcreate2(f, (name + "_codes").c_str(), H5T_STD_U64LE, csp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = cds >= 0 && H5Dwrite(cds, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, codes) >= 0;
    if (cds >= 0) H5Dclose(cds);
    H5Sclose(csp);
    if (!ok) { err = "unable to write the brick codes"; return false; }
    return true;
}

bool H5BrickWriter::write(int64_t k0, int64_t n, const float* data, std::string& err)
{
    if (dset_ < 0) { err = "writer not open"; return false; }
    if (k0 < 0 || n <= 0 || k0 + n > K_) { err = "records outside the data set"; return false; }
    ErrSilencer quiet;
    hsize_t off[4] = {(hsize_t)k0, 0, 0, 0};
    hsize_t cnt[4] = {(hsize_t)n, (hsize_t)S_, (hsize_t)S_, (hsize_t)S_};
    hid_t fsp = H5Scopy((hid_t)space_);
    H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
    hid_t msp = H5Screate_simple(4, cnt, nullptr);
    herr_t st = H5Dwrite((hid_t)dset_, H5T_IEEE_F32LE, msp, fsp, H5P_DEFAULT, data);
    H5Sclose(msp);
    H5Sclose(fsp);
    if (st < 0) { err = "unable to write bricks"; return false; }
    return true;
}

bool H5BrickWriter::close(std::string& err)
{
    bool ok = true;
    if (dset_ >= 0 && H5Dclose((hid_t)dset_) < 0) { err = "unable to close the data set"; ok = false; }
    if (space_ >= 0) H5Sclose((hid_t)space_);
    if (file_ >= 0 && H5Fclose((hid_t)file_) < 0) { err = "unable to close the file"; ok = false; }
    dset_ = space_ = file_ = -1;
    return ok;
}

H5BrickReader::~H5BrickReader() { close(); }

void H5BrickReader::close()
{
    if (space_ >= 0) H5Sclose((hid_t)space_);
    if (dset_ >= 0) H5Dclose((hid_t)dset_);
    if (file_ >= 0) H5Fclose((hid_t)file_);
    space_ = dset_ = file_ = -1;
}

bool H5BrickReader::open(const std::string& file, const std::string& name, std::string& err)
{
    ErrSilencer quiet;
    close();
    hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { err = "unable to open the HDF5 file: " + file; return false; }
    file_ = f;
    hid_t d = H5Dopen2(f, name.c_str(), H5P_DEFAULT);
    if (d < 0) { err = "unable to open the brick data set " + name; return false; }
    dset_ = d;
    hid_t sp = H5Dget_space(d);
    space_ = sp;
    hsize_t dd[4] = {0, 0, 0, 0};
    if (H5Sget_simple_extent_ndims(sp) != 4 || H5Sget_simple_extent_dims(sp, dd, nullptr) != 4 || dd[1] != dd[2] ||
        dd[1] != dd[3]) {
        err = "the brick data set is not (K, S, S, S)";
        return false;
    }
    int64_t meta[3];
    if (!get_attr_i64(d, "levelC_cubeDim_cubeInc", meta, 3) || !get_attr_i64(d, "dims", dims, 3)) {
        err = "the brick data set lacks its levelC/cubeDim/cubeInc and dims attributes";
        return false;
    }
    level_c = (int)meta[0];
    cube = (int)meta[1];
    inc = (int)meta[2];
    S = cube + 2 * inc;
    if ((hsize_t)S != dd[1]) { err = "brick side does not match cubeDim + 2 cubeInc"; return false; }
    hid_t c = H5Dopen2(f, (name + "_codes").c_str(), H5P_DEFAULT);
    if (c < 0) { err = "unable to open the brick codes " + name + "_codes"; return false; }
    hid_t csp = H5Dget_space(c);
    const hssize_t K = H5Sget_simple_extent_npoints(csp);
    H5Sclose(csp);
    bool ok = K == (hssize_t)dd[0];
    if (ok) {
        codes.assign((size_t)K, 0);
        ok = H5Dread(c, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, codes.data()) >= 0;
    }
    H5Dclose(c);
    if (!ok) { err = "unable to read the brick codes"; return false; }
    const uint64_t lo = 1ull << (3 * level_c), hi = 1ull << (3 * level_c + 3);
    for (size_t k = 0; k < codes.size(); ++k)
        if (codes[k] < lo || codes[k] >= hi || (k > 0 && codes[k] <= codes[k - 1])) {
            err = "brick codes are not ascending codes of level " + std::to_string(level_c);
            return false;
        }
    return true;
}

bool H5BrickReader::read(int64_t k, float* out, std::string& err)
{
    if (dset_ < 0) { err = "brick file not open"; return false; }
    if (k < 0 || k >= (int64_t)codes.size()) { err = "record outside the brick data set"; return false; }
    ErrSilencer quiet;
    hsize_t off[4] = {(hsize_t)k, 0, 0, 0};
    hsize_t cnt[4] = {1, (hsize_t)S, (hsize_t)S, (hsize_t)S};
    hid_t fsp = H5Scopy((hid_t)space_);
    H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, nullptr, cnt, nullptr);
    hid_t msp = H5Screate_simple(4, cnt, nullptr);
    herr_t st = H5Dread((hid_t)dset_, H5T_IEEE_F32LE, msp, fsp, H5P_DEFAULT, out);
    H5Sclose(msp);
    H5Sclose(fsp);
    if (st < 0) { err = "unable to read brick record " + std::to_string(k); return false; }
    return true;
}

}  // namespace vt
