/* vt_oracle_h5.c -- test infrastructure only: the reference's HDF5 box read
 * (FileManager.cu:11-37 open, :86-167 readHDF5_Voxel_Array) with libhdf5, as
 * the brick source of the CPU baseline (vto_render_cached). */
#include <hdf5.h>
#include <stdlib.h>
#include <string.h>

#include "vt_oracle.h"

struct vto_h5 {
    hid_t file, dset, space;
    hsize_t dims[3];
};

vto_h5 *vto_h5_open(const char *file, const char *dataset)
{
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return NULL;
    hid_t d = H5Dopen2(f, dataset, H5P_DEFAULT);
    if (d < 0) {
        H5Fclose(f);
        return NULL;
    }
    hid_t sp = H5Dget_space(d);
    if (sp < 0 || H5Sget_simple_extent_ndims(sp) != 3) {
        if (sp >= 0) H5Sclose(sp);
        H5Dclose(d);
        H5Fclose(f);
        return NULL;
    }
    vto_h5 *h = (vto_h5 *)malloc(sizeof *h);
    h->file = f;
    h->dset = d;
    h->space = sp;
    H5Sget_simple_extent_dims(sp, h->dims, NULL);
    return h;
}

void vto_h5_close(vto_h5 *h)
{
    if (!h) return;
    H5Sclose(h->space);
    H5Dclose(h->dset);
    H5Fclose(h->file);
    free(h);
}

/* FileManager.cu:86-167: zeros, then (unless the box is entirely outside)
 * the clipped hyperslab of the file into the offset hyperslab of the box. */
int vto_h5_box_reader(void *user, const int32_t s[3], const int32_t e[3], float *out)
{
    vto_h5 *h = (vto_h5 *)user;
    hsize_t dim[3] = {(hsize_t)abs(e[0] - s[0]), (hsize_t)abs(e[1] - s[1]), (hsize_t)abs(e[2] - s[2])};
    memset(out, 0, (size_t)(dim[0] * dim[1] * dim[2]) * sizeof(float));
    if (s[0] >= (int)h->dims[0] || s[1] >= (int)h->dims[1] || s[2] >= (int)h->dims[2] || e[0] < 0 || e[1] < 0 ||
        e[2] < 0)
        return 0;
    hsize_t off_out[3], off[3], dimR[3];
    for (int k = 0; k < 3; ++k) {
        off_out[k] = s[k] < 0 ? (hsize_t)(-s[k]) : 0;
        off[k] = s[k] < 0 ? 0 : (hsize_t)s[k];
        dimR[k] = e[k] > (int)h->dims[k] ? h->dims[k] - off[k] : (hsize_t)e[k] - off[k];
    }
    hid_t fsp = H5Scopy(h->space);   /* each read selects in its own copy of the file space */
    if (fsp < 0) return -1;
    int rc = 0;
    hid_t msp = H5Screate_simple(3, dim, NULL);
    if (msp < 0 || H5Sselect_hyperslab(fsp, H5S_SELECT_SET, off, NULL, dimR, NULL) < 0 ||
        H5Sselect_hyperslab(msp, H5S_SELECT_SET, off_out, NULL, dimR, NULL) < 0 ||
        H5Dread(h->dset, H5T_IEEE_F32LE, msp, fsp, H5P_DEFAULT, out) < 0)
        rc = -1;
    if (msp >= 0) H5Sclose(msp);
    H5Sclose(fsp);
    return rc;
}
