/* FileManager.hpp -- HDF5 volume access (FileManager.hpp of the reference). */
#ifndef VT_FILEMANAGER_HPP
#define VT_FILEMANAGER_HPP

#include "config.hpp"

class FileManager
{
    private:
        vt_file *file;
    public:
        unsigned long long dims[3];   /* hsize_t dims[3] of the dataset */

        /* Opens read-only; throws std::runtime_error where the reference
         * calls exit(0) (FileManager.cu:11-37). */
        FileManager(const char *file_name, const char *dataset_name);
        ~FileManager();
        int readHDF5(float *volume);
        /* Box [s,e), zero outside the dataset, C order. */
        void readHDF5_Voxel_Array(int3 s, int3 e, float *data);
        int CreateFile(float *volume, char *newname, char *dataset_name, int x, int y, int z);
};

#endif
