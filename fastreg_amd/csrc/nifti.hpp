// Minimal NIfTI-1 (.nii, single file) reader/writer for the CLI: the inputs and outputs of the
// registration app (the reference reads and writes them with ITK, src/apps/demo.cpp:155-257).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nifti {

// The 348-byte NIfTI-1 header, field for field.
#pragma pack(push, 1)
struct Header {
    int32_t sizeof_hdr;
    char data_type[10];
    char db_name[18];
    int32_t extents;
    int16_t session_error;
    char regular;
    char dim_info;
    int16_t dim[8];
    float intent_p1, intent_p2, intent_p3;
    int16_t intent_code;
    int16_t datatype;
    int16_t bitpix;
    int16_t slice_start;
    float pixdim[8];
    float vox_offset;
    float scl_slope, scl_inter;
    int16_t slice_end;
    char slice_code;
    char xyzt_units;
    float cal_max, cal_min;
    float slice_duration;
    float toffset;
    int32_t glmax, glmin;
    char descrip[80];
    char aux_file[24];
    int16_t qform_code, sform_code;
    float quatern_b, quatern_c, quatern_d;
    float qoffset_x, qoffset_y, qoffset_z;
    float srow_x[4], srow_y[4], srow_z[4];
    char intent_name[16];
    char magic[4];
};
#pragma pack(pop)
static_assert(sizeof(Header) == 348, "NIfTI-1 header must be 348 bytes");

enum Datatype : int16_t {
    DT_UINT8 = 2,
    DT_INT16 = 4,
    DT_INT32 = 8,
    DT_FLOAT32 = 16,
    DT_FLOAT64 = 64,
    DT_INT8 = 256,
    DT_UINT16 = 512,
    DT_UINT32 = 768,
};

struct Image {
    Header hdr{};
    int dims[3] = {0, 0, 0};
    std::vector<float> data;   // voxel values converted to float (scl_slope/inter applied when set)
};

// Read a .nii file; returns false and sets `err` on failure.  Only 3-D (or 3-D with trailing
// singleton dims) scalar images are accepted.
bool read(const std::string &path, Image &img, std::string &err);

// Write `values` (x-fastest) as a .nii with the geometry of `like`, as float32 or int16.
bool write_float(const std::string &path, const Header &like, const float *values, std::string &err);
bool write_int16(const std::string &path, const Header &like, const int16_t *values, std::string &err);

// A fresh header for an X x Y x Z image with unit spacing (used when no input header exists).
Header make_header(int X, int Y, int Z, int16_t datatype);

}  // namespace nifti
