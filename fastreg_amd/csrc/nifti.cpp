// NIfTI-1 single-file reader/writer (see nifti.hpp).
#include "nifti.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace nifti {

static void swap_bytes(void *p, size_t n) {
    unsigned char *c = static_cast<unsigned char *>(p);
    for (size_t i = 0; i < n / 2; ++i) std::swap(c[i], c[n - 1 - i]);
}

static void swap_header(Header &h) {
    swap_bytes(&h.sizeof_hdr, 4);
    swap_bytes(&h.extents, 4);
    swap_bytes(&h.session_error, 2);
    for (auto &d : h.dim) swap_bytes(&d, 2);
    swap_bytes(&h.intent_p1, 4);
    swap_bytes(&h.intent_p2, 4);
    swap_bytes(&h.intent_p3, 4);
    swap_bytes(&h.intent_code, 2);
    swap_bytes(&h.datatype, 2);
    swap_bytes(&h.bitpix, 2);
    swap_bytes(&h.slice_start, 2);
    for (auto &p : h.pixdim) swap_bytes(&p, 4);
    swap_bytes(&h.vox_offset, 4);
    swap_bytes(&h.scl_slope, 4);
    swap_bytes(&h.scl_inter, 4);
    swap_bytes(&h.slice_end, 2);
    swap_bytes(&h.cal_max, 4);
    swap_bytes(&h.cal_min, 4);
    swap_bytes(&h.slice_duration, 4);
    swap_bytes(&h.toffset, 4);
    swap_bytes(&h.glmax, 4);
    swap_bytes(&h.glmin, 4);
    swap_bytes(&h.qform_code, 2);
    swap_bytes(&h.sform_code, 2);
    for (float *f : {&h.quatern_b, &h.quatern_c, &h.quatern_d, &h.qoffset_x, &h.qoffset_y, &h.qoffset_z}) swap_bytes(f, 4);
    for (int k = 0; k < 4; ++k) {
        swap_bytes(&h.srow_x[k], 4);
        swap_bytes(&h.srow_y[k], 4);
        swap_bytes(&h.srow_z[k], 4);
    }
}

static int bytes_of(int16_t dt) {
    switch (dt) {
        case DT_UINT8:
        case DT_INT8: return 1;
        case DT_INT16:
        case DT_UINT16: return 2;
        case DT_INT32:
        case DT_UINT32:
        case DT_FLOAT32: return 4;
        case DT_FLOAT64: return 8;
        default: return 0;
    }
}

// a value of type V from raw bytes (no alignment or aliasing assumptions)
template <typename V> static V load(const unsigned char *p) {
    V v;
    std::memcpy(&v, p, sizeof v);
    return v;
}

bool read(const std::string &path, Image &img, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    Header h;
    if (!f.read(reinterpret_cast<char *>(&h), sizeof h)) {
        err = path + ": short header";
        return false;
    }
    bool swapped = false;
    if (h.sizeof_hdr != 348) {
        swap_header(h);
        swapped = true;
        if (h.sizeof_hdr != 348) {
            err = path + ": not a NIfTI-1 file (sizeof_hdr != 348)";
            return false;
        }
    }
    if (std::strncmp(h.magic, "n+1", 3) != 0) {
        err = path + ": only single-file NIfTI-1 (.nii, magic n+1) is supported";
        return false;
    }
    int nd = h.dim[0];
    if (nd < 3 || nd > 7) {
        err = path + ": expected a 3-D image";
        return false;
    }
    for (int k = 4; k <= nd; ++k)
        if (h.dim[k] > 1) {
            err = path + ": expected a 3-D scalar image (found more than 3 non-singleton dims)";
            return false;
        }
    int bpv = bytes_of(h.datatype);
    if (!bpv) {
        err = path + ": unsupported NIfTI datatype " + std::to_string(h.datatype);
        return false;
    }
    for (int k = 0; k < 3; ++k) img.dims[k] = h.dim[k + 1];
    if (img.dims[0] < 1 || img.dims[1] < 1 || img.dims[2] < 1) {
        err = path + ": bad dimensions";
        return false;
    }
    size_t n = (size_t)img.dims[0] * img.dims[1] * img.dims[2];
    if (n > ((size_t)1 << 31)) {  // beyond what the solver accepts; refuse before allocating
        err = path + ": image of " + std::to_string(n) + " voxels is too large";
        return false;
    }
    if (!std::isfinite(h.vox_offset) || h.vox_offset < 348.f || h.vox_offset > 1e12f) {
        err = path + ": bad vox_offset";
        return false;
    }
    std::vector<unsigned char> raw(n * bpv);
    f.seekg((std::streamoff)h.vox_offset);
    if (!f.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)raw.size())) {
        err = path + ": truncated voxel data";
        return false;
    }
    img.data.resize(n);
    for (size_t i = 0; i < n; ++i) {
        unsigned char *p = raw.data() + i * bpv;
        if (swapped) swap_bytes(p, bpv);
        double v = 0;
        switch (h.datatype) {
            case DT_UINT8: v = load<uint8_t>(p); break;
            case DT_INT8: v = load<int8_t>(p); break;
            case DT_INT16: v = load<int16_t>(p); break;
            case DT_UINT16: v = load<uint16_t>(p); break;
            case DT_INT32: v = load<int32_t>(p); break;
            case DT_UINT32: v = load<uint32_t>(p); break;
            case DT_FLOAT32: v = load<float>(p); break;
            case DT_FLOAT64: v = load<double>(p); break;
        }
        if (h.scl_slope != 0.f && std::isfinite(h.scl_slope) && !(h.scl_slope == 1.f && h.scl_inter == 0.f))
            v = v * h.scl_slope + h.scl_inter;
        img.data[i] = (float)v;
    }
    img.hdr = h;   // native byte order; outputs are written in native order
    return true;
}

static bool write_raw(const std::string &path, const Header &like, int16_t dt, const void *values, size_t bytes,
                      std::string &err) {
    Header h = like;
    h.sizeof_hdr = 348;
    h.datatype = dt;
    h.bitpix = (int16_t)(8 * bytes_of(dt));
    h.dim[0] = 3;
    for (int k = 4; k < 8; ++k) h.dim[k] = 1;
    h.vox_offset = 352.f;
    h.scl_slope = 1.f;
    h.scl_inter = 0.f;
    std::memcpy(h.magic, "n+1\0", 4);
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    f.write(reinterpret_cast<const char *>(&h), sizeof h);
    const char ext[4] = {0, 0, 0, 0};
    f.write(ext, 4);
    f.write(reinterpret_cast<const char *>(values), (std::streamsize)bytes);
    if (!f) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

bool write_float(const std::string &path, const Header &like, const float *values, std::string &err) {
    size_t n = (size_t)like.dim[1] * like.dim[2] * like.dim[3];
    return write_raw(path, like, DT_FLOAT32, values, n * sizeof(float), err);
}

bool write_int16(const std::string &path, const Header &like, const int16_t *values, std::string &err) {
    size_t n = (size_t)like.dim[1] * like.dim[2] * like.dim[3];
    return write_raw(path, like, DT_INT16, values, n * sizeof(int16_t), err);
}

Header make_header(int X, int Y, int Z, int16_t datatype) {
    Header h;
    std::memset(&h, 0, sizeof h);
    h.sizeof_hdr = 348;
    h.regular = 'r';
    h.dim[0] = 3;
    h.dim[1] = (int16_t)X;
    h.dim[2] = (int16_t)Y;
    h.dim[3] = (int16_t)Z;
    for (int k = 4; k < 8; ++k) h.dim[k] = 1;
    h.datatype = datatype;
    h.bitpix = (int16_t)(8 * bytes_of(datatype));
    for (int k = 0; k < 8; ++k) h.pixdim[k] = 1.f;
    h.vox_offset = 352.f;
    h.scl_slope = 1.f;
    h.xyzt_units = 2;   // mm
    h.sform_code = 0;
    h.qform_code = 0;
    std::memcpy(h.magic, "n+1\0", 4);
    return h;
}

}  // namespace nifti
