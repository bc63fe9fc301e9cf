// fastreg::hip::Solver and the device field containers it works on — the reference's
// fastreg::cuda::Solver (include/solver.hpp:64-113), Volume (include/cuda/volume.hpp:19-46) and
// DeformationField (include/cuda/vector_fields.hpp:53-67), as thin C++ wrappers over the C-ABI
// in fastreg_c.h.  Header-only; link libfastreg_hip.so.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

#include "fastreg_c.h"
#include "params.hpp"

namespace fastreg {

// C-ABI status -> exception (the reference prints and exits; callers here decide)
inline void check(int rc) {
    if (rc != FR_OK) throw std::runtime_error(std::string("fastreg: ") + fr_last_error());
}

namespace hip {

// device buffer of `count` elements of T (RAII over fr_dev_malloc / fr_dev_free)
template <typename T>
class DeviceBuffer {
public:
    explicit DeviceBuffer(size_t count) : n_(count) {
        void *p = nullptr;
        check(fr_dev_malloc(&p, count * sizeof(T)));
        p_ = static_cast<T *>(p);
    }
    ~DeviceBuffer() {
        if (p_) fr_dev_free(p_);
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    T *data() { return p_; }
    const T *data() const { return p_; }
    size_t size() const { return n_; }
    void upload(const T *host) { check(fr_memcpy_htod(p_, host, n_ * sizeof(T))); }
    void download(T *host) const { check(fr_memcpy_dtoh(host, p_, n_ * sizeof(T))); }

private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// An image (float intensities) and its int16 labels on the device.  The reference packs both
// into one float2 volume; the solver here keeps them as separate arrays.
class Volume {
public:
    explicit Volume(const std::array<int, 3> &dims)
        : dims_(dims), img_((size_t)dims[0] * dims[1] * dims[2]), lab_((size_t)dims[0] * dims[1] * dims[2]) {}
    // copy intensities and labels from device arrays (Volume::initialise, volume.cpp:30-36)
    void initialise(const float *d_img, const int16_t *d_lab) {
        check(fr_memcpy_dtod(img_.data(), d_img, img_.size() * sizeof(float)));
        check(fr_memcpy_dtod(lab_.data(), d_lab, lab_.size() * sizeof(int16_t)));
    }
    float *image() { return img_.data(); }
    int16_t *labels() { return lab_.data(); }
    const std::array<int, 3> &get_dims() const { return dims_; }

private:
    std::array<int, 3> dims_;
    DeviceBuffer<float> img_;
    DeviceBuffer<int16_t> lab_;
};

// psi or psi_inv: float4 (x, y, z, 0) per voxel in absolute voxel coordinates
class DeformationField {
public:
    explicit DeformationField(const std::array<int, 3> &dims)
        : dims_(dims), buf_(4 * (size_t)dims[0] * dims[1] * dims[2]) {}
    float *data() { return buf_.data(); }
    const float *data() const { return buf_.data(); }
    const std::array<int, 3> &get_dims() const { return dims_; }
    void download(float *host_xyzw) const { buf_.download(host_xyzw); }

    // DeformationField::get_inverse (vector_fields.cpp:96-103): 48 fixed-point sweeps of
    // psi_inv = -trilerp(*this, psi_inv), starting from psi_inv's current values
    void get_inverse(DeformationField &psi_inv) const {
        check(fr_get_inverse(dims_[0], dims_[1], dims_[2], data(), psi_inv.data(), nullptr));
    }

    // interpolate_gradient (vector_fields.cu:194-225): out = trilerp(field, *this) per voxel
    void interpolate(const DeformationField &field, DeformationField &out) const {
        check(fr_interpolate_gradient(dims_[0], dims_[1], dims_[2], field.data(), out.data(), data(), nullptr));
    }

private:
    std::array<int, 3> dims_;
    DeviceBuffer<float> buf_;
};

// a scalar field on the device (ScalarField, scalar_fields.hpp:20-65)
class ScalarField {
public:
    explicit ScalarField(const std::array<int, 3> &dims) : dims_(dims), buf_((size_t)dims[0] * dims[1] * dims[2]) {}
    float *data() { return buf_.data(); }
    // ScalarField::sum (scalar_fields.cpp:29-33), accumulated in FP64
    double sum() const {
        double s = 0.0;
        check(fr_sum(buf_.data(), (long long)buf_.size(), &s, nullptr));
        return s;
    }

private:
    std::array<int, 3> dims_;
    DeviceBuffer<float> buf_;
};

class Solver {
public:
    explicit Solver(const Params &params) : params_(params) {
        fr_params q = params.to_c();
        fr_ctx *c = nullptr;
        check(fr_create(&q, &c));
        ctx_ = c;
    }
    ~Solver() { fr_destroy(ctx_); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;

    // every estimate_psi starts from a fresh state; kept for interface parity (solver.cpp:69-81)
    void clear() {}

    // estimate_psi (solver.cpp:83-115): register phi_n to phi_global; on return
    // phi_global_psi_inv = phi_global o psi_inv, phi_n_psi = phi_n o psi, and psi / psi_inv hold
    // the deformation and its inverse
    void estimate_psi(const std::shared_ptr<Volume> &phi_global, std::shared_ptr<Volume> phi_global_psi_inv,
                      const std::shared_ptr<Volume> &phi_n, std::shared_ptr<Volume> phi_n_psi,
                      std::shared_ptr<DeformationField> psi, std::shared_ptr<DeformationField> psi_inv) {
        check(fr_solver_init(ctx_, phi_global->image(), phi_global->labels(), phi_n->image(), phi_n->labels()));
        check(fr_solver_run(ctx_));
        check(fr_solver_finalize(ctx_, phi_global_psi_inv->image(), phi_global_psi_inv->labels(), phi_n_psi->image(),
                                 phi_n_psi->labels()));
        if (psi) check(fr_get_deformation(ctx_, psi->data()));
        if (psi_inv) check(fr_get_inverse_deformation(ctx_, psi_inv->data()));
    }

    // Reductor::voxel_max_energy (reductor.cpp:82-89) of the current state: (energy, voxel)
    std::pair<float, long long> voxel_max_energy() {
        float e = 0.f;
        long long v = 0;
        check(fr_voxel_max_energy(ctx_, &e, &v));
        return {e, v};
    }

    fr_ctx *handle() { return ctx_; }
    const Params &params() const { return params_; }

private:
    Params params_;
    fr_ctx *ctx_ = nullptr;
};

}  // namespace hip
}  // namespace fastreg
