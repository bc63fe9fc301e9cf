// fastreg::cuda::Solver and the device containers it works on, under the reference's names and
// signatures: Solver (include/solver.hpp:60-113), Volume (include/cuda/volume.hpp:19-46),
// VectorField and DeformationField (include/cuda/vector_fields.hpp:21-67), ScalarField
// (include/cuda/scalar_fields.hpp:20-65).  The namespace keeps the reference's name `cuda` so that
// callers written against the reference compile unchanged; the code behind it is HIP on gfx950,
// reached through the C-ABI of fastreg_c.h (header-only; link libfastreg_hip.so).
//
// Layouts differ from the reference's where that saves HBM traffic: a Volume holds float
// intensities and int16 labels as two arrays (the reference packs a float2 per voxel); vector
// fields keep the reference's float4 (x, y, z, 0) per voxel.
#pragma once

#include <cmath>
#include <cstdint>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fastreg_c.h"
#include "params.hpp"

namespace fastreg {

// C-ABI status -> exception (the reference prints and exits; callers here decide)
inline void check(int rc) {
    if (rc != FR_OK) throw std::runtime_error(std::string("fastreg: ") + fr_last_error());
}

namespace cuda {

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
    void zero() { check(fr_dev_memset(p_, 0, n_ * sizeof(T))); }

private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

inline size_t voxels_of(const Vec3i &d) { return (size_t)d[0] * d[1] * d[2]; }

// An image (float intensities) and its int16 labels on the device (cuda::Volume,
// volume.hpp:19-46).  A new volume is cleared, as the reference's constructor does.
class Volume {
public:
    explicit Volume(const Vec3i &dims) : dims_(dims), img_(voxels_of(dims)), lab_(voxels_of(dims)) { clear(); }

    Vec3i get_dims() const { return dims_; }

    // copy intensities and labels from device arrays of no_voxels elements (volume.cpp:30-35)
    void initialise(float *image_data, short *labels_data) {
        check(fr_memcpy_dtod(img_.data(), image_data, img_.size() * sizeof(float)));
        check(fr_memcpy_dtod(lab_.data(), labels_data, lab_.size() * sizeof(int16_t)));
    }
    // intensities and labels to 0 (volume.cpp:37-42)
    void clear() {
        img_.zero();
        lab_.zero();
    }
    // every voxel's intensity and label, x outermost, in the reference's format (volume.cpp:44-61)
    void print() const {
        std::vector<float> im(img_.size());
        std::vector<int16_t> lb(lab_.size());
        img_.download(im.data());
        lab_.download(lb.data());
        std::cout << "--- IMAGE ---" << std::endl;
        for (int i = 0; i < dims_[0]; i++)
            for (int j = 0; j < dims_[1]; j++)
                for (int k = 0; k < dims_[2]; k++) {
                    const size_t v = i + (size_t)dims_[0] * (j + (size_t)dims_[1] * k);
                    std::cout << "(x,y,z)=(" << i << ", " << j << ", " << k << "), intensity: " << im[v]
                              << ", label: " << (float)lb[v] << std::endl;
                }
    }

    float *image() { return img_.data(); }
    const float *image() const { return img_.data(); }
    int16_t *labels() { return lab_.data(); }
    const int16_t *labels() const { return lab_.data(); }

private:
    Vec3i dims_;
    DeviceBuffer<float> img_;
    DeviceBuffer<int16_t> lab_;
};

// a float4 (x, y, z, 0) field on the device (cuda::VectorField, vector_fields.hpp:21-48); a new
// field is cleared to 0
class VectorField {
public:
    explicit VectorField(const Vec3i &dims) : dims_(dims), buf_(4 * voxels_of(dims)) { buf_.zero(); }
    virtual ~VectorField() = default;

    Vec3i get_dims() const { return dims_; }
    float *data() { return buf_.data(); }
    const float *data() const { return buf_.data(); }
    void upload(const float *host_xyzw) { buf_.upload(host_xyzw); }
    void download(float *host_xyzw) const { buf_.download(host_xyzw); }

    // every component to 0 (vector_fields.cpp:24-29)
    virtual void clear() { buf_.zero(); }
    // the voxels with a component above 1e-5 in magnitude, x outermost (vector_fields.cpp:31-55)
    void print() const {
        std::vector<float> h(buf_.size());
        download(h.data());
        std::cout << "--- FIELD ---" << std::endl;
        for (int i = 0; i < dims_[0]; i++)
            for (int j = 0; j < dims_[1]; j++)
                for (int k = 0; k < dims_[2]; k++) {
                    const size_t v = 4 * (i + (size_t)dims_[0] * (j + (size_t)dims_[1] * k));
                    const float u = h[v], w1 = h[v + 1], w2 = h[v + 2];
                    if (std::fabs(u) > 1e-5f || std::fabs(w1) > 1e-5f || std::fabs(w2) > 1e-5f)
                        std::cout << "(x,y,z)=(" << i << ", " << j << ", " << k << "), (u,v,w)=(" << u << ", " << w1
                                  << "," << w2 << ")" << std::endl;
                }
    }
    // voxels with a NaN among x, y, z (vector_fields.cpp:57-79)
    int get_no_nans() const {
        std::vector<float> h(buf_.size());
        download(h.data());
        int n = 0;
        for (size_t v = 0; v < h.size(); v += 4)
            if (std::isnan(h[v]) || std::isnan(h[v + 1]) || std::isnan(h[v + 2])) ++n;
        return n;
    }

protected:
    Vec3i dims_;
    DeviceBuffer<float> buf_;
};

typedef VectorField VolumeGradient;
typedef VectorField Laplacian;
typedef VectorField PotentialGradient;
typedef VectorField Velocity;
typedef VectorField DVV;

// psi or psi_inv in absolute voxel coordinates (cuda::DeformationField, vector_fields.hpp:53-67);
// a new field is the identity
class DeformationField : public VectorField {
public:
    explicit DeformationField(const Vec3i &dims) : VectorField(dims) { clear(); }

    // the identity, (x, y, z, 0) at voxel (x, y, z) (vector_fields.cpp:87-92)
    void clear() override { check(fr_field_identity(dims_[0], dims_[1], dims_[2], data(), nullptr)); }

    // phi_psi = phi o *this: trilinear intensities and floor-corner labels at every voxel
    // (vector_fields.cpp:103-113)
    void apply(std::shared_ptr<Volume> phi, std::shared_ptr<Volume> phi_psi) const {
        const Vec3i d = phi->get_dims();
        check(fr_warp(d[0], d[1], d[2], phi->image(), phi->labels(), data(), phi_psi->image(), phi_psi->labels(),
                      nullptr));
    }

    // 48 fixed-point sweeps of psi_inv = -trilerp(*this, psi_inv), starting from psi_inv's current
    // values (vector_fields.cpp:94-101)
    void get_inverse(DeformationField &psi_inv) const {
        check(fr_get_inverse(dims_[0], dims_[1], dims_[2], data(), psi_inv.data(), nullptr));
    }

    // out = trilerp(field, *this) per voxel (interpolate_gradient, vector_fields.cu:194-225)
    void interpolate(const VectorField &field, VectorField &out) const {
        check(fr_interpolate_gradient(dims_[0], dims_[1], dims_[2], field.data(), out.data(), data(), nullptr));
    }
};

// a float scalar field on the device (cuda::ScalarField, scalar_fields.hpp:20-65), cleared to 0
class ScalarField {
public:
    explicit ScalarField(const Vec3i &dims) : dims_(dims), buf_(voxels_of(dims)) { buf_.zero(); }
    Vec3i get_dims() const { return dims_; }
    float *data() { return buf_.data(); }
    void clear() { buf_.zero(); }
    // ScalarField::sum (scalar_fields.cpp:29-33), accumulated in FP64
    double sum() const {
        double s = 0.0;
        check(fr_sum(buf_.data(), (long long)buf_.size(), &s, nullptr));
        return s;
    }

private:
    Vec3i dims_;
    DeviceBuffer<float> buf_;
};

typedef ScalarField Density;
typedef ScalarField Divergence;

class Solver {
public:
    // As the reference's Solver, an instance keeps its velocity from one estimate_psi to the next
    // (until clear()); the deformation it iterates on is the caller's psi and psi_inv.
    explicit Solver(const Params &params) : params_(params) {
        fr_params q = params.to_c();
        fr_ctx *c = nullptr;
        check(fr_create(&q, &c));
        ctx_ = c;
        check(fr_set_keep_state(ctx_, 1));
    }
    ~Solver() { fr_destroy(ctx_); }
    Solver(const Solver &) = delete;
    Solver &operator=(const Solver &) = delete;

    // Solver::clear (solver.cpp:69-83): the next estimate_psi starts from v = 0 (rho = rho_0 and
    // the derivative fields, which the reference also clears, are set or recomputed by every
    // estimate_psi here)
    void clear() { clear_ = true; }

    // estimate_psi (solver.cpp:83-115): register phi_n to phi_global, starting from the caller's
    // psi and psi_inv and the velocity the previous call left; on return
    // phi_global_psi_inv = phi_global o psi_inv, phi_n_psi = phi_n o psi, and psi / psi_inv hold the
    // deformation and its inverse (either may be null: the context's own is then used)
    void estimate_psi(std::shared_ptr<Volume> phi_global, std::shared_ptr<Volume> phi_global_psi_inv,
                      std::shared_ptr<Volume> phi_n, std::shared_ptr<Volume> phi_n_psi,
                      std::shared_ptr<DeformationField> psi, std::shared_ptr<DeformationField> psi_inv) {
        if (clear_) check(fr_set_keep_state(ctx_, 0));  // a fresh state: v = 0
        check(fr_solver_init(ctx_, phi_global->image(), phi_global->labels(), phi_n->image(), phi_n->labels()));
        if (clear_) {
            check(fr_set_keep_state(ctx_, 1));
            clear_ = false;
        }
        if (psi || psi_inv)
            check(fr_solver_set_state(ctx_, nullptr, nullptr, psi ? psi->data() : nullptr,
                                      psi_inv ? psi_inv->data() : nullptr));
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

    fr_stats stats() const {
        fr_stats st{};
        check(fr_get_stats(ctx_, &st));
        return st;
    }

    fr_ctx *handle() { return ctx_; }
    const Params &params() const { return params_; }

private:
    Params params_;
    fr_ctx *ctx_ = nullptr;
    bool clear_ = false;
};

}  // namespace cuda

// the same classes under the name of the platform they run on
namespace hip = cuda;

}  // namespace fastreg
