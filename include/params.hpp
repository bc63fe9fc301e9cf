// Params — the solver parameters of the reference's params.ini (include/params.hpp:7-26 of the
// reference).  The reference types volume_dims / volume_size as OpenCV's cv::Vec3i / cv::Vec3f;
// here they are fastreg::Vec3i / Vec3f, small 3-vectors with element access that also convert
// from any indexable 3-vector (a cv::Vec3i, a std::array, an int[3]), so that a caller built
// against the reference (with or without OpenCV) assigns and reads them unchanged.
#pragma once

#include <utility>

#include "fastreg_c.h"

namespace fastreg {

template <typename T>
struct Vec3 {
    T val[3];

    Vec3() : val{T(0), T(0), T(0)} {}
    Vec3(T a, T b, T c) : val{a, b, c} {}
    // from any 3-vector with operator[] (cv::Vec3i, std::array<int, 3>, ...)
    template <typename V, typename = decltype(std::declval<const V &>()[0])>
    Vec3(const V &v) : val{T(v[0]), T(v[1]), T(v[2])} {}

    T &operator[](int i) { return val[i]; }
    const T &operator[](int i) const { return val[i]; }
    bool operator==(const Vec3 &o) const { return val[0] == o.val[0] && val[1] == o.val[1] && val[2] == o.val[2]; }
    bool operator!=(const Vec3 &o) const { return !(*this == o); }
};

typedef Vec3<int> Vec3i;
typedef Vec3<float> Vec3f;

}  // namespace fastreg

struct Params {
    fastreg::Vec3i volume_dims{0, 0, 0}; /* volume dimensions in voxels */
    fastreg::Vec3f volume_size{1.f, 1.f, 1.f}; /* volume size (read, not used by the solver) */

    int reset_iter = -1; /* iteration at which t restarts (-1: never) */
    int verbosity = 0;   /* solver verbosity: 0, 1 (--verbose) or 2 (--vverbose) */

    int no_labels = 6;   /* label values 0..no_labels-1 are counted in Dice */
    int max_iter = 2048; /* iteration limit of the solver */
    float max_update_norm = -1e-10f, tau = 0.2f, rho_0 = 1.f, alpha = 0.004f, w_reg = 0.2f, p = 2.f, C = 0.25f;

    int no_voxels() const { return volume_dims[0] * volume_dims[1] * volume_dims[2]; }

    fastreg::Vec3f voxel_sizes() const {
        return fastreg::Vec3f(volume_size[0] / volume_dims[0], volume_size[1] / volume_dims[1],
                              volume_size[2] / volume_dims[2]);
    }

    fr_params to_c() const {
        fr_params q;
        fr_params_default(&q);
        for (int k = 0; k < 3; ++k) {
            q.vol_dims[k] = volume_dims[k];
            q.vol_size[k] = volume_size[k];
        }
        q.no_labels = no_labels;
        q.max_iter = max_iter;
        q.max_update_norm = max_update_norm;
        q.rho_0 = rho_0;
        q.tau = tau;
        q.alpha = alpha;
        q.w_reg = w_reg;
        q.p = p;
        q.C = C;
        q.reset_iter = reset_iter;
        q.verbosity = verbosity;
        return q;
    }

    static Params from_c(const fr_params &q) {
        Params r;
        for (int k = 0; k < 3; ++k) {
            r.volume_dims[k] = q.vol_dims[k];
            r.volume_size[k] = q.vol_size[k];
        }
        r.no_labels = q.no_labels;
        r.max_iter = q.max_iter;
        r.max_update_norm = q.max_update_norm;
        r.rho_0 = q.rho_0;
        r.tau = q.tau;
        r.alpha = q.alpha;
        r.w_reg = q.w_reg;
        r.p = q.p;
        r.C = q.C;
        r.reset_iter = q.reset_iter;
        r.verbosity = q.verbosity;
        return r;
    }
};
