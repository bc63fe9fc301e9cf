// Params — the solver parameters of the reference's params.ini (include/params.hpp:7-26 of the
// reference), with standard-library types in place of cv::Vec3i / cv::Vec3f.
#pragma once

#include <array>

#include "fastreg_c.h"

struct Params {
    std::array<int, 3> volume_dims{{0, 0, 0}};   /* volume dimensions in voxels */
    std::array<float, 3> volume_size{{1.f, 1.f, 1.f}};   /* volume size (read, not used by the solver) */

    int reset_iter = -1; /* iteration when to reset t in order to avoid underdamping */
    int verbosity = 0;   /* solver verbosity */

    int no_labels = 6;          /* no. of segmentations */
    int max_iter = 2048;        /* max. no. of iterations of the solver */
    float max_update_norm = -1e-10f, tau = 0.2f, rho_0 = 1.f, alpha = 0.004f, w_reg = 0.2f, p = 2.f, C = 0.25f;

    int no_voxels() const { return volume_dims[0] * volume_dims[1] * volume_dims[2]; }

    std::array<float, 3> voxel_sizes() const {
        return {{volume_size[0] / volume_dims[0], volume_size[1] / volume_dims[1], volume_size[2] / volume_dims[2]}};
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
