// FastReg — the registration pipeline class of the reference (include/fastreg.hpp:12-38,
// src/fastreg/fastreg.cpp), as a thin C++ wrapper over the C-ABI in fastreg_c.h.  Header-only;
// link libfastreg_hip.so.
#pragma once

#include <memory>
#include <cstring>
#include <string>

#include "fastreg_c.h"
#include "params.hpp"
#include "solver.hpp"

class FastReg {
public:
    // As the reference's FastReg, an instance carries its solver state (v, psi, psi_inv) from one
    // operator() call to the next (fr_set_keep_state); construct a new one to start afresh.
    explicit FastReg(const Params &params) : params_(params) {
        fr_params q = params.to_c();
        fr_ctx *c = nullptr;
        fastreg::check(fr_create(&q, &c));
        ctx_ = c;
        fastreg::check(fr_set_keep_state(ctx_, 1));
        psi_ = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
    }
    ~FastReg() { fr_destroy(ctx_); }
    FastReg(const FastReg &) = delete;
    FastReg &operator=(const FastReg &) = delete;

    Params &getParams() { return params_; }

    // psi of the last registration (float4 per voxel, absolute voxel coordinates)
    std::shared_ptr<fastreg::cuda::DeformationField> getDeformationField() {
        if (ran_) fastreg::check(fr_get_deformation(ctx_, psi_->data()));
        return psi_;
    }

    // Register (image_data_2, labels_data_2) = source to (image_data_1, labels_data_1) = target,
    // all device pointers of no_voxels elements, and overwrite them in place with
    // target o psi_inv and source o psi (fastreg.cpp:58-75).  Throws on failure.
    bool operator()(float *image_data_1, short *labels_data_1, float *image_data_2, short *labels_data_2) {
        fastreg::check(fr_register(ctx_, image_data_1, reinterpret_cast<int16_t *>(labels_data_1), image_data_2,
                                   reinterpret_cast<int16_t *>(labels_data_2)));
        ran_ = true;
        return ++frame_counter_, true;
    }

    // messages of the last run in the reference's format
    std::string log() const {
        std::string s(1 << 20, '\0');
        if (fr_get_log(ctx_, &s[0], s.size()) != FR_OK) return std::string();
        s.resize(strlen(s.c_str()));
        return s;
    }

    fr_stats stats() const {
        fr_stats st{};
        fastreg::check(fr_get_stats(ctx_, &st));
        return st;
    }

    fr_ctx *handle() { return ctx_; }

private:
    int frame_counter_ = 0;
    bool ran_ = false;
    Params params_;
    fr_ctx *ctx_ = nullptr;
    std::shared_ptr<fastreg::cuda::DeformationField> psi_;
};
