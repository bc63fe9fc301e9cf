// A caller written against the reference's C++ interface (include/fastreg.hpp, include/solver.hpp,
// include/params.hpp, include/cuda/volume.hpp, include/cuda/vector_fields.hpp of the reference):
// the same class names, namespaces and calls, compiled against this repository's headers and
// run by tests/test_gpu_cpp_api.py.  Only the device allocation differs (fr_dev_malloc here,
// cudaMalloc in a CUDA build).
//
//   reference_caller X Y Z MAX_ITER NO_LABELS IN_DIR OUT_DIR
//
// IN_DIR holds tgt.f32, tgt_lab.i16, src.f32, src_lab.i16 (raw, x-fastest); OUT_DIR receives
//   fastreg_{img1,lab1,img2,lab2}, fastreg_psi   FastReg::operator() in place, getDeformationField()
//   fastreg2_{img1,lab1,img2,lab2}, fastreg2_psi a second operator() call on the same object and
//                                                the same inputs: it continues from the state the
//                                                first left (fastreg.cpp:58-75, solver.cu:317)
//   solver_{img1,lab1,img2,lab2}, solver_psi, solver_psi_inv   fastreg::cuda::Solver::estimate_psi
//   solver2_{img,lab,n_img,n_lab}, solver2_psi   a second estimate_psi on the same solver, from the
//                                                psi and psi_inv the first returned
//   solver3_{img,lab,n_img,n_lab}                estimate_psi after Solver::clear(), from the identity
//   apply_{img,lab}                              DeformationField::apply(phi_n, .) with the solver's psi
//   identity_{img,lab}, identity_psi             the same after DeformationField::clear()
//   cleared_{img,lab}                            a Volume after Volume::clear()
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "fastreg.hpp"

template <typename T>
static std::vector<T> read_raw(const std::string &path, size_t n) {
    std::vector<T> v(n);
    FILE *f = fopen(path.c_str(), "rb");
    if (!f || fread(v.data(), sizeof(T), n, f) != n) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
    return v;
}

template <typename T>
static void write_raw(const std::string &path, const std::vector<T> &v) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f || fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
}

template <typename T>
static T *device_copy(const std::vector<T> &h) {
    void *d = nullptr;
    fastreg::check(fr_dev_malloc(&d, h.size() * sizeof(T)));
    fastreg::check(fr_memcpy_htod(d, h.data(), h.size() * sizeof(T)));
    return static_cast<T *>(d);
}

template <typename T>
static std::vector<T> host_copy(const T *d, size_t n) {
    std::vector<T> h(n);
    fastreg::check(fr_memcpy_dtoh(h.data(), d, n * sizeof(T)));
    return h;
}

static void save_volume(const std::string &stem, const fastreg::cuda::Volume &vol, size_t n) {
    write_raw(stem + "_img", host_copy(vol.image(), n));
    write_raw(stem + "_lab", host_copy(vol.labels(), n));
}

int main(int argc, char **argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s X Y Z MAX_ITER NO_LABELS IN_DIR OUT_DIR\n", argv[0]);
        return 2;
    }
    Params params;
    params.volume_dims[0] = atoi(argv[1]);
    params.volume_dims[1] = atoi(argv[2]);
    params.volume_dims[2] = atoi(argv[3]);
    params.max_iter = atoi(argv[4]);
    params.no_labels = atoi(argv[5]);
    const std::string in = argv[6], out = argv[7];
    const size_t n = (size_t)params.no_voxels();

    const std::vector<float> tgt = read_raw<float>(in + "/tgt.f32", n), src = read_raw<float>(in + "/src.f32", n);
    const std::vector<short> tgt_lab = read_raw<short>(in + "/tgt_lab.i16", n),
                             src_lab = read_raw<short>(in + "/src_lab.i16", n);

    // FastReg::operator() overwrites its four device arrays in place (fastreg.cpp:58-75)
    {
        float *img1 = device_copy(tgt), *img2 = device_copy(src);
        short *lab1 = device_copy(tgt_lab), *lab2 = device_copy(src_lab);
        FastReg fastreg(params);
        for (const std::string stem : {"fastreg", "fastreg2"}) {
            fastreg::check(fr_memcpy_htod(img1, tgt.data(), n * sizeof(float)));
            fastreg::check(fr_memcpy_htod(lab1, tgt_lab.data(), n * sizeof(short)));
            fastreg::check(fr_memcpy_htod(img2, src.data(), n * sizeof(float)));
            fastreg::check(fr_memcpy_htod(lab2, src_lab.data(), n * sizeof(short)));
            if (!fastreg(img1, lab1, img2, lab2)) return 1;
            write_raw(out + "/" + stem + "_img1", host_copy(img1, n));
            write_raw(out + "/" + stem + "_lab1", host_copy(lab1, n));
            write_raw(out + "/" + stem + "_img2", host_copy(img2, n));
            write_raw(out + "/" + stem + "_lab2", host_copy(lab2, n));
            std::shared_ptr<fastreg::cuda::DeformationField> psi = fastreg.getDeformationField();
            std::vector<float> h(4 * n);
            psi->download(h.data());
            write_raw(out + "/" + stem + "_psi", h);
        }
        if (fastreg.getParams().max_iter != params.max_iter) return 1;
        fr_dev_free(img1), fr_dev_free(img2), fr_dev_free(lab1), fr_dev_free(lab2);
    }

    // the solver on the reference's containers (solver.hpp:64-78)
    float *d_tgt = device_copy(tgt), *d_src = device_copy(src);
    short *d_tgt_lab = device_copy(tgt_lab), *d_src_lab = device_copy(src_lab);
    auto phi_global = std::make_shared<fastreg::cuda::Volume>(params.volume_dims);
    auto phi_global_psi_inv = std::make_shared<fastreg::cuda::Volume>(params.volume_dims);
    auto phi_n = std::make_shared<fastreg::cuda::Volume>(params.volume_dims);
    auto phi_n_psi = std::make_shared<fastreg::cuda::Volume>(params.volume_dims);
    phi_global->initialise(d_tgt, d_tgt_lab);
    phi_global_psi_inv->initialise(d_tgt, d_tgt_lab);
    phi_n->initialise(d_src, d_src_lab);
    phi_n_psi->initialise(d_src, d_src_lab);
    auto psi = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
    auto psi_inv = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
    auto solver = std::make_shared<fastreg::cuda::Solver>(params);
    solver->clear();
    solver->estimate_psi(phi_global, phi_global_psi_inv, phi_n, phi_n_psi, psi, psi_inv);
    save_volume(out + "/solver", *phi_global_psi_inv, n);
    save_volume(out + "/solver_n", *phi_n_psi, n);
    std::vector<float> h(4 * n);
    psi->download(h.data());
    write_raw(out + "/solver_psi", h);
    psi_inv->download(h.data());
    write_raw(out + "/solver_psi_inv", h);

    // a second estimate_psi continues from the velocity the first left and the psi / psi_inv it
    // returned (here copies of them); after clear() the velocity starts from 0 again
    {
        auto psi2 = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
        auto psi_inv2 = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
        fastreg::check(fr_memcpy_dtod(psi2->data(), psi->data(), 4 * n * sizeof(float)));
        fastreg::check(fr_memcpy_dtod(psi_inv2->data(), psi_inv->data(), 4 * n * sizeof(float)));
        solver->estimate_psi(phi_global, phi_global_psi_inv, phi_n, phi_n_psi, psi2, psi_inv2);
        save_volume(out + "/solver2", *phi_global_psi_inv, n);
        save_volume(out + "/solver2_n", *phi_n_psi, n);
        psi2->download(h.data());
        write_raw(out + "/solver2_psi", h);
        auto psi3 = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
        auto psi_inv3 = std::make_shared<fastreg::cuda::DeformationField>(params.volume_dims);
        solver->clear();
        solver->estimate_psi(phi_global, phi_global_psi_inv, phi_n, phi_n_psi, psi3, psi_inv3);
        save_volume(out + "/solver3", *phi_global_psi_inv, n);
        save_volume(out + "/solver3_n", *phi_n_psi, n);
    }

    // DeformationField::apply with the estimated psi (vector_fields.cpp:103-113)
    auto warped = std::make_shared<fastreg::cuda::Volume>(phi_n->get_dims());
    psi->apply(phi_n, warped);
    save_volume(out + "/apply", *warped, n);

    // DeformationField::clear() is the identity: applying it copies the volume
    psi->clear();
    psi->apply(phi_n, warped);
    save_volume(out + "/identity", *warped, n);
    psi->download(h.data());
    write_raw(out + "/identity_psi", h);

    // Volume::clear()
    warped->clear();
    save_volume(out + "/cleared", *warped, n);

    fr_dev_free(d_tgt), fr_dev_free(d_src), fr_dev_free(d_tgt_lab), fr_dev_free(d_src_lab);
    printf("reference_caller ok\n");
    return 0;
}
