// fastreg_app — the params.ini-driven registration CLI (the reference's src/apps/demo.cpp):
//
//   fastreg_app [--verbose | --vverbose] <params.ini> <target> <target labels> <source>
//               <source labels> <out dir>
//
// Reads the four NIfTI-1 volumes (float image, int16 labels), registers the source to the target
// on the GPU through the C-ABI, prints the reference's messages and writes
//   <out>/target_psi_inv_<source name>.nii, <out>/source_psi_<source name>.nii,
//   <out>/target_psi_inv_labels_<source name>.nii, <out>/source_psi_labels_<source name>.nii.
// Unlike the reference (which exits 0 on CUDA errors), every failure prints an `error:` line and
// exits nonzero.

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "fastreg.hpp"
#include "nifti.hpp"

namespace {

// file name from a path (demo.cpp:29-42)
std::string get_file_name(const std::string &s) {
    size_t i = s.rfind('/', s.length());
    if (i != std::string::npos) return s.substr(i + 1, s.length() - i);
    return s;
}

int die(const std::string &msg, int code = 1) {
    std::cerr << "error: " << msg << std::endl;
    return code;
}

struct Args {
    std::string params_path, target_path, target_labels_path, source_path, source_labels_path, out_path;
    bool verbose = false, vverbose = false, help = false;
    int positional = 0;
};

// the reference's flag parsing (demo.cpp:347-391)
Args parse_flags(const std::vector<std::string> &args) {
    Args a;
    for (const auto &arg : args) {
        if (arg == "-h" || arg == "--help") {
            a.help = true;
        } else if (arg == "--verbose") {
            a.verbose = true;
        } else if (arg == "--vverbose") {
            a.vverbose = true;
        } else {
            std::string *slots[] = {&a.params_path, &a.target_path, &a.target_labels_path, &a.source_path,
                                    &a.source_labels_path, &a.out_path};
            if (a.positional < 6) *slots[a.positional] = arg;
            a.positional++;
        }
    }
    return a;
}

void usage() {
    std::cout << "USAGE: fastreg [OPTIONS] <ini path> <target path> <target labels path> <source path> "
                 "<source labels path> <out path>"
              << std::endl;
    std::cout << "\t--help -h:    display help" << std::endl;
    std::cout << "\t--verbose: low verbosity" << std::endl;
    std::cout << "\t--vverbose: high verbosity" << std::endl;
}

}  // namespace

int main(int argc, char *argv[]) {
    std::vector<std::string> args(argv + 1, argv + argc);
    Args a = parse_flags(args);
    if (a.help) {
        usage();
        return EXIT_SUCCESS;
    }

    char info[512];
    if (fr_device_info(info, sizeof info) == FR_OK) std::cout << "Device: " << info << std::endl;

    if (argc < 6 || a.positional < 6)
        return die("incorrect number of arguments; please supply path to source and target data and .ini file; "
                   "exiting...",
                   255);

    fr_params q;
    fr_params_default(&q);
    if (a.verbose) q.verbosity = 1;
    else if (a.vverbose) q.verbosity = 2;
    if (fr_params_load_ini(a.params_path.c_str(), &q) != FR_OK) return die(fr_last_error());
    Params params = Params::from_c(q);

    std::cout << "-------" << std::endl;
    std::cout << "TARGET: " << a.target_path << ", " << a.target_labels_path << std::endl;
    std::cout << "SOURCE: " << a.source_path << ", " << a.source_labels_path << std::endl;
    std::cout << "-------" << std::endl;
    std::cout << "RUNNING NON-RIGID REGISTRATION" << std::endl;

    nifti::Image img1, lab1, img2, lab2;
    std::string err;
    if (!nifti::read(a.target_path, img1, err) || !nifti::read(a.target_labels_path, lab1, err) ||
        !nifti::read(a.source_path, img2, err) || !nifti::read(a.source_labels_path, lab2, err))
        return die(err);
    for (const nifti::Image *im : {&img1, &lab1, &img2, &lab2})
        for (int k = 0; k < 3; ++k)
            if (im->dims[k] != params.volume_dims[k])
                return die("volume dimensions " + std::to_string(im->dims[0]) + "x" + std::to_string(im->dims[1]) + "x" +
                           std::to_string(im->dims[2]) + " do not match VOL_DIMS in " + a.params_path);

    const size_t n = (size_t)params.no_voxels();
    std::vector<int16_t> hl1(n), hl2(n);
    for (size_t i = 0; i < n; ++i) {
        hl1[i] = (int16_t)lab1.data[i];
        hl2[i] = (int16_t)lab2.data[i];
    }

    try {
        FastReg fastreg(params);
        fastreg::cuda::DeviceBuffer<float> d_img1(n), d_img2(n);
        fastreg::cuda::DeviceBuffer<int16_t> d_lab1(n), d_lab2(n);
        d_img1.upload(img1.data.data());
        d_img2.upload(img2.data.data());
        d_lab1.upload(hl1.data());
        d_lab2.upload(hl2.data());

        auto begin = std::chrono::steady_clock::now();
        fastreg(d_img1.data(), reinterpret_cast<short *>(d_lab1.data()), d_img2.data(),
                reinterpret_cast<short *>(d_lab2.data()));
        auto end = std::chrono::steady_clock::now();
        std::cout << fastreg.log();
        std::cout << "--- REGISTRATION TIME: " << std::chrono::duration_cast<std::chrono::milliseconds>(end - begin).count()
                  << "ms" << std::endl;

        std::vector<float> o1(n), o2(n);
        std::vector<int16_t> ol1(n), ol2(n);
        d_img1.download(o1.data());
        d_img2.download(o2.data());
        d_lab1.download(ol1.data());
        d_lab2.download(ol2.data());

        const std::string src = get_file_name(a.source_path);
        const std::string out = a.out_path;
        if (!nifti::write_float(out + "/target_psi_inv_" + src + ".nii", img1.hdr, o1.data(), err) ||
            !nifti::write_float(out + "/source_psi_" + src + ".nii", img1.hdr, o2.data(), err) ||
            !nifti::write_int16(out + "/target_psi_inv_labels_" + src + ".nii", lab1.hdr, ol1.data(), err) ||
            !nifti::write_int16(out + "/source_psi_labels_" + src + ".nii", lab1.hdr, ol2.data(), err))
            return die(err);
        std::cout << "--- SAVED OUTPUT" << std::endl;
    } catch (const std::exception &e) {
        return die(e.what());
    }
    return 0;
}
