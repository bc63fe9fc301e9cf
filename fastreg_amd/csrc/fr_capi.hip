// C-ABI of the fastreg solver (include/fastreg_c.h): contexts, the iteration loop, the
// reference's messages, params.ini parsing and device-memory plumbing.

#include "../../include/fastreg_c.h"
#include "fr_internal.hpp"

#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

using frk::Dims;

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int frk::set_error(int code, const char *msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(FR_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)

extern "C" const char *fr_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------------

struct fr_ctx {
    fr_params prm{};
    Dims d{};
    int n = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    // state
    float *T = nullptr, *S = nullptr;
    int16_t *LT = nullptr, *LS = nullptr;
    float *rho[2] = {nullptr, nullptr};
    float *v[2] = {nullptr, nullptr};
    float *pinv[2] = {nullptr, nullptr};
    float *psi = nullptr;
    int st = frk::ST_F32;                  // storage of v and rho
    int arith = frk::ARITH_REF;            // arithmetic of pass A
    unsigned storage_gen = 0;              // bumped by every fr_set_storage / fr_set_arith (batches check it)
    unsigned *vr[2] = {nullptr, nullptr};  // 16-bit (v, rho) records, 2 unsigned per voxel
    int cur = 0;
    int zc[3] = {0, 8, 2};       // interior z-slices per block of pass A ([1]) and pass B ([2])
    int zc_auto[3] = {0, 8, 2};  // the automatic choice for the current volume and storage
    bool zc_user[3] = {false, false, false};  // set by fr_set_chunking (else automatic)
    bool initialised = false;
    // fr_solver_init keeps v, psi and psi_inv of an initialised context (new images, rho reset),
    // as the reference's FastReg keeps its solver state from one call to the next
    bool keep_state = false;
    // loop control and history
    frk::Ctl *ctl = nullptr;
    unsigned long long *maxpack = nullptr;  // max_iter x NSLOT slots
    double *energy = nullptr;               // 3 * max_iter: energy, e_data, e_reg
    double *partials = nullptr;
    int nparts = 0;
    // finalisation
    unsigned *dice_counts = nullptr;  // DICE_COPIES x (overlap[no_labels], voxels[no_labels], bad)
    float *det_parts = nullptr;
    int ndet = 0;
    int16_t *lab_tmp = nullptr;
    unsigned long long *max_slot = nullptr;  // fr_voxel_max_energy
    // the iteration loop of a run as a hipGraph, per buffer parity at its start (captured at the
    // first run from that parity; dropped when the storage or the chunking changes)
    hipGraph_t run_graph[2] = {nullptr, nullptr};
    hipGraphExec_t run_exec[2] = {nullptr, nullptr};
    unsigned long long *maxrun = nullptr;  // max_iter per-iteration maxima (the slots reduced)
    // pinned host copies of the control word, the per-iteration maxima and the energies, written
    // by the copies each run queues after its loop
    struct RunHost {
        frk::Ctl ctl;
        unsigned long long *mp;
        double *e;
    };
    RunHost *h_run = nullptr;  // hipHostMalloc'd: the struct, then max_iter maxima, 3 max_iter energies
    // host copies of the last run
    std::vector<float> h_max;
    std::vector<long long> h_argmax;
    std::vector<double> h_energy, h_edata, h_ereg;
    std::vector<long long> h_overlap, h_voxels;
    long long h_bad = 0;
    int run_start_buf = 1;
    fr_stats stats{};
    std::string log;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool have_run = false, have_final = false;
    // the last fr_solver_run was recorded into a caller's graph capture: every replay of that
    // graph reruns the loop, so the run's results are read from the device again at each request
    bool captured = false;
};

static int collect_run(fr_ctx *c);

static int check_params(const fr_params *p) {
    for (int k = 0; k < 3; ++k)
        if (p->vol_dims[k] < 3) return fail(FR_ERR_INVALID, "VOL_DIMS must be >= 3 along every axis (got %d %d %d)",
                                            p->vol_dims[0], p->vol_dims[1], p->vol_dims[2]);
    long long n = (long long)p->vol_dims[0] * p->vol_dims[1] * p->vol_dims[2];
    if ((long long)p->vol_dims[0] * p->vol_dims[1] >= (1LL << 24))
        return fail(FR_ERR_INVALID, "an x-y slice of %lld voxels is too large (X*Y must be < 2^24)",
                    (long long)p->vol_dims[0] * p->vol_dims[1]);
    // byte offsets into the packed 3-vector fields (12 B per voxel) are 32-bit buffer offsets
    if (12 * n >= (1LL << 31))
        return fail(FR_ERR_INVALID, "volume of %lld voxels is too large (12*N bytes must be < 2^31)", n);
    if (p->max_iter < 0) return fail(FR_ERR_INVALID, "MAX_ITER must be >= 0");
    if (p->no_labels < 1 || p->no_labels > 32768) return fail(FR_ERR_INVALID, "NO_LABELS must be in [1, 32768]");
    if (p->verbosity < 0 || p->verbosity > 2) return fail(FR_ERR_INVALID, "verbosity must be 0, 1 or 2");
    const float fv[6] = {p->rho_0, p->tau, p->alpha, p->w_reg, p->p, p->C};
    const char *fn[6] = {"RHO_0", "TAU", "ALPHA", "W_REG", "p", "C"};
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(fv[k])) return fail(FR_ERR_INVALID, "%s must be finite (got %g)", fn[k], (double)fv[k]);
    if (std::isnan(p->max_update_norm)) return fail(FR_ERR_INVALID, "MAX_UPDATE_NORM must not be NaN");
    return FR_OK;
}

extern "C" int fr_params_default(fr_params *p) {
    if (!p) return fail(FR_ERR_INVALID, "null params");
    memset(p, 0, sizeof *p);
    p->vol_dims[0] = 193;
    p->vol_dims[1] = 229;
    p->vol_dims[2] = 193;
    p->vol_size[0] = p->vol_size[1] = p->vol_size[2] = 1.f;
    p->no_labels = 6;
    p->max_iter = 2048;
    p->max_update_norm = -1e-10f;
    p->rho_0 = 1.f;
    p->tau = 0.2f;
    p->alpha = 0.004f;
    p->w_reg = 0.2f;
    p->p = 2.f;
    p->C = 0.25f;
    p->reset_iter = -1;
    p->verbosity = 0;
    return FR_OK;
}

static void drop_run_graphs(fr_ctx *c) {
    for (int k = 0; k < 2; ++k) {
        if (c->run_exec[k]) (void)hipGraphExecDestroy(c->run_exec[k]);
        if (c->run_graph[k]) (void)hipGraphDestroy(c->run_graph[k]);
        c->run_exec[k] = nullptr;
        c->run_graph[k] = nullptr;
    }
}

static void free_ctx(fr_ctx *c) {
    if (!c) return;
    drop_run_graphs(c);
    if (c->h_run) (void)hipHostFree(c->h_run);
    void *bufs[] = {c->T, c->S, c->LT, c->LS, c->rho[0], c->rho[1], c->v[0], c->v[1], c->pinv[0], c->pinv[1], c->psi,
                    c->ctl, c->maxpack, c->energy, c->partials, c->dice_counts, c->det_parts, c->lab_tmp,
                    c->vr[0], c->vr[1], c->max_slot, c->maxrun};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

template <typename T>
static int dalloc(T **p, size_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, (count ? count : 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(FR_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    return FR_OK;
}

// z-slices per block of the iteration kernels.  Both passes march long column chunks with the
// block's rows in lockstep and stage the planes around the current one in LDS, so their chunks
// are sized for the whole grid to be resident at once: every block then marches one long chunk,
// the halo planes a chunk re-reads are amortised over it and no second wave of blocks runs
// behind a partly idle first one.
static int chunk_for(Dims d, int xs, int ty, long long slots) {
    const long long tiles = (long long)((d.X - 2 + xs - 1) / xs) * ((d.Y - 2 + ty - 1) / ty);
    const int interior = d.Z - 2;
    long long chunks = slots / tiles;
    if (chunks < 1) chunks = 1;
    if (chunks > interior) chunks = interior;
    return (int)((interior + chunks - 1) / chunks);
}

// Pass B on a volume whose psi and v_new (24 B per voxel) exceed ~100 MB streams them from HBM,
// and each CU's time is about the planes its blocks march, counting the 3 planes a block stages
// before its first step.  Among the chunkings whose blocks are all resident at once with at least
// 2 blocks on the busiest CU (16 waves to hide latency), take the one with the fewest planes on
// the busiest CU: min over c of ceil(tiles c / CUs) x (ceil((Z-2) / c) + 3).  Measured: 30 slices
// at 182x218x182 (-1.5 us against filling every slot with 20) and 23 at 160x192x160 (-0.5 us);
// smaller volumes, served largely from the 256 MB Infinity Cache, run faster with every slot
// filled (128^3: 8 slices, 1.9 us faster than the 13 this rule would give), as pass A does.
static int chunk_for_b(Dims d, long long n_cu, long long occ, int storage) {
    // 16-bit storage reads v_new as 8-byte records and unpacks it: less bound by memory, and
    // faster with every slot filled (at 182x218x182, 1.8 us faster with 20 slices than with 30)
    if (storage != frk::ST_F32 || 24LL * d.X * d.Y * d.Z < 100LL * 1000 * 1000)
        return chunk_for(d, frk::XS_B, frk::TY_B, n_cu * occ);
    const long long tiles = (long long)((d.X - 2 + frk::XS_B - 1) / frk::XS_B) * ((d.Y - 2 + frk::TY_B - 1) / frk::TY_B);
    const int interior = d.Z - 2;
    long long best_cost = -1;
    int best_zc = 0;
    for (int ch = 1; ch <= interior; ++ch) {
        const long long per_cu = (tiles * ch + n_cu - 1) / n_cu;
        if (per_cu > occ) break;
        if (per_cu < 2) continue;
        const int zc = (interior + ch - 1) / ch;
        const long long cost = per_cu * (zc + 3);
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best_zc = zc;
        }
    }
    return best_zc > 0 ? best_zc : chunk_for(d, frk::XS_B, frk::TY_B, n_cu * occ);
}

static int auto_chunking(fr_ctx *c) {
    int dev = 0, cus = 0, occ_a = 0, occ_b = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(frk::pass_occupancy(1, &occ_a));
    HIP_TRY(frk::pass_occupancy(2, &occ_b));
    const long long n_cu = cus > 0 ? cus : 1;
    c->zc_auto[0] = 0;
    c->zc_auto[1] = chunk_for(c->d, frk::xs_a(c->d), frk::TY_A, n_cu * (occ_a > 0 ? occ_a : 1));
    c->zc_auto[2] = chunk_for_b(c->d, n_cu, occ_b > 0 ? occ_b : 1, c->st);
    for (int k = 0; k < 3; ++k)
        if (!c->zc_user[k]) c->zc[k] = c->zc_auto[k];
    return FR_OK;
}

extern "C" int fr_create(const fr_params *p, fr_ctx **out) {
    if (!p || !out) return fail(FR_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    fr_ctx *c = new fr_ctx();
    c->prm = *p;
    c->d = Dims{p->vol_dims[0], p->vol_dims[1], p->vol_dims[2]};
    c->n = p->vol_dims[0] * p->vol_dims[1] * p->vol_dims[2];
    size_t n = (size_t)c->n;
    int mi = p->max_iter > 0 ? p->max_iter : 1;
    c->nparts = frk::energy_partials_count(c->d);
    c->ndet = frk::min_det_partials_count(c->d);
    if ((rc = dalloc(&c->T, n)) || (rc = dalloc(&c->S, n)) || (rc = dalloc(&c->LT, n)) || (rc = dalloc(&c->LS, n)) ||
        (rc = dalloc(&c->rho[0], n)) || (rc = dalloc(&c->rho[1], n)) || (rc = dalloc(&c->v[0], 3 * n)) ||
        (rc = dalloc(&c->v[1], 3 * n)) || (rc = dalloc(&c->pinv[0], 3 * n)) || (rc = dalloc(&c->pinv[1], 3 * n)) ||
        (rc = dalloc(&c->psi, 3 * n)) || (rc = dalloc(&c->ctl, 1)) || (rc = dalloc(&c->maxpack, (size_t)mi * frk::NSLOT)) ||
        (rc = dalloc(&c->energy, 3 * (size_t)mi)) || (rc = dalloc(&c->partials, 2 * (size_t)c->nparts)) ||
        (rc = dalloc(&c->dice_counts, frk::DICE_COPIES * (2 * (size_t)p->no_labels + 1))) || (rc = dalloc(&c->det_parts, 2 * (size_t)c->ndet)) ||
        (rc = dalloc(&c->lab_tmp, n)) || (rc = dalloc(&c->max_slot, 1)) || (rc = dalloc(&c->maxrun, (size_t)mi))) {
        free_ctx(c);
        return rc;
    }
    {
        void *h = nullptr;
        const size_t hb = sizeof(fr_ctx::RunHost) + sizeof(unsigned long long) * mi + sizeof(double) * 3 * (size_t)mi;
        if (hipHostMalloc(&h, hb, hipHostMallocDefault) != hipSuccess) {
            free_ctx(c);
            return fail(FR_ERR_ALLOC, "hipHostMalloc(%zu bytes) failed", hb);
        }
        c->h_run = static_cast<fr_ctx::RunHost *>(h);
        c->h_run->mp = reinterpret_cast<unsigned long long *>(c->h_run + 1);
        c->h_run->e = reinterpret_cast<double *>(c->h_run->mp + mi);
    }
    hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        free_ctx(c);
        return fail(FR_ERR_HIP, "stream/event creation: %s", hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    if ((rc = auto_chunking(c))) {
        free_ctx(c);
        return rc;
    }
    *out = c;
    return FR_OK;
}

extern "C" int fr_destroy(fr_ctx *c) {
    if (!c) return FR_OK;
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return FR_OK;
}

extern "C" int fr_get_params(const fr_ctx *c, fr_params *p) {
    if (!c || !p) return fail(FR_ERR_INVALID, "null argument");
    *p = c->prm;
    return FR_OK;
}

extern "C" int fr_set_stream(fr_ctx *c, void *s) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return FR_OK;
}

extern "C" int fr_set_chunking(fr_ctx *c, int zc_a, int zc_b) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (zc_a < 0 || zc_b < 0) return fail(FR_ERR_INVALID, "z-chunk sizes must be >= 0 (0 = automatic)");
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->zc[1] = zc_a ? zc_a : c->zc_auto[1];
    c->zc[2] = zc_b ? zc_b : c->zc_auto[2];
    c->zc_user[1] = zc_a != 0;
    c->zc_user[2] = zc_b != 0;
    drop_run_graphs(c);  // the captured launches have the old tile grids
    return FR_OK;
}

extern "C" int fr_get_chunking(const fr_ctx *c, int *zc_a, int *zc_b) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (zc_a) *zc_a = c->zc[1];
    if (zc_b) *zc_b = c->zc[2];
    return FR_OK;
}

extern "C" int fr_set_storage(fr_ctx *c, int storage) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (storage != FR_STORAGE_F32 && storage != FR_STORAGE_BF16 && storage != FR_STORAGE_F16)
        return fail(FR_ERR_INVALID, "storage must be FR_STORAGE_F32, FR_STORAGE_BF16 or FR_STORAGE_F16");
    if (storage != FR_STORAGE_F32 && c->arith != frk::ARITH_REF)
        return fail(FR_ERR_INVALID, "the contracted arithmetic (FR_ARITH_FMA) runs on FP32 storage only");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (storage != FR_STORAGE_F32) {
        // allocated on first use; a failed second allocation leaves the first for the next call
        int rc;
        if (!c->vr[0] && (rc = dalloc(&c->vr[0], 2 * (size_t)c->n))) return rc;
        if (!c->vr[1] && (rc = dalloc(&c->vr[1], 2 * (size_t)c->n))) return rc;
    }
    c->st = storage == FR_STORAGE_BF16 ? frk::ST_BF16 : storage == FR_STORAGE_F16 ? frk::ST_F16 : frk::ST_F32;
    ++c->storage_gen;  // a batch laid out before this call holds the old buffers and kernels
    drop_run_graphs(c);  // and so do the context's own captured runs
    if (int rc = auto_chunking(c)) return rc;  // the automatic chunking depends on the storage
    c->initialised = false;  // the state must be initialised in the new storage
    c->have_run = c->have_final = false;
    return FR_OK;
}

extern "C" int fr_set_keep_state(fr_ctx *c, int keep) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (keep != 0 && keep != 1) return fail(FR_ERR_INVALID, "keep must be 0 or 1");
    c->keep_state = keep == 1;
    return FR_OK;
}

extern "C" int fr_get_keep_state(const fr_ctx *c, int *keep) {
    if (!c || !keep) return fail(FR_ERR_INVALID, "null argument");
    *keep = c->keep_state ? 1 : 0;
    return FR_OK;
}

extern "C" int fr_set_arith(fr_ctx *c, int arith) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (arith != FR_ARITH_REFERENCE && arith != FR_ARITH_FMA)
        return fail(FR_ERR_INVALID, "arith must be FR_ARITH_REFERENCE or FR_ARITH_FMA");
    if (arith == FR_ARITH_FMA && c->st != frk::ST_F32)
        return fail(FR_ERR_INVALID, "the contracted arithmetic (FR_ARITH_FMA) runs on FP32 storage only");
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int a = arith == FR_ARITH_FMA ? frk::ARITH_FMA : frk::ARITH_REF;
    if (a != c->arith) {
        c->arith = a;
        ++c->storage_gen;    // a batch laid out before this call must not run this context
        drop_run_graphs(c);  // the captured loops launch the other kernel
    }
    return FR_OK;
}

extern "C" int fr_get_arith(const fr_ctx *c, int *arith) {
    if (!c || !arith) return fail(FR_ERR_INVALID, "null argument");
    *arith = c->arith == frk::ARITH_FMA ? FR_ARITH_FMA : FR_ARITH_REFERENCE;
    return FR_OK;
}

extern "C" int fr_get_storage(const fr_ctx *c, int *storage) {
    if (!c || !storage) return fail(FR_ERR_INVALID, "null argument");
    *storage = c->st == frk::ST_BF16 ? FR_STORAGE_BF16 : c->st == frk::ST_F16 ? FR_STORAGE_F16 : FR_STORAGE_F32;
    return FR_OK;
}

// ---------------------------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------------------------

extern "C" int fr_solver_init(fr_ctx *c, const float *tgt, const int16_t *tgt_lab, const float *src,
                              const int16_t *src_lab) {
    if (!c || !tgt || !tgt_lab || !src || !src_lab) return fail(FR_ERR_INVALID, "null argument");
    if (c->keep_state && c->initialised) {
        // the state of the last run is the one kept: an eager run not yet collected settles its
        // buffer parity first (it may have stopped early), which a stream capture cannot wait for
        if (c->have_run && !c->captured) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            HIP_TRY(hipStreamIsCapturing(c->stream, &cs));
            if (cs != hipStreamCaptureStatusNone)
                return fail(FR_ERR_STATE, "fr_solver_init keeping the state of a run not yet collected, inside a "
                                          "stream capture: read the run's results (fr_get_stats) before capturing");
            if (int rc = collect_run(c)) return rc;
        }
        HIP_TRY(frk::launch_reinit(c->d, tgt, tgt_lab, src, src_lab, c->T, c->LT, c->S, c->LS, c->st, c->rho[0],
                                   c->rho[1], c->vr[0], c->vr[1], c->prm.rho_0, c->stream));
        c->have_run = c->captured;
        c->have_final = false;
        return FR_OK;
    }
    HIP_TRY(frk::launch_init(c->d, tgt, tgt_lab, src, src_lab, c->T, c->LT, c->S, c->LS, c->rho[0], c->rho[1], c->v[0],
                             c->v[1], c->pinv[0], c->pinv[1], c->psi, c->prm.rho_0, c->stream));
    if (c->st != frk::ST_F32) HIP_TRY(frk::launch_vr_init(c->st, c->vr[0], c->vr[1], c->n, c->prm.rho_0, c->stream));
    c->cur = 0;
    c->initialised = true;
    c->have_run = c->captured;  // a replay of the caller's graph follows
    c->have_final = false;
    return FR_OK;
}

static void iter_args(fr_ctx *c, frk::IterArgs &a) {
    a.d = c->d;
    a.nbx = a.nby = a.nbz = a.nblocks = a.zc = a.xs = a.ty = 0;  // set per launch
    a.rho_o = c->rho[c->cur];
    a.rho_n = c->rho[c->cur ^ 1];
    a.v_o = c->v[c->cur];
    a.v_n = c->v[c->cur ^ 1];
    if (c->st != frk::ST_F32) {  // the 16-bit (v, rho) records
        a.v_o = reinterpret_cast<const float *>(c->vr[c->cur]);
        a.v_n = reinterpret_cast<float *>(c->vr[c->cur ^ 1]);
        a.rho_o = nullptr;
        a.rho_n = nullptr;
    }
    a.pinv_o = c->pinv[c->cur];
    a.pinv_n = c->pinv[c->cur ^ 1];
    a.psi = c->psi;
    a.S = c->S;
    a.T = c->T;
    a.alpha = c->prm.alpha;
    a.tau = c->prm.tau;
    a.w_reg = c->prm.w_reg;
    a.thr = c->prm.max_update_norm;
    a.ctl = c->ctl;
    a.maxpack = c->maxpack;
    a.energy = c->energy;
    a.check_energy = 0;
    a.check_conv = c->prm.max_update_norm >= 0.f;  // a norm is >= 0: a negative threshold never stops
    a.iter = 1;
}

// per-iteration scalars of update_velocity_kernel (solver.cu:198-199): (p+1)/t and
// C * p^2 * t^(p-2), in float as the reference computes them
static void vel_scalars(const fr_params &p, float t, float *m1t, float *m2) {
    float m1 = p.p + 1.f;
    *m1t = m1 / t;
    *m2 = p.C * powf(p.p, 2.f) * powf(t, p.p - 2.f);
}

static bool verbose_now(const fr_params &p, int iter) {
    return (p.verbosity == 1 && (iter == 1 || iter % 100 == 0 || iter == p.max_iter)) || p.verbosity == 2;
}

// one iteration's launches: [energy], pass A, pass B
static int launch_iteration(fr_ctx *c, int iter, float t, bool with_energy) {
    frk::IterArgs a;
    iter_args(c, a);
    a.iter = iter;
    vel_scalars(c->prm, t, &a.m1t, &a.m2);
    if (with_energy) {
        HIP_TRY(frk::launch_energy(c->d, c->S, c->T, c->psi, c->partials, c->nparts, c->stream));
        double *e = c->energy;
        int mi = c->prm.max_iter > 0 ? c->prm.max_iter : 1;
        HIP_TRY(frk::launch_energy_finalize(c->partials, c->nparts, (double)c->prm.w_reg, e + mi + (iter - 1),
                                            e + 2 * mi + (iter - 1), e + (iter - 1), c->stream));
        a.check_energy = 1;
    }
    HIP_TRY(frk::launch_pass_a(a, c->zc[1], c->st, c->stream, c->arith));
    HIP_TRY(frk::launch_pass_b(a, c->zc[2], c->st, c->stream));
    c->cur ^= 1;
    return FR_OK;
}

static int reset_control(fr_ctx *c) {
    int mi = c->prm.max_iter > 0 ? c->prm.max_iter : 1;
    // one kernel, not memsets: a caller capturing fr_solver_run gets no memset nodes in its graph
    HIP_TRY(frk::launch_zero3(c->ctl, sizeof(frk::Ctl), c->maxpack, sizeof(unsigned long long) * mi * frk::NSLOT,
                              c->energy, sizeof(double) * 3 * mi, c->stream));
    return FR_OK;
}

static unsigned long long slots_max(const unsigned long long *sl) {
    unsigned long long m = 0;
    for (int j = 0; j < frk::NSLOT; ++j) m = sl[j] > m ? sl[j] : m;
    return m;
}

static void unpack_max(unsigned long long mp, float *m, long long *idx) {
    float nr = 0.f;
    unsigned bits = (unsigned)(mp >> 32);
    memcpy(&nr, &bits, 4);
    *m = nr;
    *idx = (nr > 0.f) ? (long long)(0xFFFFFFFFu - (unsigned)(mp & 0xFFFFFFFFu)) : 0;
}

extern "C" int fr_solver_step(fr_ctx *c, float t, float *max_norm, long long *max_voxel) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_solver_step before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    c->captured = false;
    if ((rc = reset_control(c))) return rc;
    if ((rc = launch_iteration(c, 1, t, false))) return rc;
    unsigned long long sl[frk::NSLOT];
    HIP_TRY(hipMemcpyAsync(sl, c->maxpack, sizeof sl, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    unsigned long long mp = slots_max(sl);
    float m;
    long long idx;
    unpack_max(mp, &m, &idx);
    if (max_norm) *max_norm = m;
    if (max_voxel) *max_voxel = idx;
    return FR_OK;
}

static void logf_(std::string &s, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    s += buf;
}

// the launches of a run's iteration loop from the current buffer parity
static int run_launches(fr_ctx *c) {
    const fr_params &p = c->prm;
    float alpha = p.alpha;
    float t = alpha;
    for (int iter = 1; iter <= p.max_iter; ++iter) {
        if (int rc = launch_iteration(c, iter, t, verbose_now(p, iter))) return rc;
        // the t schedule of solver.cu:428-432
        t += alpha;
        if (p.reset_iter > 0 && (iter + 1) % p.reset_iter == 0) t = p.alpha;
    }
    return FR_OK;
}

extern "C" int fr_solver_run(fr_ctx *c) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_solver_run before fr_solver_init");
    int rc;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(c->stream, &cs));
    // an earlier run not yet collected (it may have stopped early) settles the buffer parity first;
    // a caller's capture cannot wait for it (its replays start after fr_solver_init, which writes
    // both buffers)
    if (cs == hipStreamCaptureStatusNone && c->have_run && (rc = collect_run(c))) return rc;
    // the control words are cleared by a kernel: eagerly before the context's own captured loop,
    // as a kernel node of a caller's capture
    if ((rc = reset_control(c))) return rc;
    const fr_params &p = c->prm;
    const int par = c->cur;
    c->run_start_buf = par;  // ping-pong index at the start of the run
    if (cs != hipStreamCaptureStatusNone) {
        // the caller is capturing this stream: the launches go into the caller's graph
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        if ((rc = run_launches(c))) return rc;
        c->captured = true;
    } else {
        c->captured = false;
        // The loop is the same sequence of launches on every run from one buffer parity: it is
        // captured once into a hipGraph and replayed (one launch instead of 2 per iteration).
        if (!c->run_exec[par]) {
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = run_launches(c);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            c->cur = par;
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(FR_ERR_HIP, "capture of the iteration loop: %s", hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiate(&c->run_exec[par], g, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                (void)hipGraphDestroy(g);
                c->run_exec[par] = nullptr;
                return fail(FR_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
            }
            c->run_graph[par] = g;
        }
        // the run's time (ev0 -> ev1) is the replay's, without the one-time capture above
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        HIP_TRY(hipGraphLaunch(c->run_exec[par], c->stream));
        c->cur = par ^ (p.max_iter & 1);  // as the launches leave it
    }
    HIP_TRY(hipEventRecord(c->ev1, c->stream));
    c->have_run = true;
    c->have_final = false;
    c->stats = fr_stats{};
    return FR_OK;
}

// collect the run's history and replay the reference's messages (solver.cu:330-426): the slots
// of each iteration reduced on the device, then the control word, maxima and energies copied into
// pinned host memory
static int collect_run(fr_ctx *c) {
    const fr_params &p = c->prm;
    int mi = p.max_iter > 0 ? p.max_iter : 1;
    if (p.max_iter > 0) {
        HIP_TRY(frk::launch_slots_reduce(c->maxpack, p.max_iter, c->maxrun, c->stream));
        HIP_TRY(hipMemcpyAsync(c->h_run->mp, c->maxrun, sizeof(unsigned long long) * mi, hipMemcpyDeviceToHost,
                               c->stream));
    }
    HIP_TRY(hipMemcpyAsync(&c->h_run->ctl, c->ctl, sizeof(frk::Ctl), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->h_run->e, c->energy, sizeof(double) * 3 * mi, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
        // events recorded inside a caller's graph may not time: the run's time is then unknown
        if (!c->captured) return fail(FR_ERR_HIP, "hipEventElapsedTime of the run failed");
        (void)hipGetLastError();
        ms = 0.f;
    }
    const frk::Ctl ctl = c->h_run->ctl;
    const unsigned long long *mp = c->h_run->mp;
    const double *e = c->h_run->e;

    int done = p.max_iter, reason = 0;
    if (ctl.stopped) {
        done = ctl.stop_iter;
        reason = ctl.stop_reason;
    } else if (p.max_iter > 0) {
        float m;
        long long idx;
        unpack_max(mp[p.max_iter - 1], &m, &idx);
        if (m <= p.max_update_norm) reason = 1;
    }
    // the ping-pong index after `done` completed iterations
    c->cur = c->run_start_buf ^ (done & 1);

    c->h_max.assign(done, 0.f);
    c->h_argmax.assign(done, 0);
    for (int k = 0; k < done; ++k) unpack_max(mp[k], &c->h_max[k], &c->h_argmax[k]);
    int ne = (reason == 2) ? done + 1 : done;
    c->h_energy.assign(ne, std::numeric_limits<double>::infinity());
    c->h_edata.assign(ne, 0.0);
    c->h_ereg.assign(ne, 0.0);
    for (int k = 0; k < ne && k < mi; ++k)
        if (verbose_now(p, k + 1)) {
            c->h_energy[k] = e[k];
            c->h_edata[k] = e[mi + k];
            c->h_ereg[k] = e[2 * mi + k];
        }

    std::string &L = c->log;
    L.clear();
    double last_e = std::numeric_limits<double>::infinity();
    for (int k = 1; k <= ne && k <= p.max_iter; ++k) {
        if (k == 1 || k % 100 == 0) logf_(L, "iter. no. %d\n", k);
        bool vn = verbose_now(p, k);
        if (vn) {
            last_e = c->h_energy[k - 1];
            logf_(L, "data energy + w_reg * reg energy = %g + %g * %g = %g\n", c->h_edata[k - 1], (double)p.w_reg,
                  c->h_ereg[k - 1], last_e);
        }
        if (reason == 2 && k == ne) break;
        if (vn) {
            long long am = c->h_argmax[k - 1];
            long long XY = (long long)c->d.X * c->d.Y;
            long long zz = am / XY, yy = (am - zz * XY) / c->d.X, xx = am - c->d.X * (yy + c->d.Y * zz);
            logf_(L, "max. update norm %g at voxel (%lld, %lld, %lld)\n", (double)c->h_max[k - 1], xx, yy, zz);
        }
        if (reason == 1 && k == done) {
            logf_(L, "SOLVER CONVERGED AFTER %d ITERATIONS\n", k);
            break;
        }
        if (k == p.max_iter) {
            logf_(L, "SOLVER REACHED MAX. NO. OF ITERATIONS WITHOUT CONVERGING\n");
            break;
        }
    }
    fr_stats &st = c->stats;
    st.iterations = done;
    st.stop_reason = reason;
    st.last_max_update = done > 0 ? c->h_max[done - 1] : 0.f;
    st.last_max_voxel = done > 0 ? c->h_argmax[done - 1] : 0;
    st.last_energy = last_e;
    st.solver_ms = ms;
    st.min_det_true = st.min_det_ref = std::numeric_limits<double>::quiet_NaN();
    st.dice_all = std::numeric_limits<double>::quiet_NaN();
    c->have_run = false;
    return FR_OK;
}

extern "C" int fr_solver_finalize(fr_ctx *c, float *img1, int16_t *lab1, float *img2, int16_t *lab2) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_solver_finalize before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    const fr_params &p = c->prm;
    // S o psi labels drive Dice; write them to the caller's buffer or a scratch one
    int16_t *ls = lab2 ? lab2 : c->lab_tmp;
    HIP_TRY(frk::launch_warp(c->d, c->S, c->LS, c->psi, 3, img2, ls, c->stream));
    if (img1 || lab1) HIP_TRY(frk::launch_warp(c->d, c->T, c->LT, c->pinv[c->cur], 3, img1, lab1, c->stream));
    const size_t ncnt = 2 * (size_t)p.no_labels + 1;
    HIP_TRY(hipMemsetAsync(c->dice_counts, 0, sizeof(unsigned) * frk::DICE_COPIES * ncnt, c->stream));
    HIP_TRY(frk::launch_dice(c->d, ls, c->LT, p.no_labels, c->dice_counts, c->stream));
    HIP_TRY(frk::launch_min_det(c->d, c->psi, c->det_parts, c->det_parts + c->ndet, c->stream));
    std::vector<unsigned> cnt_copies(frk::DICE_COPIES * ncnt), cnt(ncnt, 0u);
    std::vector<float> dp(2 * (size_t)c->ndet);
    HIP_TRY(hipMemcpyAsync(cnt_copies.data(), c->dice_counts, sizeof(unsigned) * cnt_copies.size(),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(dp.data(), c->det_parts, sizeof(float) * dp.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < frk::DICE_COPIES; ++k)
        for (size_t j = 0; j < ncnt; ++j) cnt[j] += cnt_copies[k * ncnt + j];
    c->h_overlap.assign(p.no_labels, 0);
    c->h_voxels.assign(p.no_labels, 0);
    for (int l = 0; l < p.no_labels; ++l) {
        c->h_overlap[l] = cnt[l];
        c->h_voxels[l] = cnt[p.no_labels + l];
    }
    c->h_bad = cnt[2 * p.no_labels];
    float mt = INFINITY, mr = INFINITY;
    for (int j = 0; j < c->ndet; ++j) {
        mt = fminf(mt, dp[j]);
        mr = fminf(mr, dp[c->ndet + j]);
    }
    c->stats.min_det_true = mt;
    c->stats.min_det_ref = mr;
    // Dice messages in the reference's float arithmetic (solver.cu:37-47)
    std::string &L = c->log;
    logf_(L, "POST-REGISTRATION\nDICE COEFFICIENT\n");
    float num = 0.f, den = 0.f;
    long long onum = 0, oden = 0;
    for (int l = 1; l < p.no_labels; ++l) {
        num += (float)c->h_overlap[l];
        den += (float)c->h_voxels[l];
        onum += c->h_overlap[l];
        oden += c->h_voxels[l];
        logf_(L, "label %d: %g\n", l, (double)(2.f * (float)c->h_overlap[l] / (float)c->h_voxels[l]));
    }
    logf_(L, "all labels: %g\n", (double)(2.f * num / den));
    c->stats.dice_all = oden > 0 ? 2.0 * (double)onum / (double)oden : std::numeric_limits<double>::quiet_NaN();
    c->have_final = true;
    return FR_OK;
}

extern "C" int fr_register(fr_ctx *c, float *img1, int16_t *lab1, float *img2, int16_t *lab2) {
    if (!c || !img1 || !lab1 || !img2 || !lab2) return fail(FR_ERR_INVALID, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    if ((rc = fr_solver_init(c, img1, lab1, img2, lab2))) return rc;
    if ((rc = fr_solver_run(c))) return rc;
    if ((rc = fr_solver_finalize(c, img1, lab1, img2, lab2))) return rc;
    auto t1 = std::chrono::steady_clock::now();
    c->stats.total_ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    return FR_OK;
}

extern "C" int fr_get_stats(const fr_ctx *c0, fr_stats *st) {
    fr_ctx *c = const_cast<fr_ctx *>(c0);
    if (!c || !st) return fail(FR_ERR_INVALID, "null argument");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    *st = c->stats;
    return FR_OK;
}

extern "C" int fr_get_history(const fr_ctx *c0, int n, float *max_norm, double *energy) {
    fr_ctx *c = const_cast<fr_ctx *>(c0);
    if (!c) return fail(FR_ERR_INVALID, "null context");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    for (int k = 0; k < n; ++k) {
        if (max_norm) max_norm[k] = k < (int)c->h_max.size() ? c->h_max[k] : 0.f;
        if (energy) energy[k] = k < (int)c->h_energy.size() ? c->h_energy[k] : std::numeric_limits<double>::infinity();
    }
    return FR_OK;
}

extern "C" int fr_get_log(const fr_ctx *c0, char *buf, size_t len) {
    fr_ctx *c = const_cast<fr_ctx *>(c0);
    if (!c || !buf || !len) return fail(FR_ERR_INVALID, "null argument");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    size_t m = c->log.size() < len - 1 ? c->log.size() : len - 1;
    memcpy(buf, c->log.data(), m);
    buf[m] = 0;
    return FR_OK;
}

extern "C" int fr_get_dice(const fr_ctx *c, int n, double *dice, long long *overlap, long long *voxels) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->have_final) return fail(FR_ERR_STATE, "no finalised registration");
    for (int l = 0; l < n; ++l) {
        bool in = l < (int)c->h_overlap.size();
        long long o = in ? c->h_overlap[l] : 0, v = in ? c->h_voxels[l] : 0;
        if (overlap) overlap[l] = o;
        if (voxels) voxels[l] = v;
        if (dice) dice[l] = v ? 2.0 * (double)o / (double)v : std::numeric_limits<double>::quiet_NaN();
    }
    return FR_OK;
}

extern "C" int fr_solver_set_state(fr_ctx *c, const float *v4, const float *rho, const float *psi4,
                                   const float *pinv4) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_solver_set_state before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    size_t n = (size_t)c->n;
    if (c->st != frk::ST_F32) {
        if (v4 || rho) HIP_TRY(frk::launch_vr_set(c->st, c->vr[0], c->vr[1], v4, rho, c->n, c->stream));
    } else {
        if (v4) HIP_TRY(frk::launch_pack4to3(v4, c->v[0], c->v[1], c->n, c->stream));
        if (rho) {
            HIP_TRY(hipMemcpyAsync(c->rho[0], rho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
            HIP_TRY(hipMemcpyAsync(c->rho[1], rho, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
        }
    }
    if (psi4) HIP_TRY(frk::launch_pack4to3(psi4, c->psi, nullptr, c->n, c->stream));
    if (pinv4) HIP_TRY(frk::launch_pack4to3(pinv4, c->pinv[0], c->pinv[1], c->n, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FR_OK;
}

extern "C" int fr_solver_get_state(fr_ctx *c, float *v4, float *rho, float *psi4, float *pinv4) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_solver_get_state before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    if (c->st != frk::ST_F32) {
        if (v4 || rho) HIP_TRY(frk::launch_vr_get(c->st, c->vr[c->cur], v4, rho, c->n, c->stream));
    } else {
        if (v4) HIP_TRY(frk::launch_unpack3to4(c->v[c->cur], v4, c->n, c->stream));
        if (rho)
            HIP_TRY(hipMemcpyAsync(rho, c->rho[c->cur], (size_t)c->n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    }
    if (psi4) HIP_TRY(frk::launch_unpack3to4(c->psi, psi4, c->n, c->stream));
    if (pinv4) HIP_TRY(frk::launch_unpack3to4(c->pinv[c->cur], pinv4, c->n, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FR_OK;
}

extern "C" int fr_get_deformation(fr_ctx *c, float *d_psi4) {
    if (!c || !d_psi4) return fail(FR_ERR_INVALID, "null argument");
    return fr_solver_get_state(c, nullptr, nullptr, d_psi4, nullptr);
}

extern "C" int fr_get_inverse_deformation(fr_ctx *c, float *d_pinv4) {
    if (!c || !d_pinv4) return fail(FR_ERR_INVALID, "null argument");
    return fr_solver_get_state(c, nullptr, nullptr, nullptr, d_pinv4);
}

extern "C" int fr_solver_energy(fr_ctx *c, double *e_data, double *e_reg) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_solver_energy before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    double *tmp = nullptr;
    if ((rc = dalloc(&tmp, 3))) return rc;
    hipError_t e = frk::launch_energy(c->d, c->S, c->T, c->psi, c->partials, c->nparts, c->stream);
    if (e == hipSuccess)
        e = frk::launch_energy_finalize(c->partials, c->nparts, (double)c->prm.w_reg, tmp + 1, tmp + 2, tmp, c->stream);
    double h[3] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, tmp, sizeof h, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(FR_ERR_HIP, "energy: %s", hipGetErrorString(e));
    if (e_data) *e_data = h[1];
    if (e_reg) *e_reg = h[2];
    return FR_OK;
}

extern "C" int fr_warp(int X, int Y, int Z, const float *img, const int16_t *lab, const float *psi4, float *out_img,
                       int16_t *out_lab, void *stream) {
    if (X < 1 || Y < 1 || Z < 1 || !psi4 || (!out_img && !out_lab) || (out_img && !img) || (out_lab && !lab))
        return fail(FR_ERR_INVALID, "bad fr_warp arguments");
    if (4LL * X * Y * Z >= (1LL << 31)) return fail(FR_ERR_INVALID, "volume too large (4*N must be < 2^31)");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(frk::launch_warp(Dims{X, Y, Z}, img, lab, psi4, 4, out_img, out_lab, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FR_OK;
}

extern "C" int fr_field_identity(int X, int Y, int Z, float *field4, void *stream) {
    if (X < 1 || Y < 1 || Z < 1 || !field4) return fail(FR_ERR_INVALID, "bad fr_field_identity arguments");
    if (16LL * X * Y * Z >= (1LL << 31)) return fail(FR_ERR_INVALID, "volume too large (16*N must be < 2^31)");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(frk::launch_identity4(Dims{X, Y, Z}, field4, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FR_OK;
}

// ---------------------------------------------------------------------------------------------
// batches: many volumes advanced together, one launch per pass per iteration, run as a hipGraph
// ---------------------------------------------------------------------------------------------

struct fr_batch {
    std::vector<fr_ctx *> ctx;
    std::vector<unsigned> storage_gen;  // each context's storage generation when the layout was made
    int st = frk::ST_F32;               // the storage of every volume at that time
    frk::IterArgs *d_base = nullptr;  // 3 per volume: (unused), pass A, pass B arguments at buffer parity 0
    int *d_starts = nullptr;          // 3 x (n + 1): first block of each volume in each pass's grid
    int total[3] = {0, 0, 0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, done = nullptr;
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // the whole run, per start parity
    hipGraph_t graph[2] = {nullptr, nullptr};
};

static void free_batch(fr_batch *b) {
    for (int k = 0; k < 2; ++k) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    }
    if (b->d_base) (void)hipFree(b->d_base);
    if (b->d_starts) (void)hipFree(b->d_starts);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

// every volume must have one storage (the batched kernels are instantiated per storage)
// and the batched kernels perform the reference's operation order only
static int batch_same_storage(fr_ctx *const *ctxs, int n) {
    for (int q = 0; q < n; ++q) {
        if (ctxs[q]->st != ctxs[0]->st)
            return fail(FR_ERR_INVALID, "context %d: the storage of v and rho must be the same for every volume of a "
                                        "batch", q);
        if (ctxs[q]->arith != frk::ARITH_REF)
            return fail(FR_ERR_INVALID, "context %d: batches run the reference-order arithmetic "
                                        "(fr_set_arith FR_ARITH_REFERENCE)", q);
    }
    return FR_OK;
}

// A run uses the buffers and kernels of the storage each volume had when the layout was made: a
// later fr_set_storage or fr_set_arith on a volume (new buffers, other kernels) makes the batch
// unusable until fr_batch_relayout.
static int batch_check_layout(const fr_batch *b) {
    for (size_t q = 0; q < b->ctx.size(); ++q)
        if (b->ctx[q]->storage_gen != b->storage_gen[q] || b->ctx[q]->st != b->st)
            return fail(FR_ERR_STATE, "batch volume %d: its storage or arithmetic was set after the batch was laid "
                                      "out; call fr_batch_relayout", (int)q);
    return FR_OK;
}

// the launch-time arguments of every volume and the block ranges of each pass
static int batch_layout(fr_batch *b) {
    const int n = (int)b->ctx.size();
    if (int rc = batch_same_storage(b->ctx.data(), n)) return rc;
    std::vector<frk::IterArgs> base(3 * (size_t)n);
    std::vector<int> starts(3 * (size_t)(n + 1));
    starts[(size_t)n] = 0;  // slot 0 (no pass) is empty
    b->total[0] = 0;
    for (int pass = 1; pass < 3; ++pass) {
        long long at = 0;
        for (int q = 0; q < n; ++q) {
            fr_ctx *c = b->ctx[q];
            frk::IterArgs a;
            const int save = c->cur;
            c->cur = 0;
            iter_args(c, a);
            c->cur = save;
            a = frk::pass_args(a, pass, c->zc[pass]);
            base[3 * (size_t)q + pass] = a;
            starts[(size_t)pass * (n + 1) + q] = (int)at;
            at += (a.nblocks + 7) / 8 * 8;  // each volume's run starts at a multiple of 8 blocks
            if (at >= INT_MAX / 2) return fail(FR_ERR_INVALID, "batch grid too large");
        }
        starts[(size_t)pass * (n + 1) + n] = (int)at;
        b->total[pass] = (int)at;
    }
    HIP_TRY(hipMemcpy(b->d_base, base.data(), sizeof(frk::IterArgs) * base.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(b->d_starts, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice));
    b->st = b->ctx[0]->st;
    b->storage_gen.resize((size_t)n);
    for (int q = 0; q < n; ++q) b->storage_gen[(size_t)q] = b->ctx[q]->storage_gen;
    for (int k = 0; k < 2; ++k) {  // captured runs refer to the old layout
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
    }
    return FR_OK;
}

extern "C" int fr_batch_create(fr_ctx *const *ctxs, int n, fr_batch **out) {
    if (!ctxs || n < 1 || !out) return fail(FR_ERR_INVALID, "bad arguments");
    *out = nullptr;
    for (int q = 0; q < n; ++q) {
        if (!ctxs[q]) return fail(FR_ERR_INVALID, "null context %d", q);
        for (int r = 0; r < q; ++r)
            if (ctxs[r] == ctxs[q]) return fail(FR_ERR_INVALID, "context %d appears twice", q);
        const fr_params &p = ctxs[q]->prm, &p0 = ctxs[0]->prm;
        if (p.max_iter != p0.max_iter || p.alpha != p0.alpha || p.reset_iter != p0.reset_iter || p.p != p0.p ||
            p.C != p0.C || p.verbosity != p0.verbosity)
            return fail(FR_ERR_INVALID, "context %d: MAX_ITER, ALPHA, RESET_ITER, p, C and verbosity must be the same for "
                                        "every volume of a batch", q);
    }
    if (int rc = batch_same_storage(ctxs, n)) return rc;
    fr_batch *b = new fr_batch();
    b->ctx.assign(ctxs, ctxs + n);
    int rc;
    if ((rc = dalloc(&b->d_base, 3 * (size_t)n)) || (rc = dalloc(&b->d_starts, 3 * (size_t)(n + 1)))) {
        free_batch(b);
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        free_batch(b);
        return fail(FR_ERR_HIP, "batch stream/events: %s", hipGetErrorString(e));
    }
    if ((rc = batch_layout(b))) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return FR_OK;
}

extern "C" int fr_batch_destroy(fr_batch *b) {
    if (!b) return FR_OK;
    (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    return FR_OK;
}

// Re-read every volume's z-chunking (after fr_set_chunking on its context).
extern "C" int fr_batch_relayout(fr_batch *b) {
    if (!b) return fail(FR_ERR_INVALID, "null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return batch_layout(b);
}

// the launches of the whole run from buffer parity `par`, in the order fr_solver_run issues them
static int batch_launches(fr_batch *b, int par) {
    const fr_params &p = b->ctx[0]->prm;
    const int n = (int)b->ctx.size();
    float t = p.alpha;
    for (int iter = 1; iter <= p.max_iter; ++iter) {
        frk::BatchIter it;
        it.iter = iter;
        it.cur = par ^ ((iter - 1) & 1);
        it.check_energy = verbose_now(p, iter) ? 1 : 0;
        it.no_stop = 0;
        vel_scalars(p, t, &it.m1t, &it.m2);
        if (it.check_energy) {
            const int mi = p.max_iter > 0 ? p.max_iter : 1;
            for (int q = 0; q < n; ++q) {
                fr_ctx *c = b->ctx[q];
                double *e = c->energy;
                HIP_TRY(frk::launch_energy(c->d, c->S, c->T, c->psi, c->partials, c->nparts, b->stream));
                HIP_TRY(frk::launch_energy_finalize(c->partials, c->nparts, (double)c->prm.w_reg, e + mi + (iter - 1),
                                                    e + 2 * mi + (iter - 1), e + (iter - 1), b->stream));
            }
        }
        for (int pass = 1; pass < 3; ++pass)
            HIP_TRY(frk::launch_pass_batch(pass, b->d_base, b->d_starts + (size_t)pass * (n + 1), n, b->total[pass], it,
                                           b->st, b->stream));
        t += p.alpha;
        if (p.reset_iter > 0 && (iter + 1) % p.reset_iter == 0) t = p.alpha;
    }
    return FR_OK;
}

extern "C" int fr_batch_run(fr_batch *b) {
    if (!b) return fail(FR_ERR_INVALID, "null batch");
    const int n = (int)b->ctx.size();
    int rc;
    if ((rc = batch_check_layout(b))) return rc;
    const int par = b->ctx[0]->cur;
    for (int q = 0; q < n; ++q) {
        fr_ctx *c = b->ctx[q];
        if (!c->initialised) return fail(FR_ERR_STATE, "batch volume %d: fr_batch_run before fr_solver_init", q);
        if (c->have_run && (rc = collect_run(c))) return rc;  // settles c->cur of an earlier run
        if (c->cur != par)
            return fail(FR_ERR_STATE, "batch volumes are at different buffer parities (volume %d); re-initialise them", q);
    }
    for (int q = 0; q < n; ++q)  // control words cleared eagerly, before the graph's launch
        if ((rc = reset_control(b->ctx[q]))) return rc;
    for (int q = 0; q < n; ++q) HIP_TRY(hipStreamSynchronize(b->ctx[q]->stream));
    if (!b->exec[par]) {
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        rc = batch_launches(b, par);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(FR_ERR_HIP, "batch capture: %s", hipGetErrorString(e));
        b->graph[par] = g;
        HIP_TRY(hipGraphInstantiate(&b->exec[par], g, nullptr, nullptr, 0));
    }
    for (int q = 0; q < n; ++q) HIP_TRY(hipEventRecord(b->ctx[q]->ev0, b->stream));
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec[par], b->stream));
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    for (int q = 0; q < n; ++q) HIP_TRY(hipEventRecord(b->ctx[q]->ev1, b->stream));
    HIP_TRY(hipEventRecord(b->done, b->stream));
    const fr_params &p = b->ctx[0]->prm;
    for (int q = 0; q < n; ++q) {
        fr_ctx *c = b->ctx[q];
        HIP_TRY(hipStreamWaitEvent(c->stream, b->done, 0));  // later calls on the context see the run
        c->run_start_buf = par;
        c->captured = false;
        c->cur = par ^ (p.max_iter & 1);
        c->have_run = true;
        c->have_final = false;
        c->stats = fr_stats{};
    }
    return FR_OK;
}

extern "C" int fr_batch_time_iterations(fr_batch *b, int iters, float *ms) {
    if (!b || iters < 1 || !ms) return fail(FR_ERR_INVALID, "bad arguments");
    const int n = (int)b->ctx.size();
    const int mi = b->ctx[0]->prm.max_iter;
    if (iters > mi) return fail(FR_ERR_INVALID, "iters must be <= MAX_ITER");
    int rc;
    if ((rc = batch_check_layout(b))) return rc;
    for (int q = 0; q < n; ++q) {
        fr_ctx *c = b->ctx[q];
        if (!c->initialised) return fail(FR_ERR_STATE, "batch volume %d not initialised", q);
        if (c->have_run && (rc = collect_run(c))) return rc;
        if ((rc = reset_control(c))) return rc;
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    const int par = b->ctx[0]->cur;
    for (int q = 0; q < n; ++q)
        if (b->ctx[q]->cur != par) return fail(FR_ERR_STATE, "batch volumes are at different buffer parities");
    const fr_params &p = b->ctx[0]->prm;
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    float t = p.alpha;
    for (int iter = 1; iter <= iters; ++iter) {
        frk::BatchIter it;
        it.iter = iter;
        it.cur = par ^ ((iter - 1) & 1);
        it.check_energy = 0;
        it.no_stop = 1;
        vel_scalars(p, t, &it.m1t, &it.m2);
        for (int pass = 1; pass < 3; ++pass)
            HIP_TRY(frk::launch_pass_batch(pass, b->d_base, b->d_starts + (size_t)pass * (n + 1), n, b->total[pass], it,
                                           b->st, b->stream));
        t += p.alpha;
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipEventSynchronize(b->ev1));
    HIP_TRY(hipEventElapsedTime(ms, b->ev0, b->ev1));
    for (int q = 0; q < n; ++q) {  // the volumes' state has advanced: they must be re-initialised
        b->ctx[q]->cur = par ^ (iters & 1);
        b->ctx[q]->initialised = false;
    }
    return FR_OK;
}

extern "C" int fr_batch_time(fr_batch *b, float *ms) {
    if (!b || !ms) return fail(FR_ERR_INVALID, "bad arguments");
    HIP_TRY(hipEventSynchronize(b->ev1));
    HIP_TRY(hipEventElapsedTime(ms, b->ev0, b->ev1));
    return FR_OK;
}

// ---------------------------------------------------------------------------------------------
// the reference's further public functions
// ---------------------------------------------------------------------------------------------

extern "C" int fr_get_inverse(int X, int Y, int Z, const float *psi4, float *pinv4, void *stream) {
    if (X < 3 || Y < 3 || Z < 3 || !psi4 || !pinv4) return fail(FR_ERR_INVALID, "bad fr_get_inverse arguments");
    if (16LL * X * Y * Z >= (1LL << 31)) return fail(FR_ERR_INVALID, "volume too large (16*N must be < 2^31)");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(frk::launch_get_inverse(Dims{X, Y, Z}, psi4, pinv4, 48, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FR_OK;
}

extern "C" int fr_interpolate_gradient(int X, int Y, int Z, const float *field4, float *out4, const float *psi4,
                                       void *stream) {
    if (X < 1 || Y < 1 || Z < 1 || !field4 || !out4 || !psi4 || field4 == out4)
        return fail(FR_ERR_INVALID, "bad fr_interpolate_gradient arguments");
    if (16LL * X * Y * Z >= (1LL << 31)) return fail(FR_ERR_INVALID, "volume too large (16*N must be < 2^31)");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(frk::launch_interpolate_field4(Dims{X, Y, Z}, field4, psi4, out4, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FR_OK;
}

extern "C" int fr_voxel_max_energy(fr_ctx *c, float *max_energy, long long *voxel) {
    if (!c) return fail(FR_ERR_INVALID, "null context");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_voxel_max_energy before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    HIP_TRY(hipMemsetAsync(c->max_slot, 0, sizeof(unsigned long long), c->stream));
    HIP_TRY(frk::launch_voxel_max_energy(c->d, c->T, c->S, c->psi, c->prm.w_reg, c->max_slot, c->stream));
    unsigned long long mp = 0;
    HIP_TRY(hipMemcpyAsync(&mp, c->max_slot, sizeof mp, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float m;
    long long idx;
    unpack_max(mp, &m, &idx);
    if (max_energy) *max_energy = m;
    if (voxel) *voxel = idx;
    return FR_OK;
}

extern "C" int fr_sum(const float *d, long long n, double *sum, void *stream) {
    if (!d || n < 0 || !sum) return fail(FR_ERR_INVALID, "bad fr_sum arguments");
    hipStream_t s = (hipStream_t)stream;
    double *tmp = nullptr;
    const int np = frk::sum_partials_count(n);
    int rc;
    if ((rc = dalloc(&tmp, (size_t)np + 1))) return rc;
    hipError_t e = frk::launch_sum(d, n, tmp, tmp + np, s);
    double h = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, tmp + np, sizeof h, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(FR_ERR_HIP, "fr_sum: %s", hipGetErrorString(e));
    *sum = h;
    return FR_OK;
}

// ---------------------------------------------------------------------------------------------
// fr_sum_noalloc: the sum's scratch in a static device array instead of a per-call allocation
// ---------------------------------------------------------------------------------------------

// The partial sums of up to SUM_SCRATCH_PARTIALS blocks and the total.  A __device__ array exists
// once on every GPU the library runs on; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host-side stand-in).
constexpr int SUM_SCRATCH_PARTIALS = 1024;
static __device__ __align__(256) double scratch[SUM_SCRATCH_PARTIALS + 1];

constexpr int SCRATCH_MAX_DEVICES = 64;
static double *g_scratch_ptr[SCRATCH_MAX_DEVICES];  // per device, looked up at its first call
static std::mutex g_scratch_mu;

// the device address of `scratch` on the current device (looked up once, with that device current)
static int scratch_on_current_device(double **out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES)
        return fail(FR_ERR_INVALID, "device %d is beyond the %d devices fr_sum_noalloc supports", dev,
                    SCRATCH_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return FR_OK;
}

extern "C" int fr_sum_noalloc(const float *d, long long n, double *sum, void *stream) {
    if (!d || n < 0 || !sum) return fail(FR_ERR_INVALID, "bad fr_sum_noalloc arguments");
    const int np = frk::sum_partials_count(n);
    if (np > SUM_SCRATCH_PARTIALS)
        return fail(FR_ERR_INVALID, "fr_sum_noalloc: %d partial sums exceed its scratch of %d", np,
                    SUM_SCRATCH_PARTIALS);
    double *ptr = nullptr;
    if (int rc = scratch_on_current_device(&ptr)) return rc;
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(frk::launch_sum(d, n, ptr, ptr + np, s));  // every partial is written: nothing to clear
    double h = 0.0;
    HIP_TRY(hipMemcpyAsync(&h, ptr + np, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *sum = h;
    return FR_OK;
}

static constexpr int TIME_KERNELS_WARMUP = 150;

extern "C" int fr_time_kernels(fr_ctx *c, int reps, float *ms_a, float *ms_b) {
    if (!c || reps < 1 || reps > 10000) return fail(FR_ERR_INVALID, "bad arguments");
    if (!c->initialised) return fail(FR_ERR_STATE, "fr_time_kernels before fr_solver_init");
    int rc;
    if (c->have_run && (rc = collect_run(c))) return rc;
    if ((rc = reset_control(c))) return rc;
    float *psi_save = nullptr;
    if ((rc = dalloc(&psi_save, 3 * (size_t)c->n))) return rc;
    frk::IterArgs a;
    iter_args(c, a);
    a.iter = 1;
    vel_scalars(c->prm, c->prm.alpha, &a.m1t, &a.m2);
    // reps iterations in sequence (A, B, A, B, ...) with an event between consecutive launches, so
    // each kernel runs with the caches as the iteration loop leaves them.  The iteration is not
    // advanced: pass A re-reads the same old buffers, pass B accumulates into psi, which is
    // restored afterwards.
    std::vector<hipEvent_t> ev(2 * (size_t)reps + 1, nullptr);
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < ev.size() && e == hipSuccess; ++k) e = hipEventCreate(&ev[k]);
    if (e == hipSuccess)
        e = hipMemcpyAsync(psi_save, c->psi, 3 * (size_t)c->n * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    // every iteration reduces its max norm into iteration slots, as the loop does; the stop tests
    // are off
    a.check_conv = 0;
    a.check_energy = 0;
    const int mi = c->prm.max_iter > 0 ? c->prm.max_iter : 1;
    // untimed iterations first: after the GPU has idled (this call's own copies and allocation
    // leave it idle), its clock takes ~25 ms of continuous work to settle, and pass A runs ~15 %
    // slower until then
    for (int r = 0; r < TIME_KERNELS_WARMUP && e == hipSuccess; ++r) {
        a.iter = 1 + r % mi;
        e = frk::launch_pass_a(a, c->zc[1], c->st, c->stream, c->arith);
        if (e == hipSuccess) e = frk::launch_pass_b(a, c->zc[2], c->st, c->stream);
    }
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        a.iter = 1 + (r + 1) % mi;
        e = hipEventRecord(ev[2 * r], c->stream);
        if (e == hipSuccess) e = frk::launch_pass_a(a, c->zc[1], c->st, c->stream, c->arith);
        if (e == hipSuccess) e = hipEventRecord(ev[2 * r + 1], c->stream);
        if (e == hipSuccess) e = frk::launch_pass_b(a, c->zc[2], c->st, c->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(ev[2 * (size_t)reps], c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ev[2 * (size_t)reps]);
    double sum[3] = {0, 0, 0};
    for (int r = 0; r < reps && e == hipSuccess; ++r)
        for (int k = 0; k < 2 && e == hipSuccess; ++k) {
            float ms = 0.f;
            e = hipEventElapsedTime(&ms, ev[2 * r + k], ev[2 * r + k + 1]);
            sum[k + 1] += ms;
        }
    if (e == hipSuccess)
        e = hipMemcpyAsync(c->psi, psi_save, 3 * (size_t)c->n * sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    (void)hipFree(psi_save);
    if (e != hipSuccess) return fail(FR_ERR_HIP, "fr_time_kernels: %s", hipGetErrorString(e));
    if (ms_a) *ms_a = (float)(sum[1] / reps);
    if (ms_b) *ms_b = (float)(sum[2] / reps);
    return FR_OK;
}

// ---------------------------------------------------------------------------------------------
// params.ini (src/apps/demo.cpp:96-149)
// ---------------------------------------------------------------------------------------------

static std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

static bool parse_int(const std::string &s, int *out) {
    char *end = nullptr;
    errno = 0;
    long v = strtol(s.c_str(), &end, 10);
    if (s.empty() || *end || errno || v < INT32_MIN || v > INT32_MAX) return false;
    *out = (int)v;
    return true;
}

static bool parse_float(const std::string &s, float *out) {
    char *end = nullptr;
    errno = 0;
    float v = strtof(s.c_str(), &end);
    if (s.empty() || *end) return false;
    *out = v;
    return true;
}

extern "C" int fr_params_load_ini(const char *path, fr_params *p) {
    if (!path || !p) return fail(FR_ERR_INVALID, "null argument");
    FILE *f = fopen(path, "r");
    if (!f) return fail(FR_ERR_IO, "cannot open %s", path);
    char line[4096];
    int lineno = 0;
    fr_params q = *p;
    while (fgets(line, sizeof line, f)) {
        ++lineno;
        std::string s = trim(line);
        if (s.empty() || s[0] == '#' || s[0] == ';') continue;
        size_t eq = s.find('=');
        if (eq == std::string::npos) {
            fclose(f);
            return fail(FR_ERR_PARSE, "%s:%d: expected KEY=VALUE", path, lineno);
        }
        std::string key = trim(s.substr(0, eq)), val = trim(s.substr(eq + 1));
        size_t hash = val.find('#');
        if (hash != std::string::npos) val = trim(val.substr(0, hash));
        bool ok = true;
        if (key == "VOL_DIMS_X") ok = parse_int(val, &q.vol_dims[0]);
        else if (key == "VOL_DIMS_Y") ok = parse_int(val, &q.vol_dims[1]);
        else if (key == "VOL_DIMS_Z") ok = parse_int(val, &q.vol_dims[2]);
        else if (key == "VOL_SIZE_X") ok = parse_float(val, &q.vol_size[0]);
        else if (key == "VOL_SIZE_Y") ok = parse_float(val, &q.vol_size[1]);
        else if (key == "VOL_SIZE_Z") ok = parse_float(val, &q.vol_size[2]);
        else if (key == "NO_LABELS") ok = parse_int(val, &q.no_labels);
        else if (key == "MAX_ITER") ok = parse_int(val, &q.max_iter);
        else if (key == "MAX_UPDATE_NORM") ok = parse_float(val, &q.max_update_norm);
        else if (key == "RHO_0") ok = parse_float(val, &q.rho_0);
        else if (key == "RESET_ITER") ok = parse_int(val, &q.reset_iter);
        else if (key == "TAU") ok = parse_float(val, &q.tau);
        else if (key == "ALPHA") ok = parse_float(val, &q.alpha);
        else if (key == "W_REG") ok = parse_float(val, &q.w_reg);
        else if (key == "p") ok = parse_float(val, &q.p);
        else if (key == "C") ok = parse_float(val, &q.C);
        else {
            fclose(f);
            return fail(FR_ERR_PARSE, "%s:%d: unrecognised option '%s'", path, lineno, key.c_str());
        }
        if (!ok) {
            fclose(f);
            return fail(FR_ERR_PARSE, "%s:%d: invalid value '%s' for %s", path, lineno, val.c_str(), key.c_str());
        }
    }
    fclose(f);
    *p = q;
    return FR_OK;
}

// ---------------------------------------------------------------------------------------------
// device memory plumbing
// ---------------------------------------------------------------------------------------------

extern "C" int fr_dev_malloc(void **d, size_t bytes) {
    if (!d) return fail(FR_ERR_INVALID, "null argument");
    hipError_t e = hipMalloc(d, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(FR_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return FR_OK;
}
extern "C" int fr_dev_free(void *d) {
    HIP_TRY(hipFree(d));
    return FR_OK;
}
extern "C" int fr_memcpy_htod(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FR_OK;
}
extern "C" int fr_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return FR_OK;
}
extern "C" int fr_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return FR_OK;
}
extern "C" int fr_dev_memset(void *d, int value, size_t bytes) {
    HIP_TRY(hipMemset(d, value, bytes));
    return FR_OK;
}
extern "C" int fr_dev_sync(void) {
    HIP_TRY(hipDeviceSynchronize());
    return FR_OK;
}
extern "C" int fr_device_info(char *buf, size_t len) {
    if (!buf || !len) return fail(FR_ERR_INVALID, "null argument");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    HIP_TRY(hipGetDeviceProperties(&pr, dev));
    snprintf(buf, len, "%s (%s), %d CUs, %.1f GiB, %.0f MHz", pr.name, pr.gcnArchName, pr.multiProcessorCount,
             pr.totalGlobalMem / 1073741824.0, pr.clockRate / 1000.0);
    return FR_OK;
}
