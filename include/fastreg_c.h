/*
 * fastreg_c.h — C-ABI of the MI355X (gfx950) fastreg registration solver.
 *
 * This is the drop-in boundary for fastreg's optimisation loop.  The reference exposes it as C++
 * classes (include/fastreg.hpp:12-38, include/solver.hpp:64-113); here every entry point is
 * plain C (pointers and sizes, no HIP or C++ types), and the C++ classes in fastreg.hpp /
 * solver.hpp of this repository are thin wrappers over it.  Device pointers are HIP device
 * allocations of the caller (or from fr_dev_malloc); volumes are x-fastest,
 * index = x + X*(y + Y*z).
 *
 * Every function returns 0 (FR_OK) on success or an FR_ERR_* code; fr_last_error() gives the
 * message of the calling thread's last failure.  Nothing here calls exit() (the reference's
 * cudaSafeCall does, include/safe_call.hpp:14-26).
 */
#ifndef FASTREG_C_H
#define FASTREG_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FR_OK 0
#define FR_ERR_INVALID 1 /* bad argument, dims or parameter */
#define FR_ERR_HIP 2     /* a HIP runtime call failed */
#define FR_ERR_ALLOC 3   /* device or host allocation failed */
#define FR_ERR_IO 4      /* file could not be read or written */
#define FR_ERR_PARSE 5   /* params.ini syntax error or unknown key */
#define FR_ERR_STATE 6   /* call out of order (e.g. step before init) */

/* Solver parameters.  Field meanings and defaults follow the reference's Params
 * (include/params.hpp:7-26) and params.ini keys (src/apps/demo.cpp:96-149, params.ini:1-25). */
typedef struct fr_params {
    int vol_dims[3];       /* VOL_DIMS_X, VOL_DIMS_Y, VOL_DIMS_Z */
    float vol_size[3];     /* VOL_SIZE_X/Y/Z (read and kept; the solver uses unit voxel spacing,
                              as the reference does) */
    int no_labels;         /* NO_LABELS: label values 0..no_labels-1 are counted in Dice */
    int max_iter;          /* MAX_ITER */
    float max_update_norm; /* MAX_UPDATE_NORM: stop when the max update norm is <= this */
    float rho_0;           /* RHO_0 */
    float tau;             /* TAU */
    float alpha;           /* ALPHA */
    float w_reg;           /* W_REG */
    float p;               /* p */
    float C;               /* C */
    int reset_iter;        /* RESET_ITER (default -1: never reset t) */
    int verbosity;         /* 0, 1 (--verbose) or 2 (--vverbose) */
} fr_params;

/* Statistics of the last fr_register / fr_solver_run. */
typedef struct fr_stats {
    int iterations;            /* iterations whose PDE updates ran */
    int stop_reason;           /* 0 = reached MAX_ITER, 1 = converged, 2 = NaN energy */
    float last_max_update;     /* max update norm of the last iteration that ran */
    long long last_max_voxel;  /* its linear voxel index */
    double last_energy;        /* last evaluated energy (verbosity > 0), else +inf */
    double min_det_true;       /* min over the interior of det(D psi) (central differences) */
    double min_det_ref;        /* the same with the reference's det() formula */
    double dice_all;           /* Dice over labels 1..no_labels-1, the reference's convention */
    float solver_ms;           /* HIP-event time of the iteration loop */
    float total_ms;            /* host wall time of fr_register */
} fr_stats;

typedef struct fr_ctx fr_ctx;

/* message of the calling thread's last failed call ("" if none) */
const char *fr_last_error(void);

/* ---- parameters -------------------------------------------------------------------------- */

/* The reference's params.ini defaults (params.ini:1-25) plus reset_iter = -1, verbosity = 0
 * (include/params.hpp:11-12). */
int fr_params_default(fr_params *p);

/* Parse KEY=VALUE lines ('#' comments, blank lines) with the reference's key set
 * (src/apps/demo.cpp:96-149).  Unknown keys and malformed values fail with FR_ERR_PARSE, as the
 * reference's Boost parser does.  Keys absent from the file keep the values already in *p. */
int fr_params_load_ini(const char *path, fr_params *p);

/* ---- pipeline: the FastReg class (include/fastreg.hpp:12-38, src/fastreg/fastreg.cpp) ------ */

/* Allocate all solver state on the current HIP device for p->vol_dims.  FastReg::FastReg
 * (src/fastreg/fastreg.cpp:24-46) + Solver::Solver (src/fastreg/solver.cpp:7-65). */
int fr_create(const fr_params *p, fr_ctx **out);
int fr_destroy(fr_ctx *ctx);
int fr_get_params(const fr_ctx *ctx, fr_params *p);

/* Run work on `hip_stream` (a hipStream_t) instead of the context's own stream; NULL restores
 * the context's own stream. */
int fr_set_stream(fr_ctx *ctx, void *hip_stream);

/* Interior z-slices per thread block of the two iteration kernels (pass A, pass B); 0 restores
 * the choice fr_create made for the device and volume.  Results do not depend on it (every voxel
 * is computed the same way whatever the chunking).  No counterpart in the reference: it is a
 * launch-shape setting of this implementation, for tuning. */
int fr_set_chunking(fr_ctx *ctx, int zc_a, int zc_b);
int fr_get_chunking(const fr_ctx *ctx, int *zc_a, int *zc_b);

/* Storage of the velocity and density fields: FP32 (the default, the reference's precision), or
 * one 8-byte record per voxel holding v.x, v.y, v.z, rho in bf16 or fp16.  All arithmetic stays
 * FP32: every value of v and rho is rounded to the storage format (nearest even) when it is
 * computed, and the rounded value is the one used afterwards; energies are summed in FP64.
 * Setting the storage discards the state (call fr_solver_init / fr_register next).  No
 * counterpart in the reference, which stores both as FP32 (include/cuda/vector_fields.hpp:21-78,
 * include/cuda/scalar_fields.hpp:20-65): it is the reduced-precision form of those fields. */
#define FR_STORAGE_F32 0
#define FR_STORAGE_BF16 1
#define FR_STORAGE_F16 2
int fr_set_storage(fr_ctx *ctx, int storage);
int fr_get_storage(const fr_ctx *ctx, int *storage);

/* Arithmetic of the iteration: FR_ARITH_REFERENCE (the default) performs the reference's FP32
 * operations in its order, every product rounded before the sum it feeds, bit for bit the oracle's
 * results.  FR_ARITH_FMA contracts those multiply-adds into fused ones (as nvcc does by default for
 * the reference's kernels, src/fastreg/cuda/solver.cu:56-80,101-173,188-211,263-279,
 * src/fastreg/cuda/vector_fields.cu:352-404,503-535) and divides by |rho| + 1e-5 with the hardware
 * reciprocal (the reference's division there is not the IEEE one either, solver.cu:205-207): about
 * 13 % fewer instructions in pass A, results a few ulps per step from the reference order.  FP32
 * storage only (FR_ERR_INVALID otherwise); batches run the reference order and refuse contexts set
 * to FR_ARITH_FMA.  The state is kept.  No counterpart in the reference: it is a precision setting
 * of this implementation. */
#define FR_ARITH_REFERENCE 0
#define FR_ARITH_FMA 1
int fr_set_arith(fr_ctx *ctx, int arith);
int fr_get_arith(const fr_ctx *ctx, int *arith);

/* Whether a context keeps its state from one registration to the next.  The reference's FastReg
 * does: each operator() call re-initialises the four volumes and the density (rho = rho_0,
 * src/fastreg/fastreg.cpp:58-63, src/fastreg/cuda/solver.cu:317) but continues from the velocity,
 * psi and psi_inv the previous call left (include/solver.hpp:64-113 keeps v as a member; FastReg
 * keeps psi and psi_inv).  With keep = 1, fr_solver_init (and so fr_register) on a context that
 * holds a state replaces T, S and their labels, resets rho and keeps v, psi and psi_inv; the first
 * initialisation, and the first after fr_set_storage, starts from the identity as always.  The
 * kept state is that of the last run, read back first if it has not been; inside a stream capture
 * that read is not possible and fr_solver_init returns FR_ERR_STATE.  The default, keep = 0,
 * starts every registration fresh (the same inputs give the same result). */
int fr_set_keep_state(fr_ctx *ctx, int keep);
int fr_get_keep_state(const fr_ctx *ctx, int *keep);

/* FastReg::operator() (src/fastreg/fastreg.cpp:58-75): register the source (img2, lab2) to the
 * target (img1, lab1), all device pointers of no_voxels elements, and overwrite them in place:
 * img1/lab1 <- target o psi_inv, img2/lab2 <- source o psi.  Every call starts from a fresh
 * state (psi = psi_inv = id, v = 0, rho = rho_0), or with fr_set_keep_state from the state the
 * previous call left (rho = rho_0).  Blocks until the outputs are written. */
int fr_register(fr_ctx *ctx, float *d_img1, int16_t *d_lab1, float *d_img2, int16_t *d_lab2);

/* FastReg::getDeformationField (include/fastreg.hpp:22): psi as float4 (x, y, z, 0) per voxel,
 * absolute voxel coordinates, into device memory of 4*no_voxels floats. */
int fr_get_deformation(fr_ctx *ctx, float *d_psi_xyzw);
/* psi_inv in the same layout */
int fr_get_inverse_deformation(fr_ctx *ctx, float *d_psi_inv_xyzw);

int fr_get_stats(const fr_ctx *ctx, fr_stats *st);

/* Dice of the last registration (src/fastreg/cuda/solver.cu:9-50): per-label 2*overlap/voxels
 * for labels 0..n-1 (NaN when a label is absent) and the raw counts; any pointer may be NULL. */
int fr_get_dice(const fr_ctx *ctx, int n, double *dice, long long *overlap, long long *voxels);

/* Per-iteration history of the last run: max update norm and energy (+inf where the energy
 * was not evaluated), n entries from iteration 1. */
int fr_get_history(const fr_ctx *ctx, int n, float *max_norm, double *energy);

/* The reference's stdout messages for the last run (src/fastreg/cuda/solver.cu:330-447,
 * src/fastreg/reductor.cpp:56-59), NUL-terminated, truncated to len. */
int fr_get_log(const fr_ctx *ctx, char *buf, size_t len);

/* ---- solver: Solver::estimate_psi (src/fastreg/cuda/solver.cu:295-444) in pieces ---------- */

/* Fresh state from the four device inputs (copied; the inputs are not modified):
 * FastReg::operator()'s initialise calls (src/fastreg/fastreg.cpp:59-63) plus estimate_psi's
 * pre-loop (solver.cu:308,317). */
int fr_solver_init(fr_ctx *ctx, const float *d_tgt, const int16_t *d_tgt_lab, const float *d_src,
                   const int16_t *d_src_lab);

/* One iteration body (solver.cu:335-407) with the given t; waits and returns the iteration's max
 * update norm and its voxel (either pointer may be NULL). */
int fr_solver_step(fr_ctx *ctx, float t, float *max_norm, long long *max_voxel);

/* The iteration loop (solver.cu:327-433) for MAX_ITER iterations from the current state with the
 * reference's t schedule, stopping rules and messages; asynchronous on the context stream
 * (fr_get_stats / fr_get_history / fr_get_log synchronise).  The loop is captured into a hipGraph
 * on the first run from each ping-pong parity and replayed afterwards.
 * When the caller is capturing the context stream (fr_set_stream, then hipStreamBeginCapture),
 * the loop's kernels are recorded into the caller's graph instead (kernel nodes only, no memset
 * nodes) and nothing runs yet.  Each replay of that graph reruns the loop on this context's
 * buffers; call fr_solver_init before every replay (outside the capture), and the statistics,
 * history, log, state and fr_solver_finalize then give that replay's results (solver_ms may be
 * 0).  The next fr_solver_run or fr_solver_step outside a capture ends this mode.  The caller's
 * graph refers to this context's buffers and events: destroy it before the context. */
int fr_solver_run(fr_ctx *ctx);

/* Final warps and figures (solver.cu:435-443, fastreg.cpp:71-72): img1/lab1 <- target o psi_inv,
 * img2/lab2 <- source o psi, Dice, min det J.  Any output pointer may be NULL. */
int fr_solver_finalize(fr_ctx *ctx, float *d_img1, int16_t *d_lab1, float *d_img2, int16_t *d_lab2);

/* State in the reference's layouts: v, psi, psi_inv float4 (x, y, z, 0) per voxel, rho float.
 * NULL skips a field.  set_state also sets the boundary voxels from the given arrays. */
int fr_solver_set_state(fr_ctx *ctx, const float *d_v4, const float *d_rho, const float *d_psi4,
                        const float *d_psi_inv4);
int fr_solver_get_state(fr_ctx *ctx, float *d_v4, float *d_rho, float *d_psi4, float *d_psi_inv4);

/* Energy of the current state (src/fastreg/reductor.cpp:45-62 with solver.cu:357-363):
 * e_data = 0.5 sum (T - S o psi)^2, e_reg = 0.5 sum ||D(psi - id)||_F^2, summed in FP64. */
int fr_solver_energy(fr_ctx *ctx, double *e_data, double *e_reg);

/* ---- batches: many registrations advanced together ---------------------------------------- */

/* A batch runs the iteration loops of n contexts together, each volume of its own size: one
 * launch per pass per iteration covers the tiles of every volume, and the whole run is captured
 * once into a hipGraph (per start parity) and replayed.  The contexts keep their own state, stop
 * rules (each volume stops on its own), logs and statistics; fr_solver_init / fr_get_stats /
 * fr_get_log / fr_solver_finalize are used on them as for a single run.  MAX_ITER, ALPHA,
 * RESET_ITER, p, C and verbosity must be the same for all n (the iteration schedule is shared);
 * TAU, W_REG, RHO_0 and MAX_UPDATE_NORM may differ.  No counterpart in the reference, which
 * registers one pair per FastReg call (src/fastreg/fastreg.cpp:58-75): it is the batched form
 * of that call.  The batch keeps the context pointers; destroy it before them. */
typedef struct fr_batch fr_batch;
int fr_batch_create(fr_ctx *const *ctxs, int n, fr_batch **out);
int fr_batch_destroy(fr_batch *b);
/* Re-read each context's z-chunking (fr_set_chunking) and storage into the batch's launch layout.
 * A batch runs the buffers and kernels of the storage its volumes had when the layout was made:
 * after fr_set_storage or fr_set_arith on any of them, fr_batch_run and fr_batch_time_iterations
 * fail with FR_ERR_STATE until this is called (and it fails while the volumes' storages differ or
 * one of them is set to FR_ARITH_FMA). */
int fr_batch_relayout(fr_batch *b);
/* MAX_ITER iterations of every volume from its current state (all volumes must have been
 * initialised and be at the same point: fresh, or after the same number of runs); asynchronous */
int fr_batch_run(fr_batch *b);
/* device time of the last fr_batch_run (waits for it) */
int fr_batch_time(fr_batch *b, float *ms);
/* Device time of `iters` iterations of every volume with the stop tests off (a fixed amount of
 * work, for choosing launch shapes).  The volumes' state advances: they are marked uninitialised
 * and must be initialised again (fr_solver_init) before any other use. */
int fr_batch_time_iterations(fr_batch *b, int iters, float *ms);

/* ---- kernels ----------------------------------------------------------------------------- */

/* apply (src/fastreg/cuda/vector_fields.cu:81-107): out_img = trilinear img at psi(p), out_lab =
 * lab at the floor corner of clamp(psi(p)), all voxels; psi float4 per voxel.  Either label
 * pointer may be NULL.  hip_stream may be NULL (default stream). */
int fr_warp(int X, int Y, int Z, const float *d_img, const int16_t *d_lab, const float *d_psi4, float *d_out_img,
            int16_t *d_out_lab, void *hip_stream);

/* init_identity_kernel (src/fastreg/cuda/vector_fields.cu:64-79, DeformationField::clear,
 * src/fastreg/vector_fields.cpp:87-92): field(p) = (x, y, z, 0) at every voxel p = (x, y, z) of a
 * float4 field.  hip_stream may be NULL (default stream).  Waits. */
int fr_field_identity(int X, int Y, int Z, float *d_field4, void *hip_stream);

/* ---- the reference's further public functions (not used by its registration loop) -------- */

/* DeformationField::get_inverse (src/fastreg/vector_fields.cpp:96-103, src/fastreg/cuda/
 * vector_fields.cu:108-136): 48 sweeps of psi_inv(p) = -trilerp(psi, psi_inv(p)) over all voxels,
 * in place, starting from the given d_psi_inv4; float4 (x, y, z, w) fields, w written 0.  Waits. */
int fr_get_inverse(int X, int Y, int Z, const float *d_psi4, float *d_psi_inv4, void *hip_stream);

/* interpolate_gradient (src/fastreg/cuda/vector_fields.cu:194-225): out(p) = trilerp(field,
 * psi(p)) over all voxels; float4 fields (d_out4 distinct from d_field4).  Waits. */
int fr_interpolate_gradient(int X, int Y, int Z, const float *d_field4, float *d_out4, const float *d_psi4,
                            void *hip_stream);

/* Reductor::voxel_max_energy (src/fastreg/reductor.cpp:82-89, src/fastreg/cuda/reductor.cu:637-760) on
 * the context's current state: the largest per-voxel (T - S o psi)^2 + W_REG * (sum of the squared
 * rows of D(psi - id)) and its voxel (the lowest index on equal energies).  Waits. */
int fr_voxel_max_energy(fr_ctx *ctx, float *max_energy, long long *voxel);

/* ScalarField::sum (src/fastreg/scalar_fields.cpp:29-33, src/fastreg/cuda/scalar_fields.cu:75-295): the
 * sum of n device floats, accumulated in FP64 (the reference accumulates float block sums).  Waits. */
int fr_sum(const float *d_data, long long n, double *sum, void *hip_stream);

/* fr_sum without any device allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same kernels and the same result bits as fr_sum; the partial sums of its at most
 * 1024 blocks and the total live in a static device array of 1025 doubles (8.2 KB on each GPU the
 * library runs on, held until the process ends), so every n >= 0 fits (a partial count above the
 * array would be refused with FR_ERR_INVALID).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other threads, other streams) race and must be serialised by
 * the caller.  Waits. */
int fr_sum_noalloc(const float *d_data, long long n, double *sum, void *hip_stream);

/* Average device time per launch of the solver's two iteration kernels (pass A: -Laplacian of psi
 * and its warp, density, velocity and psi_inv updates; pass B: psi update and max norm), from HIP
 * events between consecutive launches of `reps` iterations run in sequence on the context stream,
 * as the iteration loop runs them, after 150 untimed iterations that let the GPU clock settle (the
 * state is restored afterwards).  Any output pointer may be NULL.  No counterpart in the
 * reference: it is the per-kernel measurement of this solver. */
int fr_time_kernels(fr_ctx *ctx, int reps, float *ms_pass_a, float *ms_pass_b);

/* ---- the differentiated solver: K unrolled iterations in FP64 and their adjoint ---------------
 *
 * No counterpart in the reference, which has no gradients: these differentiate its iteration loop
 * (src/fastreg/cuda/solver.cu:295-444, one iteration = density update solver.cu:101-173, J_inv,
 * -Laplacian(psi) and its warp vector_fields.cu:231-300,503-535, potential gradient solver.cu:
 * 56-80, (Dv)v and the velocity Laplacian vector_fields.cu:307-404, velocity / psi / psi_inv
 * updates solver.cu:188-279) run for exactly K iterations (no convergence test) from
 * rho = RHO_0, v = v0, psi = psi_inv = identity, in FP64 with the reference's operation order.
 * The objective of the final state is
 *     E = 0.5 sum_p (T(p) - S o psi(p))^2
 *       + label_weight * 0.5 sum_{l=1}^{NO_LABELS-1} sum_p (1[LT(p) = l] - trilerp(1[LS = l], psi(p)))^2
 * (the SSD data energy of solver.cu:357-363 plus a label-overlap surrogate).  The backward pass
 * gives dE/dv0 (per voxel) and dE/dW_REG, dE/dTAU, with checkpoints of the state every `ckpt`
 * iterations (the segments are recomputed during the backward pass).  Branches (upwind, min/max,
 * floor, clamp) are differentiated within the branch the forward pass took.  Vector fields are
 * 3 doubles per voxel (x, y, z), device pointers; every call waits for its work. */
typedef struct fr_adjoint fr_adjoint;

/* Uses VOL_DIMS, ALPHA, RHO_0, p, C, RESET_ITER and NO_LABELS of `p` (W_REG and TAU are arguments
 * of the forward pass).  1 <= ckpt <= K. */
int fr_adjoint_create(const fr_params *p, int K, int ckpt, fr_adjoint **out);
int fr_adjoint_destroy(fr_adjoint *h);
/* target / source image and labels (device pointers, N voxels each), copied into the context */
int fr_adjoint_set_images(fr_adjoint *h, const float *d_tgt, const int16_t *d_tgt_lab, const float *d_src,
                          const int16_t *d_src_lab);
/* K iterations from v0 (d_v0: N*3 doubles, or NULL for v0 = 0); *energy receives E */
int fr_adjoint_forward(fr_adjoint *h, const double *d_v0, double w_reg, double tau, double label_weight,
                       double *energy);
/* gradients of E of the last forward pass: d_grad_v0 (N*3 doubles, may be NULL), *grad_w_reg, *grad_tau */
int fr_adjoint_backward(fr_adjoint *h, double *d_grad_v0, double *grad_w_reg, double *grad_tau);
/* the state after k iterations of the last forward pass, k a multiple of ckpt below K or k = K:
 * v and psi, psi_inv N*3 doubles, rho N doubles; any pointer may be NULL */
int fr_adjoint_get_state(fr_adjoint *h, int k, double *d_v, double *d_rho, double *d_psi, double *d_psi_inv);
/* device bytes held by the context; device time of the last forward and backward calls */
int fr_adjoint_info(fr_adjoint *h, size_t *device_bytes, float *forward_ms, float *backward_ms);
/* z-planes per block of the backward pass's two record gathers (the transposes of the trilinear
 * samples), which march tiles of tile_x x tile_y voxels; 0 restores the automatic choice (about
 * two blocks per CU).  Results do not depend on it.  No counterpart in the reference: a launch
 * shape of this implementation, for tests and tuning. */
int fr_adjoint_set_gather_chunk(fr_adjoint *h, int zc);
int fr_adjoint_get_gather_chunk(const fr_adjoint *h, int *tile_x, int *tile_y, int *zc);

/* One solver iteration as a forward/backward pair on caller states (device pointers; rho N
 * doubles, v / psi / psi_inv N*3), with the images of h, the given W_REG and TAU and the t of
 * iteration k (t_1 = ALPHA, RESET_ITER as the loop does).  fr_step_fwd writes the next state;
 * fr_step_bwd takes the adjoint of the next state in d_g_* and leaves there the adjoint of the
 * state, and returns dE/dW_REG and dE/dTAU of this iteration.  Both reuse h's work buffers, so
 * fr_adjoint_backward needs a new fr_adjoint_forward afterwards.  fr_step_fwd's outputs must not
 * overlap its inputs or each other, and fr_step_bwd's four adjoint arrays must not overlap each
 * other (FR_ERR_INVALID otherwise: the kernels read the inputs' neighbourhoods while writing).  No counterpart in the reference
 * (one iteration of src/fastreg/cuda/solver.cu:335-407, differentiated). */
int fr_step_fwd(fr_adjoint *h, int k, double w_reg, double tau, const double *d_rho, const double *d_v,
                const double *d_psi, const double *d_psi_inv, double *d_rho1, double *d_v1, double *d_psi1,
                double *d_psi_inv1);
int fr_step_bwd(fr_adjoint *h, int k, double w_reg, double tau, const double *d_rho, const double *d_v,
                const double *d_psi, const double *d_psi_inv, double *d_g_rho, double *d_g_v, double *d_g_psi,
                double *d_g_psi_inv, double *grad_w_reg, double *grad_tau);

/* Single operators of the iteration as FP64 forward/backward pairs over an X x Y x Z grid (device
 * pointers; every call waits).  The backward functions add into their outputs (clear them first
 * for a plain gradient).  No counterpart in the reference, which has no gradients; each cites the
 * forward operator it differentiates.
 * Trilinear sample at positions (N*3), clamped into the grid, all voxels (cuda_utils.hpp:60-136):
 * out(p) = trilerp(f, pos(p)); backward: d_grad_f += the transpose (FP64 atomics), d_grad_pos +=
 * d_grad_out * d trilerp / d pos (0 along clamped axes). */
int fr_warp_fwd(int X, int Y, int Z, const double *d_f, const double *d_pos, double *d_out);
int fr_warp_bwd(int X, int Y, int Z, const double *d_f, const double *d_pos, const double *d_grad_out,
                double *d_grad_f, double *d_grad_pos);
/* central-difference gradient of a scalar volume, (f(+1) - f(-1)) / 2 per axis on the interior,
 * 0 elsewhere (N*3 out; vector_fields.cu:155-192) */
int fr_gradient_fwd(int X, int Y, int Z, const double *d_f, double *d_grad);
int fr_gradient_bwd(int X, int Y, int Z, const double *d_grad_grad, double *d_grad_f);
/* -Laplacian of a vector field (N*3) on the interior, 0 elsewhere (vector_fields.cu:275-300) */
int fr_laplacian_fwd(int X, int Y, int Z, const double *d_psi, double *d_L);
int fr_laplacian_bwd(int X, int Y, int Z, const double *d_grad_L, double *d_grad_psi);

/* ---- device memory plumbing (for callers without a HIP binding of their own) -------------- */

int fr_dev_malloc(void **d_ptr, size_t bytes);
int fr_dev_free(void *d_ptr);
int fr_memcpy_htod(void *d_dst, const void *h_src, size_t bytes);
int fr_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes);
int fr_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes);
int fr_dev_memset(void *d_ptr, int value, size_t bytes);
int fr_dev_sync(void);
/* name, arch, CU count and memory of the current device, NUL-terminated */
int fr_device_info(char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif /* FASTREG_C_H */
