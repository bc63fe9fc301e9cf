// Internal declarations shared by the HIP kernels (fr_kernels.hip), the C-ABI (fr_capi.hip) and
// the differentiated solver (fr_adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frk {

// Grid dimensions; index = x + X*(y + Y*z).  3*X*Y*Z must fit in int (checked at create).
struct Dims {
    int X, Y, Z;
};

// Device-side control words of one run of the iteration loop.
struct Ctl {
    int stopped;      // set by the first launch that sees a stop condition
    int stop_iter;    // iterations completed when stopped
    int stop_reason;  // 1 = converged, 2 = NaN energy
    int pad;
};

// Tile of the iteration kernels: one 64-lane wave per x-row and TY_* rows per block, marching
// over a chunk of z-slices.  Pass A runs 8 rows per block (one block of 8 waves per CU), so that
// the rows a wave loads as y-neighbours are loaded into the CU's L1 once for all eight waves and
// the block can stage the region its trilinear taps read in LDS (pass B likewise).  Pass A takes
// x-neighbours from the adjacent lanes (x+-2) and pass B's taps reach x-1..x+1, so a wave computes
// the XS_* interior lanes of its row and keeps the edge lanes as a halo; consecutive tiles are
// XS_* columns apart.
constexpr int WAVE = 64;
constexpr int TY_A = 8, TY_B = 8;
constexpr int XS_A = WAVE - 4;
constexpr int XS_B = WAVE - 2;
// Pass A's columns per wave for a volume: XS_A, or WAVE - 2 when the interior row has 61 or 62
// columns (X = 63, 64).  Then one tile spans the row with lane l at x = l: its edge lanes are the
// boundary columns, and the only lanes next to them, x = 1 and x = X - 2, lie outside the upwind
// velocity Laplacian's range (x in [2, X - 3]), so no computed lane needs a neighbour two lanes
// past the wave.  A 64-wide volume thus runs half the x tiles, and its z-chunks are half as long.
inline int xs_a(Dims d) { return d.X - 2 > XS_A && d.X - 2 <= WAVE - 2 ? WAVE - 2 : XS_A; }
// Storage of velocity and density: FP32 (3 + 1 floats in two buffers), or one 8-byte record
// (v.x, v.y, v.z, rho) per voxel in bf16 or fp16.
constexpr int ST_F32 = 0, ST_BF16 = 1, ST_F16 = 2;
// Arithmetic of pass A: the reference's FP32 operation order (every product rounded before its
// sum, as the oracle computes it), or with its multiply-adds contracted into fma and the
// density's division as a reciprocal and a product (FP32 storage, single contexts only).
constexpr int ARITH_REF = 0, ARITH_FMA = 1;

// Pass B spreads its per-iteration 64-bit atomic maxima over NSLOT addresses (a multiple of the
// wave size: one wave reads them back).
constexpr int NSLOT = 256;

struct IterArgs {
    Dims d;
    int nbx, nby, nbz, nblocks;  // tile grid: x tiles, y tiles, z chunks (set by the launcher)
    int zc;                      // interior z-slices per chunk (set by the launcher)
    int xs;                      // x-columns computed per wave (set by the launcher)
    int ty;                      // rows (waves) per block (set by the launcher)
    const float *rho_o;          // density, read
    float *rho_n;                // density, written
    const float *v_o;            // velocity xyz-packed, read
    float *v_n;                  // velocity xyz-packed, written
    const float *pinv_o;         // psi_inv xyz-packed, read
    float *pinv_n;               // psi_inv xyz-packed, written
    float *psi;                  // psi xyz-packed (pass B updates in place)
    const float *S, *T;          // source / target intensity
    float alpha, m1t, m2, tau, w_reg, thr;
    int iter;                    // 1-based iteration number
    int check_energy;            // stop on NaN energy of this iteration
    int check_conv;              // test the previous iteration's max norm against thr
    Ctl *ctl;
    unsigned long long *maxpack; // per-iteration packed (norm bits, ~index) maxima, NSLOT each
    const double *energy;        // per-iteration energy
};

// Record `msg` as the calling thread's last error (fr_last_error) and return `code`.
int set_error(int code, const char *msg);

// launchers (all asynchronous on `s`)
hipError_t launch_init(Dims d, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab,
                       float *T, int16_t *LT, float *S, int16_t *LS, float *rho0, float *rho1, float *v0, float *v1,
                       float *pinv0, float *pinv1, float *psi, float rho_0, hipStream_t s);
// The iteration kernels over blocks of XS_* x TY_* columns by `zc` interior z-slices (the
// launch sets a's tile-grid fields).
// st: storage of v and rho (ST_F32, ST_BF16, ST_F16); in the 16-bit forms a.v_o / a.v_n point to
// the 8-byte (v, rho) records and a.rho_o / a.rho_n are unused
hipError_t launch_pass_a(const IterArgs &a, int zc, int st, hipStream_t s, int arith = ARITH_REF);
hipError_t launch_pass_b(const IterArgs &a, int zc, int st, hipStream_t s);
// new images and rho = rho_0 into a kept state (v, psi, psi_inv unchanged); the 16-bit storages
// reset rho in the (v, rho) records vr0/vr1, FP32 in rho0/rho1
hipError_t launch_reinit(Dims d, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab,
                         float *T, int16_t *LT, float *S, int16_t *LS, int st, float *rho0, float *rho1, unsigned *vr0,
                         unsigned *vr1, float rho_0, hipStream_t s);
// resident blocks per CU of pass A (1) or pass B (2)
hipError_t pass_occupancy(int which, int *blocks_per_cu);

// The arguments of one pass (1 = A, 2 = B) of a volume with its tile grid set for zc.
IterArgs pass_args(const IterArgs &a, int pass, int zc);

// Values of one iteration shared by every volume of a batch.
struct BatchIter {
    int iter;          // 1-based iteration number
    int cur;           // buffer parity: 0 = the launch-time arguments' old/new buffers, 1 = exchanged
    int check_energy;  // stop on NaN energy of this iteration
    int no_stop;       // 1: no stop test (timing runs of a fixed amount of work)
    float m1t, m2;
};

// One pass (1 = A, 2 = B) over a batch of volumes: d_base holds 3 arguments per volume (slot 0
// unused, then pass A, pass B; parity 0, grid set), d_starts the first block of each volume in this
// pass's grid (npairs + 1 entries, multiples of 8), nblocks the grid size.
hipError_t launch_pass_batch(int pass, const IterArgs *d_base, const int *d_starts, int npairs, int nblocks,
                             const BatchIter &it, int st, hipStream_t s);
hipError_t launch_energy(Dims d, const float *S, const float *T, const float *psi, double *partials, int nparts,
                         hipStream_t s);
hipError_t launch_energy_finalize(const double *partials, int nparts, double w_reg, double *e_data, double *e_reg,
                                  double *energy, hipStream_t s);
int energy_partials_count(Dims d);
// warp: def has `stride` floats per voxel (3 = packed xyz, 4 = float4)
hipError_t launch_warp(Dims d, const float *img, const int16_t *lab, const float *def, int stride, float *out_img,
                       int16_t *out_lab, hipStream_t s);
// Dice counts into DICE_COPIES copies of (overlap[no_labels], voxels[no_labels], bad), each block
// adding into copy blockIdx % DICE_COPIES (cleared by the caller; the caller sums the copies)
constexpr int DICE_COPIES = 32;
hipError_t launch_dice(Dims d, const int16_t *la, const int16_t *lb, int no_labels, unsigned *counts, hipStream_t s);
int min_det_partials_count(Dims d);
hipError_t launch_min_det(Dims d, const float *psi, float *part_true, float *part_ref, hipStream_t s);
hipError_t launch_pack4to3(const float *src4, float *dst3a, float *dst3b, int n, hipStream_t s);
// the reference's further public functions (float4 fields, 16 B per voxel)
hipError_t launch_get_inverse(Dims d, const float *psi4, float *pinv4, int sweeps, hipStream_t s);
hipError_t launch_interpolate_field4(Dims d, const float *f4, const float *psi4, float *out4, hipStream_t s);
// psi xyz-packed; *slot receives the packed (energy bits, ~index) maximum (atomic max: clear it first)
hipError_t launch_voxel_max_energy(Dims d, const float *T, const float *S, const float *psi, float w_reg,
                                   unsigned long long *slot, hipStream_t s);
int sum_partials_count(long long n);
hipError_t launch_sum(const float *a, long long n, double *partials, double *out, hipStream_t s);
// 16-bit (v, rho) records (2 unsigned per voxel): initial state (v = 0, rho = rho_0), replace v
// (float4 per voxel) and/or rho in both buffers, read v (float4) and/or rho
hipError_t launch_vr_init(int st, unsigned *r0, unsigned *r1, int n, float rho_0, hipStream_t s);
hipError_t launch_vr_set(int st, unsigned *r0, unsigned *r1, const float *v4, const float *rho, int n, hipStream_t s);
hipError_t launch_vr_get(int st, const unsigned *r, float *v4, float *rho, int n, hipStream_t s);
hipError_t launch_unpack3to4(const float *src3, float *dst4, int n, hipStream_t s);
// zeros into three buffers of b0, b1, b2 bytes (multiples of 4) in one kernel
hipError_t launch_zero3(void *p0, size_t b0, void *p1, size_t b1, void *p2, size_t b2, hipStream_t s);
// (x, y, z, 0) at every voxel of a float4 field
hipError_t launch_identity4(Dims d, float *dst4, hipStream_t s);
// out[k] = the maximum of iteration k's NSLOT slots, k < iters
hipError_t launch_slots_reduce(const unsigned long long *slots, int iters, unsigned long long *out, hipStream_t s);

}  // namespace frk
