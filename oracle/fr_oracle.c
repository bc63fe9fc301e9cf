/*
 * fastreg oracle — TEST INFRASTRUCTURE ONLY.  A CPU restatement of the reference's
 * per-iteration solver, compiled twice: *_f32 (FP32, the reference's operation order) and *_f64
 * (the same algorithm in double).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product library never links it.
 *
 * Reference functions restated (paths relative to the reference root):
 *   src/fastreg/cuda/volume.cu:29-46            k_initialise
 *   src/fastreg/cuda/vector_fields.cu:64-79     k_init_identity
 *   src/fastreg/cuda/vector_fields.cu:81-99     k_apply (+ include/cuda/cuda_utils.hpp:60-84)
 *   src/fastreg/cuda/vector_fields.cu:155-192   k_volume_gradient
 *   src/fastreg/cuda/vector_fields.cu:231-247   k_interpolate_laplacian (+ cuda_utils.hpp:86-107)
 *   src/fastreg/cuda/vector_fields.cu:275-300   k_laplacian
 *   src/fastreg/cuda/vector_fields.cu:307-346   k_laplacian_velocity
 *   src/fastreg/cuda/vector_fields.cu:352-404   k_dv_v
 *   src/fastreg/cuda/vector_fields.cu:503-535   k_jacobian
 *   src/fastreg/cuda/vector_fields.cu:537-585   k_deformation_jacobian
 *   src/fastreg/cuda/solver.cu:56-80            k_potential_gradient (+ cuda_utils.hpp:389-396)
 *   src/fastreg/cuda/solver.cu:101-144          k_divergence
 *   src/fastreg/cuda/solver.cu:159-173          k_update_density
 *   src/fastreg/cuda/solver.cu:188-211          k_update_velocity
 *   src/fastreg/cuda/solver.cu:230-247          k_update_psi (+ cuda_utils.hpp:109-136)
 *   src/fastreg/cuda/solver.cu:263-279          k_update_psi_inv (+ cuda_utils.hpp:379-385)
 *   src/fastreg/cuda/solver.cu:295-444          or_run
 *   src/fastreg/cuda/solver.cu:9-50             k_dice_counts
 *   src/fastreg/cuda/reductor.cu:426-539, src/fastreg/reductor.cpp:107-120   k_reduce_max
 *   src/fastreg/cuda/reductor.cu:11-112,224-324, src/fastreg/reductor.cpp:45-62  k_energy
 *   src/fastreg/fastreg.cpp:3-22,58-75          or_init / or_outputs
 *   (not in the reference)                      k_min_det: min over the interior of the true det and
 *                                               of det_ref of the central-difference D psi
 *
 * How each function is pinned (tests/test_oracle_kat.py; the reference ships no tests, golden
 * vectors or fixtures, and cannot be compiled here: it is CUDA with Boost/OpenCV/ITK):
 *   k_apply            identity warp = exact copy; integer shift = exact clamped shift; random
 *                      positions vs scipy.ndimage.map_coordinates(order=1, mode='nearest') and
 *                      labels vs floor(clip(c)) by brute force
 *   k_volume_gradient  linear ramp -> exact slope on the interior, 0 on the faces; numpy gradient
 *   k_jacobian         identity -> I; linear map A x -> A on the interior
 *   k_deformation_jacobian  identity -> 0 everywhere; affine psi -> A - I incl. mirrored faces
 *   k_laplacian        -Laplacian(identity) = 0 exactly; quadratic field -> constant -6a... (closed form)
 *   k_interpolate_laplacian  identity psi -> L copied; map_coordinates on each component
 *   k_laplacian_velocity  linear v with vy,vz >= 0 -> 0; vy < 0 -> 4*slope (the reference's
 *                      index bug; a textbook upwind stencil gives 0); ranges vs brute force
 *   k_dv_v             constant v -> 0; 1-D Burgers flux closed form on a ramp
 *   det_ref            det_ref(I)=1; diagonal -> product; a matrix with only M20 set -> -M20*...
 *                      (second-order term that the true det lacks)
 *   k_divergence       constant v and rho -> 0; a single moving voxel conserves mass (sum of div
 *                      over the grid = 0 away from the faces)
 *   solver             iteration-1 closed form v1 = -alpha*m2/(rho0+1e-5)*(S-T)*grad S,
 *                      psi1 = id + alpha v1, psi_inv1 = id - alpha v1; S = T is a fixed point;
 *                      FP32 vs FP64 runs agree to rounding; t schedule with RESET_ITER
 *   k_update_density   one donor-cell step of a single moving voxel along each axis, both signs:
 *                      exactly alpha u rho0 moves between p and p+e in the upwind direction, mass
 *                      conserved; the FP32 result is alpha*div + rho rounded once (a velocity for
 *                      which rounding the product first gives another float)
 *   k_energy           S = T at psi = id -> 0; ramp S = T = x/8 with psi = id + (1,0,0) ->
 *                      e_data = 0.5 (X-1) Y Z / 64; psi = id + A x -> e_reg = 0.5 sum_a |A col a|^2
 *                      (dim_a - 2) N / dim_a (mirrored faces); random images at psi = id vs numpy;
 *                      the logged energy of iteration 1
 *   k_min_det          psi = A x -> det(A) exactly (and det_ref(A) != det(A)); psi_x = x + b x^2 ->
 *                      1 + 2b(X-2), the minimum over the interior only
 *   k_dice_counts      identical labels -> 1; last slice per axis ignored (brute force)
 *   k_reduce_max       brute force over norms with first-index ties
 */

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Rounding of the 16-bit storage forms (velocity and density stored as bf16 or fp16, arithmetic
 * in FP32): round to nearest even, as IEEE 754 conversion does.  Pinned in the tests against
 * numpy's float16 and torch's bfloat16 conversions. */
static uint16_t or_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t s = (x >> 16) & 0x8000u, a = x & 0x7FFFFFFFu;
    if (a > 0x7F800000u) return (uint16_t) (s | 0x7E00u | ((a >> 13) & 0x3FFu)); /* NaN (quiet) */
    if (a >= 0x477FF000u) return (uint16_t) (s | 0x7C00u); /* rounds to infinity (>= 65520) */
    if (a >= 0x38800000u) {                                /* normal half: >= 2^-14 */
        const uint32_t r = a - 0x38000000u;                /* exponent rebiased from 127 to 15 */
        uint32_t h = r >> 13;
        const uint32_t rem = r & 0x1FFFu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        return (uint16_t) (s | h);
    }
    const uint32_t e = a >> 23;
    if (e < 102) return (uint16_t) s; /* below 2^-25: rounds to zero */
    /* subnormal half: the value in units of 2^-24 is m * 2^(e - 126), m with its leading 1 */
    const uint32_t m = (a & 0x7FFFFFu) | 0x800000u, sh = 126 - e;
    uint32_t h = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
    if (rem > half || (rem == half && (h & 1u))) ++h;
    return (uint16_t) (s | h);
}

static float or_f16_to_f32(uint16_t h) {
    const uint32_t s = (uint32_t) (h & 0x8000u) << 16, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    uint32_t x;
    if (e == 0x1Fu) x = s | 0x7F800000u | (m << 13);
    else if (e != 0) x = s | ((e + 112u) << 23) | (m << 13);
    else if (m == 0) x = s;
    else { /* subnormal half: m * 2^-24, exact in float */
        float f = (float) m * 0x1p-24f;
        memcpy(&x, &f, 4);
        x |= s;
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

static float or_bf16_round(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) x = (x | 0x00400000u) & 0xFFFF0000u; /* NaN (quiet) */
    else x = (x + 0x7FFFu + ((x >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&f, &x, 4);
    return f;
}

/* storage 0: FP32 (no rounding), 1: bf16, 2: fp16 */
static float or_q16(int storage, float f) {
    return storage == 1 ? or_bf16_round(f) : storage == 2 ? or_f16_to_f32(or_f32_to_f16(f)) : f;
}

/* the rounding applied to an array (tests pin it) */
void or_round16(int storage, float *a, long n) {
    for (long i = 0; i < n; ++i) a[i] = or_q16(storage, a[i]);
}

#define REAL float
#define SFX _f32
#include "fr_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "fr_oracle_impl.h"
#undef REAL
#undef SFX

/* run with the log going to stdout (ctypes callers cannot pass a FILE*) */
int or_run_stdout_f32(void *h, const or_params_f32 *prm, float *max_hist, double *energy_hist, int *stop_reason, int log) {
    int r = or_run_f32(h, prm, log ? stdout : NULL, max_hist, energy_hist, stop_reason);
    fflush(stdout);
    return r;
}
int or_run_stdout_f64(void *h, const or_params_f64 *prm, double *max_hist, double *energy_hist, int *stop_reason, int log) {
    int r = or_run_f64(h, prm, log ? stdout : NULL, max_hist, energy_hist, stop_reason);
    fflush(stdout);
    return r;
}

int or_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
