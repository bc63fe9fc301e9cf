/*
 * CPU restatement of fastreg's per-iteration registration solver — TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: the checker the HIP product path is compared against, and the CPU
 * baseline that bench.py times.  It is never linked into the product library, and only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * It is included twice by fr_oracle.c: once with REAL=float / SFX=_f32 ("FP32 ref-order mode":
 * the reference's kernels one loop each, unfused, in the reference's operation order and with
 * its quirks), once with REAL=double / SFX=_f64 (the same algorithm in double, the FP64 oracle).
 *
 * Layouts follow the reference: volumes are (intensity, label-as-float) pairs, vector fields
 * are 4-wide with w = 0, Jacobians are 4x4 row-major with rows = components and columns =
 * (d/dx, d/dy, d/dz); index = x + X*(y + Y*z).
 *
 * Floating point: built with -ffp-contract=off, so every a*b+c below rounds twice unless it is
 * written as FMA().  The reference's explicit __f*_rn ops are plain ops here, its explicit
 * fma() (lerp) is FMA(), and `rho += alpha*div` is written as FMA() (the contraction nvcc makes
 * of that statement).  Division and sqrt are IEEE; norm() is sqrt rounded toward zero, built
 * from the correctly rounded sqrt and an exact FMA residual.
 *
 * Parity pins for each function are listed in the header comment of fr_oracle.c.
 */

#ifndef REAL
#error "define REAL and SFX before including fr_oracle_impl.h"
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define N_(name) CAT(name, SFX)

#define V2 N_(or_v2)
#define V4 N_(or_v4)
#define M4 N_(or_m4)

typedef struct { REAL x, y; } V2;
typedef struct { REAL x, y, z, w; } V4;
typedef struct { V4 r[4]; } M4;

/* ---------------------------------------------------------------------------------------------
 * scalar helpers
 * ------------------------------------------------------------------------------------------- */

static inline REAL N_(fma_)(REAL a, REAL b, REAL c) { return (sizeof(REAL) == 4) ? (REAL) fmaf((float) a, (float) b, (float) c) : (REAL) fma((double) a, (double) b, (double) c); }
#define FMA N_(fma_)

static inline REAL N_(sqrt_)(REAL a) { return (sizeof(REAL) == 4) ? (REAL) sqrtf((float) a) : (REAL) sqrt((double) a); }
static inline REAL N_(nextdown_)(REAL a) { return (sizeof(REAL) == 4) ? (REAL) nextafterf((float) a, 0.f) : (REAL) nextafter((double) a, 0.0); }
static inline REAL N_(fabs_)(REAL a) { return a < 0 ? -a : a; }
static inline REAL N_(fmax_)(REAL a, REAL b) { return (sizeof(REAL) == 4) ? (REAL) fmaxf((float) a, (float) b) : (REAL) fmax((double) a, (double) b); }
static inline REAL N_(fmin_)(REAL a, REAL b) { return (sizeof(REAL) == 4) ? (REAL) fminf((float) a, (float) b) : (REAL) fmin((double) a, (double) b); }
static inline REAL N_(pow_)(REAL a, REAL b) { return (sizeof(REAL) == 4) ? (REAL) powf((float) a, (float) b) : (REAL) pow((double) a, (double) b); }

/* lerp(v0, v1, t) = t*v0 + (1-t)*v1 as fma(t, v0, fma(-t, v1, v1)) — cuda_utils.hpp:43-46 */
static inline REAL N_(lerp)(REAL v0, REAL v1, REAL t) { return FMA(t, v0, FMA(-t, v1, v1)); }

/* sqrt rounded toward zero: the correctly rounded sqrt, stepped down one ulp when its square
 * exceeds the argument (the residual s*s - a is exact in one FMA).  cuda_utils.hpp:251-253 uses
 * a round-down sqrt for norm(float4). */
static inline REAL N_(sqrt_rd)(REAL a) {
    REAL s = N_(sqrt_)(a);
    if (s > 0 && FMA(s, s, -a) > 0) s = N_(nextdown_)(s);
    return s;
}

/* ---------------------------------------------------------------------------------------------
 * 4-wide vector ops, w forced to 0 (cuda_utils.hpp:207-257)
 * ------------------------------------------------------------------------------------------- */

static inline V4 N_(v4)(REAL x, REAL y, REAL z) { V4 r = {x, y, z, 0}; return r; }
static inline V4 N_(add4)(V4 a, V4 b) { return N_(v4)(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V4 N_(sub4)(V4 a, V4 b) { return N_(v4)(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V4 N_(mul4)(REAL m, V4 a) { return N_(v4)(a.x * m, a.y * m, a.z * m); }
static inline REAL N_(norm4)(V4 a) { return N_(sqrt_rd)(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline REAL N_(norm_sq4)(V4 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

/* Mat4f * float4: rows 0-2, ((a*x + b*y) + c*z) — cuda_utils.hpp:379-385 */
static inline V4 N_(matvec)(const M4 *M, V4 v) {
    return N_(v4)((M->r[0].x * v.x + M->r[0].y * v.y) + M->r[0].z * v.z,
                  (M->r[1].x * v.x + M->r[1].y * v.y) + M->r[1].z * v.z,
                  (M->r[2].x * v.x + M->r[2].y * v.y) + M->r[2].z * v.z);
}

/* The reference's det(Mat4f) exactly as parenthesised in cuda_utils.hpp:389-396.  It is NOT the
 * determinant:  M00 M11 M22 + M10 M21 M02 + M20 M01 M12 - [(M02 M11 + M20)(M12 M21 M00) + M22 M01 M10]. */
static inline REAL N_(det_ref)(const M4 *M) {
    REAL a = ((M->r[0].x * M->r[1].y) * M->r[2].z + (M->r[1].x * M->r[2].y) * M->r[0].z) + (M->r[2].x * M->r[0].y) * M->r[1].z;
    REAL b = ((M->r[0].z * M->r[1].y) + M->r[2].x) * ((M->r[1].z * M->r[2].y) * M->r[0].x) + (M->r[2].z * M->r[0].y) * M->r[1].x;
    return a - b;
}

/* the true determinant of rows 0-2 (used only for the min-det-J report, not in the solver) */
static inline REAL N_(det_true)(const M4 *M) {
    return M->r[0].x * (M->r[1].y * M->r[2].z - M->r[1].z * M->r[2].y) -
           M->r[0].y * (M->r[1].x * M->r[2].z - M->r[1].z * M->r[2].x) +
           M->r[0].z * (M->r[1].x * M->r[2].y - M->r[1].y * M->r[2].x);
}

/* ---------------------------------------------------------------------------------------------
 * grid
 * ------------------------------------------------------------------------------------------- */

typedef struct { int X, Y, Z; } N_(or_dims);
#define DIMS N_(or_dims)

static inline long N_(ix)(DIMS d, int x, int y, int z) { return (long) x + (long) d.X * ((long) y + (long) d.Y * z); }
#define IX N_(ix)

/* get_global_idx: out-of-range -> voxel 0 (cuda_utils.hpp:31-37) */
static inline long N_(gix0)(DIMS d, int x, int y, int z) {
    if (x < 0 || x >= d.X || y < 0 || y >= d.Y || z < 0 || z >= d.Z) return 0;
    return IX(d, x, y, z);
}

/* Clamp a sample position into [0, dim-1], floor it, and return the corner plus weights.  The
 * reference reads the +1 neighbour without a bounds check (by linear address, or voxel 0 via
 * get_global_idx), but its weight is then exactly 0; callers clamp that neighbour's index into
 * the grid, which gives the same result for finite data (cuda_utils.hpp:60-136). */
typedef struct { int gx, gy, gz, hx, hy, hz; REAL a, b, c; } N_(or_corner);
#define CORNER N_(or_corner)

static inline CORNER N_(corner)(DIMS d, REAL px, REAL py, REAL pz) {
    CORNER k;
    REAL cx = N_(fmin_)(N_(fmax_)(0, px), (REAL) (d.X - 1));
    REAL cy = N_(fmin_)(N_(fmax_)(0, py), (REAL) (d.Y - 1));
    REAL cz = N_(fmin_)(N_(fmax_)(0, pz), (REAL) (d.Z - 1));
    k.gx = (int) floor((double) cx);
    k.gy = (int) floor((double) cy);
    k.gz = (int) floor((double) cz);
    k.a = cx - (REAL) k.gx;
    k.b = cy - (REAL) k.gy;
    k.c = cz - (REAL) k.gz;
    k.hx = k.gx + 1 < d.X ? k.gx + 1 : k.gx;
    k.hy = k.gy + 1 < d.Y ? k.gy + 1 : k.gy;
    k.hz = k.gz + 1 < d.Z ? k.gz + 1 : k.gz;
    return k;
}

/* trilinear blend of 8 corner values fXYZ (X,Y,Z in {0,1}: 1 = the +1 neighbour), in the
 * reference's nesting: lerp over a of (lerp over b of (lerp over c)) — cuda_utils.hpp:75-80 */
static inline REAL N_(tri)(REAL f111, REAL f110, REAL f101, REAL f100, REAL f011, REAL f010, REAL f001, REAL f000,
                           REAL a, REAL b, REAL c) {
    return N_(lerp)(N_(lerp)(N_(lerp)(f111, f110, c), N_(lerp)(f101, f100, c), b),
                    N_(lerp)(N_(lerp)(f011, f010, c), N_(lerp)(f001, f000, c), b), a);
}

/* interpolate_intensity: trilinear .x, label .y at the floor corner (cuda_utils.hpp:60-84) */
static inline V2 N_(interp_intensity)(DIMS d, const V2 *vol, REAL px, REAL py, REAL pz) {
    CORNER k = N_(corner)(d, px, py, pz);
    V2 r;
    r.x = N_(tri)(vol[IX(d, k.hx, k.hy, k.hz)].x, vol[IX(d, k.hx, k.hy, k.gz)].x, vol[IX(d, k.hx, k.gy, k.hz)].x,
                  vol[IX(d, k.hx, k.gy, k.gz)].x, vol[IX(d, k.gx, k.hy, k.hz)].x, vol[IX(d, k.gx, k.hy, k.gz)].x,
                  vol[IX(d, k.gx, k.gy, k.hz)].x, vol[IX(d, k.gx, k.gy, k.gz)].x, k.a, k.b, k.c);
    r.y = vol[IX(d, k.gx, k.gy, k.gz)].y;
    return r;
}

/* interpolate_field for 4-wide fields, component-wise, w = 0 (cuda_utils.hpp:86-136) */
static inline V4 N_(interp_field)(DIMS d, const V4 *f, REAL px, REAL py, REAL pz) {
    CORNER k = N_(corner)(d, px, py, pz);
    const V4 *p111 = &f[IX(d, k.hx, k.hy, k.hz)], *p110 = &f[IX(d, k.hx, k.hy, k.gz)];
    const V4 *p101 = &f[IX(d, k.hx, k.gy, k.hz)], *p100 = &f[IX(d, k.hx, k.gy, k.gz)];
    const V4 *p011 = &f[IX(d, k.gx, k.hy, k.hz)], *p010 = &f[IX(d, k.gx, k.hy, k.gz)];
    const V4 *p001 = &f[IX(d, k.gx, k.gy, k.hz)], *p000 = &f[IX(d, k.gx, k.gy, k.gz)];
    V4 r;
    r.x = N_(tri)(p111->x, p110->x, p101->x, p100->x, p011->x, p010->x, p001->x, p000->x, k.a, k.b, k.c);
    r.y = N_(tri)(p111->y, p110->y, p101->y, p100->y, p011->y, p010->y, p001->y, p000->y, k.a, k.b, k.c);
    r.z = N_(tri)(p111->z, p110->z, p101->z, p100->z, p011->z, p010->z, p001->z, p000->z, k.a, k.b, k.c);
    r.w = 0;
    return r;
}

static inline int N_(interior)(DIMS d, int x, int y, int z) {
    return x >= 1 && x <= d.X - 2 && y >= 1 && y <= d.Y - 2 && z >= 1 && z <= d.Z - 2;
}

#define FOR_ALL(d) \
    _Pragma("omp parallel for schedule(static)") for (int z = 0; z < (d).Z; ++z) for (int y = 0; y < (d).Y; ++y) for (int x = 0; x < (d).X; ++x)
#define FOR_INTERIOR(d) \
    _Pragma("omp parallel for schedule(static)") for (int z = 1; z <= (d).Z - 2; ++z) for (int y = 1; y <= (d).Y - 2; ++y) for (int x = 1; x <= (d).X - 2; ++x)

/* ---------------------------------------------------------------------------------------------
 * kernels, one per reference kernel
 * ------------------------------------------------------------------------------------------- */

/* initialise_kernel: pack (intensity, (float) label) — volume.cu:29-46 */
static void N_(k_initialise)(DIMS d, V2 *vol, const float *img, const int16_t *lab) {
    FOR_ALL(d) {
        long i = IX(d, x, y, z);
        vol[i].x = (REAL) img[i];
        vol[i].y = (REAL) lab[i];
    }
}

/* init_identity_kernel: psi(p) = (x, y, z, 0) — vector_fields.cu:64-79 */
static void N_(k_init_identity)(DIMS d, V4 *psi) {
    FOR_ALL(d) { psi[IX(d, x, y, z)] = N_(v4)((REAL) x, (REAL) y, (REAL) z); }
}

/* apply_kernel: out(p) = interpolate_intensity(phi, psi(p)) over all voxels — vector_fields.cu:81-99 */
static void N_(k_apply)(DIMS d, const V2 *phi, V2 *out, const V4 *psi) {
    FOR_ALL(d) {
        long i = IX(d, x, y, z);
        out[i] = N_(interp_intensity)(d, phi, psi[i].x, psi[i].y, psi[i].z);
    }
}

/* VolumeDifferentiator: central differences of .x / 2, interior only — vector_fields.cu:155-192 */
static void N_(k_volume_gradient)(DIMS d, const V2 *vol, V4 *grad) {
    FOR_INTERIOR(d) {
        REAL nx = (vol[IX(d, x + 1, y, z)].x - vol[IX(d, x - 1, y, z)].x) / (REAL) 2;
        REAL ny = (vol[IX(d, x, y + 1, z)].x - vol[IX(d, x, y - 1, z)].x) / (REAL) 2;
        REAL nz = (vol[IX(d, x, y, z + 1)].x - vol[IX(d, x, y, z - 1)].x) / (REAL) 2;
        grad[IX(d, x, y, z)] = N_(v4)(nx, ny, nz);
    }
}

/* Differentiator::jacobian: rows = components, cols = (dx, dy, dz), interior — vector_fields.cu:503-535 */
static void N_(k_jacobian)(DIMS d, const V4 *psi, M4 *J) {
    FOR_INTERIOR(d) {
        V4 jx = N_(mul4)((REAL) 0.5, N_(sub4)(psi[IX(d, x + 1, y, z)], psi[IX(d, x - 1, y, z)]));
        V4 jy = N_(mul4)((REAL) 0.5, N_(sub4)(psi[IX(d, x, y + 1, z)], psi[IX(d, x, y - 1, z)]));
        V4 jz = N_(mul4)((REAL) 0.5, N_(sub4)(psi[IX(d, x, y, z + 1)], psi[IX(d, x, y, z - 1)]));
        M4 m;
        m.r[0] = N_(v4)(jx.x, jy.x, jz.x);
        m.r[1] = N_(v4)(jx.y, jy.y, jz.y);
        m.r[2] = N_(v4)(jx.z, jy.z, jz.z);
        m.r[3] = N_(v4)(0, 0, 0);
        J[IX(d, x, y, z)] = m;
    }
}

/* Differentiator::deformation_jacobian: D(psi - id) with mirrored neighbours at the faces, all
 * voxels — vector_fields.cu:537-585 */
static void N_(k_deformation_jacobian)(DIMS d, const V4 *psi, M4 *J) {
    FOR_ALL(d) {
        int x1 = x + 1, x2 = x - 1, y1 = y + 1, y2 = y - 1, z1 = z + 1, z2 = z - 1;
        if (x == 0) x2 = x + 1; else if (x == d.X - 1) x1 = x - 1;
        if (y == 0) y2 = y + 1; else if (y == d.Y - 1) y1 = y - 1;
        if (z == 0) z2 = z + 1; else if (z == d.Z - 1) z1 = z - 1;
#define DISP(xx, yy, zz) N_(sub4)(psi[IX(d, xx, yy, zz)], N_(v4)((REAL) (xx), (REAL) (yy), (REAL) (zz)))
        V4 jx = N_(mul4)((REAL) 0.5, N_(sub4)(DISP(x1, y, z), DISP(x2, y, z)));
        V4 jy = N_(mul4)((REAL) 0.5, N_(sub4)(DISP(x, y1, z), DISP(x, y2, z)));
        V4 jz = N_(mul4)((REAL) 0.5, N_(sub4)(DISP(x, y, z1), DISP(x, y, z2)));
#undef DISP
        M4 m;
        m.r[0] = N_(v4)(jx.x, jy.x, jz.x);
        m.r[1] = N_(v4)(jx.y, jy.y, jz.y);
        m.r[2] = N_(v4)(jx.z, jy.z, jz.z);
        m.r[3] = N_(v4)(0, 0, 0);
        J[IX(d, x, y, z)] = m;
    }
}

/* SecondOrderDifferentiator::laplacian: L = -((((((-6 psi + psi(x+1)) + psi(x-1)) + psi(y+1)) +
 * psi(y-1)) + psi(z+1)) + psi(z-1)), interior — vector_fields.cu:275-300 */
static void N_(k_laplacian)(DIMS d, const V4 *psi, V4 *L) {
    FOR_INTERIOR(d) {
        V4 s = N_(mul4)((REAL) -6, psi[IX(d, x, y, z)]);
        s = N_(add4)(s, psi[IX(d, x + 1, y, z)]);
        s = N_(add4)(s, psi[IX(d, x - 1, y, z)]);
        s = N_(add4)(s, psi[IX(d, x, y + 1, z)]);
        s = N_(add4)(s, psi[IX(d, x, y - 1, z)]);
        s = N_(add4)(s, psi[IX(d, x, y, z + 1)]);
        s = N_(add4)(s, psi[IX(d, x, y, z - 1)]);
        L[IX(d, x, y, z)] = N_(mul4)((REAL) -1, s);
    }
}

/* interpolate_laplacian_kernel: L_o(p) = trilerp(L, psi(p)), all voxels — vector_fields.cu:231-247 */
static void N_(k_interpolate_laplacian)(DIMS d, const V4 *L, V4 *L_o, const V4 *psi) {
    FOR_ALL(d) {
        long i = IX(d, x, y, z);
        L_o[i] = N_(interp_field)(d, L, psi[i].x, psi[i].y, psi[i].z);
    }
}

/* SecondOrderDifferentiator::laplacian_velocity: one-sided second differences, with the
 * reference's indices (y-1 and z-1 where +1 is meant) and its ranges x in [2,X-3],
 * y in [2,Y-2], z in [2,Z-3]; at y = Y-2 the (y+2) neighbour is the linear address
 * idx(x, Y, z) = idx(x, 0, z+1) — vector_fields.cu:307-346 */
static void N_(k_laplacian_velocity)(DIMS d, const V4 *v, V4 *L) {
#pragma omp parallel for schedule(static)
    for (int z = 2; z <= d.Z - 3; ++z)
        for (int y = 2; y <= d.Y - 2; ++y)
            for (int x = 2; x <= d.X - 3; ++x) {
                V4 vel = v[IX(d, x, y, z)];
                V4 vxx = (vel.x < 0)
                             ? N_(add4)(N_(sub4)(v[IX(d, x + 2, y, z)], N_(mul4)((REAL) 2, v[IX(d, x + 1, y, z)])), vel)
                             : N_(add4)(N_(sub4)(vel, N_(mul4)((REAL) 2, v[IX(d, x - 1, y, z)])), v[IX(d, x - 2, y, z)]);
                V4 vyy = (vel.y < 0)
                             ? N_(add4)(N_(sub4)(v[IX(d, x, y + 2, z)], N_(mul4)((REAL) 2, v[IX(d, x, y - 1, z)])), vel)
                             : N_(add4)(N_(sub4)(vel, N_(mul4)((REAL) 2, v[IX(d, x, y - 1, z)])), v[IX(d, x, y - 2, z)]);
                V4 vzz = (vel.z < 0)
                             ? N_(add4)(N_(sub4)(v[IX(d, x, y, z + 2)], N_(mul4)((REAL) 2, v[IX(d, x, y, z - 1)])), vel)
                             : N_(add4)(N_(sub4)(vel, N_(mul4)((REAL) 2, v[IX(d, x, y, z - 1)])), v[IX(d, x, y, z - 2)]);
                L[IX(d, x, y, z)] = N_(mul4)((REAL) -1, N_(add4)(N_(add4)(vxx, vyy), vzz));
            }
}

/* calc_dv_v_kernel: (Dv)v, entropy scheme on the diagonal, upwind off it, interior —
 * vector_fields.cu:352-404 */
static void N_(k_dv_v)(DIMS d, const V4 *v, V4 *out) {
    FOR_INTERIOR(d) {
        V4 c = v[IX(d, x, y, z)];
        V4 xp = v[IX(d, x + 1, y, z)], xm = v[IX(d, x - 1, y, z)];
        V4 yp = v[IX(d, x, y + 1, z)], ym = v[IX(d, x, y - 1, z)];
        V4 zp = v[IX(d, x, y, z + 1)], zm = v[IX(d, x, y, z - 1)];
#define MX(a) N_(fmax_)((a), (REAL) 0)
#define MN(a) N_(fmin_)((a), (REAL) 0)
        REAL ax1 = ((MX(c.x) * MX(c.x) - MN(c.x) * MN(c.x)) + MN(xp.x) * MN(xp.x)) - MX(xp.x) * MX(xp.x);
        REAL ax2 = (c.y > 0) ? c.y * (c.x - ym.x) : c.y * (yp.x - c.x);
        REAL ax3 = (c.z > 0) ? c.z * (c.x - zm.x) : c.z * (zp.x - c.x);
        REAL ay1 = (c.x > 0) ? c.x * (c.y - xm.y) : c.x * (xp.y - c.y);
        REAL ay2 = ((MX(c.y) * MX(c.y) - MN(c.y) * MN(c.y)) + MN(yp.y) * MN(yp.y)) - MX(yp.y) * MX(yp.y);
        REAL ay3 = (c.z > 0) ? c.z * (c.y - zm.y) : c.z * (zp.y - c.y);
        REAL az1 = (c.x > 0) ? c.x * (c.z - xm.z) : c.x * (xp.z - c.z);
        REAL az2 = (c.y > 0) ? c.y * (c.z - ym.z) : c.y * (yp.z - c.z);
        REAL az3 = ((MX(c.z) * MX(c.z) - MN(c.z) * MN(c.z)) + MN(zp.z) * MN(zp.z)) - MX(zp.z) * MX(zp.z);
#undef MX
#undef MN
        out[IX(d, x, y, z)] = N_(v4)(((REAL) 0.5 * ax1 + ax2) + ax3, (ay1 + (REAL) 0.5 * ay2) + ay3, (az1 + az2) + (REAL) 0.5 * az3);
    }
}

/* calculate_potential_gradient_kernel: ((S - T o psi_inv) * grad S + w_reg * L_o) * det_ref(J_inv),
 * interior — solver.cu:56-80 */
static void N_(k_potential_gradient)(DIMS d, const V2 *phi_n, const V2 *phi_g_psi_inv, const V4 *nabla, const M4 *J_inv,
                                     const V4 *L_o, V4 *nabla_U, REAL w_reg) {
    FOR_INTERIOR(d) {
        long i = IX(d, x, y, z);
        REAL det = N_(det_ref)(&J_inv[i]);
        REAL diff = phi_n[i].x - phi_g_psi_inv[i].x;
        nabla_U[i] = N_(mul4)(det, N_(add4)(N_(mul4)(diff, nabla[i]), N_(mul4)(w_reg, L_o[i])));
    }
}

/* calc_divergence_kernel: upwind flux divergence, summed left to right, interior —
 * solver.cu:101-144 */
static void N_(k_divergence)(DIMS d, const REAL *rho, const V4 *v, REAL *div) {
    FOR_INTERIOR(d) {
        long i = IX(d, x, y, z);
        V4 vel = v[i];
        REAL p = rho[i];
        REAL t11 = (vel.x > 0) ? -vel.x * p : -vel.x * rho[IX(d, x + 1, y, z)];
        REAL vxm = v[IX(d, x - 1, y, z)].x;
        REAL t12 = (vxm > 0) ? vxm * rho[IX(d, x - 1, y, z)] : vxm * p;
        REAL t21 = (vel.y > 0) ? -vel.y * p : -vel.y * rho[IX(d, x, y + 1, z)];
        REAL vym = v[IX(d, x, y - 1, z)].y;
        REAL t22 = (vym > 0) ? vym * rho[IX(d, x, y - 1, z)] : vym * p;
        REAL t31 = (vel.z > 0) ? -vel.z * p : -vel.z * rho[IX(d, x, y, z + 1)];
        REAL vzm = v[IX(d, x, y, z - 1)].z;
        REAL t32 = (vzm > 0) ? vzm * rho[IX(d, x, y, z - 1)] : vzm * p;
        div[i] = ((((t11 + t12) + t21) + t22) + t31) + t32;
    }
}

/* update_density_kernel: rho += alpha * div (one fused multiply-add), interior — solver.cu:159-173 */
static void N_(k_update_density)(DIMS d, REAL *rho, const REAL *div, REAL alpha) {
    FOR_INTERIOR(d) {
        long i = IX(d, x, y, z);
        rho[i] = FMA(alpha, div[i], rho[i]);
    }
}

/* update_velocity_kernel, interior — solver.cu:188-211.  m1t = (p+1)/t and
 * m2 = C * p^2 * t^(p-2) are per-launch scalars. */
static void N_(k_update_velocity)(DIMS d, V4 *v, const V4 *dv_v, const V4 *delta_v, const REAL *rho, const V4 *nabla_U,
                                  REAL alpha, REAL m1t, REAL tau, REAL m2) {
    FOR_INTERIOR(d) {
        long i = IX(d, x, y, z);
        V4 friction = N_(mul4)(m1t, v[i]);
        V4 diffusion = N_(mul4)(tau, delta_v[i]);
        REAL q = m2 / (N_(fabs_)(rho[i]) + (REAL) 1e-5f);
        V4 potential = N_(mul4)(q, nabla_U[i]);
        V4 upd = N_(mul4)(alpha, N_(add4)(N_(sub4)(N_(add4)(friction, dv_v[i]), diffusion), potential));
        v[i].x = v[i].x - upd.x;
        v[i].y = v[i].y - upd.y;
        v[i].z = v[i].z - upd.z;
    }
}

/* update_psi_kernel: u = alpha * trilerp(v, psi(p)); psi += u; updates = u; interior —
 * solver.cu:230-247 */
static void N_(k_update_psi)(DIMS d, V4 *psi, const V4 *v, V4 *updates, REAL alpha) {
    FOR_INTERIOR(d) {
        long i = IX(d, x, y, z);
        V4 u = N_(mul4)(alpha, N_(interp_field)(d, v, psi[i].x, psi[i].y, psi[i].z));
        psi[i].x = psi[i].x + u.x;
        psi[i].y = psi[i].y + u.y;
        psi[i].z = psi[i].z + u.z;
        updates[i] = u;
    }
}

/* update_psi_inv_kernel: psi_inv -= alpha * (J_inv * v), interior — solver.cu:263-279 */
static void N_(k_update_psi_inv)(DIMS d, V4 *psi_inv, const M4 *J, const V4 *v, REAL alpha) {
    FOR_INTERIOR(d) {
        long i = IX(d, x, y, z);
        V4 u = N_(mul4)(alpha, N_(matvec)(&J[i], v[i]));
        psi_inv[i].x = psi_inv[i].x - u.x;
        psi_inv[i].y = psi_inv[i].y - u.y;
        psi_inv[i].z = psi_inv[i].z - u.z;
    }
}

/* reduce_max over all voxels: max of the round-down norm, first (lowest) index on ties —
 * reductor.cu:426-539 with reductor.cpp:107-120 */
static REAL N_(k_reduce_max)(DIMS d, const V4 *upd, long *argmax) {
    long n = (long) d.X * d.Y * d.Z;
    REAL best = 0;
    long bi = 0;
#pragma omp parallel
    {
        REAL lb = 0;
        long li = 0;
#pragma omp for schedule(static)
        for (long i = 0; i < n; ++i) {
            REAL nr = N_(norm4)(upd[i]);
            if (nr > lb) { lb = nr; li = i; }
        }
#pragma omp critical
        {
            if (lb > best || (lb == best && lb > 0 && li < bi)) { best = lb; bi = li; }
        }
    }
    if (argmax) *argmax = bi;
    return best;
}

/* energy: 0.5 * sum (T - S o psi)^2 and 0.5 * sum ||J rows||^2 over all voxels, accumulated in
 * double slice by slice, slices summed in z order — reductor.cu:11-112, 224-324,
 * reductor.cpp:45-62 */
static void N_(k_energy)(DIMS d, const V2 *tgt, const V2 *src_psi, const M4 *J, double *e_data, double *e_reg) {
    double *sd = (double *) calloc((size_t) d.Z, sizeof(double));
    double *sr = (double *) calloc((size_t) d.Z, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int z = 0; z < d.Z; ++z) {
        double a = 0, b = 0;
        for (int y = 0; y < d.Y; ++y)
            for (int x = 0; x < d.X; ++x) {
                long i = IX(d, x, y, z);
                double df = (double) tgt[i].x - (double) src_psi[i].x;
                a += df * df;
                const M4 *m = &J[i];
                for (int r = 0; r < 3; ++r) b += (double) m->r[r].x * m->r[r].x + (double) m->r[r].y * m->r[r].y + (double) m->r[r].z * m->r[r].z;
            }
        sd[z] = a;
        sr[z] = b;
    }
    double A = 0, B = 0;
    for (int z = 0; z < d.Z; ++z) { A += sd[z]; B += sr[z]; }
    free(sd);
    free(sr);
    *e_data = 0.5 * A;
    *e_reg = 0.5 * B;
}

/* estimate_inverse_kernel, one sweep over all voxels: psi_inv(p) = -trilerp(psi, psi_inv(p)), in
 * place (a voxel reads only its own psi_inv) — vector_fields.cu:108-122; DeformationField::
 * get_inverse runs 48 sweeps (vector_fields.cu:124-136, vector_fields.cpp:96-103) */
static void N_(k_estimate_inverse)(DIMS d, const V4 *psi, V4 *psi_inv) {
    FOR_ALL(d) {
        long i = IX(d, x, y, z);
        V4 r = N_(interp_field)(d, psi, psi_inv[i].x, psi_inv[i].y, psi_inv[i].z);
        psi_inv[i] = N_(mul4)((REAL) -1, r);
    }
}

/* interpolate_gradient_kernel: out(p) = trilerp(grad, psi(p)), all voxels — vector_fields.cu:194-225 */
static void N_(k_interpolate_gradient)(DIMS d, const V4 *grad, V4 *out, const V4 *psi) {
    FOR_ALL(d) {
        long i = IX(d, x, y, z);
        out[i] = N_(interp_field)(d, grad, psi[i].x, psi[i].y, psi[i].z);
    }
}

/* squared norm of a Jacobian row: (x*x + y*y) + z*z, products rounded separately —
 * cuda_utils.hpp:255-257 */
static inline REAL N_(norm_sq3)(V4 r) { return (r.x * r.x + r.y * r.y) + r.z * r.z; }

/* reduce_voxel_max_energy: the largest per-voxel energy (T - S o psi)^2 + w_reg * ((|J r0|^2 +
 * |J r1|^2) + |J r2|^2) over all voxels and its voxel; on equal energies the lowest index —
 * reductor.cu:637-760, reductor.cpp:82-89, 107-120 */
static REAL N_(k_voxel_max_energy)(DIMS d, const V2 *tgt, const V2 *src_psi, const M4 *J, REAL w_reg, long *arg) {
    long n = (long) d.X * d.Y * d.Z, best_i = 0;
    REAL best = 0;
    int found = 0;
    for (long i = 0; i < n; ++i) {
        REAL df = tgt[i].x - src_psi[i].x;
        REAL e = df * df + w_reg * ((N_(norm_sq3)(J[i].r[0]) + N_(norm_sq3)(J[i].r[1])) + N_(norm_sq3)(J[i].r[2]));
        if (e == e && (!found || e > best)) {
            best = e;
            best_i = i;
            found = 1;
        }
    }
    *arg = best_i;
    return best;
}

/* Dice counts over [0,X-2]x[0,Y-2]x[0,Z-2] (the last slice of each axis is skipped), labels
 * taken as array indices — solver.cu:9-50.  Labels outside [0, no_labels) are counted in
 * `bad` instead of indexing out of bounds. */
static void N_(k_dice_counts)(DIMS d, const V2 *a, const V2 *b, int no_labels, long *overlap, long *voxels, long *bad) {
    for (int l = 0; l < no_labels; ++l) overlap[l] = voxels[l] = 0;
    long nb = 0;
    for (int z = 0; z < d.Z - 1; ++z)
        for (int y = 0; y < d.Y - 1; ++y)
            for (int x = 0; x < d.X - 1; ++x) {
                long i = IX(d, x, y, z);
                int l1 = (int) a[i].y, l2 = (int) b[i].y;
                if (l1 < 0 || l1 >= no_labels || l2 < 0 || l2 >= no_labels) { nb++; continue; }
                voxels[l1]++;
                voxels[l2]++;
                if (a[i].y == b[i].y) overlap[l1]++;
            }
    *bad = nb;
}

/* min over the interior of the true det and of det_ref of the central-difference Jacobian of
 * psi (not in the reference; the acceptance figure "min det J") */
static void N_(k_min_det)(DIMS d, const V4 *psi, double *min_true, double *min_ref) {
    double mt = INFINITY, mr = INFINITY;
#pragma omp parallel for schedule(static) reduction(min : mt, mr)
    for (int z = 1; z <= d.Z - 2; ++z)
        for (int y = 1; y <= d.Y - 2; ++y)
            for (int x = 1; x <= d.X - 2; ++x) {
                V4 jx = N_(mul4)((REAL) 0.5, N_(sub4)(psi[IX(d, x + 1, y, z)], psi[IX(d, x - 1, y, z)]));
                V4 jy = N_(mul4)((REAL) 0.5, N_(sub4)(psi[IX(d, x, y + 1, z)], psi[IX(d, x, y - 1, z)]));
                V4 jz = N_(mul4)((REAL) 0.5, N_(sub4)(psi[IX(d, x, y, z + 1)], psi[IX(d, x, y, z - 1)]));
                M4 m;
                m.r[0] = N_(v4)(jx.x, jy.x, jz.x);
                m.r[1] = N_(v4)(jx.y, jy.y, jz.y);
                m.r[2] = N_(v4)(jx.z, jy.z, jz.z);
                m.r[3] = N_(v4)(0, 0, 0);
                double t = (double) N_(det_true)(&m), r = (double) N_(det_ref)(&m);
                if (t < mt) mt = t;
                if (r < mr) mr = r;
            }
    *min_true = mt;
    *min_ref = mr;
}

#include "fr_oracle_solver.h"

#undef FOR_ALL
#undef FOR_INTERIOR
#undef V2
#undef V4
#undef M4
#undef DIMS
#undef IX
#undef CORNER
#undef FMA
