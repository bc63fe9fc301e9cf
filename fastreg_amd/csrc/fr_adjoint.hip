// The differentiated solver: K iterations of the registration solver in FP64 and their adjoint,
// for gradients of a final objective with respect to the initial velocity and the regularisation
// weights (w_reg, tau).
//
// Forward: the reference's iteration (solver.cu:295-444) in FP64 and in the reference's operation
// order, one thread per voxel, as three kernels: F3 (f_laplacian), F1 F2 F4-F7 F9 (f_step, which
// stores the intermediates the adjoint reads only when the backward pass recomputes a segment) and
// F8 (f_psi):
//   F1 rho1 = rho + alpha * div(rho, v)                           (solver.cu:101-173)
//   F2 J = D psi_inv                                              (vector_fields.cu:503-535)
//   F3 L = -Laplacian(psi)                                        (vector_fields.cu:275-300)
//   F4 Lo = trilerp(L, psi_inv), Tp = trilerp(T, psi_inv)         (vector_fields.cu:231-247, 81-99)
//   F5 gU = ((S - Tp) grad S + w_reg Lo) det_ref(J)               (solver.cu:56-80)
//   F6 dvv = (Dv)v, dlv = upwind Laplacian of v                   (vector_fields.cu:307-404)
//   F7 v1 = v - alpha (m1t v + dvv - tau dlv + q gU), q = m2/(|rho1| + 1e-5)   (solver.cu:188-211)
//   F8 psi1 = psi + alpha trilerp(v1, psi)                        (solver.cu:230-247)
//   F9 psi_inv1 = psi_inv - alpha J v1                            (solver.cu:263-279)
// Objective of the final state: E = 0.5 sum (T - S o psi)^2 + lw * 0.5 sum_l sum (1[LT = l] -
// trilerp(1[LS = l], psi))^2 over labels l = 1..NO_LABELS-1 (a label-overlap surrogate).
//
// Backward: the transposes of F9..F1 in reverse order, fused into 8 kernels per iteration, each
// written as a gather (every thread writes only its own voxel); only sources displaced by a voxel
// or more fall back to FP64 atomic scatters.  The derivative of trilerp with respect to its
// position is that of the piecewise-trilinear function (zero where the position is clamped);
// the upwind / min / max branches are differentiated within the branch the forward took.
// Checkpoints of the state every `ckpt` iterations; the backward pass recomputes each segment's
// states and intermediates from its checkpoint, then runs the segment's adjoint.

#include "../../include/fastreg_c.h"
#include "fr_internal.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <vector>

namespace frd {

struct D3 {
    double x, y, z;
};

__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 ld(const double *p, int i) { return d3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st(double *p, int i, D3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
// FP64 atomic add in hardware (global_atomic_add_f64; every buffer here is ordinary device memory)
__device__ __forceinline__ void aadd(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void add_at(double *p, int i, D3 v) {
    aadd(&p[3 * i], v.x);
    aadd(&p[3 * i + 1], v.y);
    aadd(&p[3 * i + 2], v.z);
}
__device__ __forceinline__ double comp(D3 v, int c) { return c == 0 ? v.x : c == 1 ? v.y : v.z; }

struct Dims {
    int X, Y, Z;
};
// voxel indices are 32-bit: the library accepts volumes with 9 N < 2^31 (the Jacobian buffers)
__device__ __forceinline__ int lin(Dims d, int x, int y, int z) { return x + d.X * (y + d.Y * z); }
// the adjoint of the Jacobian of psi_inv, entry k = 3 r + c of voxel i: one plane of N entries per
// k (the neighbours' entries that step 8 gathers are then contiguous across a wave)
__device__ __forceinline__ int gj_at(Dims d, int i, int k) { return k * (d.X * d.Y * d.Z) + i; }
__device__ __forceinline__ bool interior(Dims d, int x, int y, int z) {
    return x >= 1 && x <= d.X - 2 && y >= 1 && y <= d.Y - 2 && z >= 1 && z <= d.Z - 2;
}
#define FOR_VOXELS(d, i)                                                                                    \
    for (int i = blockIdx.x * blockDim.x + threadIdx.x, n_ = (d).X * (d).Y * (d).Z; i < n_; \
         i += gridDim.x * blockDim.x)
#define XYZ(d, i)                                                                 \
    const int x = (int)((unsigned)(i) % (unsigned)(d).X);                         \
    const int y = (int)(((unsigned)(i) / (unsigned)(d).X) % (unsigned)(d).Y);     \
    const int z = (int)((unsigned)(i) / ((unsigned)(d).X * (unsigned)(d).Y))

// lerp(v0, v1, t) = fma(t, v0, fma(-t, v1, v1)) (cuda_utils.hpp:43-46), in FP64
__device__ __forceinline__ double lerp(double v0, double v1, double t) { return fma(t, v0, fma(-t, v1, v1)); }

// clamped corner and weights (cuda_utils.hpp:60-136); `in` marks the axes along which the
// position is inside the grid (the derivative with respect to a clamped coordinate is 0)
struct Corner {
    int g[8];  // linear indices of the 8 corners, order 111, 110, 101, 100, 011, 010, 001, 000
    double a, b, c;
    int gx, gy, gz;  // the floor corner
    bool inx, iny, inz;
};
__device__ __forceinline__ Corner corner(Dims d, double px, double py, double pz) {
    Corner k;
    const double cx = fmin(fmax(0.0, px), (double)(d.X - 1)), cy = fmin(fmax(0.0, py), (double)(d.Y - 1)),
                 cz = fmin(fmax(0.0, pz), (double)(d.Z - 1));
    const int gx = (int)floor(cx), gy = (int)floor(cy), gz = (int)floor(cz);
    k.gx = gx;
    k.gy = gy;
    k.gz = gz;
    k.a = cx - gx;
    k.b = cy - gy;
    k.c = cz - gz;
    const int hx = gx + 1 < d.X ? gx + 1 : gx, hy = gy + 1 < d.Y ? gy + 1 : gy, hz = gz + 1 < d.Z ? gz + 1 : gz;
    k.g[0] = lin(d, hx, hy, hz);
    k.g[1] = lin(d, hx, hy, gz);
    k.g[2] = lin(d, hx, gy, hz);
    k.g[3] = lin(d, hx, gy, gz);
    k.g[4] = lin(d, gx, hy, hz);
    k.g[5] = lin(d, gx, hy, gz);
    k.g[6] = lin(d, gx, gy, hz);
    k.g[7] = lin(d, gx, gy, gz);
    k.inx = px > 0.0 && px < (double)(d.X - 1);
    k.iny = py > 0.0 && py < (double)(d.Y - 1);
    k.inz = pz > 0.0 && pz < (double)(d.Z - 1);
    return k;
}
// trilerp of 8 corner values in the reference's nesting
__device__ __forceinline__ double tri(const double f[8], double a, double b, double c) {
    return lerp(lerp(lerp(f[0], f[1], c), lerp(f[2], f[3], c), b), lerp(lerp(f[4], f[5], c), lerp(f[6], f[7], c), b),
                a);
}
// weights of the 8 corners (the derivative of trilerp with respect to the corner values)
__device__ __forceinline__ void tri_weights(double a, double b, double c, double w[8]) {
    w[0] = a * b * c;
    w[1] = a * b * (1 - c);
    w[2] = a * (1 - b) * c;
    w[3] = a * (1 - b) * (1 - c);
    w[4] = (1 - a) * b * c;
    w[5] = (1 - a) * b * (1 - c);
    w[6] = (1 - a) * (1 - b) * c;
    w[7] = (1 - a) * (1 - b) * (1 - c);
}
// derivative of trilerp with respect to the position (zero along clamped axes)
__device__ __forceinline__ D3 tri_dpos(const double f[8], const Corner &k) {
    const double a = k.a, b = k.b, c = k.c;
    const double c11 = c * f[0] + (1 - c) * f[1], c10 = c * f[2] + (1 - c) * f[3], c01 = c * f[4] + (1 - c) * f[5],
                 c00 = c * f[6] + (1 - c) * f[7];
    const double da = (b * c11 + (1 - b) * c10) - (b * c01 + (1 - b) * c00);
    const double db = a * (c11 - c10) + (1 - a) * (c01 - c00);
    const double dc = a * (b * (f[0] - f[1]) + (1 - b) * (f[2] - f[3])) + (1 - a) * (b * (f[4] - f[5]) + (1 - b) * (f[6] - f[7]));
    return d3(k.inx ? da : 0.0, k.iny ? db : 0.0, k.inz ? dc : 0.0);
}
__device__ __forceinline__ void gather1(const double *f, const Corner &k, double out[8]) {
    for (int j = 0; j < 8; ++j) out[j] = f[k.g[j]];
}
__device__ __forceinline__ void gather3(const double *f, const Corner &k, int c, double out[8]) {
    for (int j = 0; j < 8; ++j) out[j] = f[3 * k.g[j] + c];
}

// det() of the reference (not the determinant; cuda_utils.hpp:389-396), m[r][c], and its gradient
__device__ __forceinline__ double det_ref(const double m[3][3]) {
    const double a = ((m[0][0] * m[1][1]) * m[2][2] + (m[1][0] * m[2][1]) * m[0][2]) + (m[2][0] * m[0][1]) * m[1][2];
    const double b = ((m[0][2] * m[1][1]) + m[2][0]) * ((m[1][2] * m[2][1]) * m[0][0]) + (m[2][2] * m[0][1]) * m[1][0];
    return a - b;
}
__device__ __forceinline__ void det_ref_grad(const double m[3][3], double g[3][3]) {
    // a = m00 m11 m22 + m10 m21 m02 + m20 m01 m12; b = (m02 m11 + m20)(m12 m21 m00) + m22 m01 m10
    const double P = m[0][2] * m[1][1] + m[2][0], Q = m[1][2] * m[2][1] * m[0][0];
    g[0][0] = m[1][1] * m[2][2] - P * (m[1][2] * m[2][1]);
    g[0][1] = m[2][0] * m[1][2] - m[2][2] * m[1][0];
    g[0][2] = m[1][0] * m[2][1] - m[1][1] * Q;
    g[1][0] = m[2][1] * m[0][2] - m[2][2] * m[0][1];
    g[1][1] = m[0][0] * m[2][2] - m[0][2] * Q;
    g[1][2] = m[2][0] * m[0][1] - P * (m[2][1] * m[0][0]);
    g[2][0] = m[0][1] * m[1][2] - Q;
    g[2][1] = m[1][0] * m[0][2] - P * (m[1][2] * m[0][0]);
    g[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
}

// tiles of the z-marching one-thread-per-column kernels (64 x 4 columns per block)
constexpr int MT_X = 64, MT_Y = 4, MT_BLOCKS_PER_CU = 8;

// ---- forward steps ----------------------------------------------------------------------------

// The forward steps of one voxel.  F3 (L) needs psi at the neighbours and F8 (psi1) v1 at the
// taps, so they are kernels of their own; the other steps of a voxel read only the old state, L
// and the images, and run in one kernel (f_step) that keeps their results in registers.

// The z-window of one column for the forward step: v at z-2..z+2, psi_inv and rho at z-1..z+1
// (zeros outside the grid; no branch reads them there)
struct FWin {
    D3 v[5], p[3];
    double r[3];
};

// F1: rho1 = rho + alpha * div on the interior; rho elsewhere (the z-neighbours from the window)
__device__ __forceinline__ double fv_density(Dims d, int x, int y, int z, const double *rho, const double *v,
                                             const FWin &w, double alpha) {
    const double p = w.r[1];
    if (!interior(d, x, y, z)) return p;
    const D3 vel = w.v[2];
    const double t11 = vel.x > 0 ? -vel.x * p : -vel.x * rho[lin(d, x + 1, y, z)];
    const double vxm = v[3 * lin(d, x - 1, y, z)];
    const double t12 = vxm > 0 ? vxm * rho[lin(d, x - 1, y, z)] : vxm * p;
    const double t21 = vel.y > 0 ? -vel.y * p : -vel.y * rho[lin(d, x, y + 1, z)];
    const double vym = v[3 * lin(d, x, y - 1, z) + 1];
    const double t22 = vym > 0 ? vym * rho[lin(d, x, y - 1, z)] : vym * p;
    const double t31 = vel.z > 0 ? -vel.z * p : -vel.z * w.r[2];
    const double vzm = w.v[1].z;
    const double t32 = vzm > 0 ? vzm * w.r[0] : vzm * p;
    const double div = ((((t11 + t12) + t21) + t22) + t31) + t32;
    return fma(alpha, div, p);
}

// F2: J = D psi_inv (m[3 r + c]: row r = component, column c = d/dx, d/dy, d/dz), interior; 0
// elsewhere
__device__ __forceinline__ void jac_of(D3 xp, D3 xm, D3 yp, D3 ym, D3 zp, D3 zm, double m[9]) {
    const D3 jx = d3((xp.x - xm.x) * 0.5, (xp.y - xm.y) * 0.5, (xp.z - xm.z) * 0.5);
    const D3 jy = d3((yp.x - ym.x) * 0.5, (yp.y - ym.y) * 0.5, (yp.z - ym.z) * 0.5);
    const D3 jz = d3((zp.x - zm.x) * 0.5, (zp.y - zm.y) * 0.5, (zp.z - zm.z) * 0.5);
    m[0] = jx.x; m[1] = jy.x; m[2] = jz.x;
    m[3] = jx.y; m[4] = jy.y; m[5] = jz.y;
    m[6] = jx.z; m[7] = jy.z; m[8] = jz.z;
}
__device__ __forceinline__ void fv_jacobian(Dims d, int x, int y, int z, const double *pinv, double m[9]) {
    for (int k = 0; k < 9; ++k) m[k] = 0.0;
    if (!interior(d, x, y, z)) return;
    jac_of(ld(pinv, lin(d, x + 1, y, z)), ld(pinv, lin(d, x - 1, y, z)), ld(pinv, lin(d, x, y + 1, z)),
           ld(pinv, lin(d, x, y - 1, z)), ld(pinv, lin(d, x, y, z + 1)), ld(pinv, lin(d, x, y, z - 1)), m);
}
// the same with psi_inv at z+-1 from the window
__device__ __forceinline__ void fv_jacobian_w(Dims d, int x, int y, int z, const double *pinv, const FWin &w,
                                              double m[9]) {
    for (int k = 0; k < 9; ++k) m[k] = 0.0;
    if (!interior(d, x, y, z)) return;
    jac_of(ld(pinv, lin(d, x + 1, y, z)), ld(pinv, lin(d, x - 1, y, z)), ld(pinv, lin(d, x, y + 1, z)),
           ld(pinv, lin(d, x, y - 1, z)), w.p[2], w.p[0], m);
}

// F3: L = -((((((-6 psi + psi(x+1)) + psi(x-1)) + psi(y+1)) + psi(y-1)) + psi(z+1)) + psi(z-1)), interior
__global__ void f_laplacian(Dims d, const double *psi, double *L) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        D3 o = d3(0, 0, 0);
        if (interior(d, x, y, z)) {
            const int nb[6] = {lin(d, x + 1, y, z), lin(d, x - 1, y, z), lin(d, x, y + 1, z),
                                lin(d, x, y - 1, z), lin(d, x, y, z + 1), lin(d, x, y, z - 1)};
            const D3 c = ld(psi, i);
            D3 s = d3(c.x * -6.0, c.y * -6.0, c.z * -6.0);
            for (int k = 0; k < 6; ++k) {
                const D3 q = ld(psi, nb[k]);
                s = d3(s.x + q.x, s.y + q.y, s.z + q.z);
            }
            o = d3(s.x * -1.0, s.y * -1.0, s.z * -1.0);
        }
        st(L, i, o);
    }
}

// F4: Lo = trilerp(L, psi_inv), Tp = trilerp(T, psi_inv), every voxel
__device__ __forceinline__ void fv_interp(Dims d, int i, const double *L, const double *T, const double *pinv, D3 *Lo,
                                          double *Tp) {
    const D3 p = ld(pinv, i);
    const Corner k = corner(d, p.x, p.y, p.z);
    double f[8];
    gather3(L, k, 0, f);
    Lo->x = tri(f, k.a, k.b, k.c);
    gather3(L, k, 1, f);
    Lo->y = tri(f, k.a, k.b, k.c);
    gather3(L, k, 2, f);
    Lo->z = tri(f, k.a, k.b, k.c);
    gather1(T, k, f);
    *Tp = tri(f, k.a, k.b, k.c);
}

// F5: gU = ((S - Tp) grad S + w_reg Lo) det_ref(J), interior; 0 elsewhere
__device__ __forceinline__ D3 fv_potential(Dims d, int x, int y, int z, int i, const double *S, const double *gS,
                                           double Tp, D3 lo, const double J[9], double w_reg) {
    if (!interior(d, x, y, z)) return d3(0, 0, 0);
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = J[3 * r + c];
    const double det = det_ref(m), diff = S[i] - Tp;
    const D3 g = ld(gS, i);
    return d3((g.x * diff + lo.x * w_reg) * det, (g.y * diff + lo.y * w_reg) * det, (g.z * diff + lo.z * w_reg) * det);
}

__device__ __forceinline__ double mx(double q) { return fmax(q, 0.0); }
__device__ __forceinline__ double mn(double q) { return fmin(q, 0.0); }

// F6: dvv = (Dv)v on the interior (0 elsewhere), dlv = the reference's upwind Laplacian of v on
// its range (0 elsewhere)
// (the z-neighbours from the window)
__device__ __forceinline__ void fv_vel_terms(Dims d, int x, int y, int z, int i, const double *v, const FWin &w,
                                             D3 *dvv, D3 *dlv) {
    D3 o = d3(0, 0, 0), l = d3(0, 0, 0);
    if (interior(d, x, y, z)) {
        const D3 c = w.v[2], xp = ld(v, lin(d, x + 1, y, z)), xm = ld(v, lin(d, x - 1, y, z)),
                 yp = ld(v, lin(d, x, y + 1, z)), ym = ld(v, lin(d, x, y - 1, z)), zp = w.v[3], zm = w.v[1];
        const double ax1 = ((mx(c.x) * mx(c.x) - mn(c.x) * mn(c.x)) + mn(xp.x) * mn(xp.x)) - mx(xp.x) * mx(xp.x);
        const double ax2 = c.y > 0 ? c.y * (c.x - ym.x) : c.y * (yp.x - c.x);
        const double ax3 = c.z > 0 ? c.z * (c.x - zm.x) : c.z * (zp.x - c.x);
        const double ay1 = c.x > 0 ? c.x * (c.y - xm.y) : c.x * (xp.y - c.y);
        const double ay2 = ((mx(c.y) * mx(c.y) - mn(c.y) * mn(c.y)) + mn(yp.y) * mn(yp.y)) - mx(yp.y) * mx(yp.y);
        const double ay3 = c.z > 0 ? c.z * (c.y - zm.y) : c.z * (zp.y - c.y);
        const double az1 = c.x > 0 ? c.x * (c.z - xm.z) : c.x * (xp.z - c.z);
        const double az2 = c.y > 0 ? c.y * (c.z - ym.z) : c.y * (yp.z - c.z);
        const double az3 = ((mx(c.z) * mx(c.z) - mn(c.z) * mn(c.z)) + mn(zp.z) * mn(zp.z)) - mx(zp.z) * mx(zp.z);
        o = d3((0.5 * ax1 + ax2) + ax3, (ay1 + 0.5 * ay2) + ay3, (az1 + az2) + 0.5 * az3);
    }
    if (x >= 2 && x <= d.X - 3 && y >= 2 && y <= d.Y - 2 && z >= 2 && z <= d.Z - 3) {
        const D3 c = w.v[2];
        const int sx = 1, sy = d.X;
        // one-sided second differences with the reference's y-1 / z-1 where +1 is meant
        const int ax = c.x < 0 ? i + 2 * sx : i - 2 * sx, bx = c.x < 0 ? i + sx : i - sx;
        const int ay = c.y < 0 ? i + 2 * sy : i - 2 * sy, by = i - sy;
        const D3 az = c.z < 0 ? w.v[4] : w.v[0], bz = w.v[1];
        D3 s = d3(0, 0, 0);
        for (int q = 0; q < 3; ++q) {
            const double cv = comp(c, q);
            const double vxx = c.x < 0 ? (v[3 * ax + q] - 2.0 * v[3 * bx + q]) + cv : (cv - 2.0 * v[3 * bx + q]) + v[3 * ax + q];
            const double vyy = c.y < 0 ? (v[3 * ay + q] - 2.0 * v[3 * by + q]) + cv : (cv - 2.0 * v[3 * by + q]) + v[3 * ay + q];
            const double vzz = c.z < 0 ? (comp(az, q) - 2.0 * comp(bz, q)) + cv : (cv - 2.0 * comp(bz, q)) + comp(az, q);
            const double r = -1.0 * ((vxx + vyy) + vzz);
            if (q == 0) s.x = r; else if (q == 1) s.y = r; else s.z = r;
        }
        l = s;
    }
    *dvv = o;
    *dlv = l;
}

// F7: v1 = v - alpha (((m1t v + dvv) - tau dlv) + q gU), q = m2 / (|rho1| + 1e-5f), interior; v
// elsewhere
__device__ __forceinline__ D3 fv_velocity(Dims d, int x, int y, int z, D3 vv, D3 a, D3 l, double rho1, D3 g,
                                          double alpha, double m1t, double tau, double m2) {
    if (!interior(d, x, y, z)) return vv;
    const double q = m2 / (fabs(rho1) + (double)1e-5f);
    const D3 fr = d3(vv.x * m1t, vv.y * m1t, vv.z * m1t), df = d3(l.x * tau, l.y * tau, l.z * tau),
             po = d3(g.x * q, g.y * q, g.z * q);
    const D3 up = d3((((fr.x + a.x) - df.x) + po.x) * alpha, (((fr.y + a.y) - df.y) + po.y) * alpha,
                     (((fr.z + a.z) - df.z) + po.z) * alpha);
    return d3(vv.x - up.x, vv.y - up.y, vv.z - up.z);
}

// F9: psi_inv1 = psi_inv - alpha J v1, interior; psi_inv elsewhere
__device__ __forceinline__ D3 fv_psi_inv(Dims d, int x, int y, int z, D3 o, const double m[9], D3 w, double alpha) {
    if (!interior(d, x, y, z)) return o;
    const D3 u = d3(((m[0] * w.x + m[1] * w.y) + m[2] * w.z) * alpha, ((m[3] * w.x + m[4] * w.y) + m[5] * w.z) * alpha,
                    ((m[6] * w.x + m[7] * w.y) + m[8] * w.z) * alpha);
    return d3(o.x - u.x, o.y - u.y, o.z - u.z);
}

__device__ __forceinline__ double ld1_or0(const double *__restrict__ p, int i, bool ok) { return ok ? p[i] : 0.0; }
__device__ __forceinline__ D3 ld_or0(const double *__restrict__ p, int i, bool ok) {
    return ok ? ld(p, i) : d3(0, 0, 0);
}

// F1, F2, F4, F5, F6, F7, F9 of every voxel: the new rho, v and psi_inv from the old state, L and
// the images; with `record`, also the intermediates the adjoint reads (gU, dlv; the adjoint
// recomputes J from psi_inv and Lo, Tp from L and T at psi_inv, which it gathers anyway).  One
// thread per (x, y) column of a 64 x 4 tile marches a chunk of zc planes, holding the column's
// z-neighbours of v, psi_inv and rho in registers.
__global__ __launch_bounds__(256) void f_step(Dims d, int tiles_x, int tiles_y, int zc, const double *rho,
                                              const double *v, const double *pinv, const double *L,
                                              const double *T, const double *S, const double *gS, double *rho1,
                                              double *v1, double *pinv1, double *gU_out, double *dlv_out, int record,
                                              double alpha, double w_reg, double m1t, double tau, double m2) {
    const int b = blockIdx.x;
    const int x = (b % tiles_x) * MT_X + (int)(threadIdx.x % MT_X);
    const int y = ((b / tiles_x) % tiles_y) * MT_Y + (int)(threadIdx.x / MT_X);
    const int z0 = (b / (tiles_x * tiles_y)) * zc, z1 = min(z0 + zc, d.Z);
    if (x >= d.X || y >= d.Y) return;
    const int sz = d.X * d.Y, c0 = x + d.X * y;
    FWin w;
    for (int k = 0; k < 5; ++k) {  // v at z0-2 .. z0+2
        const int p = z0 - 2 + k;
        w.v[k] = ld_or0(v, c0 + sz * p, p >= 0 && p < d.Z);
    }
    for (int k = 0; k < 3; ++k) {  // psi_inv and rho at z0-1 .. z0+1
        const int p = z0 - 1 + k;
        const bool ok = p >= 0 && p < d.Z;
        w.p[k] = ld_or0(pinv, c0 + sz * p, ok);
        w.r[k] = ld1_or0(rho, c0 + sz * p, ok);
    }
    for (int z = z0; z < z1; ++z) {
        // the entries of the next step, loaded a step ahead
        const D3 nv = ld_or0(v, c0 + sz * (z + 3), z + 3 < d.Z), np = ld_or0(pinv, c0 + sz * (z + 2), z + 2 < d.Z);
        const double nr = ld1_or0(rho, c0 + sz * (z + 2), z + 2 < d.Z);
        const int i = c0 + sz * z;
        const double r1 = fv_density(d, x, y, z, rho, v, w, alpha);
        double J[9];
        fv_jacobian_w(d, x, y, z, pinv, w, J);
        D3 lo;
        double tp;
        fv_interp(d, i, L, T, pinv, &lo, &tp);
        const D3 gu = fv_potential(d, x, y, z, i, S, gS, tp, lo, J, w_reg);
        D3 dvv, dlv;
        fv_vel_terms(d, x, y, z, i, v, w, &dvv, &dlv);
        const D3 wv = fv_velocity(d, x, y, z, w.v[2], dvv, dlv, r1, gu, alpha, m1t, tau, m2);
        const D3 pi = fv_psi_inv(d, x, y, z, w.p[1], J, wv, alpha);
        rho1[i] = r1;
        st(v1, i, wv);
        st(pinv1, i, pi);
        if (record) {
            st(gU_out, i, gu);
            st(dlv_out, i, dlv);
        }
        for (int k = 0; k < 4; ++k) w.v[k] = w.v[k + 1];
        w.v[4] = nv;
        w.p[0] = w.p[1];
        w.p[1] = w.p[2];
        w.p[2] = np;
        w.r[0] = w.r[1];
        w.r[1] = w.r[2];
        w.r[2] = nr;
    }
}

// F8: psi1 = psi + alpha trilerp(v1, psi), interior
__global__ void f_psi(Dims d, const double *psi, const double *v1, double *psi1, double alpha) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        D3 o = ld(psi, i);
        if (interior(d, x, y, z)) {
            const Corner k = corner(d, o.x, o.y, o.z);
            double f[8];
            gather3(v1, k, 0, f);
            const double ux = tri(f, k.a, k.b, k.c) * alpha;
            gather3(v1, k, 1, f);
            const double uy = tri(f, k.a, k.b, k.c) * alpha;
            gather3(v1, k, 2, f);
            const double uz = tri(f, k.a, k.b, k.c) * alpha;
            o = d3(o.x + ux, o.y + uy, o.z + uz);
        }
        st(psi1, i, o);
    }
}

// objective terms per voxel, summed in FP64 per block
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    __syncthreads();
    return s;
}

// E = 0.5 sum (T - S o psi)^2 + lw * 0.5 sum_l (1[LT = l] - trilerp(1[LS = l], psi))^2; block partial sums
__global__ __launch_bounds__(256) void f_objective(Dims d, const double *T, const double *S, const int16_t *LT,
                                                   const int16_t *LS, int no_labels, double lw, const double *psi,
                                                   double *partials) {
    __shared__ double sh[4];
    double acc = 0;
    FOR_VOXELS(d, i) {
        const D3 p = ld(psi, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double f[8];
        gather1(S, k, f);
        const double r = T[i] - tri(f, k.a, k.b, k.c);
        acc += 0.5 * r * r;
        if (lw != 0.0) {
            int lab[8];
            for (int j = 0; j < 8; ++j) lab[j] = LS[k.g[j]];
            for (int l = 1; l < no_labels; ++l) {
                for (int j = 0; j < 8; ++j) f[j] = lab[j] == l ? 1.0 : 0.0;
                const double rl = (LT[i] == l ? 1.0 : 0.0) - tri(f, k.a, k.b, k.c);
                acc += lw * 0.5 * rl * rl;
            }
        }
    }
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ---- backward steps ---------------------------------------------------------------------------

// dE/dpsi of the objective
__global__ void b_objective(Dims d, const double *T, const double *S, const int16_t *LT, const int16_t *LS,
                            int no_labels, double lw, const double *psi, double *gpsi) {
    FOR_VOXELS(d, i) {
        const D3 p = ld(psi, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double f[8];
        gather1(S, k, f);
        const double r = T[i] - tri(f, k.a, k.b, k.c);
        const D3 dp = tri_dpos(f, k);
        D3 g = d3(-r * dp.x, -r * dp.y, -r * dp.z);
        if (lw != 0.0) {
            int lab[8];
            for (int j = 0; j < 8; ++j) lab[j] = LS[k.g[j]];
            for (int l = 1; l < no_labels; ++l) {
                for (int j = 0; j < 8; ++j) f[j] = lab[j] == l ? 1.0 : 0.0;
                const double rl = (LT[i] == l ? 1.0 : 0.0) - tri(f, k.a, k.b, k.c);
                const D3 dl = tri_dpos(f, k);
                g = d3(g.x - lw * rl * dl.x, g.y - lw * rl * dl.y, g.z - lw * rl * dl.z);
            }
        }
        st(gpsi, i, g);
    }
}

// Every transpose below is a gather: voxel i sums, over the voxels j whose forward stencil reads
// i, the partial of j's output with respect to i (j's branch re-evaluated from the forward state),
// so each thread writes only its own voxel and the result does not depend on thread order.  The
// transpose of a trilinear sample at a moved position (a scatter to the 8 corners of the
// position's cell) takes two steps: each source voxel writes a record (its cell relative to
// itself, the cell fractions, the adjoint to spread), then each target voxel gathers from the 27
// sources around it whose cell it is a corner of.  A source whose cell is not within one voxel of
// it (a displacement of a voxel or more) is marked irregular and scattered with FP64 atomics by a
// separate pass that runs before the gather.

constexpr int RC_VALID = 1, RC_IRREG = 2;  // record code bits; bits 2, 3, 4: cell offset + 1 along x, y, z

__device__ __forceinline__ int rec_code(const Corner &k, int x, int y, int z) {
    const int ox = k.gx - x, oy = k.gy - y, oz = k.gz - z;
    if (ox < -1 || ox > 0 || oy < -1 || oy > 0 || oz < -1 || oz > 0) return RC_VALID | RC_IRREG;
    return RC_VALID | ((ox + 1) << 2) | ((oy + 1) << 3) | ((oz + 1) << 4);
}
__device__ __forceinline__ void put_rec(int *code, double *abc, double *gw, int i, int c, const Corner &k, D3 g) {
    code[i] = c;
    st(abc, i, d3(k.a, k.b, k.c));
    st(gw, i, g);
}

// The gathers of the records run over tiles of 64 (x) x 4 (y) targets, each block marching a chunk
// of z-planes.  The records of the tile's neighbourhood (x0-1 .. x0+64, y0-1 .. y0+4) for the three
// planes z-1, z, z+1 sit in LDS (a ring of 3 planes; records outside the grid, and the cell data of
// sources that are not regular, are stored as code 0 / zeros), so a target reads its 27 sources
// from LDS and every record is read from memory about 1.6 times instead of 27.
constexpr int GT_X = 64, GT_Y = 4, GT_THREADS = GT_X * GT_Y;
constexpr int GS_X = GT_X + 2, GS_Y = GT_Y + 2, GS_N = GS_X * GS_Y;
constexpr int GS_PER_THREAD = (GS_N + GT_THREADS - 1) / GT_THREADS;
struct RecStage {
    int code[3][GS_N];
    double v[6][3][GS_N];  // cell fractions a, b, c; adjoint g.x, g.y, g.z
};
struct RecRegs {
    int code[GS_PER_THREAD];
    double v[6][GS_PER_THREAD];
};

// this thread's share of plane zz of the tile's neighbourhood, from memory into registers
__device__ __forceinline__ void rec_fetch(Dims d, int x0, int y0, int zz, const int *code, const double *abc,
                                          const double *gw, RecRegs &r) {
    for (int q = 0; q < GS_PER_THREAD; ++q) {
        const int e = threadIdx.x + q * GT_THREADS;
        const int gx = x0 - 1 + e % GS_X, gy = y0 - 1 + e / GS_X;
        int c = 0;
        D3 f = d3(0, 0, 0), g = d3(0, 0, 0);
        if (e < GS_N && gx >= 0 && gx < d.X && gy >= 0 && gy < d.Y && zz >= 0 && zz < d.Z) {
            const int i = lin(d, gx, gy, zz);
            c = code[i];
            if ((c & (RC_VALID | RC_IRREG)) == RC_VALID) {
                f = ld(abc, i);
                g = ld(gw, i);
            } else {
                c = 0;
            }
        }
        r.code[q] = c;
        r.v[0][q] = f.x;
        r.v[1][q] = f.y;
        r.v[2][q] = f.z;
        r.v[3][q] = g.x;
        r.v[4][q] = g.y;
        r.v[5][q] = g.z;
    }
}
__device__ __forceinline__ void rec_put(RecStage &st, int slot, const RecRegs &r) {
    for (int q = 0; q < GS_PER_THREAD; ++q) {
        const int e = threadIdx.x + q * GT_THREADS;
        if (e < GS_N) {
            st.code[slot][e] = r.code[q];
            for (int k = 0; k < 6; ++k) st.v[k][slot][e] = r.v[k][q];
        }
    }
}

// sum over the regular sources within one voxel of target (x0 + lx, y0 + ly, z) of weight x
// adjoint, in the order sz, sy, sx = -1, 0, 1 (sources outside the grid have code 0)
__device__ __forceinline__ D3 rec_gather(const RecStage &st, int lx, int ly, int z) {
    D3 acc = d3(0, 0, 0);
    for (int sz = -1; sz <= 1; ++sz) {
        const int slot = (z + sz + 3) % 3;
        for (int sy = -1; sy <= 1; ++sy) {
            for (int sx = -1; sx <= 1; ++sx) {
                const int e = (ly + 1 + sy) * GS_X + (lx + 1 + sx);
                const int c = st.code[slot][e];
                if (c == 0) continue;
                // the target is the source's (x, y, z) + t; corner sides: t == o (floor), t == o + 1
                const int ox = ((c >> 2) & 1) - 1, oy = ((c >> 3) & 1) - 1, oz = ((c >> 4) & 1) - 1;
                const int tx = -sx, ty = -sy, tz = -sz;
                if ((tx != ox && tx != ox + 1) || (ty != oy && ty != oy + 1) || (tz != oz && tz != oz + 1)) continue;
                const D3 f = d3(st.v[0][slot][e], st.v[1][slot][e], st.v[2][slot][e]),
                         g = d3(st.v[3][slot][e], st.v[4][slot][e], st.v[5][slot][e]);
                const double wa = tx == ox ? 1 - f.x : f.x, wb = ty == oy ? 1 - f.y : f.y, wc = tz == oz ? 1 - f.z : f.z;
                const double w = wa * wb * wc;
                acc = d3(acc.x + w * g.x, acc.y + w * g.y, acc.z + w * g.z);
            }
        }
    }
    return acc;
}

// Block b of a gather launch: tile (b % tiles_x, (b / tiles_x) % tiles_y), z-planes
// [zc * (b / (tiles_x tiles_y)), + zc).  body(i, x, y, z, sum) runs for every target of the tile in
// the grid; every thread of the block runs the same number of loop steps (block-wide barriers).
template <typename Body>
__device__ __forceinline__ void rec_gather_tile(Dims d, int tiles_x, int tiles_y, int zc, const int *code,
                                                const double *abc, const double *gw, Body body) {
    __shared__ RecStage st;
    const int b = blockIdx.x;
    const int x0 = (b % tiles_x) * GT_X, y0 = ((b / tiles_x) % tiles_y) * GT_Y;
    const int z0 = (b / (tiles_x * tiles_y)) * zc, z1 = min(z0 + zc, d.Z);
    const int lx = threadIdx.x % GT_X, ly = threadIdx.x / GT_X;
    const int x = x0 + lx, y = y0 + ly;
    RecRegs r;
    rec_fetch(d, x0, y0, z0 - 1, code, abc, gw, r);
    rec_put(st, (z0 + 2) % 3, r);
    rec_fetch(d, x0, y0, z0, code, abc, gw, r);
    rec_put(st, z0 % 3, r);
    rec_fetch(d, x0, y0, z0 + 1, code, abc, gw, r);
    for (int z = z0; z < z1; ++z) {
        rec_put(st, (z + 1) % 3, r);  // plane z + 1 (its slot held plane z - 2, no longer read)
        __syncthreads();
        if (z + 2 <= z1) rec_fetch(d, x0, y0, z + 2, code, abc, gw, r);  // in flight during the gather
        if (x < d.X && y < d.Y) {
            const D3 sc = rec_gather(st, lx, ly, z);
            body(lin(d, x, y, z), x, y, z, sc);
        }
        __syncthreads();
    }
}

// the irregular sources' scatter: out(corners of pos(i)) += weights x gw(i)
__global__ void b_scatter_irregular(Dims d, const int *code, const double *pos, const double *gw, double *out) {
    FOR_VOXELS(d, i) {
        if ((code[i] & RC_IRREG) == 0) continue;
        const D3 p = ld(pos, i), g = ld(gw, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double w[8];
        tri_weights(k.a, k.b, k.c, w);
        for (int j = 0; j < 8; ++j) add_at(out, k.g[j], d3(w[j] * g.x, w[j] * g.y, w[j] * g.z));
    }
}

// Step 1 (adjoints of F9 and F8, own voxel), interior:
//   psi_inv1 = psi_inv - alpha J v1:         gv1 += -alpha J^T gpinv1 (and gJ = -alpha gpinv1 (x) v1,
//                                              formed in step 5 with the rest of gJ)
//   psi1 = psi + alpha trilerp(v1, psi):      gpsi += alpha (d trilerp / d pos)^T gpsi1,
//                                              record (cell of psi, alpha gpsi1) for the corners of v1
// gpsi, gv1 are updated in place (off the interior they pass through); gpinv is read only.
// J is recomputed from psi_inv (F2's own function: the same values) rather than stored.
__global__ void b_step1(Dims d, const double *__restrict__ pinv, const double *__restrict__ v1,
                        const double *__restrict__ psi, const double *__restrict__ gpinv,
                        double *__restrict__ gv1, double *__restrict__ gpsi, int *__restrict__ code,
                        double *__restrict__ abc, double *__restrict__ gw, double alpha) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        if (!interior(d, x, y, z)) {
            code[i] = 0;
            continue;
        }
        const D3 g9 = ld(gpinv, i);
        double m[9];
        fv_jacobian(d, x, y, z, pinv, m);  // J of the forward step, recomputed from psi_inv
        const D3 gv = ld(gv1, i);
        st(gv1, i, d3(gv.x - alpha * (m[0] * g9.x + m[3] * g9.y + m[6] * g9.z),
                      gv.y - alpha * (m[1] * g9.x + m[4] * g9.y + m[7] * g9.z),
                      gv.z - alpha * (m[2] * g9.x + m[5] * g9.y + m[8] * g9.z)));
        const D3 g8 = ld(gpsi, i), p = ld(psi, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double f[8];
        D3 o = g8;
        for (int q = 0; q < 3; ++q) {
            gather3(v1, k, q, f);
            const D3 dp = tri_dpos(f, k);
            const double gq = alpha * comp(g8, q);
            o = d3(o.x + gq * dp.x, o.y + gq * dp.y, o.z + gq * dp.z);
        }
        st(gpsi, i, o);
        put_rec(code, abc, gw, i, rec_code(k, x, y, z), k, d3(alpha * g8.x, alpha * g8.y, alpha * g8.z));
    }
}

// Step 3: gv1 completed by the records' gather, then the adjoint of F7
//   v1 = v - alpha (m1t v + dvv - tau dlv + q gU), q = m2 / (|rho1| + 1e-5f)   (interior; v1 = v elsewhere):
//   gv = (1 - alpha m1t) gv1, and gdvv = -alpha gv1, gdlv = alpha tau gv1, ggU = -alpha q gv1, which
//   steps 4 and 5 form from gv1 itself (stored here as gg) where they read them,
//   grho1 += -alpha (gv1 . gU) dq/drho1, dE/dtau += alpha gv1 . dlv (one partial sum per block into
//   gtau[blockIdx.x], added up in block order by k_sum_accumulate).
// gv holds gv1 on entry and gv on exit.
__global__ __launch_bounds__(GT_THREADS) void b_step3(Dims d, int tiles_x, int tiles_y, int zc, const int *code,
                                                      const double *abc, const double *gw, const double *rho1,
                                                      const double *dlv, const double *gU, double *gv, double *gg,
                                                      double *grho1, double alpha, double m1t, double m2,
                                                      double *gtau) {
    __shared__ double sh[4];
    double acc = 0;
    rec_gather_tile(d, tiles_x, tiles_y, zc, code, abc, gw, [&](int i, int x, int y, int z, D3 sc) {
        const D3 own = ld(gv, i);
        const D3 g = d3(own.x + sc.x, own.y + sc.y, own.z + sc.z);
        if (interior(d, x, y, z)) {
            const double den = fabs(rho1[i]) + (double)1e-5f;
            const D3 u = ld(gU, i), l = ld(dlv, i);
            st(gv, i, d3((1 - alpha * m1t) * g.x, (1 - alpha * m1t) * g.y, (1 - alpha * m1t) * g.z));
            const double gq = -alpha * (g.x * u.x + g.y * u.y + g.z * u.z);
            const double sgn = rho1[i] > 0 ? 1.0 : rho1[i] < 0 ? -1.0 : 0.0;
            grho1[i] += gq * (-m2 * sgn / (den * den));
            acc += alpha * (g.x * l.x + g.y * l.y + g.z * l.z);
        } else {
            st(gv, i, g);
        }
        st(gg, i, g);
    });
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) gtau[blockIdx.x] = s;
}

// voxel (x, y, z) in the range of the velocity Laplacian (x in [2, X-3], y in [2, Y-2], z in [2, Z-3])
__device__ __forceinline__ bool in_lap(Dims d, int x, int y, int z) {
    return x >= 2 && x <= d.X - 3 && y >= 2 && y <= d.Y - 2 && z >= 2 && z <= d.Z - 3;
}

// Step 4: transposes of F6 gathered into gv: dvv = (Dv)v (interior) and dlv, the velocity
// Laplacian (its range; centre 3 times, "far" neighbours (+-2) once, "near" ones (+-1) -2 times,
// chosen by the sign of the centre's component, with y-1 / z-1 as the near neighbour always).
// The z-window of one column for step 4: v and gg (gv1 after step 3's gather) at z-2..z+2 (zeros
// outside the grid; no branch below reads them there)
struct B4Win {
    D3 v[5], gg[5];
};
// the adjoints of (Dv)v and of the velocity Laplacian at a voxel j, from gv1(j) as step 3's
// transpose of F7 forms them: gdvv = -alpha gv1, gdlv = alpha tau gv1 (read only at interior j)
__device__ __forceinline__ D3 dvv_of(D3 g, double alpha) { return d3(-alpha * g.x, -alpha * g.y, -alpha * g.z); }
__device__ __forceinline__ D3 dlv_of(D3 g, double alpha, double tau) {
    return d3(alpha * tau * g.x, alpha * tau * g.y, alpha * tau * g.z);
}
// the transposes of step 4 gathered for voxel (x, y, z): its z-neighbours from the window, the
// x- and y-neighbours from memory (rows the block's other lanes and waves load at the same time)
__device__ __forceinline__ D3 b4_voxel(Dims d, int x, int y, int z, int i, const double *__restrict__ v,
                                       const double *__restrict__ gg, double alpha, double tau, const B4Win &w) {
    const int sy = d.X;
    const D3 vi = w.v[2];
    D3 a = d3(0, 0, 0);
    // (Dv)v, voxel i as the centre
    if (interior(d, x, y, z)) {
        const D3 c = vi, xp = ld(v, i + 1), xm = ld(v, i - 1), yp = ld(v, i + sy), ym = ld(v, i - sy),
                 zp = w.v[3], zm = w.v[1], G = dvv_of(w.gg[2], alpha);
        a.x += 0.5 * G.x * (2 * mx(c.x) - 2 * mn(c.x));
        if (c.y > 0) { a.y += G.x * (c.x - ym.x); a.x += G.x * c.y; }
        else { a.y += G.x * (yp.x - c.x); a.x -= G.x * c.y; }
        if (c.z > 0) { a.z += G.x * (c.x - zm.x); a.x += G.x * c.z; }
        else { a.z += G.x * (zp.x - c.x); a.x -= G.x * c.z; }
        if (c.x > 0) { a.x += G.y * (c.y - xm.y); a.y += G.y * c.x; }
        else { a.x += G.y * (xp.y - c.y); a.y -= G.y * c.x; }
        a.y += 0.5 * G.y * (2 * mx(c.y) - 2 * mn(c.y));
        if (c.z > 0) { a.z += G.y * (c.y - zm.y); a.y += G.y * c.z; }
        else { a.z += G.y * (zp.y - c.y); a.y -= G.y * c.z; }
        if (c.x > 0) { a.x += G.z * (c.z - xm.z); a.z += G.z * c.x; }
        else { a.x += G.z * (xp.z - c.z); a.z -= G.z * c.x; }
        if (c.y > 0) { a.y += G.z * (c.z - ym.z); a.z += G.z * c.y; }
        else { a.y += G.z * (yp.z - c.z); a.z -= G.z * c.y; }
        a.z += 0.5 * G.z * (2 * mx(c.z) - 2 * mn(c.z));
    }
    // voxel i as the x+1 / x-1 / y+1 / y-1 / z+1 / z-1 neighbour of an interior voxel j
    if (interior(d, x - 1, y, z)) {
        const int j = i - 1;
        const D3 cj = ld(v, j), G = dvv_of(ld(gg, j), alpha);
        a.x += 0.5 * G.x * (2 * mn(vi.x) - 2 * mx(vi.x));
        if (!(cj.x > 0)) { a.y += G.y * cj.x; a.z += G.z * cj.x; }
    }
    if (interior(d, x + 1, y, z)) {
        const int j = i + 1;
        const D3 cj = ld(v, j), G = dvv_of(ld(gg, j), alpha);
        if (cj.x > 0) { a.y -= G.y * cj.x; a.z -= G.z * cj.x; }
    }
    if (interior(d, x, y - 1, z)) {
        const int j = i - sy;
        const D3 cj = ld(v, j), G = dvv_of(ld(gg, j), alpha);
        if (!(cj.y > 0)) { a.x += G.x * cj.y; a.z += G.z * cj.y; }
        a.y += 0.5 * G.y * (2 * mn(vi.y) - 2 * mx(vi.y));
    }
    if (interior(d, x, y + 1, z)) {
        const int j = i + sy;
        const D3 cj = ld(v, j), G = dvv_of(ld(gg, j), alpha);
        if (cj.y > 0) { a.x -= G.x * cj.y; a.z -= G.z * cj.y; }
    }
    if (interior(d, x, y, z - 1)) {
        const D3 cj = w.v[1], G = dvv_of(w.gg[1], alpha);
        if (!(cj.z > 0)) { a.x += G.x * cj.z; a.y += G.y * cj.z; }
        a.z += 0.5 * G.z * (2 * mn(vi.z) - 2 * mx(vi.z));
    }
    if (interior(d, x, y, z + 1)) {
        const D3 cj = w.v[3], G = dvv_of(w.gg[3], alpha);
        if (cj.z > 0) { a.x -= G.x * cj.z; a.y -= G.y * cj.z; }
    }
    // the velocity Laplacian: dlv(j) = -(sum over axes of far - 2 near + centre)
    const auto add_v = [&](D3 g, double f) { a = d3(a.x + f * g.x, a.y + f * g.y, a.z + f * g.z); };
    const auto add = [&](int j, double f) { add_v(dlv_of(ld(gg, j), alpha, tau), f); };
    if (in_lap(d, x, y, z)) add_v(dlv_of(w.gg[2], alpha, tau), -3.0);
    if (in_lap(d, x - 2, y, z) && v[3 * (i - 2)] < 0) add(i - 2, -1.0);
    if (in_lap(d, x + 2, y, z) && !(v[3 * (i + 2)] < 0)) add(i + 2, -1.0);
    if (in_lap(d, x - 1, y, z) && v[3 * (i - 1)] < 0) add(i - 1, 2.0);
    if (in_lap(d, x + 1, y, z) && !(v[3 * (i + 1)] < 0)) add(i + 1, 2.0);
    // far y neighbour: j = i - 2 sy by linear address (for y = 0 that is (x, Y-2, z-1))
    if (y >= 2 ? in_lap(d, x, y - 2, z) : (y == 0 && in_lap(d, x, d.Y - 2, z - 1)))
        if (v[3 * (i - 2 * sy) + 1] < 0) add(i - 2 * sy, -1.0);
    if (in_lap(d, x, y + 2, z) && !(v[3 * (i + 2 * sy) + 1] < 0)) add(i + 2 * sy, -1.0);
    if (in_lap(d, x, y + 1, z)) add(i + sy, 2.0);
    if (in_lap(d, x, y, z - 2) && w.v[0].z < 0) add_v(dlv_of(w.gg[0], alpha, tau), -1.0);
    if (in_lap(d, x, y, z + 2) && !(w.v[4].z < 0)) add_v(dlv_of(w.gg[4], alpha, tau), -1.0);
    if (in_lap(d, x, y, z + 1)) add_v(dlv_of(w.gg[3], alpha, tau), 2.0);
    return a;
}

// Step 4: transposes of F6 gathered into gv: dvv = (Dv)v (interior) and dlv, the velocity
// Laplacian (its range; centre 3 times, "far" neighbours (+-2) once, "near" ones (+-1) -2 times,
// chosen by the sign of the centre's component, with y-1 / z-1 as the near neighbour always).
// One thread per (x, y) column of a 64 x 4 tile marching a chunk of zc planes, its column's
// z-neighbours in registers, so that each plane of v and gg is read from memory once per column
// (the one-thread-per-voxel form re-read them from beyond the L2 about 3 times).
__global__ __launch_bounds__(256) void b_step4(Dims d, int tiles_x, int tiles_y, int zc,
                                                const double *__restrict__ v, const double *__restrict__ gg,
                                                double alpha, double tau, double *__restrict__ gv) {
    const int b = blockIdx.x;
    const int x = (b % tiles_x) * MT_X + (int)(threadIdx.x % MT_X);
    const int y = ((b / tiles_x) % tiles_y) * MT_Y + (int)(threadIdx.x / MT_X);
    const int z0 = (b / (tiles_x * tiles_y)) * zc, z1 = min(z0 + zc, d.Z);
    if (x >= d.X || y >= d.Y) return;
    const int sz = d.X * d.Y, c0 = x + d.X * y;
    B4Win w;
    for (int k = 0; k < 5; ++k) {  // planes z0-2 .. z0+2
        const int p = z0 - 2 + k;
        const bool ok = p >= 0 && p < d.Z;
        w.v[k] = ld_or0(v, c0 + sz * p, ok);
        w.gg[k] = ld_or0(gg, c0 + sz * p, ok);
    }
    for (int z = z0; z < z1; ++z) {
        // the entries of the next step, plane z+3, loaded a step ahead
        const D3 nv = ld_or0(v, c0 + sz * (z + 3), z + 3 < d.Z), ng = ld_or0(gg, c0 + sz * (z + 3), z + 3 < d.Z);
        const int i = c0 + sz * z;
        const D3 a = b4_voxel(d, x, y, z, i, v, gg, alpha, tau, w);
        const D3 o = ld(gv, i);
        st(gv, i, d3(o.x + a.x, o.y + a.y, o.z + a.z));
        for (int k = 0; k < 4; ++k) {
            w.v[k] = w.v[k + 1];
            w.gg[k] = w.gg[k + 1];
        }
        w.v[4] = nv;
        w.gg[4] = ng;
    }
}

// Step 5 (own voxel, interior): the adjoint of F5, gU = ((S - Tp) gS + w Lo) det_ref(J):
//   gLo = w det ggU, gTp = -det (ggU . gS), gJ = -alpha gpinv1 (x) v1 (F9's part, step 1's
//   transpose) + (ggU . ((S - Tp) gS + w Lo)) d det_ref / dJ, written once (0 off the interior),
//   dE/dw_reg += det (ggU . Lo) (one partial sum per block into gwsum[blockIdx.x], added up in block
//   order by k_sum_accumulate);
// then of F4 at this voxel's position psi_inv: gpinv += (d trilerp(L) / d pos)^T gLo +
// (d trilerp(T) / d pos)^T gTp, and the record (cell of psi_inv, gLo) for the corners of L.
// gL is cleared here (the irregular scatter and the gather fill it).
__global__ __launch_bounds__(256) void b_step5(Dims d, const double *__restrict__ S, const double *__restrict__ gS,
                             const double *__restrict__ T, const double *__restrict__ L,
                             const double *__restrict__ pinv, const double *__restrict__ v1,
                             const double *__restrict__ rho1, const double *__restrict__ gg,
                             double *__restrict__ gJ, double *__restrict__ gpinv, double *__restrict__ gL,
                             int *__restrict__ code, double *__restrict__ abc, double *__restrict__ gw, double alpha,
                             double m2, double w_reg, double *__restrict__ gwsum) {
    __shared__ double sh[4];
    double acc = 0;
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        st(gL, i, d3(0, 0, 0));
        if (!interior(d, x, y, z)) {
            for (int k = 0; k < 9; ++k) gJ[gj_at(d, i, k)] = 0.0;
            code[i] = 0;
            continue;
        }
        const D3 g9 = ld(gpinv, i), w1 = ld(v1, i);  // gpinv still holds gpinv1: steps 1-4 read it only
        double m[3][3], dg[3][3], J[9];
        fv_jacobian(d, x, y, z, pinv, J);  // recomputed from psi_inv, as in step 1
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) m[r][c] = J[3 * r + c];
        // the taps of L and T at psi_inv, read once: Lo and Tp of the forward step (F4, the
        // operations of fv_interp) and the position derivatives of both samples
        const D3 p = ld(pinv, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double f[8];
        D3 lo, dpl[3];
        gather3(L, k, 0, f);
        lo.x = tri(f, k.a, k.b, k.c);
        dpl[0] = tri_dpos(f, k);
        gather3(L, k, 1, f);
        lo.y = tri(f, k.a, k.b, k.c);
        dpl[1] = tri_dpos(f, k);
        gather3(L, k, 2, f);
        lo.z = tri(f, k.a, k.b, k.c);
        dpl[2] = tri_dpos(f, k);
        gather1(T, k, f);
        const double tp = tri(f, k.a, k.b, k.c);
        const D3 dt = tri_dpos(f, k);
        const double det = det_ref(m), diff = S[i] - tp;
        // ggU = -alpha q gv1, q = m2 / (|rho1| + 1e-5f), as step 3's transpose of F7 forms it
        const double q = m2 / (fabs(rho1[i]) + (double)1e-5f);
        const D3 g1 = ld(gg, i), s = ld(gS, i);
        const D3 g = d3(-alpha * q * g1.x, -alpha * q * g1.y, -alpha * q * g1.z);
        const double gdet = g.x * (diff * s.x + w_reg * lo.x) + g.y * (diff * s.y + w_reg * lo.y) +
                            g.z * (diff * s.z + w_reg * lo.z);
        const D3 glo = d3(w_reg * det * g.x, w_reg * det * g.y, w_reg * det * g.z);
        const double gtp = -det * (g.x * s.x + g.y * s.y + g.z * s.z);
        det_ref_grad(m, dg);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                gJ[gj_at(d, i, 3 * r + c)] = -alpha * comp(g9, r) * comp(w1, c) + gdet * dg[r][c];
        acc += det * (g.x * lo.x + g.y * lo.y + g.z * lo.z);
        D3 o = d3(0, 0, 0);
        for (int q = 0; q < 3; ++q) {
            const double gq = comp(glo, q);
            o = d3(o.x + gq * dpl[q].x, o.y + gq * dpl[q].y, o.z + gq * dpl[q].z);
        }
        o = d3(o.x + gtp * dt.x, o.y + gtp * dt.y, o.z + gtp * dt.z);
        const D3 gp = ld(gpinv, i);
        st(gpinv, i, d3(gp.x + o.x, gp.y + o.y, gp.z + o.z));
        put_rec(code, abc, gw, i, rec_code(k, x, y, z), k, glo);
    }
    const double sm = block_sum(acc, sh);
    if (threadIdx.x == 0) gwsum[blockIdx.x] = sm;
}

// Step 7: gL += the records' gather (the transpose of Lo = trilerp(L, psi_inv))
__global__ __launch_bounds__(GT_THREADS) void b_step7(Dims d, int tiles_x, int tiles_y, int zc, const int *code,
                                                      const double *abc, const double *gw, double *gL) {
    rec_gather_tile(d, tiles_x, tiles_y, zc, code, abc, gw, [&](int i, int, int, int, D3 g) {
        const D3 o = ld(gL, i);
        st(gL, i, d3(o.x + g.x, o.y + g.y, o.z + g.z));
    });
}

// The z-window of one column for step 8: gL, the z-column (entries 2, 5, 8) of gJ and grho1 at
// z-1, z, z+1, rho at z and z+1, v.z at z-1 (zeros outside the grid; no branch reads them there)
struct B8Win {
    D3 gl[3], gj[3];
    double gr[3], r0, rp, vzm;
};
__device__ __forceinline__ D3 ld_jz(Dims d, const double *__restrict__ gJ, int i, bool ok) {
    return ok ? d3(gJ[gj_at(d, i, 2)], gJ[gj_at(d, i, 5)], gJ[gj_at(d, i, 8)]) : d3(0, 0, 0);
}

// Step 8: transposes of F3, F2 and F1 gathered into voxel i:
//   L(j) = 6 psi(j) - (sum of its 6 neighbours), interior j:       gpsi(i)
//   J(j)[r][a] = (psi_inv(j + e_a) - psi_inv(j - e_a))[r] / 2:      gpinv(i)
//   rho1(j) = rho(j) + alpha div(rho, v)(j) (upwind fluxes):        grho(i) = grho1(i) + ..., gv(i)
// One thread per (x, y) column of a 64 x 4 tile marching a chunk of zc planes, the z-neighbours
// from the column's window and the x- and y-neighbours from memory.
__global__ __launch_bounds__(256) void b_step8(Dims d, int tiles_x, int tiles_y, int zc,
                                                const double *__restrict__ gL, const double *__restrict__ gJ,
                                                const double *__restrict__ rho, const double *__restrict__ v,
                                                const double *__restrict__ grho1, double *__restrict__ gpsi,
                                                double *__restrict__ gpinv, double *__restrict__ grho,
                                                double *__restrict__ gv, double alpha) {
    const int b = blockIdx.x;
    const int x = (b % tiles_x) * MT_X + (int)(threadIdx.x % MT_X);
    const int y = ((b / tiles_x) % tiles_y) * MT_Y + (int)(threadIdx.x / MT_X);
    const int z0 = (b / (tiles_x * tiles_y)) * zc, z1 = min(z0 + zc, d.Z);
    if (x >= d.X || y >= d.Y) return;
    const int sy = d.X, sz = d.X * d.Y, c0 = x + d.X * y;
    B8Win w;
    for (int k = 0; k < 3; ++k) {  // planes z0-1 .. z0+1
        const int p = z0 - 1 + k;
        const bool ok = p >= 0 && p < d.Z;
        w.gl[k] = ld_or0(gL, c0 + sz * p, ok);
        w.gj[k] = ld_jz(d, gJ, c0 + sz * p, ok);
        w.gr[k] = ld1_or0(grho1, c0 + sz * p, ok);
    }
    w.r0 = rho[c0 + sz * z0];
    w.rp = ld1_or0(rho, c0 + sz * (z0 + 1), z0 + 1 < d.Z);
    w.vzm = ld1_or0(v, 3 * (c0 + sz * (z0 - 1)) + 2, z0 >= 1);
    for (int z = z0; z < z1; ++z) {
        const int i = c0 + sz * z;
        // the entries of the next step (plane z+2), loaded a step ahead
        const bool okn = z + 2 < d.Z;
        const int in2 = i + 2 * sz;
        const D3 n_gl = ld_or0(gL, in2, okn), n_gj = ld_jz(d, gJ, in2, okn);
        const double n_gr = ld1_or0(grho1, in2, okn), n_r = ld1_or0(rho, in2, okn);
        const int nb[6] = {i + 1, i - 1, i + sy, i - sy, i + sz, i - sz};
        const bool nin[6] = {interior(d, x + 1, y, z), interior(d, x - 1, y, z), interior(d, x, y + 1, z),
                             interior(d, x, y - 1, z), interior(d, x, y, z + 1), interior(d, x, y, z - 1)};
        const bool in = interior(d, x, y, z);
        // F3
        D3 gp = ld(gpsi, i);
        if (in) {
            const D3 g = w.gl[1];
            gp = d3(gp.x + 6 * g.x, gp.y + 6 * g.y, gp.z + 6 * g.z);
        }
        for (int k = 0; k < 6; ++k)
            if (nin[k]) {
                const D3 g = k == 4 ? w.gl[2] : k == 5 ? w.gl[0] : ld(gL, nb[k]);
                gp = d3(gp.x - g.x, gp.y - g.y, gp.z - g.z);
            }
        st(gpsi, i, gp);
        // F2: i = j - e_a (j = nb[2a]) gives -1/2, i = j + e_a (j = nb[2a + 1]) gives +1/2 of gJ(j)[.][a]
        D3 gq = ld(gpinv, i);
        for (int ax = 0; ax < 3; ++ax) {
            if (nin[2 * ax]) {
                const int j = nb[2 * ax];
                const D3 c = ax == 2 ? w.gj[2]
                                     : d3(gJ[gj_at(d, j, ax)], gJ[gj_at(d, j, 3 + ax)], gJ[gj_at(d, j, 6 + ax)]);
                gq = d3(gq.x - 0.5 * c.x, gq.y - 0.5 * c.y, gq.z - 0.5 * c.z);
            }
            if (nin[2 * ax + 1]) {
                const int j = nb[2 * ax + 1];
                const D3 c = ax == 2 ? w.gj[0]
                                     : d3(gJ[gj_at(d, j, ax)], gJ[gj_at(d, j, 3 + ax)], gJ[gj_at(d, j, 6 + ax)]);
                gq = d3(gq.x + 0.5 * c.x, gq.y + 0.5 * c.y, gq.z + 0.5 * c.z);
            }
        }
        st(gpinv, i, gq);
        // F1: the fluxes t_a1(j) = v(j).a > 0 ? -v(j).a rho(j) : -v(j).a rho(j + e_a) and
        // t_a2(j) = w > 0 ? w rho(j - e_a) : w rho(j), w = v(j - e_a).a
        double gr = w.gr[1];
        D3 ga = d3(0, 0, 0);
        const D3 vi = ld(v, i);
        const double ri = w.r0;
        const int step[3] = {1, sy, sz};
        for (int ax = 0; ax < 3; ++ax) {
            const double via = comp(vi, ax);
            double gax = 0;
            if (in) {  // j = i
                const double g = alpha * w.gr[1];
                if (via > 0) {
                    gr += -g * via;
                    gax += -g * ri;
                } else {
                    gax += -g * (ax == 2 ? w.rp : rho[i + step[ax]]);
                }
                const double wv = ax == 2 ? w.vzm : v[3 * (i - step[ax]) + ax];
                if (!(wv > 0)) gr += g * wv;
            }
            if (nin[2 * ax + 1]) {  // j = i - e_a: i is j + e_a in t_a1(j)
                const int j = i - step[ax];
                const double g = alpha * (ax == 2 ? w.gr[0] : grho1[j]), vj = ax == 2 ? w.vzm : v[3 * j + ax];
                if (!(vj > 0)) gr += -g * vj;
            }
            if (nin[2 * ax]) {  // j = i + e_a: i is j - e_a in t_a2(j), w = v(i).a
                const int j = i + step[ax];
                const double g = alpha * (ax == 2 ? w.gr[2] : grho1[j]);
                if (via > 0) {
                    gr += g * via;
                    gax += g * ri;
                } else {
                    gax += g * (ax == 2 ? w.rp : rho[j]);
                }
            }
            if (ax == 0) ga.x = gax; else if (ax == 1) ga.y = gax; else ga.z = gax;
        }
        grho[i] = gr;
        const D3 o = ld(gv, i);
        st(gv, i, d3(o.x + ga.x, o.y + ga.y, o.z + ga.z));
        // slide to plane z+1
        w.gl[0] = w.gl[1];
        w.gl[1] = w.gl[2];
        w.gl[2] = n_gl;
        w.gj[0] = w.gj[1];
        w.gj[1] = w.gj[2];
        w.gj[2] = n_gj;
        w.gr[0] = w.gr[1];
        w.gr[1] = w.gr[2];
        w.gr[2] = n_gr;
        w.r0 = w.rp;
        w.rp = n_r;
        w.vzm = vi.z;
    }
}

// grad S on the interior ((S(+1) - S(-1)) / 2 per axis), 0 elsewhere; FP64 copies of the inputs
__global__ void f_setup(Dims d, const float *tgt, const float *src, double *T, double *S, double *gS) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        T[i] = tgt[i];
        S[i] = src[i];
        D3 g = d3(0, 0, 0);
        if (interior(d, x, y, z)) {
            g.x = ((double)src[lin(d, x + 1, y, z)] - (double)src[lin(d, x - 1, y, z)]) / 2.0;
            g.y = ((double)src[lin(d, x, y + 1, z)] - (double)src[lin(d, x, y - 1, z)]) / 2.0;
            g.z = ((double)src[lin(d, x, y, z + 1)] - (double)src[lin(d, x, y, z - 1)]) / 2.0;
        }
        st(gS, i, g);
    }
}

// initial state: rho = rho_0, v = v0 (or 0), psi = psi_inv = identity
__global__ void f_init_state(Dims d, double *rho, double *v, double *psi, double *pinv, const double *v0, double rho_0) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        rho[i] = rho_0;
        st(v, i, v0 ? ld(v0, i) : d3(0, 0, 0));
        st(psi, i, d3(x, y, z));
        st(pinv, i, d3(x, y, z));
    }
}

// *out = sum of np partials, one block of 256 threads, in a fixed order (deterministic)
__global__ __launch_bounds__(256) void k_sum_final(const double *partials, int np, double *out) {
    __shared__ double sh[256];
    double s = 0;
    for (int j = threadIdx.x; j < np; j += 256) s += partials[j];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// *out += the sum of np partials, in a fixed order (the same bits on every run, unlike atomics)
__global__ __launch_bounds__(256) void k_sum_accumulate(const double *partials, int np, double *out) {
    __shared__ double sh[256];
    double s = 0;
    for (int j = threadIdx.x; j < np; j += 256) s += partials[j];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out += sh[0];
}

// ---- single operators and their adjoints (the per-operator forward/backward pairs) ------------

// out(p) = trilerp(f, pos(p)) over all voxels, positions clamped into the grid
__global__ void op_warp_fwd(Dims d, const double *f, const double *pos, double *out) {
    FOR_VOXELS(d, i) {
        const D3 p = ld(pos, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double t[8];
        gather1(f, k, t);
        out[i] = tri(t, k.a, k.b, k.c);
    }
}
// gf(corners of pos(p)) += weight x gout(p) (FP64 atomics); gpos(p) += gout(p) d trilerp / d pos
__global__ void op_warp_bwd(Dims d, const double *f, const double *pos, const double *gout, double *gf,
                            double *gpos) {
    FOR_VOXELS(d, i) {
        const double g = gout[i];
        const D3 p = ld(pos, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        double w[8], t[8];
        tri_weights(k.a, k.b, k.c, w);
        for (int j = 0; j < 8; ++j) aadd(&gf[k.g[j]], w[j] * g);
        gather1(f, k, t);
        const D3 dp = tri_dpos(t, k), o = ld(gpos, i);
        st(gpos, i, d3(o.x + g * dp.x, o.y + g * dp.y, o.z + g * dp.z));
    }
}
// grad(p) = (f(+1) - f(-1)) / 2 per axis on the interior, 0 elsewhere (vector_fields.cu:155-192)
__global__ void op_grad_fwd(Dims d, const double *f, double *g) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        D3 o = d3(0, 0, 0);
        if (interior(d, x, y, z)) {
            const int sy = d.X, sz = d.X * d.Y;
            o = d3((f[i + 1] - f[i - 1]) / 2.0, (f[i + sy] - f[i - sy]) / 2.0, (f[i + sz] - f[i - sz]) / 2.0);
        }
        st(g, i, o);
    }
}
// gf(i) += sum over the interior voxels j whose difference reads i: +gg(j).a / 2 for i = j + e_a,
// -gg(j).a / 2 for i = j - e_a (a gather)
__global__ void op_grad_bwd(Dims d, const double *gg, double *gf) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        const int sy = d.X, sz = d.X * d.Y;
        double a = 0;
        if (interior(d, x - 1, y, z)) a += gg[3 * (i - 1)] / 2.0;
        if (interior(d, x + 1, y, z)) a -= gg[3 * (i + 1)] / 2.0;
        if (interior(d, x, y - 1, z)) a += gg[3 * (i - sy) + 1] / 2.0;
        if (interior(d, x, y + 1, z)) a -= gg[3 * (i + sy) + 1] / 2.0;
        if (interior(d, x, y, z - 1)) a += gg[3 * (i - sz) + 2] / 2.0;
        if (interior(d, x, y, z + 1)) a -= gg[3 * (i + sz) + 2] / 2.0;
        gf[i] += a;
    }
}
// gpsi(i) += 6 gL(i) (i interior) - sum of gL over i's interior neighbours (the transpose of
// L = 6 psi - sum of the neighbours, a gather)
__global__ void op_laplacian_bwd(Dims d, const double *gL, double *gpsi) {
    FOR_VOXELS(d, i) {
        XYZ(d, i);
        const int sy = d.X, sz = d.X * d.Y;
        D3 gp = ld(gpsi, i);
        if (interior(d, x, y, z)) {
            const D3 g = ld(gL, i);
            gp = d3(gp.x + 6 * g.x, gp.y + 6 * g.y, gp.z + 6 * g.z);
        }
        const int nb[6] = {i + 1, i - 1, i + sy, i - sy, i + sz, i - sz};
        const bool nin[6] = {interior(d, x + 1, y, z), interior(d, x - 1, y, z), interior(d, x, y + 1, z),
                             interior(d, x, y - 1, z), interior(d, x, y, z + 1), interior(d, x, y, z - 1)};
        for (int k = 0; k < 6; ++k)
            if (nin[k]) {
                const D3 g = ld(gL, nb[k]);
                gp = d3(gp.x - g.x, gp.y - g.y, gp.z - g.z);
            }
        st(gpsi, i, gp);
    }
}

}  // namespace frd

// ---- host driver --------------------------------------------------------------------------------

namespace {

using frd::Dims;

constexpr int TPB = 256;  // threads per block of every kernel above (the block sums assume 256)

// One solver state: density, velocity, psi, psi_inv (1 + 3 + 3 + 3 doubles per voxel).
struct State {
    double *rho, *v, *psi, *pinv;
};
// The intermediates of one iteration that its adjoint reads.
struct Inter {
    double *L, *gU, *dlv;
};

int err(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return frk::set_error(code, buf);
}

#define ADJ_TRY(expr)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (expr);                                                                              \
        if (e_ != hipSuccess) return err(FR_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                                         __LINE__);                                                          \
    } while (0)

}  // namespace

struct fr_adjoint {
    Dims d{};
    long n = 0;
    int K = 0, ckpt = 0, nseg = 0;
    double alpha = 0, rho_0 = 0, p = 0, C = 0;
    int reset_iter = -1, no_labels = 1;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    size_t bytes = 0;
    // inputs (FP64 copies of the images, grad S, labels)
    double *T = nullptr, *S = nullptr, *gS = nullptr;
    int16_t *LT = nullptr, *LS = nullptr;
    // checkpoints: ck[s] is the state before iteration s*ckpt + 1; fin the state after K iterations
    std::vector<State> ck;
    State fin{};
    // the segment being differentiated: nxt[j] is the state after its (j+1)-th iteration,
    // inter[j] that iteration's intermediates (the plain forward uses nxt[0..1] and inter[0])
    std::vector<State> nxt;
    std::vector<Inter> inter;
    // adjoint of the state (grho double-buffered: its kernel scatters into the older iteration's)
    double *grho[2] = {nullptr, nullptr}, *gv = nullptr, *gpsi = nullptr, *gpinv = nullptr;
    // adjoints of one iteration's intermediates
    double *gJ = nullptr, *gL = nullptr, *gg = nullptr;
    // records of the trilinear transposes: per source voxel a code, the cell fractions, the adjoint
    int *code = nullptr;
    double *abc = nullptr, *gw = nullptr;
    double *scal = nullptr;  // [0] objective, [1] dE/dw_reg, [2] dE/dtau
    double *partials = nullptr;
    double *part_tau = nullptr, *part_w = nullptr;  // per-block partials of dE/dtau, dE/dw_reg of one iteration
    int grid = 0;
    // launch shape of the record gathers: 64 x 4 tiles, zc z-planes per block (gt_zc_auto: the
    // choice made at creation, about two blocks per CU)
    int gt_tiles_x = 0, gt_tiles_y = 0, gt_zc = 0, gt_blocks = 0, gt_zc_auto = 0;
    // launch shape of the z-marching kernels (f_step, b_step4, b_step8): the same 64 x 4 tiles,
    // mt_zc planes per block
    int mt_tiles_x = 0, mt_tiles_y = 0, mt_zc = 0, mt_blocks = 0;
    // per-iteration scalars of the last forward (index k-1 for iteration k)
    std::vector<double> m1t, m2;
    double w_reg = 0, tau = 0, lw = 0, energy = 0;
    bool have_inputs = false, have_forward = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    float fwd_ms = 0, bwd_ms = 0;

    template <class T>
    int alloc(T **ptr, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
        if (e != hipSuccess) {
            *ptr = nullptr;
            return err(FR_ERR_ALLOC, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        }
        allocs.push_back(q);
        bytes += count * sizeof(T);
        *ptr = static_cast<T *>(q);
        return FR_OK;
    }
    int alloc_state(State *s) {
        double *b = nullptr;
        if (int rc = alloc(&b, 10 * (size_t)n)) return rc;
        s->rho = b;
        s->v = b + n;
        s->psi = b + 4 * n;
        s->pinv = b + 7 * n;
        return FR_OK;
    }
    int alloc_inter(Inter *t) {
        double *b = nullptr;
        if (int rc = alloc(&b, 9 * (size_t)n)) return rc;
        t->L = b;
        t->gU = b + 3 * n;
        t->dlv = b + 6 * n;
        return FR_OK;
    }
    ~fr_adjoint() {
        for (void *q : allocs) (void)hipFree(q);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// t of iteration k (1-based) as the solver loop advances it (solver.cu:423-428): t_1 = alpha,
// t_k = t_{k-1} + alpha, reset to alpha when k is a multiple of RESET_ITER; then
// m1t = (p + 1) / t and m2 = C p^2 t^(p - 2) (solver.cu:198-207), all in double.
void scalars_of(const fr_adjoint *h, double t, double *m1t, double *m2) {
    const double m1 = h->p + 1.0;
    *m1t = m1 / t;
    *m2 = h->C * pow(h->p, 2.0) * pow(t, h->p - 2.0);
}
double t_of(const fr_adjoint *h, int k) {
    double t = h->alpha;
    for (int q = 2; q <= k; ++q) {
        t += h->alpha;
        if (h->reset_iter > 0 && q % h->reset_iter == 0) t = h->alpha;
    }
    return t;
}
void schedule(fr_adjoint *h) {
    h->m1t.assign(h->K, 0.0);
    h->m2.assign(h->K, 0.0);
    double t = h->alpha;
    for (int k = 1; k <= h->K; ++k) {
        if (k >= 2) {
            t += h->alpha;
            if (h->reset_iter > 0 && k % h->reset_iter == 0) t = h->alpha;
        }
        scalars_of(h, t, &h->m1t[k - 1], &h->m2[k - 1]);
    }
}

// one iteration with velocity scalars m1t, m2: state `in` -> state `out`, intermediates into `it`
int forward_iteration(fr_adjoint *h, const State &in, const State &out, const Inter &it, double m1t, double m2,
                      bool record) {
    const Dims d = h->d;
    const int g = h->grid;
    hipStream_t s = h->stream;
    const double a = h->alpha;
    frd::f_laplacian<<<g, TPB, 0, s>>>(d, in.psi, it.L);
    frd::f_step<<<h->mt_blocks, 256, 0, s>>>(d, h->mt_tiles_x, h->mt_tiles_y, h->mt_zc, in.rho, in.v, in.pinv, it.L,
                                             h->T, h->S, h->gS, out.rho, out.v, out.pinv, it.gU, it.dlv,
                                             record ? 1 : 0, a, h->w_reg, m1t, h->tau, m2);
    frd::f_psi<<<g, TPB, 0, s>>>(d, in.psi, out.v, out.psi, a);
    ADJ_TRY(hipGetLastError());
    return FR_OK;
}

// adjoint of one iteration (velocity scalars m1t, m2): the adjoint of state `nx` (grho[cur], gv,
// gpsi, gpinv) becomes that of `in`
int backward_iteration(fr_adjoint *h, const State &in, const State &nx, const Inter &it, double m1t, double m2,
                       int *cur) {
    const Dims d = h->d;
    const int g = h->grid;
    hipStream_t s = h->stream;
    const double a = h->alpha;
    double *gr1 = h->grho[*cur], *gr0 = h->grho[1 - *cur];
    frd::b_step1<<<g, TPB, 0, s>>>(d, in.pinv, nx.v, in.psi, h->gpinv, h->gv, h->gpsi, h->code, h->abc, h->gw, a);
    frd::b_scatter_irregular<<<g, TPB, 0, s>>>(d, h->code, in.psi, h->gw, h->gv);
    frd::b_step3<<<h->gt_blocks, frd::GT_THREADS, 0, s>>>(d, h->gt_tiles_x, h->gt_tiles_y, h->gt_zc, h->code, h->abc,
                                                        h->gw, nx.rho, it.dlv, it.gU, h->gv, h->gg, gr1, a, m1t,
                                                        m2, h->part_tau);
    frd::k_sum_accumulate<<<1, 256, 0, s>>>(h->part_tau, h->gt_blocks, h->scal + 2);
    frd::b_step4<<<h->mt_blocks, 256, 0, s>>>(d, h->mt_tiles_x, h->mt_tiles_y, h->mt_zc, in.v, h->gg, a, h->tau,
                                              h->gv);
    frd::b_step5<<<g, TPB, 0, s>>>(d, h->S, h->gS, h->T, it.L, in.pinv, nx.v, nx.rho, h->gg, h->gJ, h->gpinv,
                                   h->gL, h->code, h->abc, h->gw, a, m2, h->w_reg, h->part_w);
    frd::k_sum_accumulate<<<1, 256, 0, s>>>(h->part_w, g, h->scal + 1);
    frd::b_scatter_irregular<<<g, TPB, 0, s>>>(d, h->code, in.pinv, h->gw, h->gL);
    frd::b_step7<<<h->gt_blocks, frd::GT_THREADS, 0, s>>>(d, h->gt_tiles_x, h->gt_tiles_y, h->gt_zc, h->code, h->abc,
                                                        h->gw, h->gL);
    frd::b_step8<<<h->mt_blocks, 256, 0, s>>>(d, h->mt_tiles_x, h->mt_tiles_y, h->mt_zc, h->gL, h->gJ, in.rho, in.v, gr1,
                                              h->gpsi, h->gpinv, gr0, h->gv, a);
    ADJ_TRY(hipGetLastError());
    *cur = 1 - *cur;
    return FR_OK;
}

int copy_state(fr_adjoint *h, const State &st, double *v, double *rho, double *psi, double *pinv) {
    const size_t n = (size_t)h->n;
    if (v) ADJ_TRY(hipMemcpyAsync(v, st.v, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (rho) ADJ_TRY(hipMemcpyAsync(rho, st.rho, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (psi) ADJ_TRY(hipMemcpyAsync(psi, st.psi, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    if (pinv) ADJ_TRY(hipMemcpyAsync(pinv, st.pinv, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    ADJ_TRY(hipStreamSynchronize(h->stream));
    return FR_OK;
}

}  // namespace

extern "C" int fr_adjoint_create(const fr_params *p, int K, int ckpt, fr_adjoint **out) {
    if (!p || !out) return err(FR_ERR_INVALID, "null argument");
    *out = nullptr;
    for (int k = 0; k < 3; ++k)
        if (p->vol_dims[k] < 3)
            return err(FR_ERR_INVALID, "VOL_DIMS must be >= 3 along every axis (got %d %d %d)", p->vol_dims[0],
                       p->vol_dims[1], p->vol_dims[2]);
    if (K < 1) return err(FR_ERR_INVALID, "K (unrolled iterations) must be >= 1 (got %d)", K);
    if (ckpt < 1 || ckpt > K) return err(FR_ERR_INVALID, "checkpoint interval must be in [1, K] (got %d, K = %d)", ckpt, K);
    if (p->no_labels < 1 || p->no_labels > 32768) return err(FR_ERR_INVALID, "NO_LABELS must be in [1, 32768]");
    if (!std::isfinite(p->alpha) || !std::isfinite(p->rho_0) || !std::isfinite(p->p) || !std::isfinite(p->C))
        return err(FR_ERR_INVALID, "ALPHA, RHO_0, p and C must be finite");
    if (9LL * p->vol_dims[0] * p->vol_dims[1] * p->vol_dims[2] >= (1LL << 31))
        return err(FR_ERR_INVALID, "volume of %lld voxels is too large for the differentiated solver (9*N must be < 2^31)",
                   (long long)p->vol_dims[0] * p->vol_dims[1] * p->vol_dims[2]);
    fr_adjoint *h = new (std::nothrow) fr_adjoint;
    if (!h) return err(FR_ERR_ALLOC, "out of host memory");
    h->d = Dims{p->vol_dims[0], p->vol_dims[1], p->vol_dims[2]};
    h->n = (long)p->vol_dims[0] * p->vol_dims[1] * p->vol_dims[2];
    h->K = K;
    h->ckpt = ckpt;
    h->nseg = (K + ckpt - 1) / ckpt;
    h->alpha = (double)p->alpha;
    h->rho_0 = (double)p->rho_0;
    h->p = (double)p->p;
    h->C = (double)p->C;
    h->reset_iter = p->reset_iter;
    h->no_labels = p->no_labels;
    const long nb = (h->n + TPB - 1) / TPB;
    h->grid = (int)(nb < 8192 ? nb : 8192);
    {
        // about two gather blocks per CU (each holds a 62 KB stage), every block a run of z-planes
        int dev = 0, n_cu = 256;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu < 1)
            n_cu = 256;
        h->gt_tiles_x = (h->d.X + frd::GT_X - 1) / frd::GT_X;
        h->gt_tiles_y = (h->d.Y + frd::GT_Y - 1) / frd::GT_Y;
        const long tiles = (long)h->gt_tiles_x * h->gt_tiles_y;
        long chunks = 2L * n_cu / tiles;
        chunks = chunks < 1 ? 1 : chunks > h->d.Z ? h->d.Z : chunks;
        h->gt_zc = h->gt_zc_auto = (int)((h->d.Z + chunks - 1) / chunks);
        h->gt_blocks = (int)(tiles * ((h->d.Z + h->gt_zc - 1) / h->gt_zc));
        // about MT_BLOCKS_PER_CU blocks per CU for the z-marching kernels
        h->mt_tiles_x = (h->d.X + frd::MT_X - 1) / frd::MT_X;
        h->mt_tiles_y = (h->d.Y + frd::MT_Y - 1) / frd::MT_Y;
        const long mtiles = (long)h->mt_tiles_x * h->mt_tiles_y;
        long mchunks = (long)frd::MT_BLOCKS_PER_CU * n_cu / mtiles;
        mchunks = mchunks < 1 ? 1 : mchunks > h->d.Z ? h->d.Z : mchunks;
        h->mt_zc = (int)((h->d.Z + mchunks - 1) / mchunks);
        h->mt_blocks = (int)(mtiles * ((h->d.Z + h->mt_zc - 1) / h->mt_zc));
    }
    const size_t n = (size_t)h->n;
    int rc = FR_OK;
    auto fail = [&](int code) {
        delete h;
        return code;
    };
    if (hipStreamCreate(&h->stream) != hipSuccess || hipEventCreate(&h->ev[0]) != hipSuccess ||
        hipEventCreate(&h->ev[1]) != hipSuccess)
        return fail(err(FR_ERR_HIP, "stream / event creation failed"));
    if ((rc = h->alloc(&h->T, n)) || (rc = h->alloc(&h->S, n)) || (rc = h->alloc(&h->gS, 3 * n)) ||
        (rc = h->alloc(&h->LT, n)) || (rc = h->alloc(&h->LS, n)))
        return fail(rc);
    h->ck.resize(h->nseg);
    for (State &st : h->ck)
        if ((rc = h->alloc_state(&st))) return fail(rc);
    if ((rc = h->alloc_state(&h->fin))) return fail(rc);
    h->nxt.resize(ckpt < 2 ? 2 : ckpt);
    for (State &st : h->nxt)
        if ((rc = h->alloc_state(&st))) return fail(rc);
    h->inter.resize(ckpt);
    for (Inter &it : h->inter)
        if ((rc = h->alloc_inter(&it))) return fail(rc);
    if ((rc = h->alloc(&h->grho[0], n)) || (rc = h->alloc(&h->grho[1], n)) || (rc = h->alloc(&h->gv, 3 * n)) ||
        (rc = h->alloc(&h->gpsi, 3 * n)) || (rc = h->alloc(&h->gpinv, 3 * n)) || (rc = h->alloc(&h->gJ, 9 * n)) ||
        (rc = h->alloc(&h->gL, 3 * n)) || (rc = h->alloc(&h->code, n)) || (rc = h->alloc(&h->abc, 3 * n)) ||
        (rc = h->alloc(&h->gw, 3 * n)) ||
        (rc = h->alloc(&h->gg, 3 * n)) ||
        (rc = h->alloc(&h->scal, 4)) || (rc = h->alloc(&h->partials, (size_t)h->grid)) ||
        (rc = h->alloc(&h->part_tau, (size_t)h->gt_tiles_x * h->gt_tiles_y * h->d.Z)) || (rc = h->alloc(&h->part_w, (size_t)h->grid)))
        return fail(rc);
    *out = h;
    return FR_OK;
}

extern "C" int fr_adjoint_destroy(fr_adjoint *h) {
    if (h) {
        (void)hipStreamSynchronize(h->stream);
        delete h;
    }
    return FR_OK;
}

extern "C" int fr_adjoint_set_images(fr_adjoint *h, const float *d_tgt, const int16_t *d_tgt_lab, const float *d_src,
                                     const int16_t *d_src_lab) {
    if (!h || !d_tgt || !d_tgt_lab || !d_src || !d_src_lab) return err(FR_ERR_INVALID, "null argument");
    ADJ_TRY(hipDeviceSynchronize());  // the inputs may have been written on any stream
    frd::f_setup<<<h->grid, TPB, 0, h->stream>>>(h->d, d_tgt, d_src, h->T, h->S, h->gS);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipMemcpyAsync(h->LT, d_tgt_lab, (size_t)h->n * sizeof(int16_t), hipMemcpyDeviceToDevice, h->stream));
    ADJ_TRY(hipMemcpyAsync(h->LS, d_src_lab, (size_t)h->n * sizeof(int16_t), hipMemcpyDeviceToDevice, h->stream));
    ADJ_TRY(hipStreamSynchronize(h->stream));
    h->have_inputs = true;
    h->have_forward = false;
    return FR_OK;
}

extern "C" int fr_adjoint_forward(fr_adjoint *h, const double *d_v0, double w_reg, double tau, double label_weight,
                                  double *energy) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (!h->have_inputs) return err(FR_ERR_STATE, "fr_adjoint_forward before fr_adjoint_set_images");
    if (!std::isfinite(w_reg) || !std::isfinite(tau) || !std::isfinite(label_weight))
        return err(FR_ERR_INVALID, "w_reg, tau and label_weight must be finite");
    if (d_v0) ADJ_TRY(hipDeviceSynchronize());  // v0 may have been written on any stream
    h->w_reg = w_reg;
    h->tau = tau;
    h->lw = label_weight;
    h->have_forward = false;
    schedule(h);
    hipStream_t s = h->stream;
    ADJ_TRY(hipEventRecord(h->ev[0], s));
    frd::f_init_state<<<h->grid, TPB, 0, s>>>(h->d, h->ck[0].rho, h->ck[0].v, h->ck[0].psi, h->ck[0].pinv, d_v0,
                                              h->rho_0);
    ADJ_TRY(hipGetLastError());
    State in = h->ck[0];
    for (int k = 1; k <= h->K; ++k) {
        const State out = k == h->K ? h->fin : (k % h->ckpt == 0 ? h->ck[k / h->ckpt] : h->nxt[(k - 1) % 2]);
        if (int rc = forward_iteration(h, in, out, h->inter[0], h->m1t[k - 1], h->m2[k - 1], false)) return rc;
        in = out;
    }
    frd::f_objective<<<h->grid, TPB, 0, s>>>(h->d, h->T, h->S, h->LT, h->LS, h->no_labels, h->lw, h->fin.psi,
                                             h->partials);
    frd::k_sum_final<<<1, 256, 0, s>>>(h->partials, h->grid, h->scal);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipEventRecord(h->ev[1], s));
    ADJ_TRY(hipMemcpyAsync(&h->energy, h->scal, sizeof(double), hipMemcpyDeviceToHost, s));
    ADJ_TRY(hipStreamSynchronize(s));
    ADJ_TRY(hipEventElapsedTime(&h->fwd_ms, h->ev[0], h->ev[1]));
    if (energy) *energy = h->energy;
    h->have_forward = true;
    return FR_OK;
}

extern "C" int fr_adjoint_backward(fr_adjoint *h, double *d_grad_v0, double *grad_w_reg, double *grad_tau) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (!h->have_forward) return err(FR_ERR_STATE, "fr_adjoint_backward before fr_adjoint_forward");
    hipStream_t s = h->stream;
    const size_t n = (size_t)h->n;
    ADJ_TRY(hipEventRecord(h->ev[0], s));
    ADJ_TRY(hipMemsetAsync(h->scal + 1, 0, 2 * sizeof(double), s));
    int cur = 0;
    ADJ_TRY(hipMemsetAsync(h->grho[cur], 0, n * sizeof(double), s));
    ADJ_TRY(hipMemsetAsync(h->gv, 0, 3 * n * sizeof(double), s));
    ADJ_TRY(hipMemsetAsync(h->gpinv, 0, 3 * n * sizeof(double), s));
    frd::b_objective<<<h->grid, TPB, 0, s>>>(h->d, h->T, h->S, h->LT, h->LS, h->no_labels, h->lw, h->fin.psi, h->gpsi);
    ADJ_TRY(hipGetLastError());
    for (int sg = h->nseg - 1; sg >= 0; --sg) {
        const int start = sg * h->ckpt;
        const int len = h->K - start < h->ckpt ? h->K - start : h->ckpt;
        // recompute the segment's states and intermediates from its checkpoint
        for (int j = 0; j < len; ++j) {
            const State in = j == 0 ? h->ck[sg] : h->nxt[j - 1];
            const int k = start + j + 1;
            if (int rc = forward_iteration(h, in, h->nxt[j], h->inter[j], h->m1t[k - 1], h->m2[k - 1], true))
                return rc;
        }
        for (int j = len - 1; j >= 0; --j) {
            const State in = j == 0 ? h->ck[sg] : h->nxt[j - 1];
            const int k = start + j + 1;
            if (int rc = backward_iteration(h, in, h->nxt[j], h->inter[j], h->m1t[k - 1], h->m2[k - 1], &cur))
                return rc;
        }
    }
    ADJ_TRY(hipEventRecord(h->ev[1], s));
    if (d_grad_v0) ADJ_TRY(hipMemcpyAsync(d_grad_v0, h->gv, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    double sc[2] = {0, 0};
    ADJ_TRY(hipMemcpyAsync(sc, h->scal + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    ADJ_TRY(hipStreamSynchronize(s));
    ADJ_TRY(hipEventElapsedTime(&h->bwd_ms, h->ev[0], h->ev[1]));
    if (grad_w_reg) *grad_w_reg = sc[0];
    if (grad_tau) *grad_tau = sc[1];
    return FR_OK;
}

extern "C" int fr_adjoint_get_state(fr_adjoint *h, int k, double *d_v, double *d_rho, double *d_psi, double *d_psi_inv) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (!h->have_forward) return err(FR_ERR_STATE, "fr_adjoint_get_state before fr_adjoint_forward");
    if (k == h->K) return copy_state(h, h->fin, d_v, d_rho, d_psi, d_psi_inv);
    if (k < 0 || k > h->K || k % h->ckpt != 0)
        return err(FR_ERR_INVALID, "state %d is not kept (kept: multiples of %d below %d, and %d)", k, h->ckpt, h->K,
                   h->K);
    return copy_state(h, h->ck[k / h->ckpt], d_v, d_rho, d_psi, d_psi_inv);
}

extern "C" int fr_adjoint_info(fr_adjoint *h, size_t *device_bytes, float *forward_ms, float *backward_ms) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (device_bytes) *device_bytes = h->bytes;
    if (forward_ms) *forward_ms = h->fwd_ms;
    if (backward_ms) *backward_ms = h->bwd_ms;
    return FR_OK;
}

extern "C" int fr_adjoint_set_gather_chunk(fr_adjoint *h, int zc) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (zc < 0) return err(FR_ERR_INVALID, "z-planes per gather block must be >= 0 (0 = automatic), got %d", zc);
    ADJ_TRY(hipStreamSynchronize(h->stream));
    h->gt_zc = zc == 0 ? h->gt_zc_auto : (zc > h->d.Z ? h->d.Z : zc);
    h->gt_blocks = h->gt_tiles_x * h->gt_tiles_y * ((h->d.Z + h->gt_zc - 1) / h->gt_zc);
    return FR_OK;
}

extern "C" int fr_adjoint_get_gather_chunk(const fr_adjoint *h, int *tile_x, int *tile_y, int *zc) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (tile_x) *tile_x = frd::GT_X;
    if (tile_y) *tile_y = frd::GT_Y;
    if (zc) *zc = h->gt_zc;
    return FR_OK;
}

// ---- per-operator forward/backward pairs ---------------------------------------------------------

namespace {

int op_dims(int X, int Y, int Z, Dims *d, long *n, int *grid) {
    if (X < 3 || Y < 3 || Z < 3) return err(FR_ERR_INVALID, "dims must be >= 3 along every axis (got %d %d %d)", X, Y, Z);
    if (9LL * X * Y * Z >= (1LL << 31))
        return err(FR_ERR_INVALID, "volume of %lld voxels is too large (9*N must be < 2^31)", (long long)X * Y * Z);
    *d = Dims{X, Y, Z};
    *n = (long)X * Y * Z;
    const long nb = (*n + TPB - 1) / TPB;
    *grid = (int)(nb < 8192 ? nb : 8192);
    return FR_OK;
}

}  // namespace

#define OP_PROLOGUE(...)                                                      \
    Dims d;                                                                   \
    long n;                                                                   \
    int g;                                                                    \
    if (int rc = op_dims(X, Y, Z, &d, &n, &g)) return rc;                     \
    if (!(__VA_ARGS__)) return err(FR_ERR_INVALID, "null device pointer");    \
    (void)n

extern "C" int fr_warp_fwd(int X, int Y, int Z, const double *d_f, const double *d_pos, double *d_out) {
    OP_PROLOGUE(d_f && d_pos && d_out);
    frd::op_warp_fwd<<<g, TPB>>>(d, d_f, d_pos, d_out);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipDeviceSynchronize());
    return FR_OK;
}

extern "C" int fr_warp_bwd(int X, int Y, int Z, const double *d_f, const double *d_pos, const double *d_grad_out,
                           double *d_grad_f, double *d_grad_pos) {
    OP_PROLOGUE(d_f && d_pos && d_grad_out && d_grad_f && d_grad_pos);
    frd::op_warp_bwd<<<g, TPB>>>(d, d_f, d_pos, d_grad_out, d_grad_f, d_grad_pos);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipDeviceSynchronize());
    return FR_OK;
}

extern "C" int fr_gradient_fwd(int X, int Y, int Z, const double *d_f, double *d_grad) {
    OP_PROLOGUE(d_f && d_grad);
    frd::op_grad_fwd<<<g, TPB>>>(d, d_f, d_grad);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipDeviceSynchronize());
    return FR_OK;
}

extern "C" int fr_gradient_bwd(int X, int Y, int Z, const double *d_grad_grad, double *d_grad_f) {
    OP_PROLOGUE(d_grad_grad && d_grad_f);
    frd::op_grad_bwd<<<g, TPB>>>(d, d_grad_grad, d_grad_f);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipDeviceSynchronize());
    return FR_OK;
}

extern "C" int fr_laplacian_fwd(int X, int Y, int Z, const double *d_psi, double *d_L) {
    OP_PROLOGUE(d_psi && d_L);
    frd::f_laplacian<<<g, TPB>>>(d, d_psi, d_L);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipDeviceSynchronize());
    return FR_OK;
}

extern "C" int fr_laplacian_bwd(int X, int Y, int Z, const double *d_grad_L, double *d_grad_psi) {
    OP_PROLOGUE(d_grad_L && d_grad_psi);
    frd::op_laplacian_bwd<<<g, TPB>>>(d, d_grad_L, d_grad_psi);
    ADJ_TRY(hipGetLastError());
    ADJ_TRY(hipDeviceSynchronize());
    return FR_OK;
}

namespace {

int step_check(fr_adjoint *h, int k, double w_reg, double tau) {
    if (!h) return err(FR_ERR_INVALID, "null handle");
    if (!h->have_inputs) return err(FR_ERR_STATE, "step before fr_adjoint_set_images");
    if (k < 1) return err(FR_ERR_INVALID, "iteration number must be >= 1 (got %d)", k);
    if (!std::isfinite(w_reg) || !std::isfinite(tau)) return err(FR_ERR_INVALID, "w_reg and tau must be finite");
    ADJ_TRY(hipDeviceSynchronize());  // the caller's arrays may have been written on any stream
    return FR_OK;
}

// Whether arrays a and b of wa and wb doubles per voxel (rho: 1; v, psi, psi_inv: 3) share memory.
bool overlap(const void *a, int wa, const void *b, int wb, long n) {
    const char *pa = static_cast<const char *>(a), *pb = static_cast<const char *>(b);
    return pa < pb + (size_t)wb * n * sizeof(double) && pb < pa + (size_t)wa * n * sizeof(double);
}

// whether any two of the m arrays share memory
bool any_overlap(const void *const *ptrs, const int *widths, int m, long n) {
    for (int a = 0; a < m; ++a)
        for (int b = a + 1; b < m; ++b)
            if (overlap(ptrs[a], widths[a], ptrs[b], widths[b], n)) return true;
    return false;
}

}  // namespace

extern "C" int fr_step_fwd(fr_adjoint *h, int k, double w_reg, double tau, const double *d_rho, const double *d_v,
                           const double *d_psi, const double *d_psi_inv, double *d_rho1, double *d_v1, double *d_psi1,
                           double *d_psi_inv1) {
    if (int rc = step_check(h, k, w_reg, tau)) return rc;
    if (!d_rho || !d_v || !d_psi || !d_psi_inv || !d_rho1 || !d_v1 || !d_psi1 || !d_psi_inv1)
        return err(FR_ERR_INVALID, "null device pointer");
    {
        // the kernels read neighbourhoods of the inputs while writing the outputs
        const void *p[8] = {d_rho, d_v, d_psi, d_psi_inv, d_rho1, d_v1, d_psi1, d_psi_inv1};
        const int w[8] = {1, 3, 3, 3, 1, 3, 3, 3};
        for (int o = 4; o < 8; ++o)
            for (int a = 0; a < 4; ++a)
                if (overlap(p[o], w[o], p[a], w[a], h->n))
                    return err(FR_ERR_INVALID, "fr_step_fwd: output array %d overlaps input array %d", o - 4, a);
        if (any_overlap(p + 4, w + 4, 4, h->n)) return err(FR_ERR_INVALID, "fr_step_fwd: the output arrays overlap");
    }
    h->w_reg = w_reg;
    h->tau = tau;
    h->have_forward = false;  // the work buffers are reused
    double m1t, m2;
    scalars_of(h, t_of(h, k), &m1t, &m2);
    const State in{const_cast<double *>(d_rho), const_cast<double *>(d_v), const_cast<double *>(d_psi),
                   const_cast<double *>(d_psi_inv)};
    const State out{d_rho1, d_v1, d_psi1, d_psi_inv1};
    if (int rc = forward_iteration(h, in, out, h->inter[0], m1t, m2, false)) return rc;
    ADJ_TRY(hipStreamSynchronize(h->stream));
    return FR_OK;
}

extern "C" int fr_step_bwd(fr_adjoint *h, int k, double w_reg, double tau, const double *d_rho, const double *d_v,
                           const double *d_psi, const double *d_psi_inv, double *d_g_rho, double *d_g_v,
                           double *d_g_psi, double *d_g_psi_inv, double *grad_w_reg, double *grad_tau) {
    if (int rc = step_check(h, k, w_reg, tau)) return rc;
    if (!d_rho || !d_v || !d_psi || !d_psi_inv || !d_g_rho || !d_g_v || !d_g_psi || !d_g_psi_inv)
        return err(FR_ERR_INVALID, "null device pointer");
    {
        // the adjoint arrays are read in, then written back one after another
        const void *p[4] = {d_g_rho, d_g_v, d_g_psi, d_g_psi_inv};
        const int w[4] = {1, 3, 3, 3};
        if (any_overlap(p, w, 4, h->n)) return err(FR_ERR_INVALID, "fr_step_bwd: the adjoint arrays overlap");
    }
    h->w_reg = w_reg;
    h->tau = tau;
    h->have_forward = false;  // the work buffers are reused
    double m1t, m2;
    scalars_of(h, t_of(h, k), &m1t, &m2);
    hipStream_t s = h->stream;
    const size_t n = (size_t)h->n;
    const State in{const_cast<double *>(d_rho), const_cast<double *>(d_v), const_cast<double *>(d_psi),
                   const_cast<double *>(d_psi_inv)};
    const State nx = h->nxt[0];
    if (int rc = forward_iteration(h, in, nx, h->inter[0], m1t, m2, true)) return rc;
    int cur = 0;
    ADJ_TRY(hipMemcpyAsync(h->grho[cur], d_g_rho, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemcpyAsync(h->gv, d_g_v, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemcpyAsync(h->gpsi, d_g_psi, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemcpyAsync(h->gpinv, d_g_psi_inv, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemsetAsync(h->scal + 1, 0, 2 * sizeof(double), s));
    if (int rc = backward_iteration(h, in, nx, h->inter[0], m1t, m2, &cur)) return rc;
    ADJ_TRY(hipMemcpyAsync(d_g_rho, h->grho[cur], n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemcpyAsync(d_g_v, h->gv, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemcpyAsync(d_g_psi, h->gpsi, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    ADJ_TRY(hipMemcpyAsync(d_g_psi_inv, h->gpinv, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    double sc[2] = {0, 0};
    ADJ_TRY(hipMemcpyAsync(sc, h->scal + 1, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    ADJ_TRY(hipStreamSynchronize(s));
    if (grad_w_reg) *grad_w_reg = sc[0];
    if (grad_tau) *grad_tau = sc[1];
    return FR_OK;
}
