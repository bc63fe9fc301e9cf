// HIP kernels of the fastreg registration solver for gfx950 (MI355X).
//
// One solver iteration is two launches:
//   pass A  (per interior voxel, reads the old state, writes the new one into the other buffer of
//           a ping-pong pair): upwind divergence and density update, J_inv = D psi_inv and the
//           reference's det(), L = -Laplacian(psi) evaluated at the 8 corners of psi_inv(p) and
//           interpolated there, T o psi_inv, grad S, the potential gradient, (Dv)v, the upwind
//           velocity Laplacian, the velocity update and the psi_inv update.
//   pass B  (per interior voxel, after all of v_new exists): psi += alpha * v_new o psi, and the
//           max update norm of the iteration, reduced in the wave, the block and one 64-bit
//           atomic max per block into a per-iteration slot.
// The reference runs the same iteration as 13 launches on 5 streams with 12 float4 fields and
// two 64-byte Jacobians in HBM (src/fastreg/cuda/solver.cu:295-444); here the only iterated
// state is rho, v, psi, psi_inv (xyz-packed, 12 B per vector) plus the read-only S and T.
//
// Arithmetic follows the reference kernel by kernel in the same operation order.  The library
// is compiled with -ffp-contract=off, so a multiply-add rounds twice unless written fmaf():
// the reference's explicit fma (lerp) and `rho += alpha * div` use fmaf, everything else is
// separately rounded, as the reference's __f*_rn float4 operators are.  Division and sqrt are
// IEEE (correctly rounded); the max-norm uses sqrt rounded toward zero.

#include "fr_internal.hpp"

#include <math.h>

namespace frk {

struct __attribute__((aligned(4))) F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 ld3(const float *__restrict__ p, int i) {
    return *reinterpret_cast<const F3 *>(p + 3 * i);
}
__device__ __forceinline__ void st3(float *__restrict__ p, int i, F3 v) { *reinterpret_cast<F3 *>(p + 3 * i) = v; }

__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add3(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul3(float m, F3 a) { return f3(a.x * m, a.y * m, a.z * m); }

// lerp(v0, v1, t) = fma(t, v0, fma(-t, v1, v1)) — cuda_utils.hpp:43-46
__device__ __forceinline__ float lerp(float v0, float v1, float t) { return fmaf(t, v0, fmaf(-t, v1, v1)); }

// nesting of cuda_utils.hpp:75-80: lerp over a of (lerp over b of (lerp over c)); fXYZ with
// 1 = the +1 neighbour along that axis
__device__ __forceinline__ float tri(float f111, float f110, float f101, float f100, float f011, float f010,
                                     float f001, float f000, float a, float b, float c) {
    return lerp(lerp(lerp(f111, f110, c), lerp(f101, f100, c), b), lerp(lerp(f011, f010, c), lerp(f001, f000, c), b),
                a);
}

// Clamp into [0, dim-1], floor, weights; the +1 neighbour index is clamped into the grid (its
// weight is exactly 0 whenever that clamp changes it) — cuda_utils.hpp:60-136.
struct Corner {
    int gx, gy, gz, hx, hy, hz;
    float a, b, c;
};

__device__ __forceinline__ Corner corner(const Dims &d, float px, float py, float pz) {
    Corner k;
    float cx = fminf(fmaxf(0.f, px), (float)(d.X - 1));
    float cy = fminf(fmaxf(0.f, py), (float)(d.Y - 1));
    float cz = fminf(fmaxf(0.f, pz), (float)(d.Z - 1));
    k.gx = (int)floorf(cx);
    k.gy = (int)floorf(cy);
    k.gz = (int)floorf(cz);
    k.a = cx - (float)k.gx;
    k.b = cy - (float)k.gy;
    k.c = cz - (float)k.gz;
    k.hx = min(k.gx + 1, d.X - 1);
    k.hy = min(k.gy + 1, d.Y - 1);
    k.hz = min(k.gz + 1, d.Z - 1);
    return k;
}

__device__ __forceinline__ int lin(const Dims &d, int x, int y, int z) { return x + d.X * (y + d.Y * z); }

// sqrt rounded toward zero, from the correctly rounded sqrt and the exact FMA residual
__device__ __forceinline__ float sqrt_rd(float a) {
    float s = sqrtf(a);
    if (s > 0.f && fmaf(s, s, -a) > 0.f) s = __uint_as_float(__float_as_uint(s) - 1u);
    return s;
}

// det(Mat4f) exactly as parenthesised in cuda_utils.hpp:389-396 (not the determinant);
// m_rc = row r (component), column c (d/dx, d/dy, d/dz)
__device__ __forceinline__ float det_ref(float m00, float m01, float m02, float m10, float m11, float m12, float m20,
                                         float m21, float m22) {
    float a = ((m00 * m11) * m22 + (m10 * m21) * m02) + (m20 * m01) * m12;
    float b = ((m02 * m11) + m20) * ((m12 * m21) * m00) + (m22 * m01) * m10;
    return a - b;
}

__device__ __forceinline__ float det_true(float m00, float m01, float m02, float m10, float m11, float m12, float m20,
                                          float m21, float m22) {
    return m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
}

// -Laplacian(psi) at q (vector_fields.cu:275-300); 0 off the interior, where the reference's
// L buffer is never written
__device__ __forceinline__ F3 lap_at(const Dims &d, const float *__restrict__ psi, int qx, int qy, int qz) {
    if (qx < 1 || qx > d.X - 2 || qy < 1 || qy > d.Y - 2 || qz < 1 || qz > d.Z - 2) return f3(0.f, 0.f, 0.f);
    int q = lin(d, qx, qy, qz);
    int sy = d.X, sz = d.X * d.Y;
    F3 s = mul3(-6.f, ld3(psi, q));
    s = add3(s, ld3(psi, q + 1));
    s = add3(s, ld3(psi, q - 1));
    s = add3(s, ld3(psi, q + sy));
    s = add3(s, ld3(psi, q - sy));
    s = add3(s, ld3(psi, q + sz));
    s = add3(s, ld3(psi, q - sz));
    return mul3(-1.f, s);
}

// XCD-aware tile order: the hardware hands consecutive workgroups to the 8 XCDs round-robin;
// remap so that each XCD works on one contiguous run of tiles (a z-slab), keeping stencil and
// gather neighbours in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    int per = nb >> 3, rem = nb & 7, x = b & 7, s = b >> 3;
    return (x < rem) ? x * (per + 1) + s : rem * (per + 1) + (x - rem) * per + s;
}

// stop rules checked at the start of each launch of iteration a.iter (uniform per block):
// converged at iter-1 (solver.cu:418-421) or NaN energy at iter (solver.cu:366-368)
__device__ __forceinline__ bool should_stop(const IterArgs &a, bool record) {
    if (__builtin_nontemporal_load(&a.ctl->stopped)) return true;
    if (a.iter > 1) {
        unsigned long long mp = a.maxpack[a.iter - 2];
        float m = __uint_as_float((unsigned)(mp >> 32));
        if (m <= a.thr) {
            if (record) {
                a.ctl->stop_iter = a.iter - 1;
                a.ctl->stop_reason = 1;
                a.ctl->stopped = 1;
            }
            return true;
        }
    }
    if (a.check_energy && isnan(a.energy[a.iter - 1])) {
        if (record) {
            a.ctl->stop_iter = a.iter - 1;
            a.ctl->stop_reason = 2;
            a.ctl->stopped = 1;
        }
        return true;
    }
    return false;
}

// --------------------------------------------------------------------------------------------
// pass A
// --------------------------------------------------------------------------------------------

__global__ __launch_bounds__(BLOCK) void k_pass_a(IterArgs a) {
    const int tid = threadIdx.x;
    const int b = xcd_remap(blockIdx.x, a.nblocks);
    if (should_stop(a, b == 0 && tid == 0)) return;
    const Dims d = a.d;
    const int bx = b % a.nbx, r = b / a.nbx, by = r % a.nby, z = r / a.nby;
    const int x = bx * TILE_X + (tid & (TILE_X - 1));
    const int y = by * TILE_Y + (tid / TILE_X);
    if (x < 1 || x > d.X - 2 || y < 1 || y > d.Y - 2 || z < 1 || z > d.Z - 2) return;

    const int sy = d.X, sz = d.X * d.Y;
    const int i = lin(d, x, y, z);
    const float *__restrict__ rho = a.rho_o;
    const float *__restrict__ v = a.v_o;

    // 1. density: upwind divergence of rho v, then rho += alpha * div (solver.cu:101-173)
    const F3 vc = ld3(v, i);
    const float rc = rho[i];
    const float t11 = (vc.x > 0.f) ? -vc.x * rc : -vc.x * rho[i + 1];
    const float vxm = v[3 * (i - 1) + 0];
    const float t12 = (vxm > 0.f) ? vxm * rho[i - 1] : vxm * rc;
    const float t21 = (vc.y > 0.f) ? -vc.y * rc : -vc.y * rho[i + sy];
    const float vym = v[3 * (i - sy) + 1];
    const float t22 = (vym > 0.f) ? vym * rho[i - sy] : vym * rc;
    const float t31 = (vc.z > 0.f) ? -vc.z * rc : -vc.z * rho[i + sz];
    const float vzm = v[3 * (i - sz) + 2];
    const float t32 = (vzm > 0.f) ? vzm * rho[i - sz] : vzm * rc;
    const float div = ((((t11 + t12) + t21) + t22) + t31) + t32;
    const float rho_n = fmaf(a.alpha, div, rc);

    // 2. J_inv = D psi_inv by central differences, and det() (vector_fields.cu:503-535)
    const float *__restrict__ pinv = a.pinv_o;
    const F3 jx = mul3(0.5f, sub3(ld3(pinv, i + 1), ld3(pinv, i - 1)));
    const F3 jy = mul3(0.5f, sub3(ld3(pinv, i + sy), ld3(pinv, i - sy)));
    const F3 jz = mul3(0.5f, sub3(ld3(pinv, i + sz), ld3(pinv, i - sz)));
    const float det = det_ref(jx.x, jy.x, jz.x, jx.y, jy.y, jz.y, jx.z, jy.z, jz.z);

    // 3-4. L = -Laplacian(psi) interpolated at psi_inv(p), and T o psi_inv at the same point
    // (vector_fields.cu:231-300, 81-99)
    const F3 pc = ld3(pinv, i);
    const Corner k = corner(d, pc.x, pc.y, pc.z);
    const float *__restrict__ psi = a.psi;
    const F3 l111 = lap_at(d, psi, k.hx, k.hy, k.hz), l110 = lap_at(d, psi, k.hx, k.hy, k.gz);
    const F3 l101 = lap_at(d, psi, k.hx, k.gy, k.hz), l100 = lap_at(d, psi, k.hx, k.gy, k.gz);
    const F3 l011 = lap_at(d, psi, k.gx, k.hy, k.hz), l010 = lap_at(d, psi, k.gx, k.hy, k.gz);
    const F3 l001 = lap_at(d, psi, k.gx, k.gy, k.hz), l000 = lap_at(d, psi, k.gx, k.gy, k.gz);
    const F3 Lo = f3(tri(l111.x, l110.x, l101.x, l100.x, l011.x, l010.x, l001.x, l000.x, k.a, k.b, k.c),
                     tri(l111.y, l110.y, l101.y, l100.y, l011.y, l010.y, l001.y, l000.y, k.a, k.b, k.c),
                     tri(l111.z, l110.z, l101.z, l100.z, l011.z, l010.z, l001.z, l000.z, k.a, k.b, k.c));
    const float *__restrict__ T = a.T;
    const float Tp = tri(T[lin(d, k.hx, k.hy, k.hz)], T[lin(d, k.hx, k.hy, k.gz)], T[lin(d, k.hx, k.gy, k.hz)],
                         T[lin(d, k.hx, k.gy, k.gz)], T[lin(d, k.gx, k.hy, k.hz)], T[lin(d, k.gx, k.hy, k.gz)],
                         T[lin(d, k.gx, k.gy, k.hz)], T[lin(d, k.gx, k.gy, k.gz)], k.a, k.b, k.c);

    // 5. grad S (vector_fields.cu:155-192) and the potential gradient (solver.cu:56-80)
    const float *__restrict__ S = a.S;
    const F3 gS = f3((S[i + 1] - S[i - 1]) * 0.5f, (S[i + sy] - S[i - sy]) * 0.5f, (S[i + sz] - S[i - sz]) * 0.5f);
    const float diff = S[i] - Tp;
    const F3 nU = mul3(det, add3(mul3(diff, gS), mul3(a.w_reg, Lo)));

    // 6. (Dv)v: entropy scheme on the diagonal, upwind off it (vector_fields.cu:352-404)
    const F3 xp = ld3(v, i + 1), xm = ld3(v, i - 1);
    const F3 yp = ld3(v, i + sy), ym = ld3(v, i - sy);
    const F3 zp = ld3(v, i + sz), zm = ld3(v, i - sz);
#define MX(q) fmaxf((q), 0.f)
#define MN(q) fminf((q), 0.f)
    const float ax1 = ((MX(vc.x) * MX(vc.x) - MN(vc.x) * MN(vc.x)) + MN(xp.x) * MN(xp.x)) - MX(xp.x) * MX(xp.x);
    const float ax2 = (vc.y > 0.f) ? vc.y * (vc.x - ym.x) : vc.y * (yp.x - vc.x);
    const float ax3 = (vc.z > 0.f) ? vc.z * (vc.x - zm.x) : vc.z * (zp.x - vc.x);
    const float ay1 = (vc.x > 0.f) ? vc.x * (vc.y - xm.y) : vc.x * (xp.y - vc.y);
    const float ay2 = ((MX(vc.y) * MX(vc.y) - MN(vc.y) * MN(vc.y)) + MN(yp.y) * MN(yp.y)) - MX(yp.y) * MX(yp.y);
    const float ay3 = (vc.z > 0.f) ? vc.z * (vc.y - zm.y) : vc.z * (zp.y - vc.y);
    const float az1 = (vc.x > 0.f) ? vc.x * (vc.z - xm.z) : vc.x * (xp.z - vc.z);
    const float az2 = (vc.y > 0.f) ? vc.y * (vc.z - ym.z) : vc.y * (yp.z - vc.z);
    const float az3 = ((MX(vc.z) * MX(vc.z) - MN(vc.z) * MN(vc.z)) + MN(zp.z) * MN(zp.z)) - MX(zp.z) * MX(zp.z);
#undef MX
#undef MN
    const F3 dvv = f3((0.5f * ax1 + ax2) + ax3, (ay1 + 0.5f * ay2) + ay3, (az1 + az2) + 0.5f * az3);

    // 7. upwind velocity Laplacian with the reference's indices and ranges: x in [2,X-3],
    // y in [2,Y-2], z in [2,Z-3]; (y+2) at y = Y-2 is the linear address idx(x, 0, z+1)
    // (vector_fields.cu:307-346); 0 elsewhere
    F3 dlt = f3(0.f, 0.f, 0.f);
    if (x >= 2 && x <= d.X - 3 && y >= 2 && y <= d.Y - 2 && z >= 2 && z <= d.Z - 3) {
        const F3 vxx = (vc.x < 0.f) ? add3(sub3(ld3(v, i + 2), mul3(2.f, xp)), vc)
                                    : add3(sub3(vc, mul3(2.f, xm)), ld3(v, i - 2));
        const F3 vyy = (vc.y < 0.f) ? add3(sub3(ld3(v, i + 2 * sy), mul3(2.f, ym)), vc)
                                    : add3(sub3(vc, mul3(2.f, ym)), ld3(v, i - 2 * sy));
        const F3 vzz = (vc.z < 0.f) ? add3(sub3(ld3(v, i + 2 * sz), mul3(2.f, zm)), vc)
                                    : add3(sub3(vc, mul3(2.f, zm)), ld3(v, i - 2 * sz));
        dlt = mul3(-1.f, add3(add3(vxx, vyy), vzz));
    }

    // 8. velocity update with this iteration's rho (solver.cu:188-211)
    const F3 fr = mul3(a.m1t, vc);
    const F3 df = mul3(a.tau, dlt);
    const float q = a.m2 / (fabsf(rho_n) + 1e-5f);
    const F3 po = mul3(q, nU);
    const F3 up = mul3(a.alpha, add3(sub3(add3(fr, dvv), df), po));
    const F3 vn = f3(vc.x - up.x, vc.y - up.y, vc.z - up.z);

    // 10. psi_inv -= alpha * (J_inv v_new) (solver.cu:263-279, cuda_utils.hpp:379-385)
    const F3 Jv = f3((jx.x * vn.x + jy.x * vn.y) + jz.x * vn.z, (jx.y * vn.x + jy.y * vn.y) + jz.y * vn.z,
                     (jx.z * vn.x + jy.z * vn.y) + jz.z * vn.z);
    const F3 ui = mul3(a.alpha, Jv);
    const F3 pn = f3(pc.x - ui.x, pc.y - ui.y, pc.z - ui.z);

    a.rho_n[i] = rho_n;
    st3(a.v_n, i, vn);
    st3(a.pinv_n, i, pn);
}

// --------------------------------------------------------------------------------------------
// pass B
// --------------------------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    lo = __shfl_xor(lo, m, 64);
    hi = __shfl_xor(hi, m, 64);
    return ((unsigned long long)hi << 32) | lo;
}

__global__ __launch_bounds__(BLOCK) void k_pass_b(IterArgs a) {
    const int tid = threadIdx.x;
    const int b = xcd_remap(blockIdx.x, a.nblocks);
    if (__builtin_nontemporal_load(&a.ctl->stopped)) return;
    const Dims d = a.d;
    const int bx = b % a.nbx, r = b / a.nbx, by = r % a.nby, z = r / a.nby;
    const int x = bx * TILE_X + (tid & (TILE_X - 1));
    const int y = by * TILE_Y + (tid / TILE_X);
    unsigned long long pack = 0;
    if (x >= 1 && x <= d.X - 2 && y >= 1 && y <= d.Y - 2 && z >= 1 && z <= d.Z - 2) {
        // psi += alpha * v_new o psi; updates = that increment (solver.cu:230-247)
        const int i = lin(d, x, y, z);
        const F3 p = ld3(a.psi, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        const float *__restrict__ v = a.v_n;
        const F3 v111 = ld3(v, lin(d, k.hx, k.hy, k.hz)), v110 = ld3(v, lin(d, k.hx, k.hy, k.gz));
        const F3 v101 = ld3(v, lin(d, k.hx, k.gy, k.hz)), v100 = ld3(v, lin(d, k.hx, k.gy, k.gz));
        const F3 v011 = ld3(v, lin(d, k.gx, k.hy, k.hz)), v010 = ld3(v, lin(d, k.gx, k.hy, k.gz));
        const F3 v001 = ld3(v, lin(d, k.gx, k.gy, k.hz)), v000 = ld3(v, lin(d, k.gx, k.gy, k.gz));
        const F3 w = f3(tri(v111.x, v110.x, v101.x, v100.x, v011.x, v010.x, v001.x, v000.x, k.a, k.b, k.c),
                        tri(v111.y, v110.y, v101.y, v100.y, v011.y, v010.y, v001.y, v000.y, k.a, k.b, k.c),
                        tri(v111.z, v110.z, v101.z, v100.z, v011.z, v010.z, v001.z, v000.z, k.a, k.b, k.c));
        const F3 u = mul3(a.alpha, w);
        st3(a.psi, i, f3(p.x + u.x, p.y + u.y, p.z + u.z));
        // norm with a round-down sqrt (cuda_utils.hpp:251-253); NaN never wins (strict >,
        // reductor.cu:445-456)
        const float nr = sqrt_rd((u.x * u.x + u.y * u.y) + u.z * u.z);
        if (!isnan(nr)) pack = ((unsigned long long)__float_as_uint(nr) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)i);
    }
    // max (norm, lowest index) over the block, then one atomic per block into the iteration slot
    for (int m = 32; m >= 1; m >>= 1) {
        unsigned long long o = shfl_xor_u64(pack, m);
        pack = o > pack ? o : pack;
    }
    __shared__ unsigned long long red[BLOCK / 64];
    if ((tid & 63) == 0) red[tid >> 6] = pack;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < BLOCK / 64; ++w) pack = red[w] > pack ? red[w] : pack;
        unsigned long long *slot = &a.maxpack[a.iter - 1];
        if (pack > __builtin_nontemporal_load(slot)) atomicMax(slot, pack);
    }
}

// --------------------------------------------------------------------------------------------
// init, energy, warp, Dice, min det J, layout conversion
// --------------------------------------------------------------------------------------------

__global__ void k_init(Dims d, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab,
                       float *T, int16_t *LT, float *S, int16_t *LS, float *rho0, float *rho1, float *v0, float *v1,
                       float *pinv0, float *pinv1, float *psi, float rho_0) {
    int n = d.X * d.Y * d.Z;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T[i] = tgt[i];
        LT[i] = tgt_lab[i];
        S[i] = src[i];
        LS[i] = src_lab[i];
        rho0[i] = rho_0;
        rho1[i] = rho_0;
        int x = i % d.X, r = i / d.X, y = r % d.Y, z = r / d.Y;
        F3 id = f3((float)x, (float)y, (float)z);
        st3(v0, i, f3(0.f, 0.f, 0.f));
        st3(v1, i, f3(0.f, 0.f, 0.f));
        st3(pinv0, i, id);
        st3(pinv1, i, id);
        st3(psi, i, id);
    }
}

constexpr int RED_BLOCK = 256;

__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    int tid = threadIdx.x;
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) s += sh[w];
    __syncthreads();
    return s;
}

// energy terms per voxel (all voxels): (T - S o psi)^2 and ||D(psi - id)||_F^2 with mirrored
// neighbours at the faces (reductor.cu:11-112, 224-324; vector_fields.cu:537-585)
__global__ __launch_bounds__(RED_BLOCK) void k_energy(Dims d, const float *__restrict__ S,
                                                      const float *__restrict__ T, const float *__restrict__ psi,
                                                      double *partials) {
    __shared__ double sh[RED_BLOCK / 64];
    int n = d.X * d.Y * d.Z;
    double ea = 0.0, eb = 0.0;
    for (int i = blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += gridDim.x * RED_BLOCK) {
        int x = i % d.X, r = i / d.X, y = r % d.Y, z = r / d.Y;
        F3 p = ld3(psi, i);
        Corner k = corner(d, p.x, p.y, p.z);
        float sp = tri(S[lin(d, k.hx, k.hy, k.hz)], S[lin(d, k.hx, k.hy, k.gz)], S[lin(d, k.hx, k.gy, k.hz)],
                       S[lin(d, k.hx, k.gy, k.gz)], S[lin(d, k.gx, k.hy, k.hz)], S[lin(d, k.gx, k.hy, k.gz)],
                       S[lin(d, k.gx, k.gy, k.hz)], S[lin(d, k.gx, k.gy, k.gz)], k.a, k.b, k.c);
        double df = (double)T[i] - (double)sp;
        ea += df * df;
        int x1 = x + 1, x2 = x - 1, y1 = y + 1, y2 = y - 1, z1 = z + 1, z2 = z - 1;
        if (x == 0) x2 = x + 1; else if (x == d.X - 1) x1 = x - 1;
        if (y == 0) y2 = y + 1; else if (y == d.Y - 1) y1 = y - 1;
        if (z == 0) z2 = z + 1; else if (z == d.Z - 1) z1 = z - 1;
#define DISP(xx, yy, zz) sub3(ld3(psi, lin(d, xx, yy, zz)), f3((float)(xx), (float)(yy), (float)(zz)))
        F3 jx = mul3(0.5f, sub3(DISP(x1, y, z), DISP(x2, y, z)));
        F3 jy = mul3(0.5f, sub3(DISP(x, y1, z), DISP(x, y2, z)));
        F3 jz = mul3(0.5f, sub3(DISP(x, y, z1), DISP(x, y, z2)));
#undef DISP
        // rows of J are the components: (jx.c, jy.c, jz.c)
        eb += ((double)jx.x * jx.x + (double)jy.x * jy.x) + (double)jz.x * jz.x;
        eb += ((double)jx.y * jx.y + (double)jy.y * jy.y) + (double)jz.y * jz.y;
        eb += ((double)jx.z * jx.z + (double)jy.z * jy.z) + (double)jz.z * jz.z;
    }
    double sa = block_sum(ea, sh);
    double sb = block_sum(eb, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sa;
        partials[2 * blockIdx.x + 1] = sb;
    }
}

__global__ __launch_bounds__(RED_BLOCK) void k_energy_finalize(const double *partials, int nparts, double w_reg,
                                                               double *e_data, double *e_reg, double *energy) {
    __shared__ double sh[RED_BLOCK / 64];
    double a = 0.0, b = 0.0;
    for (int j = threadIdx.x; j < nparts; j += RED_BLOCK) {
        a += partials[2 * j];
        b += partials[2 * j + 1];
    }
    double sa = block_sum(a, sh);
    double sb = block_sum(b, sh);
    if (threadIdx.x == 0) {
        double ed = 0.5 * sa, er = 0.5 * sb;
        *e_data = ed;
        *e_reg = er;
        *energy = ed + w_reg * er;
    }
}

// apply (vector_fields.cu:81-99): trilinear intensity and floor-corner label at def(p)
__global__ void k_warp(Dims d, const float *__restrict__ img, const int16_t *__restrict__ lab,
                       const float *__restrict__ def, int stride, float *__restrict__ out_img,
                       int16_t *__restrict__ out_lab) {
    int n = d.X * d.Y * d.Z;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *q = def + (long)stride * i;
        Corner k = corner(d, q[0], q[1], q[2]);
        if (out_img)
            out_img[i] = tri(img[lin(d, k.hx, k.hy, k.hz)], img[lin(d, k.hx, k.hy, k.gz)], img[lin(d, k.hx, k.gy, k.hz)],
                             img[lin(d, k.hx, k.gy, k.gz)], img[lin(d, k.gx, k.hy, k.hz)], img[lin(d, k.gx, k.hy, k.gz)],
                             img[lin(d, k.gx, k.gy, k.hz)], img[lin(d, k.gx, k.gy, k.gz)], k.a, k.b, k.c);
        if (out_lab && lab) out_lab[i] = lab[lin(d, k.gx, k.gy, k.gz)];
    }
}

// Dice counts over [0,X-2]x[0,Y-2]x[0,Z-2] (solver.cu:21-35); labels outside [0, no_labels)
// go to *bad instead of indexing out of bounds
constexpr int DICE_LDS_LABELS = 2048;

__global__ void k_dice(Dims d, const int16_t *__restrict__ la, const int16_t *__restrict__ lb, int no_labels,
                       unsigned *overlap, unsigned *voxels, unsigned *bad) {
    __shared__ unsigned s_ov[DICE_LDS_LABELS], s_vx[DICE_LDS_LABELS];
    const bool use_lds = no_labels <= DICE_LDS_LABELS;
    if (use_lds)
        for (int l = threadIdx.x; l < no_labels; l += blockDim.x) s_ov[l] = s_vx[l] = 0u;
    __syncthreads();
    int n = d.X * d.Y * d.Z;
    unsigned nb = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int x = i % d.X, r = i / d.X, y = r % d.Y, z = r / d.Y;
        if (x >= d.X - 1 || y >= d.Y - 1 || z >= d.Z - 1) continue;
        int l1 = la[i], l2 = lb[i];
        if (l1 < 0 || l1 >= no_labels || l2 < 0 || l2 >= no_labels) {
            nb++;
            continue;
        }
        if (use_lds) {
            atomicAdd(&s_vx[l1], 1u);
            atomicAdd(&s_vx[l2], 1u);
            if (l1 == l2) atomicAdd(&s_ov[l1], 1u);
        } else {
            atomicAdd(&voxels[l1], 1u);
            atomicAdd(&voxels[l2], 1u);
            if (l1 == l2) atomicAdd(&overlap[l1], 1u);
        }
    }
    if (nb) atomicAdd(bad, nb);
    __syncthreads();
    if (use_lds)
        for (int l = threadIdx.x; l < no_labels; l += blockDim.x) {
            if (s_vx[l]) atomicAdd(&voxels[l], s_vx[l]);
            if (s_ov[l]) atomicAdd(&overlap[l], s_ov[l]);
        }
}

// min over the interior of det and det_ref of the central-difference Jacobian of psi
__global__ __launch_bounds__(RED_BLOCK) void k_min_det(Dims d, const float *__restrict__ psi, float *part_true,
                                                       float *part_ref) {
    __shared__ float sh[2][RED_BLOCK / 64];
    int n = d.X * d.Y * d.Z;
    float mt = INFINITY, mr = INFINITY;
    for (int i = blockIdx.x * RED_BLOCK + threadIdx.x; i < n; i += gridDim.x * RED_BLOCK) {
        int x = i % d.X, r = i / d.X, y = r % d.Y, z = r / d.Y;
        if (x < 1 || x > d.X - 2 || y < 1 || y > d.Y - 2 || z < 1 || z > d.Z - 2) continue;
        int sy = d.X, sz = d.X * d.Y;
        F3 jx = mul3(0.5f, sub3(ld3(psi, i + 1), ld3(psi, i - 1)));
        F3 jy = mul3(0.5f, sub3(ld3(psi, i + sy), ld3(psi, i - sy)));
        F3 jz = mul3(0.5f, sub3(ld3(psi, i + sz), ld3(psi, i - sz)));
        float t = det_true(jx.x, jy.x, jz.x, jx.y, jy.y, jz.y, jx.z, jy.z, jz.z);
        float f = det_ref(jx.x, jy.x, jz.x, jx.y, jy.y, jz.y, jx.z, jy.z, jz.z);
        mt = fminf(mt, t);
        mr = fminf(mr, f);
    }
    for (int m = 32; m >= 1; m >>= 1) {
        mt = fminf(mt, __shfl_xor(mt, m, 64));
        mr = fminf(mr, __shfl_xor(mr, m, 64));
    }
    int tid = threadIdx.x;
    if ((tid & 63) == 0) {
        sh[0][tid >> 6] = mt;
        sh[1][tid >> 6] = mr;
    }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < RED_BLOCK / 64; ++w) {
            mt = fminf(mt, sh[0][w]);
            mr = fminf(mr, sh[1][w]);
        }
        part_true[blockIdx.x] = mt;
        part_ref[blockIdx.x] = mr;
    }
}

__global__ void k_pack4to3(const float *__restrict__ src4, float *dst3a, float *dst3b, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        F3 v = f3(src4[4 * i], src4[4 * i + 1], src4[4 * i + 2]);
        st3(dst3a, i, v);
        if (dst3b) st3(dst3b, i, v);
    }
}

__global__ void k_unpack3to4(const float *__restrict__ src3, float *__restrict__ dst4, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        F3 v = ld3(src3, i);
        dst4[4 * i] = v.x;
        dst4[4 * i + 1] = v.y;
        dst4[4 * i + 2] = v.z;
        dst4[4 * i + 3] = 0.f;
    }
}

// --------------------------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------------------------

static int grid_for(int n, int block, int cap = 65536) {
    int g = (n + block - 1) / block;
    return g < 1 ? 1 : (g > cap ? cap : g);
}

hipError_t launch_init(Dims d, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab,
                       float *T, int16_t *LT, float *S, int16_t *LS, float *rho0, float *rho1, float *v0, float *v1,
                       float *pinv0, float *pinv1, float *psi, float rho_0, hipStream_t s) {
    int n = d.X * d.Y * d.Z;
    hipLaunchKernelGGL(k_init, dim3(grid_for(n, 256)), dim3(256), 0, s, d, tgt, tgt_lab, src, src_lab, T, LT, S, LS,
                       rho0, rho1, v0, v1, pinv0, pinv1, psi, rho_0);
    return hipGetLastError();
}

hipError_t launch_pass_a(const IterArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_pass_a, dim3(a.nblocks), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_pass_b(const IterArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_pass_b, dim3(a.nblocks), dim3(BLOCK), 0, s, a);
    return hipGetLastError();
}

int energy_partials_count(Dims d) { return grid_for(d.X * d.Y * d.Z, RED_BLOCK, 4096); }

hipError_t launch_energy(Dims d, const float *S, const float *T, const float *psi, double *partials, int nparts,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_energy, dim3(nparts), dim3(RED_BLOCK), 0, s, d, S, T, psi, partials);
    return hipGetLastError();
}

hipError_t launch_energy_finalize(const double *partials, int nparts, double w_reg, double *e_data, double *e_reg,
                                  double *energy, hipStream_t s) {
    hipLaunchKernelGGL(k_energy_finalize, dim3(1), dim3(RED_BLOCK), 0, s, partials, nparts, w_reg, e_data, e_reg,
                       energy);
    return hipGetLastError();
}

hipError_t launch_warp(Dims d, const float *img, const int16_t *lab, const float *def, int stride, float *out_img,
                       int16_t *out_lab, hipStream_t s) {
    int n = d.X * d.Y * d.Z;
    hipLaunchKernelGGL(k_warp, dim3(grid_for(n, 256)), dim3(256), 0, s, d, img, lab, def, stride, out_img, out_lab);
    return hipGetLastError();
}

hipError_t launch_dice(Dims d, const int16_t *la, const int16_t *lb, int no_labels, unsigned *overlap, unsigned *voxels,
                       unsigned *bad, hipStream_t s) {
    int n = d.X * d.Y * d.Z;
    hipLaunchKernelGGL(k_dice, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d, la, lb, no_labels, overlap, voxels,
                       bad);
    return hipGetLastError();
}

int min_det_partials_count(Dims d) { return grid_for(d.X * d.Y * d.Z, RED_BLOCK, 4096); }

hipError_t launch_min_det(Dims d, const float *psi, float *part_true, float *part_ref, hipStream_t s) {
    hipLaunchKernelGGL(k_min_det, dim3(min_det_partials_count(d)), dim3(RED_BLOCK), 0, s, d, psi, part_true, part_ref);
    return hipGetLastError();
}

hipError_t launch_pack4to3(const float *src4, float *dst3a, float *dst3b, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_pack4to3, dim3(grid_for(n, 256)), dim3(256), 0, s, src4, dst3a, dst3b, n);
    return hipGetLastError();
}

hipError_t launch_unpack3to4(const float *src3, float *dst4, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_unpack3to4, dim3(grid_for(n, 256)), dim3(256), 0, s, src3, dst4, n);
    return hipGetLastError();
}

}  // namespace frk
