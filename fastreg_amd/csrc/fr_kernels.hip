// HIP kernels of the fastreg registration solver for gfx950 (MI355X).
//
// One solver iteration is two launches over tiles of 64-lane rows, each thread marching a chunk
// of z-slices with its column's z-neighbours in registers:
//   pass A  (reads the old state, writes the new one into the other buffer of a ping-pong pair):
//           L = -Laplacian(psi) and T on the block's neighbourhood, staged in LDS, interpolated at
//           psi_inv(p); upwind divergence and density update, J_inv = D psi_inv and the
//           reference's det(), grad S, the potential gradient, (Dv)v, the upwind velocity
//           Laplacian, the velocity update and the psi_inv update.
//   pass B  (after all of v_new exists): psi += alpha * v_new o psi, and the max update norm of the
//           iteration, reduced in the wave, the block and one 64-bit atomic max per block into
//           one of NSLOT per-iteration slots.
// The reference runs the same iteration as 13 launches on 5 streams with 12 float4 fields and
// two 64-byte Jacobians in HBM (src/fastreg/cuda/solver.cu:295-444); here the iterated state is
// rho, v, psi, psi_inv (xyz-packed, 12 B per vector), plus the read-only S and T.
//
// Arithmetic follows the reference kernel by kernel in the same operation order.  The library
// is compiled with -ffp-contract=off, so a multiply-add rounds twice unless written fmaf():
// the reference's explicit fma (lerp) and `rho += alpha * div` use fmaf, everything else is
// separately rounded, as the reference's __f*_rn float4 operators are.  Division and sqrt are
// IEEE (correctly rounded); the max-norm uses sqrt rounded toward zero.

#include "fr_internal.hpp"

#include <climits>
#include <cstddef>
#include <type_traits>
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
// a - 2b with one rounding: 2b is exact, so this equals the separately rounded a - (2 * b) bit for
// bit (zero signs included) unless 2b overflows (|b| >= 2^127)
__device__ __forceinline__ F3 sub_twice(F3 a, F3 b) {
    return f3(fmaf(-2.f, b.x, a.x), fmaf(-2.f, b.y, a.y), fmaf(-2.f, b.z, a.z));
}

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

// The clamp fminf(fmaxf(0, p), dim - 1) is one v_med3_f32: the median of (p, 0, dim - 1) is the
// clamped value for every number, and for a NaN p med3 returns the minimum of the other two, 0,
// which is what fmaxf(0, NaN) gives.
__device__ __forceinline__ Corner corner(Dims d, float px, float py, float pz) {
    Corner k;
    float cx = __builtin_amdgcn_fmed3f(px, 0.f, (float)(d.X - 1));
    float cy = __builtin_amdgcn_fmed3f(py, 0.f, (float)(d.Y - 1));
    float cz = __builtin_amdgcn_fmed3f(pz, 0.f, (float)(d.Z - 1));
    // cx >= 0 after the clamp: truncation is floor, and cx - floor(cx) is exact (v_fract_f32)
    k.gx = (int)cx;
    k.gy = (int)cy;
    k.gz = (int)cz;
    k.a = __builtin_amdgcn_fractf(cx);
    k.b = __builtin_amdgcn_fractf(cy);
    k.c = __builtin_amdgcn_fractf(cz);
    k.hx = min(k.gx + 1, d.X - 1);
    k.hy = min(k.gy + 1, d.Y - 1);
    k.hz = min(k.gz + 1, d.Z - 1);
    return k;
}

// MX(q)^2 - MN(q)^2 with MX = fmaxf(., 0), MN = fminf(., 0), bit for bit: one of the two squares
// is +0 and the other the rounded q^2, so the difference is q |q| when that is nonzero, and +0
// when it is zero (q = +-0, or |q| so small that q^2 underflows: q |q| would be -0 for a negative
// q) or NaN (fmaxf and fminf drop the NaN).  One ordered not-equal compare covers both cases.
__device__ __forceinline__ float sq_signed(float q) {
    const float s = q * fabsf(q);
    return (s < 0.f || s > 0.f) ? s : 0.f;
}

__device__ __forceinline__ int lin(Dims d, int x, int y, int z) { return x + d.X * (y + d.Y * z); }

// linear index of a corner (coordinates in the grid) with 24-bit multiplies: X and X*Y are below
// 2^24 (checked when a context is created), and the products' low 32 bits are the index
__device__ __forceinline__ unsigned corner_index(Dims d, const Corner &k) {
    return (unsigned)k.gx + __umul24((unsigned)d.X, (unsigned)k.gy) + __umul24((unsigned)(d.X * d.Y), (unsigned)k.gz);
}

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

// XCD-aware tile order: the hardware hands consecutive workgroups to the 8 XCDs round-robin;
// remap so that each XCD works on one contiguous run of tiles (a z-slab), keeping stencil and
// gather neighbours in that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    int per = nb >> 3, rem = nb & 7, x = b & 7, s = b >> 3;
    return (x < rem) ? x * (per + 1) + s : rem * (per + 1) + (x - rem) * per + s;
}

// max over the NSLOT slots of one iteration
__device__ __forceinline__ unsigned long long shfl_xor_u64(unsigned long long v, int m) {
    unsigned lo = (unsigned)v, hi = (unsigned)(v >> 32);
    lo = __shfl_xor(lo, m, 64);
    hi = __shfl_xor(hi, m, 64);
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        unsigned long long o = shfl_xor_u64(v, m);
        v = o > v ? o : v;
    }
    return v;
}

// (all 64 lanes of the calling wave must be active)
__device__ __forceinline__ unsigned long long slots_max(const unsigned long long *slots) {
    const int lane = threadIdx.x & 63;
    unsigned long long v = 0ull;
    for (int j = lane; j < NSLOT; j += 64) {
        const unsigned long long s = __builtin_nontemporal_load(&slots[j]);
        v = s > v ? s : v;
    }
    return wave_max_u64(v);
}

// stop rules checked at the start of each launch of iteration a.iter, by a whole wave (uniform
// over it):
// converged at iter-1 (solver.cu:418-421) or NaN energy at iter (solver.cu:366-368)
__device__ __forceinline__ bool should_stop(const IterArgs &a, bool record) {
    if (__builtin_nontemporal_load(&a.ctl->stopped)) return true;
    if (a.check_conv && a.iter > 1) {
        unsigned long long mp = slots_max(a.maxpack + (size_t)(a.iter - 2) * NSLOT);
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

// Buffer resources (raw, stride 0): loads at byte offsets >= num_records return 0.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int RSRC_FLAGS = 0x00020000;

__device__ __forceinline__ Rsrc rsrc(const float *p, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, bytes, RSRC_FLAGS);
}
__device__ __forceinline__ float bld1(Rsrc r, unsigned voff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}
// Three dword loads, which the compiler merges into one buffer_load_dwordx3.  (This compiler lowers
// the vector-returning __builtin_amdgcn_raw_buffer_load_b96 / _b128 to a single-dword load and
// leaves the other components undefined, so they are not used.)
__device__ __forceinline__ F3 bld3(Rsrc r, unsigned voff) {
    return f3(bld1(r, voff), bld1(r, voff + 4u), bld1(r, voff + 8u));
}

// Buffer stores; a store at an offset past num_records is dropped.  Byte offsets stay below 2^31
// (the library rejects volumes with 12*N >= 2^31), so NO_STORE is past every buffer.
constexpr unsigned NO_STORE = 0x80000000u;
__device__ __forceinline__ void bst1(Rsrc r, unsigned voff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, 0, 0);
}
__device__ __forceinline__ void bst3(Rsrc r, unsigned voff, F3 v) {
    bst1(r, voff, v.x);
    bst1(r, voff + 4u, v.y);
    bst1(r, voff + 8u, v.z);
}
// Streaming stores (cache policy nt): pass A's outputs are not read again during the pass, and
// marking them non-temporal leaves the caches to the planes and taps still being read (measured:
// pass A -2 us, and pass B, which gathers the v_new they write, -2 us at 182x218x182).
constexpr int CP_NT = 2;
__device__ __forceinline__ void bst1_nt(Rsrc r, unsigned voff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, 0, CP_NT);
}
__device__ __forceinline__ void bst3_nt(Rsrc r, unsigned voff, F3 v) {
    bst1_nt(r, voff, v.x);
    bst1_nt(r, voff + 4u, v.y);
    bst1_nt(r, voff + 8u, v.z);
}

// Trilinear blend of a field at corner g with the +1 neighbours at byte offsets dx, dy, dz from
// it.  The +1 neighbours are not clamped: when one leaves the grid its weight is exactly 0 and the
// address either holds another voxel of the field (finite) or lies past the end of the buffer
// (read as 0), so the blend equals the reference's (cuda_utils.hpp:60-136).
__device__ __forceinline__ float tri_at1(Rsrc r, unsigned g, unsigned dx, unsigned dy, unsigned dz, const Corner &k) {
    return tri(bld1(r, g + dx + dy + dz), bld1(r, g + dx + dy), bld1(r, g + dx + dz), bld1(r, g + dx),
               bld1(r, g + dy + dz), bld1(r, g + dy), bld1(r, g + dz), bld1(r, g), k.a, k.b, k.c);
}
__device__ __forceinline__ F3 tri_at3(Rsrc r, unsigned g, unsigned dx, unsigned dy, unsigned dz, const Corner &k) {
    const F3 f111 = bld3(r, g + dx + dy + dz), f110 = bld3(r, g + dx + dy), f101 = bld3(r, g + dx + dz),
             f100 = bld3(r, g + dx), f011 = bld3(r, g + dy + dz), f010 = bld3(r, g + dy), f001 = bld3(r, g + dz),
             f000 = bld3(r, g);
    return f3(tri(f111.x, f110.x, f101.x, f100.x, f011.x, f010.x, f001.x, f000.x, k.a, k.b, k.c),
              tri(f111.y, f110.y, f101.y, f100.y, f011.y, f010.y, f001.y, f000.y, k.a, k.b, k.c),
              tri(f111.z, f110.z, f101.z, f100.z, f011.z, f010.z, f001.z, f000.z, k.a, k.b, k.c));
}

// Values of the adjacent lanes of the wave (wave-wide DPP shifts; the lanes of a wave are
// consecutive x): lane_next gives lane l the value of lane l+1 (x+1), lane_prev that of lane l-1
// (x-1).  The end lanes receive 0 and are never computed.
__device__ __forceinline__ float lane_next(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_prev(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ F3 lane_next3(F3 v) { return f3(lane_next(v.x), lane_next(v.y), lane_next(v.z)); }
__device__ __forceinline__ F3 lane_prev3(F3 v) { return f3(lane_prev(v.x), lane_prev(v.y), lane_prev(v.z)); }

// Thread -> voxel column.  A block is a.ty waves, one per row y; a wave computes a.xs
// consecutive interior columns in its middle lanes and holds (WAVE - a.xs) / 2 halo lanes on each
// side, which load their column but store nothing.  `own` marks a lane that stores.
struct TileCoord {
    int x, y, zs, ze;
    bool own;
};

// blk: the block's index within the launch grid of this volume (blockIdx.x for a single volume;
// in a batch, the index within the volume's run of blocks, which starts at a multiple of 8 so that
// the XCD of a block is still blk % 8)
__device__ __forceinline__ TileCoord tile_of(const IterArgs &a, int blk) {
    const int b = xcd_remap(blk, a.nblocks);
    const int bx = b % a.nbx, r = b / a.nbx, by = r % a.nby, bz = r / a.nby;
    const int lane = threadIdx.x & (WAVE - 1), off = (WAVE - a.xs) >> 1;
    TileCoord t;
    t.x = 1 + bx * a.xs + lane - off;
    t.y = 1 + by * a.ty + threadIdx.x / WAVE;
    t.own = lane >= off && lane < off + a.xs && t.x <= a.d.X - 2;
    t.zs = 1 + bz * a.zc;
    t.ze = min(t.zs + a.zc, a.d.Z - 1);
    return t;
}

// Storage of velocity and density (ST_F32: v as 3 floats and rho as a float in separate buffers,
// 12 + 4 B per voxel; ST_BF16 / ST_F16: one 8-byte record per voxel holding v.x, v.y, v.z, rho as
// bf16 / fp16, low half first).  All arithmetic is FP32; a value of v or rho is rounded to the
// storage format (nearest even) as soon as it is computed, and only the rounded value is used
// afterwards, so the 16-bit forms are exact FP32 computations on representable values.
__device__ __forceinline__ unsigned f32_to_bf16(float f) {
    return __builtin_bit_cast(unsigned short, (__bf16)f);  // v_cvt_pk_bf16_f32: IEEE round to nearest even
}
__device__ __forceinline__ float bf16_to_f32(unsigned h) { return __builtin_bit_cast(float, h << 16); }
__device__ __forceinline__ unsigned f32_to_f16(float f) {
    return __builtin_bit_cast(unsigned short, (_Float16)f);  // IEEE round to nearest even
}
__device__ __forceinline__ float f16_to_f32(unsigned h) {
    return (float)__builtin_bit_cast(_Float16, (unsigned short)h);
}

template <int ST> __device__ __forceinline__ unsigned h_of(float f) {
    return ST == ST_BF16 ? f32_to_bf16(f) : f32_to_f16(f);
}
template <int ST> __device__ __forceinline__ float f_of(unsigned h) {
    return ST == ST_BF16 ? bf16_to_f32(h) : f16_to_f32(h);
}
// the value after the storage's rounding.  The FP32 value is pinned in a register first: the
// compiler would otherwise fuse an FMA that produces it with the fp16 conversion
// (v_fma_mix*_f16), rounding once from the exact result instead of from the FP32 value.
template <int ST> __device__ __forceinline__ float q16(float f) {
    if constexpr (ST == ST_F32) {
        return f;
    } else {
        asm volatile("" : "+v"(f));
        return f_of<ST>(h_of<ST>(f));
    }
}

// (v, rho) of one voxel as loaded: floats, or the raw 8-byte record
template <int ST> struct VR {
    unsigned lo, hi;
};
template <> struct VR<ST_F32> {
    F3 v;
    float r;
};

// the velocity/density buffers of one kernel: ST_F32 uses v and r, the 16-bit forms v only
struct VRBuf {
    Rsrc v, r;
};

template <int ST> __device__ __forceinline__ VRBuf vr_buf(const float *v, const float *r, int N) {
    VRBuf b;
    b.v = rsrc(v, (ST == ST_F32 ? 12 : 8) * N);
    b.r = ST == ST_F32 ? rsrc(r, 4 * N) : b.v;
    return b;
}

template <int ST> __device__ __forceinline__ VR<ST> vr_load(const VRBuf &b, int i) {
    VR<ST> q;
    if constexpr (ST == ST_F32) {
        q.v = bld3(b.v, 12u * i);
        q.r = bld1(b.r, 4u * i);
    } else {
        q.lo = __builtin_amdgcn_raw_buffer_load_b32(b.v, 8u * i, 0, 0);
        q.hi = __builtin_amdgcn_raw_buffer_load_b32(b.v, 8u * i + 4u, 0, 0);
    }
    return q;
}
// v only (ST_F32 skips the density load)
template <int ST> __device__ __forceinline__ VR<ST> v_load(const VRBuf &b, int i) {
    if constexpr (ST == ST_F32) {
        VR<ST> q;
        q.v = bld3(b.v, 12u * i);
        q.r = 0.f;
        return q;
    } else {
        return vr_load<ST>(b, i);
    }
}
template <int ST> __device__ __forceinline__ F3 vr_v(const VR<ST> &q) {
    if constexpr (ST == ST_F32) return q.v;
    else return f3(f_of<ST>(q.lo & 0xFFFFu), f_of<ST>(q.lo >> 16), f_of<ST>(q.hi & 0xFFFFu));
}
template <int ST> __device__ __forceinline__ float vr_r(const VR<ST> &q) {
    if constexpr (ST == ST_F32) return q.r;
    else return f_of<ST>(q.hi >> 16);
}
// store (v, rho), already rounded to the storage format; a dropped store when !store
template <int ST> __device__ __forceinline__ void vr_store(const VRBuf &b, int i, bool store, F3 v, float r) {
    if constexpr (ST == ST_F32) {
        bst1_nt(b.r, store ? 4u * i : NO_STORE, r);
        bst3_nt(b.v, store ? 12u * i : NO_STORE, v);
    } else {
        const unsigned o = store ? 8u * i : NO_STORE;
        __builtin_amdgcn_raw_buffer_store_b32(h_of<ST>(v.x) | (h_of<ST>(v.y) << 16), b.v, o, 0, CP_NT);
        __builtin_amdgcn_raw_buffer_store_b32(h_of<ST>(v.z) | (h_of<ST>(r) << 16), b.v, o + 4u, 0, CP_NT);
    }
}

// pass A: steps 1-8 and 10 of one iteration per interior voxel.  One thread per (x, y) column,
// marching the chunk's z range; the column's z-neighbours (v at z-2..z+2, rho and psi_inv at
// z-1..z+2, S at z-1..z+1) slide through registers and the x-neighbours come from the adjacent
// lanes.  The loads of plane z+1 (y-neighbours, the corner of psi_inv) and the stage's rows of the
// planes ahead are issued before plane z is computed, so that their latency overlaps the
// arithmetic: software pipelining across the z-march.
// the y-neighbours of the voxel being computed
struct PlaneNb {
    F3 v_yp, v_ym;  // v at y+-1
    F3 v_y2;        // v at y+-2 on the side the upwind Laplacian reads
    F3 p_yp, p_ym;  // psi_inv at y+-1
    float r_yp, r_ym, s_yp, s_ym;  // rho and S at y+-1
};

// The y-neighbours, exchanged through LDS: a ring of 2 planes (the plane being computed is read
// while the next one is written) of the 12 rows y-2 .. y+9 around the block's 8 rows, holding
// (v, rho) and (psi_inv, S) per voxel.  Every wave writes its own row from its column window;
// the 4 halo rows are loaded one per wave by waves 4-7 (waves 0-3 carry the stages' extra psi and
// L/T rows; this split keeps the waves, which step in lockstep, evenly loaded).
constexpr int YS_ROWS = TY_A + 4, YS_PLANES = 2;
struct YStage {
    float4 vr[YS_PLANES][YS_ROWS][WAVE];
    float4 ps[YS_PLANES][YS_ROWS][WAVE];
};
// a halo row as loaded: FP32 values, or in the 16-bit forms the (v, rho) record, unpacked when the
// row is written to the stage a step after its load, so that the unpacking does not wait for it
struct YRow {
    F3 v, p;
    float r, s;
};
template <int ST> struct YRow16 {
    VR<ST> vr;
    F3 p;
    float s;
};
template <int ST> using YRowOf = typename std::conditional<ST == ST_F32, YRow, YRow16<ST>>::type;

// L = -Laplacian(psi) and T on the planes around the one being computed, staged in LDS for the
// whole block: a ring of 4 planes (z-1, z, z+1 are read while z+2 is written) of the 12 rows
// y-2 .. y+9 around the block's 8 rows (row j holds y = yb - 1 + j; rows 1..10, the block's rows
// and one on either side, hold L; rows 0 and 11 are written and never read), 64 columns (the
// lanes of a wave), (L.x, L.y, L.z, T) per voxel.  L is not a field in memory: the block computes
// it in the stage in the reference's operation order (vector_fields.cu:275-300), 0 off the
// interior.  Each wave holds two psi rows of the stage's 12 and their planes z+1..z+3 in
// registers: the z-neighbours of L's 7-point stencil come from there, the x-neighbours from the
// adjacent lanes, and only the y-neighbours (rows of other waves) from a 2-plane psi ring in LDS.
// The 8 taps of a voxel whose psi_inv lies within one voxel of it are all in the L/T stage, so
// the trilinear taps are LDS reads and the block loads each plane of psi and T once, in whole
// rows.
constexpr int LT_ROWS = TY_A + 2, LT_PLANES = 4;  // rows holding L
constexpr int ST_ROWS = TY_A + 4, PS_PLANES = 2;  // rows of both stages
struct LTStage {
    float4 v[LT_PLANES][ST_ROWS][WAVE];
};
struct PsiStage {
    float4 v[PS_PLANES][ST_ROWS][WAVE];  // (x, y, z, unused): 16-byte entries for aligned LDS access
};
// one psi row of the stage at the planes around the L plane being computed: zz-1, zz, zz+1
struct PsiWin {
    F3 m, c, p;
};

struct PassAView {
    Dims d;
    VRBuf rvr, wvr;       // (v, rho): old (loads), new (stores)
    Rsrc rp, rs, rt, rq;  // loads: psi_inv (old), S, T, psi
    Rsrc wp;              // stores: psi_inv (new)
    int X, XY;
    unsigned X12, XY12, X4, XY4;
    int xb, yb;           // the L/T stage's first column (x of lane 0) and first row (the block's y - 1)
    LTStage *lt;
    PsiStage *ps;         // row j holds y = yb - 1 + j (as the L/T stage)
    YStage *ys;           // row j holds y = yb - 1 + j
};

template <int ST> __device__ __forceinline__ YRowOf<ST> yrow_load(const PassAView &w, int i) {
    const VR<ST> q = vr_load<ST>(w.rvr, i);
    YRowOf<ST> h;
    if constexpr (ST == ST_F32) {
        h.v = vr_v<ST>(q);
        h.r = vr_r<ST>(q);
    } else {
        h.vr = q;
    }
    h.p = bld3(w.rp, 12u * i);
    h.s = bld1(w.rs, 4u * i);
    return h;
}
__device__ __forceinline__ void ys_store(const PassAView &w, int zz, int row, int lane, F3 v, float r, F3 p, float s) {
    w.ys->vr[zz & (YS_PLANES - 1)][row][lane] = make_float4(v.x, v.y, v.z, r);
    w.ys->ps[zz & (YS_PLANES - 1)][row][lane] = make_float4(p.x, p.y, p.z, s);
}
template <int ST> __device__ __forceinline__ void ys_store_row(const PassAView &w, int zz, int row, int lane,
                                                              const YRowOf<ST> &h) {
    if constexpr (ST == ST_F32) ys_store(w, zz, row, lane, h.v, h.r, h.p, h.s);
    else ys_store(w, zz, row, lane, vr_v<ST>(h.vr), vr_r<ST>(h.vr), h.p, h.s);
}
// (x, y, z) of a 16-byte stage entry as one 12-byte LDS read (ds_read_b96): the post-barrier burst of
// stage reads is the critical path of a step, and the fourth component is not needed here
__device__ __forceinline__ F3 xyz_of(const float4 &e) {
    const float *p = &e.x;
    return f3(p[0], p[1], p[2]);
}
// the y-neighbours of plane zz for the wave whose row is j, the upwind side chosen by vneg (v.y < 0)
__device__ __forceinline__ PlaneNb ys_read(const PassAView &w, int zz, int j, int lane, bool vneg) {
    const int sl = zz & (YS_PLANES - 1);
    const float4 a = w.ys->vr[sl][j + 1][lane], b = w.ys->vr[sl][j - 1][lane], e = w.ys->ps[sl][j + 1][lane],
                 f = w.ys->ps[sl][j - 1][lane];
    const F3 c = xyz_of(w.ys->vr[sl][vneg ? j + 2 : j - 2][lane]);  // rho unused: a 12-byte read
    PlaneNb n;
    n.v_yp = f3(a.x, a.y, a.z);
    n.r_yp = a.w;
    n.v_ym = f3(b.x, b.y, b.z);
    n.r_ym = b.w;
    n.v_y2 = f3(c.x, c.y, c.z);
    n.p_yp = f3(e.x, e.y, e.z);
    n.s_yp = e.w;
    n.p_ym = f3(f.x, f.y, f.z);
    n.s_ym = f.w;
    return n;
}

// Rows for the stages, loaded by a wave at its lanes' columns.  Positions outside the grid read
// other voxels of the fields or, past the buffers, 0: finite values that no in-stage tap of a
// stored voxel uses.
__device__ __forceinline__ F3 psi_row_load(const PassAView &w, int x, int y, int zz) {
    return bld3(w.rq, 12u * (x + w.X * y + w.XY * zz));
}
__device__ __forceinline__ float t_row_load(const PassAView &w, int x, int y, int zz) {
    return bld1(w.rt, 4u * (x + w.X * y + w.XY * zz));
}
__device__ __forceinline__ void psi_store(const PassAView &w, int zz, int row, int lane, F3 p) {
    w.ps->v[zz & (PS_PLANES - 1)][row][lane] = make_float4(p.x, p.y, p.z, 0.f);
}
__device__ __forceinline__ PsiWin psi_win(F3 m, F3 c, F3 p) { return PsiWin{m, c, p}; }
__device__ __forceinline__ F3 ps_at(const PassAView &w, int zz, int row, int lane) {
    return xyz_of(w.ps->v[zz & (PS_PLANES - 1)][row][lane]);
}

// L of stage row j (y = yb - 1 + j) of plane zz, with T of that row, into the L/T stage: psi of the
// row at planes zz-1..zz+1 from the wave's registers (pw), its x-neighbours from the adjacent
// lanes, its y-neighbours (rows j-1, j+1, clamped into the stage for the rows that are never
// read) from the psi ring.  Every lane of the wave runs it; the end lanes' L (x-neighbours
// missing) is never used.
__device__ __forceinline__ void lt_compute(const PassAView &w, int zz, int j, int lane, float t, const PsiWin &pw) {
    const F3 c = pw.c, ym = ps_at(w, zz, max(j - 1, 0), lane), yp = ps_at(w, zz, min(j + 1, ST_ROWS - 1), lane),
             zm = pw.m, zp = pw.p;
    F3 s = mul3(-6.f, c);
    s = add3(s, lane_next3(c));
    s = add3(s, lane_prev3(c));
    s = add3(s, yp);
    s = add3(s, ym);
    s = add3(s, zp);
    s = add3(s, zm);
    // The empty asm keeps the sum out of the exec-masked block that selects the interior lanes
    // (the compiler would sink it there), so that the lane shifts fold into the adds as DPP
    // operands: 6 fewer VALU instructions per row.
    asm volatile("" : "+v"(s.x), "+v"(s.y), "+v"(s.z));
    const F3 L = mul3(-1.f, s);
    const int x = w.xb + lane, y = w.yb - 1 + j;
    const Dims d = w.d;
    const bool in = x >= 1 && x <= d.X - 2 && y >= 1 && y <= d.Y - 2 && zz >= 1 && zz <= d.Z - 2;
    w.lt->v[zz & (LT_PLANES - 1)][j][lane] = in ? make_float4(L.x, L.y, L.z, t) : make_float4(0.f, 0.f, 0.f, t);
}

// L at grid point (x, y, z) from psi in memory (0 off the interior and outside the grid)
__device__ __forceinline__ F3 l_global(const PassAView &w, int x, int y, int z) {
    const Dims d = w.d;
    if (!(x >= 1 && x <= d.X - 2 && y >= 1 && y <= d.Y - 2 && z >= 1 && z <= d.Z - 2)) return f3(0.f, 0.f, 0.f);
    const unsigned o = 12u * (x + w.X * y + w.XY * z);
    F3 s = mul3(-6.f, bld3(w.rq, o));
    s = add3(s, bld3(w.rq, o + 12u));
    s = add3(s, bld3(w.rq, o - 12u));
    s = add3(s, bld3(w.rq, o + w.X12));
    s = add3(s, bld3(w.rq, o - w.X12));
    s = add3(s, bld3(w.rq, o + w.XY12));
    s = add3(s, bld3(w.rq, o - w.XY12));
    return mul3(-1.f, s);
}

// Whether the 8 taps of corner k (of the voxel at plane z) are all in the L/T stage: columns 1..62
// (an end lane lacks an x-neighbour, so its L is valid only where the lane lies on or past the
// volume's edge column, L being 0 there: lane 0 when it is x = 0, lane 63 when it is x >= X - 1,
// where a tap's weight is 0 past x = X - 1), rows y = yb .. yb + 9 of the corner and its +1 row
// (stage rows 1..10), planes z-1..z+1.  A voxel displaced by a voxel or more fails it.
__device__ __forceinline__ bool taps_in_stage(const PassAView &w, const Corner &k, int z) {
    const int cx = k.gx - w.xb, ry = k.gy - w.yb, dz = k.gz - z;
    const int lo = w.xb == 0 ? 0 : 1, hi = w.xb + WAVE - 1 >= w.d.X - 1 ? WAVE - 2 : WAVE - 3;
    return cx >= lo && cx <= hi && ry >= 0 && ry <= LT_ROWS - 2 && (dz == -1 || dz == 0);
}

// L and T at psi_inv(p) (corner k) of the voxel at plane z from the stage, the indices clamped
// into it (the values are those of the reference when taps_in_stage holds)
__device__ __forceinline__ void lt_stage_taps(const PassAView &w, const Corner &k, int z, F3 &Lo, float &Tv) {
    const int dz = k.gz - z;
    const int cx = min(max(k.gx - w.xb, 0), WAVE - 2);
    const int ry = min(max(k.gy - w.yb, 0), LT_ROWS - 2) + 1;  // the corner's stage row
    const int zg = dz == -1 ? z - 1 : z;
    const int s0 = zg & (LT_PLANES - 1), s1 = (zg + 1) & (LT_PLANES - 1);
    const LTStage &st = *w.lt;
    const float4 f000 = st.v[s0][ry][cx], f001 = st.v[s1][ry][cx], f010 = st.v[s0][ry + 1][cx],
                 f011 = st.v[s1][ry + 1][cx], f100 = st.v[s0][ry][cx + 1], f101 = st.v[s1][ry][cx + 1],
                 f110 = st.v[s0][ry + 1][cx + 1], f111 = st.v[s1][ry + 1][cx + 1];
    Lo = f3(tri(f111.x, f110.x, f101.x, f100.x, f011.x, f010.x, f001.x, f000.x, k.a, k.b, k.c),
            tri(f111.y, f110.y, f101.y, f100.y, f011.y, f010.y, f001.y, f000.y, k.a, k.b, k.c),
            tri(f111.z, f110.z, f101.z, f100.z, f011.z, f010.z, f001.z, f000.z, k.a, k.b, k.c));
    Tv = tri(f111.w, f110.w, f101.w, f100.w, f011.w, f010.w, f001.w, f000.w, k.a, k.b, k.c);
}

// L and T at psi_inv(p) (corner k) from memory: L at each corner from psi by l_global.  The +1
// corners are not clamped: where the reference clamps one its weight is exactly 0, and its L here
// is 0 (outside the interior), a finite value.
__device__ __forceinline__ void lt_memory_taps(const PassAView &w, const Corner &k, F3 &Lo, float &Tv) {
    const F3 l111 = l_global(w, k.gx + 1, k.gy + 1, k.gz + 1), l110 = l_global(w, k.gx + 1, k.gy + 1, k.gz),
             l101 = l_global(w, k.gx + 1, k.gy, k.gz + 1), l100 = l_global(w, k.gx + 1, k.gy, k.gz),
             l011 = l_global(w, k.gx, k.gy + 1, k.gz + 1), l010 = l_global(w, k.gx, k.gy + 1, k.gz),
             l001 = l_global(w, k.gx, k.gy, k.gz + 1), l000 = l_global(w, k.gx, k.gy, k.gz);
    Lo = f3(tri(l111.x, l110.x, l101.x, l100.x, l011.x, l010.x, l001.x, l000.x, k.a, k.b, k.c),
            tri(l111.y, l110.y, l101.y, l100.y, l011.y, l010.y, l001.y, l000.y, k.a, k.b, k.c),
            tri(l111.z, l110.z, l101.z, l100.z, l011.z, l010.z, l001.z, l000.z, k.a, k.b, k.c));
    const unsigned g = corner_index(w.d, k);
    Tv = tri_at1(w.rt, 4u * g, 4u, w.X4, w.XY4, k);
}

// own-column z-windows; rho and psi_inv run one plane further ahead than they are used, so that
// the corner of the plane being prefetched comes from a value loaded a whole step earlier
struct ColumnWindow {
    F3 v_m2, v_m1, v_0, v_p1, v_p2, p_m1, p_0, p_p1, p_p2;
    float r_m1, r_0, r_p1, r_p2, s_m1, s_0, s_p1;
};

// Everything one voxel's update reads: rho, v, psi_inv and S at the voxel and its stencil
// neighbours (v at +-2 along x and z, and at y+-2 on the upwind side in y, chosen by v.y < 0), and
// L and T sampled at psi_inv.
struct VoxelIn {
    float rc, r_xp, r_xm, r_yp, r_ym, r_zp, r_zm;
    F3 vc, v_xp, v_xm, v_xpp, v_xmm, v_yp, v_ym, v_y2, v_zp, v_zm, v_zpp, v_zmm;
    F3 pc, p_xp, p_xm, p_yp, p_ym, p_zp, p_zm;
    float sc, s_xp, s_xm, s_yp, s_ym, s_zp, s_zm;
    F3 Lo;
    float Tv;
};

// steps 1-8 and 10 of one iteration for one interior voxel (lap: inside the range of the upwind
// velocity Laplacian), the values rounded to the storage format as they are computed
template <int ST>
__device__ __forceinline__ void voxel_math(const IterArgs &a, const VoxelIn &q, bool lap, float &rho_n, F3 &vn,
                                           F3 &pinv_n) {
    // 1. density (solver.cu:101-173)
    const F3 vc = q.vc;
    const float rc = q.rc;
    const float t11 = (vc.x > 0.f) ? -vc.x * rc : -vc.x * q.r_xp;
    const float vxm = q.v_xm.x;
    const float t12 = (vxm > 0.f) ? vxm * q.r_xm : vxm * rc;
    const float t21 = (vc.y > 0.f) ? -vc.y * rc : -vc.y * q.r_yp;
    const float vym = q.v_ym.y;
    const float t22 = (vym > 0.f) ? vym * q.r_ym : vym * rc;
    const float t31 = (vc.z > 0.f) ? -vc.z * rc : -vc.z * q.r_zp;
    const float vzm = q.v_zm.z;
    const float t32 = (vzm > 0.f) ? vzm * q.r_zm : vzm * rc;
    const float div = ((((t11 + t12) + t21) + t22) + t31) + t32;
    rho_n = q16<ST>(fmaf(a.alpha, div, rc));

    // 2. J_inv and det() (vector_fields.cu:503-535)
    const F3 jx = mul3(0.5f, sub3(q.p_xp, q.p_xm));
    const F3 jy = mul3(0.5f, sub3(q.p_yp, q.p_ym));
    const F3 jz = mul3(0.5f, sub3(q.p_zp, q.p_zm));
    const float det = det_ref(jx.x, jy.x, jz.x, jx.y, jy.y, jz.y, jx.z, jy.z, jz.z);

    // 3-5. L o psi_inv and T o psi_inv (vector_fields.cu:231-247, 81-99; in q), grad S and the
    // potential gradient (vector_fields.cu:155-192, solver.cu:56-80)
    const F3 gS = f3((q.s_xp - q.s_xm) * 0.5f, (q.s_yp - q.s_ym) * 0.5f, (q.s_zp - q.s_zm) * 0.5f);
    const float diff = q.sc - q.Tv;
    const F3 nU = mul3(det, add3(mul3(diff, gS), mul3(a.w_reg, q.Lo)));

    // 6. (Dv)v (vector_fields.cu:352-404)
    const F3 xp = q.v_xp, xm = q.v_xm, yp = q.v_yp, ym = q.v_ym;
    const F3 zp = q.v_zp, zm = q.v_zm;
    // The reference's diagonal terms ((MX(u)^2 - MN(u)^2) + MN(w)^2) - MX(w)^2, MX = fmaxf(., 0),
    // MN = fminf(., 0), are sq(u) - sq(w) bit for bit, zero signs included: sq(u) is the first
    // difference exactly (+0 when it is zero), adding +0 = MN(w)^2 or subtracting +0 = MX(w)^2
    // changes nothing, and a - (-b) == a + b.
    const float ax1 = sq_signed(vc.x) - sq_signed(xp.x);
    const float ax2 = vc.y * ((vc.y > 0.f) ? (vc.x - ym.x) : (yp.x - vc.x));
    const float ax3 = vc.z * ((vc.z > 0.f) ? (vc.x - zm.x) : (zp.x - vc.x));
    const float ay1 = vc.x * ((vc.x > 0.f) ? (vc.y - xm.y) : (xp.y - vc.y));
    const float ay2 = sq_signed(vc.y) - sq_signed(yp.y);
    const float ay3 = vc.z * ((vc.z > 0.f) ? (vc.y - zm.y) : (zp.y - vc.y));
    const float az1 = vc.x * ((vc.x > 0.f) ? (vc.z - xm.z) : (xp.z - vc.z));
    const float az2 = vc.y * ((vc.y > 0.f) ? (vc.z - ym.z) : (yp.z - vc.z));
    const float az3 = sq_signed(vc.z) - sq_signed(zp.z);
    const F3 dvv = f3((0.5f * ax1 + ax2) + ax3, (ay1 + 0.5f * ay2) + ay3, (az1 + az2) + 0.5f * az3);

    // 7. upwind velocity Laplacian, the reference's indices and ranges (vector_fields.cu:307-346);
    // both branches have the form (A - 2B) + C.  At y = Y-2 the (y+2) read is the linear address
    // (x, 0, z+1), inside the buffer.
    // evaluated on every lane and selected, so that the z-march has no branch
    const bool nx = vc.x < 0.f, ny = vc.y < 0.f, nz = vc.z < 0.f;
    const F3 v_x2 = nx ? q.v_xpp : q.v_xmm;
    const F3 Ax = nx ? v_x2 : vc, Bx = nx ? xp : xm, Cx = nx ? vc : v_x2;
    const F3 Ay = ny ? q.v_y2 : vc, Cy = ny ? vc : q.v_y2;
    const F3 Az = nz ? q.v_zpp : vc, Cz = nz ? vc : q.v_zmm;
    const F3 vxx = add3(sub_twice(Ax, Bx), Cx);
    const F3 vyy = add3(sub_twice(Ay, ym), Cy);
    const F3 vzz = add3(sub_twice(Az, zm), Cz);
    const F3 dl = mul3(-1.f, add3(add3(vxx, vyy), vzz));
    const F3 dlt = lap ? dl : f3(0.f, 0.f, 0.f);

    // 8. velocity update with this iteration's rho (solver.cu:188-211)
    const F3 fr = mul3(a.m1t, vc);
    const F3 df = mul3(a.tau, dlt);
    const float qv = a.m2 / (fabsf(rho_n) + 1e-5f);
    const F3 po = mul3(qv, nU);
    const F3 up = mul3(a.alpha, add3(sub3(add3(fr, dvv), df), po));
    vn = f3(q16<ST>(vc.x - up.x), q16<ST>(vc.y - up.y), q16<ST>(vc.z - up.z));

    // 10. psi_inv -= alpha * (J_inv v_new) (solver.cu:263-279)
    const F3 pc = q.pc;
    const F3 Jv = f3((jx.x * vn.x + jy.x * vn.y) + jz.x * vn.z, (jx.y * vn.x + jy.y * vn.y) + jz.y * vn.z,
                     (jx.z * vn.x + jy.z * vn.y) + jz.z * vn.z);
    const F3 ui = mul3(a.alpha, Jv);
    pinv_n = f3(pc.x - ui.x, pc.y - ui.y, pc.z - ui.z);
}

// The same steps with the reference's multiply-adds contracted into fma (each product feeding a
// sum is not rounded on its own), the (Dv)v diagonal terms as fma(u, |u|, -w|w|), and
// m2 / (|rho| + 1e-5) as m2 times the hardware reciprocal (the reference's own division there is
// not the IEEE one).  The results differ from the reference order by a few ulps per step (the
// arithmetic mode ARITH_FMA, FP32 storage).
__device__ __forceinline__ void voxel_math_fma(const IterArgs &a, const VoxelIn &q, bool lap, float &rho_n, F3 &vn,
                                               F3 &pinv_n) {
    // 1. density: the six upwind fluxes summed in the reference's order
    const F3 vc = q.vc;
    const float rc = q.rc;
    const float vxm = q.v_xm.x, vym = q.v_ym.y, vzm = q.v_zm.z;
    float div = -vc.x * ((vc.x > 0.f) ? rc : q.r_xp);
    div = fmaf(vxm, (vxm > 0.f) ? q.r_xm : rc, div);
    div = fmaf(-vc.y, (vc.y > 0.f) ? rc : q.r_yp, div);
    div = fmaf(vym, (vym > 0.f) ? q.r_ym : rc, div);
    div = fmaf(-vc.z, (vc.z > 0.f) ? rc : q.r_zp, div);
    div = fmaf(vzm, (vzm > 0.f) ? q.r_zm : rc, div);
    rho_n = fmaf(a.alpha, div, rc);

    // 2. J_inv and the reference's det()
    const F3 jx = mul3(0.5f, sub3(q.p_xp, q.p_xm));
    const F3 jy = mul3(0.5f, sub3(q.p_yp, q.p_ym));
    const F3 jz = mul3(0.5f, sub3(q.p_zp, q.p_zm));
    const float m00 = jx.x, m01 = jy.x, m02 = jz.x, m10 = jx.y, m11 = jy.y, m12 = jz.y, m20 = jx.z, m21 = jy.z,
                m22 = jz.z;
    const float da = fmaf(m20 * m01, m12, fmaf(m00 * m11, m22, (m10 * m21) * m02));
    const float db = fmaf(fmaf(m02, m11, m20), (m12 * m21) * m00, (m22 * m01) * m10);
    const float det = da - db;

    // 3-5. grad S and the potential gradient
    const F3 gS = f3((q.s_xp - q.s_xm) * 0.5f, (q.s_yp - q.s_ym) * 0.5f, (q.s_zp - q.s_zm) * 0.5f);
    const float diff = q.sc - q.Tv;
    const F3 nU = f3(det * fmaf(diff, gS.x, a.w_reg * q.Lo.x), det * fmaf(diff, gS.y, a.w_reg * q.Lo.y),
                     det * fmaf(diff, gS.z, a.w_reg * q.Lo.z));

    // 6. (Dv)v
    const F3 xp = q.v_xp, xm = q.v_xm, yp = q.v_yp, ym = q.v_ym, zp = q.v_zp, zm = q.v_zm;
    const float ax1 = fmaf(vc.x, fabsf(vc.x), -(xp.x * fabsf(xp.x)));
    const float ax2 = vc.y * ((vc.y > 0.f) ? (vc.x - ym.x) : (yp.x - vc.x));
    const float ax3 = vc.z * ((vc.z > 0.f) ? (vc.x - zm.x) : (zp.x - vc.x));
    const float ay1 = vc.x * ((vc.x > 0.f) ? (vc.y - xm.y) : (xp.y - vc.y));
    const float ay2 = fmaf(vc.y, fabsf(vc.y), -(yp.y * fabsf(yp.y)));
    const float ay3 = vc.z * ((vc.z > 0.f) ? (vc.y - zm.y) : (zp.y - vc.y));
    const float az1 = vc.x * ((vc.x > 0.f) ? (vc.z - xm.z) : (xp.z - vc.z));
    const float az2 = vc.y * ((vc.y > 0.f) ? (vc.z - ym.z) : (yp.z - vc.z));
    const float az3 = fmaf(vc.z, fabsf(vc.z), -(zp.z * fabsf(zp.z)));
    const F3 dvv = f3(fmaf(0.5f, ax1, ax2) + ax3, fmaf(0.5f, ay2, ay1) + ay3, fmaf(0.5f, az3, az1 + az2));

    // 7. upwind velocity Laplacian (the reference's indices and ranges, as voxel_math)
    const bool nx = vc.x < 0.f, ny = vc.y < 0.f, nz = vc.z < 0.f;
    const F3 v_x2 = nx ? q.v_xpp : q.v_xmm;
    const F3 Ax = nx ? v_x2 : vc, Bx = nx ? xp : xm, Cx = nx ? vc : v_x2;
    const F3 Ay = ny ? q.v_y2 : vc, Cy = ny ? vc : q.v_y2;
    const F3 Az = nz ? q.v_zpp : vc, Cz = nz ? vc : q.v_zmm;
    const F3 lsum = add3(add3(add3(sub_twice(Ax, Bx), Cx), add3(sub_twice(Ay, ym), Cy)), add3(sub_twice(Az, zm), Cz));
    const F3 lt = lap ? lsum : f3(0.f, 0.f, 0.f);  // -(the Laplacian term)

    // 8. velocity update: v - alpha ((m1t v + (Dv)v) + tau lsum + qv nU)
    const float qv = a.m2 * __builtin_amdgcn_rcpf(fabsf(rho_n) + 1e-5f);
    F3 u = f3(fmaf(a.m1t, vc.x, dvv.x), fmaf(a.m1t, vc.y, dvv.y), fmaf(a.m1t, vc.z, dvv.z));
    u = f3(fmaf(a.tau, lt.x, u.x), fmaf(a.tau, lt.y, u.y), fmaf(a.tau, lt.z, u.z));
    u = f3(fmaf(qv, nU.x, u.x), fmaf(qv, nU.y, u.y), fmaf(qv, nU.z, u.z));
    vn = f3(fmaf(-a.alpha, u.x, vc.x), fmaf(-a.alpha, u.y, vc.y), fmaf(-a.alpha, u.z, vc.z));

    // 10. psi_inv -= alpha * (J_inv v_new)
    const F3 pc = q.pc;
    const F3 Jv = f3(fmaf(jz.x, vn.z, fmaf(jy.x, vn.y, jx.x * vn.x)), fmaf(jz.y, vn.z, fmaf(jy.y, vn.y, jx.y * vn.x)),
                     fmaf(jz.z, vn.z, fmaf(jy.z, vn.y, jx.z * vn.x)));
    pinv_n = f3(fmaf(-a.alpha, Jv.x, pc.x), fmaf(-a.alpha, Jv.y, pc.y), fmaf(-a.alpha, Jv.z, pc.z));
}

// One voxel of the z-march: every lane of the wave runs it (the x-neighbours are exchanged between
// lanes); `store` is false on halo lanes and on the voxel past the chunk.  The taps of L and T come
// from the stage.  A voxel whose taps are not all there (a displacement of a voxel or more) is
// reported (the return value) and its lane recomputes it from memory after the march, storing the
// result over what the march stored for it (computed with clamped taps; the same thread writes the
// same address later in program order).  The march itself thus has no branch and no register
// pressure from that rare path, and its stores need no per-voxel predicate.
template <int ST, int AR = ARITH_REF>
__device__ __forceinline__ bool voxel_a(const IterArgs &a, const PassAView &w, const PlaneNb &n,
                                        const ColumnWindow &c, int i, int z, bool lap, bool store) {
    VoxelIn q;
    // x-neighbours of rho, v, psi_inv, S from the adjacent lanes, y-neighbours from the LDS ring,
    // z-neighbours from the column window
    q.rc = c.r_0;
    q.r_xp = lane_next(c.r_0);
    q.r_xm = lane_prev(c.r_0);
    q.r_yp = n.r_yp;
    q.r_ym = n.r_ym;
    q.r_zp = c.r_p1;
    q.r_zm = c.r_m1;
    q.vc = c.v_0;
    q.v_xp = lane_next3(c.v_0);
    q.v_xm = lane_prev3(c.v_0);
    q.v_xpp = lane_next3(q.v_xp);
    q.v_xmm = lane_prev3(q.v_xm);
    q.v_yp = n.v_yp;
    q.v_ym = n.v_ym;
    q.v_y2 = n.v_y2;
    q.v_zp = c.v_p1;
    q.v_zm = c.v_m1;
    q.v_zpp = c.v_p2;
    q.v_zmm = c.v_m2;
    q.pc = c.p_0;
    q.p_xp = lane_next3(c.p_0);
    q.p_xm = lane_prev3(c.p_0);
    q.p_yp = n.p_yp;
    q.p_ym = n.p_ym;
    q.p_zp = c.p_p1;
    q.p_zm = c.p_m1;
    q.sc = c.s_0;
    q.s_xp = lane_next(c.s_0);
    q.s_xm = lane_prev(c.s_0);
    q.s_yp = n.s_yp;
    q.s_ym = n.s_ym;
    q.s_zp = c.s_p1;
    q.s_zm = c.s_m1;
    const Corner k = corner(w.d, c.p_0.x, c.p_0.y, c.p_0.z);
    const bool in = taps_in_stage(w, k, z);
    lt_stage_taps(w, k, z, q.Lo, q.Tv);
    float rho_n;
    F3 vn, pn;
    if constexpr (AR == ARITH_FMA) voxel_math_fma(a, q, lap, rho_n, vn, pn);
    else voxel_math<ST>(a, q, lap, rho_n, vn, pn);
    // a voxel outside the chunk (the odd step of the last trip) stores at an offset past every
    // buffer's range, which the buffer unit drops
    vr_store<ST>(w.wvr, i, store, vn, rho_n);
    bst3_nt(w.wp, store ? 12u * i : NO_STORE, pn);
    return store && !in;
}

// The same voxel from memory alone (the recovery after the march): the old state's stencil
// neighbours, L at the 8 corners of psi_inv from psi, T gathered; the same arithmetic.
template <int ST, int AR = ARITH_REF>
__device__ __forceinline__ void voxel_a_memory(const IterArgs &a, const PassAView &w, int i, int z, bool lap) {
    const int X = w.X, XY = w.XY;
    VoxelIn q;
    const VR<ST> r0 = vr_load<ST>(w.rvr, i);
    q.vc = vr_v<ST>(r0);
    q.rc = vr_r<ST>(r0);
    const VR<ST> rxp = vr_load<ST>(w.rvr, i + 1), rxm = vr_load<ST>(w.rvr, i - 1), ryp = vr_load<ST>(w.rvr, i + X),
                 rym = vr_load<ST>(w.rvr, i - X), rzp = vr_load<ST>(w.rvr, i + XY), rzm = vr_load<ST>(w.rvr, i - XY);
    q.v_xp = vr_v<ST>(rxp);
    q.r_xp = vr_r<ST>(rxp);
    q.v_xm = vr_v<ST>(rxm);
    q.r_xm = vr_r<ST>(rxm);
    q.v_yp = vr_v<ST>(ryp);
    q.r_yp = vr_r<ST>(ryp);
    q.v_ym = vr_v<ST>(rym);
    q.r_ym = vr_r<ST>(rym);
    q.v_zp = vr_v<ST>(rzp);
    q.r_zp = vr_r<ST>(rzp);
    q.v_zm = vr_v<ST>(rzm);
    q.r_zm = vr_r<ST>(rzm);
    q.v_xpp = vr_v<ST>(v_load<ST>(w.rvr, i + 2));
    q.v_xmm = vr_v<ST>(v_load<ST>(w.rvr, i - 2));
    q.v_y2 = vr_v<ST>(v_load<ST>(w.rvr, q.vc.y < 0.f ? i + 2 * X : i - 2 * X));
    q.v_zpp = vr_v<ST>(v_load<ST>(w.rvr, i + 2 * XY));
    q.v_zmm = vr_v<ST>(v_load<ST>(w.rvr, i - 2 * XY));
    q.pc = bld3(w.rp, 12u * i);
    q.p_xp = bld3(w.rp, 12u * (i + 1));
    q.p_xm = bld3(w.rp, 12u * (i - 1));
    q.p_yp = bld3(w.rp, 12u * (i + X));
    q.p_ym = bld3(w.rp, 12u * (i - X));
    q.p_zp = bld3(w.rp, 12u * (i + XY));
    q.p_zm = bld3(w.rp, 12u * (i - XY));
    q.sc = bld1(w.rs, 4u * i);
    q.s_xp = bld1(w.rs, 4u * (i + 1));
    q.s_xm = bld1(w.rs, 4u * (i - 1));
    q.s_yp = bld1(w.rs, 4u * (i + X));
    q.s_ym = bld1(w.rs, 4u * (i - X));
    q.s_zp = bld1(w.rs, 4u * (i + XY));
    q.s_zm = bld1(w.rs, 4u * (i - XY));
    lt_memory_taps(w, corner(w.d, q.pc.x, q.pc.y, q.pc.z), q.Lo, q.Tv);
    float rho_n;
    F3 vn, pn;
    if constexpr (AR == ARITH_FMA) voxel_math_fma(a, q, lap, rho_n, vn, pn);
    else voxel_math<ST>(a, q, lap, rho_n, vn, pn);
    vr_store<ST>(w.wvr, i, true, vn, rho_n);
    bst3_nt(w.wp, 12u * i, pn);
}

// the window entries entering at the next step, loaded at the start of a step (before its stores,
// which the compiler cannot move loads across) so that the slide at its end does not wait
template <int ST> struct WindowIn {
    VR<ST> vr;  // (v, rho) at z+3
    F3 p;       // psi_inv at z+3
    float s;    // S at z+2
};

template <int ST> __device__ __forceinline__ WindowIn<ST> window_in(const PassAView &w, int i, int XY) {
    WindowIn<ST> n;
    n.vr = vr_load<ST>(w.rvr, i + 3 * XY);
    n.p = bld3(w.rp, 12u * (i + 3 * XY));
    n.s = bld1(w.rs, 4u * (i + 2 * XY));
    return n;
}

// The newest (v, rho) entry: FP32 values enter the window at once; a 16-bit record is kept as
// loaded (pend) and unpacked at the start of the next step (take), a whole step after its load, so
// that the unpacking at the end of a step does not wait for a load issued in that step.
template <int ST> __device__ __forceinline__ void slide(ColumnWindow &c, const WindowIn<ST> &n, VR<ST> &pend) {
    c.v_m2 = c.v_m1;
    c.v_m1 = c.v_0;
    c.v_0 = c.v_p1;
    c.v_p1 = c.v_p2;
    if constexpr (ST == ST_F32) c.v_p2 = n.vr.v;
    c.r_m1 = c.r_0;
    c.r_0 = c.r_p1;
    c.r_p1 = c.r_p2;
    if constexpr (ST == ST_F32) c.r_p2 = n.vr.r;
    else pend = n.vr;
    c.p_m1 = c.p_0;
    c.p_0 = c.p_p1;
    c.p_p1 = c.p_p2;
    c.p_p2 = n.p;
    c.s_m1 = c.s_0;
    c.s_0 = c.s_p1;
    c.s_p1 = n.s;
}

template <int ST> __device__ __forceinline__ void take(ColumnWindow &c, const VR<ST> &pend) {
    if constexpr (ST != ST_F32) {
        c.v_p2 = vr_v<ST>(pend);
        c.r_p2 = vr_r<ST>(pend);
    }
}

// Scheduling groups of one half-step: a few loads first (the window entries, whose moves close the
// step), then one load after every VALU_PER_LOAD VALU instructions.
#define VALU_PER_LOAD 11
#define SG_LOAD() __builtin_amdgcn_sched_group_barrier(0x0020, 1, 0)
#define SG_VALU() __builtin_amdgcn_sched_group_barrier(0x0002, VALU_PER_LOAD, 0)
#define SG_PAIR() SG_LOAD(); SG_VALU()
#define SG_PAIR4() SG_PAIR(); SG_PAIR(); SG_PAIR(); SG_PAIR()
#define INTERLEAVE_LOADS() SG_PAIR4(); SG_PAIR4(); SG_PAIR4(); SG_PAIR4(); SG_PAIR4(); SG_PAIR4(); SG_PAIR4(); \
    SG_PAIR4(); SG_PAIR4(); SG_PAIR(); SG_PAIR()

template <int ST, int AR = ARITH_REF> __device__ __forceinline__ void pass_a(const IterArgs &a, int blk) {
    static_assert(AR == ARITH_REF || ST == ST_F32, "the contracted arithmetic is built for FP32 storage");
    // The block's waves step through z together (a bare s_barrier per half-step, no memory
    // fence), so that the rows one wave loads are still in L1 when the adjacent rows' waves read
    // them as y-neighbours, and adjacent rows gather from overlapping regions at the same time.
    // Every wave therefore runs the whole march: the stop test is taken once per block, and rows
    // past the interior compute without storing.
    __shared__ int stop_block;
    if (threadIdx.x < 64) {  // the whole first wave: the slot maximum is a wave reduction
        const bool st = should_stop(a, blk == 0 && threadIdx.x == 0);
        if (threadIdx.x == 0) stop_block = st;
    }
    __syncthreads();
    if (stop_block) return;  // uniform over the block
    const Dims d = a.d;
    const int N = d.X * d.Y * d.Z, XY = d.X * d.Y;
    const TileCoord tc = tile_of(a, blk);
    const int x = tc.x, y = tc.y;
    const bool own = tc.own && y <= d.Y - 2;
    PassAView w;
    w.d = d;
    w.rvr = vr_buf<ST>(a.v_o, a.rho_o, N);
    w.wvr = vr_buf<ST>(a.v_n, a.rho_n, N);
    w.rp = rsrc(a.pinv_o, 12 * N);
    w.rs = rsrc(a.S, 4 * N);
    w.rt = rsrc(a.T, 4 * N);
    w.rq = rsrc(a.psi, 12 * N);
    w.wp = rsrc(a.pinv_n, 12 * N);
    w.X = d.X;
    w.XY = XY;
    w.X12 = 12u * d.X;
    w.XY12 = 12u * XY;
    w.X4 = 4u * d.X;
    w.XY4 = 4u * XY;
    const bool lap_xy = x >= 2 && x <= d.X - 3 && y >= 2 && y <= d.Y - 2;
    // The stages.  This wave holds psi rows j0 = wave + 1 (the block's rows) and j1 of the stages
    // (rows 9, 10, 11, 0 from waves 0..3; waves 4..7 load rows 9 and 10 as well, so that the loads
    // need no branch) and computes L/T for its own row and, on waves 0 and 1, for rows 9 and 10 (the
    // prologue computes every row it loads, writing the same values where two waves hold one row).
    // Waves 4..7 carry the y-stage's halo rows instead.
    __shared__ LTStage lt_stage;
    __shared__ PsiStage psi_stage;
    __shared__ YStage y_stage;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    w.lt = &lt_stage;
    w.ps = &psi_stage;
    w.ys = &y_stage;
    w.xb = x - lane;
    w.yb = y - wave - 1;
    const int j0 = wave + 1;
    const int j1 = wave < 3 ? TY_A + 1 + wave : wave == 3 ? 0 : TY_A + 1 + (wave & 1);
    const int py0 = w.yb - 1 + j0, py1 = w.yb - 1 + j1;  // the y of both rows (psi and T)
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    const bool has_l1 = wave_u < 2, has_p1 = wave_u < 4;
    const int zs = tc.zs;
    // prologue: L/T planes zs-1, zs, zs+1 from psi planes zs-2..zs+2, each psi plane written to the
    // ring before the L plane whose y-neighbours it holds; psi planes zs+1..zs+3 of both rows stay
    // in registers for the first step, with T of plane zs+2
    PsiWin w0, w1;
    {
        const F3 a0 = psi_row_load(w, x, py0, zs - 2), a1 = psi_row_load(w, x, py1, zs - 2);
        const F3 b0 = psi_row_load(w, x, py0, zs - 1), b1 = psi_row_load(w, x, py1, zs - 1);
        const F3 c0 = psi_row_load(w, x, py0, zs), c1 = psi_row_load(w, x, py1, zs);
        psi_store(w, zs - 1, j0, lane, b0);
        psi_store(w, zs - 1, j1, lane, b1);
        __syncthreads();
        float t0 = t_row_load(w, x, py0, zs - 1), t1 = t_row_load(w, x, py1, zs - 1);
        lt_compute(w, zs - 1, j0, lane, t0, psi_win(a0, b0, c0));
        lt_compute(w, zs - 1, j1, lane, t1, psi_win(a1, b1, c1));
        const F3 d0 = psi_row_load(w, x, py0, zs + 1), d1 = psi_row_load(w, x, py1, zs + 1);
        psi_store(w, zs, j0, lane, c0);
        psi_store(w, zs, j1, lane, c1);
        __syncthreads();
        t0 = t_row_load(w, x, py0, zs);
        t1 = t_row_load(w, x, py1, zs);
        lt_compute(w, zs, j0, lane, t0, psi_win(b0, c0, d0));
        lt_compute(w, zs, j1, lane, t1, psi_win(b1, c1, d1));
        const F3 e0 = psi_row_load(w, x, py0, zs + 2), e1 = psi_row_load(w, x, py1, zs + 2);
        psi_store(w, zs + 1, j0, lane, d0);
        psi_store(w, zs + 1, j1, lane, d1);
        __syncthreads();
        t0 = t_row_load(w, x, py0, zs + 1);
        t1 = t_row_load(w, x, py1, zs + 1);
        lt_compute(w, zs + 1, j0, lane, t0, psi_win(c0, d0, e0));
        lt_compute(w, zs + 1, j1, lane, t1, psi_win(c1, d1, e1));
        const F3 f0 = psi_row_load(w, x, py0, zs + 3), f1 = psi_row_load(w, x, py1, zs + 3);
        psi_store(w, zs + 2, j0, lane, e0);
        psi_store(w, zs + 2, j1, lane, e1);
        w0 = psi_win(d0, e0, f0);
        w1 = psi_win(d1, e1, f1);
    }
    float tn0 = t_row_load(w, x, py0, zs + 2), tn1 = t_row_load(w, x, py1, zs + 2);

    int i = x + d.X * y + XY * tc.zs;
    ColumnWindow c;
    VR<ST> pend;  // the 16-bit forms' newest (v, rho) record, unpacked at the start of a step
    c.v_m2 = vr_v<ST>(v_load<ST>(w.rvr, i - 2 * XY));
    {
        const VR<ST> q_m1 = vr_load<ST>(w.rvr, i - XY), q_0 = vr_load<ST>(w.rvr, i),
                     q_p1 = vr_load<ST>(w.rvr, i + XY), q_p2 = vr_load<ST>(w.rvr, i + 2 * XY);
        if constexpr (ST != ST_F32) pend = q_p2;
        c.v_m1 = vr_v<ST>(q_m1);
        c.v_0 = vr_v<ST>(q_0);
        c.v_p1 = vr_v<ST>(q_p1);
        c.v_p2 = vr_v<ST>(q_p2);
        c.r_m1 = vr_r<ST>(q_m1);
        c.r_0 = vr_r<ST>(q_0);
        c.r_p1 = vr_r<ST>(q_p1);
        c.r_p2 = vr_r<ST>(q_p2);
    }
    c.p_m1 = bld3(w.rp, 12u * (i - XY));
    c.p_0 = bld3(w.rp, 12u * i);
    c.p_p1 = bld3(w.rp, 12u * (i + XY));
    c.p_p2 = bld3(w.rp, 12u * (i + 2 * XY));
    c.s_m1 = bld1(w.rs, 4u * (i - XY));
    c.s_0 = bld1(w.rs, 4u * i);
    c.s_p1 = bld1(w.rs, 4u * (i + XY));

    // the y-neighbour stage: this wave's row is jo; the halo row jh (y-1, y+8, y-2 or y+9 of the
    // block) is loaded by waves 4-7 (has_h, uniform over the wave); plane zs now (every wave loads
    // a halo row here), the halo of plane zs+1 into registers
    const int jo = wave + 2, jh = (wave & 3) == 0 ? 1 : (wave & 3) == 1 ? TY_A + 2 : (wave & 3) == 2 ? 0 : TY_A + 3;
    const bool has_h = __builtin_amdgcn_readfirstlane(wave) >= 4;
    const int ih = x + d.X * (w.yb - 1 + jh);
    {
        const YRowOf<ST> h0 = yrow_load<ST>(w, ih + XY * zs);
        ys_store(w, zs, jo, lane, c.v_0, c.r_0, c.p_0, c.s_0);
        ys_store_row<ST>(w, zs, jh, lane, h0);
    }
    YRowOf<ST> hA = yrow_load<ST>(w, ih + XY * (zs + 1)), hB = hA;
    // drain the prologue's loads, so that the loop's memory-counter waits are set by the loop's
    // own steady state and not by the prologue's last loads; the stage's first planes are written
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    // Two steps per trip, alternating the halo-row registers between hA and hB.  The trip has no
    // branch: the loads run ahead of the chunk unconditionally (range-checked buffer loads; the
    // values are never stored) and the odd step of the last trip drops its stores, so the memory
    // counter waits in the loop are the exact distances of one step of pipelining.
    const int nz = tc.ze - tc.zs;
    int fb_lo = INT_MAX, fb_hi = -1;  // the first and last plane of this lane's voxels left to the recovery
    for (int s = 0, z = tc.zs; s < nz; s += 2, z += 2, i += 2 * XY) {
        // (each half-step interleaves the next plane's loads with the current plane's arithmetic,
        // so that a wave keeps computing while its loads queue for the memory pipeline.  It first
        // computes L/T of plane z+2 from psi planes z+1..z+3 (registers; the y-neighbours of plane
        // z+2 from the ring), then loads psi of plane z+4 (it enters the registers at the end) and
        // T of plane z+3 (used by the next half-step), and writes psi of plane z+3 into the ring;
        // the LDS writes complete before the barrier that precedes their readers.)
        take<ST>(c, pend);
        const PlaneNb n0 = ys_read(w, z, jo, lane, c.v_0.y < 0.f);
        lt_compute(w, z + 2, j0, lane, tn0, w0);
        if (has_l1) lt_compute(w, z + 2, j1, lane, tn1, w1);
        const F3 qa0 = psi_row_load(w, x, py0, z + 4), qa1 = psi_row_load(w, x, py1, z + 4);
        tn0 = t_row_load(w, x, py0, z + 3);
        tn1 = t_row_load(w, x, py1, z + 3);
        if (has_h) hB = yrow_load<ST>(w, ih + XY * (z + 2));
        const WindowIn<ST> in0 = window_in<ST>(w, i, XY);
        const bool f0 = voxel_a<ST, AR>(a, w, n0, c, i, z, lap_xy && z >= 2 && z <= d.Z - 3, own);
        fb_lo = min(fb_lo, f0 ? z : INT_MAX);
        fb_hi = f0 ? z : fb_hi;
        psi_store(w, z + 3, j0, lane, w0.p);
        if (has_p1) psi_store(w, z + 3, j1, lane, w1.p);
        ys_store(w, z + 1, jo, lane, c.v_p1, c.r_p1, c.p_p1, c.s_p1);
        if (has_h) ys_store_row<ST>(w, z + 1, jh, lane, hA);
        w0 = psi_win(w0.c, w0.p, qa0);
        w1 = psi_win(w1.c, w1.p, qa1);
        slide<ST>(c, in0, pend);
        INTERLEAVE_LOADS();
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        take<ST>(c, pend);
        const PlaneNb n1 = ys_read(w, z + 1, jo, lane, c.v_0.y < 0.f);
        lt_compute(w, z + 3, j0, lane, tn0, w0);
        if (has_l1) lt_compute(w, z + 3, j1, lane, tn1, w1);
        const F3 qb0 = psi_row_load(w, x, py0, z + 5), qb1 = psi_row_load(w, x, py1, z + 5);
        tn0 = t_row_load(w, x, py0, z + 4);
        tn1 = t_row_load(w, x, py1, z + 4);
        if (has_h) hA = yrow_load<ST>(w, ih + XY * (z + 3));
        const WindowIn<ST> in1 = window_in<ST>(w, i + XY, XY);
        const bool f1 = voxel_a<ST, AR>(a, w, n1, c, i + XY, z + 1, lap_xy && z + 1 >= 2 && z + 1 <= d.Z - 3,
                                    own && s + 1 < nz);
        fb_lo = min(fb_lo, f1 ? z + 1 : INT_MAX);
        fb_hi = f1 ? z + 1 : fb_hi;
        psi_store(w, z + 4, j0, lane, w0.p);
        if (has_p1) psi_store(w, z + 4, j1, lane, w1.p);
        ys_store(w, z + 2, jo, lane, c.v_p1, c.r_p1, c.p_p1, c.s_p1);
        if (has_h) ys_store_row<ST>(w, z + 2, jh, lane, hB);
        w0 = psi_win(w0.c, w0.p, qb0);
        w1 = psi_win(w1.c, w1.p, qb1);
        slide<ST>(c, in1, pend);
        INTERLEAVE_LOADS();
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
    }
    // The recovery: this lane's voxels whose taps left the stage, recomputed from memory (the old
    // state, which this pass does not write) and stored over the march's values for them.  They are
    // rare (displacements of a voxel or more).
    for (int zz = fb_lo; zz <= fb_hi; ++zz) {
        const int ii = x + d.X * y + XY * zz;
        const F3 p = bld3(w.rp, 12u * ii);
        if (!taps_in_stage(w, corner(d, p.x, p.y, p.z), zz))
            voxel_a_memory<ST, AR>(a, w, ii, zz, lap_xy && zz >= 2 && zz <= d.Z - 3);
    }
}

// --------------------------------------------------------------------------------------------
// pass B
// --------------------------------------------------------------------------------------------

// v_new of the planes around the one being computed, staged in LDS for the block: a ring of 4
// planes (z-1, z, z+1 are read while z+2 is written), TY_B + 2 rows (the block's rows and one on
// either side), 64 columns (the lanes of a wave), (v.x, v.y, v.z, unused) per voxel.  The 8 taps
// of a voxel whose psi lies within one voxel of it are all in the stage.
constexpr int VS_ROWS = TY_B + 2, VS_PLANES = 4;
struct VStage {
    float4 v[VS_PLANES][VS_ROWS][WAVE];
};

// One thread per (x, y) column of a 64 x TY_B tile (XS_B computed columns), the block's waves
// marching `zc` z-slices together: psi += alpha * v_new o psi (solver.cu:230-247), the trilinear
// taps read from the stage, and the running max of the update norm.
template <int ST> __device__ __forceinline__ void pass_b(const IterArgs &a, int blk) {
    __shared__ VStage vs;
    __shared__ unsigned long long red[TY_B];
    if (__builtin_nontemporal_load(&a.ctl->stopped)) return;  // set before this launch: uniform
    const Dims d = a.d;
    const int N = d.X * d.Y * d.Z, XY = d.X * d.Y;
    const TileCoord tc = tile_of(a, blk);
    const int x = tc.x, y = tc.y;
    const bool own = tc.own && y <= d.Y - 2;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const int xb = x - lane, yb = y - wave - 1;
    const VRBuf vb = vr_buf<ST>(a.v_n, a.rho_n, N);
    const Rsrc rq = rsrc(a.psi, 12 * N), wq = rq;
    // this wave stages rows `wave` and (waves 0 and 1) `wave + TY_B`; the others stage the last row
    // as well, writing the same values, so that no wave branches
    const int row0 = wave, row1 = min(wave + TY_B, VS_ROWS - 1);
    const int sy0 = yb + row0, sy1 = yb + row1;
    auto stage_load = [&](int yy, int zz) { return vr_v<ST>(v_load<ST>(vb, x + d.X * yy + XY * zz)); };
    auto stage_store = [&](int zz, int row, F3 v) {
        vs.v[zz & (VS_PLANES - 1)][row][lane] = make_float4(v.x, v.y, v.z, 0.f);
    };
    for (int zz = tc.zs - 1; zz <= tc.zs + 1; ++zz) {
        const F3 v0 = stage_load(sy0, zz), v1 = stage_load(sy1, zz);
        stage_store(zz, row0, v0);
        stage_store(zz, row1, v1);
    }
    const unsigned o0 = 12u * (x + d.X * y + XY * tc.zs);
    F3 pn = bld3(rq, o0);  // psi of the next plane to compute
    __syncthreads();
    unsigned long long pack = 0;
    for (int z = tc.zs; z < tc.ze; ++z) {
        const unsigned o = 12u * (x + d.X * y + XY * z);
        const F3 v0 = stage_load(sy0, z + 2), v1 = stage_load(sy1, z + 2);
        const F3 p = pn;
        pn = bld3(rq, o + 12u * XY);
        const Corner k = corner(d, p.x, p.y, p.z);
        int cx = k.gx - xb, ry = k.gy - yb;
        const int dz = k.gz - z;
        const bool in = cx >= 0 && cx <= WAVE - 2 && ry >= 0 && ry <= VS_ROWS - 2 && (dz == -1 || dz == 0);
        cx = min(max(cx, 0), WAVE - 2);
        ry = min(max(ry, 0), VS_ROWS - 2);
        const int zg = dz == -1 ? z - 1 : z;
        const int s0 = zg & (VS_PLANES - 1), s1 = (zg + 1) & (VS_PLANES - 1);
        const F3 f000 = xyz_of(vs.v[s0][ry][cx]), f001 = xyz_of(vs.v[s1][ry][cx]), f010 = xyz_of(vs.v[s0][ry + 1][cx]),
                 f011 = xyz_of(vs.v[s1][ry + 1][cx]), f100 = xyz_of(vs.v[s0][ry][cx + 1]),
                 f101 = xyz_of(vs.v[s1][ry][cx + 1]), f110 = xyz_of(vs.v[s0][ry + 1][cx + 1]),
                 f111 = xyz_of(vs.v[s1][ry + 1][cx + 1]);
        F3 w = f3(tri(f111.x, f110.x, f101.x, f100.x, f011.x, f010.x, f001.x, f000.x, k.a, k.b, k.c),
                  tri(f111.y, f110.y, f101.y, f100.y, f011.y, f010.y, f001.y, f000.y, k.a, k.b, k.c),
                  tri(f111.z, f110.z, f101.z, f100.z, f011.z, f010.z, f001.z, f000.z, k.a, k.b, k.c));
        if (own && !in) {  // a voxel displaced out of the stage: gather v_new from memory
            const int g = (int)corner_index(d, k);
            const F3 g111 = vr_v<ST>(v_load<ST>(vb, g + 1 + d.X + XY)), g110 = vr_v<ST>(v_load<ST>(vb, g + 1 + d.X)),
                     g101 = vr_v<ST>(v_load<ST>(vb, g + 1 + XY)), g100 = vr_v<ST>(v_load<ST>(vb, g + 1)),
                     g011 = vr_v<ST>(v_load<ST>(vb, g + d.X + XY)), g010 = vr_v<ST>(v_load<ST>(vb, g + d.X)),
                     g001 = vr_v<ST>(v_load<ST>(vb, g + XY)), g000 = vr_v<ST>(v_load<ST>(vb, g));
            w = f3(tri(g111.x, g110.x, g101.x, g100.x, g011.x, g010.x, g001.x, g000.x, k.a, k.b, k.c),
                   tri(g111.y, g110.y, g101.y, g100.y, g011.y, g010.y, g001.y, g000.y, k.a, k.b, k.c),
                   tri(g111.z, g110.z, g101.z, g100.z, g011.z, g010.z, g001.z, g000.z, k.a, k.b, k.c));
        }
        const F3 u = mul3(a.alpha, w);
        bst3(wq, own ? o : NO_STORE, f3(p.x + u.x, p.y + u.y, p.z + u.z));
        // norm with a round-down sqrt (cuda_utils.hpp:251-253); NaN never wins (strict >,
        // reductor.cu:445-456)
        const float nr = sqrt_rd((u.x * u.x + u.y * u.y) + u.z * u.z);
        const int i = x + d.X * y + XY * z;
        const unsigned long long pk =
            ((unsigned long long)__float_as_uint(nr) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)i);
        if (own && !isnan(nr) && pk > pack) pack = pk;
        stage_store(z + 2, row0, v0);
        stage_store(z + 2, row1, v1);
        __syncthreads();
    }
    // max (norm, lowest index) over the block, then one atomic per block into one of the
    // iteration's NSLOT slots
    pack = wave_max_u64(pack);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = pack;
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < TY_B; ++w) pack = red[w] > pack ? red[w] : pack;
        unsigned long long *slot = &a.maxpack[(size_t)(a.iter - 1) * NSLOT + (blk % NSLOT)];
        if (pack > __builtin_nontemporal_load(slot)) atomicMax(slot, pack);
    }
}

// --------------------------------------------------------------------------------------------
// entry points: one volume, or a batch of volumes in one launch per pass
// --------------------------------------------------------------------------------------------

template <int ST, int AR = ARITH_REF>
__global__ __launch_bounds__(WAVE * TY_A) void k_pass_a(IterArgs a) { pass_a<ST, AR>(a, blockIdx.x); }
template <int ST> __global__ __launch_bounds__(WAVE * TY_B) void k_pass_b(IterArgs a) { pass_b<ST>(a, blockIdx.x); }

// The volume a block of a batched launch works on: the last q with starts[q] <= b (starts has
// npairs + 1 entries, increasing).
__device__ __forceinline__ int batch_volume(const int *__restrict__ starts, int npairs, int b) {
    int lo = 0, hi = npairs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The arguments of volume q for pass `pass` at this iteration: the launch-time arguments (buffer
// parity 0) with the old/new buffers exchanged on odd parities and the iteration's values set.
__device__ __forceinline__ IterArgs batch_args(const IterArgs *__restrict__ base, int q, int pass, const BatchIter &it) {
    IterArgs a = base[3 * q + pass];
    if (it.cur) {
        const float *r = a.rho_o, *v = a.v_o, *p = a.pinv_o;
        a.rho_o = a.rho_n;
        a.v_o = a.v_n;
        a.pinv_o = a.pinv_n;
        a.rho_n = const_cast<float *>(r);
        a.v_n = const_cast<float *>(v);
        a.pinv_n = const_cast<float *>(p);
    }
    a.iter = it.iter;
    a.m1t = it.m1t;
    a.m2 = it.m2;
    a.check_energy = it.check_energy;
    if (it.no_stop) a.check_conv = 0;
    return a;
}

#define BATCH_ENTRY(pass_fn, pass_idx)                                                            \
    const int q = batch_volume(starts, npairs, blockIdx.x);                                      \
    const IterArgs a = batch_args(base, q, pass_idx, it);                                        \
    const int blk = blockIdx.x - starts[q];                                                      \
    if (blk >= a.nblocks) return; /* padding up to a multiple of 8 blocks: the whole block */   \
    pass_fn(a, blk)

template <int ST>
__global__ __launch_bounds__(WAVE * TY_A) void k_pass_a_batch(const IterArgs *base, const int *starts, int npairs,
                                                              BatchIter it) {
    BATCH_ENTRY(pass_a<ST>, 1);
}
template <int ST>
__global__ __launch_bounds__(WAVE * TY_B) void k_pass_b_batch(const IterArgs *base, const int *starts, int npairs,
                                                              BatchIter it) {
    BATCH_ENTRY(pass_b<ST>, 2);
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

// New images into a context that keeps its state (the reference's FastReg::operator() re-initialises
// the four volumes and, in estimate_psi, the density, fastreg.cpp:58-63, solver.cu:317): T, LT, S,
// LS from the caller's arrays and rho = rho_0 in both buffers (FP32 storage; rho0/rho1 null in the
// 16-bit forms, whose records are reset by k_vr_reset_rho); v, psi and psi_inv are kept.
__global__ void k_reinit(Dims d, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab,
                         float *T, int16_t *LT, float *S, int16_t *LS, float *rho0, float *rho1, float rho_0) {
    const int n = d.X * d.Y * d.Z;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        T[i] = tgt[i];
        LT[i] = tgt_lab[i];
        S[i] = src[i];
        LS[i] = src_lab[i];
        if (rho0) {
            rho0[i] = rho_0;
            rho1[i] = rho_0;
        }
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

// Adds to cnt[label] the number of active lanes holding that label: one atomic per distinct label
// of the wave (neighbouring voxels mostly share a label, so per-lane atomics on one address would
// serialise).  `act` is the ballot of the lanes taking part; `lab` must be in range on them.
__device__ __forceinline__ void wave_count(unsigned long long act, int lab, unsigned *cnt) {
    const int lane = threadIdx.x & (WAVE - 1);
    while (act) {
        const int leader = __builtin_ctzll(act);
        const int l = __builtin_amdgcn_readlane(lab, leader);
        const unsigned long long m = __ballot(((act >> lane) & 1ull) && lab == l);
        if (lane == leader) atomicAdd(&cnt[l], (unsigned)__popcll(m));
        act &= ~m;
    }
}

// Per-label counts of the Dice coefficient (solver.cu:9-50) over x < X-1, y < Y-1, z < Z-1: voxels
// of each label in la and in lb, and voxels where both agree; labels outside [0, no_labels) are
// counted in `bad`.  Each wave takes whole x-rows (64 voxels per step); the counts go through LDS
// when the labels fit there.  A block adds into copy blockIdx % DICE_COPIES of the counts, so that
// the blocks' closing atomics spread over that many addresses per label.
__global__ void k_dice(Dims d, const int16_t *__restrict__ la, const int16_t *__restrict__ lb, int no_labels,
                       unsigned *counts) {
    __shared__ unsigned s_ov[DICE_LDS_LABELS], s_vx[DICE_LDS_LABELS];
    unsigned *overlap = counts + (size_t)(blockIdx.x % DICE_COPIES) * (2 * (size_t)no_labels + 1);
    unsigned *voxels = overlap + no_labels, *bad = overlap + 2 * no_labels;
    const bool use_lds = no_labels <= DICE_LDS_LABELS;
    if (use_lds)
        for (int l = threadIdx.x; l < no_labels; l += blockDim.x) s_ov[l] = s_vx[l] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1), wpb = blockDim.x / WAVE;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ny = d.Y - 1, nrows = (d.Y - 1) * (d.Z - 1);
    unsigned nb = 0;
    for (int r = blockIdx.x * wpb + wave; r < nrows; r += gridDim.x * wpb) {
        const int y = r % ny, z = r / ny;
        const int row = d.X * (y + d.Y * z);
        for (int x0 = 0; x0 < d.X - 1; x0 += WAVE) {
            const int x = x0 + lane;
            const bool in = x < d.X - 1;
            const int l1 = in ? la[row + x] : 0, l2 = in ? lb[row + x] : 0;
            const bool ok = in && l1 >= 0 && l1 < no_labels && l2 >= 0 && l2 < no_labels;
            nb += in && !ok;
            const unsigned long long both = __ballot(ok), same = __ballot(ok && l1 == l2);
            if (use_lds) {  // uniform
                wave_count(both, l1, s_vx);
                wave_count(both, l2, s_vx);
                wave_count(same, l1, s_ov);
            } else {
                wave_count(both, l1, voxels);
                wave_count(both, l2, voxels);
                wave_count(same, l1, overlap);
            }
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
    float mt = INFINITY, mr = INFINITY;
    // each wave takes whole interior x-rows, 64 voxels per step
    const int lane = threadIdx.x & (WAVE - 1), wpb = RED_BLOCK / WAVE;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int ny = d.Y - 2, nrows = (d.Y - 2) * (d.Z - 2);
    const int sy = d.X, sz = d.X * d.Y;
    for (int r = blockIdx.x * wpb + wave; r < nrows; r += gridDim.x * wpb) {
        const int y = 1 + r % ny, z = 1 + r / ny;
        const int row = d.X * (y + d.Y * z);
        for (int x = 1 + lane; x <= d.X - 2; x += WAVE) {
            const int i = row + x;
            F3 jx = mul3(0.5f, sub3(ld3(psi, i + 1), ld3(psi, i - 1)));
            F3 jy = mul3(0.5f, sub3(ld3(psi, i + sy), ld3(psi, i - sy)));
            F3 jz = mul3(0.5f, sub3(ld3(psi, i + sz), ld3(psi, i - sz)));
            float t = det_true(jx.x, jy.x, jz.x, jx.y, jy.y, jz.y, jx.z, jy.z, jz.z);
            float f = det_ref(jx.x, jy.x, jz.x, jx.y, jy.y, jz.y, jx.z, jy.z, jz.z);
            mt = fminf(mt, t);
            mr = fminf(mr, f);
        }
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

// the maximum of each iteration's NSLOT slots: one wave per iteration
__global__ __launch_bounds__(256) void k_slots_reduce(const unsigned long long *__restrict__ slots, int iters,
                                                      unsigned long long *__restrict__ out) {
    const int k = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (k >= iters) return;  // uniform over the wave
    const unsigned long long m = slots_max(slots + (size_t)k * NSLOT);
    if ((threadIdx.x & 63) == 0) out[k] = m;
}

// init_identity_kernel (vector_fields.cu:64-79): (x, y, z, 0) at voxel (x, y, z)
// zeros into three word arrays (n0, n1, n2 32-bit words): clears memory without memset nodes, so
// that a caller's graph capture of the loop holds kernels only
__global__ void k_zero3(unsigned *__restrict__ p0, long long n0, unsigned *__restrict__ p1, long long n1,
                        unsigned *__restrict__ p2, long long n2) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n0 + n1 + n2; k += stride) {
        if (k < n0) p0[k] = 0u;
        else if (k < n0 + n1) p1[k - n0] = 0u;
        else p2[k - n0 - n1] = 0u;
    }
}

__global__ void k_identity4(Dims d, float *__restrict__ dst4) {
    const int n = d.X * d.Y * d.Z;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = i % d.X, r = i / d.X, y = r % d.Y, z = r / d.Y;
        dst4[4 * i] = (float)x;
        dst4[4 * i + 1] = (float)y;
        dst4[4 * i + 2] = (float)z;
        dst4[4 * i + 3] = 0.f;
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
// the reference's further public functions: inverse by fixed point, field interpolation, the
// largest per-voxel energy, sum of a scalar field
// --------------------------------------------------------------------------------------------

// trilinear blend of the x, y, z of a float4 field (16-byte records) at point (px, py, pz)
__device__ __forceinline__ F3 tri_field4(Rsrc r, Dims d, float px, float py, float pz) {
    const Corner k = corner(d, px, py, pz);
    const unsigned g = 16u * corner_index(d, k), X16 = 16u * d.X, XY16 = 16u * d.X * d.Y;
    return tri_at3(r, g, 16u, X16, XY16, k);
}

// one sweep of estimate_inverse_kernel: psi_inv(p) = -trilerp(psi, psi_inv(p)), all voxels, in
// place (vector_fields.cu:108-122)
__global__ void k_estimate_inverse(Dims d, const float *__restrict__ psi4, float *__restrict__ pinv4) {
    const int n = d.X * d.Y * d.Z;
    const Rsrc r = rsrc(psi4, 16 * n);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const F3 q = tri_field4(r, d, pinv4[4 * i], pinv4[4 * i + 1], pinv4[4 * i + 2]);
        pinv4[4 * i] = -1.f * q.x;
        pinv4[4 * i + 1] = -1.f * q.y;
        pinv4[4 * i + 2] = -1.f * q.z;
        pinv4[4 * i + 3] = 0.f;
    }
}

// interpolate_gradient_kernel: out(p) = trilerp(field, psi(p)), all voxels (vector_fields.cu:194-225)
__global__ void k_interpolate_field4(Dims d, const float *__restrict__ f4, const float *__restrict__ psi4,
                                     float *__restrict__ out4) {
    const int n = d.X * d.Y * d.Z;
    const Rsrc r = rsrc(f4, 16 * n);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const F3 q = tri_field4(r, d, psi4[4 * i], psi4[4 * i + 1], psi4[4 * i + 2]);
        out4[4 * i] = q.x;
        out4[4 * i + 1] = q.y;
        out4[4 * i + 2] = q.z;
        out4[4 * i + 3] = 0.f;
    }
}

// reduce_voxel_max_energy_kernel: per voxel (T - S o psi)^2 + w_reg * ((|r0|^2 + |r1|^2) + |r2|^2)
// with r the rows of J = D(psi - id) (central differences, mirrored neighbours at the faces) and
// |r|^2 = (x*x + y*y) + z*z; the largest, lowest voxel on ties, into *slot as a packed (energy bits,
// ~index) key (reductor.cu:637-760, vector_fields.cu:537-585, cuda_utils.hpp:255-257)
__global__ __launch_bounds__(256) void k_voxel_max_energy(Dims d, const float *__restrict__ T,
                                                         const float *__restrict__ S,
                                                         const float *__restrict__ psi, float w_reg,
                                                         unsigned long long *slot) {
    const int n = d.X * d.Y * d.Z;
    const Rsrc rs = rsrc(S, 4 * n);
    unsigned long long pack = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int x = i % d.X, r = i / d.X, y = r % d.Y, z = r / d.Y;
        const F3 p = ld3(psi, i);
        const Corner k = corner(d, p.x, p.y, p.z);
        const float sp = tri_at1(rs, 4u * corner_index(d, k), 4u, 4u * d.X, 4u * d.X * d.Y, k);
        const int x1 = x == d.X - 1 ? x - 1 : x + 1, x2 = x == 0 ? x + 1 : x - 1;
        const int y1 = y == d.Y - 1 ? y - 1 : y + 1, y2 = y == 0 ? y + 1 : y - 1;
        const int z1 = z == d.Z - 1 ? z - 1 : z + 1, z2 = z == 0 ? z + 1 : z - 1;
        const F3 ax1 = sub3(ld3(psi, lin(d, x1, y, z)), f3((float)x1, (float)y, (float)z));
        const F3 ax2 = sub3(ld3(psi, lin(d, x2, y, z)), f3((float)x2, (float)y, (float)z));
        const F3 ay1 = sub3(ld3(psi, lin(d, x, y1, z)), f3((float)x, (float)y1, (float)z));
        const F3 ay2 = sub3(ld3(psi, lin(d, x, y2, z)), f3((float)x, (float)y2, (float)z));
        const F3 az1 = sub3(ld3(psi, lin(d, x, y, z1)), f3((float)x, (float)y, (float)z1));
        const F3 az2 = sub3(ld3(psi, lin(d, x, y, z2)), f3((float)x, (float)y, (float)z2));
        const F3 jx = mul3(0.5f, sub3(ax1, ax2)), jy = mul3(0.5f, sub3(ay1, ay2)), jz = mul3(0.5f, sub3(az1, az2));
        const float n0 = (jx.x * jx.x + jy.x * jy.x) + jz.x * jz.x;
        const float n1 = (jx.y * jx.y + jy.y * jy.y) + jz.y * jz.y;
        const float n2 = (jx.z * jx.z + jy.z * jy.z) + jz.z * jz.z;
        const float df = T[i] - sp;
        const float e = df * df + w_reg * ((n0 + n1) + n2);
        const unsigned long long pk =
            ((unsigned long long)__float_as_uint(e) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)i);
        if (!isnan(e) && pk > pack) pack = pk;
    }
    pack = wave_max_u64(pack);
    __shared__ unsigned long long red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = pack;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) pack = red[w] > pack ? red[w] : pack;
        atomicMax(slot, pack);
    }
}

// sum of a float array in FP64: per-block partial sums, then one block over the partials
__global__ __launch_bounds__(256) void k_sum_partials(const float *__restrict__ a, long long n, double *partials) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        acc += (double)a[i];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void k_sum_final(const double *partials, int np, double *out) {
    __shared__ double sh[4];
    double acc = 0.0;
    for (int j = threadIdx.x; j < np; j += 256) acc += partials[j];
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = acc;
}

// (v, rho) records of the 16-bit storage forms: initial state, set from and read into FP32 arrays
template <int ST> __global__ void k_vr_init(unsigned *r0, unsigned *r1, int n, float rho_0) {
    const unsigned hi = h_of<ST>(rho_0) << 16;  // v = 0, rho = rho_0 (rounded)
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        r0[2 * i] = 0u;
        r0[2 * i + 1] = hi;
        r1[2 * i] = 0u;
        r1[2 * i + 1] = hi;
    }
}

// replace v (float4 per voxel) and/or rho (either may be null) in both buffers of records
// rho = rho_0 in every (v, rho) record of both buffers, v kept
template <int ST> __global__ void k_vr_reset_rho(unsigned *r0, unsigned *r1, int n, float rho_0) {
    const unsigned h = h_of<ST>(rho_0) << 16;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        r0[2 * i + 1] = (r0[2 * i + 1] & 0xFFFFu) | h;
        r1[2 * i + 1] = (r1[2 * i + 1] & 0xFFFFu) | h;
    }
}

template <int ST> __global__ void k_vr_set(unsigned *r0, unsigned *r1, const float *v4, const float *rho, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned lo = r0[2 * i], hi = r0[2 * i + 1];
        if (v4) {
            lo = h_of<ST>(v4[4 * i]) | (h_of<ST>(v4[4 * i + 1]) << 16);
            hi = (hi & 0xFFFF0000u) | h_of<ST>(v4[4 * i + 2]);
        }
        if (rho) hi = (hi & 0xFFFFu) | (h_of<ST>(rho[i]) << 16);
        r0[2 * i] = lo;
        r0[2 * i + 1] = hi;
        r1[2 * i] = lo;
        r1[2 * i + 1] = hi;
    }
}

template <int ST> __global__ void k_vr_get(const unsigned *r, float *v4, float *rho, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const unsigned lo = r[2 * i], hi = r[2 * i + 1];
        if (v4) {
            v4[4 * i] = f_of<ST>(lo & 0xFFFFu);
            v4[4 * i + 1] = f_of<ST>(lo >> 16);
            v4[4 * i + 2] = f_of<ST>(hi & 0xFFFFu);
            v4[4 * i + 3] = 0.f;
        }
        if (rho) rho[i] = f_of<ST>(hi >> 16);
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

// the tile grid of one launch over the interior: xs x ty columns by zc slices per block
static IterArgs chunked(IterArgs a, int zc, int xs, int ty) {
    a.zc = zc < 1 ? 1 : zc;
    a.xs = xs;
    a.ty = ty;
    a.nbx = (a.d.X - 2 + xs - 1) / xs;
    a.nby = (a.d.Y - 2 + ty - 1) / ty;
    a.nbz = (a.d.Z - 2 + a.zc - 1) / a.zc;
    a.nblocks = a.nbx * a.nby * a.nbz;
    return a;
}


template <int ST, int AR = ARITH_REF> static void pass_a_kernel(const IterArgs &a, hipStream_t s) {
    hipLaunchKernelGGL((k_pass_a<ST, AR>), dim3(a.nblocks), dim3(WAVE * TY_A), 0, s, a);
}
template <int ST> static void pass_b_kernel(const IterArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(k_pass_b<ST>, dim3(a.nblocks), dim3(WAVE * TY_B), 0, s, a);
}

hipError_t launch_pass_a(const IterArgs &a0, int zc, int st, hipStream_t s, int arith) {
    const IterArgs a = chunked(a0, zc, xs_a(a0.d), TY_A);
    if (arith == ARITH_FMA) {
        if (st != ST_F32) return hipErrorInvalidValue;
        pass_a_kernel<ST_F32, ARITH_FMA>(a, s);
    } else if (st == ST_BF16) pass_a_kernel<ST_BF16>(a, s);
    else if (st == ST_F16) pass_a_kernel<ST_F16>(a, s);
    else pass_a_kernel<ST_F32>(a, s);
    return hipGetLastError();
}

hipError_t launch_pass_b(const IterArgs &a0, int zc, int st, hipStream_t s) {
    const IterArgs a = chunked(a0, zc, XS_B, TY_B);
    if (st == ST_BF16) pass_b_kernel<ST_BF16>(a, s);
    else if (st == ST_F16) pass_b_kernel<ST_F16>(a, s);
    else pass_b_kernel<ST_F32>(a, s);
    return hipGetLastError();
}

IterArgs pass_args(const IterArgs &a, int pass, int zc) {
    return pass == 1 ? chunked(a, zc, xs_a(a.d), TY_A) : chunked(a, zc, XS_B, TY_B);
}

template <int ST>
static void pass_batch_kernel(int pass, const IterArgs *d_base, const int *d_starts, int npairs, int nblocks,
                              const BatchIter &it, hipStream_t s) {
    if (pass == 1)
        hipLaunchKernelGGL(k_pass_a_batch<ST>, dim3(nblocks), dim3(WAVE * TY_A), 0, s, d_base, d_starts, npairs, it);
    else
        hipLaunchKernelGGL(k_pass_b_batch<ST>, dim3(nblocks), dim3(WAVE * TY_B), 0, s, d_base, d_starts, npairs, it);
}

hipError_t launch_pass_batch(int pass, const IterArgs *d_base, const int *d_starts, int npairs, int nblocks,
                             const BatchIter &it, int st, hipStream_t s) {
    if (nblocks <= 0) return hipSuccess;
    if (st == ST_BF16) pass_batch_kernel<ST_BF16>(pass, d_base, d_starts, npairs, nblocks, it, s);
    else if (st == ST_F16) pass_batch_kernel<ST_F16>(pass, d_base, d_starts, npairs, nblocks, it, s);
    else pass_batch_kernel<ST_F32>(pass, d_base, d_starts, npairs, nblocks, it, s);
    return hipGetLastError();
}

hipError_t pass_occupancy(int which, int *blocks_per_cu) {
    if (which != 1 && which != 2) return hipErrorInvalidValue;
    const void *k = which == 1 ? (const void *)k_pass_a<ST_F32> : (const void *)k_pass_b<ST_F32>;
    const int ty = which == 1 ? TY_A : TY_B;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, WAVE * ty, 0);
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

hipError_t launch_dice(Dims d, const int16_t *la, const int16_t *lb, int no_labels, unsigned *counts, hipStream_t s) {
    const int rows = (d.Y - 1) * (d.Z - 1);  // at least 2 rows per wave
    hipLaunchKernelGGL(k_dice, dim3(grid_for(rows, 4 * 2, 2048)), dim3(256), 0, s, d, la, lb, no_labels, counts);
    return hipGetLastError();
}

int min_det_partials_count(Dims d) { return grid_for(d.X * d.Y * d.Z, RED_BLOCK, 4096); }

hipError_t launch_min_det(Dims d, const float *psi, float *part_true, float *part_ref, hipStream_t s) {
    hipLaunchKernelGGL(k_min_det, dim3(min_det_partials_count(d)), dim3(RED_BLOCK), 0, s, d, psi, part_true, part_ref);
    return hipGetLastError();
}

hipError_t launch_get_inverse(Dims d, const float *psi4, float *pinv4, int sweeps, hipStream_t s) {
    const int n = d.X * d.Y * d.Z;
    for (int k = 0; k < sweeps; ++k) {
        hipLaunchKernelGGL(k_estimate_inverse, dim3(grid_for(n, 256)), dim3(256), 0, s, d, psi4, pinv4);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_interpolate_field4(Dims d, const float *f4, const float *psi4, float *out4, hipStream_t s) {
    const int n = d.X * d.Y * d.Z;
    hipLaunchKernelGGL(k_interpolate_field4, dim3(grid_for(n, 256)), dim3(256), 0, s, d, f4, psi4, out4);
    return hipGetLastError();
}

hipError_t launch_voxel_max_energy(Dims d, const float *T, const float *S, const float *psi, float w_reg,
                                   unsigned long long *slot, hipStream_t s) {
    const int n = d.X * d.Y * d.Z;
    hipLaunchKernelGGL(k_voxel_max_energy, dim3(grid_for(n, 256, 2048)), dim3(256), 0, s, d, T, S, psi, w_reg, slot);
    return hipGetLastError();
}

int sum_partials_count(long long n) { return (int)(n / 256 + 1 < 1024 ? n / 256 + 1 : 1024); }

hipError_t launch_sum(const float *a, long long n, double *partials, double *out, hipStream_t s) {
    const int np = sum_partials_count(n);
    hipLaunchKernelGGL(k_sum_partials, dim3(np), dim3(256), 0, s, a, n, partials);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, s, partials, np, out);
    return hipGetLastError();
}

hipError_t launch_vr_init(int st, unsigned *r0, unsigned *r1, int n, float rho_0, hipStream_t s) {
    if (st == ST_BF16) hipLaunchKernelGGL(k_vr_init<ST_BF16>, dim3(grid_for(n, 256)), dim3(256), 0, s, r0, r1, n, rho_0);
    else hipLaunchKernelGGL(k_vr_init<ST_F16>, dim3(grid_for(n, 256)), dim3(256), 0, s, r0, r1, n, rho_0);
    return hipGetLastError();
}

hipError_t launch_reinit(Dims d, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab,
                         float *T, int16_t *LT, float *S, int16_t *LS, int st, float *rho0, float *rho1, unsigned *vr0,
                         unsigned *vr1, float rho_0, hipStream_t s) {
    const int n = d.X * d.Y * d.Z;
    const bool f32 = st == ST_F32;
    hipLaunchKernelGGL(k_reinit, dim3(grid_for(n, 256)), dim3(256), 0, s, d, tgt, tgt_lab, src, src_lab, T, LT, S, LS,
                       f32 ? rho0 : nullptr, f32 ? rho1 : nullptr, rho_0);
    if (st == ST_BF16) hipLaunchKernelGGL(k_vr_reset_rho<ST_BF16>, dim3(grid_for(n, 256)), dim3(256), 0, s, vr0, vr1, n, rho_0);
    else if (st == ST_F16) hipLaunchKernelGGL(k_vr_reset_rho<ST_F16>, dim3(grid_for(n, 256)), dim3(256), 0, s, vr0, vr1, n, rho_0);
    return hipGetLastError();
}

hipError_t launch_vr_set(int st, unsigned *r0, unsigned *r1, const float *v4, const float *rho, int n, hipStream_t s) {
    if (st == ST_BF16) hipLaunchKernelGGL(k_vr_set<ST_BF16>, dim3(grid_for(n, 256)), dim3(256), 0, s, r0, r1, v4, rho, n);
    else hipLaunchKernelGGL(k_vr_set<ST_F16>, dim3(grid_for(n, 256)), dim3(256), 0, s, r0, r1, v4, rho, n);
    return hipGetLastError();
}

hipError_t launch_vr_get(int st, const unsigned *r, float *v4, float *rho, int n, hipStream_t s) {
    if (st == ST_BF16) hipLaunchKernelGGL(k_vr_get<ST_BF16>, dim3(grid_for(n, 256)), dim3(256), 0, s, r, v4, rho, n);
    else hipLaunchKernelGGL(k_vr_get<ST_F16>, dim3(grid_for(n, 256)), dim3(256), 0, s, r, v4, rho, n);
    return hipGetLastError();
}

hipError_t launch_pack4to3(const float *src4, float *dst3a, float *dst3b, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_pack4to3, dim3(grid_for(n, 256)), dim3(256), 0, s, src4, dst3a, dst3b, n);
    return hipGetLastError();
}

hipError_t launch_slots_reduce(const unsigned long long *slots, int iters, unsigned long long *out, hipStream_t s) {
    if (iters <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_slots_reduce, dim3((iters + 3) / 4), dim3(256), 0, s, slots, iters, out);
    return hipGetLastError();
}

hipError_t launch_zero3(void *p0, size_t b0, void *p1, size_t b1, void *p2, size_t b2, hipStream_t s) {
    if ((b0 | b1 | b2) & 3u) return hipErrorInvalidValue;
    const long long n = (long long)((b0 + b1 + b2) / 4);
    const int grid = (int)(n < 256LL * 1024 ? (n + 255) / 256 : 1024);
    hipLaunchKernelGGL(k_zero3, dim3(grid > 0 ? grid : 1), dim3(256), 0, s, static_cast<unsigned *>(p0),
                       (long long)(b0 / 4), static_cast<unsigned *>(p1), (long long)(b1 / 4),
                       static_cast<unsigned *>(p2), (long long)(b2 / 4));
    return hipGetLastError();
}

hipError_t launch_identity4(Dims d, float *dst4, hipStream_t s) {
    const int n = d.X * d.Y * d.Z;
    hipLaunchKernelGGL(k_identity4, dim3(grid_for(n, 256)), dim3(256), 0, s, d, dst4);
    return hipGetLastError();
}

hipError_t launch_unpack3to4(const float *src3, float *dst4, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_unpack3to4, dim3(grid_for(n, 256)), dim3(256), 0, s, src3, dst4, n);
    return hipGetLastError();
}

}  // namespace frk
