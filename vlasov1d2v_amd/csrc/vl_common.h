// Device-side helpers shared by the gfx950 kernels of the 1D2V DG path.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vl {

constexpr int kMaxN = 10;          // Lobatto nodes per axis supported (2..10)
constexpr int kBlock = 256;        // threads per workgroup of the tile kernels
constexpr int kNumXcd = 8;         // MI355X: 8 XCDs, workgroups dealt round-robin

// Operator tables passed by value as a kernel argument (read through the
// scalar cache; uniform across the launch).
struct OpTables {
    double up[kMaxN * kMaxN];      // volume operator, up[i*n+j] (main/basis.py:176-188)
    double xi0[kMaxN];             // face lift, left column  (main/basis.py:110)
    double xi1[kMaxN];             // face lift, right column
    double wq[kMaxN * kMaxN];      // moment weights (w_iu w_iv)/(Ju Jv) (main/grid.py:297-301)
    float upf[kMaxN * kMaxN];      // FP32 copies for the reduced-precision forms
    float xi0f[kMaxN];
    float xi1f[kMaxN];
    // Even-odd split of the centro-antisymmetric volume operator
    // (up[n-1-i][n-1-j] = -up[i][j]), h = n/2, m = n/2 (middle node, n odd):
    //   eoP[i][j] = (up[i][j] + up[n-1-i][j]) / 2,  eoQ[i][j] = (up[i][j] - up[n-1-i][j]) / 2
    //   eoM[i] = up[i][m],  eoR[j] = up[m][j]          (i, j < h)
    double eoP[(kMaxN / 2) * (kMaxN / 2)], eoQ[(kMaxN / 2) * (kMaxN / 2)], eoM[kMaxN / 2], eoR[kMaxN / 2];
    float eoPf[(kMaxN / 2) * (kMaxN / 2)], eoQf[(kMaxN / 2) * (kMaxN / 2)], eoMf[kMaxN / 2], eoRf[kMaxN / 2];
    // factored moment weights of the 16-bit forms: wq[iu][iv] ~ wu[iu] * wv[iv]
    float wu[kMaxN];               // w_iu / (Ju Jv)
    float wv[kMaxN];               // w_iv
};

// Raw buffer resource over `bytes` bytes from p (gfx9 data-format word).
// Loads at offsets >= bytes return 0 and stores there are dropped, so
// invalid slots use kOob instead of a branch.
constexpr int kOob = 0x7fffffff;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
typedef unsigned int vl_u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int vl_u32x4 __attribute__((ext_vector_type(4)));
typedef double vl_f64x2 __attribute__((ext_vector_type(2)));

// Component k (a compile-time constant after unrolling) of a 16-byte vector.
// Named components, not v[k]: with the subscript form hipcc has read lane 0
// for every k of an FP32 vector returned by a buffer load.
__device__ __forceinline__ unsigned int lane(const vl_u32x4 &v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
__device__ __forceinline__ void set_lane(vl_u32x4 &v, int k, unsigned int b) {
    if (k == 0) v.x = b; else if (k == 1) v.y = b; else if (k == 2) v.z = b; else v.w = b;
}

// Layouts of a state array.  kRefLayout is the reference's C-order
// (NX+2, n, NU+2, n, NV+2, n).  kElemLayout ("element-major") keeps the n^3
// nodes of each element contiguous, elements ordered (ex, eu, ev) and padded
// to a pitch of EP = round_up(n^3, 8) values so that every element starts
// 16-byte aligned for all storage types.  Node order inside an element:
// FP64 / FP32 states put x fastest, node (ix, iu, iv) at (iu n + iv) n + ix,
// so that an element's u-face (iu fixed) is one run of n^2 values and its
// v-face nine runs of n: those are the faces a stage tile reads from the
// neighbouring pencils (the x halo comes from the tile's own walk).  BF16 /
// FP16 states keep (ix n + iu) n + iv, the order their matrix-core kernel
// builds its operand blocks from.
enum { kRefLayout = 0, kElemLayout = 1 };
__host__ __device__ constexpr int elem_pitch(int n) { return ((n * n * n + 7) / 8) * 8; }
// storage codes as in Store<S> below: 0 FP64, 1 FP32, 2 BF16, 3 FP16
__host__ __device__ constexpr bool elem_x_minor(int storage) { return storage == 0 || storage == 1; }
__host__ __device__ constexpr int elem_node(bool x_minor, int n, int ix, int iu, int iv) {
    return x_minor ? (iu * n + iv) * n + ix : (ix * n + iu) * n + iv;
}
__device__ __forceinline__ vl_u32x4 buf_ld16(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
}
__device__ __forceinline__ void buf_st16(__amdgpu_buffer_rsrc_t r, int off, vl_u32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, 0);
}
// Streaming forms (cache policy NT): for data a kernel touches once -- the
// stage output and f^n -- so that they do not push the re-read stage input
// (halo rows of neighbouring tiles) out of the caches.
constexpr int kCachePolicyNT = 2;
__device__ __forceinline__ vl_u32x4 buf_ld16_nt(__amdgpu_buffer_rsrc_t r, int off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kCachePolicyNT);
}
__device__ __forceinline__ void buf_st16_nt(__amdgpu_buffer_rsrc_t r, int off, vl_u32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, kCachePolicyNT);
}

// Storage traits of the distribution function: element type E in HBM,
// compute type C, whether the stored value is the deviation g = (f - f0) *
// scale from the x-independent background f0 (reduced-precision forms).
template <int S> struct Store;
template <> struct Store<0> {          // FP64: f itself
    using E = double;
    using C = double;
    static constexpr bool kDelta = false;
    __device__ static C bld(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t r, int off, C v) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(vl_u32x2, v), r, off, 0, 0);
    }
    __device__ static C ld(const E *p, int64_t i) { return p[i]; }
    __device__ static void st(E *p, int64_t i, C v) { p[i] = v; }
    __device__ static C round(C v) { return v; }
    static constexpr int VW = 2;       // values per 16-byte vector
    __device__ static C unpack(const vl_u32x4 &v, int k) {
        return __builtin_bit_cast(double, k == 0 ? vl_u32x2{v.x, v.y} : vl_u32x2{v.z, v.w});
    }
    __device__ static void pack(vl_u32x4 &v, int k, C x) {
        const vl_u32x2 b = __builtin_bit_cast(vl_u32x2, x);
        if (k == 0) { v.x = b.x; v.y = b.y; } else { v.z = b.x; v.w = b.y; }
    }
};
template <> struct Store<1> {          // FP32 deviation
    using E = float;
    using C = float;
    static constexpr bool kDelta = true;
    __device__ static C bld(__amdgpu_buffer_rsrc_t r, int off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t r, int off, C v) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, off, 0, 0);
    }
    __device__ static C ld(const E *p, int64_t i) { return p[i]; }
    __device__ static void st(E *p, int64_t i, C v) { p[i] = v; }
    __device__ static C round(C v) { return v; }
    static constexpr int VW = 4;
    __device__ static C unpack(const vl_u32x4 &v, int k) { return __builtin_bit_cast(float, lane(v, k)); }
    __device__ static void pack(vl_u32x4 &v, int k, C x) { set_lane(v, k, __builtin_bit_cast(unsigned int, x)); }
};
template <> struct Store<2> {          // BF16 deviation (round to nearest even)
    using E = uint16_t;
    using C = float;
    static constexpr bool kDelta = true;
    __device__ static C ld(const E *p, int64_t i) { return (float)__builtin_bit_cast(__bf16, p[i]); }
    __device__ static C bld(__amdgpu_buffer_rsrc_t r, int off) {
        return (float)__builtin_bit_cast(__bf16, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t r, int off, C v) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, (__bf16)v), r, off, 0, 0);
    }
    __device__ static void st(E *p, int64_t i, C v) { p[i] = __builtin_bit_cast(uint16_t, (__bf16)v); }
    __device__ static C round(C v) { return (float)(__bf16)v; }
    static constexpr int VW = 8;
    __device__ static C unpack(const vl_u32x4 &v, int k) {
        return (float)__builtin_bit_cast(__bf16, (uint16_t)(lane(v, k >> 1) >> ((k & 1) * 16)));
    }
    __device__ static void pack(vl_u32x4 &v, int k, C x) {
        const unsigned int b = __builtin_bit_cast(uint16_t, (__bf16)x);
        const unsigned int w = lane(v, k >> 1);
        set_lane(v, k >> 1, (k & 1) ? ((w & 0xffffu) | (b << 16)) : ((w & 0xffff0000u) | b));
    }
};
template <> struct Store<3> {          // FP16 deviation, scaled (round to nearest even)
    using E = uint16_t;
    using C = float;
    static constexpr bool kDelta = true;
    __device__ static C ld(const E *p, int64_t i) { return (float)__builtin_bit_cast(_Float16, p[i]); }
    __device__ static C bld(__amdgpu_buffer_rsrc_t r, int off) {
        return (float)__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0));
    }
    __device__ static void bst(__amdgpu_buffer_rsrc_t r, int off, C v) {
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, (_Float16)v), r, off, 0, 0);
    }
    __device__ static void st(E *p, int64_t i, C v) { p[i] = __builtin_bit_cast(uint16_t, (_Float16)v); }
    __device__ static C round(C v) { return (float)(_Float16)v; }
    static constexpr int VW = 8;
    __device__ static C unpack(const vl_u32x4 &v, int k) {
        return (float)__builtin_bit_cast(_Float16, (uint16_t)(lane(v, k >> 1) >> ((k & 1) * 16)));
    }
    __device__ static void pack(vl_u32x4 &v, int k, C x) {
        const unsigned int b = __builtin_bit_cast(uint16_t, (_Float16)x);
        const unsigned int w = lane(v, k >> 1);
        set_lane(v, k >> 1, (k & 1) ? ((w & 0xffffu) | (b << 16)) : ((w & 0xffff0000u) | b));
    }
};

// The tables as seen through the constant address space (scalar loads).
using ConstOpTables = const __attribute__((address_space(4))) OpTables;

template <typename C> __device__ __forceinline__ C op_up(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_up<double>(ConstOpTables &t, int k) { return t.up[k]; }
template <> __device__ __forceinline__ float op_up<float>(ConstOpTables &t, int k) { return t.upf[k]; }
template <typename C> __device__ __forceinline__ C op_xi0(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_xi0<double>(ConstOpTables &t, int k) { return t.xi0[k]; }
template <> __device__ __forceinline__ float op_xi0<float>(ConstOpTables &t, int k) { return t.xi0f[k]; }
template <typename C> __device__ __forceinline__ C op_eoP(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_eoP<double>(ConstOpTables &t, int k) { return t.eoP[k]; }
template <> __device__ __forceinline__ float op_eoP<float>(ConstOpTables &t, int k) { return t.eoPf[k]; }
template <typename C> __device__ __forceinline__ C op_eoQ(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_eoQ<double>(ConstOpTables &t, int k) { return t.eoQ[k]; }
template <> __device__ __forceinline__ float op_eoQ<float>(ConstOpTables &t, int k) { return t.eoQf[k]; }
template <typename C> __device__ __forceinline__ C op_eoM(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_eoM<double>(ConstOpTables &t, int k) { return t.eoM[k]; }
template <> __device__ __forceinline__ float op_eoM<float>(ConstOpTables &t, int k) { return t.eoMf[k]; }
template <typename C> __device__ __forceinline__ C op_eoR(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_eoR<double>(ConstOpTables &t, int k) { return t.eoR[k]; }
template <> __device__ __forceinline__ float op_eoR<float>(ConstOpTables &t, int k) { return t.eoRf[k]; }
template <typename C> __device__ __forceinline__ C op_xi1(ConstOpTables &t, int k);
template <> __device__ __forceinline__ double op_xi1<double>(ConstOpTables &t, int k) { return t.xi1[k]; }
template <> __device__ __forceinline__ float op_xi1<float>(ConstOpTables &t, int k) { return t.xi1f[k]; }

// Geometry / stage scalars of one tile-kernel launch.
struct StageArgs {
    int nx, nu, nv;                // interior elements
    int nchunk;                    // v-chunks per (ex, eu)
    int xseg;                      // x elements per pencil segment (stage kernel)
    int reverse;                   // 1: stage kernels take their pencils last to first
    int periodic_x;                // 1: x-neighbours taken periodically from the interior
    int write_partials;            // 1: emit moment partials of the written values
    double jx, ju, jv;             // Jacobians 2/dx
    double acc, bfield;            // flux coefficient, magnetic field
    double c0, c1, c2dt;           // out = (c0*fn + c1*fin) + c2dt*L
    int64_t sx, sxe, su, sue;      // strides (elements) of ix, ex, iu, eu; iv stride 1, ev stride n
    double2 *partials;             // [(ex-1)*tiles_per_ex + tile][ix] (hi, lo)
    // reduced-precision forms: stored g = (f - f0) * scale
    const double *bg;              // [3][(nu+2)*n][(nv+2)*n]: G+, G-, Hv background tables
    // same index space, FP32, operator images without speed or Jacobian:
    // scale * (A+ f0 along u, A- f0 along u, A f0 along v upwinded by a_v)
    const float *bg32;
    // matrix-core operator fragments of the row-split kernel: [format][lane][4]
    // (format 0 BF16, 1 FP16; words hi.x, hi.y, lo.x, lo.y)
    const unsigned *mxfrag;
    // rounding of the 16-bit stored values (row-split kernel): 0 nearest-even,
    // 1 stochastic with random bits hashed from (*seed, value index); *seed
    // is a workspace word advanced on the device before every stage
    const unsigned *seed;
    int rounding;
    double scale, inv_scale;       // power of two
    // device time step (adaptive runs, the kernels' ADT variants): c2dt = c2 * dt_dev[0]
    const double *dt_dev;
    double c2;
};

// Stage coefficient of the time step: c2 * dt, dt from device memory in
// adaptive runs (the same product the host forms for a fixed dt).  The
// product is wave-uniform; readfirstlane keeps it in scalar registers like
// the kernel-argument value it replaces.
__device__ __forceinline__ void stage_dt_from_device(StageArgs &s) {
    if (s.dt_dev) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, s.c2 * s.dt_dev[0]);
        const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
        const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
        s.c2dt = __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    }
}

// Map a hardware workgroup id to a logical tile so that consecutive logical
// tiles run on the same XCD (the hardware deals workgroup b to XCD b % 8).
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int xcd = b % kNumXcd;
    const int local = b / kNumXcd;
    const int per = G / kNumXcd;
    const int rem = G % kNumXcd;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Logical pencil of this workgroup of a stage kernel: the XCD-aware remap,
// and with s.reverse the pencils taken last to first.  The RK stages
// alternate the order, so that each stage starts on the pencils the previous
// one wrote last, which the device's last-level cache may still hold.
#define VL_STAGE_PENCIL(s) ((s).reverse ? (int)gridDim.x - 1 - xcd_remap(blockIdx.x, gridDim.x) \
                                        : xcd_remap(blockIdx.x, gridDim.x))

// Error-free transformation a + b = s + e (Knuth two-sum).  Products that
// feed it are rounded separately (no contraction in the callers).
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Double-double accumulator (hi, lo).
struct DD {
    double hi, lo;
    __device__ __forceinline__ DD() : hi(0.0), lo(0.0) {}
    __device__ __forceinline__ void add(double x) {
#pragma clang fp contract(off)
        double s, e;
        two_sum(hi, x, s, e);
        hi = s;
        lo += e;
    }
    __device__ __forceinline__ void add(const DD &o) {
#pragma clang fp contract(off)
        double s, e;
        two_sum(hi, o.hi, s, e);
        hi = s;
        lo += e + o.lo;
    }
    __device__ __forceinline__ double value() const {
#pragma clang fp contract(off)
        return hi + lo;
    }
};

// 32-bit integer hash (xorshift-multiply "lowbias32" constants, C. Wellons);
// counter-based random bits for stochastic rounding.
__host__ __device__ __forceinline__ unsigned hash32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

}  // namespace vl
