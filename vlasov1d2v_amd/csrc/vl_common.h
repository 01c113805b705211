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
};

// Geometry / stage scalars of one tile-kernel launch.
struct StageArgs {
    int nx, nu, nv;                // interior elements
    int nchunk;                    // v-chunks per (ex, eu)
    int periodic_x;                // 1: x-neighbours taken periodically from the interior
    int write_partials;            // 1: emit moment partials of the written values
    double jx, ju, jv;             // Jacobians 2/dx
    double acc, bfield;            // flux coefficient, magnetic field
    double c0, c1, c2dt;           // out = (c0*fn + c1*fin) + c2dt*L
    int64_t sx, sxe, su, sue;      // strides (elements) of ix, ex, iu, eu; iv stride 1, ev stride n
    double2 *partials;             // [(ex-1)*tiles_per_ex + tile][ix] (hi, lo)
};

// Map a hardware workgroup id to a logical tile so that consecutive logical
// tiles run on the same XCD (the hardware deals workgroup b to XCD b % 8).
__device__ __forceinline__ int xcd_remap(int b, int G) {
    const int xcd = b % kNumXcd;
    const int local = b / kNumXcd;
    const int per = G / kNumXcd;
    const int rem = G % kNumXcd;
    return xcd * per + (xcd < rem ? xcd : rem) + local;
}

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

__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}

}  // namespace vl
