// Element-wise passes over whole states in the reference layout
// (NX+2, n, NU+2, n, NV+2, n), ghosts included: the linear SSP-RK
// combinations (main/timestep.py:189-229), separable initial states (the ring
// of main/grid.py:244-279 times an x profile), and the conversions into and
// out of the reduced storage forms (g = (f - f0) scale, f = f0 + g / scale).
//
// One workgroup row per (x element, x node) plane (blockIdx.y), threads over
// the plane's (NU+2) n (NV+2) n velocity nodes, as in the run-file kernels:
// the velocity index j addresses the plane-invariant tables (ring, f0)
// without a modulo.  HBM-bound streams.
#pragma once
#include "vl_common.h"

namespace vl {

struct PlaneShape {
    int64_t plane;     // velocity nodes per x plane, (NU+2) n (NV+2) n
    int row;           // velocity nodes per u node, (NV+2) n
    int n, nu, nv;
};

// Velocity node j = (eu n + iu) row + ev n + iv of a plane lies in a u or v
// ghost element.
__device__ __forceinline__ bool vel_ghost(const PlaneShape &p, int64_t j) {
    const int ju = (int)(j / p.row);                       // u node eu n + iu
    const int eu = ju / p.n;
    const int ev = (int)(j - (int64_t)ju * p.row) / p.n;
    return eu == 0 || eu == p.nu + 1 || ev == 0 || ev == p.nv + 1;
}

// y = a y + b x on every node, or only on the interior nodes (no ghost
// element on any axis; blockIdx.y then starts at the first interior plane).
// a == 0: y is not read; b == 0: x is not read (so neither can carry a NaN in).
__global__ __launch_bounds__(kBlock) void k_axpby(double *__restrict__ y, const double *__restrict__ x, double a,
                                                  double b, PlaneShape p, int plane0, int interior) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= p.plane) return;
    if (interior && vel_ghost(p, j)) return;
    const int64_t i = (int64_t)(plane0 + (int)blockIdx.y) * p.plane + j;
    double v = 0.0;
    if (a != 0.0) v = a * y[i];
    if (b != 0.0) v = __builtin_fma(b, x[i], v);
    y[i] = v;
}

// f[X, j] = store(((xp[X] vel[j]) scale)): the state of the storage form S
// whose value (FP64) or stored deviation (reduced forms, xp the deviation's x
// profile) is the separable product.  Rounded to nearest even; the 16-bit
// forms round the FP64 value to FP32 first (as an FP64 -> 16-bit tensor
// conversion does).
template <int S>
__global__ __launch_bounds__(kBlock) void k_init_separable(typename Store<S>::E *__restrict__ f,
                                                           const double *__restrict__ vel,
                                                           const double *__restrict__ xp, double scale,
                                                           int64_t plane) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= plane) return;
    const int64_t i = (int64_t)blockIdx.y * plane + j;
    const double v = (xp[blockIdx.y] * vel[j]) * scale;
    if constexpr (S == 0)
        f[i] = v;
    else
        Store<S>::st(f, i, (float)v);
}

// g[i] = store(d[i] scale): the stored deviation of an FP64 deviation d = f - f0.
template <int S>
__global__ __launch_bounds__(kBlock) void k_deviation_to_storage(const double *__restrict__ d,
                                                                 typename Store<S>::E *__restrict__ g,
                                                                 double scale, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    Store<S>::st(g, i, (float)(d[i] * scale));
}

// out = f (FP64 form) or f0 + g / scale (reduced forms; scale is a power of
// two, so g * inv_scale is the quotient).
template <int S>
__global__ __launch_bounds__(kBlock) void k_total_f64(const typename Store<S>::E *__restrict__ f,
                                                      const double *__restrict__ f0, double *__restrict__ out,
                                                      int64_t plane, double inv_scale) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= plane) return;
    const int64_t i = (int64_t)blockIdx.y * plane + j;
    if constexpr (Store<S>::kDelta)
        out[i] = f0[j] + (double)Store<S>::ld(f, i) * inv_scale;
    else
        out[i] = f[i];
}

// Deviation of a density from its mean over all (NX, n) nodes (the plain
// mean of the array, as cp.mean): out = mean - rho (sign > 0, the charge
// density the time steps pass to the Poisson solve, main/timestep.py:155,203)
// or rho - mean (sign < 0, main/main.py:131).  One workgroup; the sum runs in
// a fixed order (strided partial sums, then a tree), so it is deterministic.
__global__ __launch_bounds__(kBlock) void k_density_deviation(const double *__restrict__ rho,
                                                              double *__restrict__ out, int count, int sign) {
    __shared__ double part[kBlock];
    const int t = threadIdx.x;
    double a = 0.0;
    for (int i = t; i < count; i += kBlock) a += rho[i];
    part[t] = a;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) part[t] += part[t + w];
        __syncthreads();
    }
    const double mean = part[0] / (double)count;
    for (int i = t; i < count; i += kBlock) out[i] = sign > 0 ? mean - rho[i] : rho[i] - mean;
}

// Streaming probe of the stage kernels' HBM access shape: two arrays read
// once and one written once, 16-byte vectors, with non-temporal loads and
// stores (NT, the policy of the stage output and f^n) or the default policy.
// out = a + b over doubles; a grid-stride loop with four vectors in flight per
// thread.  Its best rate is the practical ceiling the stage kernels are
// quoted against, beside the 8 TB/s peak.
template <bool NT>
__device__ __forceinline__ vl_f64x2 probe_ld(const vl_f64x2 *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void probe_st(vl_f64x2 v, vl_f64x2 *p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool NT>
__global__ __launch_bounds__(kBlock) void k_stream_probe(const vl_f64x2 *__restrict__ a,
                                                         const vl_f64x2 *__restrict__ b,
                                                         vl_f64x2 *__restrict__ out, int64_t nvec) {
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + 3 * stride < nvec; i += 4 * stride) {
        vl_f64x2 x[4], y[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[k] = probe_ld<NT>(a + i + k * stride);
            y[k] = probe_ld<NT>(b + i + k * stride);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) probe_st<NT>(x[k] + y[k], out + i + k * stride);
    }
    for (; i < nvec; i += stride) probe_st<NT>(probe_ld<NT>(a + i) + probe_ld<NT>(b + i), out + i);
}

}  // namespace vl
