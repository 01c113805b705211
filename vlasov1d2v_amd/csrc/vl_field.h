// FP64 field-side kernels: deterministic moment reduction, spectral Poisson
// solve (high-order nodal DFT), field energy, max |acc E|, periodic x-ghost
// copy.  All of them are small (O(nx n M) work) and latency-bound.
//
// Reference semantics (crewsdw/Vlasov1D2V):
//   moment_zero_of_arr      main/grid.py:319-321
//   mean(density)-density   main/timestep.py:155,172
//   Elliptic.poisson2       main/elliptic.py:117-145
//   fourier_basis           main/grid.py:78-82
//   sum_fourier             main/grid.py:84-88
//   electric_energy         main/elliptic.py:150-151
//   get_max_speeds          main/timestep.py:235-242
//   swap_ghost_cells        main/grid.py:187-189
#pragma once
#include "vl_common.h"

namespace vl {

constexpr int kFieldBlock = 1024;

// rho[(ex-1)*n + ix] = sum over the tiles of x-element ex of partials, in a
// fixed order (thread-strided double-double sums, then a fixed tree).
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double2 *__restrict__ partials,
                                                            int n, int tiles_per_ex,
                                                            double *__restrict__ rho) {
    __shared__ double sh[kBlock];
    __shared__ double sl[kBlock];
    const int ex0 = blockIdx.x;   // 0-based interior x element
    const int tid = threadIdx.x;
    for (int ix = 0; ix < n; ++ix) {
        DD acc;
        for (int t = tid; t < tiles_per_ex; t += kBlock) {
            const double2 p = partials[((int64_t)ex0 * tiles_per_ex + t) * n + ix];
            DD d;
            d.hi = p.x;
            d.lo = p.y;
            acc.add(d);
        }
        sh[tid] = acc.hi;
        sl[tid] = acc.lo;
        __syncthreads();
        for (int w = kBlock / 2; w > 0; w >>= 1) {
            if (tid < w) {
                DD a, b;
                a.hi = sh[tid];
                a.lo = sl[tid];
                b.hi = sh[tid + w];
                b.lo = sl[tid + w];
                a.add(b);
                sh[tid] = a.hi;
                sl[tid] = a.lo;
            }
            __syncthreads();
        }
        if (tid == 0) rho[ex0 * n + ix] = sh[0] + sl[0];
        __syncthreads();
    }
}

struct FieldArgs {
    int nx, n, M;
    int from_rho;            // 1: charge density = mean(rho) - rho; 0: given
    double poisson_coef;
    double dx, length;
    double acc;
};

// Block-wide double-double sum (kFieldBlock threads); returns the rounded
// value to every thread.
__device__ __forceinline__ double block_sum_dd(DD v, double *sh, double *sl) {
    const int tid = threadIdx.x;
    sh[tid] = v.hi;
    sl[tid] = v.lo;
    __syncthreads();
    for (int w = kFieldBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            DD a, b;
            a.hi = sh[tid];
            a.lo = sl[tid];
            b.hi = sh[tid + w];
            b.lo = sl[tid + w];
            a.add(b);
            sh[tid] = a.hi;
            sl[tid] = a.lo;
        }
        __syncthreads();
    }
    const double r = sh[0] + sl[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double *sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int w = kFieldBlock / 2; w > 0; w >>= 1) {
        if (tid < w) sh[tid] = fmax(sh[tid], sh[tid + w]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Spectral Poisson solve + energy (+ max|acc E|) in one workgroup.
//   cd   = from_rho ? mean(rho) - rho : rho                 (nx*n)
//   c_m  = (sum_k (poisson_coef*cd_k) T[k][m]) * dx / length
//   Ehat = -c/(i k), phihat = -c/k^2, k = 0 -> 0
//   E[1:-1] = Re sum_m Ehat_m phases[m][k]; periodic ghosts
//   energy = 0.5 * sum w_j E^2
__global__ __launch_bounds__(kFieldBlock) void k_field(
    const double *__restrict__ rho_or_cd, const double2 *__restrict__ T,   // [k][m]
    const double2 *__restrict__ phases,                                   // [m][k]
    const double *__restrict__ kwave, const double *__restrict__ wts,     // M, n
    double *__restrict__ cd_ws, double2 *__restrict__ coef_ws,            // nx*n, 2*M
    double *__restrict__ E, double *__restrict__ phi, double *__restrict__ energy,
    double *__restrict__ maxabs, FieldArgs a) {
    __shared__ double sh[kFieldBlock];
    __shared__ double sl[kFieldBlock];
    const int tid = threadIdx.x;
    const int K = a.nx * a.n;

    // 1. charge density
    if (a.from_rho) {
        DD s;
        for (int k = tid; k < K; k += kFieldBlock) s.add(rho_or_cd[k]);
        const double mean = block_sum_dd(s, sh, sl) / (double)K;
        for (int k = tid; k < K; k += kFieldBlock) cd_ws[k] = mean - rho_or_cd[k];
    } else {
        for (int k = tid; k < K; k += kFieldBlock) cd_ws[k] = rho_or_cd[k];
    }
    __syncthreads();

    // 2. forward nodal DFT and spectral solve
    for (int m = tid; m < a.M; m += kFieldBlock) {
        double re = 0.0, im = 0.0;
        for (int k = 0; k < K; ++k) {
            const double r = a.poisson_coef * cd_ws[k];
            const double2 t = T[(int64_t)k * a.M + m];
            re += r * t.x;
            im += r * t.y;
        }
        re = re * a.dx / a.length;
        im = im * a.dx / a.length;
        const double kw = kwave[m];
        double2 eh = make_double2(0.0, 0.0), ph = make_double2(0.0, 0.0);
        if (kw != 0.0) {
            eh = make_double2(-(im / kw), re / kw);
            const double k2 = kw * kw;
            ph = make_double2(-(re / k2), -(im / k2));
        }
        coef_ws[m] = eh;
        coef_ws[a.M + m] = ph;
    }
    __syncthreads();

    // 3. inverse sums at the nodes, ghosts, energy, max |acc E|
    DD en;
    double mx = 0.0;
    for (int k = tid; k < K; k += kFieldBlock) {
        double e = 0.0, p = 0.0;
        for (int m = 0; m < a.M; ++m) {
            const double2 ph = phases[(int64_t)m * K + k];
            const double2 c = coef_ws[m];
            e += c.x * ph.x - c.y * ph.y;
            if (phi) {
                const double2 cp = coef_ws[a.M + m];
                p += cp.x * ph.x - cp.y * ph.y;
            }
        }
        const int el = k / a.n;
        const int j = k - el * a.n;
        E[(el + 1) * a.n + j] = e;
        if (el == a.nx - 1) E[j] = e;                      // E[0] = E[-2]
        if (el == 0) E[(a.nx + 1) * a.n + j] = e;          // E[-1] = E[1]
        if (phi) phi[k] = p;
        en.add(mul_rn(mul_rn(e, e), wts[j]));
        mx = fmax(mx, fabs(a.acc * e));
    }
    const double tot = block_sum_dd(en, sh, sl);
    const double gmx = block_max(mx, sh);
    if (tid == 0) {
        if (energy) energy[0] = 0.5 * tot;
        if (maxabs) maxabs[0] = gmx;
    }
}

// energy[0] = 0.5 * sum_interior w_j E^2 (main/elliptic.py:150-151)
__global__ __launch_bounds__(kFieldBlock) void k_energy(const double *__restrict__ E,
                                                        const double *__restrict__ wts, int nx,
                                                        int n, double *__restrict__ energy) {
    __shared__ double sh[kFieldBlock];
    __shared__ double sl[kFieldBlock];
    DD en;
    for (int k = threadIdx.x; k < nx * n; k += kFieldBlock) {
        const double e = E[n + k];
        en.add(mul_rn(mul_rn(e, e), wts[k % n]));
    }
    const double tot = block_sum_dd(en, sh, sl);
    if (threadIdx.x == 0) energy[0] = 0.5 * tot;
}

// out[0] = max |acc * E| over the ghosted field (main/timestep.py:237)
__global__ __launch_bounds__(kFieldBlock) void k_max_abs_acc_e(const double *__restrict__ E, int count,
                                                               double acc, double *__restrict__ out) {
    __shared__ double sh[kFieldBlock];
    double mx = 0.0;
    for (int k = threadIdx.x; k < count; k += kFieldBlock) mx = fmax(mx, fabs(acc * E[k]));
    const double g = block_max(mx, sh);
    if (threadIdx.x == 0) out[0] = g;
}

// Periodic x-ghost copy, in place: plane 0 <- plane nx, plane nx+1 <- plane 1.
// Planes are contiguous (the x element is the outermost axis).
template <typename W>
__global__ __launch_bounds__(kBlock) void k_swap_ghosts(W *__restrict__ f, int64_t plane_words, int nx) {
    const int64_t total = 2 * plane_words;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kBlock) {
        const int side = i >= plane_words;
        const int64_t o = i - side * plane_words;
        if (side == 0)
            f[o] = f[(int64_t)nx * plane_words + o];
        else
            f[(int64_t)(nx + 1) * plane_words + o] = f[plane_words + o];
    }
}

}  // namespace vl
