// Stabilized central-flux DG Poisson solve (Elliptic.poisson,
// main/elliptic.py:86-115), the alternative to the spectral solve.  The
// operator is assembled and inverted on the host (as the reference does,
// main/elliptic.py:24-84); the device applies it:
//   rhs   = pc * (charge_density . mass^T), with a zero gauge entry
//   phi   = (inv_op rhs)[:-1] / J^2               -> anti-aliased (optional)
//   E     = -J * grad_op^T phi                     -> anti-aliased (optional)
//   anti-alias: c_m = (sum_k f_k T[k][m]) dx / L, f_k = Re sum_m c_m phases[m][k]
//   (Grid1D.fourier_basis / sum_fourier, main/grid.py:78-88)
#pragma once
#include "vl_common.h"
#include "vl_field.h"

namespace vl {

// rhs[e n + i] = pc * sum_j cd[e n + j] mass[i n + j], rhs[K] = 0
__global__ __launch_bounds__(kBlock) void k_cf_rhs(const double *__restrict__ cd, const double *__restrict__ mass,
                                                   double *__restrict__ rhs, int nx, int n, double pc) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int K = nx * n;
    if (k > K) return;
    if (k == K) {
        rhs[K] = 0.0;
        return;
    }
    const int e = k / n, i = k - (k / n) * n;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += cd[e * n + j] * mass[i * n + j];
    rhs[k] = pc * s;
}

// y[r + out_off] = (sum_c A[r cols + c] x[c]) * scale for r < rows: one
// workgroup per row, thread partial sums in a fixed order, then a tree.
__global__ __launch_bounds__(kBlock) void k_matvec(const double *__restrict__ A, const double *__restrict__ x,
                                                   double *__restrict__ y, int rows, int cols, double scale,
                                                   int out_off) {
    __shared__ double sh[kBlock];
    const int r = blockIdx.x;
    if (r >= rows) return;
    const double *a = A + (int64_t)r * cols;
    double s = 0.0;
    for (int c = threadIdx.x; c < cols; c += kBlock) s += a[c] * x[c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) y[r + out_off] = sh[0] * scale;
}

// c_m = (sum_k f[k] T[k][m]) * dx / length, kDftModes modes per workgroup
__global__ __launch_bounds__(kBlock) void k_dft_plain(const double *__restrict__ f, const double2 *__restrict__ T,
                                                      double2 *__restrict__ coef, int K, int M, double dx,
                                                      double length) {
    __shared__ double pr[kBlock], pi[kBlock];
    constexpr int G = kBlock / kDftModes;
    const int tid = threadIdx.x;
    const int mm = tid % kDftModes, g = tid / kDftModes;
    const int m = blockIdx.x * kDftModes + mm;
    double re = 0.0, im = 0.0;
    if (m < M)
        for (int k = g; k < K; k += G) {
            const double2 t = T[(int64_t)k * M + m];
            re += f[k] * t.x;
            im += f[k] * t.y;
        }
    pr[tid] = re;
    pi[tid] = im;
    __syncthreads();
    for (int w = G / 2; w > 0; w >>= 1) {
        if (g < w) {
            pr[tid] += pr[tid + w * kDftModes];
            pi[tid] += pi[tid + w * kDftModes];
        }
        __syncthreads();
    }
    if (g == 0 && m < M) coef[m] = make_double2(pr[tid] * dx / length, pi[tid] * dx / length);
}

// out[k + out_off] = Re sum_m coef[m] phases[m][k], k < K; with ghost_n > 0
// also the periodic ghosts of a ghosted (nx+2)*n field (out_off = n)
__global__ __launch_bounds__(kBlock) void k_dft_plain_sum(const double2 *__restrict__ coef,
                                                          const double2 *__restrict__ phases,
                                                          double *__restrict__ out, int K, int M, int out_off,
                                                          int ghost_n) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= K) return;
    double s = 0.0;
    for (int m = 0; m < M; ++m) {
        const double2 c = coef[m], p = phases[(int64_t)m * K + k];
        s += c.x * p.x - c.y * p.y;
    }
    out[k + out_off] = s;
    if (ghost_n > 0) {
        if (k >= K - ghost_n) out[k - (K - ghost_n)] = s;     // E[0] = E[-2]
        if (k < ghost_n) out[K + ghost_n + k] = s;            // E[-1] = E[1]
    }
}

// periodic ghosts of a ghosted (nx+2)*n field whose interior is set
__global__ void k_field_ghosts(double *__restrict__ E, int nx, int n) {
    const int j = threadIdx.x;
    if (j < n) {
        E[j] = E[nx * n + j];
        E[(nx + 1) * n + j] = E[n + j];
    }
}

}  // namespace vl
