// FP64 field-side kernels: deterministic moment reduction, spectral Poisson
// solve (high-order nodal DFT), field energy, max |acc E|, periodic x-ghost
// copy.  All of them are small (O(nx n M) work) and latency-bound; they are
// spread over several workgroups so that the 2x nx*n*M complex tables are
// streamed by many CUs.
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
// The spectral tables (2 x nx*n*M complex) are read once per stage; the
// launches below spread them over ~80 workgroups with several independent
// loads in flight per thread, so that the solve costs a few load latencies.
constexpr int kDftModes = 4;      // modes per workgroup of the forward DFT
constexpr int kInvNodes = 8;      // nodes per workgroup of the inverse sum
constexpr int kInvGroups = 32;    // mode groups per node in the inverse sum

// Block-wide double-double sum over blockDim.x (a power of two) threads;
// returns the rounded value to every thread.
__device__ __forceinline__ double block_sum_dd(DD v, double *sh, double *sl) {
    const int tid = threadIdx.x;
    sh[tid] = v.hi;
    sl[tid] = v.lo;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
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
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w) sh[tid] = fmax(sh[tid], sh[tid + w]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// rho[(ex-1)*n + ix] = sum over the tiles of x-element ex of the moment
// partials, in a fixed order: thread t owns tiles t, t+256, ... (double-double),
// then a fixed tree over the 256 threads.  One workgroup per x element.
// When `seed` is given, workgroup 0 also advances it by one (the stage kernel
// that follows reads it as the seed of its stochastic rounding).
// Body for x element ex0 (sh, sl: N * kBlock doubles of LDS each).
template <int N>
__device__ __forceinline__ void reduce_partials_body(const double2 *__restrict__ partials, int tiles_per_ex,
                                                     double *__restrict__ rho, unsigned *__restrict__ seed,
                                                     int ex0, double *sh_, double *sl_) {
    if (seed && ex0 == 0 && threadIdx.x == 0) seed[0] += 1u;
    double (*sh)[kBlock] = reinterpret_cast<double (*)[kBlock]>(sh_);
    double (*sl)[kBlock] = reinterpret_cast<double (*)[kBlock]>(sl_);
    const int tid = threadIdx.x;
    DD acc[N];
    const double2 *base = partials + (int64_t)ex0 * tiles_per_ex * N;
    for (int t = tid; t < tiles_per_ex; t += kBlock) {
        double2 p[N];
#pragma unroll
        for (int ix = 0; ix < N; ++ix) p[ix] = base[(int64_t)t * N + ix];
#pragma unroll
        for (int ix = 0; ix < N; ++ix) {
            DD d;
            d.hi = p[ix].x;
            d.lo = p[ix].y;
            acc[ix].add(d);
        }
    }
#pragma unroll
    for (int ix = 0; ix < N; ++ix) {
        sh[ix][tid] = acc[ix].hi;
        sl[ix][tid] = acc[ix].lo;
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        for (int idx = tid; idx < N * w; idx += kBlock) {
            const int ix = idx / w;
            const int j = idx - ix * w;
            DD a, b;
            a.hi = sh[ix][j];
            a.lo = sl[ix][j];
            b.hi = sh[ix][j + w];
            b.lo = sl[ix][j + w];
            a.add(b);
            sh[ix][j] = a.hi;
            sl[ix][j] = a.lo;
        }
        __syncthreads();
    }
    if (tid < N) rho[ex0 * N + tid] = sh[tid][0] + sl[tid][0];
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double2 *__restrict__ partials,
                                                            int tiles_per_ex, double *__restrict__ rho,
                                                            unsigned *__restrict__ seed) {
    __shared__ double sh[N * kBlock];
    __shared__ double sl[N * kBlock];
    reduce_partials_body<N>(partials, tiles_per_ex, rho, seed, blockIdx.x, sh, sl);
}

struct FieldArgs {
    int nx, n, M;
    int from_rho;            // 1: charge density = mean(rho) - rho; 0: given
    double poisson_coef;
    double dx, length;
    double acc;
    // adaptive time step (main/timestep.py:231-242): when dt_next is set, the
    // inverse sums also write courant / (s0/dx + (max|acc E| + bv)/du + s2/dv)
    // there, with s0 = max|u|, bv = |acc B max|v||, s2 = |acc |B max|u|||
    double *dt_next;
    double cfl_courant, cfl_s0, cfl_bv, cfl_s2, cfl_dx, cfl_du, cfl_dv;
    // when set, the inverse sums reduce the field energy and max|acc E| and
    // set status[0] = 1 (sticky) if either is not finite: the state or the
    // field has overflowed or holds NaN (the reference has no such check; its
    // NaN test is commented out, main/timestep.py:118-121)
    unsigned int *status;
};

// The reference's adaptive step from max|acc E| (main/timestep.py:231-232,
// 235-242): the same divisions and left-to-right sum.
__device__ __forceinline__ double cfl_dt(const FieldArgs &a, double max_acc_e) {
    const double s1 = max_acc_e + a.cfl_bv;
    return a.cfl_courant / ((a.cfl_s0 / a.cfl_dx) + (s1 / a.cfl_du) + (a.cfl_s2 / a.cfl_dv));
}

// Forward nodal DFT and spectral solve for kDftModes modes per workgroup.
//   cd   = from_rho ? mean(rho) - rho : rho          (every workgroup forms it
//          identically: same fixed-order double-double sum)
//   c_m  = (sum_k (poisson_coef*cd_k) T[k][m]) * dx / length
//   Ehat = -c/(i k), phihat = -c/k^2, k = 0 -> 0     (main/elliptic.py:124-128)
// Body for mode group bx (rhs: nx*n doubles, sh, sl: kBlock doubles of LDS).
__device__ __forceinline__ void dft_forward_body(const double *__restrict__ rho_or_cd,
                                                 const double2 *__restrict__ T,   // [k][m]
                                                 const double *__restrict__ kwave,
                                                 double2 *__restrict__ coef,      // [2][M]
                                                 const FieldArgs &a, int bx, double *rhs, double *sh,
                                                 double *sl) {
    const int tid = threadIdx.x;
    const int K = a.nx * a.n;
    double mean = 0.0;
    if (a.from_rho) {
        DD s;
        for (int k = tid; k < K; k += kBlock) s.add(rho_or_cd[k]);
        mean = block_sum_dd(s, sh, sl) / (double)K;
    }
    for (int k = tid; k < K; k += kBlock) {
        const double cd = a.from_rho ? mean - rho_or_cd[k] : rho_or_cd[k];
        rhs[k] = a.poisson_coef * cd;
    }
    __syncthreads();
    // thread (group g, mode offset mm): k = g, g + G, ...
    constexpr int G = kBlock / kDftModes;
    const int mm = tid % kDftModes;
    const int g = tid / kDftModes;
    const int m = bx * kDftModes + mm;
    double re = 0.0, im = 0.0;
    if (m < a.M) {
        // batches of 8 independent table loads, then the multiply-adds in k order
        constexpr int U = 8;
        for (int k0 = g; k0 < K; k0 += U * G) {
            double2 t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * G;
                t[u] = k < K ? T[(int64_t)k * a.M + m] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = k0 + u * G;
                if (k < K) {
                    re += rhs[k] * t[u].x;
                    im += rhs[k] * t[u].y;
                }
            }
        }
    }
    // fixed-order tree over the G groups of each mode
    double *pr = sh, *pi = sl;
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
    if (g == 0 && m < a.M) {
        re = pr[tid] * a.dx / a.length;
        im = pi[tid] * a.dx / a.length;
        const double kw = kwave[m];
        double2 eh = make_double2(0.0, 0.0), ph = make_double2(0.0, 0.0);
        if (kw != 0.0) {
            eh = make_double2(-(im / kw), re / kw);
            const double k2 = kw * kw;
            ph = make_double2(-(re / k2), -(im / k2));
        }
        coef[m] = eh;
        coef[a.M + m] = ph;
    }
}

__global__ __launch_bounds__(kBlock) void k_dft_forward(const double *__restrict__ rho_or_cd,
                                                        const double2 *__restrict__ T,
                                                        const double *__restrict__ kwave,
                                                        double2 *__restrict__ coef, FieldArgs a) {
    extern __shared__ double dyn[];                   // nx*n values
    __shared__ double sh[kBlock];
    __shared__ double sl[kBlock];
    dft_forward_body(rho_or_cd, T, kwave, coef, a, blockIdx.x, dyn, sh, sl);
}

// Inverse sums at the nodes: E[1:-1] = Re sum_m Ehat_m phases[m][k] (and phi),
// periodic ghosts, and -- when energy/maxabs are requested -- the field energy
// 0.5 sum w E^2 and max|acc E|, reduced in a fixed order by all threads of the
// last workgroup to finish (ticket counter in the workspace; it returns to zero).
// LDS of the inverse sums besides the 2*M coefficients
struct InvLds {
    double pe[kBlock], pp[kBlock];
    double sh[kInvNodes], sl[kInvNodes], smx[kInvNodes];
    int is_last;
};

// Body for node group bx of nblocks (cdyn: 2*M complex of LDS).
__device__ __forceinline__ void dft_inverse_body(
    const double2 *__restrict__ coef, const double2 *__restrict__ phases,   // [2][M], [m][k]
    const double *__restrict__ wts, double *__restrict__ E, double *__restrict__ phi,
    double2 *__restrict__ en_part, double *__restrict__ mx_part, unsigned int *__restrict__ ticket,
    double *__restrict__ energy, double *__restrict__ maxabs, const FieldArgs &a, int bx, int nblocks,
    double2 *cdyn, InvLds &ls) {
    double *pe = ls.pe, *pp = ls.pp, *sh = ls.sh, *sl = ls.sl, *smx = ls.smx;
    int &is_last = ls.is_last;
    const int tid = threadIdx.x;
    const int K = a.nx * a.n;
    const int M = a.M;
    for (int i = tid; i < 2 * M; i += kBlock) cdyn[i] = coef[i];
    __syncthreads();
    const int kk = tid % kInvNodes;
    const int g = tid / kInvNodes;
    const int k = bx * kInvNodes + kk;
    double e = 0.0, p = 0.0;
    if (k < K) {
        // batches of 8 independent phase loads, then the sums in m order
        constexpr int U = 8;
        for (int m0 = g; m0 < M; m0 += U * kInvGroups) {
            double2 ph[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int m = m0 + u * kInvGroups;
                ph[u] = m < M ? phases[(int64_t)m * K + k] : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int m = m0 + u * kInvGroups;
                if (m < M) {
                    const double2 c = cdyn[m];
                    e += c.x * ph[u].x - c.y * ph[u].y;
                    if (phi) {
                        const double2 cp = cdyn[M + m];
                        p += cp.x * ph[u].x - cp.y * ph[u].y;
                    }
                }
            }
        }
    }
    pe[tid] = e;
    pp[tid] = p;
    __syncthreads();
    for (int w = kInvGroups / 2; w > 0; w >>= 1) {
        if (g < w) {
            pe[tid] += pe[tid + w * kInvNodes];
            pp[tid] += pp[tid + w * kInvNodes];
        }
        __syncthreads();
    }
    const bool want_red = energy || maxabs || a.dt_next || a.status;
    if (g == 0) {
        DD en;
        double mx = 0.0;
        if (k < K) {
            e = pe[tid];
            const int el = k / a.n;
            const int j = k - el * a.n;
            E[(el + 1) * a.n + j] = e;
            if (el == a.nx - 1) E[j] = e;                      // E[0] = E[-2]
            if (el == 0) E[(a.nx + 1) * a.n + j] = e;          // E[-1] = E[1]
            if (phi) phi[k] = pp[tid];
            en.add(mul_rn(mul_rn(e, e), wts[j]));
            mx = fabs(a.acc * e);
        }
        sh[kk] = en.hi;
        sl[kk] = en.lo;
        smx[kk] = mx;
    }
    if (!want_red) return;
    __syncthreads();
    if (tid == 0) {
        DD s;
        double mx = 0.0;
        for (int i = 0; i < kInvNodes; ++i) {
            DD d;
            d.hi = sh[i];
            d.lo = sl[i];
            s.add(d);
            mx = fmax(mx, smx[i]);
        }
        en_part[bx] = make_double2(s.hi, s.lo);
        mx_part[bx] = mx;
        __threadfence();
        const unsigned int prev = atomicAdd(ticket, 1u);
        is_last = (prev == (unsigned)nblocks - 1);
    }
    __syncthreads();
    if (!is_last) return;
    // the last workgroup: thread t sums partials t, t + kBlock, ... (double-double),
    // then a fixed tree over the threads (pe / pp are free again)
    __threadfence();
    DD s;
    double mx = 0.0;
    const volatile double *vp = reinterpret_cast<const volatile double *>(en_part);
    const volatile double *vm = mx_part;
    for (int b = tid; b < nblocks; b += kBlock) {
        DD d;
        d.hi = vp[2 * b];
        d.lo = vp[2 * b + 1];
        s.add(d);
        mx = fmax(mx, vm[b]);
    }
    const double tot = block_sum_dd(s, pe, pp);
    mx = block_max(mx, pe);
    if (tid == 0) {
        if (energy) energy[0] = 0.5 * tot;
        if (maxabs) maxabs[0] = mx;
        if (a.dt_next) a.dt_next[0] = cfl_dt(a, mx);
        // fmax drops NaN operands, so the energy (a sum of squares) carries
        // the check for NaN and the maximum only adds infinities
        if (a.status && !(isfinite(tot) && isfinite(mx))) a.status[0] = 1u;
        atomicExch(ticket, 0u);
    }
}

__global__ __launch_bounds__(kBlock) void k_dft_inverse(
    const double2 *__restrict__ coef, const double2 *__restrict__ phases, const double *__restrict__ wts,
    double *__restrict__ E, double *__restrict__ phi, double2 *__restrict__ en_part,
    double *__restrict__ mx_part, unsigned int *__restrict__ ticket, double *__restrict__ energy,
    double *__restrict__ maxabs, FieldArgs a) {
    extern __shared__ double2 cdyn[];                 // 2*M coefficients
    __shared__ InvLds ls;
    dft_inverse_body(coef, phases, wts, E, phi, en_part, mx_part, ticket, energy, maxabs, a, blockIdx.x,
                     gridDim.x, cdyn, ls);
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

// Clock of an adaptive run in the workspace: clock[0] = dt of the next step,
// clock[1] = time, clock[2] = dt computed from the latest stage-3 field.
// Initial dt from the set-up field (main/timestep.py:88-89): max |acc E|
// over the ghosted field, then the CFL rule.
__global__ __launch_bounds__(kFieldBlock) void k_clock_init(const double *__restrict__ E, int count,
                                                            FieldArgs a, double time,
                                                            double *__restrict__ clock) {
    __shared__ double sh[kFieldBlock];
    double mx = 0.0;
    for (int k = threadIdx.x; k < count; k += kFieldBlock) mx = fmax(mx, fabs(a.acc * E[k]));
    const double g = block_max(mx, sh);
    if (threadIdx.x == 0) {
        const double dt = cfl_dt(a, g);
        clock[0] = dt;
        clock[1] = time;
        clock[2] = dt;
    }
}

// After a step (main/timestep.py:184 then 96): the adapted dt becomes the
// step's dt and the time advances by it; out (may be null) = (time, dt).
__global__ void k_clock_advance(double *__restrict__ clock, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double dt = clock[2];
        const double t = clock[1] + dt;
        clock[0] = dt;
        clock[1] = t;
        if (out) {
            out[0] = t;
            out[1] = dt;
        }
    }
}

// Fill a small buffer with zeros (used instead of a memset so that the call
// sequence stays capturable in a HIP graph).
__global__ void k_zero_words(unsigned int *__restrict__ p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0u;
}

// Periodic x-ghost copy, in place: plane 0 <- plane nx, plane nx+1 <- plane 1.
// Planes are contiguous (the x element is the outermost axis).
template <typename W>
__device__ __forceinline__ void swap_ghosts_body(W *__restrict__ f, int64_t plane_words, int nx, int bx,
                                                 int nblocks) {
    const int64_t total = 2 * plane_words;
    for (int64_t i = (int64_t)bx * kBlock + threadIdx.x; i < total; i += (int64_t)nblocks * kBlock) {
        const int side = i >= plane_words;
        const int64_t o = i - side * plane_words;
        if (side == 0)
            f[o] = f[(int64_t)nx * plane_words + o];
        else
            f[(int64_t)(nx + 1) * plane_words + o] = f[plane_words + o];
    }
}

template <typename W>
__global__ __launch_bounds__(kBlock) void k_swap_ghosts(W *__restrict__ f, int64_t plane_words, int nx) {
    swap_ghosts_body<W>(f, plane_words, nx, blockIdx.x, gridDim.x);
}

}  // namespace vl
