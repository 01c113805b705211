// FP64 tile kernel: semi-discrete upwind DG RHS of the 1D2V Vlasov equation,
// optionally fused with the Shu-Osher stage combination and the zeroth-moment
// partial sums of the values it writes.
//
// Reference semantics (crewsdw/Vlasov1D2V):
//   DGFlux.semi_discrete_rhs  main/fluxes.py:59-105
//   x_flux / u_flux / v_flux  main/fluxes.py:107-153
//   spatial/velocity_flux     main/fluxes.py:156-207 (upwind by sign(speed), >=0 positive)
//   stage combination         main/timestep.py:158-163,180-182
//   moment_zero_of_arr        main/grid.py:319-321
//
// Work decomposition.  One workgroup of 256 threads owns a tile
// (ex, eu, TV consecutive v-elements) with all n^3 nodes of each element,
// TV = floor(256 / n^2) so that every direction has n*P <= 256 lines of n
// nodes (P = TV*n v-nodes).  The tile plus its one-node halos is staged in
// LDS with loads that are contiguous along v (the innermost axis of the
// state).  Each thread then contracts one x-line, one u-line and one v-line
// with the n x n volume operator held in registers/kernarg memory and
// accumulates J_d * L_d into an LDS output tile in the reference's order
// ((Lx Jx + Lu Ju) + Lv Jv).  The epilogue combines with the stage inputs,
// writes coalesced rows, and reduces the moment partial of the tile.
#pragma once
#include "vl_common.h"

namespace vl {

enum StageMode { kModeRhs = 0, kModeStage1 = 1, kModeStage = 2 };

template <int N>
struct TileShape {
    static constexpr int TV = (kBlock / (N * N)) > 0 ? (kBlock / (N * N)) : 1;
    static constexpr int P = TV * N;        // v-nodes per tile
    static constexpr int PP = P + 2;        // with one halo node each side
};

// One 1-D upwind DG line operator (main/fluxes.py:107-207 for one line):
//   F_j = f_j * a
//   L_i = sum_j up[i][j] F_j - (xi0[i] * nfL + xi1[i] * nfR)
//   nfL = -(a >= 0 ? F_R(e-1) : F_L(e)),  nfR = (a >= 0 ? F_R(e) : F_L(e+1))
template <int N>
__device__ __forceinline__ void line_op(const double (&f)[N], double fl_nb, double fr_nb,
                                        double a, const OpTables &tab, double (&out)[N]) {
    double F[N];
#pragma unroll
    for (int j = 0; j < N; ++j) F[j] = f[j] * a;
    const double FL = fl_nb * a;   // neighbour-left element's right-face flux
    const double FR = fr_nb * a;   // neighbour-right element's left-face flux
    double nfl, nfr;
    if (a >= 0.0) {
        nfl = -FL;
        nfr = F[N - 1];
    } else {
        nfl = -F[0];
        nfr = FR;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double vol = 0.0;
#pragma unroll
        for (int j = 0; j < N; ++j) vol += tab.up[i * N + j] * F[j];
        const double lift = tab.xi0[i] * nfl + tab.xi1[i] * nfr;
        out[i] = vol - lift;
    }
}

template <int N, int MODE>
__global__ __launch_bounds__(kBlock) void k_stage_f64(
    const double *__restrict__ fin,    // stage input (interior; ghosts too when !periodic_x)
    const double *__restrict__ fgh,    // array whose u/v ghost planes are used
    const double *__restrict__ fn,     // f^n for MODE == kModeStage
    double *__restrict__ fout,
    const double *__restrict__ E,      // (nx+2)*N ghosted field
    const double *__restrict__ unodes, // (nu+2)*N
    const double *__restrict__ vnodes, // (nv+2)*N
    StageArgs s, OpTables tab) {
    using TS = TileShape<N>;
    constexpr int TV = TS::TV, P = TS::P, PP = TS::PP;
    __shared__ double ft[N * N * PP];      // tile rows (ix,iu) x v-nodes with halo
    __shared__ double fxh[2 * N * P];      // x-halo [side][iu][q]; reused for moment
    __shared__ double fuh[2 * N * P];      // u-halo [side][ix][q]
    __shared__ double acc_l[N * N * P];    // output accumulator rows (ix,iu) x q

    const int tid = threadIdx.x;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile % s.nchunk;
    const int t2 = tile / s.nchunk;
    const int eu = 1 + t2 % s.nu;
    const int ex = 1 + t2 / s.nu;
    const int v0 = 1 + chunk * TV;                       // first v element of the tile
    const int nvalid = min(TV, s.nv + 1 - v0);           // valid v elements
    const int pvalid = nvalid * N;
    const int gv0 = v0 * N;                              // first v node of the tile

    const int64_t base_x = (int64_t)ex * s.sxe;
    const int64_t base_u = (int64_t)eu * s.sue;

    // ---- stage the tile (+ v halo nodes) and the x/u face halos ----------
    // All global loads of the phase are issued before the first LDS store so
    // that they are in flight together.
    constexpr int NT = N * N * PP;
    constexpr int IT_T = (NT + kBlock - 1) / kBlock;
    constexpr int NH = 2 * N * P;
    constexpr int IT_H = (NH + kBlock - 1) / kBlock;
    double rt[IT_T], rx[IT_H], ru[IT_H];
#pragma unroll
    for (int it = 0; it < IT_T; ++it) {
        const int idx = tid + it * kBlock;
        rt[it] = 0.0;
        if (idx < NT) {
            const int r = idx / PP;
            const int q = idx - r * PP;
            if (q <= pvalid + 1) {
                const int ix = r / N;
                const int iu = r - ix * N;
                const int gv = gv0 - 1 + q;
                const int ev = gv / N;
                const double *src = (ev == 0 || ev == s.nv + 1) ? fgh : fin;
                rt[it] = src[base_x + ix * s.sx + base_u + iu * s.su + gv];
            }
        }
    }
    int exl = ex - 1, exr = ex + 1;
    if (s.periodic_x) {
        if (exl == 0) exl = s.nx;
        if (exr == s.nx + 1) exr = 1;
    }
#pragma unroll
    for (int it = 0; it < IT_H; ++it) {
        const int idx = tid + it * kBlock;
        rx[it] = 0.0;
        ru[it] = 0.0;
        if (idx < NH) {
            const int side = idx / (N * P);
            const int rem = idx - side * N * P;
            const int c = rem / P;          // iu for the x halo, ix for the u halo
            const int q = rem - c * P;
            if (q < pvalid) {
                const int64_t e = side ? exr : exl;
                const int ixn = side ? 0 : N - 1;
                rx[it] = fin[e * s.sxe + ixn * s.sx + base_u + c * s.su + gv0 + q];
                const int eun = side ? eu + 1 : eu - 1;
                const int iun = side ? 0 : N - 1;
                const double *src = (eun == 0 || eun == s.nu + 1) ? fgh : fin;
                ru[it] = src[base_x + c * s.sx + (int64_t)eun * s.sue + iun * s.su + gv0 + q];
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT_T; ++it) {
        const int idx = tid + it * kBlock;
        if (idx < NT) ft[idx] = rt[it];
    }
#pragma unroll
    for (int it = 0; it < IT_H; ++it) {
        const int idx = tid + it * kBlock;
        if (idx < NH) {
            fxh[idx] = rx[it];
            fuh[idx] = ru[it];
        }
    }
    __syncthreads();

    // ---- x-lines: (iu, q), line over ix ----------------------------------
    if (tid < N * P) {
        const int iu = tid / P;
        const int q = tid - iu * P;
        if (q < pvalid) {
            const double a = unodes[eu * N + iu];
            double f[N], o[N];
#pragma unroll
            for (int j = 0; j < N; ++j) f[j] = ft[(j * N + iu) * PP + q + 1];
            line_op<N>(f, fxh[iu * P + q], fxh[(N + iu) * P + q], a, tab, o);
#pragma unroll
            for (int i = 0; i < N; ++i) acc_l[(i * N + iu) * P + q] = o[i] * s.jx;
        }
    }
    __syncthreads();
    // ---- u-lines: (ix, q), line over iu ----------------------------------
    if (tid < N * P) {
        const int ix = tid / P;
        const int q = tid - ix * P;
        if (q < pvalid) {
            const int gv = gv0 + q;
            const double a = s.acc * (E[ex * N + ix] + s.bfield * vnodes[gv]);
            double f[N], o[N];
#pragma unroll
            for (int j = 0; j < N; ++j) f[j] = ft[(ix * N + j) * PP + q + 1];
            line_op<N>(f, fuh[ix * P + q], fuh[(N + ix) * P + q], a, tab, o);
#pragma unroll
            for (int i = 0; i < N; ++i) acc_l[(ix * N + i) * P + q] += o[i] * s.ju;
        }
    }
    __syncthreads();
    // ---- v-lines: (ix, iu, ev_local), line over iv -----------------------
    if (tid < N * N * TV) {
        const int r = tid / TV;
        const int el = tid - r * TV;
        if (el < nvalid) {
            const int iu = r - (r / N) * N;
            const double a = s.acc * (-s.bfield * unodes[eu * N + iu]);
            const double *row = &ft[r * PP + el * N];
            double f[N], o[N];
#pragma unroll
            for (int j = 0; j < N; ++j) f[j] = row[j + 1];
            line_op<N>(f, row[0], row[N + 1], a, tab, o);
#pragma unroll
            for (int i = 0; i < N; ++i) acc_l[r * P + el * N + i] += o[i] * s.jv;
        }
    }
    __syncthreads();

    // ---- epilogue: combine and write rows (coalesced along v) ------------
    // fn and fout may be the same array: each element is read and then
    // written by the same thread.
    constexpr int NO = N * N * P;
    constexpr int IT_O = (NO + kBlock - 1) / kBlock;
    double rn[IT_O];
    if (MODE == kModeStage) {
#pragma unroll
        for (int it = 0; it < IT_O; ++it) {
            const int idx = tid + it * kBlock;
            rn[it] = 0.0;
            if (idx < NO) {
                const int r = idx / P;
                const int q = idx - r * P;
                if (q < pvalid) {
                    const int ix = r / N;
                    const int iu = r - ix * N;
                    rn[it] = fn[base_x + ix * s.sx + base_u + iu * s.su + gv0 + q];
                }
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT_O; ++it) {
        const int idx = tid + it * kBlock;
        if (idx < NO) {
            const int r = idx / P;
            const int q = idx - r * P;
            if (q < pvalid) {
                const int ix = r / N;
                const int iu = r - ix * N;
                const int64_t addr = base_x + ix * s.sx + base_u + iu * s.su + gv0 + q;
                const double L = acc_l[idx];
                double val;
                if (MODE == kModeRhs) {
                    val = L;
                } else if (MODE == kModeStage1) {
                    val = ft[r * PP + q + 1] + s.c2dt * L;
                } else {
                    val = (s.c0 * rn[it] + s.c1 * ft[r * PP + q + 1]) + s.c2dt * L;
                }
                fout[addr] = val;
                acc_l[idx] = val;
            }
        }
    }
    if (MODE == kModeRhs || !s.write_partials) return;
    __syncthreads();

    // ---- moment partial of the written values: sum over (iu, q) per ix ---
    double *red_hi = fxh;            // N*P entries
    double *red_lo = fxh + N * P;    // N*P entries
    if (tid < N * P) {
        const int ix = tid / P;
        const int q = tid - ix * P;
        DD acc;
        if (q < pvalid) {
            const int iv = q - (q / N) * N;
#pragma unroll
            for (int iu = 0; iu < N; ++iu)
                acc.add(mul_rn(acc_l[(ix * N + iu) * P + q], tab.wq[iu * N + iv]));
        }
        red_hi[tid] = acc.hi;
        red_lo[tid] = acc.lo;
    }
    __syncthreads();
    if (tid < N) {
        DD acc;
        for (int q = 0; q < P; ++q) {
            DD t;
            t.hi = red_hi[tid * P + q];
            t.lo = red_lo[tid * P + q];
            acc.add(t);
        }
        const int tiles_per_ex = s.nu * s.nchunk;
        const int64_t slot = ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + chunk) * N + tid;
        s.partials[slot] = make_double2(acc.hi, acc.lo);
    }
}

// Moment partials of an FP64 state array (no RHS): same tiling as the stage
// kernel so the partial layout is shared.  main/grid.py:303-321.
template <int N>
__global__ __launch_bounds__(kBlock) void k_moment_partials_f64(const double *__restrict__ f,
                                                                StageArgs s, OpTables tab) {
    using TS = TileShape<N>;
    constexpr int TV = TS::TV, P = TS::P;
    __shared__ double red_hi[N * P];
    __shared__ double red_lo[N * P];
    const int tid = threadIdx.x;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile % s.nchunk;
    const int t2 = tile / s.nchunk;
    const int eu = 1 + t2 % s.nu;
    const int ex = 1 + t2 / s.nu;
    const int v0 = 1 + chunk * TV;
    const int nvalid = min(TV, s.nv + 1 - v0);
    const int pvalid = nvalid * N;
    const int gv0 = v0 * N;
    if (tid < N * P) {
        const int ix = tid / P;
        const int q = tid - ix * P;
        DD acc;
        if (q < pvalid) {
            const int iv = q - (q / N) * N;
            const int64_t base = (int64_t)ex * s.sxe + ix * s.sx + (int64_t)eu * s.sue + gv0 + q;
#pragma unroll
            for (int iu = 0; iu < N; ++iu) acc.add(mul_rn(f[base + iu * s.su], tab.wq[iu * N + iv]));
        }
        red_hi[tid] = acc.hi;
        red_lo[tid] = acc.lo;
    }
    __syncthreads();
    if (tid < N) {
        DD acc;
        for (int q = 0; q < P; ++q) {
            DD t;
            t.hi = red_hi[tid * P + q];
            t.lo = red_lo[tid * P + q];
            acc.add(t);
        }
        const int tiles_per_ex = s.nu * s.nchunk;
        const int64_t slot = ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + chunk) * N + tid;
        s.partials[slot] = make_double2(acc.hi, acc.lo);
    }
}

}  // namespace vl
