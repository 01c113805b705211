// Stage kernel: semi-discrete upwind DG RHS of the 1D2V Vlasov equation,
// optionally fused with the Shu-Osher stage combination and the zeroth-moment
// partial sums of the values it writes.  One template serves the FP64 form
// (f stored and computed in FP64, the reference's arithmetic) and the
// reduced-precision forms (FP32 / BF16 / FP16 storage of g = (f - f0) * scale,
// FP32 arithmetic, FP64 background terms and moments).
//
// Reference semantics (crewsdw/Vlasov1D2V):
//   DGFlux.semi_discrete_rhs  main/fluxes.py:59-105
//   x_flux / u_flux / v_flux  main/fluxes.py:107-153
//   spatial/velocity_flux     main/fluxes.py:156-207 (upwind by sign(speed), >=0 positive)
//   stage combination         main/timestep.py:158-163,180-182
//   moment_zero_of_arr        main/grid.py:319-321
//
// Work decomposition.  A tile is (ex, eu, TV consecutive v-elements) with all
// n^3 nodes of each element, TV = floor(256 / n^2), so every direction has
// n*P <= 256 lines of n nodes (P = TV*n v-nodes).  One workgroup of 256
// threads owns a pencil of tiles (eu, v-chunk, x-segment) and walks it along
// x, in the direction of the pencil's x-line speeds (the u nodes of element
// eu): the upwind x halo of each tile is then the end x-plane of the previous
// tile of the walk, a register per thread, and no x halo is loaded.  (A pencil
// whose element holds both signs walks towards +x and loads the right x-plane
// at the start of each tile.)  While it computes a tile from LDS, the loads
// of the next tile of the walk and of its u-face halo are in flight in
// registers.  u-halos come from the neighbouring pencils (eu +- 1), which run
// on the same XCD at about the same time.
//
// Per tile each thread contracts one u-line, one v-line and one x-line with
// the n x n volume operator (kernel-argument memory, read through the scalar
// cache) and accumulates J_d * L_d into an LDS output tile as
// ((Lu Ju + Lv Jv) + Lx Jx).  The epilogue combines with the stage inputs,
// writes rows that are contiguous along v, and reduces the moment partial of
// the tile.
//
// Background of the reduced forms.  f0(u, v) is x-independent, so
// L(f) = L(g)/scale + L(f0) with L_x(f0) = 0, L_u(f0) = a_u * G+-(u, v)
// (G+ when a_u >= 0, G- otherwise; a_u = acc (E + B v) is constant along the
// u-line) and L_v(f0) = Hv(u, v).  G+, G-, Hv are FP64 tables; the slab of a
// pencil is loaded into LDS once.
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

// The operator tables are re-read from the scalar cache in every phase: the
// opaque copy of the pointer keeps the compiler from hoisting the n^2 FP64
// entries out of the x loop into vector registers.
__device__ __forceinline__ ConstOpTables &fresh(const OpTables *p) {
    ConstOpTables *c = (ConstOpTables *)p;
    asm volatile("" : "+s"(c));
    return *c;
}

// One 1-D upwind DG line operator (main/fluxes.py:107-207 for one line):
//   F_j = f_j * a
//   L_i = sum_j up[i][j] F_j - (xi0[i] * nfL + xi1[i] * nfR)
//   nfL = -(pos ? F_R(e-1) : F_L(e)),  nfR = (pos ? F_R(e) : F_L(e+1)),  pos = a >= 0
// The volume product uses the even-odd split of the centro-antisymmetric
// up (exact for the Lobatto operator): with Fe_j = F_j + F_{n-1-j},
// Fo_j = F_j - F_{n-1-j} (j < h = n/2),
//   up.F [i]     = sum_j eoP[i][j] Fo_j + sum_j eoQ[i][j] Fe_j + eoM[i] F_m
//   up.F [n-1-i] = sum_j eoP[i][j] Fo_j - sum_j eoQ[i][j] Fe_j - eoM[i] F_m
//   up.F [m]     = sum_j eoR[j] Fo_j                                (n odd)
// which halves the multiply-adds of the n x n product.
template <int N, typename C>
__device__ __forceinline__ void line_op(const C (&f)[N], C fl_nb, C fr_nb, C a, bool pos,
                                        const OpTables *tabp, C (&out)[N]) {
    constexpr int H = N / 2;
    constexpr bool ODD = (N % 2) == 1;
    constexpr int M = N / 2;
    C F[N];
#pragma unroll
    for (int j = 0; j < N; ++j) F[j] = f[j] * a;
    const C FL = fl_nb * a;   // neighbour-left element's right-face flux
    const C FR = fr_nb * a;   // neighbour-right element's left-face flux
    C nfl, nfr;
    if (pos) {
        nfl = -FL;
        nfr = F[N - 1];
    } else {
        nfl = -F[0];
        nfr = FR;
    }
    C Fe[H > 0 ? H : 1], Fo[H > 0 ? H : 1];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        Fe[j] = F[j] + F[N - 1 - j];
        Fo[j] = F[j] - F[N - 1 - j];
    }
    C vol[N];
#pragma unroll
    for (int i = 0; i < H; ++i) {
        ConstOpTables &tab = fresh(tabp);   // only this row's entries live in SGPRs
        C sp = 0, dq = 0;
#pragma unroll
        for (int j = 0; j < H; ++j) {
            sp += op_eoP<C>(tab, i * H + j) * Fo[j];
            dq += op_eoQ<C>(tab, i * H + j) * Fe[j];
        }
        if (ODD) dq += op_eoM<C>(tab, i) * F[M];
        vol[i] = sp + dq;
        vol[N - 1 - i] = sp - dq;
    }
    if (ODD) {
        ConstOpTables &tab = fresh(tabp);
        C mv = 0;
#pragma unroll
        for (int j = 0; j < H; ++j) mv += op_eoR<C>(tab, j) * Fo[j];
        vol[M] = mv;
    }
    ConstOpTables &tab = fresh(tabp);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const C lift = op_xi0<C>(tab, i) * nfl + op_xi1<C>(tab, i) * nfr;
        out[i] = vol[i] - lift;
    }
}

// Pencil geometry shared by the stage kernel and the host launcher.
struct Pencil {
    int chunk, eu, ex_begin, ex_end;
};

// Logical pencil p -> (v-chunk, eu, x-segment) of one run.
__device__ __forceinline__ Pencil decode_pencil(const StageArgs &s, int p) {
    Pencil c;
    c.chunk = p % s.nchunk;
    const int t2 = p / s.nchunk;
    c.eu = 1 + t2 % s.nu;
    const int seg = t2 / s.nu;
    c.ex_begin = 1 + seg * s.xseg;
    c.ex_end = min(s.nx, c.ex_begin + s.xseg - 1);
    return c;
}

// LDS arrays of one stage workgroup.  The single-run kernel declares them as
// separate __shared__ arrays; the ensemble kernel carves them out of one
// buffer (byte offsets below, the 16-byte aligned arrays first).
//   ft, acc  tile values and RHS accumulator, element-major like the state's
//            element-major layout: node (el, ix, iu, iv) at
//            el EP + elem_node(elem_x_minor(S), N, ix, iu, iv), so that the
//            16-byte runs of that layout move as whole vectors
//   vh       v-halo nodes of the tile's v-line rows: [side][a N + b], the row
//            (ix, iu) = (a, b), x-minor order (iu, ix) = (a, b)
//   uh       u-halo faces [side][el][ix][iv], x-minor order [side][el][iv][ix]
// (the x halos stay in registers: one value per x-line and thread)
// With one line per thread in the order (el, a, b), LDS reads of the x-lines
// and v-lines (old order) or of the x-lines and u-lines (x-minor order) are
// free of bank conflicts on 64 x 4-byte banks; the third direction has a
// two-way conflict.  The epilogue's 16-byte accesses are contiguous across
// lanes.
template <int N, int S>
struct StageLds {
    using C = typename Store<S>::C;
    static constexpr bool BG = Store<S>::kDelta;
    static constexpr int P = TileShape<N>::P;
    static constexpr int EP = elem_pitch(N);
    static constexpr int n_ft = TileShape<N>::TV * EP;   // tile values, element-major
    static constexpr int n_acc = n_ft;                   // RHS accumulator, same layout
    static constexpr int n_vh = 2 * N * N;               // v halo [side][row]
    static constexpr int n_uh = 2 * N * P;               // u halo [side][el][line]
    static constexpr int n_bgs = BG ? 3 * N * P : 1;     // background slab [table][iu][q]
    static constexpr int n_red = BG ? N * P : 1;         // moment scratch (the FP64 form reuses uh)
    static constexpr size_t o_ft = 0;
    static constexpr size_t o_acc = o_ft + (size_t)n_ft * sizeof(C);
    static constexpr size_t o_bgs = o_acc + (size_t)n_acc * sizeof(C);
    static constexpr size_t o_red = o_bgs + (size_t)n_bgs * 8;
    static constexpr size_t o_vh = o_red + (size_t)n_red * 8;
    static constexpr size_t o_uh = o_vh + (size_t)n_vh * sizeof(C);
    static constexpr size_t bytes = o_uh + (size_t)n_uh * sizeof(C);
    static_assert((EP * sizeof(C)) % 16 == 0, "ft / acc elements start 16-byte aligned");
    static_assert(BG || (o_uh % 8 == 0), "the FP64 moment scratch aliases uh");
};

// Sum of the P column partials red[ix * P + q] of each x node ix of a tile:
// groups of 4 columns, then the groups in order.  The vector-ALU stage kernel
// and k_moment_partials both reduce with it, so that the partials recomputed
// from a stored state equal those the stage kernel wrote, bit for bit (a
// step gives the same result whether it reuses the previous step's partials
// or recomputes them).  Block-wide: call from every thread once red[] is
// written; the sum is returned to threads tid < N.
template <int N, int P>
__device__ __forceinline__ double tile_column_sum(double *red, int tid) {
    constexpr int G = 4;
    constexpr int NG = (P + G - 1) / G;
    __syncthreads();
    const bool g_on = tid < N * NG;
    const int gix = g_on ? tid / NG : 0;
    const int gg = g_on ? tid - gix * NG : 0;
    double gacc = 0.0;
    if (g_on) {
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int q = gg * G + k;
            if (q < P) gacc += red[gix * P + q];
        }
    }
    __syncthreads();
    if (g_on) red[gix * P + gg] = gacc;
    __syncthreads();
    double acc = 0.0;
    if (tid < N)
        for (int g = 0; g < NG; ++g) acc += red[tid * P + g];
    return acc;
}

// Workgroups per CU the stage kernel is compiled for (a register budget; LDS
// sets its own limit).  The element-major FP64 stage-2/3 kernel at n = 9 fits
// 4 per CU (40 KB of LDS, 128 registers).  The element-major FP32 first stage
// at n = 9 (28 KB of LDS, no f^n) fits 5 within 96 registers and runs 5 %
// faster there than at 4 (its stages 2/3 spill at 96 and stay at 3).  The
// others are compiled for 3: the reference-layout kernels would spill at 128,
// and the FP64 first stage ran faster with the larger budget (measured on
// MI355X).
template <int N, int S, int L, int MODE = kModeStage>
constexpr int stage_min_blocks() {
    if (N == 9 && S == 0 && L == kElemLayout && MODE == kModeStage) return 4;
    if (N == 9 && S == 1 && L == kElemLayout && MODE == kModeStage1) return 5;
    return 3;
}

template <int N, int MODE, int S, int L, bool ADT = false>
__global__ __launch_bounds__(kBlock, (stage_min_blocks<N, S, L, MODE>())) void k_stage(
    const typename Store<S>::E *__restrict__ fin,  // stage input (interior; ghosts too when !periodic_x)
    const typename Store<S>::E *__restrict__ fgh,  // array whose u/v ghost planes are used
    const typename Store<S>::E *__restrict__ fn,   // f^n for MODE == kModeStage (may alias fout)
    typename Store<S>::E *__restrict__ fout,
    const double *__restrict__ E,      // (nx+2)*N ghosted field
    const double *__restrict__ unodes, // (nu+2)*N
    const double *__restrict__ vnodes, // (nv+2)*N
    StageArgs s, const OpTables *__restrict__ tabp) {
    __shared__ __attribute__((aligned(16))) typename Store<S>::C ft[StageLds<N, S>::n_ft];
    __shared__ __attribute__((aligned(16))) typename Store<S>::C acc_l[StageLds<N, S>::n_acc];
    __shared__ typename Store<S>::C vh[StageLds<N, S>::n_vh];
    __shared__ typename Store<S>::C uh[StageLds<N, S>::n_uh];
    __shared__ double bgs[StageLds<N, S>::n_bgs];
    __shared__ double red_own[StageLds<N, S>::n_red];
    if constexpr (ADT) stage_dt_from_device(s);   // adaptive runs: dt from the device clock
    const Pencil pc = decode_pencil(s, VL_STAGE_PENCIL(s));
    const int stage_tid = threadIdx.x;
#include "vl_stage_body.inc"
}

// Moment partials of a state array (no RHS), one workgroup per tile
// (ex, eu, v-chunk) in the partial layout the stage kernel writes.
// main/grid.py:303-321.  Reduced forms: partials of g / scale.
template <int N, int S, int L>
__global__ __launch_bounds__(kBlock) void k_moment_partials(const typename Store<S>::E *__restrict__ f,
                                                            StageArgs s, OpTables tab) {
    using ST = Store<S>;
    using TS = TileShape<N>;
    constexpr int TV = TS::TV, P = TS::P;
    constexpr int EP = elem_pitch(N);
    __shared__ double red[N * P];
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
        double acc = 0.0;
        if (q < pvalid) {
            const int iv = q - (q / N) * N;
            const int el = q / N;
#pragma unroll
            for (int iu = 0; iu < N; ++iu) {
                const int64_t off = L == kElemLayout
                    ? (int64_t)ex * s.sxe + ((int64_t)eu * (s.nv + 2) + v0 + el) * EP +
                          elem_node(elem_x_minor(S), N, ix, iu, iv)
                    : (int64_t)ex * s.sxe + ix * s.sx + (int64_t)eu * s.sue + iu * s.su + gv0 + q;
                acc += (double)ST::ld(f, off) * tab.wq[iu * N + iv];
            }
        }
        red[tid] = acc;
    }
    const double acc = tile_column_sum<N, P>(red, tid);
    if (tid < N) {
        const int tiles_per_ex = s.nu * s.nchunk;
        const int64_t slot = ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + chunk) * N + tid;
        s.partials[slot] = make_double2(ST::kDelta ? acc * s.inv_scale : acc, 0.0);
    }
}

// Layout conversion of a state array (element width W bytes via type U):
// to_elem = 1: reference -> `lay`, 0: `lay` -> reference, where `lay` is
// kElemLayout (element-major; node order by width: x fastest for the 8- and
// 4-byte FP64 / FP32 states, see elem_x_minor) or 2 (row-split, n = 8 or 9).
// Padding values of the element-wise arrays are not written.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_convert_layout(const U *__restrict__ src, U *__restrict__ dst,
                                                           int to_elem, int lay, int n, int nu, int nv,
                                                           int64_t total) {
    const int64_t ep = elem_pitch(n);
    const int64_t sxe_e = (int64_t)(nu + 2) * (nv + 2) * ep;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
        int64_t t = i;
        const int iv = (int)(t % n);
        t /= n;
        const int ev = (int)(t % (nv + 2));
        t /= (nv + 2);
        const int iu = (int)(t % n);
        t /= n;
        const int eu = (int)(t % (nu + 2));
        t /= (nu + 2);
        const int ix = (int)(t % n);
        const int64_t ex = t / n;
        const int inner = lay == kElemLayout ? elem_node(sizeof(U) >= 4, n, ix, iu, iv)
                                             : (iv < 8 ? (ix * n + iu) * 8 + iv
                                                       : 8 * n * n + (ix * n + iu) * (n - 8) + (iv - 8));
        const int64_t e = ex * sxe_e + ((int64_t)eu * (nv + 2) + ev) * ep + inner;
        if (to_elem)
            dst[e] = src[i];
        else
            dst[i] = src[e];
    }
}

// Layout conversion reference <-> element-major for the x-fastest node order
// of the FP64 / FP32 states, staged through LDS so that both sides move
// contiguous runs: one workgroup per (ex, ix-free, eu, iu) slab walks the v
// elements in chunks of kCvtEv.  The reference side of a chunk is n rows
// (ix = 0 .. n-1) of kCvtEv n consecutive values; the element side is, per
// v element, the n^2 values (iv, ix) of plane iu, one contiguous run.
// Grid: (nx + 2) (nu + 2) n workgroups.  Padding values are not written.
constexpr int kCvtEv = 8;
template <typename U, int N>
__global__ __launch_bounds__(kBlock) void k_convert_xminor(const U *__restrict__ src, U *__restrict__ dst,
                                                           int to_elem, int nu, int nv) {
    constexpr int RL = kCvtEv * N + 1;            // LDS row (one value of padding: bank spread)
    constexpr int EP = elem_pitch(N);
    __shared__ U buf[N * RL];
    const int b = blockIdx.x;
    const int iu = b % N;
    const int t2 = b / N;
    const int eu = t2 % (nu + 2);
    const int ex = t2 / (nu + 2);
    const int64_t vrow = (int64_t)(nv + 2) * N;                      // values of one (ix, eu, iu) row
    const int64_t su = vrow;                                         // ref strides
    const int64_t seu = su * N, six = seu * (nu + 2);
    const int64_t ref0 = (int64_t)ex * N * six + (int64_t)eu * seu + (int64_t)iu * su;   // ix = 0, ev = 0
    const int64_t el0 = (((int64_t)ex * (nu + 2) + eu) * (nv + 2)) * EP + (int64_t)iu * N * N;
    for (int ev0 = 0; ev0 < nv + 2; ev0 += kCvtEv) {
        const int nev = min(kCvtEv, nv + 2 - ev0);
        if (to_elem) {
            for (int idx = threadIdx.x; idx < N * kCvtEv * N; idx += kBlock) {
                const int ix = idx / (kCvtEv * N), q = idx - ix * (kCvtEv * N);
                if (q < nev * N) buf[ix * RL + q] = src[ref0 + ix * six + (int64_t)ev0 * N + q];
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < kCvtEv * N * N; idx += kBlock) {
                const int e = idx / (N * N), r = idx - e * (N * N);      // r = iv n + ix
                if (e < nev) {
                    const int iv = r / N, ix = r - iv * N;
                    dst[el0 + (int64_t)(ev0 + e) * EP + r] = buf[ix * RL + e * N + iv];
                }
            }
        } else {
            for (int idx = threadIdx.x; idx < kCvtEv * N * N; idx += kBlock) {
                const int e = idx / (N * N), r = idx - e * (N * N);
                if (e < nev) {
                    const int iv = r / N, ix = r - iv * N;
                    buf[ix * RL + e * N + iv] = src[el0 + (int64_t)(ev0 + e) * EP + r];
                }
            }
            __syncthreads();
            for (int idx = threadIdx.x; idx < N * kCvtEv * N; idx += kBlock) {
                const int ix = idx / (kCvtEv * N), q = idx - ix * (kCvtEv * N);
                if (q < nev * N) dst[ref0 + ix * six + (int64_t)ev0 * N + q] = buf[ix * RL + q];
            }
        }
        __syncthreads();
    }
}

}  // namespace vl
