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
// x.  While it computes tile ex from LDS, the loads of tile ex+1, of its u-face
// halo and of the first x-plane of tile ex+2 (the right x-halo one step ahead)
// are already in flight in registers; the left x-halo is the last x-plane of
// the previous tile, copied inside LDS.  So every element of the stage input
// is fetched from HBM once; u-halos are read from L2 because the neighbouring
// pencils (eu +- 1) run on the same XCD at the same time.
//
// Per tile each thread contracts one x-line, one u-line and one v-line with
// the n x n volume operator (kernel-argument memory, read through the scalar
// cache) and accumulates J_d * L_d into an LDS output tile in the reference's
// order ((Lx Jx + Lu Ju) + Lv Jv).  The epilogue combines with the stage
// inputs, writes rows that are contiguous along v, and reduces the moment
// partial of the tile.
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

__device__ __forceinline__ Pencil decode_pencil(const StageArgs &s) {
    const int p = xcd_remap(blockIdx.x, gridDim.x);
    Pencil c;
    c.chunk = p % s.nchunk;
    const int t2 = p / s.nchunk;
    c.eu = 1 + t2 % s.nu;
    const int seg = t2 / s.nu;
    c.ex_begin = 1 + seg * s.xseg;
    c.ex_end = min(s.nx, c.ex_begin + s.xseg - 1);
    return c;
}

template <int N, int MODE, int S, int L>
__global__ __launch_bounds__(kBlock, 3) void k_stage(
    const typename Store<S>::E *__restrict__ fin,  // stage input (interior; ghosts too when !periodic_x)
    const typename Store<S>::E *__restrict__ fgh,  // array whose u/v ghost planes are used
    const typename Store<S>::E *__restrict__ fn,   // f^n for MODE == kModeStage (may alias fout)
    typename Store<S>::E *__restrict__ fout,
    const double *__restrict__ E,      // (nx+2)*N ghosted field
    const double *__restrict__ unodes, // (nu+2)*N
    const double *__restrict__ vnodes, // (nv+2)*N
    StageArgs s, const OpTables *__restrict__ tabp) {
    using ST = Store<S>;
    using T = typename ST::E;
    using C = typename ST::C;
    constexpr bool BG = ST::kDelta;
    using TS = TileShape<N>;
    constexpr int TV = TS::TV, P = TS::P, PP = TS::PP;
    __shared__ C ft[N * N * PP];           // tile rows (ix,iu) x v-nodes with halo
    __shared__ C xl[N * P];                // left x halo  [iu][q]
    __shared__ C xr[N * P];                // right x halo [iu][q]
    __shared__ C uh[2 * N * P];            // u halo [side][ix][q]
    __shared__ C acc_l[N * N * P];         // output accumulator rows (ix,iu) x q
    __shared__ double bgs[BG ? 3 * N * P : 1];       // background slab [table][iu][q]
    __shared__ double red_own[BG ? N * P : 1];       // moment scratch (FP64 form reuses uh)
    double *red = BG ? red_own : reinterpret_cast<double *>(uh);

    const int tid = threadIdx.x;
    const Pencil pc = decode_pencil(s);
    const int eu = pc.eu;
    const int v0 = 1 + pc.chunk * TV;                    // first v element of the tile
    const int nvalid = min(TV, s.nv + 1 - v0);           // valid v elements
    const int pvalid = nvalid * N;
    const int gv0 = v0 * N;                              // first v node of the tile
    // offsets inside one x element fit in 32 bits (checked at handle creation)
    const int sx = (int)s.sx, su = (int)s.su, sue = (int)s.sue;
    const int base_u = eu * sue;
    // pencil-constant parts of the speeds (main/fluxes.py:110,120-121,143)
    const int lx_iu = tid / P, lx_q = tid - (tid / P) * P;        // x-line / u-line owner
    const int lv_r = tid / TV, lv_el = tid - (tid / TV) * TV;     // v-line owner
    const bool lx_on = tid < N * P && lx_q < pvalid;
    const bool lv_on = tid < N * N * TV && lv_el < nvalid;
    const double ax = lx_on ? unodes[eu * N + lx_iu] : 0.0;
    const double bv = lx_on ? s.bfield * vnodes[gv0 + lx_q] : 0.0;
    const double av = lv_on ? s.acc * (-s.bfield * unodes[eu * N + (lv_r - (lv_r / N) * N)]) : 0.0;
    double e_cur = lx_on ? E[pc.ex_begin * N + lx_iu] : 0.0;     // lx_iu is ix for the u-line

    auto xmap = [&](int e) {
        if (s.periodic_x) {
            if (e == 0) return s.nx;
            if (e == s.nx + 1) return 1;
        }
        return e;
    };

    constexpr int NT = N * N * PP;
    constexpr int IT_T = (NT + kBlock - 1) / kBlock;
    constexpr int NH = 2 * N * P;
    constexpr int IT_H = (NH + kBlock - 1) / kBlock;
    constexpr int NX = N * P;
    constexpr int IT_X = (NX + kBlock - 1) / kBlock;
    constexpr int NO = N * N * P;
    constexpr int IT_O = (NO + kBlock - 1) / kBlock;
    C rt[IT_T], ru[IT_H], rr[IT_X];

    // Per-thread byte offsets of every load/store slot inside one x element
    // (pencil-invariant; kOob marks an empty slot).  Tile and u-halo slots
    // whose node lies in a u/v ghost element are read from fgh instead of fin
    // (bit it of the gh masks); that only happens on pencils at the velocity
    // boundary.
    constexpr int ES = (int)sizeof(T);
    int o_t[IT_T], o_h[IT_H], o_x[IT_X], o_o[IT_O];
    unsigned gh_t = 0u, gh_h = 0u;
#pragma unroll
    for (int it = 0; it < IT_T; ++it) {
        const int idx = tid + it * kBlock;
        o_t[it] = kOob;
        if (idx < NT) {
            const int r = idx / PP;
            const int q = idx - r * PP;
            if (q <= pvalid + 1) {
                const int ix = r / N;
                const int iu = r - ix * N;
                const int gv = gv0 - 1 + q;
                const int ev = gv / N;
                o_t[it] = (ix * sx + base_u + iu * su + gv) * ES;
                if (ev == 0 || ev == s.nv + 1) gh_t |= 1u << it;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT_H; ++it) {
        const int idx = tid + it * kBlock;
        o_h[it] = kOob;
        if (idx < NH) {
            const int side = idx / (N * P);
            const int rem = idx - side * N * P;
            const int ix = rem / P;
            const int q = rem - ix * P;
            if (q < pvalid) {
                const int eun = side ? eu + 1 : eu - 1;
                const int iun = side ? 0 : N - 1;
                o_h[it] = (ix * sx + eun * sue + iun * su + gv0 + q) * ES;
                if (eun == 0 || eun == s.nu + 1) gh_h |= 1u << it;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < IT_X; ++it) {
        const int idx = tid + it * kBlock;
        o_x[it] = kOob;
        if (idx < NX) {
            const int iu = idx / P;
            const int q = idx - iu * P;
            if (q < pvalid) o_x[it] = (base_u + iu * su + gv0 + q) * ES;
        }
    }
#pragma unroll
    for (int it = 0; it < IT_O; ++it) {
        const int idx = tid + it * kBlock;
        o_o[it] = kOob;
        if (idx < NO) {
            const int r = idx / P;
            const int q = idx - r * P;
            if (q < pvalid) {
                const int ix = r / N;
                const int iu = r - ix * N;
                o_o[it] = (ix * sx + base_u + iu * su + gv0 + q) * ES;
            }
        }
    }
    // element-major layout: own elements as 16-byte vectors, v-halo values
    constexpr int EP = elem_pitch(N);
    constexpr int VW = ST::VW;
    constexpr int NVS = (TV * EP) / VW;
    constexpr int IT_V = (NVS + kBlock - 1) / kBlock;
    constexpr int NVH = 2 * N * N;
    constexpr int IT_VH = (NVH + kBlock - 1) / kBlock;
    const int vrow = s.nv + 2;                       // elements per (ex, eu) row
    int o_v[L == kElemLayout ? IT_V : 1], o_vh[L == kElemLayout ? IT_VH : 1];
    unsigned gh_vh = 0u;
    vl_u32x4 rv[L == kElemLayout ? IT_V : 1];
    C rvh[L == kElemLayout ? IT_VH : 1];
    if constexpr (L == kElemLayout) {
        const int elem0 = (eu * vrow + v0) * EP;
        const int nvs_valid = nvalid * EP / VW;
#pragma unroll
        for (int it = 0; it < IT_V; ++it) {
            const int slot = tid + it * kBlock;
            o_v[it] = slot < nvs_valid ? (elem0 + slot * VW) * ES : kOob;
        }
#pragma unroll
        for (int it = 0; it < IT_VH; ++it) {
            const int idx = tid + it * kBlock;
            o_vh[it] = kOob;
            if (idx < NVH) {
                const int side = idx / (N * N);
                const int r = idx - side * N * N;
                const int ev = side ? v0 + nvalid : v0 - 1;
                o_vh[it] = ((eu * vrow + ev) * EP + r * N + (side ? 0 : N - 1)) * ES;
                if (ev == 0 || ev == s.nv + 1) gh_vh |= 1u << it;
            }
        }
#pragma unroll
        for (int it = 0; it < IT_H; ++it) {
            const int idx = tid + it * kBlock;
            o_h[it] = kOob;
            gh_h &= ~(1u << it);
            if (idx < NH) {
                const int side = idx / (N * P);
                const int rem = idx - side * N * P;
                const int ix = rem / P;
                const int q = rem - ix * P;
                if (q < pvalid) {
                    const int el = q / N;
                    const int iv = q - el * N;
                    const int eun = side ? eu + 1 : eu - 1;
                    const int iun = side ? 0 : N - 1;
                    o_h[it] = ((eun * vrow + v0 + el) * EP + (ix * N + iun) * N + iv) * ES;
                    if (eun == 0 || eun == s.nu + 1) gh_h |= 1u << it;
                }
            }
        }
#pragma unroll
        for (int it = 0; it < IT_X; ++it) {
            const int idx = tid + it * kBlock;
            o_x[it] = kOob;
            if (idx < NX) {
                const int iu = idx / P;
                const int q = idx - iu * P;
                if (q < pvalid) {
                    const int el = q / N;
                    const int iv = q - el * N;
                    o_x[it] = ((eu * vrow + v0 + el) * EP + iu * N + iv) * ES;
                }
            }
        }
    }
    const int xplane_bytes = (L == kElemLayout ? N * N : sx) * ES;   // ix -> ix+1
    const int64_t xbytes = s.sxe * ES;              // one x element of the state
    auto rs = [&](const T *base, int e) { return make_rsrc(base + (int64_t)e * s.sxe, xbytes); };

    // tile rows of element e (+ v halo nodes)
    auto load_tile = [&](int e) {
        const __amdgpu_buffer_rsrc_t rin = rs(fin, e);
        if constexpr (L == kElemLayout) {
#pragma unroll
            for (int it = 0; it < IT_V; ++it) rv[it] = buf_ld16(rin, o_v[it]);
#pragma unroll
            for (int it = 0; it < IT_VH; ++it) rvh[it] = ST::bld(rin, o_vh[it]);
            if (gh_vh) {
                const __amdgpu_buffer_rsrc_t rgh = rs(fgh, e);
#pragma unroll
                for (int it = 0; it < IT_VH; ++it)
                    if (gh_vh & (1u << it)) rvh[it] = ST::bld(rgh, o_vh[it]);
            }
            return;
        }
#pragma unroll
        for (int it = 0; it < IT_T; ++it) rt[it] = ST::bld(rin, o_t[it]);
        if (gh_t) {
            const __amdgpu_buffer_rsrc_t rgh = rs(fgh, e);
#pragma unroll
            for (int it = 0; it < IT_T; ++it)
                if (gh_t & (1u << it)) rt[it] = ST::bld(rgh, o_t[it]);
        }
    };
    // u-face halo of element e: plane iu=N-1 of eu-1 and iu=0 of eu+1
    auto load_uhalo = [&](int e) {
        const __amdgpu_buffer_rsrc_t rin = rs(fin, e);
#pragma unroll
        for (int it = 0; it < IT_H; ++it) ru[it] = ST::bld(rin, o_h[it]);
        if (gh_h) {
            const __amdgpu_buffer_rsrc_t rgh = rs(fgh, e);
#pragma unroll
            for (int it = 0; it < IT_H; ++it)
                if (gh_h & (1u << it)) ru[it] = ST::bld(rgh, o_h[it]);
        }
    };
    // x-plane ixn of element e (interior v nodes of the tile)
    auto load_xplane = [&](int e, int ixn) {
        const __amdgpu_buffer_rsrc_t rin = rs(fin, e);
        const int add = ixn * xplane_bytes;
#pragma unroll
        for (int it = 0; it < IT_X; ++it) rr[it] = ST::bld(rin, o_x[it] == kOob ? kOob : o_x[it] + add);
    };
    auto store_tile = [&]() {
        if constexpr (L == kElemLayout) {
            // scatter the contiguous element runs into the (ix,iu) x q rows
#pragma unroll
            for (int it = 0; it < IT_V; ++it) {
                const int slot = tid + it * kBlock;
                if (slot < NVS) {
#pragma unroll
                    for (int k = 0; k < VW; ++k) {
                        const int g = slot * VW + k;
                        const int el = g / EP;
                        const int rem = g - el * EP;
                        // elements past nvalid are absent: their slot el*N + 1
                        // would overwrite the right v-halo node at pvalid + 1
                        if (rem < N * N * N && el < nvalid) {
                            const int r = rem / N;            // ix*N + iu
                            const int iv = rem - r * N;
                            ft[r * PP + el * N + iv + 1] = ST::unpack(rv[it], k);
                        }
                    }
                }
            }
#pragma unroll
            for (int it = 0; it < IT_VH; ++it) {
                const int idx = tid + it * kBlock;
                if (idx < NVH) {
                    const int side = idx / (N * N);
                    const int r = idx - side * N * N;
                    ft[r * PP + (side ? pvalid + 1 : 0)] = rvh[it];
                }
            }
        } else {
#pragma unroll
            for (int it = 0; it < IT_T; ++it) {
                const int idx = tid + it * kBlock;
                if (idx < NT) ft[idx] = rt[it];
            }
        }
#pragma unroll
        for (int it = 0; it < IT_H; ++it) {
            const int idx = tid + it * kBlock;
            if (idx < NH) uh[idx] = ru[it];
        }
    };
    auto store_plane = [&](C *dst) {
#pragma unroll
        for (int it = 0; it < IT_X; ++it) {
            const int idx = tid + it * kBlock;
            if (idx < NX) dst[idx] = rr[it];
        }
    };

    // ---- prologue: first tile, its halos, the background slab ------------
    load_tile(pc.ex_begin);
    load_uhalo(pc.ex_begin);
    load_xplane(xmap(pc.ex_begin - 1), N - 1);
    if (BG) {
        const int64_t tbl = (int64_t)(s.nu + 2) * N * (s.nv + 2) * N;
        for (int idx = tid; idx < 3 * N * P; idx += kBlock) {
            const int t = idx / (N * P);
            const int rem = idx - t * N * P;
            const int iu = rem / P;
            const int q = rem - iu * P;
            double v = 0.0;
            if (q < pvalid) v = s.bg[t * tbl + ((int64_t)eu * N + iu) * (s.nv + 2) * N + gv0 + q];
            bgs[idx] = v;
        }
    }
    store_tile();
    store_plane(xl);
    load_xplane(xmap(pc.ex_begin + 1), 0);
    store_plane(xr);
    __syncthreads();

    const int tiles_per_ex = s.nu * s.nchunk;
    for (int ex = pc.ex_begin; ex <= pc.ex_end; ++ex) {
        const bool has_next = ex < pc.ex_end;
        const __amdgpu_buffer_rsrc_t rfn = rs(fn, ex);
        const __amdgpu_buffer_rsrc_t rout = rs(fout, ex);
        // prefetch the next tile, its u halo and the right x halo of the next step
        double e_next = 0.0;
        if (has_next) {
            load_tile(ex + 1);
            load_uhalo(ex + 1);
            load_xplane(xmap(ex + 2), 0);
            if (lx_on) e_next = E[(ex + 1) * N + lx_iu];
        }
        C rn[L == kElemLayout ? 1 : IT_O];
        vl_u32x4 rnv[L == kElemLayout ? IT_V : 1];
        if (MODE == kModeStage) {
            if constexpr (L == kElemLayout) {
#pragma unroll
                for (int it = 0; it < IT_V; ++it) rnv[it] = buf_ld16(rfn, o_v[it]);
            } else {
#pragma unroll
                for (int it = 0; it < IT_O; ++it) rn[it] = ST::bld(rfn, o_o[it]);
            }
        }

        // ---- x-lines: (iu, q), line over ix ------------------------------
        if (lx_on) {
            const int iu = lx_iu;
            const int q = lx_q;
            {
                const double ad = ax;
                C f[N], o[N];
#pragma unroll
                for (int j = 0; j < N; ++j) f[j] = ft[(j * N + iu) * PP + q + 1];
                line_op<N, C>(f, xl[iu * P + q], xr[iu * P + q], (C)ad, ad >= 0.0, tabp, o);
#pragma unroll
                for (int i = 0; i < N; ++i) acc_l[(i * N + iu) * P + q] = o[i] * (C)s.jx;
            }
        }
        __syncthreads();
        // ---- u-lines: (ix, q), line over iu ------------------------------
        if (lx_on) {
            const int ix = lx_iu;
            const int q = lx_q;
            {
                const double ad = s.acc * (e_cur + bv);
                const bool pos = ad >= 0.0;
                C f[N], o[N];
#pragma unroll
                for (int j = 0; j < N; ++j) f[j] = ft[(ix * N + j) * PP + q + 1];
                line_op<N, C>(f, uh[ix * P + q], uh[(N + ix) * P + q], (C)ad, pos, tabp, o);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    C add = o[i] * (C)s.ju;
                    if (BG) add += (C)(s.scale * (ad * bgs[(pos ? 0 : N * P) + i * P + q]));
                    acc_l[(ix * N + i) * P + q] += add;
                }
            }
        }
        __syncthreads();
        // ---- v-lines: (ix, iu, ev_local), line over iv -------------------
        if (lv_on) {
            const int r = lv_r;
            const int el = lv_el;
            {
                const int iu = r - (r / N) * N;
                const double ad = av;
                const C *row = &ft[r * PP + el * N];
                C f[N], o[N];
#pragma unroll
                for (int j = 0; j < N; ++j) f[j] = row[j + 1];
                line_op<N, C>(f, row[0], row[N + 1], (C)ad, ad >= 0.0, tabp, o);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    C add = o[i] * (C)s.jv;
                    if (BG) add += (C)(s.scale * bgs[2 * N * P + iu * P + el * N + i]);
                    acc_l[r * P + el * N + i] += add;
                }
            }
        }
        __syncthreads();

        // ---- epilogue: combine and write rows (contiguous along v) -------
        // fn and fout may be the same array: each element is read and then
        // written by the same thread.
        if constexpr (L == kElemLayout) {
#pragma unroll
            for (int it = 0; it < IT_V; ++it) {
                const int slot = tid + it * kBlock;
                if (slot < NVS) {
                    vl_u32x4 w = {0u, 0u, 0u, 0u};
#pragma unroll
                    for (int k = 0; k < VW; ++k) {
                        const int g = slot * VW + k;
                        const int el = g / EP;
                        const int rem = g - el * EP;
                        if (rem < N * N * N) {
                            const int r = rem / N;
                            const int iv = rem - r * N;
                            const int q = el * N + iv;
                            const C Lr = acc_l[r * P + q];
                            C val;
                            if (MODE == kModeRhs) {
                                val = Lr;
                            } else if (MODE == kModeStage1) {
                                val = ft[r * PP + q + 1] + (C)s.c2dt * Lr;
                            } else {
                                val = ((C)s.c0 * ST::unpack(rnv[it], k) + (C)s.c1 * ft[r * PP + q + 1]) +
                                      (C)s.c2dt * Lr;
                            }
                            ST::pack(w, k, val);
                            acc_l[r * P + q] = ST::round(val);
                        }
                    }
                    buf_st16(rout, o_v[it], w);     // slots past the valid elements are dropped
                }
            }
        } else {
#pragma unroll
        for (int it = 0; it < IT_O; ++it) {
            const int idx = tid + it * kBlock;
            if (idx < NO) {
                const int r = idx / P;
                const int q = idx - r * P;
                const C Lr = acc_l[idx];
                C val;
                if (MODE == kModeRhs) {
                    val = Lr;
                } else if (MODE == kModeStage1) {
                    val = ft[r * PP + q + 1] + (C)s.c2dt * Lr;
                } else {
                    val = ((C)s.c0 * rn[it] + (C)s.c1 * ft[r * PP + q + 1]) + (C)s.c2dt * Lr;
                }
                ST::bst(rout, o_o[it], val);      // empty slots (q >= pvalid) are dropped
                acc_l[idx] = ST::round(val);     // the moment sees the value as stored
            }
        }
        }
        __syncthreads();

        // ---- moment partial of the written values; left halo of next tile
        if (MODE != kModeRhs && s.write_partials) {
            // plain FP64 inside the tile (a few hundred terms); the sums over
            // tiles are double-double in the reduction kernel
            if (tid < N * P) {
                const int ix = tid / P;
                const int q = tid - ix * P;
                double acc = 0.0;
                if (q < pvalid) {
                    const int iv = q - (q / N) * N;
                    ConstOpTables &tab = fresh(tabp);
#pragma unroll
                    for (int iu = 0; iu < N; ++iu)
                        acc += (double)acc_l[(ix * N + iu) * P + q] * tab.wq[iu * N + iv];
                }
                red[tid] = acc;
            }
            __syncthreads();
            // fixed two-level tree over the P columns of each x node
            constexpr int G = 4;
            constexpr int NG = (P + G - 1) / G;
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
            if (tid < N) {
                double acc = 0.0;
                for (int g = 0; g < NG; ++g) acc += red[tid * P + g];
                const int64_t slot =
                    ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + pc.chunk) * N + tid;
                s.partials[slot] = make_double2(BG ? acc * s.inv_scale : acc, 0.0);
            }
        }
        e_cur = e_next;
        if (has_next) {
            if (tid < N * P) {
                const int iu = tid / P;
                const int q = tid - iu * P;
                xl[tid] = ft[((N - 1) * N + iu) * PP + q + 1];
            }
            __syncthreads();
            store_tile();
            store_plane(xr);
            __syncthreads();
        }
    }
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
                    ? (int64_t)ex * s.sxe + ((int64_t)eu * (s.nv + 2) + v0 + el) * EP + (ix * N + iu) * N + iv
                    : (int64_t)ex * s.sxe + ix * s.sx + (int64_t)eu * s.sue + iu * s.su + gv0 + q;
                acc += (double)ST::ld(f, off) * tab.wq[iu * N + iv];
            }
        }
        red[tid] = acc;
    }
    __syncthreads();
    if (tid < N) {
        double acc = 0.0;
        for (int q = 0; q < P; ++q) acc += red[tid * P + q];
        const int tiles_per_ex = s.nu * s.nchunk;
        const int64_t slot = ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + chunk) * N + tid;
        s.partials[slot] = make_double2(ST::kDelta ? acc * s.inv_scale : acc, 0.0);
    }
}

// Layout conversion of a state array (element width W bytes via type U):
// to_elem = 1: reference -> `lay`, 0: `lay` -> reference, where `lay` is
// kElemLayout (element-major) or 2 (row-split, n = 8 or 9).  Padding values
// of the element-wise arrays are not written.
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
        const int inner = lay == kElemLayout ? (ix * n + iu) * n + iv
                                             : (iv < 8 ? (ix * n + iu) * 8 + iv
                                                       : 8 * n * n + (ix * n + iu) * (n - 8) + (iv - 8));
        const int64_t e = ex * sxe_e + ((int64_t)eu * (nv + 2) + ev) * ep + inner;
        if (to_elem)
            dst[e] = src[i];
        else
            dst[i] = src[e];
    }
}

}  // namespace vl
