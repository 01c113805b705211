// MFMA stage kernel of the 16-bit storage forms (BF16 / FP16 deviation
// g = (f - f0) * scale) in the element-major layout.  Same semantics as
// k_stage<N, MODE, S, kElemLayout> (fused upwind DG RHS + Shu-Osher stage
// combination + moment partials; main/fluxes.py:59-207,
// main/timestep.py:158-163,180-182, main/grid.py:319-321), with the line
// contractions on the matrix cores.
//
// Line operator as a matrix product.  For one line of n nodes with speed a
// (constant along the line), the upwind DG operator is
//   L / (J a) = up f + xi0 fL - xi1 fR,
//   fL = a >= 0 ? f_{e-1}[n-1] : f[0],   fR = a >= 0 ? f[n-1] : f_{e+1}[0].
// Over the extended line x_p (p = -1 .. n: left neighbour node, the n own
// nodes, right neighbour node) that is one row of A [xi0 | up | -xi1] against
// (fL, f_0 .. f_{n-1}, fR): fL is x_{-1} or x_0 and fR is x_{n-1} or x_n,
// chosen by the sign of a.  A 16x16x16 MFMA takes 16 lines as the columns
// of its B operand; lane group g (lanes 16g .. 16g+15) supplies the K slots
// 4g .. 4g+3, which here are the positions p = b_g + t (t = 0..3) of a
// window inside the extended line: b_g = min(4g - 1, n - 3) for the groups
// up to GR (the one whose window ends at p = n), b_g = -1 beyond.  Windows
// may overlap; the A column of a repeated position is zero.  The upwind
// choice is an address shift of two slots: p = -1 -> 0 (group 0, t = 0)
// when a < 0, p = n -> n-1 (group GR, t = 3) when a >= 0.  A is split into
// hi + lo 16-bit parts (two MFMAs into one FP32 accumulator), so the
// operator keeps about 16 more bits than the stored data; the stored
// values are exact MFMA inputs.
//
// Tile = (ex, eu, TV v-elements) as in k_stage.  The tile sits in LDS as one
// brick of 16-bit values over (x, u, V) = (-1..n, -1..n, -1..P) nodes
// (P = TV n v-nodes), so that each x-, u- and v-line of the tile is a
// uniform-stride run that includes its neighbour nodes.  MFMA results are
// scaled by J a (plus the background terms a_u G+- and Hv, FP64 tables
// kept as FP32 in LDS) and summed into an element-major FP32 accumulator in
// the reference's order ((Lx + Lu) + Lv): x writes, then u and v add, one
// barrier apart, each address read and written by one lane per phase.  The epilogue
// reads the accumulator as 16-byte vectors in the element-major order of
// the state, combines the stage, stores, and forms the moment partials of
// the stored values.
#pragma once
#include "vl_stage.h"

namespace vl {

typedef short vl_s16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 vl_h16x4 __attribute__((ext_vector_type(4)));
typedef float vl_f32x4 __attribute__((ext_vector_type(4)));
typedef float vl_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 vl_bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 vl_h16x2 __attribute__((ext_vector_type(2)));

// 16-bit MFMA format of a storage form.
template <int S> struct MxFmt;
template <> struct MxFmt<2> {          // BF16
    __device__ static vl_f32x4 mfma(vl_u32x2 a, vl_u32x2 b, vl_f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(vl_s16x4, a),
                                                          __builtin_bit_cast(vl_s16x4, b), c, 0, 0, 0);
    }
    __device__ static unsigned int enc(float x) { return __builtin_bit_cast(uint16_t, (__bf16)x); }
    __device__ static float dec(unsigned int b) { return (float)__builtin_bit_cast(__bf16, (uint16_t)b); }
    __device__ static unsigned int enc2(float a, float b) {     // round to nearest even, a in the low half
        return __builtin_bit_cast(unsigned int, __builtin_convertvector(vl_f32x2{a, b}, vl_bf16x2));
    }
};
template <> struct MxFmt<3> {          // FP16
    __device__ static vl_f32x4 mfma(vl_u32x2 a, vl_u32x2 b, vl_f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(vl_h16x4, a),
                                                     __builtin_bit_cast(vl_h16x4, b), c, 0, 0, 0);
    }
    __device__ static unsigned int enc(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
    __device__ static float dec(unsigned int b) { return (float)__builtin_bit_cast(_Float16, (uint16_t)b); }
    __device__ static unsigned int enc2(float a, float b) {
        return __builtin_bit_cast(unsigned int, __builtin_convertvector(vl_f32x2{a, b}, vl_h16x2));
    }
};

template <int N>
struct MxShape {
    static constexpr int TV = TileShape<N>::TV;
    static constexpr int P = TV * N;               // v-nodes per tile
    static constexpr int NB = N + 2;               // brick extent along x and u
    static constexpr int PB = P + 2;               // brick extent along v
    static constexpr int SX = NB * PB;             // brick strides (16-bit words)
    static constexpr int SU = PB;
    static constexpr int BRICK = NB * NB * PB;
    static constexpr int DUMMY = BRICK;            // sink for padding values
    static constexpr int BRICK_ALLOC = ((BRICK + 8 + 7) / 8) * 8;
    static constexpr int NL = TV * N * N;          // lines per direction
    static constexpr int MG = (NL + 15) / 16;      // MFMA column groups per direction
    static constexpr int MW = (MG + 3) / 4;        // groups per wave
    static constexpr int GR = (N + 1) / 4;         // lane group holding p = n  (= ceil((n-2)/4))
    static constexpr int EP = elem_pitch(N);
    static constexpr int DPG = N <= 8 ? 8 : 12;    // row pitch of the background tables
    static constexpr int VW = 8;                   // 16-bit values per 16-byte vector
    static constexpr int NVS = TV * EP / VW;       // vector slots per tile
    static constexpr int IT_V = (NVS + kBlock - 1) / kBlock;
    static constexpr int NH = 2 * N * P, IT_H = (NH + kBlock - 1) / kBlock;
    static constexpr int NX = N * P, IT_X = (NX + kBlock - 1) / kBlock;
    static constexpr int NVH = 2 * N * N, IT_VH = (NVH + kBlock - 1) / kBlock;
    static_assert(N + 2 <= 13, "extended line must fit the 16 K slots");
    static_assert(NL <= 16 * 16, "at most 4 column groups per wave");
    static_assert(BRICK_ALLOC < 32768, "brick offsets fit in 16 bits");
};

// first position of the K window of lane group g
template <int N>
__device__ __forceinline__ constexpr int mx_base(int g) {
    return g <= MxShape<N>::GR ? (4 * g - 1 < N - 3 ? 4 * g - 1 : N - 3) : -1;
}

template <int N, int MODE, int S, bool ADT = false>
__global__ __launch_bounds__(kBlock, 4) void k_stage_mx(
    const uint16_t *__restrict__ fin,   // stage input
    const uint16_t *__restrict__ fgh,   // array whose u/v ghost elements are used
    const uint16_t *fn,                 // f^n for MODE == kModeStage (may alias fout)
    uint16_t *fout,
    const double *__restrict__ E,       // (nx+2)*N ghosted field
    const double *__restrict__ unodes,  // (nu+2)*N
    const double *__restrict__ vnodes,  // (nv+2)*N
    StageArgs s, const OpTables *__restrict__ tabp) {
    static_assert(MODE == kModeStage1 || MODE == kModeStage, "stage modes only");
    if constexpr (ADT) stage_dt_from_device(s);   // adaptive runs: dt from the device clock
    using SH = MxShape<N>;
    using F = MxFmt<S>;
    constexpr int TV = SH::TV, P = SH::P, NB = SH::NB, PB = SH::PB, SX = SH::SX, SU = SH::SU;
    constexpr int NL = SH::NL, MG = SH::MG, MW = SH::MW, GR = SH::GR, EP = SH::EP, DPG = SH::DPG;
    constexpr int NVS = SH::NVS, IT_V = SH::IT_V, NH = SH::NH, IT_H = SH::IT_H, NX = SH::NX,
                  IT_X = SH::IT_X, NVH = SH::NVH, IT_VH = SH::IT_VH;
    constexpr int N2 = N * N, N3 = N * N * N;
    constexpr int ES = 2;

    constexpr int ACC_DUMMY = TV * EP;     // sink of the rows past n and of the lines past NL
    __shared__ __attribute__((aligned(16))) uint16_t bk[SH::BRICK_ALLOC];   // tile brick
    __shared__ __attribute__((aligned(16))) float acc[TV * EP + 4];        // element-major L
    __shared__ __attribute__((aligned(16))) float gs[2 * P * DPG];         // scale*G+-  [sign][q][iu]
    __shared__ __attribute__((aligned(16))) float hvs[N * TV * DPG];       // scale*Hv   [iu*TV+el][iv]
    // moment weights per position inside an element, split by whether the
    // node's ix equals the ix of the first node of its 8-value slot
    __shared__ __attribute__((aligned(16))) float wa[EP], wbw[EP];
    __shared__ double erow[N];                                             // E of the tile's x element
    __shared__ float jxs[N], jvs[N];       // Jx u(eu, iu), Jv acc (-B u(eu, iu))  (main/fluxes.py:110,143)
    __shared__ double bvs[P];              // B v(q)                               (main/fluxes.py:120-121)
    __shared__ __attribute__((aligned(16))) float mpart[2 * NVS];          // per-slot moment partials
    __shared__ double red[TV * N];

    const int tid = threadIdx.x;
    const int ln = tid & 63;
    const int wave = tid >> 6;
    const int lj = ln & 15;            // MFMA column (line within the group)
    const int lg = ln >> 4;            // MFMA K / row group
    const Pencil pc = decode_pencil(s, VL_STAGE_PENCIL(s));
    const int eu = pc.eu;
    const int v0 = 1 + pc.chunk * TV;
    const int nvalid = min(TV, s.nv + 1 - v0);
    const int pvalid = nvalid * N;
    const int gv0 = v0 * N;
    const int vrow = s.nv + 2;

    auto xmap = [&](int e) {
        if (s.periodic_x) {
            if (e == 0) return s.nx;
            if (e == s.nx + 1) return 1;
        }
        return e;
    };

    // ---- operator fragments: lane holds A[i = lj][slot 4 lg + t] ----------
    vl_u32x2 ahi, alo;
    {
        const int bg = mx_base<N>(lg);
        const int prev_end = lg == 0 ? -2 : mx_base<N>(lg - 1) + 3;
        unsigned int h[4], l[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int p = bg + t;
            const bool first = lg <= GR && p > prev_end;
            double c = 0.0;
            if (first && lj < N) {
                if (p == -1) c = tabp->xi0[lj];
                else if (p == N) c = -tabp->xi1[lj];
                else c = tabp->up[lj * N + p];
            }
            h[t] = F::enc((float)c);
            l[t] = F::enc((float)(c - (double)F::dec(h[t])));
        }
        ahi = vl_u32x2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
        alo = vl_u32x2{l[0] | (l[1] << 16), l[2] | (l[3] << 16)};
    }

    // Brick word offset of window slot t = 0 of a line, lane-constant part:
    // x-line (iu, q): cxb + iu SU + q; u-line (ix, q): cub + ix SX + q;
    // v-line (ix, iu, el): cvb + ix SX + iu SU + el N.
    const int wb = mx_base<N>(lg);
    const int cxb = (wb + 1) * SX + SU + 1;
    const int cub = SX + (wb + 1) * SU + 1;
    const int cvb = SX + SU + wb + 1;

    // ---- load slots (byte offsets inside one x element) ---------------------
    const int elem0 = (eu * vrow + v0) * EP;
    int o_v[IT_V];
    unsigned int bpos[IT_V][4];        // brick word of each 16-bit value, two per word
    {
        const int nvs_valid = nvalid * EP / 8;
#pragma unroll
        for (int it = 0; it < IT_V; ++it) {
            const int slot = tid + it * kBlock;
            o_v[it] = slot < nvs_valid ? (elem0 + slot * 8) * ES : kOob;
            const int el = (slot * 8) / EP;
            const int rem0 = slot * 8 - el * EP;
            unsigned int w[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int rem = rem0 + k;
                int pos = SH::DUMMY;
                if (slot < NVS && el < nvalid && rem < N3) {
                    const int ix = rem / N2, r2 = rem - (rem / N2) * N2;
                    const int iu = r2 / N, iv = r2 - (r2 / N) * N;
                    pos = ((ix + 1) * NB + iu + 1) * PB + el * N + iv + 1;
                }
                w[k] = (unsigned int)pos;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) bpos[it][k] = w[2 * k] | (w[2 * k + 1] << 16);
        }
    }
    int o_h[IT_H], p_h[IT_H];
    unsigned gh_h = 0u;
#pragma unroll
    for (int it = 0; it < IT_H; ++it) {
        const int idx = tid + it * kBlock;
        o_h[it] = kOob;
        p_h[it] = SH::DUMMY;
        if (idx < NH) {
            const int side = idx / (N * P);
            const int rem = idx - side * N * P;
            const int ix = rem / P;
            const int q = rem - ix * P;
            if (q < pvalid) {
                const int el = q / N, iv = q - (q / N) * N;
                const int eun = side ? eu + 1 : eu - 1;
                const int iun = side ? 0 : N - 1;
                o_h[it] = ((eun * vrow + v0 + el) * EP + (ix * N + iun) * N + iv) * ES;
                if (eun == 0 || eun == s.nu + 1) gh_h |= 1u << it;
                p_h[it] = ((ix + 1) * NB + (side ? N + 1 : 0)) * PB + q + 1;
            }
        }
    }
    int o_x[IT_X], p_xr[IT_X], p_xl[IT_X], p_xs[IT_X];
#pragma unroll
    for (int it = 0; it < IT_X; ++it) {
        const int idx = tid + it * kBlock;
        o_x[it] = kOob;
        p_xr[it] = p_xl[it] = p_xs[it] = SH::DUMMY;
        if (idx < NX) {
            const int iu = idx / P;
            const int q = idx - iu * P;
            if (q < pvalid) {
                const int el = q / N, iv = q - (q / N) * N;
                o_x[it] = ((eu * vrow + v0 + el) * EP + iu * N + iv) * ES;
                p_xr[it] = ((N + 1) * NB + iu + 1) * PB + q + 1;   // x = n   (right face)
                p_xl[it] = (0 * NB + iu + 1) * PB + q + 1;         // x = -1  (left face)
                p_xs[it] = (N * NB + iu + 1) * PB + q + 1;         // x = n-1 (source of the next left face)
            }
        }
    }
    int o_vh[IT_VH], p_vh[IT_VH];
    unsigned gh_vh = 0u;
#pragma unroll
    for (int it = 0; it < IT_VH; ++it) {
        const int idx = tid + it * kBlock;
        o_vh[it] = kOob;
        p_vh[it] = SH::DUMMY;
        if (idx < NVH) {
            const int side = idx / N2;
            const int r = idx - side * N2;
            const int ev = side ? v0 + nvalid : v0 - 1;
            o_vh[it] = ((eu * vrow + ev) * EP + r * N + (side ? 0 : N - 1)) * ES;
            if (ev == 0 || ev == s.nv + 1) gh_vh |= 1u << it;
            const int ix = r / N, iu = r - (r / N) * N;
            p_vh[it] = ((ix + 1) * NB + iu + 1) * PB + (side ? pvalid + 1 : 0);
        }
    }

    const int xplane_bytes = N2 * ES;
    const int64_t xbytes = s.sxe * ES;
    auto rs = [&](const uint16_t *base, int e) { return make_rsrc(base + (int64_t)e * s.sxe, xbytes); };

    vl_u32x4 rv[IT_V];
    unsigned int rvh[IT_VH], ru[IT_H], rr[IT_X];
    auto ld16 = [&](__amdgpu_buffer_rsrc_t r, int off) {
        return (unsigned int)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    };
    auto load_tile = [&](int e) {
        const __amdgpu_buffer_rsrc_t rin = rs(fin, e);
#pragma unroll
        for (int it = 0; it < IT_V; ++it) rv[it] = buf_ld16(rin, o_v[it]);
#pragma unroll
        for (int it = 0; it < IT_VH; ++it) rvh[it] = ld16(rin, o_vh[it]);
        if (gh_vh) {
            const __amdgpu_buffer_rsrc_t rgh = rs(fgh, e);
#pragma unroll
            for (int it = 0; it < IT_VH; ++it)
                if (gh_vh & (1u << it)) rvh[it] = ld16(rgh, o_vh[it]);
        }
    };
    auto load_uhalo = [&](int e) {
        const __amdgpu_buffer_rsrc_t rin = rs(fin, e);
#pragma unroll
        for (int it = 0; it < IT_H; ++it) ru[it] = ld16(rin, o_h[it]);
        if (gh_h) {
            const __amdgpu_buffer_rsrc_t rgh = rs(fgh, e);
#pragma unroll
            for (int it = 0; it < IT_H; ++it)
                if (gh_h & (1u << it)) ru[it] = ld16(rgh, o_h[it]);
        }
    };
    auto load_xplane = [&](int e, int ixn) {
        const __amdgpu_buffer_rsrc_t rin = rs(fin, e);
        const int add = ixn * xplane_bytes;
#pragma unroll
        for (int it = 0; it < IT_X; ++it) rr[it] = ld16(rin, o_x[it] == kOob ? kOob : o_x[it] + add);
    };
    // tile values, v halos and u halos of the loaded tile into the brick
    auto store_tile = [&]() {
#pragma unroll
        for (int it = 0; it < IT_V; ++it) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const unsigned int pw = bpos[it][k >> 1];
                const int pos = (int)((k & 1) ? (pw >> 16) : (pw & 0xffffu));
                const unsigned int w = lane(rv[it], k >> 1);
                bk[pos] = (uint16_t)((k & 1) ? (w >> 16) : w);
            }
        }
#pragma unroll
        for (int it = 0; it < IT_VH; ++it) bk[p_vh[it]] = (uint16_t)rvh[it];
#pragma unroll
        for (int it = 0; it < IT_H; ++it) bk[p_h[it]] = (uint16_t)ru[it];
    };

    // ---- prologue -----------------------------------------------------------
    load_tile(pc.ex_begin);
    load_uhalo(pc.ex_begin);
    {
        const int64_t tbl = (int64_t)(s.nu + 2) * N * (s.nv + 2) * N;
        const int64_t W = (int64_t)(s.nv + 2) * N;
        for (int idx = tid; idx < 2 * P * DPG; idx += kBlock) {
            const int sg = idx / (P * DPG);
            const int rem = idx - sg * P * DPG;
            const int q = rem / DPG, iu = rem - (rem / DPG) * DPG;
            double v = 0.0;
            if (iu < N && q < pvalid) v = s.scale * s.bg[sg * tbl + ((int64_t)eu * N + iu) * W + gv0 + q];
            gs[idx] = (float)v;
        }
        for (int idx = tid; idx < N * TV * DPG; idx += kBlock) {
            const int row = idx / DPG, iv = idx - (idx / DPG) * DPG;
            const int iu = row / TV, el = row - (row / TV) * TV;
            double v = 0.0;
            if (iv < N && el < nvalid)
                v = s.scale * s.bg[2 * tbl + ((int64_t)eu * N + iu) * W + gv0 + el * N + iv];
            hvs[idx] = (float)v;
        }
        for (int idx = tid; idx < EP; idx += kBlock) {
            float a = 0.f, b = 0.f;
            if (idx < N3) {
                const float w = (float)tabp->wq[idx - (idx / N2) * N2];
                if (idx / N2 == ((idx / 8) * 8) / N2) a = w; else b = w;
            }
            wa[idx] = a;
            wbw[idx] = b;
        }
        // padding entries are never written by the phases below: they stay
        // zero, so the padding of the stored state stays zero too
        for (int idx = tid; idx < TV * EP + 4; idx += kBlock) acc[idx] = 0.f;
        if (tid < N) {
            erow[tid] = E[pc.ex_begin * N + tid];
            const double u = unodes[eu * N + tid];
            jxs[tid] = (float)(s.jx * u);
            jvs[tid] = (float)(s.jv * (s.acc * (-s.bfield * u)));
        }
        for (int q = tid; q < P; q += kBlock) bvs[q] = s.bfield * vnodes[gv0 + min(q, pvalid - 1)];
    }
    store_tile();
    load_xplane(xmap(pc.ex_begin - 1), N - 1);
#pragma unroll
    for (int it = 0; it < IT_X; ++it) bk[p_xl[it]] = (uint16_t)rr[it];
    load_xplane(xmap(pc.ex_begin + 1), 0);
#pragma unroll
    for (int it = 0; it < IT_X; ++it) bk[p_xr[it]] = (uint16_t)rr[it];
    __syncthreads();

    const float c0 = (float)s.c0, c1 = (float)s.c1, c2 = (float)s.c2dt;
    const int tiles_per_ex = s.nu * s.nchunk;
    // window slot offsets of this lane: group 0 slot 0 takes p = 0 for a < 0,
    // group GR slot 3 takes p = n-1 for a >= 0
    const bool g0 = lg == 0, gR = lg == GR;

    auto bload = [&](int base, int stride, bool neg) {
        const int a0 = base + ((g0 && neg) ? stride : 0);
        const int a3 = base + 3 * stride - ((gR && !neg) ? stride : 0);
        const unsigned int w0 = (unsigned int)bk[a0] | ((unsigned int)bk[base + stride] << 16);
        const unsigned int w1 = (unsigned int)bk[base + 2 * stride] | ((unsigned int)bk[a3] << 16);
        return vl_u32x2{w0, w1};
    };
    auto contract = [&](vl_u32x2 b) {
        vl_f32x4 d = {0.f, 0.f, 0.f, 0.f};
        d = F::mfma(ahi, b, d);
        d = F::mfma(alo, b, d);
        return d;
    };

    // moment of one finished tile: per (el, ix) over its slots, then over el
    auto moment_l1 = [&]() {
        if (tid < TV * N) {
            const int el = tid / N, ix = tid - (tid / N) * N;
            double a = 0.0;
            if (el < nvalid) {
                const int sb = (el * EP + ix * N2) / 8;
                const int se = (el * EP + (ix + 1) * N2 - 1) / 8;
                for (int sl = sb; sl <= se; ++sl) {
                    const int sx0 = (sl * 8 - el * EP) / N2;
                    a += (double)mpart[2 * sl + (sx0 == ix ? 0 : 1)];
                }
            }
            red[tid] = a;
        }
    };
    auto moment_l2 = [&](int exm) {
        if (tid < N) {
            double a = 0.0;
            for (int el = 0; el < nvalid; ++el) a += red[el * N + tid];
            const int64_t slot = ((int64_t)(exm - 1) * tiles_per_ex + (eu - 1) * s.nchunk + pc.chunk) * N + tid;
            s.partials[slot] = make_double2(a * s.inv_scale, 0.0);
        }
    };
    const bool partials = s.write_partials != 0;
    const float ju = (float)s.ju;

    for (int ex = pc.ex_begin; ex <= pc.ex_end; ++ex) {
        const bool has_next = ex < pc.ex_end;
        const __amdgpu_buffer_rsrc_t rfn = rs(fn, ex);
        const __amdgpu_buffer_rsrc_t rout = rs(fout, ex);
        double e_next = 0.0;
        if (has_next) {
            load_tile(ex + 1);
            load_uhalo(ex + 1);
            load_xplane(xmap(ex + 2), 0);
            if (tid < N) e_next = E[(ex + 1) * N + tid];
        }
        vl_u32x4 rnv[IT_V];
        if (MODE == kModeStage) {
#pragma unroll
            for (int it = 0; it < IT_V; ++it) rnv[it] = buf_ld16(rfn, o_v[it]);
        }
        if (partials && ex > pc.ex_begin) moment_l1();      // previous tile
        // per-line addresses are recomputed in every iteration: an opaque
        // copy of the lane's column index keeps the compiler from hoisting
        // them all (12 column groups) into registers
        int ljv = lj;
        asm volatile("" : "+v"(ljv));
        // the x = n-1 layer of this tile is the left face of the next one
        unsigned int xs[IT_X];
#pragma unroll
        for (int it = 0; it < IT_X; ++it) xs[it] = bk[p_xs[it]];

        // ---- x: acc = Jx a_x Dx (rows past n and lines past NL go to a sink)
#pragma unroll
        for (int mi = 0; mi < MW; ++mi) {
            if (wave + 4 * mi < MG) {
                const int L = (wave + 4 * mi) * 16 + ljv;
                const int Lc = min(L, NL - 1);
                const int iu = Lc / P, q = Lc - (Lc / P) * P;
                const float ja = jxs[iu];
                const vl_f32x4 d = contract(bload(cxb + iu * SU + q, SX, ja < 0.f));
                const int el = q / N, iv = q - (q / N) * N;
                const int base = el * EP + iu * N + iv;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool ok = 4 * lg + t < N && L < NL;
                    acc[ok ? base + (4 * lg + t) * N2 : ACC_DUMMY] = d[t] * ja;
                }
            }
        }
        __syncthreads();
        if (partials && ex > pc.ex_begin) moment_l2(ex - 1);
        // ---- u: acc += a_u (Ju Du + scale G+-) ---------------------------
#pragma unroll
        for (int mi = 0; mi < MW; ++mi) {
            if (wave + 4 * mi < MG) {
                const int L = (wave + 4 * mi) * 16 + ljv;
                const int Lc = min(L, NL - 1);
                const int ix = Lc / P, q = Lc - (Lc / P) * P;
                const double a = s.acc * (erow[ix] + bvs[q]);
                const bool neg = a < 0.0;
                const float ja = (float)a;
                const vl_f32x4 d = contract(bload(cub + ix * SX + q, SU, neg));
                const vl_f32x4 g = *reinterpret_cast<const vl_f32x4 *>(
                    &gs[((neg ? P : 0) + q) * DPG + min(4 * lg, DPG - 4)]);
                const int el = q / N, iv = q - (q / N) * N;
                const int base = el * EP + ix * N2 + iv;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool ok = 4 * lg + t < N && L < NL;
                    const int ad = ok ? base + (4 * lg + t) * N : ACC_DUMMY;
                    acc[ad] += ja * (ju * d[t] + g[t]);
                }
            }
        }
        __syncthreads();
        // ---- v: acc += Jv a_v Dv + scale Hv ------------------------------
#pragma unroll
        for (int mi = 0; mi < MW; ++mi) {
            if (wave + 4 * mi < MG) {
                const int L = (wave + 4 * mi) * 16 + ljv;
                const int Lc = min(L, NL - 1);
                const int r = Lc / TV, el = Lc - (Lc / TV) * TV;
                const int ix = r / N, iu = r - (r / N) * N;
                const float ja = jvs[iu];
                const vl_f32x4 d = contract(bload(cvb + ix * SX + iu * SU + el * N, 1, ja < 0.f));
                const vl_f32x4 h = *reinterpret_cast<const vl_f32x4 *>(
                    &hvs[(iu * TV + el) * DPG + min(4 * lg, DPG - 4)]);
                const int base = el * EP + r * N;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const bool ok = 4 * lg + t < N && L < NL;
                    const int ad = ok ? base + 4 * lg + t : ACC_DUMMY;
                    acc[ad] += ja * d[t] + h[t];
                }
            }
        }
        __syncthreads();

        // ---- epilogue: combine, store, moment partials of stored values ----
        // padding values: state 0, accumulator 0, weights 0 -> stored 0
#pragma unroll
        for (int it = 0; it < IT_V; ++it) {
            const int slot = tid + it * kBlock;
            if (slot < NVS) {
                const int el = (slot * 8) / EP;
                const int rem0 = slot * 8 - el * EP;
                const vl_f32x4 l0 = *reinterpret_cast<const vl_f32x4 *>(&acc[slot * 8]);
                const vl_f32x4 l1 = *reinterpret_cast<const vl_f32x4 *>(&acc[slot * 8 + 4]);
                const vl_f32x4 wa0 = *reinterpret_cast<const vl_f32x4 *>(&wa[rem0]);
                const vl_f32x4 wa1 = *reinterpret_cast<const vl_f32x4 *>(&wa[rem0 + 4]);
                const vl_f32x4 wb0 = *reinterpret_cast<const vl_f32x4 *>(&wbw[rem0]);
                const vl_f32x4 wb1 = *reinterpret_cast<const vl_f32x4 *>(&wbw[rem0 + 4]);
                vl_u32x4 w;
                float m0 = 0.f, m1 = 0.f;
#pragma unroll
                for (int k2 = 0; k2 < 4; ++k2) {
                    float v2[2];
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int k = 2 * k2 + h;
                        const float Lr = k < 4 ? l0[k] : l1[k - 4];
                        const unsigned int pw = bpos[it][k2];
                        const float fi = F::dec(bk[h ? (pw >> 16) : (pw & 0xffffu)]);
                        if (MODE == kModeStage1) {
                            v2[h] = fi + c2 * Lr;
                        } else {
                            const unsigned int nw = lane(rnv[it], k2);
                            const float fnv = F::dec(h ? (nw >> 16) : (nw & 0xffffu));
                            v2[h] = (c0 * fnv + c1 * fi) + c2 * Lr;
                        }
                    }
                    const unsigned int b = F::enc2(v2[0], v2[1]);
                    set_lane(w, k2, b);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int k = 2 * k2 + h;
                        const float sv = F::dec(h ? (b >> 16) : (b & 0xffffu));
                        m0 += sv * (k < 4 ? wa0[k] : wa1[k - 4]);
                        m1 += sv * (k < 4 ? wb0[k] : wb1[k - 4]);
                    }
                }
                buf_st16(rout, o_v[it], w);     // slots past the valid elements are dropped
                mpart[2 * slot] = m0;
                mpart[2 * slot + 1] = m1;
            }
        }
        if (has_next) {
            store_tile();
#pragma unroll
            for (int it = 0; it < IT_X; ++it) {
                bk[p_xl[it]] = (uint16_t)xs[it];
                bk[p_xr[it]] = (uint16_t)rr[it];
            }
            if (tid < N) erow[tid] = e_next;
        }
        __syncthreads();
    }
    if (partials) {
        moment_l1();
        __syncthreads();
        moment_l2(pc.ex_end);
    }
}

}  // namespace vl
