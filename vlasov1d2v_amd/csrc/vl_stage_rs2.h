// Matrix-core stage kernel of the 16-bit storage forms, second layout of the
// work (row-split state layout, n = 9).  Same semantics as k_stage_rs: fused
// upwind DG RHS + Shu-Osher stage combination + moment partials of the
// stored deviation g = (f - f0) * scale (main/fluxes.py:59-207,
// main/timestep.py:158-163,180-182, main/grid.py:319-321).
//
// Where k_stage_rs accumulates all three directions in one 16 x 16 frame
// per x-node (9 x 9 of 256 outputs useful, transposes on the matrix core),
// this kernel gives every direction its natural frame: the 16 columns of a
// 16 x 16 x 16 product are 16 lines of that direction, its rows the line's
// output nodes, and K its 16-slot window [nodes 0..8 | copy of node 0 |
// upwinded faces F_L, F_R | 0..].  The line operands come from an LDS brick:
// v-lines are contiguous (plain 8-byte reads); u- and x-lines are strided
// and arrive through gfx950's LDS transpose read (ds_read_b64_tr_b16: in
// each 16-lane group, lane 4a+b receives element b of the 8-byte chunks
// addressed by lanes a, a+4, a+8, a+12 -- so lane a+4t points at four
// consecutive line columns of K slot 4g+t).  Per-line speeds (which differ
// between the columns of a product) multiply the products, 4 multiplies per
// lane and product.  Results go into an FP32 array L in storage order:
// the x-direction stores, the u-direction adds (LDS atomic add), and the
// v-direction reads that sum, adds its own scaled product and then the
// stage combination c1 fin + c0 f^n through identity-matrix products, and
// stores the stage result.  A dense epilogue (4 consecutive stored values
// per lane, 3 passes over the 183 units of an element) rounds, stores and
// forms the moment partials.
//
// Upwinding (main/fluxes.py:156-207, a >= 0 positive): the K slots 8..11
// hold (node 8, node 0, left neighbour's node 8, right neighbour's node 0);
// one byte permute of that lane's second dword makes (F_L, F_R) = (nbL, f8)
// or (f0, nbR) for its line's sign.  x: sign of u(iu) -- one per x-product;
// v: sign of a_v(iu) = -acc B u(iu) -- per column; u: sign of
// a_u = acc (E(ix) + B v(iv)) -- per column and tile.
//
// Brick (per element, 16-bit slots, rows of 12 v-slots = 6 dwords):
//   MAIN rows (xs, us < 9) at 6 (9 xs + us); x-extension rows (xs = 9..11,
//   us < 9) and u-extension rows (xs < 9, us = 9..11) follow.  Slot 9 of an
//   axis is a copy of node 0, slots 10 / 11 the neighbour elements' face
//   nodes.  Rows are linear in the block index of every pass, so the read
//   addresses of a product are a per-lane base plus a compile-time offset.
#pragma once
#include "vl_stage_rs.h"

namespace vl {

template <int N>
struct Rs2Shape {
    static_assert(N == 9, "second-layout row-split kernel: 9 nodes per axis");
    static constexpr int TV = 4;                      // elements per tile, one wave each
    static constexpr int EP = elem_pitch(N);
    static constexpr int ROWS = N * N;                // rows (= v-lines) per element
    static constexpr int NRI = (ROWS + 63) / 64;      // own rows per lane (loads)
    static constexpr int NS = 12;                     // brick slots per axis
    static constexpr int EX = NS - N;                 // extension slots (copy, 2 faces)
    static constexpr int RW = 6;                      // dwords per brick row
    static constexpr int XEXT = RW * N * N;           // x-extension rows
    static constexpr int UEXT = XEXT + RW * N * EX;   // u-extension rows
    static constexpr int BRICK = UEXT + RW * N * EX;  // dwords per element
    static constexpr int LP = 12;                     // floats per L row (v-slots 0..11)
    static constexpr int LSZ = LP * ROWS;              // floats of L per element
    static constexpr int VB = (ROWS + 15) / 16;       // v-products per element
    static constexpr int NROWU = 2 * ROWS;            // row units (4 values)
    static constexpr int PASSES = (NROWU + 63) / 64;  // epilogue passes over the row halves
    static constexpr int FNR = ROWS * 4;              // f^n staging: rows (dwords) ...
    static constexpr int FNW = ((ROWS * 4 + 64) + 3) & ~3;  // ... then the tail dwords
    static constexpr int NNB = 4 * N;                 // neighbour rows per element
    static constexpr int HT = N * 16;                 // v-background table floats (us, vs)
    static constexpr int MR = ROWS + 3;               // per-row moment sums (floats)
};
static_assert(Rs2Shape<9>::BRICK % 2 == 0, "brick rows must stay 8-byte aligned");
static_assert(Rs2Shape<9>::NROWU % 2 == 0, "the two halves of a row share a DPP lane pair");

// Brick row of slot (xs, us), dwords from the element's brick base.
template <int N>
__device__ __forceinline__ int rs2_row(int xs, int us) {
    using SH = Rs2Shape<N>;
    if (xs < N && us < N) return SH::RW * (N * xs + us);
    if (xs >= N) return SH::XEXT + SH::RW * (N * (xs - N) + us);
    return SH::UEXT + SH::RW * (SH::EX * xs + (us - N));
}

// Waves per SIMD the natural-frame kernel is compiled for (register budget).
#ifndef VL_RS2_WAVES
#define VL_RS2_WAVES 3
#endif

template <int N, int MODE, int S, bool ADT = false>
__global__ __launch_bounds__(kBlock, VL_RS2_WAVES) void k_stage_rs2(
    const uint16_t *__restrict__ fin,   // stage input
    const uint16_t *__restrict__ fgh,   // array whose u/v ghost elements are used
    const uint16_t *fn,                 // f^n for MODE == kModeStage (may alias fout)
    uint16_t *fout,
    const double *__restrict__ E,       // (nx+2)*N ghosted field
    const double *__restrict__ unodes,  // (nu+2)*N
    const double *__restrict__ vnodes,  // (nv+2)*N
    StageArgs s, const OpTables *__restrict__ tabp) {
    static_assert(MODE == kModeStage1 || MODE == kModeStage, "stage modes only");
    if constexpr (ADT) stage_dt_from_device(s);
    using SH = Rs2Shape<N>;
    using F = MxFmt<S>;
    using PK = RsPack<S>;
    constexpr int TV = SH::TV, EP = SH::EP, ROWS = SH::ROWS, NRI = SH::NRI, RW = SH::RW, LP = SH::LP,
                  EX = SH::EX, NNB = SH::NNB, FNR = SH::FNR, FNW = SH::FNW;

    __shared__ __attribute__((aligned(16))) unsigned int brick[TV * SH::BRICK];
    __shared__ __attribute__((aligned(16))) float lacc[TV * SH::LSZ];
    __shared__ __attribute__((aligned(16))) unsigned fnst[MODE == kModeStage ? TV * FNW : 4];
    __shared__ __attribute__((aligned(16))) float htab[TV * SH::HT];
    __shared__ float mrow[TV * SH::MR];
    __shared__ __attribute__((aligned(8))) float spd[TV * 2 * 16];

    const int tid = threadIdx.x;
    const int ln = tid & 63;
    const int el = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int c16 = ln & 15;            // product column
    const int g = ln >> 4;              // K group (operands) / row group (results)
    const Pencil pc = decode_pencil_rs(s, VL_STAGE_PENCIL(s));
    int eu, ev;
    const bool active = rs_wave_element<TV>(s, pc.chunk, el, eu, ev);
    const int vrow = s.nv + 2;
    const int64_t sxe = s.sxe;
    unsigned int *bk = brick + el * SH::BRICK;
    float *lw = lacc + el * SH::LSZ;
    float *ht = htab + el * SH::HT;
    float *mr = mrow + el * SH::MR;
    float *sp = spd + el * 32;
    ConstOpTables &tab = *(ConstOpTables *)tabp;

    auto ebase = [&](int ex_, int eu_, int ev_) { return (int64_t)ex_ * sxe + ((int64_t)eu_ * vrow + ev_) * EP; };
    auto xmap = [&](int e) {
        if (e == 0) return s.nx;
        if (e == s.nx + 1) return 1;
        return e;
    };

    // ---- per-lane constants -------------------------------------------------
    // operator fragments (host-built): lane holds op[i = c16][slot 4 g + t], hi / lo.  As
    // the A operand they map slots to nodes; the same registers as the B operand give
    // B[slot][node] = op[node][slot], i.e. the product with the data as A is transposed.
    vl_u32x2 ophi, oplo;
    {
        const vl_u32x4 fr = *reinterpret_cast<const vl_u32x4 *>(s.mxfrag + ((S == 2 ? 0 : 64) + ln) * 4);
        ophi = vl_u32x2{fr.x, fr.y};
        oplo = vl_u32x2{fr.z, fr.w};
    }
    // identity fragments c * I on the node slots (stage combination), hi + lo
    const float c0 = (float)s.c0, c1 = (float)s.c1;
    auto ident = [&](float c, vl_u32x2 &hi, vl_u32x2 &lo) {
        const float chi = F::dec(F::enc(c));
        const float clo = c - chi;
        unsigned h[4], l[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bool on = (4 * g + t == c16) && c16 < N;
            h[t] = on ? F::enc(chi) : 0u;
            l[t] = on ? F::enc(clo) : 0u;
        }
        hi = vl_u32x2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
        lo = vl_u32x2{l[0] | (l[1] << 16), l[2] | (l[3] << 16)};
    };
    vl_u32x2 id1h, id1l, id0h, id0l;
    ident(c1, id1h, id1l);
    ident(c0, id0h, id0l);
    const bool c1lo = F::dec(F::enc(c1)) != c1;        // uniform: the lo product is needed
    const bool c0lo = F::dec(F::enc(c0)) != c0;

    // transpose-read source role: lane (c16 = a + 4 t) points at K slot 4 g + t, v-quad a
    // (quad 3 repeats quad 2: rows 12..15 of the products are padding lines)
    const int src_a = min(c16 & 3, 2), src_t = c16 >> 2;
    const int kslot = min(4 * g + src_t, SH::NS - 1);   // slots 12..15: a finite stand-in (zero columns)
    const int u_base = (kslot < N ? RW * kslot : SH::UEXT + RW * (kslot - N)) + 2 * src_a;
    const int u_step = kslot < N ? RW * N : RW * EX;
    const int x_base = (kslot < N ? RW * N * kslot : SH::XEXT + RW * N * (kslot - N)) + 2 * src_a;
    // the lane's operand line after the transpose read: v-slot vs_a (rows 12..15 repeat 8..11)
    const int vs_a = 4 * min(c16 >> 2, 2) + (c16 & 3);
    // v-products: B operand of line (xs = c16, us = block) -- clamped to a finite row
    const int xsv = min(c16, SH::NS - 1);
    const int v_base = (xsv < N ? RW * N * xsv : SH::XEXT + RW * N * (xsv - N)) + 2 * min(g, 2);
    const int cl = min(c16, N - 1);                      // clamped column for L / f^n reads
    const bool out_ok = c16 < N && g < 3;                // lane holds results of a real line
    const unsigned selp = g == 2 ? kSelPos : kSelKeep, seln = g == 2 ? kSelNeg : kSelKeep;

    const bool sround = s.rounding != 0;
    const unsigned seedh = sround ? hash32(*s.seed * 0x9e3779b9u + 0x7f4a7c15u) : 0u;

    // uniform speeds (one per u-node, in the wave's LDS table sp[us] = (x, v)):
    // x-products c2 Jx u (main/fluxes.py:110), v-products c2 Jv a_v (main/fluxes.py:143);
    // their signs as bit masks
    unsigned xposm = 0u, vposm = 0u;
#pragma unroll
    for (int us = 0; us < N; ++us) {
        const double u = unodes[eu * N + us];
        xposm |= (u >= 0.0 ? 1u : 0u) << us;
        vposm |= (s.acc * (-s.bfield * u) >= 0.0 ? 1u : 0u) << us;
    }
    // u-products (data as A: rows = lines (xs, vs = 4 g + t), columns = u-nodes): the
    // B v part of the speed c2 Ju acc (E(xs) + B v(vs)) of each result row, and of the
    // lane's operand line (for its upwind choice).  c2 dt > 0 (checked by the host), so
    // a speed has the sign of its a_u.
    float pv4[4], gavg[4], gh[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int vs = min(4 * g + t, N - 1);
        pv4[t] = (float)(s.c2dt * s.ju * s.acc * (s.bfield * vnodes[ev * N + vs]));
        gavg[t] = 0.f;
        gh[t] = 0.f;
    }
    const float pva = (float)(s.c2dt * s.ju * s.acc * (s.bfield * vnodes[ev * N + min(vs_a, N - 1)]));
    // moment weights of the lane's row half (h = unit & 1 = ln & 1)
    float wrow[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) wrow[t] = tab.wv[4 * (ln & 1) + t];
    const float wlast = tab.wv[N - 1];

    // never-written brick slots stay zero (finite); per-pencil background tables
    if (active) {
        for (int i = ln; i < SH::BRICK; i += 64) bk[i] = 0u;
        const int64_t W = (int64_t)vrow * N;
        const int64_t tbl = (int64_t)(s.nu + 2) * N * W;
        // u background of the lane's results (us = c16, vs = 4 g + t): G(sign a_u) as
        // average + half difference, su G = su Gavg + |su| Gh (scale * A+-_u f0)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int vs = 4 * g + t;
            if (c16 < N && vs < N) {
                const int64_t o = ((int64_t)eu * N + c16) * W + (int64_t)ev * N + vs;
                const float gp = s.bg32[o], gm = s.bg32[tbl + o];
                gavg[t] = 0.5f * (gp + gm);
                gh[t] = 0.5f * (gp - gm);
            }
        }
        // v background, ht[us * 16 + vs]: scale * A_v f0 upwinded by a_v
        for (int i = ln; i < SH::HT; i += 64) {
            const int vs = i & 15, us = i >> 4;
            ht[i] = vs < N ? s.bg32[2 * tbl + ((int64_t)eu * N + us) * W + (int64_t)ev * N + vs] : 0.f;
        }
        if (ln < N) {
            const double u = unodes[eu * N + ln];
            sp[2 * ln] = (float)(s.c2dt * s.jx * u);
            sp[2 * ln + 1] = (float)(s.c2dt * s.jv * (s.acc * (-s.bfield * u)));
        }
    }

    // ---- global loads of one tile (as k_stage_rs) ------------------------------
    const int sbytes = (int)((int64_t)(s.nx + 2) * sxe * 2);
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin, sbytes);
    const __amdgpu_buffer_rsrc_t rgh = make_rsrc(fgh, sbytes);
    const __amdgpu_buffer_rsrc_t rfn = make_rsrc(fn, sbytes);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc(fout, sbytes);
    auto ld32 = [&](__amdgpu_buffer_rsrc_t rr, int off) {
        return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(rr, off, 0, 0);
    };
    const int nside = ln / N, nj = ln - (ln / N) * N;
    const bool nlane = ln < NNB;
    const int nrr = nside == 0 ? (N - 1) * N + nj : nside == 1 ? nj : nside == 2 ? nj * N + (N - 1) : nj * N;
    const bool nbr_gh = (nside == 2 && eu - 1 == 0) || (nside == 3 && eu + 1 == s.nu + 1);
    auto load = [&](int ex, RsLoads<N> &d) {
        if (!active) return;
        const int b = (int)(ebase(ex, eu, ev) * 2);
        const __amdgpu_buffer_rsrc_t rL = ev - 1 == 0 ? rgh : rin;
        const __amdgpu_buffer_rsrc_t rR = ev + 1 == s.nv + 1 ? rgh : rin;
        const int bl = (int)(ebase(ex, eu, ev - 1) * 2), br = (int)(ebase(ex, eu, ev + 1) * 2);
#pragma unroll
        for (int k = 0; k < NRI; ++k) {
            const int r = ln + 64 * k;
            const bool ok = r < ROWS;
            d.row[k] = buf_ld16(rin, ok ? b + r * 16 : kOob);
            d.tail[k] = ld32(rin, ok ? (b + (8 * ROWS + r) * 2) & ~3 : kOob);
            d.vnl[k] = ld32(rL, ok ? (bl + (8 * ROWS + r) * 2) & ~3 : kOob);
            d.vnr[k] = ld32(rR, ok ? br + r * 16 : kOob);
        }
        // the neighbour element of each side is uniform: four scalar bases and
        // a per-lane select; the neighbour rows come from fin, except on the
        // first / last u element, whose u neighbours are ghost elements read
        // from fgh by a second load of just those lanes (same registers)
        const int xl = ex - 1 == 0 ? s.nx : ex - 1, xr = ex + 1 == s.nx + 1 ? 1 : ex + 1;   // periodic x
        const int bx0 = (int)(ebase(xl, eu, ev) * 2), bx1 = (int)(ebase(xr, eu, ev) * 2);
        const int bu0 = (int)(ebase(ex, eu - 1, ev) * 2), bu1 = (int)(ebase(ex, eu + 1, ev) * 2);
        int bn = bu1;
        bn = nside == 2 ? bu0 : bn;
        bn = nside == 1 ? bx1 : bn;
        bn = nside == 0 ? bx0 : bn;
        const int orow = nlane ? bn + nrr * 16 : kOob, otail = nlane ? (bn + (8 * ROWS + nrr) * 2) & ~3 : kOob;
        d.nrow = buf_ld16(rin, orow);
        d.ntail = N == 9 ? ld32(rin, otail) : 0u;
        if (nbr_gh) {
            d.nrow = buf_ld16(rgh, orow);
            if (N == 9) d.ntail = ld32(rgh, otail);
        }
    };
    // registers -> brick
    auto fill = [&](const RsLoads<N> &d) {
        if (!active) return;
        auto put = [&](int xs, int us, const vl_u32x4 &R, unsigned w4, unsigned w5, bool faces) {
            unsigned *p = bk + rs2_row<N>(xs, us);
            *reinterpret_cast<vl_u32x2 *>(p) = vl_u32x2{R.x, R.y};
            *reinterpret_cast<vl_u32x2 *>(p + 2) = vl_u32x2{R.z, R.w};
            if (faces) *reinterpret_cast<vl_u32x2 *>(p + 4) = vl_u32x2{w4, w5};
            else p[4] = w4;
        };
        auto halves = [](unsigned a, int ha, unsigned b, int hb) {
            const unsigned sa = 2u * ha, sb = 4u + 2u * hb;
            return __builtin_amdgcn_perm(b, a, sa | ((sa + 1) << 8) | (sb << 16) | ((sb + 1) << 24));
        };
#pragma unroll
        for (int k = 0; k < NRI; ++k) {
            const int r = ln + 64 * k;
            if (r < ROWS) {
                const int ix = r / N, iu = r - (r / N) * N;
                const vl_u32x4 R = d.row[k];
                const unsigned w4 = halves(d.tail[k], r & 1, R.x, 0);      // slot 8 = node 8, slot 9 = node 0
                const unsigned w5 = halves(d.vnl[k], r & 1, d.vnr[k], 0);  // slots 10, 11: v faces
                put(ix, iu, R, w4, w5, true);
                if (iu == 0) put(ix, N, R, w4, w5, true);   // u-slot 9: copy of node 0
                if (ix == 0) put(N, iu, R, w4, w5, true);   // x-slot 9: copy of node 0
            }
        }
        if (nlane) {
            const int xs = nside == 0 ? N + 1 : nside == 1 ? N + 2 : nj;
            const int us = nside == 2 ? N + 1 : nside == 3 ? N + 2 : nj;
            const vl_u32x4 R = d.nrow;
            put(xs, us, R, halves(d.ntail, nrr & 1, R.x, 0), 0u, false);
        }
    };

    const int tiles_per_ex = s.nchunk;      // element groups per x element
    const bool partials = s.write_partials != 0;
    const vl_f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
    const float jua = (float)(s.c2dt * s.ju * s.acc);      // u speed offset per unit E
    const unsigned *fs = fnst + el * FNW;

    // moment weight of the u-node of each epilogue row (us = r % N)
    float wup[SH::PASSES];
#pragma unroll
    for (int p = 0; p < SH::PASSES; ++p) {
        const int r = (64 * p + ln) >> 1;
        wup[p] = 64 * p + ln < SH::NROWU ? tab.wu[r - N * ((r * 57) >> 9)] : 0.f;   // r % 9 (r < 128)
    }
    rs_wait_all();

    RsLoads<N> cur, nxt;
    load(pc.ex_begin, cur);
    for (int ex = pc.ex_begin; ex <= pc.ex_end; ++ex) {
        const bool has_next = ex < pc.ex_end;
        fill(cur);
        if (MODE == kModeStage && active) {
            // f^n rows of this element straight into LDS (row-split order)
            const int bfn = (int)(ebase(ex, eu, ev) * 2);
            unsigned *st0 = fnst + el * FNW;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rfn, (__attribute__((address_space(3))) void *)st0, 16,
                                                     bfn + ln * 16, 0, 0, kCachePolicyNT);
            if (ln < ROWS - 64)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rfn, (__attribute__((address_space(3))) void *)(st0 + 256),
                                                         16, bfn + (64 + ln) * 16, 0, 0, kCachePolicyNT);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rfn, (__attribute__((address_space(3))) void *)(st0 + FNR), 4,
                                                     ln < (ROWS + 1) / 2 ? bfn + 16 * ROWS + 4 * ln : kOob, 0, 0,
                                                     kCachePolicyNT);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (has_next) load(ex + 1, nxt);
        wave_sync_lds();
        if (active) {
            // ---- u: one product per x-node (block = xs), data as A: rows = lines
            // (xs, vs), columns = u-nodes; L[(xs, us = c16, vs = 4 g + t)] = Lu
#pragma unroll
            for (int b = 0; b < N; ++b) {
                const float q = __builtin_bit_cast(
                    float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, jua * (float)E[ex * N + b])));
                vl_u32x2 A = lds_tr16(bk + u_base + u_step * b);
                A.y = __builtin_amdgcn_perm(A.x, A.y, (q + pva >= 0.f) ? selp : seln);
                vl_f32x4 d = F::mfma(A, ophi, vl_f32x4{gavg[0], gavg[1], gavg[2], gavg[3]});
                d = F::mfma(A, oplo, d);                 // lane: u-node c16, lines vs = 4 g + t
                vl_f32x4 o;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float su = q + pv4[t];
                    o[t] = __builtin_fmaf(su, d[t], __builtin_fabsf(su) * gh[t]);
                }
                if (out_ok) *reinterpret_cast<vl_f32x4 *>(lw + LP * (N * b + c16) + 4 * g) = o;
            }
        }
        if (MODE == kModeStage) {
            // the staged f^n is in LDS: its loads were issued before the next tile's
            // prefetch, which may stay in flight
            if (has_next)
                rs_wait_vm<rs_prefetch_loads<N>()>();
            else
                rs_wait_vm<0>();
        }
        wave_sync_lds();
        if (active) {
            // ---- x and v: one product pair per u-node (block = us); results in the
            // layout lane = x-node c16, rows vs = 4 g + t, for both directions
#pragma unroll
            for (int b = 0; b < N; ++b) {
                vl_u32x2 A = lds_tr16(bk + x_base + RW * b);        // x-lines (us = b, vs), data as A
                A.y = __builtin_amdgcn_perm(A.x, A.y, ((xposm >> b) & 1u) ? selp : seln);
                vl_u32x2 B = *reinterpret_cast<const vl_u32x2 *>(bk + v_base + RW * b);   // v-line (c16, b)
                B.y = __builtin_amdgcn_perm(B.x, B.y, ((vposm >> b) & 1u) ? selp : seln);
                const vl_f32x4 hv = *reinterpret_cast<const vl_f32x4 *>(ht + b * 16 + 4 * g);
                vl_f32x4 dx = F::mfma(A, ophi, zero4);
                dx = F::mfma(A, oplo, dx);
                vl_f32x4 dv = F::mfma(ophi, B, hv);
                dv = F::mfma(oplo, B, dv);
                const vl_f32x4 lu = *reinterpret_cast<const vl_f32x4 *>(lw + LP * (N * cl + b) + 4 * g);
                const vl_f32x2 sxv = *reinterpret_cast<const vl_f32x2 *>(sp + 2 * b);   // (x, v) speeds of u-node b
                vl_f32x4 d;
#pragma unroll
                for (int t = 0; t < 4; ++t) d[t] = __builtin_fmaf(sxv.y, dv[t], __builtin_fmaf(sxv.x, dx[t], lu[t]));
                // stage combination: + c1 fin (+ c0 f^n) on the matrix core
                d = F::mfma(id1h, B, d);
                if (c1lo) d = F::mfma(id1l, B, d);
                if (MODE == kModeStage) {
                    const int fr = N * cl + b;
                    vl_u32x2 Bf = {0u, 0u};
                    if (g < 2) Bf = *reinterpret_cast<const vl_u32x2 *>(fs + 4 * fr + 2 * g);
                    else if (g == 2)
                        Bf.x = (unsigned)reinterpret_cast<const uint16_t *>(fs + FNR)[fr];
                    d = F::mfma(id0h, Bf, d);
                    if (c0lo) d = F::mfma(id0l, Bf, d);
                }
                if (out_ok) *reinterpret_cast<vl_f32x4 *>(lw + LP * (N * c16 + b) + 4 * g) = d;
            }
        }
        wave_sync_lds();
        // ---- epilogue: row halves (v-nodes 0..3 and 4..8) -----------------------
        const int64_t bo = ebase(ex, eu, ev);
        if (active) {
            unsigned rs = sround ? (hash32((((unsigned)((ex * (s.nu + 2) + eu) * vrow + ev)) << 6 | (unsigned)ln) ^
                                           seedh) |
                                    1u)
                                 : 0u;
            const int h = ln & 1;
#pragma unroll
            for (int p = 0; p < SH::PASSES; ++p) {
                const int u = 64 * p + ln;
                if (u < SH::NROWU) {
                    const int r = u >> 1;
                    const vl_f32x4 v = *reinterpret_cast<const vl_f32x4 *>(lw + LP * r + 4 * h);
                    unsigned w0, w1;
                    if (sround) {
                        rs = xorshift32(rs);
                        w0 = PK::sr2(v[0], v[1], rs);
                        rs = xorshift32(rs);
                        w1 = PK::sr2(v[2], v[3], rs);
                    } else {
                        w0 = F::enc2(v[0], v[1]);
                        w1 = F::enc2(v[2], v[3]);
                    }
                    __builtin_amdgcn_raw_buffer_store_b64(vl_u32x2{w0, w1}, rout, (int)(bo * 2) + 8 * u, 0,
                                                          kCachePolicyNT);
                    float m = wrow[0] * PK::lo(w0) + wrow[1] * PK::hi(w0) + wrow[2] * PK::lo(w1) + wrow[3] * PK::hi(w1);
                    if (h) {
                        const float v8 = lw[LP * r + (N - 1)];
                        unsigned t8;
                        if (sround) {
                            rs = xorshift32(rs);
                            t8 = PK::sr1(v8, rs);
                        } else {
                            t8 = F::enc(v8);
                        }
                        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)t8, rout, (int)(bo * 2) + (8 * ROWS + r) * 2,
                                                              0, 0);
                        m += wlast * PK::lo(t8);
                    }
                    // the row's two halves sit in lanes 2 r, 2 r + 1
                    m += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, m), 0xB1, 0xF,
                                                                            0xF, false));
                    if (!h) mr[r] = wup[p] * m;
                }
            }
        }
        // moment partial of this wave's element, one slot per x-node
        if (partials) {
            wave_sync_lds();
            if (ln < N) {
                double a = 0.0;
                if (active)
                    for (int iu = 0; iu < N; ++iu) a += (double)mr[ln * N + iu];
                const int64_t slot =
                    (((int64_t)(ex - 1) * tiles_per_ex + pc.chunk) * TV + el) * N + ln;
                s.partials[slot] = make_double2(a * s.inv_scale, 0.0);
            }
        }
        cur = nxt;
    }
}

}  // namespace vl
