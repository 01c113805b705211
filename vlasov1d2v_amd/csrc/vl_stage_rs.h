// Matrix-core stage kernel of the 16-bit storage forms (BF16 / FP16 deviation
// g = (f - f0) * scale) in the row-split layout, n = 8 or 9 nodes per axis.
// Same semantics as k_stage<N, MODE, S, L> (fused upwind DG RHS + Shu-Osher
// stage combination + moment partials; main/fluxes.py:59-207,
// main/timestep.py:158-163,180-182, main/grid.py:319-321).
//
// Row-split layout.  Elements are ordered and padded as in the element-major
// layout.  Inside an element the first eight nodes of every v-line (ix, iu)
// form one 16-byte row, rows ordered (ix, iu); for n = 9 the ninth node of
// every line follows in a tail block ordered (ix, iu).  Every row is one
// aligned vector in HBM and in LDS.
//
// Line slots.  Each 1-D line operator acts on a 16-slot window:
//   slots 0..7  own nodes 0..7 (n = 8: all nodes)
//   slot 8      own node n-1 (n = 9: node 8; n = 8: a copy of node 7)
//   slot 9      own node 0 (copy)
//   slot 10     left neighbour's node n-1 (nbL),  slot 11 right neighbour's node 0 (nbR)
//   slots 12..15 zero
// and the operator row i is [up[i][0..7] | (n = 9) up[i][8] | 0 | xi0[i] | -xi1[i] | 0..]
// (main/basis.py:110,176-188).  Upwinding (main/fluxes.py:181-207, a >= 0
// positive) selects the face values: a >= 0 uses (fL, fR) = (nbL, f_{n-1}),
// a < 0 uses (f_0, nbR).  Both candidates of each face sit in the same
// 16-bit pair of the same lane (slots 8..11), so the choice is one byte
// permutation of that lane's operand.  The operator is split into hi + lo
// 16-bit parts (two MFMAs into one FP32 accumulator).
//
// Brick.  Each wave owns one v-element of the tile and keeps it in LDS as a
// brick of 12 x-slots x 12 u-slots x 12 v-slots (the slot convention above
// on every axis: the face slots 10/11 of one axis hold the neighbour
// element's face planes).  One 8-byte LDS read gives a lane four v-slots of
// one (x, u) row.
//   v-direction: the rows of a fixed x-slot are the B operand of the MFMA
//     directly (columns = u-slots, K = v-slots).
//   u-direction: the same registers multiplied by the identity matrix give
//     the transposed 16 x 16 block (lane = v-slot, K = u-slots), exact; the
//     line operator then runs with the data as the A operand.  Its output
//     has the layout of the v-direction's output (lane = u-node, four
//     v-nodes), so the two add in registers.
//   x-direction: same transposition of a fixed-u-slot block (rows = x-slots);
//     results go to an FP32 LDS array X[ix][iu][v-slot] with one 16-byte
//     store per lane, and the u/v pass adds them.
// The speeds multiply the MFMA results (they are constant along each line);
// the background images of the reduced forms, scale * A+-_u f0 and
// scale * A_v f0, are the accumulator inputs of those MFMAs.
//
// Work decomposition.  A workgroup of 4 waves owns a pencil (eu, 4
// consecutive v-elements, x-segment) and walks it along x.  Waves are
// independent except for one barrier per tile for the moment partials.
// The global loads of the next tile are in flight while the current one is
// computed; the epilogue works on whole rows (16-byte loads and stores).
#pragma once
#include "vl_stage_mx.h"

namespace vl {

enum { kRowSplitLayout = 2 };

template <int N>
struct RsShape {
    static_assert(N == 8 || N == 9, "row-split stage kernel: 8 or 9 nodes per axis");
    static constexpr int TV = 4;                  // v-elements per tile, one wave each
    static constexpr int EP = elem_pitch(N);
    static constexpr int ROWS = N * N;            // 8-value rows per element
    static constexpr int NRI = (ROWS + 63) / 64;  // rows per lane
    static constexpr int RSD = 6;                 // brick row stride (dwords, 12 slots)
    static constexpr int PSD = 74;                // brick x-slot stride (dwords): 12 rows + 2
    static constexpr int ESD = 12 * PSD;          // brick of one element (dwords)
    static constexpr int XR = 12;                 // floats per X row (v-slots 0..11)
    static constexpr int XE = ROWS * XR;          // floats of X per element
    static constexpr int GMAX = N == 9 ? 2 : 1;   // last lane group holding an output node
    static constexpr int NNB = 4 * N;             // neighbour rows per tile element
};
// 8-byte reads at row stride 6 dwords and x-slot stride 74 dwords hit 24 distinct
// bank pairs for the 12 rows (or x-slots) a lane group reads.
static_assert(RsShape<9>::PSD % 32 == 10, "x-slot stride must keep 8-byte reads conflict-free");

// Byte-permute selectors for the face slots (slot 10, slot 11) of lane group 2,
// whose operand dwords are d0 = (slot 8, slot 9) and d1 = (slot 10, slot 11):
// __builtin_amdgcn_perm(d0, d1, sel).
constexpr unsigned kSelPos = 0x05040100u;   // (nbL, f_{n-1}): a >= 0
constexpr unsigned kSelNeg = 0x03020706u;   // (f_0, nbR):     a <  0
constexpr unsigned kSelKeep = 0x03020100u;  // d1 unchanged (other lane groups)

template <int S> struct RsPack;
template <> struct RsPack<2> {              // BF16
    // two FP32 values that are exact BF16 numbers -> BF16 pair (a in the low half)
    __device__ static unsigned exact2(float a, float b) {
        return __builtin_amdgcn_perm(__builtin_bit_cast(unsigned, b), __builtin_bit_cast(unsigned, a),
                                     0x07060302u);
    }
    __device__ static float lo(unsigned w) { return __builtin_bit_cast(float, w << 16); }
    __device__ static float hi(unsigned w) { return __builtin_bit_cast(float, w & 0xffff0000u); }
};
template <> struct RsPack<3> {              // FP16
    __device__ static unsigned exact2(float a, float b) {
        return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(a, b));
    }
    __device__ static float lo(unsigned w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
    __device__ static float hi(unsigned w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <int N>
struct RsLoads {
    vl_u32x4 row[RsShape<N>::NRI];   // own rows: nodes 0..7 of line (ix, iu)
    unsigned tail[RsShape<N>::NRI];  // own node 8 (n = 9)
    unsigned vnb[RsShape<N>::NRI];   // v-neighbour faces of the row: nbL | nbR << 16
    vl_u32x4 nrow;                   // one neighbour row (x or u face plane)
    unsigned ntail;
};

template <int N, int MODE, int S>
__global__ __launch_bounds__(kBlock, 4) void k_stage_rs(
    const uint16_t *__restrict__ fin,   // stage input
    const uint16_t *__restrict__ fgh,   // array whose u/v ghost elements are used
    const uint16_t *fn,                 // f^n for MODE == kModeStage (may alias fout)
    uint16_t *fout,
    const double *__restrict__ E,       // (nx+2)*N ghosted field
    const double *__restrict__ unodes,  // (nu+2)*N
    const double *__restrict__ vnodes,  // (nv+2)*N
    StageArgs s, const OpTables *__restrict__ tabp) {
    static_assert(MODE == kModeStage1 || MODE == kModeStage, "stage modes only");
    using SH = RsShape<N>;
    using F = MxFmt<S>;
    using PK = RsPack<S>;
    constexpr int TV = SH::TV, EP = SH::EP, ROWS = SH::ROWS, NRI = SH::NRI, RSD = SH::RSD, PSD = SH::PSD,
                  ESD = SH::ESD, XR = SH::XR, XE = SH::XE, GMAX = SH::GMAX, NNB = SH::NNB;

    __shared__ __attribute__((aligned(16))) unsigned int brick[TV * ESD];
    __shared__ __attribute__((aligned(16))) float xa[TV * XE];
    __shared__ float red[2][TV * ROWS];

    const int tid = threadIdx.x;
    const int ln = tid & 63;
    const int el = tid >> 6;            // this wave's element of the tile
    const int lj = ln & 15;             // MFMA column / row within a 16 x 16 block
    const int lg = ln >> 4;             // lane group: K slots / output rows 4 lg .. 4 lg + 3
    const Pencil pc = decode_pencil(s);
    const int eu = pc.eu;
    const int v0 = 1 + pc.chunk * TV;
    const int nvalid = min(TV, s.nv + 1 - v0);
    const int ev = v0 + el;
    const bool active = el < nvalid;    // wave-uniform
    const int vrow = s.nv + 2;
    const int64_t sxe = s.sxe;
    unsigned int *bk = brick + el * ESD;
    float *xw = xa + el * XE;

    auto ebase = [&](int ex_, int eu_, int ev_) { return (int64_t)ex_ * sxe + ((int64_t)eu_ * vrow + ev_) * EP; };
    auto xmap = [&](int e) {
        if (e == 0) return s.nx;
        if (e == s.nx + 1) return 1;
        return e;
    };

    // ---- per-wave constants ------------------------------------------------
    // operator fragments: lane holds op[i = lj][slot 4 lg + t], hi and lo parts
    vl_u32x2 ophi, oplo, idv;
    {
        unsigned h[4], l[4], id[4];
        const unsigned one = F::enc(1.0f);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int sl = 4 * lg + t;
            double c = 0.0;
            if (lj < N) {
                if (sl < 8) c = tabp->up[lj * N + sl];
                else if (sl == 8) c = N == 9 ? tabp->up[lj * N + 8] : 0.0;
                else if (sl == 10) c = tabp->xi0[lj];
                else if (sl == 11) c = -tabp->xi1[lj];
            }
            h[t] = F::enc((float)c);
            l[t] = F::enc((float)(c - (double)F::dec(h[t])));
            id[t] = sl == lj ? one : 0u;
        }
        ophi = vl_u32x2{h[0] | (h[1] << 16), h[2] | (h[3] << 16)};
        oplo = vl_u32x2{l[0] | (l[1] << 16), l[2] | (l[3] << 16)};
        idv = vl_u32x2{id[0] | (id[1] << 16), id[2] | (id[3] << 16)};
    }
    // background images (accumulator inputs), v-speed terms: lane column
    // iu = lj, rows iv = 4 lg + t of this wave's element
    vl_f32x4 gp = {0.f, 0.f, 0.f, 0.f}, gm = gp, hv = gp, pt = gp;
    float pl = 0.f, sv = 0.f;
    unsigned vsel = kSelKeep;
    const float jua = (float)(s.ju * s.acc);
    if (active) {
        const int64_t W = (int64_t)vrow * N;
        const int64_t tbl = (int64_t)(s.nu + 2) * N * W;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int iv = 4 * lg + t;
            if (iv < N) {
                pt[t] = (float)(s.ju * s.acc * (s.bfield * vnodes[ev * N + iv]));
                if (lj < N) {
                    const int64_t o = ((int64_t)eu * N + lj) * W + (int64_t)ev * N + iv;
                    gp[t] = s.bg32[o];
                    gm[t] = s.bg32[tbl + o];
                    hv[t] = s.bg32[2 * tbl + o];
                }
            }
        }
        const int jc = min(lj, N - 1);
        pl = (float)(s.ju * s.acc * (s.bfield * vnodes[ev * N + jc]));
        const double av = s.acc * (-s.bfield * unodes[eu * N + jc]);   // main/fluxes.py:143
        sv = (float)(s.jv * av);
        if (lg == 2) vsel = av >= 0.0 ? kSelPos : kSelNeg;
        // never-written brick slots stay zero (finite) for the whole kernel
        for (int i = ln; i < ESD; i += 64) bk[i] = 0u;
    }

    // ---- global loads of one tile ------------------------------------------
    auto load = [&](int ex, RsLoads<N> &d) {
        if (!active) return;
        const int64_t b = ebase(ex, eu, ev);
        const uint16_t *pL = ev - 1 == 0 ? fgh : fin;
        const uint16_t *pR = ev + 1 == s.nv + 1 ? fgh : fin;
        const int64_t bl = ebase(ex, eu, ev - 1), br = ebase(ex, eu, ev + 1);
#pragma unroll
        for (int k = 0; k < NRI; ++k) {
            const int r = ln + 64 * k;
            if (r < ROWS) {
                d.row[k] = *reinterpret_cast<const vl_u32x4 *>(fin + b + r * 8);
                d.tail[k] = N == 9 ? (unsigned)fin[b + 8 * ROWS + r] : 0u;
                const unsigned nl = N == 9 ? (unsigned)pL[bl + 8 * ROWS + r] : (unsigned)pL[bl + r * 8 + 7];
                const unsigned nr = (unsigned)pR[br + r * 8];
                d.vnb[k] = nl | (nr << 16);
            }
        }
        if (ln < NNB) {
            const int side = ln / N, j = ln - (ln / N) * N;
            int exn = ex, eun = eu, rr;
            const uint16_t *p = fin;
            if (side == 0) { exn = xmap(ex - 1); rr = (N - 1) * N + j; }        // x-left:  ix = n-1, iu = j
            else if (side == 1) { exn = xmap(ex + 1); rr = j; }                 // x-right: ix = 0,   iu = j
            else if (side == 2) { eun = eu - 1; rr = j * N + (N - 1); if (eun == 0) p = fgh; }        // u-left
            else { eun = eu + 1; rr = j * N; if (eun == s.nu + 1) p = fgh; }                          // u-right
            const int64_t bn = ebase(exn, eun, ev);
            d.nrow = *reinterpret_cast<const vl_u32x4 *>(p + bn + rr * 8);
            d.ntail = N == 9 ? (unsigned)p[bn + 8 * ROWS + rr] : 0u;
        }
    };
    // registers -> brick (slots 0..11 of every row that the passes read)
    auto fill = [&](const RsLoads<N> &d) {
        if (!active) return;
        auto put = [&](int xs, int us, const vl_u32x4 &R, unsigned w4, unsigned w5, bool faces) {
            unsigned *p = bk + xs * PSD + us * RSD;
            *reinterpret_cast<vl_u32x2 *>(p) = vl_u32x2{R.x, R.y};
            *reinterpret_cast<vl_u32x2 *>(p + 2) = vl_u32x2{R.z, R.w};
            if (faces) *reinterpret_cast<vl_u32x2 *>(p + 4) = vl_u32x2{w4, w5};
            else p[4] = w4;
        };
#pragma unroll
        for (int k = 0; k < NRI; ++k) {
            const int r = ln + 64 * k;
            if (r < ROWS) {
                const int ix = r / N, iu = r - (r / N) * N;
                const vl_u32x4 R = d.row[k];
                const unsigned last = N == 9 ? d.tail[k] : (R.w >> 16);
                const unsigned w4 = (last & 0xffffu) | (R.x << 16);   // slot 8 = node n-1, slot 9 = node 0
                put(ix, iu, R, w4, d.vnb[k], true);
                if (iu == 0) put(ix, 9, R, w4, d.vnb[k], true);       // u-slot 9: copy of node 0
                if (ix == 0) put(9, iu, R, w4, d.vnb[k], true);       // x-slot 9: copy of node 0
                if (N == 8 && iu == N - 1) put(ix, 8, R, w4, d.vnb[k], true);   // slot 8: copy of node n-1
                if (N == 8 && ix == N - 1) put(8, iu, R, w4, d.vnb[k], true);
            }
        }
        if (ln < NNB) {
            const int side = ln / N, j = ln - (ln / N) * N;
            const int xs = side == 0 ? 10 : side == 1 ? 11 : j;
            const int us = side == 2 ? 10 : side == 3 ? 11 : j;
            const vl_u32x4 R = d.nrow;
            const unsigned last = N == 9 ? d.ntail : (R.w >> 16);
            put(xs, us, R, (last & 0xffffu) | (R.x << 16), 0u, false);
        }
    };

    const float c0 = (float)s.c0, c1 = (float)s.c1, c2 = (float)s.c2dt;
    const int tiles_per_ex = s.nu * s.nchunk;
    const bool partials = s.write_partials != 0;
    const vl_f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};

    RsLoads<N> cur, nxt;
    load(pc.ex_begin, cur);
    int par = 0;
    for (int ex = pc.ex_begin; ex <= pc.ex_end; ++ex) {
        const bool has_next = ex < pc.ex_end;
        fill(cur);
        wave_sync_lds();
        if (has_next) load(ex + 1, nxt);
        const int64_t b = ebase(ex, eu, ev);
        vl_u32x4 fnr[NRI];
        unsigned fnt[NRI];
        if (MODE == kModeStage && active) {
#pragma unroll
            for (int k = 0; k < NRI; ++k) {
                const int r = ln + 64 * k;
                if (r < ROWS) {
                    fnr[k] = *reinterpret_cast<const vl_u32x4 *>(fn + b + r * 8);
                    fnt[k] = N == 9 ? (unsigned)fn[b + 8 * ROWS + r] : 0u;
                }
            }
        }
        if (active) {
            // ---- x: X[ix][iu][v-slot] = Jx u(iu) (A_x g) ----------------------
#pragma unroll
            for (int su = 0; su < N; ++su) {
                const double uu = unodes[eu * N + su];
                const float sxs = (float)(s.jx * uu);                       // main/fluxes.py:110
                const unsigned sel = lg == 2 ? (uu >= 0.0 ? kSelPos : kSelNeg) : kSelKeep;
                vl_u32x2 r = {0u, 0u};
                if (lj < 12 && lg < 3) r = *reinterpret_cast<const vl_u32x2 *>(bk + lj * PSD + su * RSD + 2 * lg);
                const vl_f32x4 tr = F::mfma(r, idv, zero4);   // lane: v-slot lj, x-slots 4 lg + t
                vl_u32x2 p = {PK::exact2(tr[0], tr[1]), PK::exact2(tr[2], tr[3])};
                p.y = __builtin_amdgcn_perm(p.x, p.y, sel);
                vl_f32x4 d = F::mfma(p, ophi, zero4);
                d = F::mfma(p, oplo, d);                      // lane: x-node lj, v-slots 4 lg + t
                if (lj < N && lg <= GMAX)
                    *reinterpret_cast<vl_f32x4 *>(xw + (lj * N + su) * XR + 4 * lg) = d * sxs;
            }
            wave_sync_lds();
            // ---- u and v: X += Ju a_u (A_u g + G+-) + Jv a_v (A_v g + Hv) -----
#pragma unroll
            for (int sx = 0; sx < N; ++sx) {
                const float q = jua * (float)E[ex * N + sx];
                vl_u32x2 r = {0u, 0u};
                if (lj < 12 && lg < 3) r = *reinterpret_cast<const vl_u32x2 *>(bk + sx * PSD + lj * RSD + 2 * lg);
                vl_u32x2 bvv = r;
                bvv.y = __builtin_amdgcn_perm(r.x, r.y, vsel);
                vl_f32x4 dv = F::mfma(ophi, bvv, hv);
                dv = F::mfma(oplo, bvv, dv);                  // lane: u-slot lj, v-nodes 4 lg + t
                const vl_f32x4 tr = F::mfma(r, idv, zero4);   // lane: v-slot lj, u-slots 4 lg + t
                vl_u32x2 p = {PK::exact2(tr[0], tr[1]), PK::exact2(tr[2], tr[3])};
                const float sl = q + pl;                       // Ju a_u of this lane's u-line
                p.y = __builtin_amdgcn_perm(p.x, p.y, lg == 2 ? (sl >= 0.f ? kSelPos : kSelNeg) : kSelKeep);
                vl_f32x4 su4, gsel;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    su4[t] = q + pt[t];
                    gsel[t] = su4[t] >= 0.f ? gp[t] : gm[t];
                }
                vl_f32x4 du = F::mfma(p, ophi, gsel);
                du = F::mfma(p, oplo, du);                    // lane: u-node lj, v-slots 4 lg + t
                if (lj < N && lg <= GMAX) {
                    float *xp = xw + (sx * N + lj) * XR + 4 * lg;
                    const vl_f32x4 xv = *reinterpret_cast<const vl_f32x4 *>(xp);
                    vl_f32x4 L;
#pragma unroll
                    for (int t = 0; t < 4; ++t) L[t] = __builtin_fmaf(sv, dv[t], __builtin_fmaf(su4[t], du[t], xv[t]));
                    *reinterpret_cast<vl_f32x4 *>(xp) = L;
                }
            }
            wave_sync_lds();
            // ---- epilogue: whole rows ----------------------------------------
            const int64_t bo = ebase(ex, eu, ev);
#pragma unroll
            for (int k = 0; k < NRI; ++k) {
                const int r = ln + 64 * k;
                if (r < ROWS) {
                    const int iu = r - (r / N) * N;
                    const float *xp = xw + r * XR;
                    const vl_f32x4 l0 = *reinterpret_cast<const vl_f32x4 *>(xp);
                    const vl_f32x4 l1 = *reinterpret_cast<const vl_f32x4 *>(xp + 4);
                    float Lr[9] = {l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3], N == 9 ? xp[8] : 0.f};
                    const vl_u32x4 R = cur.row[k];
                    const unsigned rw[4] = {R.x, R.y, R.z, R.w};
                    float fi[9], fv[9];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        fi[2 * h] = PK::lo(rw[h]);
                        fi[2 * h + 1] = PK::hi(rw[h]);
                    }
                    fi[8] = N == 9 ? PK::lo(cur.tail[k]) : 0.f;
                    if (MODE == kModeStage) {
                        const unsigned nw[4] = {fnr[k].x, fnr[k].y, fnr[k].z, fnr[k].w};
#pragma unroll
                        for (int h = 0; h < 4; ++h) {
                            fv[2 * h] = PK::lo(nw[h]);
                            fv[2 * h + 1] = PK::hi(nw[h]);
                        }
                        fv[8] = N == 9 ? PK::lo(fnt[k]) : 0.f;
                    }
                    float o[9];
#pragma unroll
                    for (int j = 0; j < 9; ++j)
                        o[j] = MODE == kModeStage1 ? fi[j] + c2 * Lr[j] : (c0 * fv[j] + c1 * fi[j]) + c2 * Lr[j];
                    vl_u32x4 w;
                    w.x = F::enc2(o[0], o[1]);
                    w.y = F::enc2(o[2], o[3]);
                    w.z = F::enc2(o[4], o[5]);
                    w.w = F::enc2(o[6], o[7]);
                    *reinterpret_cast<vl_u32x4 *>(fout + bo + r * 8) = w;
                    unsigned t8 = 0u;
                    if (N == 9) {
                        t8 = F::enc(o[8]);
                        fout[bo + 8 * ROWS + r] = (uint16_t)t8;
                    }
                    // moment partial of the stored values (main/grid.py:319-321)
                    const unsigned ww[4] = {w.x, w.y, w.z, w.w};
                    ConstOpTables &tab = *(ConstOpTables *)tabp;
                    float m = 0.f;
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        m += tab.wv[2 * h] * PK::lo(ww[h]);
                        m += tab.wv[2 * h + 1] * PK::hi(ww[h]);
                    }
                    if (N == 9) m += tab.wv[8] * PK::lo(t8);
                    red[par][el * ROWS + r] = tab.wu[iu] * m;
                }
            }
        }
        if (partials) {
            __syncthreads();
            if (tid < N) {
                double a = 0.0;
                for (int e2 = 0; e2 < nvalid; ++e2)
                    for (int iu = 0; iu < N; ++iu) a += (double)red[par][e2 * ROWS + tid * N + iu];
                const int64_t slot = ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + pc.chunk) * N + tid;
                s.partials[slot] = make_double2(a * s.inv_scale, 0.0);
            }
        }
        cur = nxt;
        par ^= 1;
    }
}

// Moment partials of a row-split state (first stage of a loop), in the
// partial layout of k_stage_rs: one workgroup per tile (ex, eu, 4 v-elements).
template <int N, int S>
__global__ __launch_bounds__(kBlock) void k_moment_rs(const uint16_t *__restrict__ f, StageArgs s,
                                                      const OpTables *__restrict__ tabp) {
    using SH = RsShape<N>;
    using PK = RsPack<S>;
    constexpr int TV = SH::TV, EP = SH::EP, ROWS = SH::ROWS, NRI = SH::NRI;
    __shared__ float red[TV * ROWS];
    const int tid = threadIdx.x, ln = tid & 63, el = tid >> 6;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile % s.nchunk;
    const int t2 = tile / s.nchunk;
    const int eu = 1 + t2 % s.nu;
    const int ex = 1 + t2 / s.nu;
    const int v0 = 1 + chunk * TV;
    const int nvalid = min(TV, s.nv + 1 - v0);
    ConstOpTables &tab = *(ConstOpTables *)tabp;
    if (el < nvalid) {
        const int64_t b = (int64_t)ex * s.sxe + ((int64_t)eu * (s.nv + 2) + v0 + el) * EP;
#pragma unroll
        for (int k = 0; k < NRI; ++k) {
            const int r = ln + 64 * k;
            if (r < ROWS) {
                const vl_u32x4 R = *reinterpret_cast<const vl_u32x4 *>(f + b + r * 8);
                const unsigned ww[4] = {R.x, R.y, R.z, R.w};
                float m = 0.f;
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    m += tab.wv[2 * h] * PK::lo(ww[h]);
                    m += tab.wv[2 * h + 1] * PK::hi(ww[h]);
                }
                if (N == 9) m += tab.wv[8] * PK::lo((unsigned)f[b + 8 * ROWS + r]);
                red[el * ROWS + r] = tab.wu[r - (r / N) * N] * m;
            }
        }
    }
    __syncthreads();
    if (tid < N) {
        double a = 0.0;
        for (int e2 = 0; e2 < nvalid; ++e2)
            for (int iu = 0; iu < N; ++iu) a += (double)red[e2 * ROWS + tid * N + iu];
        const int tiles_per_ex = s.nu * s.nchunk;
        const int64_t slot = ((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + chunk) * N + tid;
        s.partials[slot] = make_double2(a * s.inv_scale, 0.0);
    }
}

}  // namespace vl
