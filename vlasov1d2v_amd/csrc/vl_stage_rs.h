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
//   u-direction: a transposing LDS read (ds_read_b64_tr_b16) of the same
//     x-slot's rows gives the 16 x 16 block with lane = v-slot, K = u-slots;
//     the line operator then runs with the data as the A operand.  Its
//     output has the layout of the v-direction's output (lane = u-node,
//     four v-nodes), so the two add in registers.
//   x-direction: the same transposing read of a fixed-u-slot block (rows =
//     x-slots); results go to an FP32 LDS array X[ix][iu][v-slot] with one
//     16-byte store per lane, and the u/v pass adds them.
// The speeds multiply the MFMA results (they are constant along each line);
// the background images of the reduced forms, scale * A+-_u f0 and
// scale * A_v f0, are the accumulator inputs of those MFMAs.
//
// Work decomposition.  A workgroup of 4 waves owns a pencil (eu, 4
// consecutive v-elements, x-segment) and walks it along x.  The waves are
// independent (no barriers): each writes the moment partial of its own
// element, one slot per tile element.
// The global loads of the next tile are in flight while the current one is
// computed; the epilogue works on whole rows (16-byte loads and stores).
// Stages 2/3 read f^n of the current element with direct-to-LDS buffer
// loads issued at the start of the tile (no registers held across the
// passes); the epilogue waits for them and reads the rows from LDS.
#pragma once
#include "vl_stage_mx.h"

namespace vl {

enum { kRowSplitLayout = 2 };

// Waves per SIMD the row-split stage kernel is compiled for (register budget
// 512 / VL_RS_WAVES per lane).
#ifndef VL_RS_WAVES
#define VL_RS_WAVES 4
#endif
// 1: no instruction scheduling across the x / u-v tasks of one tile
#ifndef VL_RS_TASK_FENCE
#define VL_RS_TASK_FENCE 0
#endif
__device__ __forceinline__ void task_fence() {
#if VL_RS_TASK_FENCE
    __builtin_amdgcn_sched_barrier(0);
#endif
}

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

typedef __bf16 vl_bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 vl_f16x2 __attribute__((ext_vector_type(2)));

template <int S> struct RsPack;
template <> struct RsPack<2> {              // BF16
    // two FP32 values that are exact BF16 numbers -> BF16 pair (a in the low half)
    __device__ static unsigned exact2(float a, float b) {
        return __builtin_amdgcn_perm(__builtin_bit_cast(unsigned, b), __builtin_bit_cast(unsigned, a),
                                     0x07060302u);
    }
    __device__ static float lo(unsigned w) { return __builtin_bit_cast(float, w << 16); }
    __device__ static float hi(unsigned w) { return __builtin_bit_cast(float, w & 0xffff0000u); }
    // stochastic rounding: add 16 random bits below the BF16 mantissa and
    // truncate (rounds up with probability equal to the discarded fraction).
    // gfx950's v_cvt_sr_bf16_f32 does exactly that with the HIGH 16 bits of
    // its random operand: a takes the low half of r, b the high half.
    __device__ static unsigned sr2(float a, float b, unsigned r) {
        vl_bf16x2 v = {};
        v = __builtin_amdgcn_cvt_sr_bf16_f32(v, a, r << 16, false);
        v = __builtin_amdgcn_cvt_sr_bf16_f32(v, b, r, true);
        return __builtin_bit_cast(unsigned, v);
    }
    __device__ static unsigned sr1(float a, unsigned r) {
        vl_bf16x2 v = {};
        v = __builtin_amdgcn_cvt_sr_bf16_f32(v, a, r << 16, false);
        return __builtin_bit_cast(unsigned, v) & 0xffffu;
    }
};
template <> struct RsPack<3> {              // FP16
    __device__ static unsigned exact2(float a, float b) {
        return __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(a, b));
    }
    __device__ static float lo(unsigned w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu)); }
    __device__ static float hi(unsigned w) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16)); }
    // stochastic rounding with gfx950's v_cvt_sr_f16_f32: for a normal FP16
    // result it adds the top 13 bits of the random operand below the FP16
    // mantissa and truncates (rounds up with probability equal to the
    // discarded fraction; also unbiased in the subnormal range).  a takes
    // bits 31..19 of r, b bits 18..6.
    __device__ static unsigned sr2(float a, float b, unsigned r) {
        vl_f16x2 v = {};
        v = __builtin_amdgcn_cvt_sr_f16_f32(v, a, r, false);
        v = __builtin_amdgcn_cvt_sr_f16_f32(v, b, r << 13, true);
        return __builtin_bit_cast(unsigned, v);
    }
    __device__ static unsigned sr1(float a, unsigned r) {
        vl_f16x2 v = {};
        v = __builtin_amdgcn_cvt_sr_f16_f32(v, a, r, false);
        return __builtin_bit_cast(unsigned, v) & 0xffffu;
    }
};

typedef short vl_s16x4b __attribute__((ext_vector_type(4)));

// gfx950 LDS transpose read (ds_read_b64_tr_b16), lane layout probed on
// MI355X: in each 16-lane group, lane j = 4 a + b receives element b of the
// 8-byte chunks addressed by lanes a, a + 4, a + 8 and a + 12 of its group.
// With lane c = a + 4 t pointing at (slot 4 g + t, v-quad a) of brick lines
// (g = lane group), lane j holds slots 4 g .. 4 g + 3 of v-slot 4 a + b:
// a matrix-core A operand with lines = v-slots and K = slots, read straight
// from LDS (no identity product and repacking).
__device__ __forceinline__ vl_u32x2 lds_tr16(const unsigned *p) {
    const vl_s16x4b r = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
        (__attribute__((address_space(3))) vl_s16x4b *)(__attribute__((address_space(3))) void *)p);
    return __builtin_bit_cast(vl_u32x2, r);
}

// Marsaglia's xorshift32 (13, 17, 5): the random words of stochastic
// rounding within one lane's share of a tile.  Each lane starts from a hash of
// (stored element, lane), so the words depend on where a value is stored and
// on the seed, not on the launch shape; full-rate shifts and xors only.
__device__ __forceinline__ unsigned xorshift32(unsigned x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Vector-memory loads a wave issues for the next tile's prefetch, at least:
// per own row its 16 bytes, the two v-face words and (n = 9) the tail word,
// then the neighbour row and (n = 9) its tail.  These are issued whatever the
// lanes (masked lanes with an out-of-range offset); pencils at the u boundary
// issue the neighbour loads once more (ghost elements), so a wait for all but
// this many loads may wait a little longer there, never shorter.
template <int N>
constexpr int rs_prefetch_loads() {
    return (N == 9 ? 4 : 3) * ((N * N + 63) / 64) + (N == 9 ? 2 : 1);
}

// f^n staging into LDS: buffer loads with the LDS flag (lane l's 16 or 4
// bytes land at lds + 16 l or 4 l), written as inline asm.  The same loads
// through the builtin make the compiler wait for every outstanding load at
// the next LDS access of any array (it cannot tell the staging buffer from
// the others), which would wait for the next tile's prefetch at once; here
// the kernel waits for them itself (rs_wait_vm) before it reads the staging
// buffer.  The asm sets M0 (the LDS base of the wave) and reads no VGPR but
// the offset, so no register of it is written while the load is in flight.
__device__ __forceinline__ vl_u32x4 rsrc_words(const void *p, int bytes) {
    const unsigned long long a = (unsigned long long)p;
    return vl_u32x4{(unsigned)a, (unsigned)(a >> 32) & 0xffffu, (unsigned)bytes, 0x00020000u};   // as make_rsrc
}
__device__ __forceinline__ unsigned lds_addr(const void *p) {
    return (unsigned)(unsigned long long)(const __attribute__((address_space(3))) void *)p;
}
__device__ __forceinline__ void lds_dma16_nt(const vl_u32x4 &r, int voff, unsigned lds) {
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tbuffer_load_dwordx4 %0, %1, 0 offen nt lds" ::"v"(voff), "s"(r),
                 "s"(lds)
                 : "memory", "m0");
}
__device__ __forceinline__ void lds_dma4_nt(const vl_u32x4 &r, int voff, unsigned lds) {
    asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tbuffer_load_dword %0, %1, 0 offen nt lds" ::"v"(voff), "s"(r),
                 "s"(lds)
                 : "memory", "m0");
}

// s_waitcnt through the builtin, so that the compiler's own wait insertion
// knows which loads have landed: vector-memory count K, no lgkm / export wait
// (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int K>
__device__ __forceinline__ void rs_wait_vm() {
    static_assert(K >= 0 && K < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((K & 15) | ((K >> 4) << 14) | (7 << 4) | (15 << 8));
}
// Every outstanding load landed.  Issued once before a tile loop: loads of
// per-lane constants would otherwise be waited for at their first use inside
// the loop, where the count also covers the next tile's prefetch.
__device__ __forceinline__ void rs_wait_all() { __builtin_amdgcn_s_waitcnt(0); }

template <int N>
struct RsLoads {
    vl_u32x4 row[RsShape<N>::NRI];   // own rows: nodes 0..7 of line (ix, iu)
    unsigned tail[RsShape<N>::NRI];  // own node 8 (n = 9)
    // v-neighbour faces of the row (kept apart until the brick is filled, so
    // that no arithmetic waits for loads of a tile that is still prefetching)
    unsigned vnl[RsShape<N>::NRI], vnr[RsShape<N>::NRI];
    vl_u32x4 nrow;                   // one neighbour row (x or u face plane)
    unsigned ntail;
};

template <int N, int MODE, int S, bool ADT = false>
__global__ __launch_bounds__(kBlock, VL_RS_WAVES) void k_stage_rs(
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
    using SH = RsShape<N>;
    using F = MxFmt<S>;
    using PK = RsPack<S>;
    constexpr int TV = SH::TV, EP = SH::EP, ROWS = SH::ROWS, NRI = SH::NRI, RSD = SH::RSD, PSD = SH::PSD,
                  ESD = SH::ESD, XR = SH::XR, XE = SH::XE, GMAX = SH::GMAX, NNB = SH::NNB;

    __shared__ __attribute__((aligned(16))) unsigned int brick[TV * ESD];
    __shared__ __attribute__((aligned(16))) float xa[TV * XE];
    __shared__ float red[TV * ROWS];   // per-row moment sums, wave-private rows
    // per-lane sink of the 16-byte stores of lanes that hold no output node
    // (shared by the waves: never read)
    __shared__ __attribute__((aligned(16))) float sink[64 * 4];
    // f^n of this wave's element, staged by direct-to-LDS loads: rows (16 B
    // each) then the tail dwords
    constexpr int FNR = ROWS * 4, FNW = ((ROWS * 4 + (N == 9 ? 64 : 0)) + 3) & ~3;
    __shared__ __attribute__((aligned(16))) unsigned fnst[MODE == kModeStage ? TV * FNW : 4];

    const int tid = threadIdx.x;
    const int ln = tid & 63;
    const int el = __builtin_amdgcn_readfirstlane(tid >> 6);   // this wave's element (wave-uniform)
    const int lj = ln & 15;             // MFMA column / row within a 16 x 16 block
    const int lg = ln >> 4;             // lane group: K slots / output rows 4 lg .. 4 lg + 3
    const Pencil pc = decode_pencil(s, xcd_remap(blockIdx.x, gridDim.x));
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
    // operator fragments (host-built): lane holds op[i = lj][slot 4 lg + t],
    // hi and lo 16-bit parts
    vl_u32x2 ophi, oplo;
    {
        const vl_u32x4 fr = *reinterpret_cast<const vl_u32x4 *>(s.mxfrag + ((S == 2 ? 0 : 64) + ln) * 4);
        ophi = vl_u32x2{fr.x, fr.y};
        oplo = vl_u32x2{fr.z, fr.w};
    }
    // transposing reads of the x- and u-direction operands (lds_tr16): this
    // lane points at slot 4 lg + t (clamped to 11: slots 12..15 meet zero
    // operator columns) and v-quad a (quad 3 repeats quad 2: lines 12..15 are
    // padding); the block adds su RSD (x pass) or sx PSD (u pass)
    const int tr_k = min(4 * lg + (lj >> 2), 11), tr_q = 2 * min(lj & 3, 2);
    const int tr_x = tr_k * PSD + tr_q;     // x-slot tr_k of u-slot su
    const int tr_u = tr_k * RSD + tr_q;     // u-slot tr_k of x-slot sx
    // background images (accumulator inputs), v-speed terms: lane column
    // iu = lj, rows iv = 4 lg + t of this wave's element
    // u background as average + half difference: su * G(sign su) = su Gavg + |su| Gh
    vl_f32x4 gavg = {0.f, 0.f, 0.f, 0.f}, gh = gavg, hv = gavg, pt = gavg;
    float pl = 0.f, sv = 0.f, sxl = 0.f;
    unsigned vsel = kSelKeep;
    unsigned xposm = 0u;               // bit su: u(eu, su) >= 0 (x-line upwinding), wave-uniform
    const float jua = (float)(s.ju * s.acc);
    const bool sround = s.rounding != 0;
    const unsigned seedh = sround ? hash32(*s.seed * 0x9e3779b9u + 0x7f4a7c15u) : 0u;
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
                    const float gp = s.bg32[o], gm = s.bg32[tbl + o];
                    gavg[t] = 0.5f * (gp + gm);
                    gh[t] = 0.5f * (gp - gm);
                    hv[t] = s.bg32[2 * tbl + o];
                }
            }
        }
        const int jc = min(lj, N - 1);
        pl = (float)(s.ju * s.acc * (s.bfield * vnodes[ev * N + jc]));
        const double av = s.acc * (-s.bfield * unodes[eu * N + jc]);   // main/fluxes.py:143
        sv = (float)(s.jv * av);
        sxl = (float)(s.jx * unodes[eu * N + jc]);                      // main/fluxes.py:110
        if (lg == 2) vsel = av >= 0.0 ? kSelPos : kSelNeg;
#pragma unroll
        for (int su = 0; su < N; ++su) xposm |= (unodes[eu * N + su] >= 0.0 ? 1u : 0u) << su;
        // never-written brick slots stay zero (finite) for the whole kernel
        for (int i = ln; i < ESD; i += 64) bk[i] = 0u;
    }

    // ---- global loads of one tile ------------------------------------------
    // Raw buffer loads over whole state arrays (fewer than 2^31 bytes, checked
    // by the host): uniform resource + 32-bit byte offset per lane; offsets
    // of lanes with nothing to load are kOob and read 0 (no branches).
    const int sbytes = (int)((int64_t)(s.nx + 2) * sxe * 2);
    const __amdgpu_buffer_rsrc_t rin = make_rsrc(fin, sbytes);
    const __amdgpu_buffer_rsrc_t rgh = make_rsrc(fgh, sbytes);
    const vl_u32x4 rfnw = rsrc_words(fn, sbytes);
    const __amdgpu_buffer_rsrc_t rout = make_rsrc(fout, sbytes);
    auto ld16 = [&](__amdgpu_buffer_rsrc_t rr, int off) {
        return (unsigned)__builtin_amdgcn_raw_buffer_load_b16(rr, off, 0, 0);
    };
    // Single 16-bit values of the prefetch are read as the aligned dword that
    // holds them (offset & ~3) and picked at fill time; a 16-bit load would be
    // zero-extended at once, stalling on a load meant to stay in flight.
    auto ld32 = [&](__amdgpu_buffer_rsrc_t rr, int off) {
        return (unsigned)__builtin_amdgcn_raw_buffer_load_b32(rr, off, 0, 0);
    };
    // neighbour rows: lane -> (side, j); sides 0/1 x-left/right, 2/3 u-left/right
    const int nside = ln / N, nj = ln - (ln / N) * N;
    const bool nlane = ln < NNB;
    const int nrr = nside == 0 ? (N - 1) * N + nj : nside == 1 ? nj : nside == 2 ? nj * N + (N - 1) : nj * N;
    // neighbour rows of the u ghost elements (first / last u element) come
    // from fgh; lanes whose neighbour is one of them (wave-uniform per side)
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
            d.tail[k] = N == 9 ? ld32(rin, ok ? (b + (8 * ROWS + r) * 2) & ~3 : kOob) : 0u;
            d.vnl[k] = ld32(rL, ok ? (bl + (N == 9 ? (8 * ROWS + r) * 2 : r * 16 + 12)) & ~3 : kOob);
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
        // halves ha of a and hb of b -> one word (a's in the low half)
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
                // the 16-bit tail values sit in half (r & 1) of their dword (8 n^2 is even)
                const unsigned w4 = N == 9 ? halves(d.tail[k], r & 1, R.x, 0)    // slot 8 = node n-1,
                                           : halves(R.w, 1, R.x, 0);             // slot 9 = node 0
                const unsigned w5 = halves(d.vnl[k], N == 9 ? (r & 1) : 1, d.vnr[k], 0);   // slots 10, 11: v faces
                put(ix, iu, R, w4, w5, true);
                if (iu == 0) put(ix, 9, R, w4, w5, true);       // u-slot 9: copy of node 0
                if (ix == 0) put(9, iu, R, w4, w5, true);       // x-slot 9: copy of node 0
                if (N == 8 && iu == N - 1) put(ix, 8, R, w4, w5, true);   // slot 8: copy of node n-1
                if (N == 8 && ix == N - 1) put(8, iu, R, w4, w5, true);
            }
        }
        if (nlane) {
            const int xs = nside == 0 ? 10 : nside == 1 ? 11 : nj;
            const int us = nside == 2 ? 10 : nside == 3 ? 11 : nj;
            const vl_u32x4 R = d.nrow;
            put(xs, us, R, N == 9 ? halves(d.ntail, nrr & 1, R.x, 0) : halves(R.w, 1, R.x, 0), 0u, false);
        }
    };

    const float c0 = (float)s.c0, c1 = (float)s.c1, c2 = (float)s.c2dt;
    const int tiles_per_ex = s.nu * s.nchunk;
    const bool partials = s.write_partials != 0;
    const vl_f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};

    // Per tile: fill the brick from `cur`, prefetch the next tile into `nxt`,
    // compute, store.
    rs_wait_all();

    // The brick of a tile is filled at the end of the previous tile, after
    // its epilogue: the prefetched loads are then older than that epilogue's
    // stores, which stay in flight (and the compiler's waits see one path).
    RsLoads<N> nxt;
    {
        RsLoads<N> first;
        load(pc.ex_begin, first);
        rs_wait_all();
        fill(first);
    }
    for (int ex = pc.ex_begin; ex <= pc.ex_end; ++ex) {
        const bool has_next = ex < pc.ex_end;
        if (MODE == kModeStage && active) {
            // f^n rows of this element straight into LDS: lane l -> row l,
            // then rows 64.. (lanes < ROWS - 64); n = 9: the tail dwords
            const int bfn = (int)(ebase(ex, eu, ev) * 2);
            const unsigned st0 = lds_addr(fnst + el * FNW);
            lds_dma16_nt(rfnw, bfn + ln * 16, st0);
            if (ln < ROWS - 64) lds_dma16_nt(rfnw, bfn + (64 + ln) * 16, st0 + 256 * 4);
            if (N == 9) lds_dma4_nt(rfnw, ln < (ROWS + 1) / 2 ? bfn + 16 * ROWS + 4 * ln : kOob, st0 + FNR * 4);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (has_next) load(ex + 1, nxt);
        const int64_t b = ebase(ex, eu, ev);
        wave_sync_lds();
        if (active) {
            // Branch-free passes: lanes beyond the 12 slots read a clamped
            // (finite) brick row -- it only meets zero operator columns or
            // feeds output lines nobody keeps -- and lanes without an output
            // node store to their sink.
            const int rduv = min(lj, 11) * RSD + 2 * min(lg, 2);   // u-slot lj, v-slots 4 lg ..
            const bool wr = lj < N && lg <= GMAX;
            float *const snk = sink + ln * 4;
            const unsigned selp = lg == 2 ? kSelPos : kSelKeep, seln = lg == 2 ? kSelNeg : kSelKeep;
            // ---- x: X[ix][iu][v-slot] = A_x g (the x speed is applied in the u/v pass)
#pragma unroll
            for (int su = 0; su < N; ++su) {
                vl_u32x2 p = lds_tr16(bk + tr_x + su * RSD);   // lane: v-slot lj, x-slots 4 lg + t
                p.y = __builtin_amdgcn_perm(p.x, p.y, ((xposm >> su) & 1u) ? selp : seln);
                vl_f32x4 d = F::mfma(p, ophi, zero4);
                d = F::mfma(p, oplo, d);                      // lane: x-node lj, v-slots 4 lg + t
                float *dst = wr ? xw + (lj * N + su) * XR + 4 * lg : snk;
                *reinterpret_cast<vl_f32x4 *>(dst) = d;
                task_fence();
            }
            // E of this tile's x nodes, as the u-speed offset Ju acc E (uniform)
            float qs[N];
#pragma unroll
            for (int sx = 0; sx < N; ++sx)
                qs[sx] = __builtin_bit_cast(
                    float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, jua * (float)E[ex * N + sx])));
            wave_sync_lds();
            // ---- u and v: L = Jx a_x X + Ju a_u (A_u g + G+-) + Jv a_v (A_v g + Hv)
#pragma unroll
            for (int sx = 0; sx < N; ++sx) {
                const float q = qs[sx];
                const vl_u32x2 r = *reinterpret_cast<const vl_u32x2 *>(bk + sx * PSD + rduv);
                vl_u32x2 bvv = r;
                bvv.y = __builtin_amdgcn_perm(r.x, r.y, vsel);
                vl_f32x4 dv = F::mfma(ophi, bvv, hv);
                dv = F::mfma(oplo, bvv, dv);                  // lane: u-slot lj, v-nodes 4 lg + t
                vl_u32x2 p = lds_tr16(bk + tr_u + sx * PSD);   // lane: v-slot lj, u-slots 4 lg + t
                p.y = __builtin_amdgcn_perm(p.x, p.y, (q + pl >= 0.f) ? selp : seln);   // Ju a_u of the lane's line
                vl_f32x4 du = F::mfma(p, ophi, gavg);
                du = F::mfma(p, oplo, du);                    // lane: u-node lj, v-slots 4 lg + t
                float *xp = wr ? xw + (sx * N + lj) * XR + 4 * lg : snk;
                const vl_f32x4 xv = *reinterpret_cast<const vl_f32x4 *>(xp);
                vl_f32x4 L;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float su = q + pt[t];
                    L[t] = __builtin_fmaf(sv, dv[t],
                                          __builtin_fmaf(su, du[t], __builtin_fmaf(__builtin_fabsf(su), gh[t], sxl * xv[t])));
                }
                *reinterpret_cast<vl_f32x4 *>(xp) = L;
                task_fence();
            }
            wave_sync_lds();
            // ---- epilogue: whole rows ----------------------------------------
            const int64_t bo = ebase(ex, eu, ev);
            // stochastic rounding: this lane's word stream for this element
            unsigned rstate = sround ? (hash32(((unsigned)((ex * (s.nu + 2) + eu) * vrow + ev) << 6 | (unsigned)ln) ^
                                              seedh) |
                                       1u)
                                     : 0u;
            if (MODE == kModeStage) {
                // the staged f^n is in LDS: its loads were issued before the next tile's
                // prefetch, which may stay in flight
                if (has_next)
                    rs_wait_vm<rs_prefetch_loads<N>()>();
                else
                    rs_wait_vm<0>();
            }
            // Branch-free passes: lanes past the last row recompute it and their
            // stores go to an out-of-range offset (dropped), so every pass issues
            // the same stores (the next tile's wait for its prefetch counts them).
#pragma unroll 1
            for (int k = 0; k < NRI; ++k) {
                const int r0 = ln + 64 * k;
                const bool rok = r0 < ROWS;
                const int r = rok ? r0 : ROWS - 1;
                {
                    vl_u32x4 fnr = {0u, 0u, 0u, 0u};
                    unsigned fnt = 0u;
                    if (MODE == kModeStage) {
                        const unsigned *st0 = fnst + el * FNW;
                        fnr = *reinterpret_cast<const vl_u32x4 *>(st0 + r * 4);
                        fnt = N == 9 ? st0[FNR + (r >> 1)] : 0u;
                    }
                    const int iu = r - (r / N) * N;
                    const int ix = r / N;
                    const float *xp = xw + r * XR;
                    const unsigned *br = bk + ix * PSD + iu * RSD;    // stage input row, slots 0..8
                    const unsigned nw[4] = {fnr.x, fnr.y, fnr.z, fnr.w};
                    ConstOpTables &tab = *(ConstOpTables *)tabp;
                    // pairs of values: combine, round, moment of the stored pair
                    auto pair = [&](float l0, float l1, unsigned fiw, unsigned fnw, float &m, int j) {
                        float o0, o1;
                        if (MODE == kModeStage1) {
                            o0 = PK::lo(fiw) + c2 * l0;
                            o1 = PK::hi(fiw) + c2 * l1;
                        } else {
                            o0 = (c0 * PK::lo(fnw) + c1 * PK::lo(fiw)) + c2 * l0;
                            o1 = (c0 * PK::hi(fnw) + c1 * PK::hi(fiw)) + c2 * l1;
                        }
                        unsigned wv2;
                        if (sround) {
                            rstate = xorshift32(rstate);
                            wv2 = PK::sr2(o0, o1, rstate);
                        } else {
                            wv2 = F::enc2(o0, o1);
                        }
                        m += tab.wv[j] * PK::lo(wv2);
                        m += tab.wv[j + 1] * PK::hi(wv2);
                        return wv2;
                    };
                    float m = 0.f;
                    vl_u32x4 w;
                    {
                        const vl_f32x4 l0 = *reinterpret_cast<const vl_f32x4 *>(xp);
                        const vl_u32x2 f01 = *reinterpret_cast<const vl_u32x2 *>(br);
                        w.x = pair(l0[0], l0[1], f01.x, nw[0], m, 0);
                        w.y = pair(l0[2], l0[3], f01.y, nw[1], m, 2);
                    }
                    {
                        const vl_f32x4 l1 = *reinterpret_cast<const vl_f32x4 *>(xp + 4);
                        const vl_u32x2 f23 = *reinterpret_cast<const vl_u32x2 *>(br + 2);
                        w.z = pair(l1[0], l1[1], f23.x, nw[2], m, 4);
                        w.w = pair(l1[2], l1[3], f23.y, nw[3], m, 6);
                    }
                    buf_st16_nt(rout, rok ? (int)(bo * 2) + r * 16 : kOob, w);
                    if (N == 9) {
                        const float fi8 = PK::lo(br[4]);
                        const float fn8 = (r & 1) ? PK::hi(fnt) : PK::lo(fnt);
                        const float o8 = MODE == kModeStage1 ? fi8 + c2 * xp[8] : (c0 * fn8 + c1 * fi8) + c2 * xp[8];
                        unsigned t8;
                        if (sround) {
                            rstate = xorshift32(rstate);
                            t8 = PK::sr1(o8, rstate);
                        } else {
                            t8 = F::enc(o8);
                        }
                        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)t8, rout,
                                                              rok ? (int)(bo * 2) + (8 * ROWS + r) * 2 : kOob,
                                                              0, 0);
                        m += tab.wv[8] * PK::lo(t8);
                    }
                    // v-moment of the stored row (its u weight is applied in the sum below)
                    if (rok) red[el * ROWS + r] = m;
                }
            }
        }
        // moment partial of this wave's element (one slot per tile element,
        // zero for the unused elements of a partial chunk); no barrier
        if (partials) {
            wave_sync_lds();
            if (ln < N) {
                double a = 0.0;
                if (active) {
                    ConstOpTables &tab = *(ConstOpTables *)tabp;
#pragma unroll
                    for (int iu = 0; iu < N; ++iu) a += (double)(tab.wu[iu] * red[el * ROWS + ln * N + iu]);
                }
                const int64_t slot =
                    (((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + pc.chunk) * TV + el) * N + ln;
                s.partials[slot] = make_double2(a * s.inv_scale, 0.0);
            }
        }
        if (has_next) fill(nxt);
    }
}

// Moment partials of a row-split state (first stage of a loop), in the
// partial layout of k_stage_rs (one slot per tile element): one workgroup per
// tile (ex, eu, 4 v-elements), one wave per element.
template <int N, int S>
__global__ __launch_bounds__(kBlock) void k_moment_rs(const uint16_t *__restrict__ f, StageArgs s,
                                                      const OpTables *__restrict__ tabp) {
    using SH = RsShape<N>;
    using PK = RsPack<S>;
    constexpr int TV = SH::TV, EP = SH::EP, ROWS = SH::ROWS, NRI = SH::NRI;
    __shared__ float red[TV * ROWS];   // wave-private rows
    const int tid = threadIdx.x, ln = tid & 63, el = __builtin_amdgcn_readfirstlane(tid >> 6);
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
    wave_sync_lds();
    if (ln < N) {
        double a = 0.0;
        if (el < nvalid)
            for (int iu = 0; iu < N; ++iu) a += (double)red[el * ROWS + ln * N + iu];
        const int tiles_per_ex = s.nu * s.nchunk;
        const int64_t slot = (((int64_t)(ex - 1) * tiles_per_ex + (eu - 1) * s.nchunk + chunk) * TV + el) * N + ln;
        s.partials[slot] = make_double2(a * s.inv_scale, 0.0);
    }
}

}  // namespace vl
