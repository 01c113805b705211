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
// Brick.  Each wave owns one element of the tile and keeps it in LDS as a
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
// Work decomposition.  A workgroup of 4 waves owns a pencil (4 consecutive
// elements of the velocity plane in (eu, ev) order, x-segment) and walks it
// along x (decode_pencil_rs).  The waves are
// independent (no barriers): each writes the moment partial of its own
// element, one slot per tile element.
// The global loads of the next tile are in flight while the current one is
// computed.  Stages 2/3 read f^n of the current element with direct-to-LDS
// buffer loads issued at the start of the tile (no registers held across
// the passes); the epilogue waits for them and reads them from LDS.
//
// Epilogue.  The element's 729 values are taken as units of four values
// that are consecutive in the storage order: 162 half rows, then (n = 9) the
// tail block in 21 groups of four rows.  A wave does one unit per lane and
// pass, 3 passes (n = 9; 2 for n = 8): the stage combination, the rounding,
// an 8-byte store and the w_v-weighted moment of the stored values.  The
// per-x-node moment is then summed over the unit results by four lanes per
// node (w_u applied there) and a quad reduction.
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

// Work decomposition of the row-split kernels.  A tile is (ex, group g of TV
// consecutive elements of the velocity plane in (eu, ev) order), one wave per
// element, so that no wave idles when nv is not a multiple of TV (the groups
// of a plane row may reach into the next row; only the plane's last group
// can be partial).  s.nchunk holds the number of groups per x element; a
// pencil is (g, x-segment).
__device__ __forceinline__ Pencil decode_pencil_rs(const StageArgs &s, int p) {
    Pencil c;
    c.chunk = p % s.nchunk;
    const int seg = p / s.nchunk;
    c.eu = 0;                               // per wave: rs_wave_element
    c.ex_begin = 1 + seg * s.xseg;
    c.ex_end = min(s.nx, c.ex_begin + s.xseg - 1);
    return c;
}

// Element (eu, ev) of wave el in group g; false past the plane's last
// element (then (eu, ev) is the last element, so that addresses stay valid).
template <int TV>
__device__ __forceinline__ bool rs_wave_element(const StageArgs &s, int g, int el, int &eu, int &ev) {
    const int plane = s.nu * s.nv;
    const int q = g * TV + el;
    const int qc = min(q, plane - 1);
    eu = 1 + qc / s.nv;
    ev = 1 + qc - (qc / s.nv) * s.nv;
    return q < plane;
}

template <int N>
struct RsShape {
    static_assert(N == 8 || N == 9, "row-split stage kernel: 8 or 9 nodes per axis");
    static constexpr int TV = 4;                  // elements per tile, one wave each
    static constexpr int EP = elem_pitch(N);
    static constexpr int ROWS = N * N;            // 8-value rows per element
    static constexpr int NRI = (ROWS + 63) / 64;  // rows per lane
    static constexpr int RSD = 6;                 // brick row stride (dwords, 12 slots)
    static constexpr int PSD = 74;                // brick x-slot stride (dwords): 12 rows + 2
    static constexpr int ESD = 12 * PSD;          // brick of one element (dwords)
    static constexpr int GMAX = N == 9 ? 2 : 1;   // last lane group holding an output node
    static constexpr int NNB = 4 * N;             // neighbour rows per tile element
    // Epilogue units: four values that are consecutive in the element's
    // storage order -- half rows (row r, nodes 4q .. 4q+3: unit 2r + q),
    // then (n = 9) the tail block in groups of four rows (unit 2 n^2 + m:
    // node 8 of rows 4m .. 4m+3).  Unit u holds storage values 4u .. 4u+3.
    static constexpr int UH = 2 * ROWS;                  // half-row units
    static constexpr int UT = N == 9 ? (ROWS + 3) / 4 : 0;   // tail units
    static constexpr int NUNIT = UH + UT;                // 183 (n = 9), 128 (n = 8)
    static constexpr int NPASS = (NUNIT + 63) / 64;      // epilogue passes of one unit per lane
    static constexpr int XR = 12;                        // floats per L row (v-slots 0..11)
    static constexpr int XE = ROWS * XR;                 // floats of L per element
    static constexpr int TAIL = 8 * ROWS;                // storage index of the tail block
    static constexpr int TF = N == 9 ? 4 * UT : 4;       // tail values of the stage input (16-bit)
    // per-wave scratch (floats): 16 bytes per lane as the sink of the passes'
    // lanes that hold no output node, then the epilogue's moment sums:
    // half-row units at [0, UH), the weighted tail values of row r at TWO + r
    static constexpr int RW = 256;
    static constexpr int TWO = (UH + 3) / 4 * 4;
    static constexpr int RDUMMY = RW - 8;                // discarded moment writes
    static_assert(N == 8 || TWO + 4 * UT <= RDUMMY, "moment scratch layout");
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
        // both halves are written: the initial value is irrelevant
        vl_bf16x2 v = __builtin_bit_cast(vl_bf16x2, r);
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
        vl_f16x2 v = __builtin_bit_cast(vl_f16x2, r);   // both halves are written
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

// ---- moment partials of the stored values (stage epilogue and k_moment_rs,
// which must give the same bits for the same state) -------------------------
// Weights in a small LDS table (read where they are used, so that no
// registers hold them across the passes): [0, 8) w_v of nodes 0..7 (a lane
// takes the four of its unit's quad q = ln & 1, the same in every pass),
// [8, 16) the pairs (w_u[j], w_u[j + 4]) of the reduction lane 4 ix + j.
// Written by the first 16 threads of the workgroup, then one barrier.
constexpr int kRsWeights = 16;
__device__ __forceinline__ void rs_weights_fill(float *wt, ConstOpTables &tab, int tid) {
    if (tid < 8) wt[tid] = tab.wv[tid];
    else if (tid < 16) wt[tid] = tab.wu[(tid - 8) / 2 + 4 * ((tid - 8) & 1)];
}

// Moment of unit u0 (pass p, stored words w): w_v-weighted values; a
// half-row unit keeps their sum at rw[u0], a tail unit its four rows' values
// at rw[TWO + 4 m ..] (units past the last write to a dummy slot).
template <int N, int S>
__device__ __forceinline__ void rs_unit_moment(float *rw, int p, int u0, vl_u32x2 w, const float *wtab,
                                               ConstOpTables &tab) {
    using SH = RsShape<N>;
    using PK = RsPack<S>;
    const bool tail_pass = N == 9 && 64 * p + 63 >= SH::UH;
    vl_f32x4 wt = *reinterpret_cast<const vl_f32x4 *>(wtab + 4 * (u0 & 1));
    if (tail_pass) {
        const float wv8 = tab.wv[N - 1];
#pragma unroll
        for (int t = 0; t < 4; ++t) wt[t] = u0 >= SH::UH ? wv8 : wt[t];
    }
    vl_f32x4 mv;
    mv[0] = wt[0] * PK::lo(w.x);
    mv[1] = wt[1] * PK::hi(w.x);
    mv[2] = wt[2] * PK::lo(w.y);
    mv[3] = wt[3] * PK::hi(w.y);
    if (!tail_pass || u0 < SH::UH) rw[u0] = (mv[0] + mv[1]) + (mv[2] + mv[3]);
    if (tail_pass && u0 >= SH::UH)
        *reinterpret_cast<vl_f32x4 *>(rw + (u0 < SH::NUNIT ? SH::TWO + 4 * (u0 - SH::UH) : SH::RDUMMY)) = mv;
}

// Moment of x-node ix = ln >> 2 of the element from the unit moments (after
// a wave sync): sum over rows (ix, iu) of w_u (half-row sums + tail value),
// complete on lanes 4 ix (lanes with ix >= n return garbage).
template <int N>
__device__ __forceinline__ float rs_element_moment(const float *rw, int ln, const float *wtab, ConstOpTables &tab) {
    using SH = RsShape<N>;
    const int rix = min(ln >> 2, N - 1), rj = ln & 3;
    const int ra = rix * N + rj, rb = ra + 4;
    const vl_f32x2 ha = *reinterpret_cast<const vl_f32x2 *>(rw + 2 * ra);
    const vl_f32x2 hb = *reinterpret_cast<const vl_f32x2 *>(rw + 2 * rb);
    float ta = 0.f, tb = 0.f;
    if (N == 9) {
        ta = rw[SH::TWO + ra];
        tb = rw[SH::TWO + rb];
    }
    const vl_f32x2 wu = *reinterpret_cast<const vl_f32x2 *>(wtab + 8 + 2 * rj);
    float a = wu[0] * ((ha[0] + ha[1]) + ta) + wu[1] * ((hb[0] + hb[1]) + tb);
    if (N == 9) {   // row (ix, 8) on lane j = 0
        const int rc = rix * N + 8;
        const vl_f32x2 hc = *reinterpret_cast<const vl_f32x2 *>(rw + 2 * rc);
        const float c8 = tab.wu[N - 1] * ((hc[0] + hc[1]) + rw[SH::TWO + rc]);
        a += rj == 0 ? c8 : 0.f;
    }
    // sum over the four lanes of the x-node (quad_perm 1,0,3,2 then 2,3,0,1)
    a += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, a), 0xB1, 0xf, 0xf, false));
    a += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, a), 0x4E, 0xf, 0xf, false));
    return a;
}

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
                  ESD = SH::ESD, GMAX = SH::GMAX, NNB = SH::NNB, UH = SH::UH, NUNIT = SH::NUNIT,
                  NPASS = SH::NPASS, XR = SH::XR, XE = SH::XE, TAIL = SH::TAIL, TF = SH::TF, RW = SH::RW, TWO = SH::TWO,
                  RDUMMY = SH::RDUMMY;

    // bricks of the tile's elements, then (n = 9) node 8 of every row of the
    // stage input, rows in order, zero past the last row: the tail units'
    // stage-input values
    constexpr int TFD = TF / 2;
    __shared__ __attribute__((aligned(16))) unsigned int brick[TV * ESD + TV * TFD];
    // L of this wave's element, FP32 rows (ix, iu) of v-slots 0..11: the x pass
    // writes A_x g, the u/v pass replaces it by L (node 8 at slot 8)
    __shared__ __attribute__((aligned(16))) float xa[TV * XE];
    // per-wave scratch: sink of the passes' discarded lanes, then moment sums
    __shared__ __attribute__((aligned(16))) float rwk[TV * RW];
    // f^n of this wave's element, staged by direct-to-LDS loads: rows (16 B
    // each) then the tail dwords
    constexpr int FNR = ROWS * 4, FNW = ((ROWS * 4 + (N == 9 ? 64 : 0)) + 3) & ~3;
    __shared__ __attribute__((aligned(16))) unsigned fnst[MODE == kModeStage ? TV * FNW : 4];

    const int tid = threadIdx.x;
    const int ln = tid & 63;
    const int el = __builtin_amdgcn_readfirstlane(tid >> 6);   // this wave's element (wave-uniform)
    const int lj = ln & 15;             // MFMA column / row within a 16 x 16 block
    const int lg = ln >> 4;             // lane group: K slots / output rows 4 lg .. 4 lg + 3
    const Pencil pc = decode_pencil_rs(s, VL_STAGE_PENCIL(s));
    int eu, ev;
    const bool active = rs_wave_element<TV>(s, pc.chunk, el, eu, ev);   // wave-uniform
    const int vrow = s.nv + 2;
    const int64_t sxe = s.sxe;
    unsigned int *bk = brick + el * ESD;
    float *xw = xa + el * XE;
    uint16_t *tfw = reinterpret_cast<uint16_t *>(brick + TV * ESD + el * TFD);
    float *rw = rwk + el * RW;

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
        // never-written brick slots stay zero (finite) for the whole kernel,
        // and so do the tail inputs past the last row
        for (int i = ln; i < ESD; i += 64) bk[i] = 0u;
        if (N == 9 && ln < TF - ROWS) tfw[ROWS + ln] = 0;
    }
    __shared__ __attribute__((aligned(16))) float wtab[kRsWeights];
    rs_weights_fill(wtab, *(ConstOpTables *)tabp, tid);
    __syncthreads();

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
                if (N == 9) tfw[r] = (uint16_t)w4;               // node 8 for the tail units
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
    const int tiles_per_ex = s.nchunk;      // element groups per x element
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
            float *const snk = rw + ln * 4;
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
            // ---- epilogue: one unit of four stored values per lane and pass --
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
            const unsigned *fnw = fnst + el * FNW;
            const int obase = (int)(bo * 2);
            // the lane index as an opaque value: the per-lane addresses of the
            // epilogue are formed here, not hoisted out of the tile loop
            // (where they would hold registers across the passes)
            int lno = ln;
            asm volatile("" : "+v"(lno));
            lno &= 63;
#pragma unroll
            for (int p = 0; p < NPASS; ++p) {
                const int u0 = 64 * p + lno;
                // units past the last one recompute it; their stores are dropped
                const int u = min(u0, NUNIT - 1);
                // L: a half row is one 16-byte read at row r, slot 4q; a tail unit
                // takes slot 8 of rows 4m .. 4m+3 (rows past the last read a
                // clamped place; their values are not stored)
                vl_f32x4 l4;
                if (N == 9 && 64 * p + 63 >= UH) {
                    const bool tl = u >= UH;
                    const int ua = tl ? 4 * (u - UH) * XR + 8 : (u >> 1) * XR + 4 * (u & 1);
                    l4 = *reinterpret_cast<const vl_f32x4 *>(xw + ua);
                    const float *pb = xw + min(ua, XE - 1 - 3 * XR);
                    const float b1 = pb[XR], b2 = pb[2 * XR], b3 = pb[3 * XR];
                    l4[1] = tl ? b1 : l4[1];
                    l4[2] = tl ? b2 : l4[2];
                    l4[3] = tl ? b3 : l4[3];
                } else {
                    l4 = *reinterpret_cast<const vl_f32x4 *>(xw + (u >> 1) * XR + 4 * (u & 1));
                }
                // stage-input values: the brick row (ix, iu), slots 4q .. 4q+3,
                // of a half row; node 8 of rows 4m .. 4m+3 for a tail unit
                const int r = u >> 1, ix = r / N;
                const int fa = u < UH ? el * ESD + ix * (PSD - N * RSD) + r * RSD + 2 * (u & 1)
                                      : TV * ESD + el * TFD + 2 * (u - UH);
                const vl_u32x2 fi = *reinterpret_cast<const vl_u32x2 *>(brick + fa);
                vl_u32x2 fv = {0u, 0u};
                if (MODE == kModeStage)
                    fv = *reinterpret_cast<const vl_u32x2 *>(fnw + 2 * u);
                float o[4];
                {
                    const unsigned fiw[2] = {fi.x, fi.y}, fnv[2] = {fv.x, fv.y};
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (MODE == kModeStage1) {
                            o[2 * h] = PK::lo(fiw[h]) + c2 * l4[2 * h];
                            o[2 * h + 1] = PK::hi(fiw[h]) + c2 * l4[2 * h + 1];
                        } else {
                            o[2 * h] = (c0 * PK::lo(fnv[h]) + c1 * PK::lo(fiw[h])) + c2 * l4[2 * h];
                            o[2 * h + 1] = (c0 * PK::hi(fnv[h]) + c1 * PK::hi(fiw[h])) + c2 * l4[2 * h + 1];
                        }
                    }
                }
                vl_u32x2 w;
                if (sround) {
                    rstate = xorshift32(rstate);
                    w.x = PK::sr2(o[0], o[1], rstate);
                    rstate = xorshift32(rstate);
                    w.y = PK::sr2(o[2], o[3], rstate);
                } else {
                    w.x = F::enc2(o[0], o[1]);
                    w.y = F::enc2(o[2], o[3]);
                }
                // stores: whole units, and the one unit that holds the last
                // row's node 8 and padding past it (n = 9) as a single value
                constexpr bool has_partial = NUNIT * 4 > (N == 9 ? TAIL + ROWS : TAIL);
                const bool whole = u0 < (has_partial ? NUNIT - 1 : NUNIT);
                __builtin_amdgcn_raw_buffer_store_b64(w, rout, whole ? obase + 8 * u : kOob, 0, kCachePolicyNT);
                if (has_partial && 64 * p + 63 >= NUNIT - 1)
                    __builtin_amdgcn_raw_buffer_store_b16((uint16_t)w.x, rout,
                                                          u0 == NUNIT - 1 ? obase + 8 * u : kOob, 0, 0);
                if (partials) rs_unit_moment<N, S>(rw, p, u0, w, wtab, *(ConstOpTables *)tabp);
                // one pass at a time (hoisting the next pass's reads over this one
                // would hold more registers than the kernel has)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // moment partial of this wave's element (one slot per tile element,
        // zero for the unused elements of the plane's last group); no barrier
        if (partials) {
            wave_sync_lds();
            int lno = ln;
            asm volatile("" : "+v"(lno));
            lno &= 63;
            float a = rs_element_moment<N>(rw, lno, wtab, *(ConstOpTables *)tabp);
            const int rix = lno >> 2;
            if ((lno & 3) == 0 && rix < N) {
                const int64_t slot =
                    (((int64_t)(ex - 1) * tiles_per_ex + pc.chunk) * TV + el) * N + rix;
                s.partials[slot] = make_double2(active ? (double)a * s.inv_scale : 0.0, 0.0);
            }
        }
        if (has_next) fill(nxt);
    }
}

// Moment partials of a row-split state (first stage of a loop), in the
// partial layout of k_stage_rs (one slot per tile element) and with its
// arithmetic (rs_unit_moment, rs_element_moment), so that both give the same
// partials for the same stored values: one workgroup per tile (ex, group of
// 4 elements), one wave per element.
template <int N, int S>
__global__ __launch_bounds__(kBlock) void k_moment_rs(const uint16_t *__restrict__ f, StageArgs s,
                                                      const OpTables *__restrict__ tabp) {
    using SH = RsShape<N>;
    constexpr int TV = SH::TV, EP = SH::EP, RW = SH::RW, NPASS = SH::NPASS, NUNIT = SH::NUNIT;
    __shared__ __attribute__((aligned(16))) float rwk[TV * RW];   // wave-private moment scratch
    const int tid = threadIdx.x, ln = tid & 63, el = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int chunk = tile % s.nchunk;
    const int ex = 1 + tile / s.nchunk;
    int eu, ev;
    const bool active = rs_wave_element<TV>(s, chunk, el, eu, ev);
    __shared__ __attribute__((aligned(16))) float wtab[kRsWeights];
    rs_weights_fill(wtab, *(ConstOpTables *)tabp, tid);
    __syncthreads();
    float *rw = rwk + el * RW;
    if (active) {
        const int64_t b = (int64_t)ex * s.sxe + ((int64_t)eu * (s.nv + 2) + ev) * EP;
#pragma unroll
        for (int p = 0; p < NPASS; ++p) {
            const int u0 = 64 * p + ln;
            const int u = min(u0, NUNIT - 1);
            const vl_u32x2 w = *reinterpret_cast<const vl_u32x2 *>(f + b + 4 * u);
            rs_unit_moment<N, S>(rw, p, u0, w, wtab, *(ConstOpTables *)tabp);
        }
    }
    wave_sync_lds();
    const float a = rs_element_moment<N>(rw, ln, wtab, *(ConstOpTables *)tabp);
    const int rix = ln >> 2;
    if ((ln & 3) == 0 && rix < N) {
        const int64_t slot = (((int64_t)(ex - 1) * s.nchunk + chunk) * TV + el) * N + rix;
        s.partials[slot] = make_double2(active ? (double)a * s.inv_scale : 0.0, 0.0);
    }
}

}  // namespace vl
