// Ensemble of independent runs (different n, grids and physics) stepped
// together: every RK stage of every run is ONE launch per kernel kind, so
// that a step of the whole ensemble is 13 launches whatever the number of
// runs, and a HIP graph of K ensemble steps has 13 K nodes.
//
//   k_ens_swap      periodic x-ghost copy of every run's state
//   k_ens_reduce    moment partials -> rho, workgroup (run, x element)
//   k_ens_forward   forward nodal DFT + spectral solve, workgroup (run, mode group)
//   k_ens_inverse   inverse sums -> E (and the field energy after stage 3)
//   k_ens_stage     persistent stage kernel over the pencils of all runs
//
// The per-run bodies are the single-run kernels' (vl_field.h, vl_stage.h),
// so every run of an ensemble computes what vl_rk3_step computes for it, up
// to the compiler's choice of fused multiply-adds in each kernel (rounding
// level).  Reference semantics: main/timestep.py:141-187 (per run).
//
// Work list of the stage kernel.  An item is one pencil (v-chunk, eu,
// x-segment) of one run; the host orders the items (longest first, a run's
// pencils consecutive) and picks each run's x-segment length so that items
// carry similar work.  A workgroup takes items blockIdx.x, blockIdx.x + G, ...
// through an XCD-aware remap inside every round of G items, so the pencils
// that share u-halos run on one XCD at the same time.
#pragma once
#include "vl_field.h"
#include "vl_stage.h"

namespace vl {

constexpr int kEnsMinN = 6, kEnsMaxN = 9;   // node counts the ensemble kernels are built for

struct EnsRun {
    StageArgs s;                   // stage arguments (c0/c1/c2dt from cf, xseg from the schedule)
    double cf[3][3];               // (c0, c1, c2 dt) of stages 1..3 (main/timestep.py:37-38)
    const OpTables *tab;
    void *f, *f1, *f2;             // element-major state buffers
    double *E;
    const double *unodes, *vnodes, *wts, *kwave;
    const double2 *T, *phases;
    double *rho;
    double2 *coef, *enpart;
    double *mxpart;
    unsigned *ticket, *seed;
    FieldArgs fa;                  // from_rho = 1 (CFL fields set for adaptive runs)
    unsigned int *status;          // the run's workspace status word (non-finite field)
    int64_t plane;                 // values per x element of the element-major state
    int n, tpe, inv_blocks, swap_blocks;
    // adaptive time step (main/timestep.py:184,231-242): the run's workspace
    // clock (dt, time, next dt) and the stage factors c2 of c2 * dt
    double *clock;
    double c2[3];
    int adaptive;
};

// workgroup -> (run, local workgroup) of the field kernels
struct EnsBlock {
    int run, local;
};

template <typename W>
__global__ __launch_bounds__(kBlock) void k_ens_swap(const EnsRun *__restrict__ runs,
                                                     const EnsBlock *__restrict__ map) {
    const EnsBlock b = map[blockIdx.x];
    const EnsRun &r = runs[b.run];
    swap_ghosts_body<W>(static_cast<W *>(r.f), r.plane, r.fa.nx, b.local, r.swap_blocks);
}

__global__ __launch_bounds__(kBlock) void k_ens_reduce(const EnsRun *__restrict__ runs,
                                                       const EnsBlock *__restrict__ map) {
    __shared__ double sh[kEnsMaxN * kBlock];
    __shared__ double sl[kEnsMaxN * kBlock];
    const EnsBlock b = map[blockIdx.x];
    const EnsRun &r = runs[b.run];
    const double2 *part = r.s.partials;
    switch (r.n) {
        case 6: reduce_partials_body<6>(part, r.tpe, r.rho, r.seed, b.local, sh, sl); break;
        case 7: reduce_partials_body<7>(part, r.tpe, r.rho, r.seed, b.local, sh, sl); break;
        case 8: reduce_partials_body<8>(part, r.tpe, r.rho, r.seed, b.local, sh, sl); break;
        case 9: reduce_partials_body<9>(part, r.tpe, r.rho, r.seed, b.local, sh, sl); break;
        default: break;
    }
}

__global__ __launch_bounds__(kBlock) void k_ens_forward(const EnsRun *__restrict__ runs,
                                                        const EnsBlock *__restrict__ map) {
    extern __shared__ double dyn[];                   // max nx*n over the runs
    __shared__ double sh[kBlock];
    __shared__ double sl[kBlock];
    const EnsBlock b = map[blockIdx.x];
    const EnsRun &r = runs[b.run];
    dft_forward_body(r.rho, r.T, r.kwave, r.coef, r.fa, b.local, dyn, sh, sl);
}

// energy: null, or one value per run (written for the run's field).
// last != 0 (stage 3): adaptive runs also write their next dt from this
// field (the field the reference's adapt_time_step reads after the step).
__global__ __launch_bounds__(kBlock) void k_ens_inverse(const EnsRun *__restrict__ runs,
                                                        const EnsBlock *__restrict__ map,
                                                        double *__restrict__ energy, int last) {
    extern __shared__ double2 cdyn[];                 // max 2*M over the runs
    __shared__ InvLds ls;
    const EnsBlock b = map[blockIdx.x];
    const EnsRun &r = runs[b.run];
    FieldArgs fa = r.fa;
    fa.dt_next = (last && r.adaptive) ? r.clock + 2 : nullptr;
    fa.status = last ? r.status : nullptr;
    dft_inverse_body(r.coef, r.phases, r.wts, r.E, nullptr, r.enpart, r.mxpart, r.ticket,
                     energy ? energy + b.run : nullptr, nullptr, fa, b.local, r.inv_blocks, cdyn, ls);
}

// After a step of every run: adaptive runs take the next dt and advance their
// time by it (main/timestep.py:96,184), as k_clock_advance does for one run.
__global__ void k_ens_clock_advance(const EnsRun *__restrict__ runs, int count) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        if (!runs[i].adaptive) continue;
        double *c = runs[i].clock;
        const double dt = c[2];
        const double t = c[1] + dt;
        c[0] = dt;
        c[1] = t;
    }
}

// (time, dt) of every run's clock -> out[2 i], out[2 i + 1] (fixed-dt runs: 0, 0)
__global__ void k_ens_clocks(const EnsRun *__restrict__ runs, int count, double *__restrict__ out) {
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const bool a = runs[i].adaptive != 0;
        out[2 * i] = a ? runs[i].clock[1] : 0.0;
        out[2 * i + 1] = a ? runs[i].clock[0] : 0.0;
    }
}

// LDS of the ensemble stage kernel: the largest single-run layout
template <int S>
struct EnsStageLds {
    static constexpr size_t m67 = StageLds<6, S>::bytes > StageLds<7, S>::bytes ? StageLds<6, S>::bytes
                                                                                 : StageLds<7, S>::bytes;
    static constexpr size_t m89 = StageLds<8, S>::bytes > StageLds<9, S>::bytes ? StageLds<8, S>::bytes
                                                                                 : StageLds<9, S>::bytes;
    static constexpr size_t bytes = m67 > m89 ? m67 : m89;
};

// One pencil p of run r in RK stage `stage` (1, 2, 3) of main/timestep.py:141-187:
//   1: f1 = f + dt L(f)   2: f2 = 3/4 f + 1/4 f1 + 1/4 dt L(f1)
//   3: f  = 1/3 f + 2/3 f2 + 2/3 dt L(f2)
template <int N, int MODE, int S>
__device__ __forceinline__ void ens_stage_pencil(const EnsRun &r, int p, int stage, unsigned char *lds) {
    constexpr int L = kElemLayout;
    using T = typename Store<S>::E;
    using C = typename Store<S>::C;
    using LY = StageLds<N, S>;
    StageArgs s = r.s;
    s.c0 = r.cf[stage - 1][0];
    s.c1 = r.cf[stage - 1][1];
    s.c2dt = r.adaptive ? r.c2[stage - 1] * r.clock[0] : r.cf[stage - 1][2];   // c2 * dt as the host forms it
    T *fs = static_cast<T *>(r.f);
    const T *fin = stage == 1 ? fs : static_cast<const T *>(stage == 2 ? r.f1 : r.f2);
    const T *fgh = fs;
    const T *fn = fs;
    T *fout = static_cast<T *>(stage == 1 ? r.f1 : stage == 2 ? r.f2 : r.f);
    const double *E = r.E;
    const double *unodes = r.unodes;
    const double *vnodes = r.vnodes;
    const OpTables *tabp = r.tab;
    const Pencil pc = decode_pencil(s, p);
    C *ft = reinterpret_cast<C *>(lds + LY::o_ft);
    C *vh = reinterpret_cast<C *>(lds + LY::o_vh);
    C *uh = reinterpret_cast<C *>(lds + LY::o_uh);
    C *acc_l = reinterpret_cast<C *>(lds + LY::o_acc);
    double *bgs = reinterpret_cast<double *>(lds + LY::o_bgs);
    double *red_own = reinterpret_cast<double *>(lds + LY::o_red);
    // opaque per work item: keeps the thread-index arithmetic of the four
    // node-count bodies from being hoisted out of the item loop (it would
    // stay live across all of them)
    int stage_tid = threadIdx.x;
    asm volatile("" : "+v"(stage_tid));
#include "vl_stage_body.inc"
}

// items[i] = (run, pencil); MODE 1 for stage 1, MODE 2 for stages 2 and 3
template <int MODE, int S>
__global__ __launch_bounds__(kBlock, 3) void k_ens_stage(const EnsRun *__restrict__ runs,
                                                         const int2 *__restrict__ items, int n_items,
                                                         int stage) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[EnsStageLds<S>::bytes];
    const int G = gridDim.x;
    const int me = xcd_remap(blockIdx.x, G);
    for (int base = 0; base < n_items; base += G) {
        const int it = base + me;
        if (it >= n_items) break;                     // later rounds are shorter still
        const int2 w = items[it];
        const EnsRun &r = runs[w.x];
        __syncthreads();                              // LDS of the previous pencil is free
        switch (r.n) {
            case 6: ens_stage_pencil<6, MODE, S>(r, w.y, stage, lds); break;
            case 7: ens_stage_pencil<7, MODE, S>(r, w.y, stage, lds); break;
            case 8: ens_stage_pencil<8, MODE, S>(r, w.y, stage, lds); break;
            case 9: ens_stage_pencil<9, MODE, S>(r, w.y, stage, lds); break;
            default: break;
        }
    }
}

}  // namespace vl
