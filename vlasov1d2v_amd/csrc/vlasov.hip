// C ABI of the gfx950 1D2V nodal-DG Vlasov-Poisson SSP-RK3 path.
// Declarations and the reference interface each entry point replaces:
// include/vlasov1d2v.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "vlasov1d2v.h"
#include "vl_common.h"
#include "vl_field.h"
#include "vl_stage.h"
#include "vl_stage_mx.h"
#include "vl_stage_rs.h"
#include "vl_stage_rs2.h"
#include "vl_ensemble.h"
#include "vl_init.h"
#include "vl_central.h"
#include "vl_export.h"
#include "vl_elementwise.h"

#include <vector>

using namespace vl;

struct vl_geom {
    int n, nx, nu, nv, M, storage;
    double jx, ju, jv, acc, bfield, pcoef, dx, length;
    OpTables tab;
    int TV, nchunk, tiles_per_ex, tiles;
    int xseg, pencils;     // stage kernel: x elements per pencil segment, pencils launched
    int mfma;              // 1: matrix-core stage kernel for the 16-bit forms (element-major)
    int row_split;         // 1: the 16-bit forms with n = 8, 9 prefer the row-split layout
    int nchunk4, tiles_per_ex4, tiles4, pencils4;   // row-split kernels: groups of 4 plane elements per tile
    int xseg4;             // row-split stage kernel: x elements per pencil
    int rounding;          // 16-bit forms, row-split kernel: 0 nearest-even, 1 stochastic
    int rs_gen;            // row-split stage kernel at n = 9: 1 k_stage_rs, 2 k_stage_rs2
    int64_t sxe_elem;      // values per x element in the element-major layout
    int64_t sx, sxe, su, sue;
    size_t state_elems;
    int device;
    int flip;              // pencil order of the next step's first stage (alternates)
    double *d_xnodes, *d_unodes, *d_vnodes, *d_wts, *d_kwave;
    double2 *d_T, *d_phases;
    double *d_bg;          // reduced forms: G+, G-, Hv background tables
    float *d_bg32;         // reduced forms: scaled operator images of f0 (row-split kernel)
    unsigned *d_mxfrag;    // row-split kernel: operator fragments per lane (BF16, FP16)
    OpTables *d_tab;       // device copy of tab (read by the stage kernel)
    // central-flux Poisson (optional): inverse operator (K+1)^2, gradient
    // operator transposed K*K, mass n*n, and its scratch
    double *d_cf_inv, *d_cf_gradT, *d_cf_mass, *d_cf_rhs, *d_cf_phi, *d_cf_e;
    double2 *d_cf_coef;
    std::vector<double> f0_host;   // reduced forms: background f0 (u, v), host copy
    double *d_f0;          // reduced forms: device copy of f0 (run-file records)
    double scale, inv_scale;       // reduced forms: stored g = (f - f0) * scale
    void **stage_events;   // optional: 6 hipEvent_t recorded around the stage kernels
    // adaptive time step (main/timestep.py:231-242): Courant number, cell
    // sizes, max |u| and max |v| over the ghosted nodes
    int cfl_set;
    double cfl_courant, cfl_dx, cfl_du, cfl_dv, umax, vmax;
    // workspace layout (byte offsets)
    size_t off_partials, off_rho, off_coef, off_enpart, off_mxpart, off_ticket, off_clock, ws_bytes;
    int inv_blocks;
};

namespace {

// relative tolerance of the centro-antisymmetry check of the volume operator
constexpr double kUpSymTol = 1e-10;

// Workspace of vl_rk3_step_noalloc: one library-owned array per GPU (each
// device holds its own copy until the process ends).  In host code `scratch`
// names a host-side stand-in, not a device address: kernels and copies use
// the per-device address that hipGetSymbolAddress returns (g_scratch).
constexpr size_t kScratchBytes = (size_t)VL_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

// Resolves the device address of `scratch` for `device` (current), once.
hipError_t resolve_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = p;
    return hipSuccess;
}

void *scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch[device];
}

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int hip_fail(hipError_t e, const char *where) {
    snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
    return (int)e;
}

#define VL_HIP(call)                                  \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(e_, #call); \
    } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int tile_tv(int n) {
    const int tv = kBlock / (n * n);
    return tv > 0 ? tv : 1;
}

bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

size_t storage_bytes(int storage) {
    switch (storage) {
        case VL_F64: return 8;
        case VL_F32: return 4;
        case VL_BF16: return 2;
        case VL_F16: return 2;
        default: return 0;
    }
}

StageArgs make_stage_args(const vl_geom *h, int layout = kRefLayout) {
    StageArgs s;
    std::memset(&s, 0, sizeof(s));
    s.nx = h->nx;
    s.nu = h->nu;
    s.nv = h->nv;
    s.nchunk = layout == kRowSplitLayout ? h->nchunk4 : h->nchunk;
    s.xseg = layout == kRowSplitLayout ? h->xseg4 : h->xseg;
    s.jx = h->jx;
    s.ju = h->ju;
    s.jv = h->jv;
    s.acc = h->acc;
    s.bfield = h->bfield;
    s.sx = h->sx;
    s.sxe = layout != kRefLayout ? h->sxe_elem : h->sxe;
    s.su = h->su;
    s.sue = h->sue;
    s.bg = h->d_bg;
    s.bg32 = h->d_bg32;
    s.mxfrag = h->d_mxfrag;
    s.rounding = h->rounding;
    s.scale = h->scale;
    s.inv_scale = h->inv_scale;
    return s;
}

FieldArgs make_field_args(const vl_geom *h, int from_rho) {
    FieldArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = h->nx;
    a.n = h->n;
    a.M = h->M;
    a.from_rho = from_rho;
    a.poisson_coef = h->pcoef;
    a.dx = h->dx;
    a.length = h->length;
    a.acc = h->acc;
    return a;
}

// CFL constants of the adaptive step, formed as main/timestep.py:235-242 forms
// the speeds: s0 = max|u|, bv = |acc B max|v||, s2 = |acc |B max|u|||.
void set_cfl_args(const vl_geom *h, FieldArgs &a) {
    a.cfl_courant = h->cfl_courant;
    a.cfl_s0 = std::fabs(h->umax);
    a.cfl_bv = std::fabs(h->acc * h->bfield * h->vmax);
    a.cfl_s2 = std::fabs(h->acc * std::fabs(h->bfield * h->umax));
    a.cfl_dx = h->cfl_dx;
    a.cfl_du = h->cfl_du;
    a.cfl_dv = h->cfl_dv;
}

// ---- launch helpers (template dispatch over n) --------------------------

// Stage kernels read the time step from kernel arguments; adaptive runs
// (s.dt_dev set) launch the ADT variants, built for the resident layouts
// only: element-major FP64/FP32 here, the 16-bit kernels below.
template <int N, int MODE, int S, int L>
hipError_t launch_stage_ns(const vl_geom *h, const void *fin, const void *fgh, const void *fn, void *fout,
                           const double *E, const StageArgs &s, hipStream_t st) {
    using T = typename Store<S>::E;
    if (s.dt_dev) {
        if constexpr (L == kElemLayout && MODE != kModeRhs && (S == VL_F64 || S == VL_F32)) {
            hipLaunchKernelGGL((k_stage<N, MODE, S, L, true>), dim3(h->pencils), dim3(kBlock), 0, st,
                               (const T *)fin, (const T *)fgh, (const T *)fn, (T *)fout, E, h->d_unodes,
                               h->d_vnodes, s, h->d_tab);
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    hipLaunchKernelGGL((k_stage<N, MODE, S, L>), dim3(h->pencils), dim3(kBlock), 0, st, (const T *)fin,
                       (const T *)fgh, (const T *)fn, (T *)fout, E, h->d_unodes, h->d_vnodes, s, h->d_tab);
    return hipGetLastError();
}

// switch over the node count; EXPR may use the compile-time constant NN
#define VL_DISPATCH_N(EXPR)                                  \
    switch (h->n) {                                          \
        case 2: { constexpr int NN = 2; return EXPR; }       \
        case 3: { constexpr int NN = 3; return EXPR; }       \
        case 4: { constexpr int NN = 4; return EXPR; }       \
        case 5: { constexpr int NN = 5; return EXPR; }       \
        case 6: { constexpr int NN = 6; return EXPR; }       \
        case 7: { constexpr int NN = 7; return EXPR; }       \
        case 8: { constexpr int NN = 8; return EXPR; }       \
        case 9: { constexpr int NN = 9; return EXPR; }       \
        case 10: { constexpr int NN = 10; return EXPR; }     \
        default: return hipErrorInvalidValue;                \
    }

template <int N, int MODE, int S>
hipError_t launch_stage_mx_ns(const vl_geom *h, const void *fin, const void *fgh, const void *fn, void *fout,
                              const double *E, const StageArgs &s, hipStream_t st) {
    if (s.dt_dev)
        hipLaunchKernelGGL((k_stage_mx<N, MODE, S, true>), dim3(h->pencils), dim3(kBlock), 0, st,
                           (const uint16_t *)fin, (const uint16_t *)fgh, (const uint16_t *)fn, (uint16_t *)fout, E,
                           h->d_unodes, h->d_vnodes, s, h->d_tab);
    else
        hipLaunchKernelGGL((k_stage_mx<N, MODE, S>), dim3(h->pencils), dim3(kBlock), 0, st, (const uint16_t *)fin,
                           (const uint16_t *)fgh, (const uint16_t *)fn, (uint16_t *)fout, E, h->d_unodes,
                           h->d_vnodes, s, h->d_tab);
    return hipGetLastError();
}

template <int MODE, int S>
hipError_t launch_stage_mx(const vl_geom *h, const void *fin, const void *fgh, const void *fn, void *fout,
                           const double *E, const StageArgs &s, hipStream_t st) {
    VL_DISPATCH_N((launch_stage_mx_ns<NN, MODE, S>(h, fin, fgh, fn, fout, E, s, st)))
}

template <int N, int MODE, int S>
hipError_t launch_stage_rs_ns(const vl_geom *h, const void *fin, const void *fgh, const void *fn, void *fout,
                              const double *E, const StageArgs &s, hipStream_t st) {
    if constexpr (N == 9) {
        if (h->rs_gen == 2) {
            if (s.dt_dev)
                hipLaunchKernelGGL((k_stage_rs2<N, MODE, S, true>), dim3(h->pencils4), dim3(kBlock), 0, st,
                                   (const uint16_t *)fin, (const uint16_t *)fgh, (const uint16_t *)fn,
                                   (uint16_t *)fout, E, h->d_unodes, h->d_vnodes, s, h->d_tab);
            else
                hipLaunchKernelGGL((k_stage_rs2<N, MODE, S>), dim3(h->pencils4), dim3(kBlock), 0, st,
                                   (const uint16_t *)fin, (const uint16_t *)fgh, (const uint16_t *)fn,
                                   (uint16_t *)fout, E, h->d_unodes, h->d_vnodes, s, h->d_tab);
            return hipGetLastError();
        }
    }
    if (s.dt_dev)
        hipLaunchKernelGGL((k_stage_rs<N, MODE, S, true>), dim3(h->pencils4), dim3(kBlock), 0, st,
                           (const uint16_t *)fin, (const uint16_t *)fgh, (const uint16_t *)fn, (uint16_t *)fout, E,
                           h->d_unodes, h->d_vnodes, s, h->d_tab);
    else
        hipLaunchKernelGGL((k_stage_rs<N, MODE, S>), dim3(h->pencils4), dim3(kBlock), 0, st, (const uint16_t *)fin,
                           (const uint16_t *)fgh, (const uint16_t *)fn, (uint16_t *)fout, E, h->d_unodes,
                           h->d_vnodes, s, h->d_tab);
    return hipGetLastError();
}

template <int MODE, int S>
hipError_t launch_stage_rs(const vl_geom *h, const void *fin, const void *fgh, const void *fn, void *fout,
                           const double *E, const StageArgs &s, hipStream_t st) {
    if (h->n == 8) return launch_stage_rs_ns<8, MODE, S>(h, fin, fgh, fn, fout, E, s, st);
    if (h->n == 9) return launch_stage_rs_ns<9, MODE, S>(h, fin, fgh, fn, fout, E, s, st);
    return hipErrorInvalidValue;
}

template <int MODE, int S, int L>
hipError_t launch_stage_sl(const vl_geom *h, const void *fin, const void *fgh, const void *fn, void *fout,
                           const double *E, const StageArgs &s, hipStream_t st) {
    VL_DISPATCH_N((launch_stage_ns<NN, MODE, S, L>(h, fin, fgh, fn, fout, E, s, st)))
}

// RHS mode exists for FP64 in the reference layout only (the drop-in
// DGFlux.semi_discrete_rhs); the stage modes for every storage and layout.
template <int MODE>
hipError_t launch_stage(const vl_geom *h, int storage, int layout, const void *fin, const void *fgh,
                        const void *fn, void *fout, const double *E, const StageArgs &s, hipStream_t st) {
#define VL_STAGE_S(SS)                                                                                  \
    return layout == kElemLayout ? launch_stage_sl<MODE, SS, kElemLayout>(h, fin, fgh, fn, fout, E, s, st) \
                                 : launch_stage_sl<MODE, SS, kRefLayout>(h, fin, fgh, fn, fout, E, s, st)
    if (layout == kRowSplitLayout) {
        if (storage == VL_BF16) return launch_stage_rs<MODE, VL_BF16>(h, fin, fgh, fn, fout, E, s, st);
        if (storage == VL_F16) return launch_stage_rs<MODE, VL_F16>(h, fin, fgh, fn, fout, E, s, st);
        return hipErrorInvalidValue;
    }
    if (layout == kElemLayout && h->mfma) {
        if (storage == VL_BF16) return launch_stage_mx<MODE, VL_BF16>(h, fin, fgh, fn, fout, E, s, st);
        if (storage == VL_F16) return launch_stage_mx<MODE, VL_F16>(h, fin, fgh, fn, fout, E, s, st);
    }
    switch (storage) {
        case VL_F64: VL_STAGE_S(VL_F64);
        case VL_F32: VL_STAGE_S(VL_F32);
        case VL_BF16: VL_STAGE_S(VL_BF16);
        case VL_F16: VL_STAGE_S(VL_F16);
        default: return hipErrorInvalidValue;
    }
#undef VL_STAGE_S
}
template <>
hipError_t launch_stage<kModeRhs>(const vl_geom *h, int storage, int layout, const void *fin, const void *fgh,
                                  const void *fn, void *fout, const double *E, const StageArgs &s,
                                  hipStream_t st) {
    if (storage != VL_F64 || layout != kRefLayout) return hipErrorInvalidValue;
    return launch_stage_sl<kModeRhs, VL_F64, kRefLayout>(h, fin, fgh, fn, fout, E, s, st);
}

template <int N, int S, int L>
hipError_t launch_moment_nsl(const vl_geom *h, const void *f, const StageArgs &s, hipStream_t st) {
    using T = typename Store<S>::E;
    hipLaunchKernelGGL((k_moment_partials<N, S, L>), dim3(h->tiles), dim3(kBlock), 0, st, (const T *)f, s,
                       h->tab);
    return hipGetLastError();
}

template <int S, int L>
hipError_t launch_moment_sl(const vl_geom *h, const void *f, const StageArgs &s, hipStream_t st) {
    VL_DISPATCH_N((launch_moment_nsl<NN, S, L>(h, f, s, st)))
}

template <int N, int S>
hipError_t launch_moment_rs_ns(const vl_geom *h, const void *f, const StageArgs &s, hipStream_t st) {
    hipLaunchKernelGGL((k_moment_rs<N, S>), dim3(h->tiles4), dim3(kBlock), 0, st, (const uint16_t *)f, s,
                       h->d_tab);
    return hipGetLastError();
}

template <typename U, int N>
hipError_t launch_convert_xminor_n(const vl_geom *h, dim3 grid, const U *src, U *dst, int te, hipStream_t st) {
    hipLaunchKernelGGL((k_convert_xminor<U, N>), grid, dim3(kBlock), 0, st, src, dst, te, h->nu, h->nv);
    return hipGetLastError();
}

template <typename U>
hipError_t convert_xminor(const vl_geom *h, dim3 grid, const U *src, U *dst, int te, hipStream_t st) {
    VL_DISPATCH_N((launch_convert_xminor_n<U, NN>(h, grid, src, dst, te, st)))
}

hipError_t launch_moment(const vl_geom *h, int storage, int layout, const void *f, const StageArgs &s,
                         hipStream_t st) {
    if (layout == kRowSplitLayout) {
        if (storage == VL_BF16 && h->n == 9) return launch_moment_rs_ns<9, VL_BF16>(h, f, s, st);
        if (storage == VL_BF16 && h->n == 8) return launch_moment_rs_ns<8, VL_BF16>(h, f, s, st);
        if (storage == VL_F16 && h->n == 9) return launch_moment_rs_ns<9, VL_F16>(h, f, s, st);
        if (storage == VL_F16 && h->n == 8) return launch_moment_rs_ns<8, VL_F16>(h, f, s, st);
        return hipErrorInvalidValue;
    }
#define VL_MOM_S(SS)                                                                   \
    return layout == kElemLayout ? launch_moment_sl<SS, kElemLayout>(h, f, s, st) \
                                 : launch_moment_sl<SS, kRefLayout>(h, f, s, st)
    switch (storage) {
        case VL_F64: VL_MOM_S(VL_F64);
        case VL_F32: VL_MOM_S(VL_F32);
        case VL_BF16: VL_MOM_S(VL_BF16);
        case VL_F16: VL_MOM_S(VL_F16);
        default: return hipErrorInvalidValue;
    }
#undef VL_MOM_S
}

double2 *ws_partials(const vl_geom *h, void *ws) {
    return reinterpret_cast<double2 *>(static_cast<char *>(ws) + h->off_partials);
}
double *ws_rho(const vl_geom *h, void *ws) {
    return reinterpret_cast<double *>(static_cast<char *>(ws) + h->off_rho);
}
double2 *ws_coef(const vl_geom *h, void *ws) {
    return reinterpret_cast<double2 *>(static_cast<char *>(ws) + h->off_coef);
}
double2 *ws_enpart(const vl_geom *h, void *ws) {
    return reinterpret_cast<double2 *>(static_cast<char *>(ws) + h->off_enpart);
}
double *ws_mxpart(const vl_geom *h, void *ws) {
    return reinterpret_cast<double *>(static_cast<char *>(ws) + h->off_mxpart);
}
unsigned int *ws_ticket(const vl_geom *h, void *ws) {
    return reinterpret_cast<unsigned int *>(static_cast<char *>(ws) + h->off_ticket);
}
// rounding seed: the word after the ticket (zeroed with it by vl_workspace_init)
unsigned int *ws_seed(const vl_geom *h, void *ws) { return ws_ticket(h, ws) + 1; }
// status word (VL_STATUS_*): the second word after the ticket (zeroed with it)
unsigned int *ws_status(const vl_geom *h, void *ws) { return ws_ticket(h, ws) + 2; }
// clock of adaptive runs: (dt, time, next dt)
double *ws_clock(const vl_geom *h, void *ws) {
    return reinterpret_cast<double *>(static_cast<char *>(ws) + h->off_clock);
}

template <int N>
hipError_t launch_reduce_n(const vl_geom *h, void *ws, double *rho, int tpe, unsigned *seed, hipStream_t st) {
    hipLaunchKernelGGL((k_reduce_partials<N>), dim3(h->nx), dim3(kBlock), 0, st, ws_partials(h, ws), tpe, rho,
                       seed);
    return hipGetLastError();
}

// partial tiles per x element of a layout's stage / moment kernels
int tiles_per_ex_of(const vl_geom *h, int layout) {
    return layout == kRowSplitLayout ? h->tiles_per_ex4 * 4 : h->tiles_per_ex;   // row-split: one per element
}

hipError_t launch_reduce(const vl_geom *h, void *ws, double *rho, hipStream_t st, int layout = kRefLayout,
                         unsigned *seed = nullptr) {
    const int tpe = tiles_per_ex_of(h, layout);
    switch (h->n) {
        case 2: return launch_reduce_n<2>(h, ws, rho, tpe, seed, st);
        case 3: return launch_reduce_n<3>(h, ws, rho, tpe, seed, st);
        case 4: return launch_reduce_n<4>(h, ws, rho, tpe, seed, st);
        case 5: return launch_reduce_n<5>(h, ws, rho, tpe, seed, st);
        case 6: return launch_reduce_n<6>(h, ws, rho, tpe, seed, st);
        case 7: return launch_reduce_n<7>(h, ws, rho, tpe, seed, st);
        case 8: return launch_reduce_n<8>(h, ws, rho, tpe, seed, st);
        case 9: return launch_reduce_n<9>(h, ws, rho, tpe, seed, st);
        case 10: return launch_reduce_n<10>(h, ws, rho, tpe, seed, st);
        default: return hipErrorInvalidValue;
    }
}

// spectral solve: forward DFT (M/16 workgroups), inverse sums (K/64 workgroups)
hipError_t launch_field(const vl_geom *h, const double *rho_or_cd, int from_rho, void *ws, double *E,
                        double *phi, double *energy, double *maxabs, hipStream_t st, double *dt_next = nullptr,
                        unsigned *status = nullptr) {
    FieldArgs a = make_field_args(h, from_rho);
    if (dt_next) {
        set_cfl_args(h, a);
        a.dt_next = dt_next;
    }
    a.status = status;
    const int K = h->nx * h->n;
    const int gf = (h->M + kDftModes - 1) / kDftModes;
    hipLaunchKernelGGL(k_dft_forward, dim3(gf), dim3(kBlock), (size_t)K * sizeof(double), st, rho_or_cd,
                       h->d_T, h->d_kwave, ws_coef(h, ws), a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dft_inverse, dim3(h->inv_blocks), dim3(kBlock), (size_t)2 * h->M * sizeof(double2),
                       st, ws_coef(h, ws), h->d_phases, h->d_wts, E, phi, ws_enpart(h, ws), ws_mxpart(h, ws),
                       ws_ticket(h, ws), energy, maxabs, a);
    return hipGetLastError();
}

// partials -> rho (workspace) -> E (+phi, energy)
hipError_t field_from_partials(const vl_geom *h, void *ws, double *E, double *phi, double *energy,
                               hipStream_t st, int layout, double *dt_next = nullptr, unsigned *status = nullptr) {
    hipError_t e = launch_reduce(h, ws, ws_rho(h, ws), st, layout, ws_seed(h, ws));
    if (e != hipSuccess) return e;
    return launch_field(h, ws_rho(h, ws), 1, ws, E, phi, energy, nullptr, st, dt_next, status);
}

hipError_t record_stage_event(const vl_geom *h, int idx, hipStream_t st) {
    if (!h->stage_events || !h->stage_events[idx]) return hipSuccess;
    return hipEventRecord((hipEvent_t)h->stage_events[idx], st);
}

// 16-bit encodings on the host, round to nearest even (finite inputs)
uint16_t host_bf16(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return (uint16_t)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
float host_bf16_dec(uint16_t h) {
    const uint32_t b = (uint32_t)h << 16;
    float x;
    std::memcpy(&x, &b, 4);
    return x;
}
uint16_t host_f16(float x) {
    const _Float16 h = (_Float16)x;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
float host_f16_dec(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return (float)h;
}

// Operator fragments of the row-split kernel: lane l (column lj = l % 16,
// slot group lg = l / 16) holds op[lj][4 lg + t], t = 0..3, split into a
// 16-bit hi part and a 16-bit lo part of the remainder.  Line slots: 0..7
// nodes 0..7, 8 node n-1 (coefficient only when n = 9), 9 node 0 (copy, 0),
// 10 left face (xi0), 11 right face (-xi1); rows >= n are zero.
std::vector<unsigned> build_mxfrag(const OpTables &tab, int n) {
    std::vector<unsigned> out(2 * 64 * 4, 0u);
    for (int fmt = 0; fmt < 2; ++fmt)
        for (int l = 0; l < 64; ++l) {
            const int lj = l % 16, lg = l / 16;
            uint16_t hi[4], lo[4];
            for (int t = 0; t < 4; ++t) {
                const int sl = 4 * lg + t;
                double c = 0.0;
                if (lj < n) {
                    if (sl < 8 && sl < n) c = tab.up[lj * n + sl];
                    else if (sl == 8) c = n == 9 ? tab.up[lj * n + 8] : 0.0;
                    else if (sl == 10) c = tab.xi0[lj];
                    else if (sl == 11) c = -tab.xi1[lj];
                }
                hi[t] = fmt == 0 ? host_bf16((float)c) : host_f16((float)c);
                const double rem = c - (double)(fmt == 0 ? host_bf16_dec(hi[t]) : host_f16_dec(hi[t]));
                lo[t] = fmt == 0 ? host_bf16((float)rem) : host_f16((float)rem);
            }
            unsigned *o = &out[(size_t)(fmt * 64 + l) * 4];
            o[0] = hi[0] | ((unsigned)hi[1] << 16);
            o[1] = hi[2] | ((unsigned)hi[3] << 16);
            o[2] = lo[0] | ((unsigned)lo[1] << 16);
            o[3] = lo[2] | ((unsigned)lo[3] << 16);
        }
    return out;
}

// The row-split kernels address a whole state array through one buffer
// resource with 32-bit byte offsets (and an out-of-range offset of
// 0x7fffffff that must fall outside it): states of 2^31 bytes or more are
// refused.
bool row_split_fits(const vl_geom *h) {
    return (int64_t)(h->nx + 2) * h->sxe_elem * 2 < 2147483647LL;
}

int check_handle(const vl_geom *h) {
    if (!h) return fail(VL_E_HANDLE, "null handle");
    return 0;
}

template <typename T>
hipError_t alloc_copy(T **dst, const T *src, size_t count) {
    hipError_t e = hipMalloc((void **)dst, count * sizeof(T));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
}

// Background RHS tables of the reduced forms from f0 (host, FP64):
//   G+[eu,iu,gv] = Ju (sum_j up[iu][j] f0_j + xi0[iu] f0(eu-1, n-1) - xi1[iu] f0_{n-1})
//   G-[eu,iu,gv] = Ju (sum_j up[iu][j] f0_j + xi0[iu] f0_0 - xi1[iu] f0(eu+1, 0))
// so that L_u(f0) = a_u G+ (a_u >= 0) or a_u G- (a_u < 0) on the u-line, and
//   Hv[eu,iu,gv] = L_v(f0) with a_v = acc (-B u), the upwind choice of
// main/fluxes.py:181-207 for one v-line.  Ghost rows/columns stay zero.
int upload_background(vl_geom *h) {
    const int n = h->n, nu = h->nu, nv = h->nv;
    const int64_t W = (int64_t)(nv + 2) * n;           // row length (v nodes)
    const int64_t tbl = (int64_t)(nu + 2) * n * W;
    std::vector<double> bg((size_t)(3 * tbl), 0.0);
    // FP32 tables of the row-split kernel: scale * (A+ f0, A- f0) along u and
    // scale * A f0 along v (upwinded by the sign of a_v), no speed or Jacobian
    std::vector<float> bg32((size_t)(3 * tbl), 0.0f);
    const double *f0 = h->f0_host.data();
    auto F0 = [&](int eu, int iu, int64_t gv) { return f0[((int64_t)eu * n + iu) * W + gv]; };
    std::vector<double> unodes((size_t)(nu + 2) * n);
    VL_HIP(hipMemcpy(unodes.data(), h->d_unodes, unodes.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int eu = 1; eu <= nu; ++eu)
        for (int ev = 1; ev <= nv; ++ev)
            for (int iv = 0; iv < n; ++iv) {
                const int64_t gv = (int64_t)ev * n + iv;
                for (int i = 0; i < n; ++i) {
                    double vol = 0.0;
                    for (int j = 0; j < n; ++j) vol += h->tab.up[i * n + j] * F0(eu, j, gv);
                    const double gp = vol + h->tab.xi0[i] * F0(eu - 1, n - 1, gv) - h->tab.xi1[i] * F0(eu, n - 1, gv);
                    const double gm = vol + h->tab.xi0[i] * F0(eu, 0, gv) - h->tab.xi1[i] * F0(eu + 1, 0, gv);
                    const int64_t o = ((int64_t)eu * n + i) * W + gv;
                    bg[o] = h->ju * gp;
                    bg[tbl + o] = h->ju * gm;
                    bg32[o] = (float)(h->scale * gp);
                    bg32[tbl + o] = (float)(h->scale * gm);
                }
            }
    for (int eu = 1; eu <= nu; ++eu)
        for (int iu = 0; iu < n; ++iu) {
            const double av = h->acc * (-h->bfield * unodes[(size_t)eu * n + iu]);
            const bool pos = av >= 0.0;
            const int64_t row = ((int64_t)eu * n + iu) * W;
            for (int ev = 1; ev <= nv; ++ev) {
                const double *fr = f0 + row + (int64_t)ev * n;
                const double fl_nb = f0[row + (int64_t)ev * n - 1];
                const double fr_nb = f0[row + (int64_t)(ev + 1) * n];
                for (int i = 0; i < n; ++i) {
                    double vol = 0.0;
                    for (int j = 0; j < n; ++j) vol += h->tab.up[i * n + j] * fr[j];
                    const double face = pos ? (h->tab.xi0[i] * fl_nb - h->tab.xi1[i] * fr[n - 1])
                                            : (h->tab.xi0[i] * fr[0] - h->tab.xi1[i] * fr_nb);
                    bg[2 * tbl + row + (int64_t)ev * n + i] = h->jv * av * (vol + face);
                    bg32[2 * tbl + row + (int64_t)ev * n + i] = (float)(h->scale * (vol + face));
                }
            }
        }
    if (!h->d_bg) VL_HIP(hipMalloc((void **)&h->d_bg, bg.size() * sizeof(double)));
    VL_HIP(hipMemcpy(h->d_bg, bg.data(), bg.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!h->d_bg32) VL_HIP(hipMalloc((void **)&h->d_bg32, bg32.size() * sizeof(float)));
    VL_HIP(hipMemcpy(h->d_bg32, bg32.data(), bg32.size() * sizeof(float), hipMemcpyHostToDevice));
    return VL_OK;
}

}  // namespace

extern "C" {

const char *vl_last_error_string(void) { return g_err; }

int vl_geom_create(const vl_geom_desc *d, vl_geom_t *out) {
    if (!d || !out) return fail(VL_E_NULL, "vl_geom_create: null desc or out");
    *out = nullptr;
    if (d->n < 2 || d->n > kMaxN) return fail(VL_E_ARG, "n=%d outside [2,%d]", d->n, kMaxN);
    if (d->nx < 1 || d->nu < 1 || d->nv < 1) return fail(VL_E_ARG, "element counts must be >= 1");
    if (d->n_modes < 1) return fail(VL_E_ARG, "n_modes must be >= 1");
    if (storage_bytes(d->storage) == 0) return fail(VL_E_ARG, "unknown storage %d", d->storage);
    if (!d->up || !d->xi || !d->weights || !d->u_nodes || !d->v_nodes || !d->wave_numbers ||
        !d->transform || !d->phases)
        return fail(VL_E_NULL, "vl_geom_create: null table");
    if (!(d->x_dx > 0) || !(d->x_length > 0)) return fail(VL_E_ARG, "x_dx/x_length must be > 0");
    // LDS of the spectral kernels: nx*n doubles (forward), 2*M complex (inverse)
    if ((int64_t)d->nx * d->n > 6144) return fail(VL_E_ARG, "nx*n=%d above 6144", d->nx * d->n);
    if (d->n_modes > 1536) return fail(VL_E_ARG, "n_modes=%d above 1536", d->n_modes);

    vl_geom *h = new (std::nothrow) vl_geom();
    if (!h) return fail(VL_E_ARG, "out of host memory");
    const int n = d->n;
    h->n = n;
    h->nx = d->nx;
    h->nu = d->nu;
    h->nv = d->nv;
    h->M = d->n_modes;
    h->storage = d->storage;
    h->jx = d->jx;
    h->ju = d->ju;
    h->jv = d->jv;
    h->acc = d->acc;
    h->bfield = d->b_field;
    h->pcoef = d->poisson_coef;
    h->dx = d->x_dx;
    h->length = d->x_length;
    for (int i = 0; i < n * n; ++i) h->tab.up[i] = d->up[i];
    for (int i = 0; i < n; ++i) {
        h->tab.xi0[i] = d->xi[2 * i];
        h->tab.xi1[i] = d->xi[2 * i + 1];
    }
    // even-odd split of up.  The Lobatto volume operator is
    // centro-antisymmetric, up[n-1-i][n-1-j] = -up[i][j]; a table built in
    // floating point (the reference evaluates polynomials whose odd/even
    // coefficients are not exactly zero, main/basis.py:176-188) misses that by
    // rounding (4e-13 of max|up| at n = 9).  Tables within kUpSymTol of it are
    // accepted and made exactly antisymmetric (the mean of each entry pair),
    // which every kernel then uses; larger defects are an error.
    {
        const int H = n / 2, M = n / 2;
        double amax = 0.0;
        for (int i = 0; i < n * n; ++i) amax = std::max(amax, std::fabs(h->tab.up[i]));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double a = h->tab.up[i * n + j], b = h->tab.up[(n - 1 - i) * n + (n - 1 - j)];
                if (!std::isfinite(a) || std::fabs(a + b) > kUpSymTol * amax) {
                    delete h;
                    return fail(VL_E_ARG, "volume operator is not centro-antisymmetric at (%d,%d): %.3e",
                                i, j, std::fabs(a + b));
                }
            }
        double sym[kMaxN * kMaxN];
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                sym[i * n + j] = 0.5 * (h->tab.up[i * n + j] - h->tab.up[(n - 1 - i) * n + (n - 1 - j)]);
        for (int i = 0; i < n * n; ++i) h->tab.up[i] = sym[i];
        for (int i = 0; i < H; ++i) {
            for (int j = 0; j < H; ++j) {
                h->tab.eoP[i * H + j] = 0.5 * (h->tab.up[i * n + j] + h->tab.up[(n - 1 - i) * n + j]);
                h->tab.eoQ[i * H + j] = 0.5 * (h->tab.up[i * n + j] - h->tab.up[(n - 1 - i) * n + j]);
                h->tab.eoPf[i * H + j] = (float)h->tab.eoP[i * H + j];
                h->tab.eoQf[i * H + j] = (float)h->tab.eoQ[i * H + j];
            }
            if (n % 2) {
                h->tab.eoM[i] = h->tab.up[i * n + M];
                h->tab.eoR[i] = h->tab.up[M * n + i];
                h->tab.eoMf[i] = (float)h->tab.eoM[i];
                h->tab.eoRf[i] = (float)h->tab.eoR[i];
            }
        }
    }
    for (int i = 0; i < n * n; ++i) h->tab.upf[i] = (float)h->tab.up[i];
    for (int i = 0; i < n; ++i) {
        h->tab.xi0f[i] = (float)h->tab.xi0[i];
        h->tab.xi1f[i] = (float)h->tab.xi1[i];
    }
    // FP16 storage keeps g * 2^16 so that the deviation stays in the normal range
    h->scale = d->storage == VL_F16 ? 65536.0 : 1.0;
    h->inv_scale = 1.0 / h->scale;
    // moment weights exactly as main/grid.py:297-301 forms them:
    // (w_iu * w_iv) / (Ju * Jv)
    for (int iu = 0; iu < n; ++iu)
        for (int iv = 0; iv < n; ++iv)
            h->tab.wq[iu * n + iv] = (d->weights[iu] * d->weights[iv]) / (d->ju * d->jv);
    for (int i = 0; i < n; ++i) {
        h->tab.wu[i] = (float)(d->weights[i] / (d->ju * d->jv));
        h->tab.wv[i] = (float)d->weights[i];
    }
    for (int i = 0; i < (d->nu + 2) * n; ++i) h->umax = std::max(h->umax, std::fabs(d->u_nodes[i]));
    for (int i = 0; i < (d->nv + 2) * n; ++i) h->vmax = std::max(h->vmax, std::fabs(d->v_nodes[i]));
    h->TV = tile_tv(n);
    h->nchunk = (d->nv + h->TV - 1) / h->TV;
    h->tiles_per_ex = d->nu * h->nchunk;
    const int64_t tiles = (int64_t)d->nx * h->tiles_per_ex;
    if (tiles > 2147483647LL) {
        delete h;
        return fail(VL_E_ARG, "grid too large");
    }
    h->tiles = (int)tiles;
    // pencils of the stage kernel: about 6500 workgroups (measured best on
    // MI355X for the 116.6M-DOF grid among 2, 4, 8 and 16 x elements per
    // pencil: 8, with 4 workgroups per CU for FP64); host-tunable with
    // vl_set_tuning
    {
        const int base = d->nu * h->nchunk;
        int nseg = (6500 + base - 1) / base;
        nseg = std::max(1, std::min(nseg, d->nx));
        h->xseg = (d->nx + nseg - 1) / nseg;
        h->mfma = 1;
        h->pencils = base * ((d->nx + h->xseg - 1) / h->xseg);
    }
    // row-split kernels: tiles of 4 consecutive elements of the velocity
    // plane in (eu, ev) order (decode_pencil_rs), one wave each
    h->row_split = 1;
    h->rounding = 1;
    h->rs_gen = 1;
    h->nchunk4 = (d->nu * d->nv + 3) / 4;
    h->tiles_per_ex4 = h->nchunk4;
    h->tiles4 = (int)std::min<int64_t>((int64_t)d->nx * h->tiles_per_ex4, 2147483647LL);
    // x-segments: a power of two of them, about 2600 workgroups (16 x elements
    // per pencil, 4 segments, at 116.6M DOF: 0.5-1.5 % faster than 8 elements
    // on MI355X; segment counts that leave a short last segment, e.g. 5 of 13,
    // were slower)
    {
        const int want = std::max(1, 2600 / std::max(1, h->tiles_per_ex4));
        int nseg = 1;
        while (2 * nseg <= want) nseg *= 2;
        nseg = std::max(1, std::min(nseg, d->nx));
        h->xseg4 = (d->nx + nseg - 1) / nseg;
    }
    h->pencils4 = h->tiles_per_ex4 * ((d->nx + h->xseg4 - 1) / h->xseg4);
    h->su = (int64_t)(d->nv + 2) * n;
    h->sue = h->su * n;
    h->sx = h->sue * (d->nu + 2);
    h->sxe = h->sx * n;
    h->state_elems = (size_t)h->sxe * (d->nx + 2);
    h->sxe_elem = (int64_t)(d->nu + 2) * (d->nv + 2) * elem_pitch(n);
    // buffer addressing: one x element of the state must span < 2^31 bytes
    if (h->sxe_elem * (int64_t)storage_bytes(d->storage) >= 2147483647LL) {
        delete h;
        return fail(VL_E_ARG, "one x element of the state exceeds 2^31 bytes");
    }

    size_t off = 0;
    h->off_partials = off;
    off = align_up(off + (size_t)std::max<int64_t>(tiles, 4 * (int64_t)h->tiles4) * n * sizeof(double2), 256);
    h->off_rho = off;
    off = align_up(off + (size_t)d->nx * n * sizeof(double), 256);
    h->off_coef = off;
    off = align_up(off + (size_t)2 * h->M * sizeof(double2), 256);
    h->inv_blocks = (d->nx * n + kInvNodes - 1) / kInvNodes;
    h->off_enpart = off;
    off = align_up(off + (size_t)h->inv_blocks * sizeof(double2), 256);
    h->off_mxpart = off;
    off = align_up(off + (size_t)h->inv_blocks * sizeof(double), 256);
    h->off_ticket = off;
    off = align_up(off + 256, 256);
    h->off_clock = off;
    off = align_up(off + 4 * sizeof(double), 256);
    h->ws_bytes = off;

    hipError_t e = hipGetDevice(&h->device);
    const int K = d->nx * n;
    if (e == hipSuccess) e = alloc_copy(&h->d_tab, &h->tab, 1);
    if (e == hipSuccess) {
        const std::vector<unsigned> fr = build_mxfrag(h->tab, n);
        e = alloc_copy(&h->d_mxfrag, fr.data(), fr.size());
    }
    if (e == hipSuccess && d->x_nodes) e = alloc_copy(&h->d_xnodes, d->x_nodes, (size_t)(d->nx + 2) * n);
    if (e == hipSuccess) e = alloc_copy(&h->d_unodes, d->u_nodes, (size_t)(d->nu + 2) * n);
    if (e == hipSuccess) e = alloc_copy(&h->d_vnodes, d->v_nodes, (size_t)(d->nv + 2) * n);
    if (e == hipSuccess) e = alloc_copy(&h->d_wts, d->weights, (size_t)n);
    if (e == hipSuccess) e = alloc_copy(&h->d_kwave, d->wave_numbers, (size_t)h->M);
    if (e == hipSuccess)
        e = alloc_copy(&h->d_T, reinterpret_cast<const double2 *>(d->transform), (size_t)K * h->M);
    if (e == hipSuccess)
        e = alloc_copy(&h->d_phases, reinterpret_cast<const double2 *>(d->phases), (size_t)K * h->M);
    // set-up of the allocation-free entry point on this (current) device
    if (e == hipSuccess) e = resolve_scratch(h->device);
    if (e != hipSuccess) {
        vl_geom_destroy(h);
        return hip_fail(e, "vl_geom_create");
    }
    *out = h;
    return VL_OK;
}

int vl_geom_destroy(vl_geom_t h) {
    if (!h) return VL_OK;
    void *bufs[] = {h->d_xnodes, h->d_unodes, h->d_vnodes,  h->d_wts,     h->d_kwave,  h->d_T,
                    h->d_phases, h->d_bg,     h->d_bg32,    h->d_tab,     h->d_mxfrag, h->d_cf_inv,
                    h->d_cf_gradT, h->d_cf_mass, h->d_cf_rhs, h->d_cf_phi, h->d_cf_e,  h->d_cf_coef,
                    h->d_f0};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete h;
    return VL_OK;
}

int vl_set_physics(vl_geom_t h, double acc, double b_field, double poisson_coef) {
    if (int r = check_handle(h)) return r;
    if (!std::isfinite(acc) || !std::isfinite(b_field) || !std::isfinite(poisson_coef))
        return fail(VL_E_NOTFINITE, "vl_set_physics: non-finite coefficient");
    const bool changed = acc != h->acc || b_field != h->bfield;
    h->acc = acc;
    h->bfield = b_field;
    h->pcoef = poisson_coef;
    if (changed && !h->f0_host.empty())
        if (int r = upload_background(h)) return r;
    return VL_OK;
}

int vl_workspace_bytes(vl_geom_t h, size_t *bytes) {
    if (int r = check_handle(h)) return r;
    if (!bytes) return fail(VL_E_NULL, "null bytes");
    *bytes = h->ws_bytes;
    return VL_OK;
}

int vl_state_elems(vl_geom_t h, size_t *elems) {
    if (int r = check_handle(h)) return r;
    if (!elems) return fail(VL_E_NULL, "null elems");
    *elems = h->state_elems;
    return VL_OK;
}

int vl_rhs(vl_geom_t h, const double *f, const double *E, double *out, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!f || !E || !out) return fail(VL_E_NULL, "vl_rhs: null pointer");
    if (!aligned(f, 8) || !aligned(E, 8) || !aligned(out, 8)) return fail(VL_E_ALIGN, "vl_rhs: misaligned");
    StageArgs s = make_stage_args(h);
    s.periodic_x = 0;
    s.write_partials = 0;
    VL_HIP(launch_stage<kModeRhs>(h, VL_F64, kRefLayout, f, f, nullptr, out, E, s, (hipStream_t)stream));
    return VL_OK;
}

int vl_moment0(vl_geom_t h, const double *f, double *rho, void *ws, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!f || !rho || !ws) return fail(VL_E_NULL, "vl_moment0: null pointer");
    if (!aligned(f, 8) || !aligned(rho, 8) || !aligned(ws, 256)) return fail(VL_E_ALIGN, "vl_moment0: misaligned");
    hipStream_t st = (hipStream_t)stream;
    StageArgs s = make_stage_args(h);
    s.partials = ws_partials(h, ws);
    VL_HIP(launch_moment(h, VL_F64, kRefLayout, f, s, st));
    VL_HIP(launch_reduce(h, ws, rho, st));
    return VL_OK;
}

int vl_poisson(vl_geom_t h, const double *cd, double *E, double *phi, void *ws, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!cd || !E || !ws) return fail(VL_E_NULL, "vl_poisson: null pointer");
    if (!aligned(cd, 8) || !aligned(E, 8) || !aligned(ws, 256) || (phi && !aligned(phi, 8)))
        return fail(VL_E_ALIGN, "vl_poisson: misaligned");
    VL_HIP(launch_field(h, cd, 0, ws, E, phi, nullptr, nullptr, (hipStream_t)stream));
    return VL_OK;
}

int vl_field_energy(vl_geom_t h, const double *E, double *energy, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!E || !energy) return fail(VL_E_NULL, "vl_field_energy: null pointer");
    hipLaunchKernelGGL(k_energy, dim3(1), dim3(kFieldBlock), 0, (hipStream_t)stream, E, h->d_wts, h->nx,
                       h->n, energy);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_max_abs_acc_e(vl_geom_t h, const double *E, double *out, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!E || !out) return fail(VL_E_NULL, "vl_max_abs_acc_e: null pointer");
    hipLaunchKernelGGL(k_max_abs_acc_e, dim3(1), dim3(kFieldBlock), 0, (hipStream_t)stream, E,
                       (h->nx + 2) * h->n, h->acc, out);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int swap_ghosts_layout(vl_geom_t h, void *f, int layout, void *stream) {
    const size_t eb = storage_bytes(h->storage);
    const int64_t plane = layout != kRefLayout ? h->sxe_elem : h->sxe;
    const int64_t blocks = std::min<int64_t>((2 * plane + kBlock - 1) / kBlock, 8192);
    hipStream_t st = (hipStream_t)stream;
    if (eb == 8)
        hipLaunchKernelGGL(k_swap_ghosts<double>, dim3(blocks), dim3(kBlock), 0, st, (double *)f, plane, h->nx);
    else if (eb == 4)
        hipLaunchKernelGGL(k_swap_ghosts<float>, dim3(blocks), dim3(kBlock), 0, st, (float *)f, plane, h->nx);
    else
        hipLaunchKernelGGL(k_swap_ghosts<uint16_t>, dim3(blocks), dim3(kBlock), 0, st, (uint16_t *)f, plane,
                           h->nx);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_swap_ghosts(vl_geom_t h, void *f, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!f) return fail(VL_E_NULL, "vl_swap_ghosts: null pointer");
    if (!aligned(f, storage_bytes(h->storage))) return fail(VL_E_ALIGN, "vl_swap_ghosts: misaligned");
    return swap_ghosts_layout(h, f, kRefLayout, stream);
}

}  // extern "C"

namespace {

// One SSP-RK3 step; dt_dev = null: fixed dt, else the adaptive clock of the
// workspace (dt read by the stage kernels, next dt from the stage-3 field).
int rk3_step(vl_geom_t h, void *fv, void *f1v, void *f2v, double *E, double *phi, double dt, bool adaptive,
             double *energy_out, double *clock_out, void *ws, int flags, void *stream, const char *what) {
    if (int r = check_handle(h)) return r;
    if (!fv || !f1v || !f2v || !E || !ws) return fail(VL_E_NULL, "%s: null pointer", what);
    const size_t eb = storage_bytes(h->storage);
    if (!aligned(fv, eb) || !aligned(f1v, eb) || !aligned(f2v, eb) || !aligned(E, 8) || !aligned(ws, 256))
        return fail(VL_E_ALIGN, "%s: misaligned pointer", what);
    if (!std::isfinite(dt)) return fail(VL_E_NOTFINITE, "%s: dt not finite", what);
    if (!adaptive && !(dt > 0.0)) return fail(VL_E_ARG, "%s: dt must be > 0", what);
    if (adaptive && !h->cfl_set) return fail(VL_E_ARG, "%s: adaptive steps need vl_set_cfl first", what);
    if (adaptive && !(flags & (VL_STEP_ELEMENT_MAJOR | VL_STEP_ROW_SPLIT)))
        return fail(VL_E_ARG, "%s: adaptive steps run in a resident layout (VL_STEP_ELEMENT_MAJOR / ROW_SPLIT)", what);
    if (adaptive && (flags & VL_STEP_ELEMENT_MAJOR) && (h->storage == VL_BF16 || h->storage == VL_F16) && !h->mfma)
        return fail(VL_E_ARG, "%s: adaptive 16-bit steps need the matrix-core kernels (VL_TUNE_MFMA 1)", what);
    if (h->storage != VL_F64 && !h->d_bg)
        return fail(VL_E_ARG, "%s: reduced storage needs vl_set_background first", what);
    void *f = fv, *f1 = f1v, *f2 = f2v;
    const int S = h->storage;
    if ((flags & VL_STEP_ELEMENT_MAJOR) && (flags & VL_STEP_ROW_SPLIT))
        return fail(VL_E_ARG, "%s: VL_STEP_ELEMENT_MAJOR and VL_STEP_ROW_SPLIT are exclusive", what);
    const int LAY = (flags & VL_STEP_ROW_SPLIT) ? kRowSplitLayout
                    : (flags & VL_STEP_ELEMENT_MAJOR) ? kElemLayout : kRefLayout;
    if (LAY != kRefLayout && (!aligned(fv, 16) || !aligned(f1v, 16) || !aligned(f2v, 16)))
        return fail(VL_E_ALIGN, "%s: element-wise arrays must be 16-byte aligned", what);
    if (LAY == kRowSplitLayout && ((h->storage != VL_BF16 && h->storage != VL_F16) || (h->n != 8 && h->n != 9)))
        return fail(VL_E_ARG, "%s: the row-split layout needs BF16/FP16 storage and n = 8 or 9", what);
    if (LAY == kRowSplitLayout && !row_split_fits(h))
        return fail(VL_E_ARG, "%s: the row-split kernels address the state with 32-bit offsets; this state "
                              "has 2^31 bytes or more (use the element-major layout)", what);
    hipStream_t st = (hipStream_t)stream;
    if (int r = swap_ghosts_layout(h, f, LAY, stream)) return r;
    StageArgs s = make_stage_args(h, LAY);
    s.periodic_x = 1;
    s.write_partials = 1;
    s.partials = ws_partials(h, ws);
    s.seed = ws_seed(h, ws);
    double *clock = ws_clock(h, ws);
    if (adaptive) s.dt_dev = clock;
    if (!(flags & VL_STEP_RHO_VALID)) VL_HIP(launch_moment(h, S, LAY, f, s, st));
    // stage 1: f1 = f + dt L(f)                       (main/timestep.py:153-163)
    VL_HIP(field_from_partials(h, ws, E, nullptr, nullptr, st, LAY));
    s.c0 = 0.0;
    s.c1 = 1.0;
    s.c2 = 1.0;
    s.c2dt = dt;
    // pencil orders alternate from stage to stage and, three stages per step,
    // from step to step (the results do not depend on the order)
    const int flip = h->flip;
    h->flip ^= 1;
    s.reverse = flip;
    VL_HIP(record_stage_event(h, 0, st));
    VL_HIP(launch_stage<kModeStage1>(h, S, LAY, f, f, nullptr, f1, E, s, st));
    VL_HIP(record_stage_event(h, 1, st));
    // stage 2: f2 = 3/4 f + 1/4 f1 + 1/4 dt L(f1)     (main/timestep.py:165-182)
    VL_HIP(field_from_partials(h, ws, E, nullptr, nullptr, st, LAY));
    s.c0 = 3.0 / 4.0;
    s.c1 = 1.0 / 4.0;
    s.c2 = 1.0 / 4.0;
    s.c2dt = s.c2 * dt;
    s.reverse = flip ^ 1;
    VL_HIP(record_stage_event(h, 2, st));
    VL_HIP(launch_stage<kModeStage>(h, S, LAY, f1, f, f, f2, E, s, st));
    VL_HIP(record_stage_event(h, 3, st));
    // stage 3: f = 1/3 f + 2/3 f2 + 2/3 dt L(f2); E, phi, energy of f2's field
    // (and, adaptive, the next dt from that field: main/timestep.py:184)
    VL_HIP(field_from_partials(h, ws, E, phi, energy_out, st, LAY, adaptive ? clock + 2 : nullptr,
                               ws_status(h, ws)));
    s.c0 = 1.0 / 3.0;
    s.c1 = 2.0 / 3.0;
    s.c2 = 2.0 / 3.0;
    s.c2dt = s.c2 * dt;
    s.reverse = flip;
    VL_HIP(record_stage_event(h, 4, st));
    VL_HIP(launch_stage<kModeStage>(h, S, LAY, f2, f, f, f, E, s, st));
    VL_HIP(record_stage_event(h, 5, st));
    if (adaptive) {
        hipLaunchKernelGGL(k_clock_advance, dim3(1), dim3(64), 0, st, clock, clock_out);
        VL_HIP(hipGetLastError());
    }
    return VL_OK;
}

}  // namespace

extern "C" {

int vl_rk3_step(vl_geom_t h, void *fv, void *f1v, void *f2v, double *E, double *phi, double dt,
                double *energy_out, void *ws, int flags, void *stream) {
    return rk3_step(h, fv, f1v, f2v, E, phi, dt, false, energy_out, nullptr, ws, flags, stream, "vl_rk3_step");
}

int vl_set_cfl(vl_geom_t h, double courant, double dx, double du, double dv) {
    if (int r = check_handle(h)) return r;
    if (!(courant > 0) || !(dx > 0) || !(du > 0) || !(dv > 0) || !std::isfinite(courant) || !std::isfinite(dx) ||
        !std::isfinite(du) || !std::isfinite(dv))
        return fail(VL_E_ARG, "vl_set_cfl: courant and cell sizes must be finite and > 0");
    h->cfl_courant = courant;
    h->cfl_dx = dx;
    h->cfl_du = du;
    h->cfl_dv = dv;
    h->cfl_set = 1;
    return VL_OK;
}

int vl_clock_init(vl_geom_t h, const double *E, double time, void *ws, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!E || !ws) return fail(VL_E_NULL, "vl_clock_init: null pointer");
    if (!aligned(E, 8) || !aligned(ws, 256)) return fail(VL_E_ALIGN, "vl_clock_init: misaligned pointer");
    if (!h->cfl_set) return fail(VL_E_ARG, "vl_clock_init: call vl_set_cfl first");
    if (!std::isfinite(time)) return fail(VL_E_NOTFINITE, "vl_clock_init: time not finite");
    FieldArgs a = make_field_args(h, 1);
    set_cfl_args(h, a);
    hipLaunchKernelGGL(k_clock_init, dim3(1), dim3(kFieldBlock), 0, (hipStream_t)stream, E, (h->nx + 2) * h->n, a,
                       time, ws_clock(h, ws));
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_rk3_step_adaptive(vl_geom_t h, void *fv, void *f1v, void *f2v, double *E, double *phi, double *energy_out,
                         double *clock_out, void *ws, int flags, void *stream) {
    if (clock_out && !aligned(clock_out, 8)) return fail(VL_E_ALIGN, "vl_rk3_step_adaptive: misaligned clock_out");
    return rk3_step(h, fv, f1v, f2v, E, phi, 0.0, true, energy_out, clock_out, ws, flags, stream,
                    "vl_rk3_step_adaptive");
}

int vl_rk3_step_noalloc(vl_geom_t h, void *fv, void *f1v, void *f2v, double *E, double *phi, double dt,
                        double *energy_out, int flags, void *stream) {
    if (int r = check_handle(h)) return r;
    if (h->ws_bytes > kScratchBytes)
        return fail(VL_E_WORKSPACE, "vl_rk3_step_noalloc: this geometry needs a %zu-byte workspace, above the "
                                    "%zu bytes of the library's scratch array (use vl_rk3_step)",
                    h->ws_bytes, kScratchBytes);
    void *ws = scratch_of(h->device);
    if (!ws) return fail(VL_E_ARG, "vl_rk3_step_noalloc: no scratch array on device %d", h->device);
    // the array keeps the previous call's contents (maybe another geometry's
    // layout): the reduction ticket must start at zero
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, (hipStream_t)stream, ws_ticket(h, ws), 1);
    VL_HIP(hipGetLastError());
    return rk3_step(h, fv, f1v, f2v, E, phi, dt, false, energy_out, nullptr, ws, flags, stream,
                    "vl_rk3_step_noalloc");
}

int vl_read_status(vl_geom_t h, void *ws, int clear, unsigned *status, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!ws || !status) return fail(VL_E_NULL, "vl_read_status: null pointer");
    if (!aligned(ws, 256)) return fail(VL_E_ALIGN, "vl_read_status: misaligned workspace");
    hipStream_t st = (hipStream_t)stream;
    VL_HIP(hipMemcpyAsync(status, ws_status(h, ws), sizeof(unsigned), hipMemcpyDeviceToHost, st));
    if (clear) {
        hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, ws_status(h, ws), 1);
        VL_HIP(hipGetLastError());
    }
    VL_HIP(hipStreamSynchronize(st));
    return VL_OK;
}

int vl_moment_partials(vl_geom_t h, const void *f, void *ws, int flags, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!f || !ws) return fail(VL_E_NULL, "vl_moment_partials: null pointer");
    if ((flags & VL_STEP_ELEMENT_MAJOR) && (flags & VL_STEP_ROW_SPLIT))
        return fail(VL_E_ARG, "vl_moment_partials: VL_STEP_ELEMENT_MAJOR and VL_STEP_ROW_SPLIT are exclusive");
    const int LAY = (flags & VL_STEP_ROW_SPLIT) ? kRowSplitLayout
                    : (flags & VL_STEP_ELEMENT_MAJOR) ? kElemLayout : kRefLayout;
    const size_t eb = storage_bytes(h->storage);
    if (!aligned(f, LAY != kRefLayout ? 16 : eb) || !aligned(ws, 256))
        return fail(VL_E_ALIGN, "vl_moment_partials: misaligned pointer");
    if (LAY == kRowSplitLayout && ((h->storage != VL_BF16 && h->storage != VL_F16) || (h->n != 8 && h->n != 9) ||
                                   !row_split_fits(h)))
        return fail(VL_E_ARG, "vl_moment_partials: no row-split layout for this geometry");
    StageArgs s = make_stage_args(h, LAY);
    s.partials = ws_partials(h, ws);
    VL_HIP(launch_moment(h, h->storage, LAY, f, s, (hipStream_t)stream));
    return VL_OK;
}

int vl_state_elems_layout(vl_geom_t h, int layout, size_t *elems) {
    if (int r = check_handle(h)) return r;
    if (!elems) return fail(VL_E_NULL, "null elems");
    if (layout != VL_LAYOUT_REFERENCE && layout != VL_LAYOUT_ELEMENT_MAJOR && layout != VL_LAYOUT_ROW_SPLIT)
        return fail(VL_E_ARG, "bad layout %d", layout);
    if (layout == VL_LAYOUT_ROW_SPLIT && h->n != 8 && h->n != 9)
        return fail(VL_E_ARG, "the row-split layout needs n = 8 or 9 (n = %d)", h->n);
    *elems = layout != VL_LAYOUT_REFERENCE ? (size_t)h->sxe_elem * (h->nx + 2) : h->state_elems;
    return VL_OK;
}

int vl_convert_layout(vl_geom_t h, const void *src, int src_layout, void *dst, int dst_layout, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!src || !dst) return fail(VL_E_NULL, "vl_convert_layout: null pointer");
    if (src == dst) return fail(VL_E_ARG, "vl_convert_layout: in-place conversion not supported");
    auto elementwise = [](int l) { return l == VL_LAYOUT_ELEMENT_MAJOR || l == VL_LAYOUT_ROW_SPLIT; };
    const bool to_elem = src_layout == VL_LAYOUT_REFERENCE && elementwise(dst_layout);
    const bool from_elem = elementwise(src_layout) && dst_layout == VL_LAYOUT_REFERENCE;
    if (!to_elem && !from_elem) return fail(VL_E_ARG, "vl_convert_layout: unsupported layout pair");
    const int lay = to_elem ? dst_layout : src_layout;
    if (lay == VL_LAYOUT_ROW_SPLIT && h->n != 8 && h->n != 9)
        return fail(VL_E_ARG, "the row-split layout needs n = 8 or 9 (n = %d)", h->n);
    if (lay == VL_LAYOUT_ROW_SPLIT && !row_split_fits(h))
        return fail(VL_E_ARG, "vl_convert_layout: a row-split state of 2^31 bytes or more is not supported");
    const size_t eb = storage_bytes(h->storage);
    if (!aligned(src, eb) || !aligned(dst, eb)) return fail(VL_E_ALIGN, "vl_convert_layout: misaligned");
    const int64_t total = (int64_t)h->state_elems;
    const int64_t blocks = std::min<int64_t>((total + kBlock - 1) / kBlock, 65536);
    hipStream_t st = (hipStream_t)stream;
    const int te = to_elem ? 1 : 0;
    if (lay == VL_LAYOUT_ELEMENT_MAJOR && (eb == 8 || eb == 4)) {
        // x-fastest element order (elem_x_minor): slab-wise through LDS
        const dim3 grid((unsigned)((h->nx + 2) * (h->nu + 2) * h->n));
        if (eb == 8) {
            VL_HIP(convert_xminor<double>(h, grid, (const double *)src, (double *)dst, te, st));
        } else {
            VL_HIP(convert_xminor<float>(h, grid, (const float *)src, (float *)dst, te, st));
        }
        return VL_OK;
    }
    if (eb == 8)
        hipLaunchKernelGGL(k_convert_layout<double>, dim3(blocks), dim3(kBlock), 0, st, (const double *)src,
                           (double *)dst, te, lay, h->n, h->nu, h->nv, total);
    else if (eb == 4)
        hipLaunchKernelGGL(k_convert_layout<float>, dim3(blocks), dim3(kBlock), 0, st, (const float *)src,
                           (float *)dst, te, lay, h->n, h->nu, h->nv, total);
    else
        hipLaunchKernelGGL(k_convert_layout<uint16_t>, dim3(blocks), dim3(kBlock), 0, st,
                           (const uint16_t *)src, (uint16_t *)dst, te, lay, h->n, h->nu, h->nv, total);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_workspace_init(vl_geom_t h, void *ws, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!ws) return fail(VL_E_NULL, "vl_workspace_init: null workspace");
    if (!aligned(ws, 256)) return fail(VL_E_ALIGN, "vl_workspace_init: misaligned");
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, (hipStream_t)stream, ws_ticket(h, ws), 64);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_set_tuning(vl_geom_t h, int key, int value) {
    if (int r = check_handle(h)) return r;
    if (key == VL_TUNE_XSEG) {
        if (value < 1) return fail(VL_E_ARG, "xseg must be >= 1");
        h->xseg = std::min(value, h->nx);
        h->pencils = h->nu * h->nchunk * ((h->nx + h->xseg - 1) / h->xseg);
        h->xseg4 = h->xseg;
        h->pencils4 = h->tiles_per_ex4 * ((h->nx + h->xseg4 - 1) / h->xseg4);
        return VL_OK;
    }
    if (key == VL_TUNE_ROW_SPLIT) {
        if (value != 0 && value != 1) return fail(VL_E_ARG, "row_split must be 0 or 1");
        h->row_split = value;
        return VL_OK;
    }
    if (key == VL_TUNE_ROUNDING) {
        if (value != 0 && value != 1) return fail(VL_E_ARG, "rounding must be 0 or 1");
        h->rounding = value;
        return VL_OK;
    }
    if (key == VL_TUNE_MFMA) {
        if (value != 0 && value != 1) return fail(VL_E_ARG, "mfma must be 0 or 1");
        h->mfma = value;
        return VL_OK;
    }
    if (key == VL_TUNE_RS_KERNEL) {
        if (value != 1 && value != 2) return fail(VL_E_ARG, "row-split kernel must be 1 or 2");
        h->rs_gen = value;
        return VL_OK;
    }
    return fail(VL_E_ARG, "unknown tuning key %d", key);
}

int vl_get_tuning(vl_geom_t h, int key, int *value) {
    if (int r = check_handle(h)) return r;
    if (!value) return fail(VL_E_NULL, "vl_get_tuning: null value");
    if (key == VL_TUNE_XSEG) {
        *value = h->xseg;
        return VL_OK;
    }
    if (key == VL_TUNE_MFMA) {
        *value = h->mfma;
        return VL_OK;
    }
    if (key == VL_TUNE_ROW_SPLIT) {
        *value = h->row_split;
        return VL_OK;
    }
    if (key == VL_TUNE_ROUNDING) {
        *value = h->rounding;
        return VL_OK;
    }
    if (key == VL_TUNE_RS_KERNEL) {
        *value = h->rs_gen;
        return VL_OK;
    }
    return fail(VL_E_ARG, "unknown tuning key %d", key);
}

int vl_resident_layout(vl_geom_t h, int *layout) {
    if (int r = check_handle(h)) return r;
    if (!layout) return fail(VL_E_NULL, "vl_resident_layout: null layout");
    const bool rs = (h->storage == VL_BF16 || h->storage == VL_F16) && (h->n == 8 || h->n == 9) && h->mfma &&
                    h->row_split && row_split_fits(h);
    *layout = rs ? VL_LAYOUT_ROW_SPLIT : VL_LAYOUT_ELEMENT_MAJOR;
    return VL_OK;
}

int vl_set_stage_events(vl_geom_t h, void **events) {
    if (int r = check_handle(h)) return r;
    h->stage_events = events;
    return VL_OK;
}

int vl_set_central_flux(vl_geom_t h, const double *inv_op, const double *grad_op, const double *mass) {
    if (int r = check_handle(h)) return r;
    if (!inv_op || !grad_op || !mass) return fail(VL_E_NULL, "vl_set_central_flux: null table");
    const int K = h->nx * h->n, n = h->n;
    const size_t ni = (size_t)(K + 1) * (K + 1), ng = (size_t)K * K;
    for (size_t i = 0; i < ni; ++i)
        if (!std::isfinite(inv_op[i])) return fail(VL_E_NOTFINITE, "vl_set_central_flux: non-finite inverse");
    std::vector<double> gT(ng);
    for (int r = 0; r < K; ++r)
        for (int c = 0; c < K; ++c) gT[(size_t)c * K + r] = grad_op[(size_t)r * K + c];
    hipError_t e = hipSuccess;
    if (!h->d_cf_inv) {
        e = hipMalloc((void **)&h->d_cf_inv, ni * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&h->d_cf_gradT, ng * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&h->d_cf_mass, (size_t)n * n * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&h->d_cf_rhs, (size_t)(K + 1) * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&h->d_cf_phi, (size_t)K * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&h->d_cf_e, (size_t)K * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void **)&h->d_cf_coef, (size_t)h->M * sizeof(double2));
    }
    if (e == hipSuccess) e = hipMemcpy(h->d_cf_inv, inv_op, ni * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_cf_gradT, gT.data(), ng * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_cf_mass, mass, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "vl_set_central_flux");
    return VL_OK;
}

int vl_poisson_central(vl_geom_t h, const double *cd, double *E, double *phi, int anti_alias, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!cd || !E) return fail(VL_E_NULL, "vl_poisson_central: null pointer");
    if (!aligned(cd, 8) || !aligned(E, 8) || (phi && !aligned(phi, 8)))
        return fail(VL_E_ALIGN, "vl_poisson_central: misaligned pointer");
    if (!h->d_cf_inv) return fail(VL_E_ARG, "vl_poisson_central: call vl_set_central_flux first");
    hipStream_t st = (hipStream_t)stream;
    const int K = h->nx * h->n, n = h->n, M = h->M;
    const double J = h->jx;
    const int gd = (M + kDftModes - 1) / kDftModes, gk = (K + kBlock - 1) / kBlock;
    double *ph = phi ? phi : h->d_cf_phi;
    // rhs, then phi = (inv_op rhs)[:-1] / J^2   (main/elliptic.py:93-98)
    hipLaunchKernelGGL(k_cf_rhs, dim3((K + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, st, cd, h->d_cf_mass,
                       h->d_cf_rhs, h->nx, n, h->pcoef);
    hipLaunchKernelGGL(k_matvec, dim3(K), dim3(kBlock), 0, st, h->d_cf_inv, h->d_cf_rhs,
                       anti_alias ? h->d_cf_e : ph, K, K + 1, 1.0 / (J * J), 0);
    if (anti_alias) {                                                 // main/elliptic.py:101-103
        hipLaunchKernelGGL(k_dft_plain, dim3(gd), dim3(kBlock), 0, st, h->d_cf_e, h->d_T, h->d_cf_coef, K, M,
                           h->dx, h->length);
        hipLaunchKernelGGL(k_dft_plain_sum, dim3(gk), dim3(kBlock), 0, st, h->d_cf_coef, h->d_phases, ph, K, M, 0,
                           0);
    }
    // E = -(J grad_op . phi)   (main/elliptic.py:106-108)
    hipLaunchKernelGGL(k_matvec, dim3(K), dim3(kBlock), 0, st, h->d_cf_gradT, ph, anti_alias ? h->d_cf_e : E, K,
                       K, -J, anti_alias ? 0 : n);
    if (anti_alias) {                                                 // main/elliptic.py:111-113
        hipLaunchKernelGGL(k_dft_plain, dim3(gd), dim3(kBlock), 0, st, h->d_cf_e, h->d_T, h->d_cf_coef, K, M,
                           h->dx, h->length);
        hipLaunchKernelGGL(k_dft_plain_sum, dim3(gk), dim3(kBlock), 0, st, h->d_cf_coef, h->d_phases, E, K, M, n,
                           n);
    } else {
        hipLaunchKernelGGL(k_field_ghosts, dim3(1), dim3(64), 0, st, E, h->nx, n);
    }
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_init_eigen(vl_geom_t h, const vl_eigen_desc *d, double *f, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!d || !f) return fail(VL_E_NULL, "vl_init_eigen: null pointer");
    if (!aligned(f, 8)) return fail(VL_E_ALIGN, "vl_init_eigen: misaligned f");
    if (!h->d_xnodes) return fail(VL_E_ARG, "vl_init_eigen: the geometry was created without x nodes");
    if (d->ring_j < 0 || d->ring_j > 170) return fail(VL_E_ARG, "vl_init_eigen: ring_j=%d outside [0,170]", d->ring_j);
    if (d->terms < 0 || d->terms > kEigenMaxTerms)
        return fail(VL_E_ARG, "vl_init_eigen: terms=%d outside [0,%d]", d->terms, kEigenMaxTerms);
    if (!(d->vt > 0) || !std::isfinite(d->k1) || !std::isfinite(d->om_re) || !std::isfinite(d->om_im) ||
        !std::isfinite(d->om_pc) || !std::isfinite(d->delta_n) || (d->perturbation && d->k1 == 0.0))
        return fail(VL_E_ARG, "vl_init_eigen: vt > 0 and finite parameters (k1 != 0) required");
    EigenArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nx = h->nx;
    a.nu = h->nu;
    a.nv = h->nv;
    a.n = h->n;
    a.ring_j = d->ring_j;
    a.terms = d->terms;
    a.perturbation = d->perturbation ? 1 : 0;
    a.pert_only = d->pert_only ? 1 : 0;
    a.vt = d->vt;
    a.k1 = d->k1;
    a.om_re = d->om_re;
    a.om_im = d->om_im;
    a.om_pc = d->om_pc;
    a.delta_n = d->delta_n;
    double jfact = 1.0;
    for (int k = 2; k <= d->ring_j; ++k) jfact *= (double)k;
    // main/grid.py:253 and :213, in the reference's operation order
    a.ring_factor = 1.0 / ((std::pow(2.0, d->ring_j)) * (2.0 * M_PI) * (d->vt * d->vt) * jfact);
    a.f0_factor = 1.0 / (2.0 * M_PI * (d->vt * d->vt) * jfact);
    const int nvel = (h->nu + 2) * h->n * (h->nv + 2) * h->n;
    hipLaunchKernelGGL(k_init_eigen<kEigenMaxTerms>, dim3((nvel + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       (hipStream_t)stream, f, h->d_xnodes, h->d_unodes, h->d_vnodes, a);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_set_background(vl_geom_t h, const double *f0_host) {
    if (int r = check_handle(h)) return r;
    if (!f0_host) return fail(VL_E_NULL, "vl_set_background: null pointer");
    if (h->storage == VL_F64) return fail(VL_E_ARG, "vl_set_background: FP64 storage keeps f itself");
    const size_t cnt = (size_t)(h->nu + 2) * h->n * (h->nv + 2) * h->n;
    for (size_t i = 0; i < cnt; ++i)
        if (!std::isfinite(f0_host[i])) return fail(VL_E_NOTFINITE, "vl_set_background: non-finite f0");
    h->f0_host.assign(f0_host, f0_host + cnt);
    if (int r = upload_background(h)) return r;
    if (!h->d_f0) VL_HIP(hipMalloc((void **)&h->d_f0, cnt * sizeof(double)));
    VL_HIP(hipMemcpy(h->d_f0, f0_host, cnt * sizeof(double), hipMemcpyHostToDevice));
    return VL_OK;
}

// Run-file records (RunData.save_data, main/data_management.py:37-52): the
// state as float32 f in the reference layout, and back.
static int record_check(vl_geom_t h, const void *a, const void *b, const char *what) {
    if (int r = check_handle(h)) return r;
    if (!a || !b) return fail(VL_E_NULL, "%s: null pointer", what);
    if (h->storage != VL_F64 && !h->d_f0) return fail(VL_E_ARG, "%s: reduced storage needs vl_set_background", what);
    return VL_OK;
}

int vl_export_f32(vl_geom_t h, const void *f, float *out, void *stream) {
    if (int r = record_check(h, f, out, "vl_export_f32")) return r;
    const size_t eb = storage_bytes(h->storage);
    if (!aligned(f, eb) || !aligned(out, 4)) return fail(VL_E_ALIGN, "vl_export_f32: misaligned");
    const int64_t plane = (int64_t)(h->nu + 2) * h->n * (h->nv + 2) * h->n;
    const dim3 grid((unsigned)((plane + kBlock - 1) / kBlock), (unsigned)((h->nx + 2) * h->n));
    hipStream_t st = (hipStream_t)stream;
    switch (h->storage) {
    case VL_F64:
        hipLaunchKernelGGL(k_export_f32<0>, grid, dim3(kBlock), 0, st, (const double *)f, h->d_f0, out, plane, 1.0);
        break;
    case VL_F32:
        hipLaunchKernelGGL(k_export_f32<1>, grid, dim3(kBlock), 0, st, (const float *)f, h->d_f0, out, plane,
                           h->inv_scale);
        break;
    case VL_BF16:
        hipLaunchKernelGGL(k_export_f32<2>, grid, dim3(kBlock), 0, st, (const uint16_t *)f, h->d_f0, out, plane,
                           h->inv_scale);
        break;
    default:
        hipLaunchKernelGGL(k_export_f32<3>, grid, dim3(kBlock), 0, st, (const uint16_t *)f, h->d_f0, out, plane,
                           h->inv_scale);
        break;
    }
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_import_f32(vl_geom_t h, const float *in, void *f, void *stream) {
    if (int r = record_check(h, in, f, "vl_import_f32")) return r;
    const size_t eb = storage_bytes(h->storage);
    if (!aligned(f, eb) || !aligned(in, 4)) return fail(VL_E_ALIGN, "vl_import_f32: misaligned");
    const int64_t plane = (int64_t)(h->nu + 2) * h->n * (h->nv + 2) * h->n;
    const dim3 grid((unsigned)((plane + kBlock - 1) / kBlock), (unsigned)((h->nx + 2) * h->n));
    hipStream_t st = (hipStream_t)stream;
    switch (h->storage) {
    case VL_F64:
        hipLaunchKernelGGL(k_import_f32<0>, grid, dim3(kBlock), 0, st, in, h->d_f0, (double *)f, plane, 1.0);
        break;
    case VL_F32:
        hipLaunchKernelGGL(k_import_f32<1>, grid, dim3(kBlock), 0, st, in, h->d_f0, (float *)f, plane, h->scale);
        break;
    case VL_BF16:
        hipLaunchKernelGGL(k_import_f32<2>, grid, dim3(kBlock), 0, st, in, h->d_f0, (uint16_t *)f, plane,
                           h->scale);
        break;
    default:
        hipLaunchKernelGGL(k_import_f32<3>, grid, dim3(kBlock), 0, st, in, h->d_f0, (uint16_t *)f, plane,
                           h->scale);
        break;
    }
    VL_HIP(hipGetLastError());
    return VL_OK;
}

// ---- HIP graphs of library calls ------------------------------------------

}  // extern "C"

struct vl_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

int vl_graph_begin(void *stream) {
    if (!stream) return fail(VL_E_ARG, "vl_graph_begin: capture needs a stream other than the null stream");
    VL_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return VL_OK;
}

int vl_graph_end(void *stream, vl_graph_t *out) {
    if (!out) return fail(VL_E_NULL, "vl_graph_end: null pointer");
    *out = nullptr;
    if (!stream) return fail(VL_E_ARG, "vl_graph_end: null stream");
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(e, "vl_graph_end: hipStreamEndCapture");
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(ei, "vl_graph_end: hipGraphInstantiate");
    }
    vl_graph *r = new (std::nothrow) vl_graph;
    if (!r) {
        (void)hipGraphExecDestroy(x);
        (void)hipGraphDestroy(g);
        return hip_fail(hipErrorOutOfMemory, "vl_graph_end: host allocation");
    }
    r->graph = g;
    r->exec = x;
    *out = r;
    return VL_OK;
}

int vl_graph_launch(vl_graph_t g, void *stream) {
    if (!g) return fail(VL_E_HANDLE, "vl_graph_launch: null graph");
    VL_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return VL_OK;
}

int vl_graph_destroy(vl_graph_t g) {
    if (!g) return VL_OK;
    const hipError_t a = hipGraphExecDestroy(g->exec);
    const hipError_t b = hipGraphDestroy(g->graph);
    delete g;
    if (a != hipSuccess) return hip_fail(a, "vl_graph_destroy: hipGraphExecDestroy");
    if (b != hipSuccess) return hip_fail(b, "vl_graph_destroy: hipGraphDestroy");
    return VL_OK;
}

// ---- element-wise passes over whole states ---------------------------------

static PlaneShape plane_shape(const vl_geom *h) {
    PlaneShape p;
    p.row = (h->nv + 2) * h->n;
    p.plane = (int64_t)(h->nu + 2) * h->n * p.row;
    p.n = h->n;
    p.nu = h->nu;
    p.nv = h->nv;
    return p;
}

int vl_axpby(vl_geom_t h, double *y, const double *x, double a, double b, int interior, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!y || (b != 0.0 && !x)) return fail(VL_E_NULL, "vl_axpby: null pointer");
    if (!aligned(y, 8) || (x && !aligned(x, 8))) return fail(VL_E_ALIGN, "vl_axpby: misaligned");
    if (!std::isfinite(a) || !std::isfinite(b)) return fail(VL_E_NOTFINITE, "vl_axpby: non-finite coefficient");
    const PlaneShape p = plane_shape(h);
    const int planes = (interior ? h->nx : h->nx + 2) * h->n;
    const dim3 grid((unsigned)((p.plane + kBlock - 1) / kBlock), (unsigned)planes);
    hipLaunchKernelGGL(k_axpby, grid, dim3(kBlock), 0, (hipStream_t)stream, y, x, a, b, p, interior ? h->n : 0,
                       interior ? 1 : 0);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_init_separable(vl_geom_t h, const double *vel, const double *xprofile, void *f, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!vel || !xprofile || !f) return fail(VL_E_NULL, "vl_init_separable: null pointer");
    if (!aligned(vel, 8) || !aligned(xprofile, 8) || !aligned(f, storage_bytes(h->storage)))
        return fail(VL_E_ALIGN, "vl_init_separable: misaligned");
    const PlaneShape p = plane_shape(h);
    const dim3 grid((unsigned)((p.plane + kBlock - 1) / kBlock), (unsigned)((h->nx + 2) * h->n));
    hipStream_t st = (hipStream_t)stream;
    switch (h->storage) {
    case VL_F64:
        hipLaunchKernelGGL(k_init_separable<0>, grid, dim3(kBlock), 0, st, (double *)f, vel, xprofile, 1.0,
                           p.plane);
        break;
    case VL_F32:
        hipLaunchKernelGGL(k_init_separable<1>, grid, dim3(kBlock), 0, st, (float *)f, vel, xprofile, h->scale,
                           p.plane);
        break;
    case VL_BF16:
        hipLaunchKernelGGL(k_init_separable<2>, grid, dim3(kBlock), 0, st, (uint16_t *)f, vel, xprofile, h->scale,
                           p.plane);
        break;
    default:
        hipLaunchKernelGGL(k_init_separable<3>, grid, dim3(kBlock), 0, st, (uint16_t *)f, vel, xprofile, h->scale,
                           p.plane);
        break;
    }
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_deviation_to_storage(vl_geom_t h, const double *dev, void *g, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!dev || !g) return fail(VL_E_NULL, "vl_deviation_to_storage: null pointer");
    if (h->storage == VL_F64) return fail(VL_E_ARG, "vl_deviation_to_storage: FP64 storage keeps f itself");
    if (!aligned(dev, 8) || !aligned(g, storage_bytes(h->storage)))
        return fail(VL_E_ALIGN, "vl_deviation_to_storage: misaligned");
    const int64_t count = plane_shape(h).plane * (h->nx + 2) * h->n;
    const dim3 grid((unsigned)((count + kBlock - 1) / kBlock));
    hipStream_t st = (hipStream_t)stream;
    switch (h->storage) {
    case VL_F32:
        hipLaunchKernelGGL(k_deviation_to_storage<1>, grid, dim3(kBlock), 0, st, dev, (float *)g, h->scale, count);
        break;
    case VL_BF16:
        hipLaunchKernelGGL(k_deviation_to_storage<2>, grid, dim3(kBlock), 0, st, dev, (uint16_t *)g, h->scale,
                           count);
        break;
    default:
        hipLaunchKernelGGL(k_deviation_to_storage<3>, grid, dim3(kBlock), 0, st, dev, (uint16_t *)g, h->scale,
                           count);
        break;
    }
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_density_deviation(vl_geom_t h, const double *rho, double *out, int sign, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!rho || !out) return fail(VL_E_NULL, "vl_density_deviation: null pointer");
    if (!aligned(rho, 8) || !aligned(out, 8)) return fail(VL_E_ALIGN, "vl_density_deviation: misaligned");
    if (sign != 1 && sign != -1) return fail(VL_E_ARG, "vl_density_deviation: sign must be 1 or -1");
    hipLaunchKernelGGL(k_density_deviation, dim3(1), dim3(kBlock), 0, (hipStream_t)stream, rho, out, h->nx * h->n,
                       sign);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_stream_probe(const void *a, const void *b, void *out, size_t bytes, int grid, int nontemporal,
                    void *stream) {
    if (!a || !b || !out) return fail(VL_E_NULL, "vl_stream_probe: null pointer");
    if (!aligned(a, 16) || !aligned(b, 16) || !aligned(out, 16) || bytes % 16 != 0)
        return fail(VL_E_ALIGN, "vl_stream_probe: 16-byte aligned arrays of whole 16-byte vectors required");
    if (grid <= 0 || grid > (1 << 20)) return fail(VL_E_ARG, "vl_stream_probe: grid %d outside [1, 2^20]", grid);
    const int64_t nvec = (int64_t)(bytes / 16);
    if (nvec == 0) return VL_OK;
    if (nontemporal)
        hipLaunchKernelGGL(k_stream_probe<true>, dim3((unsigned)grid), dim3(kBlock), 0, (hipStream_t)stream,
                           (const vl_f64x2 *)a, (const vl_f64x2 *)b, (vl_f64x2 *)out, nvec);
    else
        hipLaunchKernelGGL(k_stream_probe<false>, dim3((unsigned)grid), dim3(kBlock), 0, (hipStream_t)stream,
                           (const vl_f64x2 *)a, (const vl_f64x2 *)b, (vl_f64x2 *)out, nvec);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_total_f64(vl_geom_t h, const void *f, double *out, void *stream) {
    if (int r = record_check(h, f, out, "vl_total_f64")) return r;
    if (!aligned(f, storage_bytes(h->storage)) || !aligned(out, 8)) return fail(VL_E_ALIGN, "vl_total_f64: misaligned");
    const PlaneShape p = plane_shape(h);
    const dim3 grid((unsigned)((p.plane + kBlock - 1) / kBlock), (unsigned)((h->nx + 2) * h->n));
    hipStream_t st = (hipStream_t)stream;
    switch (h->storage) {
    case VL_F64:
        hipLaunchKernelGGL(k_total_f64<0>, grid, dim3(kBlock), 0, st, (const double *)f, h->d_f0, out, p.plane, 1.0);
        break;
    case VL_F32:
        hipLaunchKernelGGL(k_total_f64<1>, grid, dim3(kBlock), 0, st, (const float *)f, h->d_f0, out, p.plane,
                           h->inv_scale);
        break;
    case VL_BF16:
        hipLaunchKernelGGL(k_total_f64<2>, grid, dim3(kBlock), 0, st, (const uint16_t *)f, h->d_f0, out, p.plane,
                           h->inv_scale);
        break;
    default:
        hipLaunchKernelGGL(k_total_f64<3>, grid, dim3(kBlock), 0, st, (const uint16_t *)f, h->d_f0, out, p.plane,
                           h->inv_scale);
        break;
    }
    VL_HIP(hipGetLastError());
    return VL_OK;
}

}  // extern "C"

// ---- ensembles of runs ----------------------------------------------------

struct vl_ensemble {
    int count, storage;
    std::vector<vl_geom *> geoms;
    std::vector<EnsRun> runs;      // host copy of the device table
    std::vector<char> bound;
    EnsRun *d_runs;
    // schedule of the stage kernel
    int work;                      // target DOF per work item (sets each run's x-segment)
    int grid;                      // workgroups of the persistent launch (0: one per item)
    int order;                     // 0: longest items first; 1: run by run
    std::vector<int2> items;
    int2 *d_items;
    size_t items_cap;
    std::vector<int2 *> retired;   // earlier item lists (freed by vl_ensemble_destroy)
    // field kernels: workgroup maps
    EnsBlock *d_map_swap, *d_map_red, *d_map_fwd, *d_map_inv;
    int nb_swap, nb_red, nb_fwd, nb_inv;
    size_t fwd_lds, inv_lds;
    bool dirty;                    // host tables changed since the last upload
    int n_adaptive;                // runs with the adaptive time step
};

namespace {

constexpr int kEnsSwapCap = 2048;     // swap workgroups per run (grid-stride beyond)

int ens_check(const vl_ensemble *e) {
    if (!e) return fail(VL_E_HANDLE, "null ensemble handle");
    return 0;
}

// x elements per pencil of run r so that a pencil carries about `work` DOF
int ens_xseg(const vl_geom *h, int work) {
    const int64_t per_x = (int64_t)h->TV * h->n * h->n * h->n;
    int64_t xs = (work + per_x / 2) / per_x;
    xs = std::max<int64_t>(1, std::min<int64_t>(xs, h->nx));
    return (int)xs;
}

// Work list and per-run x-segments for the current schedule.
void ens_build_items(vl_ensemble *e) {
    struct Item {
        int run, pencil;
        int64_t cost;
    };
    std::vector<Item> all;
    for (int r = 0; r < e->count; ++r) {
        const vl_geom *h = e->geoms[r];
        const int xs = ens_xseg(h, e->work);
        e->runs[r].s.xseg = xs;
        const int base = h->nu * h->nchunk;
        const int nseg = (h->nx + xs - 1) / xs;
        const int64_t per_x = (int64_t)h->TV * h->n * h->n * h->n;
        for (int p = 0; p < base * nseg; ++p) {
            const int seg = p / base;
            const int len = std::min(xs, h->nx - seg * xs);
            all.push_back({r, p, len * per_x});
        }
    }
    if (e->order == 0)
        std::stable_sort(all.begin(), all.end(), [](const Item &a, const Item &b) { return a.cost > b.cost; });
    e->items.resize(all.size());
    for (size_t i = 0; i < all.size(); ++i) e->items[i] = make_int2(all[i].run, all[i].pencil);
    e->dirty = true;
}

int ens_upload(vl_ensemble *e) {
    if (!e->dirty) return VL_OK;
    VL_HIP(hipMemcpy(e->d_runs, e->runs.data(), e->runs.size() * sizeof(EnsRun), hipMemcpyHostToDevice));
    if (e->items.size() > e->items_cap) {
        // a graph captured earlier may still launch over the old list: keep
        // it allocated until the ensemble is destroyed
        if (e->d_items) e->retired.push_back(e->d_items);
        e->d_items = nullptr;
        VL_HIP(hipMalloc((void **)&e->d_items, e->items.size() * sizeof(int2)));
        e->items_cap = e->items.size();
    }
    VL_HIP(hipMemcpy(e->d_items, e->items.data(), e->items.size() * sizeof(int2), hipMemcpyHostToDevice));
    e->dirty = false;
    return VL_OK;
}

int ens_ready(vl_ensemble *e, const char *what) {
    for (int r = 0; r < e->count; ++r)
        if (!e->bound[r]) return fail(VL_E_ARG, "%s: run %d has no buffers (vl_ensemble_bind)", what, r);
    return ens_upload(e);
}

template <int S>
hipError_t ens_stage_launch(const vl_ensemble *e, int stage, hipStream_t st) {
    const int n_items = (int)e->items.size();
    const int G = e->grid > 0 ? std::min(e->grid, n_items) : n_items;
    if (stage == 1)
        hipLaunchKernelGGL((k_ens_stage<kModeStage1, S>), dim3(G), dim3(kBlock), 0, st, e->d_runs, e->d_items,
                           n_items, stage);
    else
        hipLaunchKernelGGL((k_ens_stage<kModeStage, S>), dim3(G), dim3(kBlock), 0, st, e->d_runs, e->d_items,
                           n_items, stage);
    return hipGetLastError();
}

hipError_t ens_stage(const vl_ensemble *e, int stage, hipStream_t st) {
    return e->storage == VL_F64 ? ens_stage_launch<VL_F64>(e, stage, st) : ens_stage_launch<VL_F32>(e, stage, st);
}

// moment partials -> rho -> E of every run (energy: one value per run, or
// null; last: the stage-3 solve, where adaptive runs form their next dt)
hipError_t ens_field(const vl_ensemble *e, double *energy, bool last, hipStream_t st) {
    hipLaunchKernelGGL(k_ens_reduce, dim3(e->nb_red), dim3(kBlock), 0, st, e->d_runs, e->d_map_red);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_ens_forward, dim3(e->nb_fwd), dim3(kBlock), e->fwd_lds, st, e->d_runs, e->d_map_fwd);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_ens_inverse, dim3(e->nb_inv), dim3(kBlock), e->inv_lds, st, e->d_runs, e->d_map_inv,
                       energy, last ? 1 : 0);
    return hipGetLastError();
}

hipError_t ens_swap(const vl_ensemble *e, hipStream_t st) {
    if (e->storage == VL_F64)
        hipLaunchKernelGGL(k_ens_swap<double>, dim3(e->nb_swap), dim3(kBlock), 0, st, e->d_runs, e->d_map_swap);
    else
        hipLaunchKernelGGL(k_ens_swap<float>, dim3(e->nb_swap), dim3(kBlock), 0, st, e->d_runs, e->d_map_swap);
    return hipGetLastError();
}

hipError_t ens_step(const vl_ensemble *e, double *energy, hipStream_t st) {
    hipError_t err = ens_swap(e, st);
    for (int stage = 1; stage <= 3 && err == hipSuccess; ++stage) {
        err = ens_field(e, stage == 3 ? energy : nullptr, stage == 3, st);
        if (err == hipSuccess) err = ens_stage(e, stage, st);
    }
    if (err == hipSuccess && e->n_adaptive > 0) {
        hipLaunchKernelGGL(k_ens_clock_advance, dim3(1), dim3(kBlock), 0, st, e->d_runs, e->count);
        err = hipGetLastError();
    }
    return err;
}

// moment partials of every run's f (the first stage of a loop needs them)
int ens_moments(vl_ensemble *e, hipStream_t st) {
    for (int r = 0; r < e->count; ++r) {
        const vl_geom *h = e->geoms[r];
        VL_HIP(launch_moment(h, e->storage, kElemLayout, e->runs[r].f, e->runs[r].s, st));
    }
    return VL_OK;
}

}  // namespace

extern "C" {

int vl_ensemble_create(const vl_geom_t *geoms, int count, vl_ensemble_t *out) {
    if (!geoms || !out) return fail(VL_E_NULL, "vl_ensemble_create: null pointer");
    *out = nullptr;
    if (count < 1) return fail(VL_E_ARG, "vl_ensemble_create: count must be >= 1");
    const int storage = geoms[0] ? geoms[0]->storage : -1;
    if (storage != VL_F64 && storage != VL_F32)
        return fail(VL_E_ARG, "vl_ensemble_create: FP64 or FP32 storage only");
    for (int r = 0; r < count; ++r) {
        const vl_geom *h = geoms[r];
        if (!h) return fail(VL_E_HANDLE, "vl_ensemble_create: null geometry %d", r);
        if (h->storage != storage) return fail(VL_E_ARG, "vl_ensemble_create: runs must share one storage");
        if (h->n < kEnsMinN || h->n > kEnsMaxN)
            return fail(VL_E_ARG, "vl_ensemble_create: run %d has n=%d outside [%d,%d]", r, h->n, kEnsMinN,
                        kEnsMaxN);
        if (storage != VL_F64 && !h->d_bg)
            return fail(VL_E_ARG, "vl_ensemble_create: run %d needs vl_set_background first", r);
    }
    vl_ensemble *e = new (std::nothrow) vl_ensemble();
    if (!e) return fail(VL_E_ARG, "out of host memory");
    e->count = count;
    e->storage = storage;
    e->geoms.assign(geoms, geoms + count);
    e->runs.assign(count, EnsRun{});
    e->bound.assign(count, 0);
    e->work = 9000;
    e->grid = 0;
    e->order = 0;
    std::vector<EnsBlock> mswap, mred, mfwd, minv;
    size_t fwd_lds = 8, inv_lds = 16;
    for (int r = 0; r < count; ++r) {
        const vl_geom *h = geoms[r];
        EnsRun &x = e->runs[r];
        x.s = make_stage_args(h, kElemLayout);
        x.s.periodic_x = 1;
        x.s.write_partials = 1;
        x.tab = h->d_tab;
        x.unodes = h->d_unodes;
        x.vnodes = h->d_vnodes;
        x.wts = h->d_wts;
        x.kwave = h->d_kwave;
        x.T = h->d_T;
        x.phases = h->d_phases;
        x.fa = make_field_args(h, 1);
        x.plane = h->sxe_elem;
        x.n = h->n;
        x.tpe = tiles_per_ex_of(h, kElemLayout);
        x.inv_blocks = h->inv_blocks;
        x.swap_blocks = (int)std::min<int64_t>((2 * x.plane + kBlock - 1) / kBlock, kEnsSwapCap);
        for (int b = 0; b < x.swap_blocks; ++b) mswap.push_back({r, b});
        for (int b = 0; b < h->nx; ++b) mred.push_back({r, b});
        for (int b = 0; b < (h->M + kDftModes - 1) / kDftModes; ++b) mfwd.push_back({r, b});
        for (int b = 0; b < h->inv_blocks; ++b) minv.push_back({r, b});
        fwd_lds = std::max(fwd_lds, (size_t)h->nx * h->n * sizeof(double));
        inv_lds = std::max(inv_lds, (size_t)2 * h->M * sizeof(double2));
    }
    e->fwd_lds = fwd_lds;
    e->inv_lds = inv_lds;
    e->nb_swap = (int)mswap.size();
    e->nb_red = (int)mred.size();
    e->nb_fwd = (int)mfwd.size();
    e->nb_inv = (int)minv.size();
    hipError_t err = hipMalloc((void **)&e->d_runs, count * sizeof(EnsRun));
    if (err == hipSuccess) err = alloc_copy(&e->d_map_swap, mswap.data(), mswap.size());
    if (err == hipSuccess) err = alloc_copy(&e->d_map_red, mred.data(), mred.size());
    if (err == hipSuccess) err = alloc_copy(&e->d_map_fwd, mfwd.data(), mfwd.size());
    if (err == hipSuccess) err = alloc_copy(&e->d_map_inv, minv.data(), minv.size());
    if (err != hipSuccess) {
        vl_ensemble_destroy(e);
        return hip_fail(err, "vl_ensemble_create");
    }
    ens_build_items(e);
    *out = e;
    return VL_OK;
}

int vl_ensemble_destroy(vl_ensemble_t e) {
    if (!e) return VL_OK;
    void *bufs[] = {e->d_runs, e->d_items, e->d_map_swap, e->d_map_red, e->d_map_fwd, e->d_map_inv};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (int2 *b : e->retired) (void)hipFree(b);
    delete e;
    return VL_OK;
}

int vl_ensemble_bind(vl_ensemble_t e, int run, void *f, void *f1, void *f2, double *E, void *ws, double dt) {
    if (int r = ens_check(e)) return r;
    if (run < 0 || run >= e->count) return fail(VL_E_ARG, "vl_ensemble_bind: run %d outside [0,%d)", run, e->count);
    if (!f || !f1 || !f2 || !E || !ws) return fail(VL_E_NULL, "vl_ensemble_bind: null pointer");
    if (!aligned(f, 16) || !aligned(f1, 16) || !aligned(f2, 16) || !aligned(E, 8) || !aligned(ws, 256))
        return fail(VL_E_ALIGN, "vl_ensemble_bind: misaligned pointer");
    if (!std::isfinite(dt)) return fail(VL_E_NOTFINITE, "vl_ensemble_bind: dt not finite");
    const vl_geom *h = e->geoms[run];
    EnsRun &x = e->runs[run];
    x.f = f;
    x.f1 = f1;
    x.f2 = f2;
    x.E = E;
    x.s.partials = ws_partials(h, ws);
    x.s.seed = ws_seed(h, ws);
    x.rho = ws_rho(h, ws);
    x.coef = ws_coef(h, ws);
    x.enpart = ws_enpart(h, ws);
    x.mxpart = ws_mxpart(h, ws);
    x.ticket = ws_ticket(h, ws);
    x.seed = ws_seed(h, ws);
    x.clock = ws_clock(h, ws);
    x.status = ws_status(h, ws);
    x.c2[0] = 1.0;
    x.c2[1] = 1.0 / 4.0;
    x.c2[2] = 2.0 / 3.0;
    // the Shu-Osher coefficients as vl_rk3_step forms them (main/timestep.py:37-38)
    const double cf[3][3] = {{0.0, 1.0, dt}, {3.0 / 4.0, 1.0 / 4.0, (1.0 / 4.0) * dt},
                             {1.0 / 3.0, 2.0 / 3.0, (2.0 / 3.0) * dt}};
    std::memcpy(x.cf, cf, sizeof(cf));
    e->bound[run] = 1;
    e->dirty = true;
    return VL_OK;
}

int vl_ensemble_set_adaptive(vl_ensemble_t e, int run, int on) {
    if (int r = ens_check(e)) return r;
    if (run < 0 || run >= e->count) return fail(VL_E_ARG, "vl_ensemble_set_adaptive: run %d outside [0,%d)", run, e->count);
    if (!e->bound[run]) return fail(VL_E_ARG, "vl_ensemble_set_adaptive: run %d has no buffers (vl_ensemble_bind)", run);
    const vl_geom *h = e->geoms[run];
    EnsRun &x = e->runs[run];
    if (on && !h->cfl_set) return fail(VL_E_ARG, "vl_ensemble_set_adaptive: run %d needs vl_set_cfl first", run);
    if (on) set_cfl_args(h, x.fa);
    x.fa.dt_next = nullptr;                 // set per launch for the stage-3 solve
    const int was = x.adaptive;
    x.adaptive = on ? 1 : 0;
    e->n_adaptive += x.adaptive - was;
    e->dirty = true;
    return VL_OK;
}

int vl_ensemble_clocks(vl_ensemble_t e, double *out, void *stream) {
    if (int r = ens_check(e)) return r;
    if (!out) return fail(VL_E_NULL, "vl_ensemble_clocks: null pointer");
    if (!aligned(out, 8)) return fail(VL_E_ALIGN, "vl_ensemble_clocks: misaligned pointer");
    if (int r = ens_ready(e, "vl_ensemble_clocks")) return r;
    hipLaunchKernelGGL(k_ens_clocks, dim3(1), dim3(kBlock), 0, (hipStream_t)stream, e->d_runs, e->count, out);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_ensemble_set_schedule(vl_ensemble_t e, int work_per_item, int grid, int order) {
    if (int r = ens_check(e)) return r;
    if (work_per_item < 1) return fail(VL_E_ARG, "vl_ensemble_set_schedule: work_per_item must be >= 1");
    if (grid < 0) return fail(VL_E_ARG, "vl_ensemble_set_schedule: grid must be >= 0");
    if (order != 0 && order != 1) return fail(VL_E_ARG, "vl_ensemble_set_schedule: order must be 0 or 1");
    e->work = work_per_item;
    e->grid = grid;
    e->order = order;
    ens_build_items(e);
    return VL_OK;
}

int vl_ensemble_get_schedule(vl_ensemble_t e, int *work_per_item, int *grid, int *order, int *items) {
    if (int r = ens_check(e)) return r;
    if (work_per_item) *work_per_item = e->work;
    if (grid) *grid = e->grid;
    if (order) *order = e->order;
    if (items) *items = (int)e->items.size();
    return VL_OK;
}

int vl_ensemble_moments(vl_ensemble_t e, void *stream) {
    if (int r = ens_check(e)) return r;
    if (int r = ens_ready(e, "vl_ensemble_moments")) return r;
    return ens_moments(e, (hipStream_t)stream);
}

int vl_ensemble_rk3_step(vl_ensemble_t e, double *energy_out, void *stream) {
    if (int r = ens_check(e)) return r;
    if (energy_out && !aligned(energy_out, 8)) return fail(VL_E_ALIGN, "vl_ensemble_rk3_step: misaligned energy");
    if (int r = ens_ready(e, "vl_ensemble_rk3_step")) return r;
    VL_HIP(ens_step(e, energy_out, (hipStream_t)stream));
    return VL_OK;
}

int vl_ensemble_autotune(vl_ensemble_t e, int reps, float *table, int max_rows, int *rows, void *stream) {
    if (int r = ens_check(e)) return r;
    if (reps < 1) return fail(VL_E_ARG, "vl_ensemble_autotune: reps must be >= 1");
    if (int r = ens_ready(e, "vl_ensemble_autotune")) return r;
    hipStream_t st = (hipStream_t)stream;
    int cus = 256;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
    }
    const int works[] = {9000, 18000, 36000, 72000, 144000, 288000};
    const int grids[] = {0, 3 * cus, 6 * cus, 12 * cus};
    hipEvent_t t0, t1;
    VL_HIP(hipEventCreate(&t0));
    VL_HIP(hipEventCreate(&t1));
    float best = 1e30f;
    int bw = e->work, bg = e->grid, bo = e->order, nrow = 0;
    int rc = VL_OK;
    for (int o = 0; o < 2 && rc == VL_OK; ++o)
        for (int w : works)
            for (int g : grids) {
                e->work = w;
                e->grid = g;
                e->order = o;
                ens_build_items(e);
                if ((rc = ens_upload(e)) != VL_OK) break;
                // stage 2 reads f1 and f and writes f2 (and the moment
                // partials); f is not touched
                hipError_t err = ens_stage(e, 2, st);
                if (err == hipSuccess) err = hipEventRecord(t0, st);
                for (int k = 0; k < reps && err == hipSuccess; ++k) err = ens_stage(e, 2, st);
                if (err == hipSuccess) err = hipEventRecord(t1, st);
                if (err == hipSuccess) err = hipEventSynchronize(t1);
                float ms = 0.f;
                if (err == hipSuccess) err = hipEventElapsedTime(&ms, t0, t1);
                if (err != hipSuccess) {
                    rc = hip_fail(err, "vl_ensemble_autotune");
                    break;
                }
                ms /= reps;
                if (table && nrow < max_rows) {
                    table[4 * nrow + 0] = (float)w;
                    table[4 * nrow + 1] = (float)g;
                    table[4 * nrow + 2] = (float)o;
                    table[4 * nrow + 3] = ms;
                }
                ++nrow;
                if (ms < best) {
                    best = ms;
                    bw = w;
                    bg = g;
                    bo = o;
                }
            }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (rows) *rows = nrow;
    e->work = bw;
    e->grid = bg;
    e->order = bo;
    ens_build_items(e);
    if (rc != VL_OK) return rc;
    if (int r = ens_upload(e)) return r;
    // the timed stages overwrote the moment partials: those of f again
    return ens_moments(e, st);
}

}  // extern "C"
