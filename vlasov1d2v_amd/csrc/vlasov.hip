// C ABI of the gfx950 1D2V nodal-DG Vlasov-Poisson SSP-RK3 path.
// Declarations and the reference interface each entry point replaces:
// include/vlasov1d2v.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "vlasov1d2v.h"
#include "vl_common.h"
#include "vl_field.h"
#include "vl_stage_f64.h"

using namespace vl;

struct vl_geom {
    int n, nx, nu, nv, M, storage;
    double jx, ju, jv, acc, bfield, pcoef, dx, length;
    OpTables tab;
    int TV, nchunk, tiles_per_ex, tiles;
    int64_t sx, sxe, su, sue;
    size_t state_elems;
    int device;
    double *d_unodes, *d_vnodes, *d_wts, *d_kwave;
    double2 *d_T, *d_phases;
    double *d_f0;
    void **stage_events;   // optional: 6 hipEvent_t recorded around the stage kernels
    // workspace layout (byte offsets)
    size_t off_partials, off_rho, off_cd, off_coef, ws_bytes;
};

namespace {

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

StageArgs make_stage_args(const vl_geom *h) {
    StageArgs s;
    std::memset(&s, 0, sizeof(s));
    s.nx = h->nx;
    s.nu = h->nu;
    s.nv = h->nv;
    s.nchunk = h->nchunk;
    s.jx = h->jx;
    s.ju = h->ju;
    s.jv = h->jv;
    s.acc = h->acc;
    s.bfield = h->bfield;
    s.sx = h->sx;
    s.sxe = h->sxe;
    s.su = h->su;
    s.sue = h->sue;
    return s;
}

FieldArgs make_field_args(const vl_geom *h, int from_rho) {
    FieldArgs a;
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

// ---- launch helpers (template dispatch over n) --------------------------

template <int N, int MODE>
hipError_t launch_stage_n(const vl_geom *h, const double *fin, const double *fgh, const double *fn,
                          double *fout, const double *E, const StageArgs &s, hipStream_t st) {
    hipLaunchKernelGGL((k_stage_f64<N, MODE>), dim3(h->tiles), dim3(kBlock), 0, st, fin, fgh, fn, fout,
                       E, h->d_unodes, h->d_vnodes, s, h->tab);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_stage(const vl_geom *h, const double *fin, const double *fgh, const double *fn,
                        double *fout, const double *E, const StageArgs &s, hipStream_t st) {
    switch (h->n) {
        case 2: return launch_stage_n<2, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 3: return launch_stage_n<3, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 4: return launch_stage_n<4, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 5: return launch_stage_n<5, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 6: return launch_stage_n<6, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 7: return launch_stage_n<7, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 8: return launch_stage_n<8, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 9: return launch_stage_n<9, MODE>(h, fin, fgh, fn, fout, E, s, st);
        case 10: return launch_stage_n<10, MODE>(h, fin, fgh, fn, fout, E, s, st);
        default: return hipErrorInvalidValue;
    }
}

template <int N>
hipError_t launch_moment_n(const vl_geom *h, const double *f, const StageArgs &s, hipStream_t st) {
    hipLaunchKernelGGL((k_moment_partials_f64<N>), dim3(h->tiles), dim3(kBlock), 0, st, f, s, h->tab);
    return hipGetLastError();
}

hipError_t launch_moment(const vl_geom *h, const double *f, const StageArgs &s, hipStream_t st) {
    switch (h->n) {
        case 2: return launch_moment_n<2>(h, f, s, st);
        case 3: return launch_moment_n<3>(h, f, s, st);
        case 4: return launch_moment_n<4>(h, f, s, st);
        case 5: return launch_moment_n<5>(h, f, s, st);
        case 6: return launch_moment_n<6>(h, f, s, st);
        case 7: return launch_moment_n<7>(h, f, s, st);
        case 8: return launch_moment_n<8>(h, f, s, st);
        case 9: return launch_moment_n<9>(h, f, s, st);
        case 10: return launch_moment_n<10>(h, f, s, st);
        default: return hipErrorInvalidValue;
    }
}

double2 *ws_partials(const vl_geom *h, void *ws) {
    return reinterpret_cast<double2 *>(static_cast<char *>(ws) + h->off_partials);
}
double *ws_rho(const vl_geom *h, void *ws) {
    return reinterpret_cast<double *>(static_cast<char *>(ws) + h->off_rho);
}
double *ws_cd(const vl_geom *h, void *ws) {
    return reinterpret_cast<double *>(static_cast<char *>(ws) + h->off_cd);
}
double2 *ws_coef(const vl_geom *h, void *ws) {
    return reinterpret_cast<double2 *>(static_cast<char *>(ws) + h->off_coef);
}

// partials -> rho (workspace) -> E (+phi, energy)
hipError_t field_from_partials(const vl_geom *h, void *ws, double *E, double *phi, double *energy,
                               hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(h->nx), dim3(kBlock), 0, st, ws_partials(h, ws), h->n,
                       h->tiles_per_ex, ws_rho(h, ws));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const FieldArgs a = make_field_args(h, 1);
    hipLaunchKernelGGL(k_field, dim3(1), dim3(kFieldBlock), 0, st, ws_rho(h, ws), h->d_T, h->d_phases,
                       h->d_kwave, h->d_wts, ws_cd(h, ws), ws_coef(h, ws), E, phi, energy,
                       (double *)nullptr, a);
    return hipGetLastError();
}

hipError_t record_stage_event(const vl_geom *h, int idx, hipStream_t st) {
    if (!h->stage_events || !h->stage_events[idx]) return hipSuccess;
    return hipEventRecord((hipEvent_t)h->stage_events[idx], st);
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
    if (d->storage != VL_F64) return fail(VL_E_ARG, "storage %d not available in this build", d->storage);
    if (!d->up || !d->xi || !d->weights || !d->u_nodes || !d->v_nodes || !d->wave_numbers ||
        !d->transform || !d->phases)
        return fail(VL_E_NULL, "vl_geom_create: null table");
    if (!(d->x_dx > 0) || !(d->x_length > 0)) return fail(VL_E_ARG, "x_dx/x_length must be > 0");

    vl_geom *h = new (std::nothrow) vl_geom;
    if (!h) return fail(VL_E_ARG, "out of host memory");
    std::memset(h, 0, sizeof(*h));
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
    // moment weights exactly as main/grid.py:297-301 forms them:
    // (w_iu * w_iv) / (Ju * Jv)
    for (int iu = 0; iu < n; ++iu)
        for (int iv = 0; iv < n; ++iv)
            h->tab.wq[iu * n + iv] = (d->weights[iu] * d->weights[iv]) / (d->ju * d->jv);
    h->TV = tile_tv(n);
    h->nchunk = (d->nv + h->TV - 1) / h->TV;
    h->tiles_per_ex = d->nu * h->nchunk;
    const int64_t tiles = (int64_t)d->nx * h->tiles_per_ex;
    if (tiles > 2147483647LL) {
        delete h;
        return fail(VL_E_ARG, "grid too large");
    }
    h->tiles = (int)tiles;
    h->su = (int64_t)(d->nv + 2) * n;
    h->sue = h->su * n;
    h->sx = h->sue * (d->nu + 2);
    h->sxe = h->sx * n;
    h->state_elems = (size_t)h->sxe * (d->nx + 2);

    size_t off = 0;
    h->off_partials = off;
    off = align_up(off + (size_t)tiles * n * sizeof(double2), 256);
    h->off_rho = off;
    off = align_up(off + (size_t)d->nx * n * sizeof(double), 256);
    h->off_cd = off;
    off = align_up(off + (size_t)d->nx * n * sizeof(double), 256);
    h->off_coef = off;
    off = align_up(off + (size_t)2 * h->M * sizeof(double2), 256);
    h->ws_bytes = off;

    hipError_t e = hipGetDevice(&h->device);
    const int K = d->nx * n;
    if (e == hipSuccess) e = alloc_copy(&h->d_unodes, d->u_nodes, (size_t)(d->nu + 2) * n);
    if (e == hipSuccess) e = alloc_copy(&h->d_vnodes, d->v_nodes, (size_t)(d->nv + 2) * n);
    if (e == hipSuccess) e = alloc_copy(&h->d_wts, d->weights, (size_t)n);
    if (e == hipSuccess) e = alloc_copy(&h->d_kwave, d->wave_numbers, (size_t)h->M);
    if (e == hipSuccess)
        e = alloc_copy(&h->d_T, reinterpret_cast<const double2 *>(d->transform), (size_t)K * h->M);
    if (e == hipSuccess)
        e = alloc_copy(&h->d_phases, reinterpret_cast<const double2 *>(d->phases), (size_t)K * h->M);
    if (e != hipSuccess) {
        vl_geom_destroy(h);
        return hip_fail(e, "vl_geom_create");
    }
    *out = h;
    return VL_OK;
}

int vl_geom_destroy(vl_geom_t h) {
    if (!h) return VL_OK;
    void *bufs[] = {h->d_unodes, h->d_vnodes, h->d_wts, h->d_kwave, h->d_T, h->d_phases, h->d_f0};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete h;
    return VL_OK;
}

int vl_set_physics(vl_geom_t h, double acc, double b_field, double poisson_coef) {
    if (int r = check_handle(h)) return r;
    if (!std::isfinite(acc) || !std::isfinite(b_field) || !std::isfinite(poisson_coef))
        return fail(VL_E_NOTFINITE, "vl_set_physics: non-finite coefficient");
    h->acc = acc;
    h->bfield = b_field;
    h->pcoef = poisson_coef;
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
    VL_HIP(launch_stage<kModeRhs>(h, f, f, nullptr, out, E, s, (hipStream_t)stream));
    return VL_OK;
}

int vl_moment0(vl_geom_t h, const double *f, double *rho, void *ws, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!f || !rho || !ws) return fail(VL_E_NULL, "vl_moment0: null pointer");
    if (!aligned(f, 8) || !aligned(rho, 8) || !aligned(ws, 256)) return fail(VL_E_ALIGN, "vl_moment0: misaligned");
    hipStream_t st = (hipStream_t)stream;
    StageArgs s = make_stage_args(h);
    s.partials = ws_partials(h, ws);
    VL_HIP(launch_moment(h, f, s, st));
    hipLaunchKernelGGL(k_reduce_partials, dim3(h->nx), dim3(kBlock), 0, st, ws_partials(h, ws), h->n,
                       h->tiles_per_ex, rho);
    VL_HIP(hipGetLastError());
    return VL_OK;
}

int vl_poisson(vl_geom_t h, const double *cd, double *E, double *phi, void *ws, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!cd || !E || !ws) return fail(VL_E_NULL, "vl_poisson: null pointer");
    if (!aligned(cd, 8) || !aligned(E, 8) || !aligned(ws, 256) || (phi && !aligned(phi, 8)))
        return fail(VL_E_ALIGN, "vl_poisson: misaligned");
    const FieldArgs a = make_field_args(h, 0);
    hipLaunchKernelGGL(k_field, dim3(1), dim3(kFieldBlock), 0, (hipStream_t)stream, cd, h->d_T,
                       h->d_phases, h->d_kwave, h->d_wts, ws_cd(h, ws), ws_coef(h, ws), E, phi,
                       (double *)nullptr, (double *)nullptr, a);
    VL_HIP(hipGetLastError());
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

int vl_swap_ghosts(vl_geom_t h, void *f, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!f) return fail(VL_E_NULL, "vl_swap_ghosts: null pointer");
    const size_t eb = storage_bytes(h->storage);
    if (!aligned(f, eb)) return fail(VL_E_ALIGN, "vl_swap_ghosts: misaligned");
    const int64_t plane = h->sxe;
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

int vl_rk3_step(vl_geom_t h, void *fv, void *f1v, void *f2v, double *E, double *phi, double dt,
                double *energy_out, void *ws, int flags, void *stream) {
    if (int r = check_handle(h)) return r;
    if (!fv || !f1v || !f2v || !E || !ws) return fail(VL_E_NULL, "vl_rk3_step: null pointer");
    if (!aligned(fv, 8) || !aligned(f1v, 8) || !aligned(f2v, 8) || !aligned(E, 8) || !aligned(ws, 256))
        return fail(VL_E_ALIGN, "vl_rk3_step: misaligned pointer");
    if (!std::isfinite(dt)) return fail(VL_E_NOTFINITE, "vl_rk3_step: dt not finite");
    if (h->storage != VL_F64) return fail(VL_E_ARG, "vl_rk3_step: storage not available");
    double *f = (double *)fv, *f1 = (double *)f1v, *f2 = (double *)f2v;
    hipStream_t st = (hipStream_t)stream;
    if (int r = vl_swap_ghosts(h, f, stream)) return r;
    StageArgs s = make_stage_args(h);
    s.periodic_x = 1;
    s.write_partials = 1;
    s.partials = ws_partials(h, ws);
    if (!(flags & VL_STEP_RHO_VALID)) VL_HIP(launch_moment(h, f, s, st));
    // stage 1: f1 = f + dt L(f)                       (main/timestep.py:153-163)
    VL_HIP(field_from_partials(h, ws, E, nullptr, nullptr, st));
    s.c0 = 0.0;
    s.c1 = 1.0;
    s.c2dt = dt;
    VL_HIP(record_stage_event(h, 0, st));
    VL_HIP(launch_stage<kModeStage1>(h, f, f, nullptr, f1, E, s, st));
    VL_HIP(record_stage_event(h, 1, st));
    // stage 2: f2 = 3/4 f + 1/4 f1 + 1/4 dt L(f1)     (main/timestep.py:165-182)
    VL_HIP(field_from_partials(h, ws, E, nullptr, nullptr, st));
    s.c0 = 3.0 / 4.0;
    s.c1 = 1.0 / 4.0;
    s.c2dt = (1.0 / 4.0) * dt;
    VL_HIP(record_stage_event(h, 2, st));
    VL_HIP(launch_stage<kModeStage>(h, f1, f, f, f2, E, s, st));
    VL_HIP(record_stage_event(h, 3, st));
    // stage 3: f = 1/3 f + 2/3 f2 + 2/3 dt L(f2); E, phi, energy of f2's field
    VL_HIP(field_from_partials(h, ws, E, phi, energy_out, st));
    s.c0 = 1.0 / 3.0;
    s.c1 = 2.0 / 3.0;
    s.c2dt = (2.0 / 3.0) * dt;
    VL_HIP(record_stage_event(h, 4, st));
    VL_HIP(launch_stage<kModeStage>(h, f2, f, f, f, E, s, st));
    VL_HIP(record_stage_event(h, 5, st));
    return VL_OK;
}

int vl_set_stage_events(vl_geom_t h, void **events) {
    if (int r = check_handle(h)) return r;
    h->stage_events = events;
    return VL_OK;
}

int vl_set_background(vl_geom_t h, const double *f0_host) {
    if (int r = check_handle(h)) return r;
    if (!f0_host) return fail(VL_E_NULL, "vl_set_background: null pointer");
    return fail(VL_E_ARG, "vl_set_background: storage not available in this build");
}

}  // extern "C"
