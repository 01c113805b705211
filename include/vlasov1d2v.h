/*
 * vlasov1d2v.h -- C ABI of the MI355X (gfx950) 1D2V nodal-DG Vlasov-Poisson
 * SSP-RK3 path.
 *
 * Every entry point takes plain pointers and sizes.  Device pointers are HIP
 * device addresses; `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  Every function returns int: 0 = ok, negative = validation
 * error (see VL_E_*), positive = a hipError_t.  vl_last_error_string()
 * describes the most recent failure on the calling thread.
 *
 * State layout (the reference's, main/grid.py:268): f is a C-order array of
 * shape (NX+2, n, NU+2, n, NV+2, n) -- element index before node index on each
 * axis, one ghost element on each side of every axis, v-line contiguous.
 * E is (NX+2, n) with periodic ghosts; rho is (NX, n).
 *
 * Only the set-up calls allocate device memory (vl_geom_create,
 * vl_set_background, vl_set_central_flux and the vl_ensemble_* set-up; freed
 * by the matching destroy).  Scratch is caller-provided through `workspace`
 * (size from vl_workspace_bytes), or for vl_rk3_step_noalloc a library-owned
 * array.  A handle may be used from one
 * host thread at a time; calls on one stream are ordered; no call
 * synchronises the device, so sequences can be captured in a hipGraph.
 */
#ifndef VLASOV1D2V_H
#define VLASOV1D2V_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VL_OK = 0,
    VL_E_ARG = -1,        /* bad scalar argument (n, dims, dtype, ...)      */
    VL_E_NULL = -2,       /* null pointer where one is required             */
    VL_E_ALIGN = -3,      /* device pointer not aligned to its element size */
    VL_E_HANDLE = -4,     /* invalid handle                                 */
    VL_E_WORKSPACE = -5,  /* workspace too small                            */
    VL_E_NOTFINITE = -6   /* non-finite value detected (host-side checks)   */
};

/* Storage type of the distribution function. */
enum {
    VL_F64 = 0,
    VL_F32 = 1,
    VL_BF16 = 2,
    VL_F16 = 3
};

/* Geometry and operator tables of one run.  All pointers are HOST pointers to
 * FP64 data; vl_geom_create copies them to device memory.
 *
 * Replaces the setup objects Basis1D (main/basis.py:87-228: up, xi, weights),
 * Grid1D/Grid3D (main/grid.py:10-103: arr, J, wave_numbers, grid_phases,
 * spectral_transform), DGFlux.__init__ (main/fluxes.py:14-57:
 * flux_coefficients) and Elliptic (main/elliptic.py:10-22,147-148:
 * poisson_coefficient, magnetic_field). */
typedef struct vl_geom_desc {
    int32_t n;                 /* nodes per axis, 2..10                          */
    int32_t nx, nu, nv;        /* interior elements per axis (>= 1)              */
    int32_t n_modes;           /* M = number of Fourier modes (main/grid.py:46)  */
    int32_t storage;           /* VL_F64 / VL_F32 / VL_BF16 / VL_F16             */
    double jx, ju, jv;         /* element Jacobians 2/dx (main/grid.py:23)       */
    double acc;                /* flux_coefficients (main/fluxes.py:53)          */
    double b_field;            /* magnetic_field (main/elliptic.py:147)          */
    double poisson_coef;       /* poisson_coefficient (main/elliptic.py:22)      */
    double x_dx, x_length;     /* dx and length of the x grid (main/grid.py:82)  */
    const double *up;          /* n*n, up[i*n+j]  (main/basis.py:176-188)        */
    const double *xi;          /* n*2, xi[i*2+k]  (main/basis.py:110)            */
    const double *weights;     /* n Lobatto weights (main/basis.py:31-57)        */
    const double *x_nodes;     /* (nx+2)*n x coordinates (main/grid.py:55-70)   */
    const double *u_nodes;     /* (nu+2)*n                                       */
    const double *v_nodes;     /* (nv+2)*n                                       */
    const double *wave_numbers;/* M  (main/grid.py:48)                           */
    const double *transform;   /* nx*n*M complex, interleaved re,im
                                  (Grid1D.spectral_transform, main/basis.py:204)  */
    const double *phases;      /* M*nx*n complex, interleaved re,im
                                  (Grid1D.grid_phases, main/grid.py:50)           */
} vl_geom_desc;

typedef struct vl_geom *vl_geom_t;

/* Create / destroy a geometry handle (setup only; allocates the device copies
 * of the operator tables). */
int vl_geom_create(const vl_geom_desc *desc, vl_geom_t *out);
int vl_geom_destroy(vl_geom_t h);

/* Update the physics coefficients of a handle (host-side, no device work):
 * flux_coefficients (DGFlux, main/fluxes.py:53), magnetic_field
 * (Elliptic.set_magnetic_field, main/elliptic.py:147-148) and
 * poisson_coefficient (Elliptic.__init__, main/elliptic.py:22). */
int vl_set_physics(vl_geom_t h, double acc, double b_field, double poisson_coef);

/* Bytes of device workspace the calls below need for this geometry.  The
 * workspace holds moment partials, the density, DFT coefficients and a
 * completion ticket of the field-energy reduction. */
int vl_workspace_bytes(vl_geom_t h, size_t *bytes);

/* Prepare a freshly allocated workspace (zeroes its ticket word with a small
 * kernel, so it is safe inside stream capture).  Call once per workspace
 * before its first use; every entry point leaves the ticket at zero. */
int vl_workspace_init(vl_geom_t h, void *workspace, void *stream);

/* Launch-shape tuning of the stage kernel (no counterpart in the reference).
 * VL_TUNE_XSEG: x elements each workgroup walks (default: chosen so that the
 * launch has about 13000 workgroups, about 5200 for the row-split kernel;
 * a set value applies to both).
 * VL_TUNE_MFMA: 1 (default) runs the BF16/FP16 stages of the element-major
 * layout on the matrix cores, 0 on the vector ALUs.
 * VL_TUNE_ROW_SPLIT: 1 (default) makes vl_resident_layout report the
 * row-split layout for BF16/FP16 storage with n = 8 or 9 (while
 * VL_TUNE_MFMA is 1); 0 keeps the element-major layout.
 * VL_TUNE_ROUNDING: rounding of the 16-bit stored values by the row-split
 * stage kernel: 1 (default) stochastic -- unbiased, so that stage updates
 * smaller than one storage ulp are not lost (round-to-nearest damps the
 * field at small dt); the random bits are a hash of a per-stage seed kept in
 * the workspace and the value's index, so runs are reproducible; 0
 * round-to-nearest-even (as the other 16-bit kernels).
 * VL_TUNE_RS_KERNEL: the row-split stage kernel at n = 9: 1 the 9 x 9-frame
 * kernel (transposes on the matrix core), 2 the natural-frame kernel (LDS
 * transpose reads, per-direction products, dense epilogue). */
enum { VL_TUNE_XSEG = 1, VL_TUNE_MFMA = 2, VL_TUNE_ROW_SPLIT = 3, VL_TUNE_ROUNDING = 4, VL_TUNE_RS_KERNEL = 5 };
int vl_set_tuning(vl_geom_t h, int key, int value);
int vl_get_tuning(vl_geom_t h, int key, int *value);

/* Number of elements of one state array: (nx+2)*n*(nu+2)*n*(nv+2)*n. */
int vl_state_elems(vl_geom_t h, size_t *elems);

/* Layouts of a state array (no counterpart in the reference, which has only
 * the first): VL_LAYOUT_REFERENCE is the C-order layout above;
 * VL_LAYOUT_ELEMENT_MAJOR keeps the n^3 nodes of each element contiguous
 * (elements ordered (ex, eu, ev)) with a pitch of round_up(n^3, 8) values
 * per element, so that a tile of elements is one contiguous run.  Inside an
 * element node (ix, iu, iv) sits at (iu n + iv) n + ix for VL_F64 / VL_F32
 * states (x fastest) and at (ix n + iu) n + iv for VL_BF16 / VL_F16 states.
 * It is the resident layout of Stepper.main_loop. */
/* VL_LAYOUT_ROW_SPLIT (n = 8 or 9 only, states under 2^31 bytes): elements ordered and padded as in
 * the element-major layout; inside an element the first eight nodes of each
 * v-line (ix, iu) form one 16-byte row (rows in (ix, iu) order), and for
 * n = 9 the ninth nodes of all lines follow in (ix, iu) order.  It is the
 * resident layout of the matrix-core stage kernel of the 16-bit forms. */
enum { VL_LAYOUT_REFERENCE = 0, VL_LAYOUT_ELEMENT_MAJOR = 1, VL_LAYOUT_ROW_SPLIT = 2 };
int vl_state_elems_layout(vl_geom_t h, int layout, size_t *elems);

/* The layout vl_rk3_step runs fastest in for this handle's storage and n:
 * VL_LAYOUT_ROW_SPLIT or VL_LAYOUT_ELEMENT_MAJOR (no counterpart in the
 * reference, which keeps one layout). */
int vl_resident_layout(vl_geom_t h, int *layout);

/* dst = src in the other layout (out of place; element-major padding values
 * are not written). */
int vl_convert_layout(vl_geom_t h, const void *src, int src_layout, void *dst, int dst_layout,
                      void *stream);

/* out = semi-discrete RHS of f (FP64), interior entries only; ghost entries of
 * `out` are not written.  f's ghost elements are used as given.  E is the
 * ghosted field (nx+2)*n.
 * Replaces DGFlux.semi_discrete_rhs (main/fluxes.py:59-207). */
int vl_rhs(vl_geom_t h, const double *f, const double *E, double *out, void *stream);

/* rho[ex,ix] = sum_{eu,iu,ev,iv} f[interior] * w_iu w_iv / (Ju Jv), FP64.
 * Replaces Distribution.moment_zero / moment_zero_of_arr (main/grid.py:303-321). */
int vl_moment0(vl_geom_t h, const double *f, double *rho, void *workspace, void *stream);

/* Spectral Poisson solve with the given charge density (nx*n):
 * E (nx+2)*n with periodic ghosts, phi (nx*n, may be NULL).  The workspace
 * holds the DFT coefficients.
 * Replaces Elliptic.poisson2 (main/elliptic.py:117-145) with
 * Grid1D.fourier_basis / sum_fourier (main/grid.py:78-88). */
int vl_poisson(vl_geom_t h, const double *charge_density, double *E, double *phi,
               void *workspace, void *stream);

/* energy[0] = 0.5 * sum_interior w_ix E^2.
 * Replaces Elliptic.electric_energy (main/elliptic.py:150-151). */
int vl_field_energy(vl_geom_t h, const double *E, double *energy, void *stream);

/* out[0] = max |acc * E| over the ghosted field (the device part of
 * get_max_speeds, main/timestep.py:235-242). */
int vl_max_abs_acc_e(vl_geom_t h, const double *E, double *out, void *stream);

/* Periodic x-ghost copy of a state array, in place (storage dtype):
 * f[0] = f[nx], f[nx+1] = f[1].
 * Replaces Distribution.swap_ghost_cells_of_array (main/grid.py:187-189). */
int vl_swap_ghosts(vl_geom_t h, void *f, void *stream);

/* One full Shu-Osher SSP-RK3 step of Stepper.nonlinear_ssp_rk
 * (main/timestep.py:141-187) with a fixed dt > 0 (VL_E_ARG otherwise):
 *   x-ghost swap of f; rho(f) -> E; f1 = f + dt L(f);
 *   rho(f1) -> E; f2 = 3/4 f + 1/4 f1 + 1/4 dt L(f1);
 *   rho(f2) -> E; f = 1/3 f + 2/3 f2 + 2/3 dt L(f2).
 * f, f1, f2 are state arrays of the handle's storage dtype; f1/f2 need only
 * their interior (ghosts are never read: x-neighbours are taken
 * periodically, u/v ghost neighbours from f, whose u/v ghost planes stay
 * frozen as in the reference).  On return E (nx+2)*n holds the field solved
 * from f2 and phi (nx*n, may be NULL) its potential -- the reference's
 * elliptic state after a step -- and energy_out[0] (device, may be NULL)
 * = 0.5 sum w E^2 of that field (the value main/timestep.py:102 appends).
 * flags: VL_STEP_RHO_VALID -- the workspace already holds the moment partials
 * of f, left by the previous vl_rk3_step on the same f (or vl_moment_partials),
 * and nothing has written the workspace or f since (skips one read of f; the
 * caller vouches for it: a stale workspace gives a wrong first-stage field).
 * VL_STEP_ELEMENT_MAJOR -- f, f1, f2 are in the element-major layout (16-byte
 * aligned).  VL_STEP_ROW_SPLIT -- f, f1, f2 are in the row-split layout
 * (16-byte aligned; BF16/FP16 storage, n = 8 or 9).
 * For VL_F32/VL_BF16/VL_F16 storage, f holds the deviation from the
 * x-independent background `f0` (nu+2)*n*(nv+2)*n FP64 table set with
 * vl_set_background; the moment, Poisson solve and energy stay FP64. */
enum { VL_STEP_RHO_VALID = 1, VL_STEP_ELEMENT_MAJOR = 2, VL_STEP_ROW_SPLIT = 4 };
int vl_rk3_step(vl_geom_t h, void *f, void *f1, void *f2, double *E, double *phi,
                double dt, double *energy_out, void *workspace, int flags, void *stream);

/* Adaptive time step on the device (main/timestep.py:88-89,184,231-242:
 * dt = courant / (max|u|/dx + (max|acc E| + |acc B max|v||)/du
 *                 + |acc |B max|u|||/dv), recomputed from the field of the
 * last stage solve after every step, and time += the new dt), so that
 * adaptive runs need no host synchronisation and can be graph-captured.
 * max|u|, max|v| are taken over the ghosted nodes given to vl_geom_create;
 * acc and B are the handle's (vl_set_physics).
 * vl_set_cfl: the Courant number and the cell sizes dx, du, dv.
 * vl_clock_init: the workspace clock gets time `time` and the dt of the
 *   field E ((nx+2)*n, e.g. the set-up solve's) -- the loop-entry adapt.
 * vl_rk3_step_adaptive: vl_rk3_step with the clock's dt; afterwards the
 *   clock holds the adapted dt and the advanced time, and clock_out (device,
 *   may be NULL) receives (time, dt) -- the values main/timestep.py:96-98
 *   records after the step. */
int vl_set_cfl(vl_geom_t h, double courant, double dx, double du, double dv);
int vl_clock_init(vl_geom_t h, const double *E, double time, void *workspace, void *stream);
int vl_rk3_step_adaptive(vl_geom_t h, void *f, void *f1, void *f2, double *E, double *phi, double *energy_out,
                         double *clock_out, void *workspace, int flags, void *stream);

/* Allocation-free SSP-RK3 step.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  vl_rk3_step with the
 * workspace taken from a library-owned device array of
 * VL_NOALLOC_SCRATCH_BYTES per GPU (resolved when the first geometry is
 * created on that GPU, held until the process ends) instead of the caller's;
 * the caller passes f, f1, f2, E, phi and energy_out as for vl_rk3_step.  It
 * calls no hipMalloc / hipFree.  Geometries whose vl_workspace_bytes exceeds
 * the array are refused with VL_E_WORKSPACE (the 116.6M-DOF n = 9,
 * 64 x 50 x 50 grid needs 24 MB).  Every call on one GPU shares that GPU's
 * array: calls that overlap on a GPU (two streams, two threads) race, and
 * VL_STEP_RHO_VALID is valid only when the previous call on the GPU was a
 * step of the same f.  The reduction ticket is cleared by a small kernel at
 * every call; the stochastic-rounding seed of the 16-bit forms carries over
 * between calls. */
#define VL_NOALLOC_SCRATCH_BYTES (32u << 20)
int vl_rk3_step_noalloc(vl_geom_t h, void *f, void *f1, void *f2, double *E, double *phi, double dt,
                        double *energy_out, int flags, void *stream);

/* Run-time failure detection (the reference has none: its NaN check is
 * commented out, main/timestep.py:118-121).  Every step's stage-3 field solve
 * (vl_rk3_step, vl_rk3_step_adaptive, vl_ensemble_rk3_step per run) reduces
 * the field energy and max|acc E| of its field and sets VL_STATUS_NONFINITE in
 * the workspace's status word when either is not finite -- the state has
 * overflowed (an unstable time step) or holds NaN.  The bit stays set until
 * cleared.  vl_read_status copies the word to *status (host), clears it when
 * `clear` is nonzero, and synchronises `stream`: call it between graph
 * replays, not inside a capture. */
enum { VL_STATUS_NONFINITE = 1 };
int vl_read_status(vl_geom_t h, void *workspace, int clear, unsigned *status, void *stream);

/* The moment partials of f (storage dtype; layout by VL_STEP_ELEMENT_MAJOR /
 * VL_STEP_ROW_SPLIT in flags, else the reference layout) into the workspace:
 * the first pass of vl_rk3_step without VL_STEP_RHO_VALID.  A caller that
 * replays captured steps whose first step was captured with
 * VL_STEP_RHO_VALID runs it before a replay whenever the workspace or f was
 * written by anything else since the previous step (the first half of
 * Distribution.moment_zero_of_arr, main/grid.py:319-321, for the stepper's
 * own state). */
int vl_moment_partials(vl_geom_t h, const void *f, void *workspace, int flags, void *stream);

/* Optional timing hook (no counterpart in the reference, whose only timing
 * is a printed wall clock, main/timestep.py:117,139): when `events` is
 * non-NULL, every later vl_rk3_step on this handle records events[2k] just
 * before and events[2k+1] just after its k-th stage kernel (k = 0, 1, 2) on
 * the call's stream.  `events` points to 6 hipEvent_t created by the caller
 * with timing enabled; the array is read at each call, so the caller may
 * change its contents between calls.  NULL disables the hook. */
int vl_set_stage_events(vl_geom_t h, void **events);

/* Background table for the reduced-precision storage forms:
 * f_total = f0(u,v) + stored deviation.  f0 is a host FP64 array of
 * (nu+2)*n*(nv+2)*n values.  Only used when storage != VL_F64. */
int vl_set_background(vl_geom_t h, const double *f0_host);

/* Run-file records.  vl_export_f32 writes the state f (the handle's storage
 * form, reference layout, ghosts included) as float32 values of f itself
 * (f0 + g / scale for the reduced forms) into `out` (same shape): the "pdf"
 * record that RunData.save_data stores (main/data_management.py:37-52,
 * `f['pdf'][-1] = distribution`, dtype 'f').  vl_import_f32 is the inverse
 * for resuming from a record: f = in (FP64), or g = (in - f0) * scale rounded
 * to nearest (reduced forms).  The reduced forms need vl_set_background
 * first.  No counterpart for the import: the reference has no resume. */
int vl_export_f32(vl_geom_t h, const void *f, float *out, void *stream);
int vl_import_f32(vl_geom_t h, const float *in, void *f, void *stream);

/* Element-wise passes over whole states (reference layout, ghosts included),
 * so that tensors stay buffers on every product path.
 *
 * vl_axpby: y = a y + b x on every node of an FP64 state of the handle's shape,
 * or (interior != 0) only on the nodes of interior elements, leaving the
 * ghost planes of y as they are.  a == 0 does not read y, b == 0 does not
 * read x (x may then be NULL).  The combinations of the linear SSP-RK update,
 * main/timestep.py:189-229 (`arr[ng] = c0 arr[ng]`, `copy[ng] += dt rhs[ng]`,
 * `arr[ng] += c_i copy[ng]`), one fused multiply-add per node; with a = 0,
 * b = 1 and interior = 0 the copy of main/grid.py:198-199.
 *
 * vl_init_separable: f[X, (u, v)] = xprofile[X] vel[(u, v)] for the FP64 form
 * (the ring of main/grid.py:244-279 when xprofile is one and vel the ring);
 * the reduced forms store g = (xprofile[X] vel[(u, v)]) scale, the deviation
 * f - f0 of a state whose deviation is separable (xprofile = modulation - 1).
 * xprofile has (NX+2) n values, vel (NU+2) n (NV+2) n; both on the device.
 * Rounded to nearest even (16-bit forms through FP32).
 *
 * vl_deviation_to_storage (reduced forms): g = dev scale rounded to the
 * storage type, dev = f - f0 in FP64.
 *
 * vl_total_f64: out = f (FP64 form) or f0 + g / scale (reduced forms, which
 * need vl_set_background first), FP64, the whole ghosted state.
 *
 * vl_density_deviation: out = mean(rho) - rho (sign = 1), the charge density
 * the time steps hand to the Poisson solve (main/timestep.py:155,172,203,221,
 * `cp.mean(density) - density`), or rho - mean(rho) (sign = -1, the initial
 * solve of main/main.py:131); rho and out are (NX, n) FP64, mean the plain
 * mean over all NX n values (summed in a fixed order).
 *
 * No counterparts in the reference beyond the cited statements: it forms
 * these with CuPy array arithmetic. */
int vl_axpby(vl_geom_t h, double *y, const double *x, double a, double b, int interior, void *stream);
int vl_init_separable(vl_geom_t h, const double *vel, const double *xprofile, void *f, void *stream);
int vl_deviation_to_storage(vl_geom_t h, const double *dev, void *g, void *stream);
int vl_total_f64(vl_geom_t h, const void *f, double *out, void *stream);
int vl_density_deviation(vl_geom_t h, const double *rho, double *out, int sign, void *stream);

/* Measurement probe, no counterpart in the reference: the stage kernels' HBM
 * access shape alone -- out = a + b over `bytes` bytes per array (FP64 pairs,
 * 16-byte vectors; non-temporal loads and stores when `nontemporal`, else the
 * default cache policy), `grid` workgroups of 256 threads in a grid-stride
 * loop.  Its best rate (3 bytes moved per array byte) is the practical ceiling
 * the stage kernels are quoted against. */
int vl_stream_probe(const void *a, const void *b, void *out, size_t bytes, int grid, int nontemporal,
                    void *stream);

/* ---- Ensembles (no counterpart in the reference, which steps one run per
 * process through main/main.py:165 -> Stepper.main_loop).  Many independent
 * runs -- each its own geometry (n, grid, physics), state and workspace --
 * step together: every RK stage of every run is one launch per kernel kind
 * (ghost swap, moment reduction, forward DFT, inverse DFT, stage), so one
 * ensemble step is 13 launches whatever the number of runs.  Each run
 * computes what vl_rk3_step with VL_STEP_ELEMENT_MAJOR computes for it
 * (main/timestep.py:141-187 per run), the same statements up to the
 * compiler's multiply-add fusion (rounding level); the potential phi is not
 * formed.  Storage FP64 or FP32 (one for all
 * runs), n = 6..9.  The geometries must outlive the ensemble. */
/* ---- HIP graphs of library calls (no counterpart in the reference, which
 * launches every CuPy kernel from Python each step).  vl_graph_begin starts a
 * thread-local capture of `stream`; every library call made on that stream
 * until vl_graph_end is recorded instead of run (calls that synchronise or
 * allocate fail while capturing: use fixed-dt or device-clock steps, whose
 * launches neither allocate nor synchronise).  vl_graph_end ends the capture
 * (always, also when it then reports an error) and instantiates the graph;
 * vl_graph_launch replays it on a stream; vl_graph_destroy frees it.  The
 * recorded calls keep the pointers they were given: buffers must outlive the
 * graph and stay where they are. */
typedef struct vl_graph *vl_graph_t;
int vl_graph_begin(void *stream);
int vl_graph_end(void *stream, vl_graph_t *graph);
int vl_graph_launch(vl_graph_t graph, void *stream);
int vl_graph_destroy(vl_graph_t graph);

typedef struct vl_ensemble *vl_ensemble_t;
int vl_ensemble_create(const vl_geom_t *geoms, int count, vl_ensemble_t *out);
int vl_ensemble_destroy(vl_ensemble_t e);

/* Buffers and time step of run `run`: f, f1, f2 in the element-major layout
 * (16-byte aligned, vl_state_elems_layout values each), E (nx+2)*n, the run's
 * workspace (vl_workspace_bytes, initialised), and its fixed dt. */
int vl_ensemble_bind(vl_ensemble_t e, int run, void *f, void *f1, void *f2, double *E, void *workspace,
                     double dt);

/* Schedule of the ensemble stage kernel: about `work_per_item` DOF per work
 * item (one pencil of one run; sets each run's x-segment length), `grid`
 * workgroups of the persistent launch (0: one workgroup per item), item
 * `order` 0 = longest first (a run's pencils stay consecutive), 1 = run by
 * run.  get: also the number of work items. */
int vl_ensemble_set_schedule(vl_ensemble_t e, int work_per_item, int grid, int order);
int vl_ensemble_get_schedule(vl_ensemble_t e, int *work_per_item, int *grid, int *order, int *items);

/* Adaptive time step for one run of an ensemble (main/timestep.py:184,
 * 231-242, as vl_rk3_step_adaptive does for a single run): the run's dt and
 * time live in the clock of the workspace it was bound with; the stage-3
 * field solve forms its next dt and one kernel after the step makes it the
 * step's dt and advances the time.  Needs vl_set_cfl on the run's geometry
 * and vl_clock_init (with the set-up field) on its workspace; the dt given to
 * vl_ensemble_bind is then ignored.  on = 0 returns the run to its fixed dt. */
int vl_ensemble_set_adaptive(vl_ensemble_t e, int run, int on);

/* (time, dt) of every run's clock into out[2 i], out[2 i + 1] (device, FP64;
 * zeros for fixed-dt runs), on `stream`. */
int vl_ensemble_clocks(vl_ensemble_t e, double *out, void *stream);

/* Moment partials of every run's f: call once before the first step of a
 * loop (a step leaves those of its result for the next one). */
int vl_ensemble_moments(vl_ensemble_t e, void *stream);

/* One SSP-RK3 step of every run.  energy_out (device, may be NULL): count
 * values, run r's field energy of its stage-3 field (as vl_rk3_step).  The
 * first call after vl_ensemble_bind / set_schedule uploads the run table
 * (synchronous): make it outside a graph capture; later calls only launch.
 * A graph captured from these launches holds the run table and schedule of
 * its capture: after vl_ensemble_bind, vl_ensemble_set_schedule,
 * vl_ensemble_autotune or vl_ensemble_set_adaptive capture it again (earlier
 * work-item lists stay allocated until vl_ensemble_destroy, so a stale
 * replay reads valid memory, but it runs the old schedule). */
int vl_ensemble_rk3_step(vl_ensemble_t e, double *energy_out, void *stream);

/* Autotuner: times the stage-2 launch (reps times each) over x-segment work
 * targets {9000 .. 288000 DOF, factors of 2}, persistent grids {one per item, 3, 6, 12 per
 * CU} and both item orders; keeps the fastest schedule, then restores the
 * moment partials of f.  table (may be NULL): up to max_rows rows of
 * (work, grid, order, ms); *rows = rows measured.  Synchronises the stream. */
int vl_ensemble_autotune(vl_ensemble_t e, int reps, float *table, int max_rows, int *rows, void *stream);

/* Stabilized central-flux DG Poisson solve (Elliptic.poisson,
 * main/elliptic.py:86-115), the reference's alternative to poisson2.  The
 * operator is assembled and inverted on the host as the reference does
 * (build_central_flux_operator / invert, main/elliptic.py:24-84) and
 * uploaded once: inv_op ((nx n + 1)^2, row-major, gauge row/column
 * included), grad_op (the gradient operator as an (nx n) x (nx n) row-major
 * matrix indexed [(i, j), (e, k)]) and the n x n mass matrix.  Allocates its
 * device tables and scratch on the first call.
 * vl_poisson_central: E (nx+2)*n with periodic ghosts and phi nx*n (may be
 * NULL) from the charge density cd (nx*n); anti_alias = 1 re-sums phi and E
 * through the nodal Fourier transform (main/elliptic.py:101-113). */
int vl_set_central_flux(vl_geom_t h, const double *inv_op, const double *grad_op, const double *mass);
int vl_poisson_central(vl_geom_t h, const double *charge_density, double *E, double *phi, int anti_alias,
                       void *stream);

/* The reference driver's initial condition (Distribution.initialize_gpu,
 * main/grid.py:244-279 with eigenfunction, :207-242): f = ring + (when
 * `perturbation`) the Bernstein-mode eigenfunction
 *   2 delta_n Re(exp(i k1 x) (-i) exp(i beta sin phi) df_dv Upsilon / k1^2),
 *   Upsilon = sum_{n=-terms..terms} n / (n - om) J_n(beta) exp(-i n phi),
 *   beta = -k1 r om_pc,
 * on every ghosted node of f (reference layout, FP64, vl_state_elems values),
 * computed on the device (Bessel orders by backward recurrence).  om is
 * complex (om_re + i om_im).  pert_only = 1 writes the perturbation alone
 * (the stored deviation of the reduced-precision forms). */
typedef struct vl_eigen_desc {
    int32_t ring_j, terms, perturbation, pert_only;
    double vt, k1, om_re, om_im, om_pc, delta_n;
} vl_eigen_desc;
int vl_init_eigen(vl_geom_t h, const vl_eigen_desc *desc, double *f, void *stream);

/* Description of the last failure on this thread (static storage). */
const char *vl_last_error_string(void);

#ifdef __cplusplus
}
#endif

#endif /* VLASOV1D2V_H */
