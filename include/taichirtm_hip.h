/*
 * taichirtm_hip.h — C ABI of the MI355X (gfx950) elastic FD / RTM hot path.
 *
 * Drop-in boundary for mkt-kuno/TaichiRTM's data-parallel path: the Taichi
 * kernels of src/forward_modeling.py and src/backward_modeling.py and the time
 * loops that sequence them.  The reference exposes that path only as Python
 * classes over Taichi JIT kernels; the Python package `taichirtm_amd` keeps
 * those classes (same names, kwargs, return codes) and binds this ABI with
 * ctypes.  Citations are file:line in the reference's src/.
 *
 * Conventions
 *  - Grid fields are (nx, nz) float32, C order, z (j) contiguous — the
 *    reference's ti.field layout (forward_modeling.py:158-190).
 *  - All pointers named in a descriptor are DEVICE pointers unless the field
 *    name ends in _host.  The library never allocates caller-visible memory;
 *    scratch comes from the caller after trtm_workspace_bytes().
 *  - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *    asynchronous on that stream unless stated.
 *  - Return values: TRTM_OK (0), or a negative TRTM_ERR_* with a message from
 *    trtm_last_error() (thread-local).  Blow-up is not an error return: it is
 *    reported through check_flags with the reference's codes 1-3 (forward) and
 *    4-6 (backward) (forward_modeling.py:522-529, backward_modeling.py:533-537).
 *  - Thread safety: distinct descriptors may be used concurrently on distinct
 *    streams.  The one piece of global mutable state is the static device
 *    scratch of trtm_forward_steps_noalloc, shared by every call of that entry
 *    point on one device.
 */
#ifndef TAICHIRTM_HIP_H
#define TAICHIRTM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TRTM_ABI_VERSION 7

#define TRTM_OK 0
#define TRTM_ERR_ARG (-1)         /* invalid argument or shape */
#define TRTM_ERR_HIP (-2)         /* HIP runtime error */
#define TRTM_ERR_UNSUPPORTED (-3) /* request outside what this build supports */

/* wave systems (systems field): the P-SV (u, w, sxx, szz, sxz) and SH (v, syx,
 * syz) systems are decoupled (forward_modeling.py:330-347,367-384) */
#define TRTM_PSV 1
#define TRTM_SH 2

/* step kernel selection (kernel field) */
#define TRTM_KERNEL_AUTO 0
#define TRTM_KERNEL_SIMPLE 1 /* one launch per reference kernel group */
#define TRTM_KERNEL_FUSED 2  /* one LDS-tiled launch per time step */
#define TRTM_KERNEL_BLOCKED 3 /* one launch per block_steps time steps (temporal blocking) */
#define TRTM_KERNEL_WAVE 4    /* temporal blocking, one wavefront per tile, state in registers */

/*
 * Compiled shapes of a temporally blocked kernel family (TRTM_KERNEL_BLOCKED or
 * TRTM_KERNEL_WAVE) for the given wave systems: writes up to max_n rows of
 * (tile_x, tile_z, threads per workgroup, block_steps) into out and returns
 * the number of shapes (no counterpart in the reference: the autotuner's
 * search space).
 */
int trtm_block_configs(int32_t systems, int32_t kernel, int32_t *out, int32_t max_n);

/* ABI version compiled into the library (== TRTM_ABI_VERSION). */
int trtm_abi_version(void);

/* Message for the last failing call on this thread ("" if none). */
const char *trtm_last_error(void);

/*
 * Averaged moduli, densities and reciprocals.
 * Replaces ForwardModeling._compute_shear_avg / _compute_rho_avg /
 * _compute_inv_rho (forward_modeling.py:243-307; identical in
 * backward_modeling.py:246-310).  myx, myz, mxz, rho_u, rho_w are written on
 * the interior and zeroed on the border; inv_* on the full grid (1/0 = +inf on
 * the averaged borders, as in the reference).  Expression order follows the
 * reference operand for operand, so float32 overflow of mu^4 in mxz (for mu
 * above ~4.3e9) is reproduced.
 */
int trtm_prepare_medium(int32_t nx, int32_t nz, const float *mu, const float *rho,
                        float *myx, float *myz, float *mxz, float *rho_u, float *rho_w,
                        float *inv_rho, float *inv_rho_u, float *inv_rho_w, void *stream);

/*
 * Sponge taper table, host side: table_host[k] = expf((-a*a*k)*k) in float32
 * for k = 0..fw (fw+1 entries).  The step kernels evaluate the Cerjan sponge
 * of forward_modeling.py:203-241 from the cell index and this table.
 */
int trtm_sponge_table(int32_t fw, float a, float *table_host);

/*
 * Full (nx, nz) sponge coefficient array, for the reference's absorb_coeff
 * field (forward_modeling.py:203-241): left strip, then right, then bottom,
 * later writes winning; requires 0 <= fw <= min(nx, nz).  table: device copy
 * of trtm_sponge_table's output.
 */
int trtm_sponge_fill(int32_t nx, int32_t nz, int32_t fw, const float *table, float *coeff,
                     void *stream);

/*
 * Injection plan, host side.  The reference adds every source (or receiver)
 * entry with a parallel `+=` (forward_modeling.py:432-443,
 * backward_modeling.py:455-460), i.e. atomically.  This library groups entries
 * by cell so that duplicates are summed in entry order, deterministically.
 * loc_host: (n, 2) int32 (i, j).  Outputs: order_host[n] (entry indices sorted
 * by cell, stable), group_start_host[n+1]; *n_groups receives the number of
 * distinct cells.  Fails if a location is outside [0,nx) x [0,nz).
 */
int trtm_plan_injection(int32_t nx, int32_t nz, int32_t n, const int32_t *loc_host,
                        int32_t *order_host, int32_t *group_start_host, int32_t *n_groups);

/* One propagator: medium, geometry, state and I/O of a forward or backward run. */
typedef struct trtm_step_desc {
    int32_t struct_size; /* = sizeof(trtm_step_desc) */
    int32_t nx, nz, nt;
    float dt;         /* float32(1/fs) (forward_modeling.py:135,320) */
    float inv_dx;     /* float32(1/dx) */
    float inv_dz;     /* float32(1/dz) */
    float src_factor; /* float32(dt*dx*dz) (forward_modeling.py:430) */
    int32_t fw;       /* absorbing_frame */
    int32_t systems;  /* TRTM_PSV | TRTM_SH: systems that can become nonzero */
    int32_t kernel;   /* TRTM_KERNEL_* */
    int32_t tile_x;   /* blocked kernel: tile rows (0 = library choice) */
    int32_t block_steps; /* blocked kernel: time steps per launch (0 = library choice) */
    int32_t snap_base;   /* global snapshot slot held at index 0 of snap_* (usually 0) */
    const float *sponge_table; /* fw+1 floats (trtm_sponge_table) */
    /* medium (trtm_prepare_medium outputs, plus the caller's lam and mu) */
    const float *lam, *mu, *mxz, *myx, *myz, *inv_rho, *inv_rho_u, *inv_rho_w;
    const float *surface_matrix; /* NULL: flat free surface */
    /* state, updated in place */
    float *u, *v, *w, *sxx, *szz, *sxz, *syx, *syz;
    /* injection: forward = sources with wavelets (scaled by src_factor/rho);
     * backward = receivers with observed data (raw).  (n_inj, nt) each,
     * reference layout; a component pointer may be NULL when its system is
     * inactive. */
    int32_t n_inj, n_inj_groups;
    const int32_t *inj_loc;         /* (n_inj, 2) */
    const int32_t *inj_order;       /* trtm_plan_injection order */
    const int32_t *inj_group_start; /* trtm_plan_injection group starts */
    const float *inj_u, *inj_v, *inj_w;
    /* recording: forward = receivers (seismogram); backward = sources
     * (synthetic source, backward_modeling.py:462-470).  Time-major (nt, n_rec);
     * NULL to skip. */
    int32_t n_rec;
    int32_t tile_z;      /* blocked kernel: tile columns, 64 or 128 (0 = library choice) */
    const int32_t *rec_loc; /* (n_rec, 2) */
    float *rec_u, *rec_v, *rec_w;
    /* finite check every check_interval steps (it % check_interval == 0);
     * check_flags has nt/check_interval + 1 ints, zeroed by the caller; entry
     * it/check_interval receives the highest code found at step it. */
    int32_t check_interval;
    int32_t step_offset; /* batch calls: this shot runs steps [it_begin, it_end) + step_offset */
    int32_t *check_flags;
    /* snapshots: forward writes global slot s = it/isnap-1 at it % isnap == 0,
     * it != 0, into index s - snap_base when that lies in [0, num_snaps);
     * backward correlates index s - snap_base into the image when
     * isnaps_host[s - snap_base] == t.  Snapshot-major (num_snaps, nx, nz). */
    int32_t isnap, num_snaps;
    float *snap_u, *snap_v, *snap_w;
    const int32_t *isnaps_host; /* backward only: HOST array of num_snaps step indices */
    float *image_u, *image_v, *image_w;
    /* scratch of trtm_workspace_bytes() bytes (fused kernel), may be NULL otherwise */
    void *workspace;
    int64_t workspace_bytes;
    /* snapshots in another layout (blocked kernel only): slot stride in floats
     * (0: nx*nz), and snap_sponge = 1 when each stored value still lacks the
     * cell's sponge factor, which the backward image then applies (snapshots
     * taken from an FWI history: velocities before the sponge) */
    int64_t snap_stride;
    int32_t snap_sponge, reserved_snap;
} trtm_step_desc;

/* Scratch bytes trtm_forward_steps / trtm_backward_steps need for desc. */
int64_t trtm_workspace_bytes(const trtm_step_desc *desc);

/*
 * Shape the blocked kernel uses for desc (its systems, tile_x, tile_z and
 * block_steps requests): out[0] tile rows, out[1] tile columns, out[2]
 * threads per workgroup, out[3] time steps per launch.
 */
int trtm_block_config(const trtm_step_desc *desc, int32_t *out);

/* Scratch bytes of the batch-level buffer for trtm_*_steps_batch with n shots. */
int64_t trtm_batch_workspace_bytes(int32_t n);

/*
 * Forward steps it_begin <= it < it_end.  Replaces the body of
 * ForwardModeling.run (forward_modeling.py:539-572): per step, boundary
 * condition, velocity update, source injection (wavelet column it), stress
 * update, sponge, receiver record (row it of rec_*), finite check when
 * it % check_interval == 0, and, when save != 0, the snapshot at
 * it % isnap == 0, it != 0.
 */
int trtm_forward_steps(const trtm_step_desc *desc, int32_t it_begin, int32_t it_end, int32_t save,
                       void *stream);

/*
 * Backward steps it_begin <= it < it_end, t = nt-1-it.  Replaces the body of
 * BackwardModeling.run_calc (backward_modeling.py:539-576): boundary
 * condition, backward velocity update, observed-data injection (column t),
 * backward stress update, sponge, synthetic-source record (row t of rec_*),
 * zero-lag image accumulation image += snap[s] * state when t > 0,
 * t % isnap == 0, s = t/isnap - 1 < num_snaps and isnaps_host[s] == t, and the
 * finite check (codes 4-6) when it % check_interval == 0.
 */
int trtm_backward_steps(const trtm_step_desc *desc, int32_t it_begin, int32_t it_end,
                        void *stream);

/*
 * trtm_forward_steps without a caller workspace and without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace (second state buffer, per-tile lists, shot
 * table) is a static device array of TRTM_NOALLOC_SCRATCH_BYTES that each
 * device holds until the process ends; desc->workspace must be NULL and
 * desc->workspace_bytes 0.  A kernel whose workspace exceeds the scratch is
 * rejected with TRTM_ERR_ARG (a 64 MiB scratch covers the blocked kernel up to
 * about 1.8 M cells, e.g. 1200 x 1500); TRTM_KERNEL_AUTO falls back to the
 * fused, then the simple kernel.  Same results as trtm_forward_steps.  Every
 * call on a device shares that device's scratch: calls that overlap on one
 * device (two streams, two threads) race and must be serialised by the caller.
 */
#define TRTM_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)
int trtm_forward_steps_noalloc(const trtm_step_desc *desc, int32_t it_begin, int32_t it_end,
                               int32_t save, void *stream);

/*
 * Many shots on one medium in one launch per time block (blocked kernel):
 * descs[0..n) describe one shot each (own state, injection, records,
 * snapshots, image, check flags, workspace, record length nt) and share nx,
 * nz, fw, the float32 constants, systems and every medium pointer.  Shot k
 * runs steps it_begin <= it < min(it_end, descs[k].nt).  batch_workspace holds
 * trtm_batch_workspace_bytes(n) bytes.  Same per-shot semantics as
 * trtm_forward_steps / trtm_backward_steps.  No reference counterpart: the
 * reference runs one shot per ForwardModeling / BackwardModeling
 * (rtm.py:935-1000).
 */
int trtm_forward_steps_batch(const trtm_step_desc *descs, int32_t n, void *batch_workspace,
                             int64_t batch_workspace_bytes, int32_t it_begin, int32_t it_end,
                             int32_t save, void *stream);
int trtm_backward_steps_batch(const trtm_step_desc *descs, int32_t n, void *batch_workspace,
                              int64_t batch_workspace_bytes, int32_t it_begin, int32_t it_end,
                              void *stream);

/*
 * Captured time stepping: the launches that trtm_forward_steps_batch
 * (forward != 0, with save) or trtm_backward_steps_batch (forward == 0) would
 * issue for [it_begin, it_end), recorded once into a HIP graph.
 * trtm_plan_create validates the call, builds the per-tile lists and shot
 * tables in the workspaces on `stream` (and waits for it), records the
 * launches on a stream of its own and instantiates the graph; it runs no
 * step.  Each trtm_plan_launch replays the steps on `stream`, from whatever
 * state the buffers then hold, with the same results as the batch call.  The
 * plan keeps every pointer of the descriptors: the caller keeps those buffers
 * alive and leaves both workspaces untouched (they hold the lists and tables)
 * until trtm_plan_destroy, which is called once the plan's launches have
 * finished (after a synchronisation of their stream); buffer contents (state,
 * medium, wavelets, data, image, check flags) may change between launches.  Blocked and wave kernels
 * (kernel TRTM_KERNEL_AUTO, _BLOCKED or _WAVE).  No reference counterpart: it
 * replaces the per-step kernel calls of the loops in ForwardModeling.run
 * (forward_modeling.py:539-572) and BackwardModeling.run_calc
 * (backward_modeling.py:539-576) with one graph launch per range.
 */
typedef struct trtm_plan trtm_plan;
int trtm_plan_create(const trtm_step_desc *descs, int32_t n, void *batch_workspace,
                     int64_t batch_workspace_bytes, int32_t it_begin, int32_t it_end,
                     int32_t forward, int32_t save, void *stream, trtm_plan **plan);
int trtm_plan_launch(trtm_plan *plan, void *stream);
int trtm_plan_destroy(trtm_plan *plan);

/* ------------------------------------------------------------------------
 * FWI gradient: exact discrete adjoint of the forward step.
 * No reference counterpart (the reference's backward propagator is not the
 * adjoint of its forward step: it swaps the difference directions and injects
 * data unscaled, backward_modeling.py:331-345,458-460).  REAL is float
 * (real_bytes = 4) or double (real_bytes = 8, for gradient checks).
 * ------------------------------------------------------------------------ */

/* Medium preparation of trtm_prepare_medium in double precision. */
int trtm_prepare_medium_f64(int32_t nx, int32_t nz, const double *mu, const double *rho,
                            double *myx, double *myz, double *mxz, double *rho_u, double *rho_w,
                            double *inv_rho, double *inv_rho_u, double *inv_rho_w, void *stream);

typedef struct trtm_fwi_desc {
    int32_t struct_size; /* = sizeof(trtm_fwi_desc) */
    int32_t real_bytes;  /* 4 or 8: element size of every REAL array below */
    int32_t nx, nz, nt, fw;
    int32_t systems;     /* TRTM_PSV | TRTM_SH */
    int32_t n_src, n_src_groups, n_rec;
    double dt, inv_dx, inv_dz, src_factor; /* narrowed to REAL by the kernels */
    const void *sponge_table;              /* REAL[fw+1] */
    const void *lam, *mu, *mxz, *myx, *myz, *inv_rho, *inv_rho_u, *inv_rho_w; /* REAL (nx, nz) */
    void *u, *v, *w, *sxx, *szz, *sxz, *syx, *syz;          /* forward state */
    void *pu, *pv, *pw, *psxx, *pszz, *psxz, *psyx, *psyz;  /* adjoint state */
    const int32_t *src_loc, *src_order, *src_group_start;   /* trtm_plan_injection plan */
    const void *wav_u, *wav_v, *wav_w;                      /* REAL (n_src, nt) */
    const int32_t *rec_loc, *rec_order, *rec_group_start;   /* receivers and their plan */
    int32_t n_rec_groups, hist_base;
    void *rec_u, *rec_v, *rec_w;       /* forward records, REAL time-major (nt, n_rec) */
    const void *res_u, *res_v, *res_w; /* data adjoint dJ/d(record), REAL (nt, n_rec) */
    /* history of steps [hist_base, hist_base + hist_len), in one of two layouts.
     * hist_every = 0 (full): for each step the pre-step stresses and the
     * post-update velocities, REAL (hist_len, n_hist_fields, nx, nz),
     * n_hist_fields = 5 (P-SV) + 3 (SH) for the active systems in the order
     * sxx, szz, sxz, u, w, syx, syz, v; read by the one-step-per-launch adjoint.
     * hist_every = K > 0 (compact, blocked kernels only): the post-update
     * velocities of every step, REAL (hist_len, n_vel, nx, nz), n_vel = 2 (P-SV:
     * u, w) + 1 (SH: v), followed by the pre-step stresses of the steps
     * hist_base + m K, REAL (ceil(hist_len / K), n_str, nx, nz), n_str = 3 (P-SV:
     * sxx, szz, sxz) + 2 (SH: syx, syz); read by the blocked adjoint, whose
     * block_steps B must give min(B, 4) = K (it recomputes the other steps'
     * stresses). */
    int32_t hist_len, hist_every;
    void *hist;
    /* gradient accumulators in coefficient space, REAL (nx, nz), += */
    void *g_lam, *g_mu, *g_mxz, *g_myx, *g_myz, *g_ir, *g_iru, *g_irw;
    /* workspace int32 (2, nx, nz): source group / receiver group of each cell
     * (-1: none), filled by trtm_fwi_prepare for this geometry */
    int32_t *cell_map;
    /* step kernels: TRTM_KERNEL_SIMPLE (one step per launch, float or double),
     * TRTM_KERNEL_BLOCKED (float only: block_steps steps per launch, LDS
     * tiles with a halo, needs the workspace), TRTM_KERNEL_AUTO (blocked for
     * float with a workspace, else one step per launch).  tile_x / tile_z /
     * block_steps pick a compiled blocked shape (0 = default; listed by
     * trtm_fwi_block_configs).  Both give the same results bit for bit. */
    int32_t kernel, tile_x, tile_z, block_steps;
    /* blocked kernels: second state buffer of trtm_fwi_workspace_bytes() bytes */
    void *workspace;
    /* trtm_fwi_forward_scattered_batch only (NULL otherwise): this model's
     * coefficients minus those of the reference model whose records are the
     * observed data, per coefficient (lam, mu and the trtm_prepare_medium
     * outputs), REAL (nx, nz); best formed in double and rounded once, since
     * they carry the whole difference between the two models ... */
    const void *dif_lam, *dif_mu, *dif_mxz, *dif_myx, *dif_myz, *dif_inv_rho, *dif_inv_rho_u,
        *dif_inv_rho_w;
    /* ... the scattered state, this model's state minus the reference model's
     * (same field order as the state), updated in place ... */
    void *du, *dv, *dw, *dsxx, *dszz, *dsxz, *dsyx, *dsyz;
    /* ... and its second buffer, trtm_fwi_workspace_bytes() bytes */
    void *scat_workspace;
} trtm_fwi_desc;

/* Workspace of the blocked FWI kernels for d's grid (8 REAL fields). */
int64_t trtm_fwi_workspace_bytes(const trtm_fwi_desc *d);

/*
 * Compiled shapes of the blocked FWI kernels of one pass (0 forward, 1
 * adjoint, 2 scattered forward) that fit these systems: writes up to max_n rows
 * of (tile_x, tile_z, threads, block_steps) to out, returns the number.
 */
int trtm_fwi_block_configs(int32_t systems, int32_t pass, int32_t *out, int32_t max_n);

/*
 * Fills cell_map from the source and receiver plans (src_*, rec_*): the
 * forward and adjoint steps read it to inject and record per cell, in the
 * entry order of each group.  Call once per shot geometry.
 */
int trtm_fwi_prepare(const trtm_fwi_desc *d, void *stream);

/*
 * Forward steps [it_begin, it_end) of the forward scheme (same expressions as
 * trtm_forward_steps) in REAL, recording into rec_* and, when hist != NULL,
 * the history of every step inside [hist_base, hist_base + hist_len).
 */
int trtm_fwi_forward(const trtm_fwi_desc *d, int32_t it_begin, int32_t it_end, void *stream);

/*
 * Adjoint steps n = it_end-1 down to it_begin (history must hold them): add
 * res_*[n] at the receivers to the adjoint state, apply the transposes of the
 * sponge, stress update, source term and velocity update of forward step n,
 * and accumulate the gradient with respect to lam, mu, mxz, myx, myz, inv_rho,
 * inv_rho_u, inv_rho_w.  Flat free surface only.
 */
int trtm_fwi_adjoint(const trtm_fwi_desc *d, int32_t it_begin, int32_t it_end, void *stream);

/*
 * Batched forms: n shots of one grid, precision and wave systems advance
 * together, up to 8 shots per launch (the shot is the grid's third
 * dimension); each shot keeps its own state, adjoint state, history and
 * gradient accumulators, and skips the steps at or beyond its own nt.  The
 * descriptors must share hist_base, hist_len and whether hist is set.
 */
int trtm_fwi_forward_batch(const trtm_fwi_desc *descs, int32_t n, int32_t it_begin,
                           int32_t it_end, void *stream);
int trtm_fwi_adjoint_batch(const trtm_fwi_desc *descs, int32_t n, int32_t it_begin,
                           int32_t it_end, void *stream);

/*
 * Scattered-field forward steps [it_begin, it_end), float, blocked kernel (no
 * reference counterpart).  Advances this model's state exactly as
 * trtm_fwi_forward_batch does (bit for bit) and, beside it, the scattered
 * state ds = s(model) - s(reference model) by the linear recurrence
 *   dv' = A (dv + dt b_ref D ds + dt (b - b_ref) D s) + source x (b - b_ref)
 *   ds' = A (ds + dt C_ref D dv' + dt (C - C_ref) D v')
 * with the differences (b - b_ref, C - C_ref) from dif_* and b_ref = b -
 * dif, C_ref = C - dif
 * of the forward step (forward_modeling.py:309-402, 421-443), whose float32
 * rounding is relative to ds itself.  rec_* receive the scattered records,
 * d(model) - d(reference): the FWI residual against data modelled in the
 * reference model, without the cancellation of two float32 record sets.
 * hist must be NULL; up to 4 shots per launch.
 */
int trtm_fwi_forward_scattered_batch(const trtm_fwi_desc *descs, int32_t n, int32_t it_begin,
                                     int32_t it_end, void *stream);

/*
 * Coefficient-space gradient to (lam, mu, rho) through the harmonic and
 * arithmetic averages (forward_modeling.py:243-307), then to (vp, vs, rho)
 * with mu = rho vs^2, lam = rho (vp^2 - 2 vs^2) (rtm.py:901-902).  Double
 * precision inputs (nx, nz) and outputs.
 */
int trtm_fwi_chain_gradient(int32_t nx, int32_t nz, const double *mu, const double *rho,
                            const double *vp, const double *vs, const double *g_lam,
                            const double *g_mu, const double *g_mxz, const double *g_myx,
                            const double *g_myz, const double *g_ir, const double *g_iru,
                            const double *g_irw, double *g_vp, double *g_vs, double *g_rho,
                            void *stream);

/*
 * Standalone finite check of u, v, w over the full grid: NaN or |x| > 1e30
 * gives code base+1 / base+2 / base+3, the highest found is atomically
 * max-ed into *flag (device).  Replaces _check_finite
 * (forward_modeling.py:455-488 with base 0; backward_modeling.py:417-450
 * with base 3).
 */
int trtm_check_finite(int32_t nx, int32_t nz, const float *u, const float *v, const float *w,
                      int32_t base, int32_t *flag, void *stream);

/* dst (cols, rows) = transpose of src (rows, cols), float32, device. */
int trtm_transpose(int32_t rows, int32_t cols, const float *src, float *dst, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TAICHIRTM_HIP_H */
