/*
 * pequod_hip.h -- C ABI of the MI355X implementation of pequod's time-step
 * hot path: the modal Helmholtz (PV) inversion and the CABARET advection
 * step.  Library: libpequod_hip.so (built from pequod_amd/csrc).
 *
 * Conventions
 *   - Every entry point returns an int status: PQ_OK (0) or a PQ_ERR_* code;
 *     pq_last_error() returns a message for the calling thread.  The library
 *     never exits the process; the Fortran shim turns a nonzero status into
 *     the reference's labort (debug.h:10, qg3p_base.F90:383-403).
 *   - Arrays use the reference's Fortran layout: column-major, i fastest.
 *     A "g x g view" (ptr, ld1, ld2) addresses element (i, j, l) (1-based) at
 *     ptr[(i-1) + ld1*((j-1) + ld2*(l-1))], so the sdata slice
 *     data(2:G+1, 2:G+1, :) of a data(G+2, G+2, L) array is passed as
 *     (&data(2,2,1), G+2, G+2) without a copy (qg3p_base.F90:1322-1323).
 *   - Entry points whose names end in _dev take device pointers and enqueue
 *     work on the handle's stream without synchronising.  Entry points
 *     without the suffix take host pointers: if the pointer was registered
 *     with pq_mirror_register the device mirror is used in place (no PCIe
 *     traffic); otherwise the data is copied in and out and the call is
 *     synchronous (the plain drop-in).
 *   - Streams are passed as void* (a hipStream_t); NULL selects the handle's
 *     own stream and PQ_STREAM_NULL the HIP null (default) stream.
 *   - Handles are independent and re-entrant: each owns its device buffers,
 *     plans and stream (the reference's caches and CABARET work arrays are
 *     `save` variables and not re-entrant, qg3p_aux.F90:1084,1185,1256).
 */
#ifndef PEQUOD_HIP_H
#define PEQUOD_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQ_OK 0
#define PQ_ERR_ARG 1      /* invalid argument or unsupported size           */
#define PQ_ERR_HIP 2      /* HIP runtime failure                            */
#define PQ_ERR_NAN 3      /* NaN psiC (qg3p_aux.F90:504-509 labort)         */
#define PQ_WARN_LARGE 4   /* |psiC| > 1e100 (qg3p_aux.F90:510-517 warning)  */
#define PQ_ERR_NOMEM 5    /* device memory exhausted                        */

/* Stream argument meaning "the HIP null stream" (NULL means the handle's
 * own stream).                                                             */
#define PQ_STREAM_NULL ((void *)-1)

/* Arithmetic of the sine transforms.  Modal projection coefficients, the
 * eigendecomposition, psiC and the mass correction constants are FP64 in
 * every mode.                                                              */
#define PQ_PREC_FP64 0    /* FP64 storage and transforms (plain drop-in)   */
#define PQ_PREC_FP32 1    /* FP32 storage and FP32 transforms              */
#define PQ_PREC_MIXED 2   /* FP64 storage, FP32 transforms                 */
#define PQ_PREC_FP32_STORAGE 3 /* FP32 storage, FP64 transforms             */
/* FP32 storage; FP32 transforms when the layer <-> mode projection is well
 * conditioned (pq_prec_fp32_choice), else FP64 transforms; the FP32
 * transforms run on the matrix cores where PQ_PREC_FP32_MFMA exists for
 * the grid (pq_prec_fp32_auto).                                            */
#define PQ_PREC_FP32_AUTO 4
/* FP32 storage (fields, scratch and on-chip buffers), FP64 arithmetic: the
 * transforms round to FP32 only between FFT stages.                        */
#define PQ_PREC_FP32_WIDE 5
/* FP32 storage and scratch; the row sine transforms (along i) as two GEMM
 * stages on the matrix cores (fp16 hi + lo split operands, FP32
 * accumulation), the column pass as in PQ_PREC_FP32.  Transform length
 * g - 1 = 1024; the closed basin only.                                     */
#define PQ_PREC_FP32_MFMA 6

/* PQ_PREC_FP32 or PQ_PREC_FP32_STORAGE for an L-layer stratification: FP32
 * transforms when kappa_inf(modes_to_z) = |modes_to_z|_inf |z_to_modes|_inf
 * <= 32 (the FP32 transform error can reach psi amplified by it).         */
int pq_prec_fp32_choice(int layers, const double *modes_to_z,
                        const double *z_to_modes);
/* What PQ_PREC_FP32_AUTO resolves to for a g x g closed basin:
 * PQ_PREC_FP32_MFMA where pq_prec_fp32_choice picks FP32 transforms and the
 * matrix-core transforms exist for g - 1, else pq_prec_fp32_choice's pick.
 * No counterpart in the reference (its precision is fixed at build time).  */
int pq_prec_fp32_auto(int g, int layers, const double *modes_to_z,
                      const double *z_to_modes);

const char *pq_last_error(void);
/* Library/runtime information: device name and CU count into buf.         */
int pq_device_info(char *buf, int len);

/* ---- device memory helpers (for callers without their own allocator) -- */
int pq_malloc(void **dev_ptr, size_t bytes);
int pq_free(void *dev_ptr);
int pq_memcpy_h2d(void *dev, const void *host, size_t bytes);
int pq_memcpy_d2h(void *host, const void *dev, size_t bytes);
int pq_synchronize(void);

/* ---- device mirrors of host arrays (keeps Fortran state resident) ------ */
/* Registers [host, host+bytes) and allocates a device copy (uploaded now). */
int pq_mirror_register(void *host, size_t bytes);
int pq_mirror_unregister(void *host);
/* Device address mirroring host pointer p (anywhere inside a registered
 * range), or NULL.                                                        */
void *pq_mirror_device_ptr(const void *host);
int pq_mirror_push(const void *host); /* host -> device for the whole range */
int pq_mirror_pull(void *host);       /* device -> host for the whole range */

/* ======================================================================
 * Helmholtz solver: replaces qg_helmholtz's allocate(helm_c, decomp, ss),
 * helm(helm_c, zeta) and deallocate(helm_c)
 * (qg3p_helmholtz.F90:292-295, 614-769, 389; the seam is the renamed import
 * at qg3p_aux.F90:13-17).  Solves (Delta_5 - ss_l) psi_l = zeta_l in grid
 * units with homogeneous Dirichlet values, per layer, in place, with the
 * reference's output boundary semantics (column i=1 and row j=jj keep their
 * input, column i=ii and row j=1 are zeroed on the interior).
 * ii = jj = G = n + 2 with n + 1 = 2^a (the reference's sizes) or 3 * 2^a.
 * ==================================================================== */
typedef struct pq_helm pq_helm;
int pq_helm_create(pq_helm **h, int ii, int jj, int layers, const double *ss,
                   int precision);
/* The channel's helm: walls at i = 1 and i = ii, periodic in j with period
 * jj = 2 (ii - 1) (allocate(helm_c, decomp, ss) on a decomposition with
 * periodic_j, qg3p_helmholtz.F90:292-386, 423-610).  The same solve and
 * boundary semantics along i; every j is interior.                        */
int pq_helm_create_periodic(pq_helm **h, int ii, int jj, int layers, const double *ss,
                            int precision);
int pq_helm_destroy(pq_helm *h);
int pq_helm_solve(pq_helm *h, double *zeta, int ld1, int ld2);
/* zeta is a device g x g x layers view of FP64 (or FP32 for PQ_PREC_FP32). */
int pq_helm_solve_dev(pq_helm *h, void *zeta, long ld1, long ld2, void *stream);

/* ======================================================================
 * PV inversion: replaces allocate_helmholtz_cache(hcache, decomp, strat,
 * vbc) and pv_inversion(strat, vbc, zeta, hcache, psi) for the closed basin
 * (qg3p_aux.F90:152-193, 452-479): z_to_modes, helm per mode, Dirichlet
 * boundary, mean psiC and the mass-conserving correction (modes_to_z_btcu,
 * qg3p_diagnostics.F90:615-635).
 * z_to_modes and modes_to_z are L x L column-major as in the reference's
 * stratification type (qg3p_diagnostics.F90:25-47); modes_to_z(:,1) must
 * be the barotropic mode (all ones) with ss(1) = 0.
 * A handle serves up to batch_max members of the same G and L in one call.
 * ==================================================================== */
typedef struct pq_inv pq_inv;
int pq_inv_create(pq_inv **h, int g, int layers, const double *ss,
                  const double *modes_to_z, const double *z_to_modes,
                  int precision, int batch_max);
/* The periodic channel (decomposition periodic in j): replaces
 * allocate_helmholtz_cache and pv_inversion with
 * correction_function_momentum, zeta_boundary and
 * modes_to_z_correct_momentum (qg3p_aux.F90:152-193, 233-379, 452-479,
 * 524-575).  Views are gi x gj with gj = 2 (gi - 1) (walls at i = 1 and
 * i = gi, period gj along j).  The baroclinic modes are corrected so that
 * their mean boundary relative PV vanishes on both walls; bc_type as the
 * reference's velocity boundary condition (1 partial slip with alpha,
 * 2 no slip, 3 free slip).  The returned handle is used with pq_inv_run /
 * pq_inv_run_dev / pq_inv_status / pq_inv_set_tiles / pq_inv_destroy;
 * pq_inv_status's psic receives (psiC_1, psiC_2) per mode (2 L values per
 * member, those of the barotropic mode zero).                             */
int pq_chan_create(pq_inv **h, int gi, int gj, int layers, const double *ss,
                   const double *modes_to_z, const double *z_to_modes, int bc_type,
                   double alpha, int precision, int batch_max);
int pq_inv_destroy(pq_inv *h);
/* Host drop-in: zeta/psi are g x g x L views (FP64).  Synchronous unless
 * both arrays are registered mirrors.  Returns PQ_ERR_NAN / PQ_WARN_LARGE
 * like the reference's checks.                                           */
int pq_inv_run(pq_inv *h, const double *zeta, int ldz1, int ldz2, double *psi,
               int ldp1, int ldp2);
/* Device batch: zeta[b] / psi[b] are device g x g x L views (FP64, or FP32
 * for PQ_PREC_FP32), b < batch.  Enqueued on stream; no synchronisation.  */
int pq_inv_run_dev(pq_inv *h, int batch, const void *const *zeta, long ldz1,
                   long ldz2, void *const *psi, long ldp1, long ldp2,
                   void *stream);
/* Per-member stratification (no counterpart in the reference, whose cache
 * holds one): member slot `slot` of later runs uses these ss / modes_to_z /
 * z_to_modes instead of the creation ones.  For ensembles whose members
 * share a grid size and layer count but not a stratification.  Synchronous. */
int pq_inv_set_member(pq_inv *h, int slot, const double *ss,
                      const double *modes_to_z, const double *z_to_modes);
/* As pq_inv_set_member for a member of `layers` <= the handle's layer count
 * (no counterpart in the reference): one handle then serves members of
 * several layer counts on the same grid in one batch (ragged batches); its
 * psi / zeta views hold `layers` layers.  Synchronous.                     */
int pq_inv_set_member_layers(pq_inv *h, int slot, int layers, const double *ss,
                             const double *modes_to_z, const double *z_to_modes);
/* Waits for the handle's last run and returns its status (NaN / large
 * psiC); psic (may be NULL) receives batch*L values of the last run.       */
int pq_inv_status(pq_inv *h, double *psic);
/* The handle's stream (hipStream_t). */
void *pq_inv_stream(pq_inv *h);
/* Tile shapes (no counterpart in the reference; used by the autotuner):
 * rows per tile of the row passes (also the scratch block height, a power
 * of two in 2..16 dividing g-1) and columns per workgroup of the column pass
 * (a power of two in 1..8).  Rejected shapes return PQ_ERR_ARG and leave the
 * handle unchanged.  Results do not depend on the tiles.                  */
int pq_inv_set_tiles(pq_inv *h, int rows, int cols);
int pq_inv_get_tiles(pq_inv *h, int *rows, int *cols);
/* Measurement helper (no counterpart in the reference): runs the three
 * passes of one inversion separately, reps times each, on the handle's
 * stream between HIP events; ms3 receives the mean duration of the row
 * forward, column and row inverse pass in milliseconds.                   */
int pq_inv_pass_times(pq_inv *h, int batch, const void *const *zeta,
                      long ldz1, long ldz2, void *const *psi, long ldp1,
                      long ldp2, int reps, float *ms3);

/* ======================================================================
 * CABARET stages on device arrays (qg3p_aux.F90:917-1642 and the boundary
 * routines at 643-915).  Array shapes as in the reference on one domain:
 * node fields (G+2, G+2, L); i-faces (G+3, G+2, L); j-faces (G+2, G+3, L);
 * res / res_beta (G, G, L); force_wind (G, G).  All FP64.  These are
 * asynchronous on `stream`.
 * ==================================================================== */
/* cabaret_cfl (qg3p_aux.F90:920-943): writes max|u| * dt to *cfl_host
 * (synchronous; the maximum is formed in an internal buffer of the
 * stream's, allocated on its first call).                                */
int pq_cabaret_cfl_dev(int g, int layers, double dt, const double *ui,
                       const double *uj, double *cfl_host, void *stream);
int pq_cabaret_predictor_dev(int g, int layers, double dt,
                             const double *zeta_old, const double *res,
                             double *zeta, void *stream);
int pq_cabaret_beta_dev(int g, int layers, double dt, double tweight_a,
                        const double *beta, const double *ui, double *zeta,
                        double *res_beta, void *stream);
int pq_cabaret_forcing_dev(int g, int layers, double dt,
                           const double *force_wind, double viscosity,
                           double visc_bot, const double *rel, double *zeta,
                           void *stream);
int pq_cabaret_cell_boundaries_dev(int g, int layers, double tweight_u,
                                   const double *psi, double *ui, double *uj,
                                   double *uimo, double *ujmo, void *stream);
int pq_cabaret_cell_boundaries_velocity_dev(int g, int layers,
                                            const double *psi, double *ui,
                                            double *uj, void *stream);
/* cabaret_extrapolator (qg3p_aux.F90:1317-1544): zeta_fi / zeta_fj are
 * updated in place; work must hold one i-face and one j-face array
 * ((G+3)(G+2)L + (G+2)(G+3)L doubles) or be NULL: the library then uses an
 * internal buffer of the stream's, allocated on the first call and grown
 * only when a later call on that stream needs more.                      */
int pq_cabaret_extrapolator_dev(int g, int layers, double dt,
                                const double *zeta, const double *zeta_old,
                                const double *ui, const double *uj,
                                double *zeta_fi, double *zeta_fj,
                                double *work, void *stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  pq_cabaret_extrapolator_dev without a work argument:
 * it allocates nothing, keeping the new face values in a fixed 64 MiB
 * device array of the library's (one per GPU, held until the process
 * ends), so 2 (g+3)(g+2) layers <= 8388608 doubles (e.g. g = 1025 with up
 * to 3 layers); larger grids return PQ_ERR_ARG before any GPU work.  Every
 * call on a GPU shares that array: calls that overlap on one GPU (on
 * different streams or threads) race.                                     */
int pq_cabaret_extrapolator_noalloc_dev(int g, int layers, double dt,
                                        const double *zeta, const double *zeta_old,
                                        const double *ui, const double *uj,
                                        double *zeta_fi, double *zeta_fj, void *stream);
int pq_cabaret_flux_dev(int g, int layers, const double *ui, const double *uj,
                        const double *zeta_fi, const double *zeta_fj,
                        double *flux_i, double *flux_j, void *stream);
int pq_cabaret_corrector_dev(int g, int layers, double dt,
                             const double *flux_i, const double *flux_j,
                             double *res, double *zeta, void *stream);
/* rel_from_zeta_and_psi + apply_rel_bc (qg3p_aux.F90:738-915);
 * bc_type 1 partial slip (alpha), 2 no slip, 3 free slip.                 */
int pq_rel_from_zeta_and_psi_dev(int g, int layers, const double *s1,
                                 const double *s2, int bc_type, double alpha,
                                 const double *zeta, const double *psi,
                                 double *rel, void *stream);
/* The centred scheme's step (zeta_centred, qg3p_aux.F90:2027-2202, and
 * zeta_tendency / zeta_forward_euler, qg3p_aux.F90:1789-2024): dst = src +
 * T F(psi, rel, rel_p) on the interior nodes (src may equal dst); beta[L]
 * (beta_y formed inside), force_wind (G, G).                              */
int pq_zeta_step_dev(int g, int layers, double T, const double *beta, const double *s1,
                     const double *s2, const double *force_wind, double viscosity,
                     double visc_bot, const double *psi, const double *rel,
                     const double *rel_p, const double *src, double *dst, void *stream);
/* robert_asselin_filter with Williams' modification (nu = 0.02, alpha =
 * 0.5; qg3p_aux.F90:1716-1786): phi_l level n+1, phi_m level n, phi_old_l
 * level n-1; phi_old_m receives the filtered level n.                      */
int pq_ra_filter_dev(int g, int layers, double *phi_l, double *phi_m, const double *phi_old_l,
                     double *phi_old_m, void *stream);
/* update_zeta_bc (qg3p_aux.F90:643-734). */
int pq_update_zeta_bc_dev(int g, int layers, const double *s1,
                          const double *s2, const double *rel,
                          const double *psi, double *zeta, void *stream);

/* ======================================================================
 * Device-resident model: one CABARET time step in the driver's order
 * (qg3.F90:590-967), with the PV inversion, for state that stays in HBM.
 * ==================================================================== */
typedef struct pq_model pq_model;
/* Time stepping scheme of a model: the CABARET scheme (the reference's
 * default build), or the centred (leapfrog) scheme with the
 * Robert-Asselin-Williams filter and a forward Euler start, the driver built
 * without CABARET (qg3.F90:456-490, 612-660).                              */
#define PQ_SCHEME_CABARET 0
#define PQ_SCHEME_LEAPFROG 1
typedef struct {
  int g, layers;
  const double *s1, *s2;   /* [L] stratification (non-dimensional)        */
  const double *ss, *modes_to_z, *z_to_modes; /* from the host eigensolve */
  const double *beta;      /* [L] non-dimensional beta                     */
  const double *force_wind;/* (G, G), already scaled (qg3.F90:253-279)     */
  double viscosity, visc_bot;
  int bc_type;
  double alpha;
  int adapt_dt;
  double target_cfl, max_dt, max_dt_increase;
  double dt;               /* initial (non-dimensional) time step          */
  int precision;           /* PQ_PREC_* of the inversion; it also sets the
                            * field storage: FP64 for PQ_PREC_FP64 and
                            * PQ_PREC_MIXED, FP32 for the other modes (the
                            * stages compute in FP64 either way)          */
  int scheme;              /* PQ_SCHEME_* (adapt_dt needs CABARET)         */
} pq_model_params;

int pq_model_create(pq_model **m, const pq_model_params *p);
int pq_model_destroy(pq_model *m);
/* Start like START_PERTURB (qg3.F90:393-425): zeta(G,G,L) host interior
 * field is installed, boundary set to zero, inverted, rel/zeta_bc applied,
 * CABARET work arrays zeroed, zeta_old = zeta, model_time = dt/2.         */
int pq_model_init(pq_model *m, const double *zeta_sdata_host);
/* Run nsteps full steps; if final_partial, the last step stops after the
 * flux stage (the driver's exit point, qg3.F90:815-946).  A centred-scheme
 * model runs nsteps leapfrog steps of dt, alternating the time level they
 * write (1, 2, 1, ...; pq_model_init leaves level 1 next, after the
 * forward Euler start into level 2); final_partial must be 0.             */
int pq_model_step(pq_model *m, int nsteps, int final_partial);
/* Status of the model's PV inversions since the last check (the reference
 * checks psiC in every pv_inversion, qg3p_aux.F90:504-517): waits for the
 * model's work, returns PQ_ERR_NAN after a NaN psiC (the reference aborts)
 * or PQ_WARN_LARGE after |psiC| > 1e100, and clears it.  pq_model_step is
 * asynchronous and returns no such status, except that with adapt_dt, whose
 * CFL maximum comes back to the host every step, a step stops with
 * PQ_ERR_NAN once an earlier inversion gave a NaN psiC.                    */
int pq_model_status(pq_model *m);
/* Copy a field to the host: which = 0 psi, 1 zeta, 2 rel, 3 zeta_old,
 * 4 ui, 5 uj, 6 zeta_fi, 7 zeta_fj, 8 res, 9 res_beta, 10 uimo, 11 ujmo,
 * 12 flux_i, 13 flux_j (CABARET); a centred-scheme model has 0-3 for time
 * level 1 and 14 psi, 15 zeta, 16 rel, 17 zeta_old for level 2.  Sizes
 * follow the reference shapes.                                            */
int pq_model_get(pq_model *m, int which, double *host, size_t count);
int pq_model_set(pq_model *m, int which, const double *host, size_t count);
/* dt, tweight_a, tweight_u, model_time */
int pq_model_scalars(pq_model *m, double *out4);
void *pq_model_stream(pq_model *m);
/* Time mean of psi (qg3.F90:503, 672-681): with enable = 1 the sum and the
 * averaging time restart from zero, and every later step whose model time
 * (after the inversion) exceeds max_spin adds dt * psi over all G x G
 * nodes on the device; enable = 0 stops adding.  Synchronous.             */
int pq_model_average(pq_model *m, int enable, double max_spin);
/* The time mean psi_av / diagnostic_time as written to the _s_av output
 * (qg3.F90:886-898): (G, G, L) into host; diag_time (may be NULL) receives
 * the averaging time.  PQ_ERR_ARG before the first sample.                */
int pq_model_psi_av(pq_model *m, double *host, size_t count, double *diag_time);
/* energy(strat, psi, ekin, epot) of the model's current psi
 * (qg3p_diagnostics.F90:778-863, qg3.F90:744) with layer depths h[L] (the
 * model's s2 is the stratification's); FP64 reduction on the device, ekin[L]
 * and epot on the host.  PQ_ERR_NAN when a result is NaN.                 */
int pq_model_energy(pq_model *m, const double *h, double *ekin, double *epot);

/* ---- model output and checkpoints: the reference's data sets, a
 * plain-text meta-data file base[.index] and a binary data file
 * base[.index].d of Fortran-style records (4-byte byte count, doubles in
 * column-major order, byte count) in dir (qg3p_io.F90:202-317, 405-732;
 * qg3.F90:821-934; manual/manual.tex:1610-1790).  l_scale / t_scale go to
 * the meta-data (Fortran E26.17E3).                                        */
/* Snapshot qgL_G_s: the two times, then psi and zeta per layer (both time
 * levels per record for the centred scheme).  index < 0 omits the index,
 * which is the name START_PICKUP reads.  supp (CABARET only) also writes
 * qgL_G_s_supp: the two times, dt, ui, uj, zeta_fi, zeta_fj, res, res_beta
 * per layer.  Synchronous.                                                 */
int pq_model_write_snapshot(pq_model *m, const char *dir, int index, double model_time_day,
                            double total_time_day, double l_scale, double t_scale, int supp);
/* Time mean qgL_G_s_av: the two times, psi_av / diagnostic_time per layer. */
int pq_model_write_mean(pq_model *m, const char *dir, double model_time_day,
                        double total_time_day, double l_scale, double t_scale);
/* START_PICKUP (qg3.F90:316-364): reads dir/qgL_G_s.d (and qgL_G_s_supp.d
 * for CABARET), installs the fields and starts as the reference does
 * (CABARET: fluxes of the saved faces and the corrector's half step, dt
 * from the file, model_time = dt / 2; centred scheme: rel from psi,
 * model_time = 0).  The times read are returned (either may be NULL).      */
int pq_model_read_checkpoint(pq_model *m, const char *dir, double *model_time_day,
                             double *total_time_day);

/* ======================================================================
 * Diagnostics of a stream function field: replaces energy(strat, psi, ekin,
 * epot) (qg3p_diagnostics.F90:778-863) for the closed basin on one domain,
 * and the time-mean accumulation of the driver (qg3.F90:672-681).
 * Kinetic energy per layer (cell-centred squared gradient, weighted by
 * h/H / 2) and potential energy of the interfaces (trapezoid-weighted
 * s2(k) h(k) / (2H) (psi_k - psi_k+1)^2), both divided by (G-1)^2.
 * FP64, reduced on the device in a fixed order (deterministic).
 * ==================================================================== */
typedef struct pq_diag pq_diag;
/* h[L] layer depths (any unit: only h / sum(h) enters), s2[L] the
 * stratification's s2 (qg3p_diagnostics.F90:25-47); layers 1..8.          */
int pq_diag_create(pq_diag **d, int g, int layers, const double *h, const double *s2);
int pq_diag_destroy(pq_diag *d);
void *pq_diag_stream(pq_diag *d);
/* Host drop-in: psi a g x g x L view (FP64; a registered mirror is used in
 * place).  ekin[L], *epot as the reference's; synchronous; PQ_ERR_NAN when
 * a result is NaN (the reference aborts).                                 */
int pq_diag_energy(pq_diag *d, const double *psi, int ld1, int ld2, double *ekin, double *epot);
/* Device field psi (g x g x L FP64 view); out_dev[L + 1] receives
 * ekin(1..L), epot on the device.  Asynchronous on stream; allocates
 * nothing, so it may be captured into a HIP graph.                        */
int pq_diag_energy_dev(pq_diag *d, const void *psi, long ld1, long ld2, double *out_dev,
                       void *stream);
/* psi_av += dt * psi over every node of g x g x L device views (FP64);
 * asynchronous on stream (NULL: the null stream).                         */
int pq_psi_av_add_dev(int g, int layers, double dt, const void *psi, long ld1, long ld2,
                      double *psi_av, long la1, long la2, void *stream);

/* ======================================================================
 * Ensembles (no counterpart in the reference, which runs one member per
 * MPI job): independent members of mixed grid sizes, layer counts and
 * stratifications on one GPU.  Members with the same grid, leading
 * dimensions and transform precision share a ragged batched inversion
 * handle (a size class) with per-member layer counts and stratifications;
 * classes run concurrently.  The inversion of
 * all members, or one time step of all CABARET model members, can be
 * captured into a HIP graph and replayed.
 * ==================================================================== */
typedef struct pq_ens pq_ens;
/* precision: PQ_PREC_* of every class (FP64 storage, i.e. PQ_PREC_FP64 or
 * PQ_PREC_MIXED, when model members are added).                          */
int pq_ens_create(pq_ens **e, int precision);
/* A member whose zeta / psi are device g x g x layers views (types as for
 * pq_inv_run_dev); index receives its number.                            */
int pq_ens_add(pq_ens *e, int g, int layers, const double *ss,
               const double *modes_to_z, const double *z_to_modes,
               const void *zeta, long ldz1, long ldz2, void *psi, long ldp1,
               long ldp2, int *index);
/* A CABARET model member (fixed dt); the ensemble inverts its zeta into its
 * psi and steps it with pq_ens_step.  The model's field storage must be the
 * ensemble's: FP64 fields for PQ_PREC_FP64 / PQ_PREC_MIXED ensembles, FP32
 * fields for the FP32-storage modes (else PQ_ERR_ARG).                    */
int pq_ens_add_model(pq_ens *e, pq_model *m, const double *ss,
                     const double *modes_to_z, const double *z_to_modes,
                     int *index);
/* Groups the members into size classes and creates their handles.        */
int pq_ens_finalize(pq_ens *e);
int pq_ens_class_count(pq_ens *e);
pq_inv *pq_ens_class_handle(pq_ens *e, int k);  /* e.g. for pq_inv_set_tiles */
int pq_ens_member_class(pq_ens *e, int member, int *cls, int *slot);
/* Precision chosen for class k (resolves PQ_PREC_FP32_AUTO), or -1.      */
int pq_ens_class_precision(pq_ens *e, int k);
void *pq_ens_stream(pq_ens *e);
/* All members' inversions (asynchronous on the ensemble's stream).        */
int pq_ens_invert(pq_ens *e);
/* nsteps time steps of the model members (asynchronous).                  */
int pq_ens_step(pq_ens *e, int nsteps);
/* what = 0: capture the inversion, 1: capture one model step; later
 * pq_ens_invert / pq_ens_step replay the graph.                           */
int pq_ens_capture(pq_ens *e, int what);
int pq_ens_synchronize(pq_ens *e);
/* Waits and returns the worst status of the last inversions.              */
int pq_ens_status(pq_ens *e);
int pq_ens_destroy(pq_ens *e);

/* ======================================================================
 * Serving (no counterpart in the reference): members resident in HBM and a
 * stream of requests -- the PV inversion of a field member, or n CABARET
 * steps of a model member -- served with continuous batching.  A scheduler
 * thread launches every waiting inversion of a grid size as one ragged batch
 * whenever one of the size's handles is free, and a model member's advance
 * as replays of its captured one-step graph on its own stream; requests
 * that arrive meanwhile join the next launches.  Members are admitted by
 * device footprint: an open that would take the reserved bytes above the
 * cap fails with PQ_ERR_NOMEM.  Requests of one member run in order.
 * ==================================================================== */
typedef struct pq_srv pq_srv;
/* precision: PQ_PREC_FP32, _FP32_STORAGE, _MIXED or _FP64 for field
 * members (models are FP64); max_batch: members per inversion batch.      */
int pq_srv_create(pq_srv **s, int precision, size_t hbm_cap_bytes, int max_batch);
/* Field member: device g x g x layers views zeta / psi (types as for
 * pq_inv_run_dev; the caller keeps them allocated until pq_srv_close) with
 * its own stratification.                                                 */
int pq_srv_open_field(pq_srv *s, int g, int layers, const double *ss,
                      const double *modes_to_z, const double *z_to_modes,
                      const void *zeta, long ldz1, long ldz2, void *psi,
                      long ldp1, long ldp2, int *member);
/* Model member (fixed dt): created, started from zeta (as pq_model_init)
 * and owned by the engine.                                                */
int pq_srv_open_model(pq_srv *s, const pq_model_params *p,
                      const double *zeta_sdata_host, int *member);
pq_model *pq_srv_model(pq_srv *s, int member);  /* e.g. for pq_model_get */
/* kind 0: inversion of a field member; 1: `steps` steps of a model member. */
int pq_srv_submit(pq_srv *s, int member, int kind, int steps, long long *request);
/* 1 done (latency_s = completion - submission), 0 pending, or minus the
 * failure's code: -PQ_ERR_NAN when the request's inversion gave a NaN psiC
 * (the reference aborts, qg3p_aux.F90:504-509), -PQ_ERR_HIP when a launch
 * or the device failed.  A done request whose inversion gave |psiC| >
 * 1e100 returns 1 and leaves the reference's warning in pq_last_error.    */
int pq_srv_poll(pq_srv *s, long long request, double *latency_s);
/* Waits for every submitted request (at most timeout_s seconds).          */
int pq_srv_drain(pq_srv *s, double timeout_s);
/* Completed requests of kind (-1 all) numbered `first` or later: count,
 * mean, p50, p95, p99, max latency [s], completions per second (first
 * submission to last completion), failures.                               */
int pq_srv_stats(pq_srv *s, int kind, long long first, double *out8);
size_t pq_srv_used_bytes(pq_srv *s);
/* Closes an idle member and releases its footprint.                       */
int pq_srv_close(pq_srv *s, int member);
int pq_srv_destroy(pq_srv *s);

#ifdef __cplusplus
}
#endif
#endif
