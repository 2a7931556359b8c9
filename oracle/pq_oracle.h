/*
 * pq_oracle.h -- CPU FP64 restatement of pequod's time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against (tests/, __graft_entry__.smoke(), and the
 * `cpu_baseline` leg of bench.py).  Nothing in pequod_amd/ links, loads or
 * calls it; the product path fails loudly when its HIP library is missing.
 *
 * Layout conventions are those of the Fortran reference (column-major, i is
 * the fastest index, 1-based in the comments below):
 *   node field  data(G+2, G+2, L)  (qg3p_base.F90:1310-1325); the owned nodes
 *               are data(2:G+1, 2:G+1, :) ("sdata"); G = n + 2 nodes per side.
 *   i-faces     data(G+3, G+2, L)  (FLUX_II, qg3p_base.F90:2808-2810)
 *   j-faces     data(G+2, G+3, L)  (FLUX_JJ, qg3p_base.F90:2811-2813)
 *   residuals   res(G, G, L)        (qg3.F90:80)
 * All routines describe a single, non-periodic domain (p_a = p_b = p_c =
 * p_d = -1), for which the reference's halo exchanges are no-ops.
 */
#ifndef PQ_ORACLE_H
#define PQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sine transform (qg3p_helmholtz.F90:168-183, 1370-1591) ---- */
typedef struct pqo_sint pqo_sint;
pqo_sint *pqo_sint_create(int n);
void pqo_sint_destroy(pqo_sint *p);
/* Unnormalised DST-I in place on x[0..n-1]:
 *   X_k = 2 sum_{j=1..n} x_j sin(pi j k / (n+1)).                            */
void pqo_sint_apply(pqo_sint *p, double *x);

/* ---- modal Helmholtz solver (qg3p_helmholtz.F90:292-386, 614-769) ---- */
typedef struct pqo_helm pqo_helm;
pqo_helm *pqo_helm_create(int g, int layers, const double *ss);
void pqo_helm_destroy(pqo_helm *h);
/* In place on zeta(ld1, ld2, layers); uses zeta(1:g, 1:g, :).             */
void pqo_helm_solve(pqo_helm *h, double *zeta, int ld1, int ld2);

/* ---- periodic-j Helmholtz solver, the channel's (non-PH1D path:
 * poinit_j, periodic_coefficients, periodic_ge, helm;
 * qg3p_helmholtz.F90:423-610, 614-769) on gi x gj nodes, walls at i = 1
 * and i = gi, periodic in j with period gj. ---- */
typedef struct pqo_phelm pqo_phelm;
pqo_phelm *pqo_phelm_create(int gi, int gj, int layers, const double *ss);
void pqo_phelm_destroy(pqo_phelm *h);
/* In place on zeta(ld1, ld2, layers); uses zeta(1:gi, 1:gj, :).           */
void pqo_phelm_solve(pqo_phelm *h, double *zeta, int ld1, int ld2);

/* ---- channel PV inversion (qg3p_aux.F90:152-193, 233-379, 452-479,
 * 524-575): helm with the momentum-conserving correction of the
 * baroclinic modes.  bc_type 1 partial slip (alpha), 2 no slip, 3 free
 * slip. ---- */
typedef struct pqo_chan pqo_chan;
pqo_chan *pqo_chan_create(int gi, int gj, int layers, const double *ss,
                          const double *modes_to_z, const double *z_to_modes,
                          int bc_type, double alpha);
void pqo_chan_destroy(pqo_chan *v);
/* corr (gi, gj, L, 2) and inv_corr_a (2, 2, L-1) of the handle; psiC
 * (2, L-1) of the last run.                                               */
const double *pqo_chan_corr(const pqo_chan *v);
const double *pqo_chan_inv_corr_a(const pqo_chan *v);
const double *pqo_chan_psic(const pqo_chan *v);
/* zeta_boundary (qg3p_aux.F90:288-379) of psi (gi, gj, L), modal.         */
void pqo_zeta_boundary(int gi, int gj, int layers, const double *ss, int bc_type,
                       double alpha, const double *psi, double *z_b, double *z_d);
/* pv_inversion of zeta into psi (g x g views as pqo_inv_run); returns 0, 3
 * (NaN psiC) or 4 (|psiC| > 1e100).                                       */
int pqo_chan_run(pqo_chan *v, const double *zeta, int ldz1, int ldz2, double *psi,
                 int ldp1, int ldp2);

/* ---- boundary helpers (qg3p_laplace_bc.F90:55-87, 193-224) ---- */
void pqo_apply_dirichlet_bc(int g, int layers, double bc, double *zeta,
                            int ld1, int ld2);
void pqo_add_laplace_bc_rhs(int g, int layers, double bc, double *rhs,
                            int ld1, int ld2);
/* Trapezoid mean of one g x g layer (qg3p_diagnostics.F90:669-720).       */
double pqo_mean(int g, const double *psi, int ld1);

/* ---- energy diagnostics (qg3p_diagnostics.F90:778-863) ---- */
/* One closed domain of gi x gj nodes; psi is the data array
 * (gi+2, gj+2, L) with halo (sdata(i, j) = data(i+1, j+1)); h, s2 [L].   */
void pqo_energy(int gi, int gj, int layers, const double *h, const double *s2,
                const double *psi, double *ekin, double *epot);
/* psi_av(i,j,l) += dt * psi(i,j,l) over all nodes (qg3.F90:672-681); psi is
 * a data array as above, psi_av (gi, gj, L).                              */
void pqo_psi_av_add(int gi, int gj, int layers, double dt, const double *psi,
                    double *psi_av);

/* ---- stratification (qg3p_diagnostics.F90:210-422) ---- */
/* s1[l], s2[l] as in the reference's stratification type (s1(1) and s2(L)
 * unused).  Outputs: ss[L], modes_to_z[L*L], z_to_modes[L*L] (column-major).
 * Returns 0 on success.                                                    */
int pqo_strat_modes(int layers, const double *s1, const double *s2,
                    double *ss, double *modes_to_z, double *z_to_modes);
void pqo_strat_l2(double rd, double h1, double h2, double *s1, double *s2);
void pqo_strat_l3(double rd1, double rd2, double h1, double h2, double h3,
                  double *s1, double *s21, double *s22, double *s3);

/* ---- PV inversion (qg3p_aux.F90:152-230, 452-521) ---- */
typedef struct pqo_inv pqo_inv;
pqo_inv *pqo_inv_create(int g, int layers, const double *ss,
                        const double *modes_to_z, const double *z_to_modes);
void pqo_inv_destroy(pqo_inv *v);
/* zeta and psi are g x g x L views with leading dimensions (ld1, ld2).
 * Returns 0, 3 (NaN psiC) or 4 (|psiC| > 1e100, result still written).    */
int pqo_inv_run(pqo_inv *v, const double *zeta, int ldz1, int ldz2,
                double *psi, int ldp1, int ldp2);
/* Correction field corr(g, g, L) and last psiC[L].                        */
const double *pqo_inv_corr(const pqo_inv *v);
const double *pqo_inv_psic(const pqo_inv *v);

/* ---- CABARET stages (qg3p_aux.F90:643-1642) ---- */
double pqo_cabaret_cfl(int g, int layers, double dt, const double *ui,
                       const double *uj);
void pqo_cabaret_adapt_dt(int g, int layers, double target_cfl, double max_dt,
                          double max_dt_increase, const double *ui,
                          const double *uj, double *dt, double *tweight_a,
                          double *tweight_u);
void pqo_cabaret_predictor(int g, int layers, double dt,
                           const double *zeta_old, const double *res,
                           double *zeta);
void pqo_cabaret_beta(int g, int layers, double dt, double tweight_a,
                      const double *beta, const double *ui, double *zeta,
                      double *res_beta);
void pqo_cabaret_forcing(int g, int layers, double dt,
                         const double *force_wind, double viscosity,
                         double visc_bot, const double *rel, double *zeta);
void pqo_cabaret_cell_boundaries_velocity(int g, int layers, const double *psi,
                                          double *ui, double *uj);
void pqo_cabaret_cell_boundaries(int g, int layers, double tweight_u,
                                 const double *psi, double *ui, double *uj,
                                 double *uimo, double *ujmo);
void pqo_cabaret_extrapolator(int g, int layers, double dt, const double *zeta,
                              const double *zeta_old, const double *ui,
                              const double *uj, double *zeta_fi,
                              double *zeta_fj);
void pqo_cabaret_flux(int g, int layers, const double *ui, const double *uj,
                      const double *zeta_fi, const double *zeta_fj,
                      double *flux_i, double *flux_j);
void pqo_cabaret_corrector(int g, int layers, double dt, const double *flux_i,
                           const double *flux_j, double *res, double *zeta);

/* relative vorticity and boundary values (qg3p_aux.F90:643-915) */
enum { PQO_BC_PARTIAL_SLIP = 1, PQO_BC_NO_SLIP = 2, PQO_BC_FREE_SLIP = 3 };
void pqo_rel_from_zeta_and_psi(int g, int layers, const double *s1,
                               const double *s2, int bc_type, double alpha,
                               const double *zeta, const double *psi,
                               double *rel);
void pqo_apply_rel_bc(int g, int layers, int bc_type, double alpha,
                      const double *psi, double *rel);
void pqo_update_zeta_bc(int g, int layers, const double *s1, const double *s2,
                        const double *rel, const double *psi, double *zeta);

/* Wind stress curl field force_wind(g, g) (qg3.F90:253-279). */
void pqo_wind_field(int g, double tau, double wind_shift, double wind_tilt,
                    double wind_asym, double *force_wind);

/* ---- one full CABARET time step in the driver's order (qg3.F90:590-967) */
typedef struct {
  int g, layers;
  double *psi, *rel, *zeta, *zeta_old;  /* (g+2)(g+2)L */
  double *zeta_fi, *ui, *uimo, *flux_i; /* (g+3)(g+2)L */
  double *zeta_fj, *uj, *ujmo, *flux_j; /* (g+2)(g+3)L */
  double *res, *res_beta;               /* g g L */
  double *force_wind;                   /* g g   */
  double dt, tweight_a, tweight_u, model_time;
  /* time mean (qg3.F90:672-681): psi_av (g g L) or NULL when not averaging */
  double *psi_av;
  double max_spin, diag_time;
} pqo_state;

typedef struct {
  const double *beta, *s1, *s2; /* [L] */
  double viscosity, visc_bot;
  int bc_type;
  double alpha;
  int adapt_dt;
  double target_cfl, max_dt, max_dt_increase;
} pqo_params;

/* Runs stages 1-13 (adapt_dt .. flux); if do_corrector, also the corrector
 * and zeta_old <- zeta.  Returns the inversion status.                      */
int pqo_step(pqo_state *s, const pqo_params *p, pqo_inv *inv,
             int do_corrector);

/* ---- centred (leapfrog) integration, the driver without CABARET
 * (qg3.F90:456-490, 612-632; qg3p_aux.F90:1644-2202), with WIND_FORCING,
 * RA_DOUBLE_FILTER and FORWARD_EULER_START, no background flow ---- */
/* dst = src + T F(psi, rel, rel_p) on the interior nodes: zeta_centred
 * (T = 2 dt, psi and rel of level m, rel_p = rel of level l, src = dst =
 * zeta(l)) and zeta_forward_euler (T = dt, one level).  Node arrays
 * (g+2, g+2, L); force_wind (g, g); beta [L] (beta_y formed inside).       */
void pqo_zeta_step(int g, int layers, double T, const double *beta, const double *s1,
                   const double *s2, const double *force_wind, double viscosity,
                   double visc_bot, const double *psi, const double *rel,
                   const double *rel_p, const double *src, double *dst);
/* robert_asselin_filter with RA_DOUBLE_FILTER (nu = 0.02, alpha = 0.5) on
 * the owned nodes of every layer: phi_l = level n+1, phi_m = level n,
 * phi_old_l = level n-1 (all updated as the reference does).               */
void pqo_ra_filter(int g, int layers, double *phi_l, double *phi_m, const double *phi_old_l,
                   double *phi_old_m);

typedef struct {
  int g, layers;
  double *psi[2], *rel[2], *zeta[2], *zeta_old[2]; /* (g+2)(g+2)L each */
  double *force_wind;                               /* g g */
  double dt, model_time;
  int leapfrog; /* 1 or 2: the level the next step writes */
} pqo_lf_state;
/* START_PERTURB-style start from zeta(1) (interior set by the caller,
 * boundary zero) followed by the forward Euler start; returns the worst
 * inversion status.                                                        */
int pqo_lf_init(pqo_lf_state *s, const pqo_params *p, pqo_inv *inv);
/* One leapfrog step: centred step, filter, inversion, rel, zeta_bc,
 * model_time += dt.                                                        */
int pqo_lf_step(pqo_lf_state *s, const pqo_params *p, pqo_inv *inv);

#ifdef __cplusplus
}
#endif
#endif
