// pq_internal.h -- entry points shared between the library's translation
// units (inversion, CABARET model, ensemble and serving runtimes); not part
// of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "pequod_hip.h"

namespace pq {

// ---- inversion handles (pq_inversion.hip)
// Doubles in one member's coefficient block of handle h (its layer count).
size_t inv_coef_size(const pq_inv *h);
// Coefficient block of a member with `layers` <= the handle's layers and
// the given stratification, written to dev_block (device, inv_coef_size
// doubles).  Synchronous.
int inv_make_coefs(const pq_inv *h, int layers, const double *ss, const double *m2z,
                   const double *z2m, double *dev_block);
// A ragged batch whose members bring their own coefficient blocks and layer
// counts (1..the handle's layers); otherwise as pq_inv_run_dev.
int inv_run_members(pq_inv *h, int batch, const void *const *zeta, long ldz1, long ldz2,
                    void *const *psi, long ldp1, long ldp2, const double *const *coef,
                    const int *layers, hipStream_t s);
// Device bytes held by the handle (tables, scratch, coefficients).
size_t inv_device_bytes(const pq_inv *h);
// The handle's per-slot status words on the device (inv_batch_max ints):
// bit 0 NaN psiC, bit 1 |psiC| > 1e100 (qg3p_aux.F90:504-517); the
// inversion ORs into them and only pq_inv_status / the caller clears them.
int *inv_status_dev(pq_inv *h);
int inv_batch_max(const pq_inv *h);

// ---- CABARET model (pq_cabaret.hip)
int model_step_pre(pq_model *m, hipStream_t s);
int model_step_post(pq_model *m, hipStream_t s, bool last);
int model_step_invert(pq_model *m);  // the step's inversion, on the model's stream
hipStream_t model_stream(pq_model *m);
bool model_adaptive(const pq_model *m);
int model_scheme(const pq_model *m);  // PQ_SCHEME_*
// Device arrays and scalars of a model (the reference's full data arrays;
// CABARET-only arrays are null for the centred scheme and the second time
// level is null for CABARET).  The field arrays hold es-byte values (8:
// FP64, 4: FP32 storage); psi_av is FP64.
struct ModelView {
  int g, L, scheme, es;
  double *psi, *zeta, *rel, *zeta_old, *psi2, *zeta2;
  double *ui, *uj, *zfi, *zfj, *res, *res_beta, *psi_av;
  double dt, diag_time;
  hipStream_t stream;
};
ModelView model_view(pq_model *m);
// The START_PICKUP start once a checkpoint's fields are on the device.
int model_pickup(pq_model *m, double dt);
// zeta / psi sdata views (leading dimension ld) of es-byte values.
void model_views(pq_model *m, int *g, int *L, void **zeta_sdata, void **psi_sdata, long *ld, int *es);
// Host-side clocks of a model: model time and the time-mean weight.
struct ModelClock {
  double t, diag_t;
};
ModelClock model_clock(const pq_model *m);
void model_set_clock(pq_model *m, ModelClock c);
// What one captured step of the model contains: 0 no time-mean
// accumulation, 1 an accumulation in every step from the next one on, -1
// changes over the coming steps (averaging enabled, spin-up not yet over),
// so that a captured step would go stale.
int model_step_kind(const pq_model *m);
// Host-side bookkeeping of one replayed step (clocks as model_step_post).
void model_account_step(pq_model *m);
size_t model_device_bytes(const pq_model *m);
// The model inversion's status word on the device (see inv_status_dev).
int *model_status_dev(pq_model *m);

// ---- diagnostics (pq_diag.hip)
// psi of es-byte values (8: FP64, 4: FP32, read widened).
int diag_energy_sync(pq_diag *d, const void *psi, long ld1, long ld2, double *ekin, double *epot,
                     int es = 8);
int diag_energy_dev_es(pq_diag *d, const void *psi, long ld1, long ld2, double *out_dev, void *stream,
                       int es);
int psi_av_add_es(int g, int layers, double dt, const void *psi, long ld1, long ld2, double *psi_av,
                  long la1, long la2, void *stream, int es);

}  // namespace pq
