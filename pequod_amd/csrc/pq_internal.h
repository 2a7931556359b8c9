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

// ---- CABARET model (pq_cabaret.hip)
int model_step_pre(pq_model *m, hipStream_t s);
int model_step_post(pq_model *m, hipStream_t s, bool last);
int model_step_invert(pq_model *m);  // the step's inversion, on the model's stream
hipStream_t model_stream(pq_model *m);
bool model_adaptive(const pq_model *m);
void model_views(pq_model *m, int *g, int *L, double **zeta_sdata, double **psi_sdata, long *ld);
double model_time(const pq_model *m);
void model_set_time(pq_model *m, double t);
void model_account_step(pq_model *m);
size_t model_device_bytes(const pq_model *m);

}  // namespace pq
