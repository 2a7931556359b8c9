/*
 * pq_oracle.c -- CPU FP64 restatement of pequod's Helmholtz inversion and
 * CABARET step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product library never
 * links or calls it.
 *
 * Each routine cites the reference routine it restates.  Loop order and the
 * order of floating-point operations inside each statement follow the
 * reference so that the restatement is a fair CPU stand-in; the FFT inside
 * the sine transform is this file's own mixed-radix Stockham FFT (radices
 * 4, 2, 3, 5 and a direct DFT for other factors), which computes the same
 * real DFT that FFTPACK's rfftf computes.
 *
 * Index macros below are 1-based, matching the Fortran source.
 */
#include "pq_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PQO_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (forward, exp(-2 pi i jk/M))                      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int m;
  int nrad;
  int rad[64];
  double complex *tw; /* exp(-2 pi i k / m), k = 0..m-1 */
  double complex *work;
} cfft_plan;

static int cfft_init(cfft_plan *p, int m) {
  p->m = m;
  p->nrad = 0;
  int r = m;
  /* factors of 4 first, then 2, 3, 5, then odd numbers */
  while (r % 4 == 0) { p->rad[p->nrad++] = 4; r /= 4; }
  while (r % 2 == 0) { p->rad[p->nrad++] = 2; r /= 2; }
  for (int f = 3; r > 1; f += 2) {
    while (r % f == 0) { p->rad[p->nrad++] = f; r /= f; }
  }
  p->tw = (double complex *)malloc(sizeof(double complex) * (size_t)(m > 0 ? m : 1));
  p->work = (double complex *)malloc(sizeof(double complex) * (size_t)(m > 0 ? m : 1));
  if (!p->tw || !p->work) return -1;
  for (int k = 0; k < m; ++k) {
    double a = -2.0 * PQO_PI * (double)k / (double)m;
    p->tw[k] = cos(a) + I * sin(a);
  }
  return 0;
}

static void cfft_free(cfft_plan *p) {
  free(p->tw);
  free(p->work);
}

/* Stockham auto-sort FFT; result is written back to x. */
static void cfft_exec(cfft_plan *p, double complex *x) {
  const int m = p->m;
  double complex *src = x, *dst = p->work;
  int ns = 1;
  double complex v[64];
  for (int s = 0; s < p->nrad; ++s) {
    const int R = p->rad[s];
    const int mr = m / R;
    for (int j = 0; j < mr; ++j) {
      const int k = j % ns;
      /* twiddle exp(-2 pi i r k / (ns R)) = tw[r k (m / (ns R))] */
      const int step = mr / ns;
      for (int r = 0; r < R; ++r) {
        double complex a = src[j + r * mr];
        v[r] = (r == 0 || k == 0) ? a : a * p->tw[((long)r * k * step) % m];
      }
      double complex out[64];
      if (R == 2) {
        out[0] = v[0] + v[1];
        out[1] = v[0] - v[1];
      } else if (R == 4) {
        double complex a0 = v[0] + v[2], a1 = v[0] - v[2];
        double complex b0 = v[1] + v[3], b1 = (v[1] - v[3]) * (-I);
        out[0] = a0 + b0;
        out[1] = a1 + b1;
        out[2] = a0 - b0;
        out[3] = a1 - b1;
      } else {
        for (int q = 0; q < R; ++q) {
          double complex acc = 0;
          for (int r = 0; r < R; ++r)
            acc += v[r] * p->tw[((long)r * q * mr) % m];
          out[q] = acc;
        }
      }
      const int base = (j / ns) * ns * R + k;
      for (int r = 0; r < R; ++r) dst[base + r * ns] = out[r];
    }
    ns *= R;
    double complex *t = src;
    src = dst;
    dst = t;
  }
  if (src != x) memcpy(x, src, sizeof(double complex) * (size_t)m);
}

/* ------------------------------------------------------------------------ */
/* DST-I with FFTPACK sint semantics (qg3p_helmholtz.F90:168-183,1370-1426)  */
/* ------------------------------------------------------------------------ */
struct pqo_sint {
  int n;            /* transform length                       */
  int nn;           /* n + 1, the length of the real DFT      */
  double *was;      /* 2 sin(k pi / (n+1)), k = 1..n/2        */
  double *y;        /* pre-twiddled sequence, length n + 1    */
  double *re, *im;  /* real DFT of y: Re/Im of Y_k, k=0..nn/2 */
  int half;         /* nn even: complex FFT of nn/2 on packed pairs */
  cfft_plan fft;
  double complex *z;
  double complex *post; /* exp(-2 pi i k / nn), k = 0..nn/2 */
};

pqo_sint *pqo_sint_create(int n) {
  if (n < 1) return NULL;
  pqo_sint *p = (pqo_sint *)calloc(1, sizeof(pqo_sint));
  if (!p) return NULL;
  p->n = n;
  p->nn = n + 1;
  int ns2 = n / 2;
  p->was = (double *)malloc(sizeof(double) * (size_t)(ns2 > 0 ? ns2 : 1));
  p->y = (double *)malloc(sizeof(double) * (size_t)p->nn);
  p->re = (double *)malloc(sizeof(double) * (size_t)(p->nn / 2 + 1));
  p->im = (double *)malloc(sizeof(double) * (size_t)(p->nn / 2 + 1));
  for (int k = 1; k <= ns2; ++k)
    p->was[k - 1] = 2.0 * sin((double)k * (PQO_PI / (double)p->nn));
  p->half = (p->nn % 2 == 0);
  int m = p->half ? p->nn / 2 : p->nn;
  cfft_init(&p->fft, m);
  p->z = (double complex *)malloc(sizeof(double complex) * (size_t)m);
  p->post = (double complex *)malloc(sizeof(double complex) * (size_t)(p->nn / 2 + 1));
  for (int k = 0; k <= p->nn / 2; ++k) {
    double a = -2.0 * PQO_PI * (double)k / (double)p->nn;
    p->post[k] = cos(a) + I * sin(a);
  }
  return p;
}

void pqo_sint_destroy(pqo_sint *p) {
  if (!p) return;
  free(p->was);
  free(p->y);
  free(p->re);
  free(p->im);
  free(p->z);
  free(p->post);
  cfft_free(&p->fft);
  free(p);
}

/* Real DFT of y[0..nn-1]: re[k] + i im[k] = sum_j y_j exp(-2 pi i jk/nn). */
static void real_dft(pqo_sint *p) {
  const int nn = p->nn;
  if (p->half) {
    const int m = nn / 2;
    for (int j = 0; j < m; ++j) p->z[j] = p->y[2 * j] + I * p->y[2 * j + 1];
    cfft_exec(&p->fft, p->z);
    for (int k = 0; k <= m; ++k) {
      double complex zk = p->z[k % m];
      double complex zc = conj(p->z[(m - k) % m]);
      double complex e = 0.5 * (zk + zc);
      double complex o = -0.5 * I * (zk - zc);
      double complex yk = e + p->post[k] * o;
      p->re[k] = creal(yk);
      p->im[k] = cimag(yk);
    }
  } else {
    for (int j = 0; j < nn; ++j) p->z[j] = p->y[j];
    cfft_exec(&p->fft, p->z);
    for (int k = 0; k <= nn / 2; ++k) {
      p->re[k] = creal(p->z[k]);
      p->im[k] = cimag(p->z[k]);
    }
  }
}

void pqo_sint_apply(pqo_sint *p, double *x) {
  const int n = p->n, nn = p->nn, ns2 = n / 2;
  if (n == 1) {
    x[0] = x[0] + x[0];
    return;
  }
  if (n == 2) {
    const double sqrt3 = 1.73205080756888;
    double xh = sqrt3 * (x[0] + x[1]);
    x[1] = sqrt3 * (x[0] - x[1]);
    x[0] = xh;
    return;
  }
  /* pre-twiddle (sint1, qg3p_helmholtz.F90:1404-1413); x is 1-based there */
  p->y[0] = 0.0;
  for (int k = 1; k <= ns2; ++k) {
    int kc = nn - k;
    double t1 = x[k - 1] - x[kc - 1];
    double t2 = p->was[k - 1] * (x[k - 1] + x[kc - 1]);
    p->y[k] = t1 + t2;
    p->y[kc] = t2 - t1;
  }
  if (n % 2 != 0) p->y[ns2 + 1] = 4.0 * x[ns2];
  real_dft(p);
  /* post-recurrence (qg3p_helmholtz.F90:1417-1423): FFTPACK's cc(2m) is
   * Re Y_m and cc(2m+1) is Im Y_m                                          */
  x[0] = 0.5 * p->re[0];
  for (int i = 3; i <= n; i += 2) {
    int m = (i - 1) / 2;
    x[i - 1] = x[i - 3] + p->re[m];
    x[i - 2] = -p->im[m];
  }
  if (n % 2 == 0) x[n - 1] = -p->im[n / 2];
}

/* ------------------------------------------------------------------------ */
/* Helmholtz solver (allocate_helm_c, allocate_h1d_c, helm, ge_local)        */
/* ------------------------------------------------------------------------ */
struct pqo_helm {
  int g, layers;
  pqo_sint *sint;
  double *ss;
  double *ge; /* ge(i, j, layer), i, j = 1..g, reciprocal Thomas factors */
  double *row;
};

#define GE(h, i, j, l) (h)->ge[((i)-1) + (size_t)(h)->g * (((j)-1) + (size_t)(h)->g * (l))]

pqo_helm *pqo_helm_create(int g, int layers, const double *ss) {
  if (g < 3 || layers < 1) return NULL;
  pqo_helm *h = (pqo_helm *)calloc(1, sizeof(pqo_helm));
  h->g = g;
  h->layers = layers;
  h->sint = pqo_sint_create(g - 2);
  h->ss = (double *)malloc(sizeof(double) * (size_t)layers);
  memcpy(h->ss, ss, sizeof(double) * (size_t)layers);
  h->ge = (double *)calloc((size_t)g * g * layers, sizeof(double));
  h->row = (double *)malloc(sizeof(double) * (size_t)g);
  /* allocate_h1d_c (qg3p_helmholtz.F90:790-809) on a single domain:
   * max_j = jj - 1; diag(i) = -2 - ss - 4 sin^2((i-1) pi / (2 (ii-1)))     */
  const int max_j = g - 1;
  const double norm_fact = PQO_PI / (2.0 * (double)(g - 1));
  double *diag = (double *)malloc(sizeof(double) * (size_t)g);
  for (int l = 0; l < layers; ++l) {
    for (int i = 2; i <= g - 1; ++i) {
      double s = sin((double)(i - 1) * norm_fact);
      diag[i - 1] = -2.0 - ss[l] - 4.0 * (s * s);
      GE(h, i, 2, l) = diag[i - 1];
    }
    for (int j = 3; j <= max_j; ++j)
      for (int i = 2; i <= g - 1; ++i)
        GE(h, i, j, l) = diag[i - 1] - 1.0 / GE(h, i, j - 1, l);
    for (int j = 2; j <= max_j; ++j)
      for (int i = 2; i <= g - 1; ++i) GE(h, i, j, l) = 1.0 / GE(h, i, j, l);
  }
  free(diag);
  return h;
}

void pqo_helm_destroy(pqo_helm *h) {
  if (!h) return;
  pqo_sint_destroy(h->sint);
  free(h->ss);
  free(h->ge);
  free(h->row);
  free(h);
}

/* helm (qg3p_helmholtz.F90:614-769) on a single domain with PH1D: the
 * h1d call reduces to ge_local (qg3p_helmholtz.F90:1028-1059).             */
void pqo_helm_solve(pqo_helm *h, double *zeta, int ld1, int ld2) {
  const int g = h->g, n = g - 2, max_j = g - 1;
  const double cff = 0.5 / (double)(g - 1);
#define Z(i, j, l) zeta[((i)-1) + (size_t)ld1 * (((j)-1) + (size_t)ld2 * (l))]
  for (int l = 0; l < h->layers; ++l) {
    /* forward transform along i for each interior j (lines 672-681) */
    for (int j = 2; j <= g - 1; ++j) {
      for (int i = 2; i <= g - 1; ++i) Z(i, j, l) = cff * Z(i, j, l);
      for (int i = 0; i < n; ++i) h->row[i] = Z(i + 2, j, l);
      pqo_sint_apply(h->sint, h->row);
      for (int i = 0; i < n; ++i) Z(i + 2, j, l) = h->row[i];
    }
    /* ge_local (lines 1041-1057): lower sweep, upper sweep, psi(:,1) = 0 */
    for (int j = 3; j <= max_j; ++j)
      for (int i = 2; i <= g - 1; ++i)
        Z(i, j, l) = Z(i, j, l) - Z(i, j - 1, l) * GE(h, i, j - 1, l);
    for (int i = 2; i <= g - 1; ++i) Z(i, max_j, l) = Z(i, max_j, l) * GE(h, i, max_j, l);
    for (int j = max_j - 1; j >= 2; --j)
      for (int i = 2; i <= g - 1; ++i)
        Z(i, j, l) = (Z(i, j, l) - Z(i, j + 1, l)) * GE(h, i, j, l);
    for (int i = 2; i <= g - 1; ++i) Z(i, 1, l) = 0.0;
    /* inverse transform, then zeta(ii, j) = 0 (lines 757-761) */
    for (int j = 2; j <= g - 1; ++j) {
      for (int i = 0; i < n; ++i) h->row[i] = Z(i + 2, j, l);
      pqo_sint_apply(h->sint, h->row);
      for (int i = 0; i < n; ++i) Z(i + 2, j, l) = h->row[i];
      Z(g, j, l) = 0.0;
    }
  }
#undef Z
}

/* ------------------------------------------------------------------------ */
/* Periodic-j Helmholtz solver (the channel)                                 */
/* ------------------------------------------------------------------------ */
struct pqo_phelm {
  int gi, gj, layers;
  pqo_sint *sint;
  double *eig; /* eig(i, j, layer), gi x gj x L */
  double *pb, *pc;
  double *row;
};

#define PIDX(h, i, j, l) (((i)-1) + (size_t)(h)->gi * (((j)-1) + (size_t)(h)->gj * (l)))

/* poinit_j (qg3p_helmholtz.F90:423-467) and periodic_coefficients
 * (qg3p_helmholtz.F90:471-565) on one domain periodic in j: l_i = 2..ii-1,
 * l_j = 1..jj, g_ii_int = ii - 1.                                          */
pqo_phelm *pqo_phelm_create(int gi, int gj, int layers, const double *ss) {
  if (gi < 3 || gj < 3 || layers < 1) return NULL;
  pqo_phelm *h = (pqo_phelm *)calloc(1, sizeof(pqo_phelm));
  h->gi = gi;
  h->gj = gj;
  h->layers = layers;
  h->sint = pqo_sint_create(gi - 2);
  const size_t sz = (size_t)gi * gj * layers;
  h->eig = (double *)calloc(sz, sizeof(double));
  h->pb = (double *)calloc(sz, sizeof(double));
  h->pc = (double *)calloc(sz, sizeof(double));
  h->row = (double *)malloc(sizeof(double) * (size_t)gi);
#define EIG(i, j, l) h->eig[PIDX(h, i, j, l)]
#define B(i, j, l) h->pb[PIDX(h, i, j, l)]
#define C(i, j, l) h->pc[PIDX(h, i, j, l)]
  const double cff = PQO_PI / (2.0 * (double)(gi - 1));
  for (int l = 0; l < layers; ++l)
    for (int i = 2; i <= gi - 1; ++i) {
      const double s = sin((double)(i - 1) * cff);
      const double d = -2.0 - ss[l] - 4.0 * (s * s);
      for (int j = 1; j <= gj; ++j) EIG(i, j, l) = d;
    }
  for (int l = 0; l < layers; ++l)
    for (int j = 3; j <= gj - 1; ++j)
      for (int i = 2; i <= gi - 1; ++i) EIG(i, j, l) = EIG(i, j, l) - 1.0 / EIG(i, j - 1, l);
  for (int l = 0; l < layers; ++l)
    for (int j = 2; j <= gj - 1; ++j)
      for (int i = 2; i <= gi - 1; ++i) EIG(i, j, l) = 1.0 / EIG(i, j, l);
  /* periodic_coefficients */
  for (int l = 0; l < layers; ++l) {
    for (int i = 2; i <= gi - 1; ++i) {
      const double s = sin((double)(i - 1) * cff);
      const double a = -2.0 - ss[l] - 4.0 * (s * s);
      B(i, 1, l) = a;
      B(i, 2, l) = 1.0;
      for (int j = 3; j <= gj - 1; ++j) B(i, j, l) = 0.0;
      B(i, gj, l) = 1.0;
      C(i, 1, l) = 1.0;
      for (int j = 2; j <= gj - 2; ++j) C(i, j, l) = 0.0;
      C(i, gj - 1, l) = 1.0;
      C(i, gj, l) = a;
    }
    for (int j = 3; j <= gj - 1; ++j)
      for (int i = 2; i <= gi - 1; ++i) {
        B(i, j, l) = B(i, j, l) - B(i, j - 1, l) * EIG(i, j - 1, l);
        C(i, j, l) = C(i, j, l) - C(i, j - 1, l) * EIG(i, j - 1, l);
      }
    for (int i = 2; i <= gi - 1; ++i) {
      B(i, gj - 1, l) = B(i, gj - 1, l) * EIG(i, gj - 1, l);
      C(i, gj - 1, l) = C(i, gj - 1, l) * EIG(i, gj - 1, l);
    }
    for (int j = gj - 2; j >= 2; --j)
      for (int i = 2; i <= gi - 1; ++i) {
        B(i, j, l) = (B(i, j, l) - B(i, j + 1, l)) * EIG(i, j, l);
        C(i, j, l) = (C(i, j, l) - C(i, j + 1, l)) * EIG(i, j, l);
      }
    for (int i = 2; i <= gi - 1; ++i) {
      B(i, 1, l) = B(i, 1, l) - B(i, 2, l);
      C(i, 1, l) = C(i, 1, l) - C(i, 2, l);
    }
    for (int i = 2; i <= gi - 1; ++i) {
      B(i, gj, l) = B(i, gj, l) - B(i, gj - 1, l);
      C(i, gj, l) = C(i, gj, l) - C(i, gj - 1, l);
    }
    for (int i = 2; i <= gi - 1; ++i) C(i, 1, l) = C(i, 1, l) / B(i, 1, l);
    for (int j = 2; j <= gj; ++j)
      for (int i = 2; i <= gi - 1; ++i) C(i, j, l) = C(i, j, l) - C(i, 1, l) * B(i, j, l);
  }
  return h;
}

void pqo_phelm_destroy(pqo_phelm *h) {
  if (!h) return;
  pqo_sint_destroy(h->sint);
  free(h->eig);
  free(h->pb);
  free(h->pc);
  free(h->row);
  free(h);
}

/* helm (qg3p_helmholtz.F90:614-769, single domain, no PH1D) with
 * periodic_ge (qg3p_helmholtz.F90:567-610).                                */
void pqo_phelm_solve(pqo_phelm *h, double *zeta, int ld1, int ld2) {
  const int gi = h->gi, gj = h->gj, n = gi - 2;
  const double cff = 0.5 / (double)(gi - 1);
#define Z(i, j, l) zeta[((i)-1) + (size_t)ld1 * (((j)-1) + (size_t)ld2 * (l))]
  for (int l = 0; l < h->layers; ++l) {
    for (int j = 1; j <= gj; ++j) {
      for (int i = 2; i <= gi - 1; ++i) Z(i, j, l) = cff * Z(i, j, l);
      for (int i = 0; i < n; ++i) h->row[i] = Z(i + 2, j, l);
      pqo_sint_apply(h->sint, h->row);
      for (int i = 0; i < n; ++i) Z(i + 2, j, l) = h->row[i];
    }
    for (int j = 3; j <= gj - 1; ++j)
      for (int i = 2; i <= gi - 1; ++i) Z(i, j, l) = Z(i, j, l) - Z(i, j - 1, l) * EIG(i, j - 1, l);
    for (int i = 2; i <= gi - 1; ++i) Z(i, gj - 1, l) = Z(i, gj - 1, l) * EIG(i, gj - 1, l);
    for (int j = gj - 2; j >= 2; --j)
      for (int i = 2; i <= gi - 1; ++i)
        Z(i, j, l) = (Z(i, j, l) - Z(i, j + 1, l)) * EIG(i, j, l);
    /* periodic_ge */
    for (int i = 2; i <= gi - 1; ++i) Z(i, 1, l) = Z(i, 1, l) - Z(i, 2, l);
    for (int i = 2; i <= gi - 1; ++i) Z(i, gj, l) = Z(i, gj, l) - Z(i, gj - 1, l);
    for (int i = 2; i <= gi - 1; ++i) Z(i, 1, l) = Z(i, 1, l) / B(i, 1, l);
    for (int j = 2; j <= gj; ++j)
      for (int i = 2; i <= gi - 1; ++i) Z(i, j, l) = Z(i, j, l) - Z(i, 1, l) * B(i, j, l);
    for (int i = 2; i <= gi - 1; ++i) Z(i, gj, l) = Z(i, gj, l) / C(i, gj, l);
    for (int j = gj - 1; j >= 1; --j)
      for (int i = 2; i <= gi - 1; ++i) Z(i, j, l) = Z(i, j, l) - Z(i, gj, l) * C(i, j, l);
    /* inverse transform, zeta(ii, j) = 0 */
    for (int j = 1; j <= gj; ++j) {
      for (int i = 0; i < n; ++i) h->row[i] = Z(i + 2, j, l);
      pqo_sint_apply(h->sint, h->row);
      for (int i = 0; i < n; ++i) Z(i + 2, j, l) = h->row[i];
      Z(gi, j, l) = 0.0;
    }
  }
#undef Z
}
#undef EIG
#undef B
#undef C

/* ------------------------------------------------------------------------ */
/* Channel PV inversion                                                      */
/* ------------------------------------------------------------------------ */
struct pqo_chan {
  int gi, gj, layers, bc_type;
  double alpha;
  pqo_phelm *helm;
  double *ss, *m2z, *z2m;
  double *corr;       /* gi gj L 2 */
  double *inv_corr_a; /* 2 2 (L-1) */
  double *phi;        /* gi gj L */
  double *psic;       /* 2 (L-1) */
};

/* zeta_boundary (qg3p_aux.F90:288-379): p_b = p_d = -1, l_j = 1..jj,
 * g_jj_int = jj (periodic); allsum is the identity.                        */
void pqo_zeta_boundary(int gi, int gj, int L, const double *ss, int bc_type,
                       double alpha, const double *psi, double *z_b, double *z_d) {
#define P(i, j, l) psi[((i)-1) + (size_t)gi * (((j)-1) + (size_t)gj * (l))]
  for (int l = 0; l < L; ++l) {
    z_b[l] = 0.0;
    z_d[l] = 0.0;
  }
  if (bc_type == 1) {
    const double cff = alpha / (alpha + 3.0);
    for (int l = 0; l < L; ++l)
      for (int j = 1; j <= gj; ++j)
        z_b[l] = z_b[l] + cff * (-3.5 * P(1, j, l) + 4.0 * P(2, j, l) - 0.5 * P(3, j, l));
    for (int l = 0; l < L; ++l)
      for (int j = 1; j <= gj; ++j)
        z_d[l] = z_d[l] + cff * (-3.5 * P(gi, j, l) + 4.0 * P(gi - 1, j, l) - 0.5 * P(gi - 2, j, l));
  } else if (bc_type == 2) {
    for (int l = 0; l < L; ++l)
      for (int j = 1; j <= gj; ++j) z_b[l] = z_b[l] + 2.0 * (P(2, j, l) - P(1, j, l));
    for (int l = 0; l < L; ++l)
      for (int j = 1; j <= gj; ++j) z_d[l] = z_d[l] + 2.0 * (P(gi - 1, j, l) - P(gi, j, l));
  }
  for (int l = 0; l < L; ++l)
    for (int j = 1; j <= gj; ++j) z_b[l] = z_b[l] - ss[l] * P(1, j, l);
  for (int l = 0; l < L; ++l)
    for (int j = 1; j <= gj; ++j) z_d[l] = z_d[l] - ss[l] * P(gi, j, l);
  for (int l = 0; l < L; ++l) {
    z_b[l] = z_b[l] / (double)gj;
    z_d[l] = z_d[l] / (double)gj;
  }
#undef P
}

/* add_laplace_bc_rhs_periodic_layer (qg3p_laplace_bc.F90:128-150) and
 * apply_dirichlet_bc_periodic_layer_data (qg3p_laplace_bc.F90:281-303). */
static void chan_laplace_rhs(int gi, int gj, int L, double bc_b, double bc_d, double *r) {
  for (int l = 0; l < L; ++l) {
    for (int j = 1; j <= gj; ++j) r[1 + (size_t)gi * ((j - 1) + (size_t)gj * l)] -= bc_b;
    for (int j = 1; j <= gj; ++j) r[(gi - 2) + (size_t)gi * ((j - 1) + (size_t)gj * l)] -= bc_d;
  }
}
static void chan_dirichlet(int gi, int gj, int L, double bc_b, double bc_d, double *z) {
  for (int l = 0; l < L; ++l) {
    for (int j = 1; j <= gj; ++j) z[0 + (size_t)gi * ((j - 1) + (size_t)gj * l)] = bc_b;
    for (int j = 1; j <= gj; ++j) z[(gi - 1) + (size_t)gi * ((j - 1) + (size_t)gj * l)] = bc_d;
  }
}

/* inverse of a 2 x 2 matrix (lapack_interfaces.F90:319-359 calls LAPACK
 * dgetrf/dgetri; the closed form agrees to rounding), column-major.        */
static void inv2(const double *a, double *x) {
  double a11 = a[0], a21 = a[1], a12 = a[2], a22 = a[3];
  const double det = a11 * a22 - a12 * a21;
  x[0] = a22 / det;
  x[1] = -a21 / det;
  x[2] = -a12 / det;
  x[3] = a11 / det;
}

/* allocate_helmholtz_cache + correction_function_momentum
 * (qg3p_aux.F90:152-193, 233-286).                                         */
pqo_chan *pqo_chan_create(int gi, int gj, int L, const double *ss, const double *m2z,
                          const double *z2m, int bc_type, double alpha) {
  pqo_chan *v = (pqo_chan *)calloc(1, sizeof(pqo_chan));
  v->gi = gi;
  v->gj = gj;
  v->layers = L;
  v->bc_type = bc_type;
  v->alpha = alpha;
  v->helm = pqo_phelm_create(gi, gj, L, ss);
  if (!v->helm) {
    free(v);
    return NULL;
  }
  v->ss = (double *)malloc(sizeof(double) * (size_t)L);
  v->m2z = (double *)malloc(sizeof(double) * (size_t)L * L);
  v->z2m = (double *)malloc(sizeof(double) * (size_t)L * L);
  memcpy(v->ss, ss, sizeof(double) * (size_t)L);
  memcpy(v->m2z, m2z, sizeof(double) * (size_t)L * L);
  memcpy(v->z2m, z2m, sizeof(double) * (size_t)L * L);
  const size_t sz = (size_t)gi * gj * L;
  v->corr = (double *)calloc(2 * sz, sizeof(double));
  v->phi = (double *)calloc(sz, sizeof(double));
  v->inv_corr_a = (double *)calloc((size_t)4 * (L > 1 ? L - 1 : 1), sizeof(double));
  v->psic = (double *)calloc((size_t)2 * (L > 1 ? L - 1 : 1), sizeof(double));
  double *c1 = v->corr, *c2 = v->corr + sz;
  chan_laplace_rhs(gi, gj, L, 1.0, 0.0, c1);
  pqo_phelm_solve(v->helm, c1, gi, gj);
  chan_dirichlet(gi, gj, L, 1.0, 0.0, c1);
  chan_laplace_rhs(gi, gj, L, 0.0, 1.0, c2);
  pqo_phelm_solve(v->helm, c2, gi, gj);
  chan_dirichlet(gi, gj, L, 0.0, 1.0, c2);
  double *zb1 = (double *)malloc(sizeof(double) * 4 * (size_t)L);
  double *zd1 = zb1 + L, *zb2 = zb1 + 2 * L, *zd2 = zb1 + 3 * L;
  pqo_zeta_boundary(gi, gj, L, ss, bc_type, alpha, c1, zb1, zd1);
  pqo_zeta_boundary(gi, gj, L, ss, bc_type, alpha, c2, zb2, zd2);
  for (int m = 1; m < L; ++m) {
    /* z_b(mode, wall, k): rows the walls (b, d), columns the functions */
    const double a[4] = {zb1[m], zd1[m], zb2[m], zd2[m]};
    inv2(a, v->inv_corr_a + 4 * (m - 1));
  }
  free(zb1);
  return v;
}

void pqo_chan_destroy(pqo_chan *v) {
  if (!v) return;
  pqo_phelm_destroy(v->helm);
  free(v->ss);
  free(v->m2z);
  free(v->z2m);
  free(v->corr);
  free(v->phi);
  free(v->inv_corr_a);
  free(v->psic);
  free(v);
}

const double *pqo_chan_corr(const pqo_chan *v) { return v->corr; }
const double *pqo_chan_inv_corr_a(const pqo_chan *v) { return v->inv_corr_a; }
const double *pqo_chan_psic(const pqo_chan *v) { return v->psic; }

/* pv_inversion (qg3p_aux.F90:452-479) with modes_to_z_correct_momentum
 * (qg3p_aux.F90:524-575).                                                  */
int pqo_chan_run(pqo_chan *v, const double *zeta, int ldz1, int ldz2, double *psi, int ldp1,
                 int ldp2) {
  const int gi = v->gi, gj = v->gj, L = v->layers;
  const size_t sz = (size_t)gi * gj * L, gg = (size_t)gi * gj;
#define ZT(i, j, l) zeta[((i)-1) + (size_t)ldz1 * (((j)-1) + (size_t)ldz2 * (l))]
#define PS(i, j, l) psi[((i)-1) + (size_t)ldp1 * (((j)-1) + (size_t)ldp2 * (l))]
#define PH(i, j, l) v->phi[((i)-1) + (size_t)gi * ((j)-1) + gg * (l)]
#define D(m, l) v->z2m[(m) + L * (l)]
#define M(l, m) v->m2z[(l) + L * (m)]
  for (int j = 1; j <= gj; ++j)
    for (int i = 1; i <= gi; ++i)
      for (int m = 0; m < L; ++m) {
        double s = 0.0;
        for (int l = 0; l < L; ++l) s += D(m, l) * ZT(i, j, l);
        PH(i, j, m) = s;
      }
  pqo_phelm_solve(v->helm, v->phi, gi, gj);
  chan_dirichlet(gi, gj, L, 0.0, 0.0, v->phi);
  double *zb = (double *)malloc(sizeof(double) * 2 * (size_t)L), *zd = zb + L;
  pqo_zeta_boundary(gi, gj, L, v->ss, v->bc_type, v->alpha, v->phi, zb, zd);
  int status = 0;
  for (int m = 1; m < L; ++m) {
    const double *ia = v->inv_corr_a + 4 * (m - 1);
    double p1 = ia[0] * zb[m] + ia[2] * zd[m];
    double p2 = ia[1] * zb[m] + ia[3] * zd[m];
    v->psic[2 * (m - 1)] = p1;
    v->psic[2 * (m - 1) + 1] = p2;
    if (p1 != p1 || p2 != p2) {
      status = 3;
      break;
    }
    if (fabs(p1) > 1.0e100 || fabs(p2) > 1.0e100) status = 4;
    const double *c1 = v->corr + gg * m, *c2 = v->corr + sz + gg * m;
    for (int j = 1; j <= gj; ++j)
      for (int i = 1; i <= gi; ++i) {
        const size_t k = ((size_t)i - 1) + (size_t)gi * (j - 1);
        PH(i, j, m) = PH(i, j, m) - p1 * c1[k] - p2 * c2[k];
      }
  }
  free(zb);
  if (status == 3) return status;
  /* modes_to_z_rank2 (qg3p_diagnostics.F90:570-587) */
  for (int j = 1; j <= gj; ++j)
    for (int i = 1; i <= gi; ++i)
      for (int l = 0; l < L; ++l) {
        double s = 0.0;
        for (int m = 0; m < L; ++m) s += M(l, m) * PH(i, j, m);
        PS(i, j, l) = s;
      }
  return status;
#undef ZT
#undef PS
#undef PH
#undef D
#undef M
}

/* ------------------------------------------------------------------------ */
/* Boundary helpers and mean                                                 */
/* ------------------------------------------------------------------------ */
/* apply_dirichlet_bc_layer (qg3p_laplace_bc.F90:193-224) */
void pqo_apply_dirichlet_bc(int g, int layers, double bc, double *zeta, int ld1,
                            int ld2) {
#define Z(i, j, l) zeta[((i)-1) + (size_t)ld1 * (((j)-1) + (size_t)ld2 * (l))]
  for (int l = 0; l < layers; ++l) {
    for (int i = 1; i <= g; ++i) Z(i, g, l) = bc;
    for (int i = 1; i <= g; ++i) Z(i, 1, l) = bc;
    for (int j = 1; j <= g; ++j) Z(1, j, l) = bc;
    for (int j = 1; j <= g; ++j) Z(g, j, l) = bc;
  }
#undef Z
}

/* add_laplace_bc_rhs_layer (qg3p_laplace_bc.F90:55-87) */
void pqo_add_laplace_bc_rhs(int g, int layers, double bc, double *rhs, int ld1,
                            int ld2) {
#define Z(i, j, l) rhs[((i)-1) + (size_t)ld1 * (((j)-1) + (size_t)ld2 * (l))]
  for (int l = 0; l < layers; ++l) {
    for (int i = 2; i <= g - 1; ++i) Z(i, g - 1, l) = Z(i, g - 1, l) - bc;
    for (int i = 2; i <= g - 1; ++i) Z(i, 2, l) = Z(i, 2, l) - bc;
    for (int j = 2; j <= g - 1; ++j) Z(2, j, l) = Z(2, j, l) - bc;
    for (int j = 2; j <= g - 1; ++j) Z(g - 1, j, l) = Z(g - 1, j, l) - bc;
  }
#undef Z
}

/* local_integral + mean_decomp_layer (qg3p_diagnostics.F90:669-720) */
double pqo_mean(int g, const double *psi, int ld1) {
#define P(i, j) psi[((i)-1) + (size_t)ld1 * ((j)-1)]
  double s = 0.0;
  for (int j = 2; j <= g - 1; ++j)
    for (int i = 2; i <= g - 1; ++i) s += P(i, j);
  double e = 0.0;
  for (int j = 2; j <= g - 1; ++j) e += P(1, j);
  s = s + 0.5 * e;
  s = s + 0.25 * P(1, 1);
  s = s + 0.25 * P(1, g);
  e = 0.0;
  for (int j = 2; j <= g - 1; ++j) e += P(g, j);
  s = s + 0.5 * e;
  s = s + 0.25 * P(g, 1);
  s = s + 0.25 * P(g, g);
  e = 0.0;
  for (int i = 2; i <= g - 1; ++i) e += P(i, 1);
  s = s + 0.5 * e;
  e = 0.0;
  for (int i = 2; i <= g - 1; ++i) e += P(i, g);
  s = s + 0.5 * e;
  return s / ((double)(g - 1) * (double)(g - 1));
#undef P
}

/* ------------------------------------------------------------------------ */
/* Stratification (qg3p_diagnostics.F90:210-422)                              */
/* ------------------------------------------------------------------------ */
/* strat_l2 (qg3p_diagnostics.F90:256-270): returns strat%s2(1), strat%s1(2) */
void pqo_strat_l2(double rd, double h1, double h2, double *s2_1, double *s1_2) {
  double fsq = (h1 * h2) / (2.0 * (rd * rd));
  *s2_1 = fsq * 2.0 / (h1 * (h1 + h2));
  *s1_2 = fsq * 2.0 / (h2 * (h1 + h2));
}

/* strat_l3 (qg3p_diagnostics.F90:272-321): closed-form inversion of the
 * 3-layer eigenproblem for (s2(1), s1(2), s2(2), s1(3)).                    */
void pqo_strat_l3(double rd1, double rd2, double h1, double h2, double h3,
                  double *s1, double *s21, double *s22, double *s3) {
  double lam1 = -1.0 / (rd1 * rd1), lam2 = -1.0 / (rd2 * rd2);
  double b1 = -(lam1 + lam2), c1 = lam1 * lam2;
  double w1 = 2.0 / (h1 * (h1 + h2)), w2 = 2.0 / (h2 * (h1 + h2));
  double w3 = 2.0 / (h2 * (h2 + h3)), w4 = 2.0 / (h3 * (h2 + h3));
  double l = w1 * w3 + w1 * w4 + w2 * w4;
  double a2 = l * (w3 + w4), b2 = -b1 * l, c2 = c1 * (w1 + w2);
  double f2 = (-b2 - sqrt(b2 * b2 - 4.0 * a2 * c2)) / (2.0 * a2);
  double f1 = (b1 - f2 * (w3 + w4)) / (w1 + w2);
  *s1 = f1 * w1;
  *s21 = f1 * w2;
  *s22 = f2 * w3;
  *s3 = f2 * w4;
}

/* Cyclic Jacobi eigen-solver for a symmetric L x L matrix (column-major). */
static void jacobi_sym(int L, double *a, double *v, double *w) {
  for (int i = 0; i < L * L; ++i) v[i] = 0.0;
  for (int i = 0; i < L; ++i) v[i + L * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < L; ++p)
      for (int q = p + 1; q < L; ++q) off += a[p + L * q] * a[p + L * q];
    if (off < 1e-300) break;
    for (int p = 0; p < L; ++p) {
      for (int q = p + 1; q < L; ++q) {
        double apq = a[p + L * q];
        if (fabs(apq) < 1e-300) continue;
        double app = a[p + L * p], aqq = a[q + L * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < L; ++k) {
          double akp = a[k + L * p], akq = a[k + L * q];
          a[k + L * p] = c * akp - s * akq;
          a[k + L * q] = s * akp + c * akq;
        }
        for (int k = 0; k < L; ++k) {
          double apk = a[p + L * k], aqk = a[q + L * k];
          a[p + L * k] = c * apk - s * aqk;
          a[q + L * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < L; ++k) {
          double vkp = v[k + L * p], vkq = v[k + L * q];
          v[k + L * p] = c * vkp - s * vkq;
          v[k + L * q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < L; ++i) w[i] = a[i + L * i];
}

/* Gauss-Jordan inverse with partial pivoting (column-major). */
static int invert(int L, const double *a, double *inv) {
  double m[64 * 128];
  if (L > 64) return -1;
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < 2 * L; ++j)
      m[i * 2 * L + j] = j < L ? a[i + L * j] : (j - L == i ? 1.0 : 0.0);
  for (int c = 0; c < L; ++c) {
    int piv = c;
    for (int r = c + 1; r < L; ++r)
      if (fabs(m[r * 2 * L + c]) > fabs(m[piv * 2 * L + c])) piv = r;
    if (m[piv * 2 * L + c] == 0.0) return -1;
    if (piv != c)
      for (int j = 0; j < 2 * L; ++j) {
        double t = m[c * 2 * L + j];
        m[c * 2 * L + j] = m[piv * 2 * L + j];
        m[piv * 2 * L + j] = t;
      }
    double d = m[c * 2 * L + c];
    for (int j = 0; j < 2 * L; ++j) m[c * 2 * L + j] /= d;
    for (int r = 0; r < L; ++r) {
      if (r == c) continue;
      double f = m[r * 2 * L + c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * L; ++j) m[r * 2 * L + j] -= f * m[c * 2 * L + j];
    }
  }
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) inv[i + L * j] = m[i * 2 * L + L + j];
  return 0;
}

/* s_matrix + s_conversion_matrices + sort_barotropic
 * (qg3p_diagnostics.F90:342-422).  The eigenproblem is solved through the
 * diagonal similarity that makes the tridiagonal S symmetric; the final
 * inversion is independent of eigenvector scale, sign and the order of the
 * baroclinic modes, which are sorted by ascending ss here.                  */
int pqo_strat_modes(int L, const double *s1, const double *s2, double *ss,
                    double *m2z, double *z2m) {
  if (L < 1 || L > 64) return -1;
  if (L == 1) {
    ss[0] = 0.0;
    m2z[0] = 1.0;
    z2m[0] = 1.0;
    return 0;
  }
  double S[64 * 64], B[64 * 64], V[64 * 64], w[64], d[64];
  for (int i = 0; i < L * L; ++i) S[i] = 0.0;
  for (int i = 1; i <= L; ++i) {
#define SM(r, c) S[((r)-1) + L * ((c)-1)]
    if (i > 1) {
      SM(i, i - 1) = s1[i - 1];
      if (i < L) {
        SM(i, i) = -(s1[i - 1] + s2[i - 1]);
        SM(i, i + 1) = s2[i - 1];
      } else {
        SM(i, i) = -s1[i - 1];
      }
    } else if (i < L) {
      SM(i, i) = -s2[i - 1];
      SM(i, i + 1) = s2[i - 1];
    }
#undef SM
  }
  d[0] = 1.0;
  for (int i = 0; i + 1 < L; ++i) {
    if (!(s2[i] > 0.0) || !(s1[i + 1] > 0.0)) return -2;
    d[i + 1] = d[i] * sqrt(s2[i] / s1[i + 1]);
  }
  for (int c = 0; c < L; ++c)
    for (int r = 0; r < L; ++r) B[r + L * c] = d[r] * S[r + L * c] / d[c];
  /* exact symmetry of the off-diagonal */
  for (int i = 0; i + 1 < L; ++i) {
    double o = sqrt(s2[i] * s1[i + 1]);
    B[i + L * (i + 1)] = o;
    B[(i + 1) + L * i] = o;
  }
  jacobi_sym(L, B, V, w);
  int order[64];
  double sv[64];
  for (int i = 0; i < L; ++i) {
    order[i] = i;
    sv[i] = -w[i];
  }
  /* barotropic = smallest ss; the rest ascending */
  for (int a = 0; a < L; ++a)
    for (int b = a + 1; b < L; ++b)
      if (sv[order[b]] < sv[order[a]]) {
        int t = order[a];
        order[a] = order[b];
        order[b] = t;
      }
  for (int c = 0; c < L; ++c) {
    int e = order[c];
    ss[c] = c == 0 ? 0.0 : sv[e];
    for (int r = 0; r < L; ++r)
      m2z[r + L * c] = c == 0 ? 1.0 : V[r + L * e] / d[r];
  }
  return invert(L, m2z, z2m);
}

/* ------------------------------------------------------------------------ */
/* PV inversion (qg3p_aux.F90:152-230, 452-521; qg3p_diagnostics.F90:534-635) */
/* ------------------------------------------------------------------------ */
struct pqo_inv {
  int g, layers;
  pqo_helm *helm;
  double *m2z, *z2m, *ss;
  double *corr; /* g g L */
  double *phi;  /* g g L */
  double *psic;
};

pqo_inv *pqo_inv_create(int g, int L, const double *ss, const double *m2z,
                        const double *z2m) {
  pqo_inv *v = (pqo_inv *)calloc(1, sizeof(pqo_inv));
  v->g = g;
  v->layers = L;
  v->helm = pqo_helm_create(g, L, ss);
  if (!v->helm) {
    free(v);
    return NULL;
  }
  v->m2z = (double *)malloc(sizeof(double) * (size_t)L * L);
  v->z2m = (double *)malloc(sizeof(double) * (size_t)L * L);
  v->ss = (double *)malloc(sizeof(double) * (size_t)L);
  v->psic = (double *)calloc((size_t)L, sizeof(double));
  memcpy(v->m2z, m2z, sizeof(double) * (size_t)L * L);
  memcpy(v->z2m, z2m, sizeof(double) * (size_t)L * L);
  memcpy(v->ss, ss, sizeof(double) * (size_t)L);
  size_t sz = (size_t)g * g * L;
  v->corr = (double *)malloc(sizeof(double) * sz);
  v->phi = (double *)calloc(sz, sizeof(double));
  /* correction_function_mass (qg3p_aux.F90:196-230) */
  for (size_t k = 0; k < sz; ++k) v->corr[k] = 1.0;
  pqo_helm_solve(v->helm, v->corr, g, g);
  pqo_apply_dirichlet_bc(g, L, 0.0, v->corr, g, g);
  for (int l = 0; l < L; ++l) {
    double *c = v->corr + (size_t)g * g * l;
    double s = 0.0;
    for (int j = 2; j <= g - 1; ++j)
      for (int i = 2; i <= g - 1; ++i) s += c[(i - 1) + (size_t)g * (j - 1)];
    double mean_corr = s / ((double)(g - 1) * (double)(g - 1));
    for (size_t k = 0; k < (size_t)g * g; ++k)
      c[k] = (1.0 + ss[l] * c[k]) / (1.0 + ss[l] * mean_corr);
  }
  return v;
}

void pqo_inv_destroy(pqo_inv *v) {
  if (!v) return;
  pqo_helm_destroy(v->helm);
  free(v->m2z);
  free(v->z2m);
  free(v->ss);
  free(v->corr);
  free(v->phi);
  free(v->psic);
  free(v);
}

const double *pqo_inv_corr(const pqo_inv *v) { return v->corr; }
const double *pqo_inv_psic(const pqo_inv *v) { return v->psic; }

int pqo_inv_run(pqo_inv *v, const double *zeta, int ldz1, int ldz2, double *psi,
                int ldp1, int ldp2) {
  const int g = v->g, L = v->layers;
  const size_t gg = (size_t)g * g;
#define ZT(i, j, l) zeta[((i)-1) + (size_t)ldz1 * (((j)-1) + (size_t)ldz2 * (l))]
#define PS(i, j, l) psi[((i)-1) + (size_t)ldp1 * (((j)-1) + (size_t)ldp2 * (l))]
#define PH(i, j, l) v->phi[((i)-1) + (size_t)g * ((j)-1) + gg * (l)]
#define CR(i, j, l) v->corr[((i)-1) + (size_t)g * ((j)-1) + gg * (l)]
#define D(m, l) v->z2m[(m) + L * (l)]
#define M(l, m) v->m2z[(l) + L * (m)]
  /* z_to_modes_rank2 (qg3p_diagnostics.F90:534-553) */
  for (int j = 1; j <= g; ++j)
    for (int i = 1; i <= g; ++i)
      for (int m = 0; m < L; ++m) {
        double s = 0.0;
        for (int l = 0; l < L; ++l) s += D(m, l) * ZT(i, j, l);
        PH(i, j, m) = s;
      }
  pqo_helm_solve(v->helm, v->phi, g, g);
  pqo_apply_dirichlet_bc(g, L, 0.0, v->phi, g, g);
  /* modes_to_z_correct_mass (qg3p_aux.F90:482-521) */
  int status = 0;
  for (int m = 0; m < L; ++m) {
    v->psic[m] = pqo_mean(g, v->phi + gg * m, g);
    if (v->psic[m] != v->psic[m]) status = 3;
  }
  if (status) return status;
  for (int m = 0; m < L; ++m)
    if (fabs(v->psic[m]) > 1.0e100) status = 4;
  /* modes_to_z_correction_btcu (qg3p_diagnostics.F90:615-635) */
  for (int j = 1; j <= g; ++j)
    for (int i = 1; i <= g; ++i)
      for (int l = 0; l < L; ++l) {
        double s = 0.0;
        for (int m = 1; m < L; ++m)
          s += M(l, m) * (PH(i, j, m) - v->psic[m] * CR(i, j, m));
        PS(i, j, l) = PH(i, j, 0) - v->psic[0] + s;
      }
  return status;
#undef ZT
#undef PS
#undef PH
#undef CR
#undef D
#undef M
}

/* ------------------------------------------------------------------------ */
/* CABARET stages (qg3p_aux.F90:917-1642), single non-periodic domain.        */
/* ------------------------------------------------------------------------ */
#define ND(a, i, j, l) (a)[((i)-1) + (size_t)(g + 2) * (((j)-1) + (size_t)(g + 2) * (l))]
#define FI(a, i, j, l) (a)[((i)-1) + (size_t)(g + 3) * (((j)-1) + (size_t)(g + 2) * (l))]
#define FJ(a, i, j, l) (a)[((i)-1) + (size_t)(g + 2) * (((j)-1) + (size_t)(g + 3) * (l))]
#define RS(a, i, j, l) (a)[((i)-1) + (size_t)g * (((j)-1) + (size_t)g * (l))]

/* cabaret_cfl (qg3p_aux.F90:920-943): fa_i = fa_j = 2..g+1 */
double pqo_cabaret_cfl(int g, int L, double dt, const double *ui, const double *uj) {
  double mx = 0.0;
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g + 1; ++i) mx = fmax(mx, fabs(FI(ui, i, j, l)));
  double my = 0.0;
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g + 1; ++j)
      for (int i = 3; i <= g; ++i) my = fmax(my, fabs(FJ(uj, i, j, l)));
  return fmax(mx, my) * dt;
}

/* cabaret_adapt_dt (qg3p_aux.F90:946-990); fne(max_u, 0) uses the
 * tolerance of test_tools.F90:86-94 (10 * max(spacing, epsilon)).           */
void pqo_cabaret_adapt_dt(int g, int L, double target_cfl, double max_dt,
                          double max_dt_increase, const double *ui,
                          const double *uj, double *dt, double *tweight_a,
                          double *tweight_u) {
  double max_u = pqo_cabaret_cfl(g, L, 1.0, ui, uj);
  double old_dt = *dt;
  double a = fabs(max_u);
  double spacing = a == 0.0 ? 2.2250738585072014e-308 : nextafter(a, INFINITY) - a;
  double tol = 10.0 * fmax(spacing, 2.220446049250313e-16);
  if (fabs(max_u - 0.0) > tol) {
    double d = target_cfl / max_u;
    d = fmin(d, fmin(old_dt * max_dt_increase, max_dt));
    *dt = d;
  }
  *tweight_a = 0.5 * *dt / old_dt;
  *tweight_u = *dt / (old_dt + *dt);
}

/* cabaret_predictor (qg3p_aux.F90:993-1020): f_i = f_j = 3..g */
void pqo_cabaret_predictor(int g, int L, double dt, const double *zeta_old,
                           const double *res, double *zeta) {
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g; ++i)
        ND(zeta, i, j, l) = ND(zeta_old, i, j, l) + 0.5 * dt * RS(res, i - 1, j - 1, l);
}

/* cabaret_beta (qg3p_aux.F90:1023-1061): l_i = l_j = 2..g-1 in sdata */
void pqo_cabaret_beta(int g, int L, double dt, double tweight_a,
                      const double *beta, const double *ui, double *zeta,
                      double *res_beta) {
  for (int l = 0; l < L; ++l)
    for (int j = 2; j <= g - 1; ++j)
      for (int i = 2; i <= g - 1; ++i) {
        double rb = dt * (0.5 * (FI(ui, i + 2, j + 1, l) + FI(ui, i + 1, j + 1, l)) * beta[l]);
        ND(zeta, i + 1, j + 1, l) = ND(zeta, i + 1, j + 1, l) +
                                    (1.0 + tweight_a) * rb -
                                    tweight_a * RS(res_beta, i, j, l);
        RS(res_beta, i, j, l) = rb;
      }
}

/* cabaret_forcing (qg3p_aux.F90:1065-1164) with WIND_FORCING */
void pqo_cabaret_forcing(int g, int L, double dt, const double *force_wind,
                         double viscosity, double visc_bot, const double *rel,
                         double *zeta) {
#define FW(i, j) force_wind[((i)-1) + (size_t)g * ((j)-1)]
#define LAP(l)                                                                 \
  (-4.0 * ND(rel, i, j, l) + ND(rel, i - 1, j, l) + ND(rel, i + 1, j, l) +     \
   ND(rel, i, j - 1, l) + ND(rel, i, j + 1, l))
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g; ++i) {
        double rv;
        if (L == 1)
          rv = dt * (FW(i - 1, j - 1) - visc_bot * ND(rel, i, j, 0) + viscosity * LAP(0));
        else if (l == 0)
          rv = dt * (FW(i - 1, j - 1) + viscosity * LAP(0));
        else if (l < L - 1)
          rv = dt * viscosity * LAP(l);
        else
          rv = dt * (-visc_bot * ND(rel, i, j, l) + viscosity * LAP(l));
        ND(zeta, i, j, l) = ND(zeta, i, j, l) + rv;
      }
#undef LAP
#undef FW
}

/* psi_edge (qg3p_aux.F90:1206-1212, 1279-1285) */
#define PSI_EDGE(i, j, l)                                                       \
  (0.25 * (ND(psi, i, j, l) + ND(psi, i, j - 1, l) + ND(psi, i - 1, j, l) +    \
           ND(psi, i - 1, j - 1, l)))

/* cabaret_cell_boundaries_velocity (qg3p_aux.F90:1168-1231) */
void pqo_cabaret_cell_boundaries_velocity(int g, int L, const double *psi,
                                          double *ui, double *uj) {
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g + 1; ++i)
        FI(ui, i, j, l) = -PSI_EDGE(i, j + 1, l) + PSI_EDGE(i, j, l);
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g + 1; ++j)
      for (int i = 3; i <= g; ++i)
        FJ(uj, i, j, l) = +PSI_EDGE(i + 1, j, l) - PSI_EDGE(i, j, l);
}

/* cabaret_cell_boundaries_time_integrated (qg3p_aux.F90:1235-1313) */
void pqo_cabaret_cell_boundaries(int g, int L, double tw, const double *psi,
                                 double *ui, double *uj, double *uimo,
                                 double *ujmo) {
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g + 1; ++i) {
        double uim = -PSI_EDGE(i, j + 1, l) + PSI_EDGE(i, j, l);
        FI(ui, i, j, l) = (1.0 + tw) * uim - tw * FI(uimo, i, j, l);
        FI(uimo, i, j, l) = uim;
      }
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g + 1; ++j)
      for (int i = 3; i <= g; ++i) {
        double ujm = +PSI_EDGE(i + 1, j, l) - PSI_EDGE(i, j, l);
        FJ(uj, i, j, l) = (1.0 + tw) * ujm - tw * FJ(ujmo, i, j, l);
        FJ(ujmo, i, j, l) = ujm;
      }
}

static inline double max3(double a, double b, double c) {
  double m = a > b ? a : b;
  return m > c ? m : c;
}
static inline double min3(double a, double b, double c) {
  double m = a < b ? a : b;
  return m < c ? m : c;
}

/* cabaret_extrapolator with CABARET_LIMITER (qg3p_aux.F90:1317-1544).
 * fb_i_min + 1 = 3 and fb_i_max = g + 1 on a single domain.                 */
void pqo_cabaret_extrapolator(int g, int L, double dt, const double *zeta,
                              const double *zeta_old, const double *ui,
                              const double *uj, double *zeta_fi,
                              double *zeta_fj) {
  size_t nfi = (size_t)(g + 3) * (g + 2) * L, nfj = (size_t)(g + 2) * (g + 3) * L;
  double *fin = (double *)calloc(nfi, sizeof(double));
  double *fjn = (double *)calloc(nfj, sizeof(double));
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g + 1; ++i) {
        double u1 = FI(ui, i, j, l), fi1 = FI(zeta_fi, i, j, l);
        double u2, fi2, v0, psi_n, psi_o, d_f;
        if (u1 >= 0.0) {
          if (i > 3) {
            u2 = FI(ui, i - 1, j, l);
            fi2 = FI(zeta_fi, i - 1, j, l);
            v0 = (u1 + u2) / 2.0;
            psi_n = ND(zeta, i - 1, j, l);
            psi_o = ND(zeta_old, i - 1, j, l);
            d_f = -(psi_n - psi_o) * 2.0 - dt * v0 * (fi1 - fi2);
          } else {
            fi2 = ND(zeta, i - 1, j, l);
            psi_n = ND(zeta, i - 1, j, l);
            psi_o = ND(zeta_old, i - 1, j, l);
            d_f = -(psi_n - psi_o);
          }
        } else {
          if (i < g + 1) {
            u2 = FI(ui, i + 1, j, l);
            fi2 = FI(zeta_fi, i + 1, j, l);
            v0 = (u1 + u2) / 2.0;
            psi_n = ND(zeta, i, j, l);
            psi_o = ND(zeta_old, i, j, l);
            d_f = -(psi_n - psi_o) * 2.0 - dt * v0 * (fi2 - fi1);
          } else {
            fi2 = ND(zeta, i, j, l);
            psi_n = ND(zeta, i, j, l);
            psi_o = ND(zeta_old, i, j, l);
            d_f = -(psi_n - psi_o);
          }
        }
        double f = 2.0 * psi_n - fi2;
        double fmx = max3(fi1, fi2, psi_n) - d_f;
        double fmn = min3(fi1, fi2, psi_n) - d_f;
        if (f > fmx) f = fmx;
        if (f < fmn) f = fmn;
        FI(fin, i, j, l) = f;
      }
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g + 1; ++j)
      for (int i = 3; i <= g; ++i) {
        double u1 = FJ(uj, i, j, l), fi1 = FJ(zeta_fj, i, j, l);
        double u2, fi2, v0, psi_n, psi_o, d_f;
        if (u1 >= 0.0) {
          if (j > 3) {
            u2 = FJ(uj, i, j - 1, l);
            fi2 = FJ(zeta_fj, i, j - 1, l);
            v0 = (u1 + u2) / 2.0;
            psi_n = ND(zeta, i, j - 1, l);
            psi_o = ND(zeta_old, i, j - 1, l);
            d_f = -(psi_n - psi_o) * 2.0 - dt * v0 * (fi1 - fi2);
          } else {
            fi2 = ND(zeta, i, j - 1, l);
            psi_n = ND(zeta, i, j - 1, l);
            psi_o = ND(zeta_old, i, j - 1, l);
            d_f = -(psi_n - psi_o);
          }
        } else {
          if (j < g + 1) {
            u2 = FJ(uj, i, j + 1, l);
            fi2 = FJ(zeta_fj, i, j + 1, l);
            v0 = (u1 + u2) / 2.0;
            psi_n = ND(zeta, i, j, l);
            psi_o = ND(zeta_old, i, j, l);
            d_f = -(psi_n - psi_o) * 2.0 - dt * v0 * (fi2 - fi1);
          } else {
            fi2 = ND(zeta, i, j, l);
            psi_n = ND(zeta, i, j, l);
            psi_o = ND(zeta_old, i, j, l);
            d_f = -(psi_n - psi_o);
          }
        }
        double f = 2.0 * psi_n - fi2;
        double fmx = max3(fi1, fi2, psi_n) - d_f;
        double fmn = min3(fi1, fi2, psi_n) - d_f;
        if (f > fmx) f = fmx;
        if (f < fmn) f = fmn;
        FJ(fjn, i, j, l) = f;
      }
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g + 1; ++i) FI(zeta_fi, i, j, l) = FI(fin, i, j, l);
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g + 1; ++j)
      for (int i = 3; i <= g; ++i) FJ(zeta_fj, i, j, l) = FJ(fjn, i, j, l);
  free(fin);
  free(fjn);
}

/* cabaret_flux (qg3p_aux.F90:1547-1594) */
void pqo_cabaret_flux(int g, int L, const double *ui, const double *uj,
                      const double *zeta_fi, const double *zeta_fj,
                      double *flux_i, double *flux_j) {
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 2; i <= g + 2; ++i)
        FI(flux_i, i, j, l) = FI(ui, i, j, l) * FI(zeta_fi, i, j, l);
  for (int l = 0; l < L; ++l)
    for (int j = 2; j <= g + 2; ++j)
      for (int i = 3; i <= g; ++i)
        FJ(flux_j, i, j, l) = FJ(uj, i, j, l) * FJ(zeta_fj, i, j, l);
}

/* cabaret_corrector (qg3p_aux.F90:1597-1642) */
void pqo_cabaret_corrector(int g, int L, double dt, const double *flux_i,
                           const double *flux_j, double *res, double *zeta) {
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g; ++i) {
        RS(res, i - 1, j - 1, l) =
            -(FI(flux_i, i + 1, j, l) - FI(flux_i, i, j, l) + FJ(flux_j, i, j + 1, l) -
              FJ(flux_j, i, j, l));
        ND(zeta, i, j, l) = ND(zeta, i, j, l) + 0.5 * dt * RS(res, i - 1, j - 1, l);
      }
}

/* apply_rel_bc (qg3p_aux.F90:779-915): fa ranges 2..g+1 */
void pqo_apply_rel_bc(int g, int L, int bc_type, double alpha, const double *psi,
                      double *rel) {
  if (bc_type == PQO_BC_PARTIAL_SLIP) {
    double cff = alpha / (alpha + 3.0);
    for (int l = 0; l < L; ++l)
      for (int j = 2; j <= g + 1; ++j)
        ND(rel, 2, j, l) = cff * (-3.5 * ND(psi, 2, j, l) + 4.0 * ND(psi, 3, j, l) -
                                  0.5 * ND(psi, 4, j, l));
    for (int l = 0; l < L; ++l)
      for (int j = 2; j <= g + 1; ++j)
        ND(rel, g + 1, j, l) = cff * (-3.5 * ND(psi, g + 1, j, l) + 4.0 * ND(psi, g, j, l) -
                                      0.5 * ND(psi, g - 1, j, l));
    for (int l = 0; l < L; ++l)
      for (int i = 2; i <= g + 1; ++i)
        ND(rel, i, 2, l) = cff * (-3.5 * ND(psi, i, 2, l) + 4.0 * ND(psi, i, 3, l) -
                                  0.5 * ND(psi, i, 4, l));
    for (int l = 0; l < L; ++l)
      for (int i = 2; i <= g + 1; ++i)
        ND(rel, i, g + 1, l) = cff * (-3.5 * ND(psi, i, g + 1, l) + 4.0 * ND(psi, i, g, l) -
                                      0.5 * ND(psi, i, g - 1, l));
  } else if (bc_type == PQO_BC_NO_SLIP) {
    for (int l = 0; l < L; ++l)
      for (int j = 2; j <= g + 1; ++j)
        ND(rel, 2, j, l) = 2.0 * (ND(psi, 3, j, l) - ND(psi, 2, j, l));
    for (int l = 0; l < L; ++l)
      for (int j = 2; j <= g + 1; ++j)
        ND(rel, g + 1, j, l) = 2.0 * (ND(psi, g, j, l) - ND(psi, g + 1, j, l));
    for (int l = 0; l < L; ++l)
      for (int i = 2; i <= g + 1; ++i)
        ND(rel, i, 2, l) = 2.0 * (ND(psi, i, 3, l) - ND(psi, i, 2, l));
    for (int l = 0; l < L; ++l)
      for (int i = 2; i <= g + 1; ++i)
        ND(rel, i, g + 1, l) = 2.0 * (ND(psi, i, g, l) - ND(psi, i, g + 1, l));
  } else {
    for (int l = 0; l < L; ++l)
      for (int j = 2; j <= g + 1; ++j) {
        ND(rel, 2, j, l) = 0.0;
        ND(rel, g + 1, j, l) = 0.0;
      }
    for (int l = 0; l < L; ++l)
      for (int i = 2; i <= g + 1; ++i) {
        ND(rel, i, 2, l) = 0.0;
        ND(rel, i, g + 1, l) = 0.0;
      }
  }
}

/* rel_from_zeta_and_psi (qg3p_aux.F90:738-776) */
void pqo_rel_from_zeta_and_psi(int g, int L, const double *s1, const double *s2,
                               int bc_type, double alpha, const double *zeta,
                               const double *psi, double *rel) {
  if (L == 1) {
    memcpy(rel, zeta, sizeof(double) * (size_t)(g + 2) * (g + 2));
  } else {
    for (int j = 2; j <= g + 1; ++j)
      for (int i = 2; i <= g + 1; ++i)
        ND(rel, i, j, 0) = ND(zeta, i, j, 0) + s2[0] * (ND(psi, i, j, 0) - ND(psi, i, j, 1));
    for (int l = 1; l < L - 1; ++l)
      for (int j = 2; j <= g + 1; ++j)
        for (int i = 2; i <= g + 1; ++i)
          ND(rel, i, j, l) = ND(zeta, i, j, l) +
                             s1[l] * (ND(psi, i, j, l) - ND(psi, i, j, l - 1)) +
                             s2[l] * (ND(psi, i, j, l) - ND(psi, i, j, l + 1));
    for (int j = 2; j <= g + 1; ++j)
      for (int i = 2; i <= g + 1; ++i)
        ND(rel, i, j, L - 1) = ND(zeta, i, j, L - 1) +
                               s1[L - 1] * (ND(psi, i, j, L - 1) - ND(psi, i, j, L - 2));
  }
  pqo_apply_rel_bc(g, L, bc_type, alpha, psi, rel);
}

/* update_zeta_bc (qg3p_aux.F90:643-734) */
void pqo_update_zeta_bc(int g, int L, const double *s1, const double *s2,
                        const double *rel, const double *psi, double *zeta) {
#define ZBC(i, j)                                                               \
  do {                                                                          \
    if (L == 1) {                                                               \
      ND(zeta, i, j, 0) = ND(rel, i, j, 0);                                     \
    } else {                                                                    \
      ND(zeta, i, j, 0) = ND(rel, i, j, 0) - s2[0] * (ND(psi, i, j, 0) - ND(psi, i, j, 1)); \
      ND(zeta, i, j, L - 1) = ND(rel, i, j, L - 1) -                            \
                              s1[L - 1] * (ND(psi, i, j, L - 1) - ND(psi, i, j, L - 2)); \
    }                                                                           \
  } while (0)
#define ZBCM(i, j, l)                                                           \
  ND(zeta, i, j, l) = ND(rel, i, j, l) - s1[l] * (ND(psi, i, j, l) - ND(psi, i, j, l - 1)) - \
                      s2[l] * (ND(psi, i, j, l) - ND(psi, i, j, l + 1))
  /* left (j = 2), right (j = g+1), top (i = 2), bottom (i = g+1) */
  for (int i = 2; i <= g + 1; ++i) ZBC(i, 2);
  for (int l = 1; l < L - 1; ++l)
    for (int i = 2; i <= g + 1; ++i) ZBCM(i, 2, l);
  for (int i = 2; i <= g + 1; ++i) ZBC(i, g + 1);
  for (int l = 1; l < L - 1; ++l)
    for (int i = 2; i <= g + 1; ++i) ZBCM(i, g + 1, l);
  for (int j = 2; j <= g + 1; ++j) ZBC(2, j);
  for (int l = 1; l < L - 1; ++l)
    for (int j = 2; j <= g + 1; ++j) ZBCM(2, j, l);
  for (int j = 2; j <= g + 1; ++j) ZBC(g + 1, j);
  for (int l = 1; l < L - 1; ++l)
    for (int j = 2; j <= g + 1; ++j) ZBCM(g + 1, j, l);
#undef ZBC
#undef ZBCM
}

/* Wind field (qg3.F90:253-279); tau already non-dimensionalised (qg3.F90:164,
 * 236). */
void pqo_wind_field(int g, double tau, double wind_shift, double wind_tilt,
                    double wind_asym, double *fw) {
  double a_wind = -1.0 * wind_asym, b_wind = 1.0 / wind_asym;
  for (int j = 1; j <= g; ++j)
    for (int i = 1; i <= g; ++i) {
      double x = (double)(j - 1) / (double)(g - 1) - 0.5;
      double y = -(double)(i - 1) / (double)(g - 1) + 0.5;
      double y_mid = wind_shift + wind_tilt * x;
      double f;
      if (y < y_mid)
        f = a_wind * sin(PQO_PI * (0.5 + y) / (0.5 + y_mid));
      else
        f = b_wind * sin(PQO_PI * (y - y_mid) / (0.5 - y_mid));
      fw[(i - 1) + (size_t)g * (j - 1)] = tau * 2.0 * PQO_PI * f;
    }
}

/* energy (qg3p_diagnostics.F90:778-863) on one domain with every neighbour
 * absent (p_a = p_b = p_c = p_d = -1): l_i = 2..ii-1 in sdata, the cell
 * loop over data indices f_i_min..fb_i_max = 3..ii+1 (qg3p_base.F90:855-941);
 * allsum is the identity.  Statement and loop order as the reference.      */
void pqo_energy(int gi, int gj, int L, const double *h, const double *s2,
                const double *psi, double *ekin, double *epot) {
#define PD(i, j, k) psi[((i)-1) + (size_t)(gi + 2) * (((j)-1) + (size_t)(gj + 2) * (k))]
#define PS(i, j, k) PD((i) + 1, (j) + 1, k)
#define SQ(x) ((x) * (x))
  double h_total = 0.0;
  for (int k = 0; k < L; ++k) h_total += h[k];
  double fac[64];
  for (int k = 0; k < L - 1; ++k) fac[k] = 0.5 * s2[k] * h[k] / h_total;
  double ep = 0.0;
  for (int k = 0; k < L - 1; ++k) {
    for (int j = 2; j <= gj - 1; ++j)
      for (int i = 2; i <= gi - 1; ++i) ep = ep + fac[k] * SQ(PS(i, j, k) - PS(i, j, k + 1));
    /* left boundary, its corners */
    for (int i = 2; i <= gi - 1; ++i) ep = ep + 0.5 * fac[k] * SQ(PS(i, 1, k) - PS(i, 1, k + 1));
    ep = ep + 0.25 * fac[k] * SQ(PS(1, 1, k) - PS(1, 1, k + 1));
    ep = ep + 0.25 * fac[k] * SQ(PS(gi, 1, k) - PS(gi, 1, k + 1));
    /* right boundary, its corners */
    for (int i = 2; i <= gi - 1; ++i) ep = ep + 0.5 * fac[k] * SQ(PS(i, gj, k) - PS(i, gj, k + 1));
    ep = ep + 0.25 * fac[k] * SQ(PS(1, gj, k) - PS(1, gj, k + 1));
    ep = ep + 0.25 * fac[k] * SQ(PS(gi, gj, k) - PS(gi, gj, k + 1));
    /* top and bottom boundaries */
    for (int j = 2; j <= gj - 1; ++j) ep = ep + 0.5 * fac[k] * SQ(PS(1, j, k) - PS(1, j, k + 1));
    for (int j = 2; j <= gj - 1; ++j) ep = ep + 0.5 * fac[k] * SQ(PS(gi, j, k) - PS(gi, j, k + 1));
  }
  for (int k = 0; k < L; ++k) {
    double e = 0.0;
    for (int j = 3; j <= gj + 1; ++j)
      for (int i = 3; i <= gi + 1; ++i)
        e = e + 0.25 * (SQ(PD(i - 1, j, k) + PD(i, j, k) - PD(i - 1, j - 1, k) - PD(i, j - 1, k)) +
                        SQ(PD(i, j - 1, k) + PD(i, j, k) - PD(i - 1, j - 1, k) - PD(i - 1, j, k)));
    ekin[k] = e;
  }
  const double cff = 1.0 / (double)((gi - 1) * (gj - 1));
  *epot = cff * ep;
  for (int k = 0; k < L; ++k) ekin[k] = 0.5 * cff * (h[k] / h_total) * ekin[k];
#undef SQ
#undef PS
#undef PD
}

/* the time-mean accumulation of the driver (qg3.F90:672-681) */
void pqo_psi_av_add(int gi, int gj, int L, double dt, const double *psi, double *psi_av) {
  for (int l = 0; l < L; ++l)
    for (int j = 1; j <= gj; ++j)
      for (int i = 1; i <= gi; ++i) {
        double *o = &psi_av[(i - 1) + (size_t)gi * ((j - 1) + (size_t)gj * l)];
        const double p = psi[i + (size_t)(gi + 2) * (j + (size_t)(gj + 2) * l)];
        *o = *o + dt * p;
      }
}

/* zeta_centred (qg3p_aux.F90:2027-2202) / zeta_tendency with
 * zeta_forward_euler (qg3p_aux.F90:1789-2024): dst = src + T * F on
 * f_i, f_j = 3..G, with beta_y(i) = beta ((G + 1) / 2 - (i - 1))
 * (qg3.F90:302-307) and the bottom friction on rel_p
 * (CENTRED_BOTTOM_FRICTION undefined); one layer keeps the reference's
 * visc_bot inside the viscosity bracket.                                    */
void pqo_zeta_step(int g, int L, double T, const double *beta, const double *s1,
                   const double *s2, const double *force_wind, double viscosity,
                   double visc_bot, const double *psi, const double *rel,
                   const double *rel_p, const double *src, double *dst) {
#define FW(i, j) force_wind[((i)-1) + (size_t)g * ((j)-1)]
#define BY(i, l) (beta[l] * (((double)g + 1.0) / 2.0 - (double)((i)-1)))
#define P(i, j, l) ND(psi, i, j, l)
#define R(i, j, l) ND(rel, i, j, l)
#define LAPP(l)                                                                \
  (-4.0 * ND(rel_p, i, j, l) + ND(rel_p, i - 1, j, l) + ND(rel_p, i + 1, j, l) + \
   ND(rel_p, i, j - 1, l) + ND(rel_p, i, j + 1, l))
  for (int l = 0; l < L; ++l)
    for (int j = 3; j <= g; ++j)
      for (int i = 3; i <= g; ++i) {
        /* q(ii, jj): beta_y + rel + the coupling to the neighbouring layers */
#define Q(ii, jj)                                                              \
  (L == 1 ? BY(ii, l) + R(ii, jj, l)                                           \
   : l == 0 ? BY(ii, l) + R(ii, jj, l) + s2[l] * P(ii, jj, l + 1)              \
   : l == L - 1 ? BY(ii, l) + R(ii, jj, l) + s1[l] * P(ii, jj, l - 1)          \
                : BY(ii, l) + R(ii, jj, l) + s1[l] * P(ii, jj, l - 1) + s2[l] * P(ii, jj, l + 1))
        const double jac = Q(i, j + 1) * (P(i - 1, j + 1, l) - P(i + 1, j + 1, l)) -
                           Q(i, j - 1) * (P(i - 1, j - 1, l) - P(i + 1, j - 1, l)) -
                           Q(i - 1, j) * (P(i - 1, j + 1, l) - P(i - 1, j - 1, l)) +
                           Q(i + 1, j) * (P(i + 1, j + 1, l) - P(i + 1, j - 1, l));
#undef Q
        double f;
        if (L == 1)
          f = 0.25 * jac + FW(i - 1, j - 1) + viscosity * (LAPP(0) - visc_bot * ND(rel_p, i, j, 0));
        else if (l == 0)
          f = 0.25 * jac + FW(i - 1, j - 1) + viscosity * LAPP(0);
        else if (l < L - 1)
          f = 0.25 * jac + viscosity * LAPP(l);
        else
          f = 0.25 * jac + viscosity * LAPP(l) - visc_bot * ND(rel_p, i, j, l);
        ND(dst, i, j, l) = ND(src, i, j, l) + T * f;
      }
#undef LAPP
#undef R
#undef P
#undef BY
#undef FW
}

/* robert_asselin_filter_rank3 with RA_DOUBLE_FILTER (qg3p_aux.F90:44-62,
 * 1716-1755) over fa_i, fa_j = 2..G+1 (qg3p_aux.F90:1758-1786).            */
void pqo_ra_filter(int g, int L, double *phi_l, double *phi_m, const double *phi_old_l,
                   double *phi_old_m) {
  const double ra_nu = 0.02, ra_alpha = 0.5;
  const double w21 = 0.5 * ra_nu * ra_alpha, w22 = 1.0 - 2.0 * w21;
  const double w11 = -0.5 * ra_nu * (1.0 - ra_alpha), w12 = -2.0 * w11, w13 = 1.0 + w11;
  for (int l = 0; l < L; ++l)
    for (int j = 2; j <= g + 1; ++j)
      for (int i = 2; i <= g + 1; ++i) {
        const double lphi_l = ND(phi_l, i, j, l);
        ND(phi_l, i, j, l) = w11 * ND(phi_old_l, i, j, l) + w12 * ND(phi_m, i, j, l) + w13 * lphi_l;
        ND(phi_m, i, j, l) = w21 * (ND(phi_old_l, i, j, l) + lphi_l) + w22 * ND(phi_m, i, j, l);
        ND(phi_old_m, i, j, l) = ND(phi_m, i, j, l);
      }
}

static int lf_invert_level(pqo_lf_state *s, const pqo_params *p, pqo_inv *inv, int k) {
  const int g = s->g, L = s->layers, ld = g + 2;
  int st = pqo_inv_run(inv, s->zeta[k] + 1 + ld, ld, ld, s->psi[k] + 1 + ld, ld, ld);
  if (st == 3) return st;
  pqo_rel_from_zeta_and_psi(g, L, p->s1, p->s2, p->bc_type, p->alpha, s->zeta[k], s->psi[k],
                            s->rel[k]);
  pqo_update_zeta_bc(g, L, p->s1, p->s2, s->rel[k], s->psi[k], s->zeta[k]);
  return st;
}

/* qg3.F90:383-391 (level 1), 449-452, 456-490 (FORWARD_EULER_START).      */
int pqo_lf_init(pqo_lf_state *s, const pqo_params *p, pqo_inv *inv) {
  const int g = s->g, L = s->layers;
  const size_t n = (size_t)(g + 2) * (g + 2) * L;
  int st = lf_invert_level(s, p, inv, 0);
  if (st == 3) return st;
  memcpy(s->zeta_old[0], s->zeta[0], sizeof(double) * n);
  pqo_zeta_step(g, L, s->dt, p->beta, p->s1, p->s2, s->force_wind, p->viscosity, p->visc_bot,
                s->psi[0], s->rel[0], s->rel[0], s->zeta[0], s->zeta[1]);
  int st2 = lf_invert_level(s, p, inv, 1);
  if (st2 == 3) return st2;
  memcpy(s->zeta_old[1], s->zeta[1], sizeof(double) * n);
  s->model_time = 0.0 + s->dt;
  s->leapfrog = 1;
  return st > st2 ? st : st2;
}

/* One pass of DOLEAPFROG in the time loop (qg3.F90:612-660).              */
int pqo_lf_step(pqo_lf_state *s, const pqo_params *p, pqo_inv *inv) {
  const int g = s->g, L = s->layers;
  const int l = s->leapfrog - 1, m = 1 - l;
  const double T = 2.0 * s->dt;
  pqo_zeta_step(g, L, T, p->beta, p->s1, p->s2, s->force_wind, p->viscosity, p->visc_bot,
                s->psi[m], s->rel[m], s->rel[l], s->zeta[l], s->zeta[l]);
  pqo_ra_filter(g, L, s->zeta[l], s->zeta[m], s->zeta_old[l], s->zeta_old[m]);
  int st = lf_invert_level(s, p, inv, l);
  s->model_time = s->model_time + s->dt;
  s->leapfrog = s->leapfrog % 2 + 1;
  return st;
}

/* One time step in the order of the CABARET time loop (qg3.F90:590-967). */
int pqo_step(pqo_state *s, const pqo_params *p, pqo_inv *inv, int do_corrector) {
  const int g = s->g, L = s->layers;
  if (p->adapt_dt)
    pqo_cabaret_adapt_dt(g, L, p->target_cfl, p->max_dt, p->max_dt_increase, s->ui,
                         s->uj, &s->dt, &s->tweight_a, &s->tweight_u);
  pqo_cabaret_predictor(g, L, s->dt, s->zeta_old, s->res, s->zeta);
  pqo_cabaret_beta(g, L, s->dt, s->tweight_a, p->beta, s->ui, s->zeta, s->res_beta);
  const int ld = g + 2;
  int st = pqo_inv_run(inv, s->zeta + 1 + ld, ld, ld, s->psi + 1 + ld, ld, ld);
  if (st == 3) return st;
  pqo_rel_from_zeta_and_psi(g, L, p->s1, p->s2, p->bc_type, p->alpha, s->zeta, s->psi,
                            s->rel);
  pqo_update_zeta_bc(g, L, p->s1, p->s2, s->rel, s->psi, s->zeta);
  s->model_time += 0.5 * s->dt;
  if (s->psi_av && s->model_time > s->max_spin) {
    pqo_psi_av_add(g, g, L, s->dt, s->psi, s->psi_av);
    s->diag_time = s->diag_time + s->dt;
  }
  pqo_cabaret_forcing(g, L, s->dt, s->force_wind, p->viscosity, p->visc_bot, s->rel,
                      s->zeta);
  pqo_cabaret_cell_boundaries(g, L, s->tweight_u, s->psi, s->ui, s->uj, s->uimo, s->ujmo);
  pqo_cabaret_extrapolator(g, L, s->dt, s->zeta, s->zeta_old, s->ui, s->uj, s->zeta_fi,
                           s->zeta_fj);
  pqo_cabaret_flux(g, L, s->ui, s->uj, s->zeta_fi, s->zeta_fj, s->flux_i, s->flux_j);
  if (do_corrector) {
    pqo_cabaret_corrector(g, L, s->dt, s->flux_i, s->flux_j, s->res, s->zeta);
    memcpy(s->zeta_old, s->zeta, sizeof(double) * (size_t)(g + 2) * (g + 2) * L);
    s->model_time += 0.5 * s->dt;
  }
  return st;
}
