// pq_io.cpp -- the reference's model output and checkpoint files for a
// device-resident model: snapshot (qgL_G_s), CABARET supplementary
// (qgL_G_s_supp) and time-mean (qgL_G_s_av) data sets, each a plain-text
// meta-data file and a binary data file, and the START_PICKUP start from a
// checkpoint (qg3p_io.F90:202-317, 405-732; qg3.F90:316-364, 821-934;
// manual/manual.tex:1610-1790).
//
// Binary data files hold, in the order the meta-data lists them, one record
// per variable: a 4-byte integer byte count, the values as 8-byte doubles
// in column-major order (i fastest), and the byte count again (the
// reference's write_unformatted / read_unformatted on a stream file,
// qg3p_io.F90:572-732).  A serial run's file names carry no process suffix:
// base[.index] for the meta-data, base[.index].d for the data
// (qg3p_io.F90:382-423).  Fields are copied from / to the device in full and
// sliced on the host.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pq_common.h"
#include "pq_internal.h"

using namespace pq;

namespace {

// Fortran's E26.17E3 edit descriptor (real_8_chr, qg3p_params.F90:12) as
// gfortran writes it: optional sign, "0.", 17 digits, "E", exponent sign and
// 3 digits, right-justified in 26 columns.
std::string fmt_e26_17e3(double x) {
  char mant[64];
  if (x == 0.0) {
    std::snprintf(mant, sizeof(mant), "%s0.%017dE+000", std::signbit(x) ? "-" : "", 0);
  } else {
    // 17 significant digits d.dddd...e+k is 0.dddd... x 10^(k+1)
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.16e", std::fabs(x));
    char digits[18];
    digits[0] = buf[0];
    std::memcpy(digits + 1, buf + 2, 16);
    digits[17] = '\0';
    const int k = std::atoi(std::strchr(buf, 'e') + 1) + 1;
    std::snprintf(mant, sizeof(mant), "%s0.%sE%c%03d", x < 0 ? "-" : "", digits, k < 0 ? '-' : '+',
                  k < 0 ? -k : k);
  }
  char out[64];
  std::snprintf(out, sizeof(out), "%26s", mant);
  return out;
}

std::string base_name(int L, int g, const char *kind) {
  return "qg" + std::to_string(L) + "_" + std::to_string(g) + "_" + kind;
}

std::string join(const char *dir, const std::string &name) {
  std::string d = dir ? dir : ".";
  if (!d.empty() && d.back() != '/') d += '/';
  return d + name;
}

// write_parallel_header (qg3p_io.F90:426-438) for one serial domain, then
// the variable lines.
int write_meta(const std::string &path, int g, int L, double l_scale, double t_scale,
               const std::vector<std::string> &vars) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) {
    pq_set_error("Failed to open header file: %s (%s)", path.c_str(), std::strerror(errno));
    return PQ_ERR_ARG;
  }
  std::fprintf(f, "1 1 1 0 1 1\n");
  std::fprintf(f, "%d %d %d\n", g, g, L);
  std::fprintf(f, "0 0 %d %d\n", g, g);
  std::fprintf(f, "%s %s\n", fmt_e26_17e3(l_scale).c_str(), fmt_e26_17e3(t_scale).c_str());
  for (const auto &v : vars) std::fprintf(f, "%s\n", v.c_str());
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    pq_set_error("Failed to write header file: %s", path.c_str());
    return PQ_ERR_ARG;
  }
  return PQ_OK;
}

struct Out {
  FILE *f = nullptr;
  bool ok = true;
  // one record: byte count, data, byte count
  void rec(const double *p, size_t n) {
    const int32_t s = (int32_t)(n * 8);
    ok = ok && std::fwrite(&s, 4, 1, f) == 1 && std::fwrite(p, 8, n, f) == n && std::fwrite(&s, 4, 1, f) == 1;
  }
  void rec(double x) { rec(&x, 1); }
};

struct In {
  FILE *f = nullptr;
  std::string path;
  // one record of exactly n doubles
  int rec(double *p, size_t n) {
    int32_t s1 = 0, s2 = 0;
    if (std::fread(&s1, 4, 1, f) != 1 || s1 != (int32_t)(n * 8) || std::fread(p, 8, n, f) != n ||
        std::fread(&s2, 4, 1, f) != 1 || s2 != s1) {
      pq_set_error("%s: record of %zu values expected (found %d bytes)", path.c_str(), n, (int)s1);
      return PQ_ERR_ARG;
    }
    return PQ_OK;
  }
};

// Host copy of a device array of es-byte values (8: FP64, 4: FP32, widened).
int d2h(std::vector<double> &h, const double *dev, size_t n, int es = 8) {
  h.resize(n);
  if (es == 8) {
    PQ_HIP(hipMemcpy(h.data(), dev, n * sizeof(double), hipMemcpyDeviceToHost));
    return PQ_OK;
  }
  std::vector<float> f(n);
  PQ_HIP(hipMemcpy(f.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost));
  std::copy(f.begin(), f.end(), h.begin());
  return PQ_OK;
}

// Device copy of host values, rounded to FP32 for es = 4.
int h2d(double *dev, const std::vector<double> &h, int es) {
  if (es == 8) {
    PQ_HIP(hipMemcpy(dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return PQ_OK;
  }
  std::vector<float> f(h.begin(), h.end());
  PQ_HIP(hipMemcpy(dev, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
  return PQ_OK;
}

// Block [i0, i0 + ni) x [j0, j0 + nj) (0-based) of layer l of an array with
// leading dimensions (d1, d2).
void slice(const std::vector<double> &a, int d1, int d2, int l, int i0, int ni, int j0, int nj,
           std::vector<double> &out) {
  out.resize((size_t)ni * nj);
  for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i)
      out[(size_t)i + (size_t)ni * j] = a[(size_t)(i0 + i) + (size_t)d1 * ((size_t)(j0 + j) + (size_t)d2 * l)];
}

void unslice(std::vector<double> &a, int d1, int d2, int l, int i0, int ni, int j0, int nj,
             const double *in) {
  for (int j = 0; j < nj; ++j)
    for (int i = 0; i < ni; ++i)
      a[(size_t)(i0 + i) + (size_t)d1 * ((size_t)(j0 + j) + (size_t)d2 * l)] = in[(size_t)i + (size_t)ni * j];
}

// fne of the reference's test_tools (qg3p_io.F90:292-295 uses it to compare
// the supplementary file's times): |a - b| > 10 max(spacing(a), epsilon).
bool fne(double a, double b) {
  const double sp = a == 0.0 ? DBL_TRUE_MIN : std::nextafter(std::fabs(a), INFINITY) - std::fabs(a);
  return std::fabs(a - b) > 10.0 * std::fmax(sp, DBL_EPSILON);
}

}  // namespace

extern "C" {

int pq_model_write_snapshot(pq_model *m, const char *dir, int index, double model_time_day,
                            double total_time_day, double l_scale, double t_scale, int supp) {
  PQ_CHECK_ARG(m != nullptr, "pq_model_write_snapshot: null model");
  const ModelView v = model_view(m);
  if (v.g == 0) return PQ_ERR_HIP;  // model_view set the error
  const int g = v.g, L = v.L, G2 = g + 2;
  const bool lf = v.scheme == PQ_SCHEME_LEAPFROG;
  PQ_CHECK_ARG(!supp || !lf, "pq_model_write_snapshot: supplementary data are CABARET data");
  PQ_HIP(hipStreamSynchronize(v.stream));
  const std::string sfx = index >= 0 ? "." + std::to_string(index) : std::string();
  // snapshot: write_s_header (qg3p_io.F90:441-471), data (qg3.F90:836-856)
  {
    std::vector<std::string> vars = {"model_time_day", "total_time_day"};
    for (const char *nm : {"psi", "zeta"})
      for (int l = 1; l <= L; ++l)
        vars.push_back(std::string(nm) + std::to_string(l) + (lf ? "  ii jj leapfrog" : "  ii jj"));
    const std::string path = join(dir, base_name(L, g, "s") + sfx);
    int rc = write_meta(path, g, L, l_scale, t_scale, vars);
    if (rc) return rc;
    std::vector<double> a1, a2, s1, s2;
    Out o;
    o.f = std::fopen((path + ".d").c_str(), "wb");
    PQ_CHECK_ARG(o.f != nullptr, "Failed to open output data file: %s.d", path.c_str());
    o.rec(model_time_day);
    o.rec(total_time_day);
    const size_t n = (size_t)G2 * G2 * L;
    for (int which = 0; which < 2; ++which) {
      rc = d2h(a1, which == 0 ? v.psi : v.zeta, n, v.es);
      if (!rc && lf) rc = d2h(a2, which == 0 ? v.psi2 : v.zeta2, n, v.es);
      if (rc) {
        std::fclose(o.f);
        return rc;
      }
      for (int l = 0; l < L; ++l) {
        slice(a1, G2, G2, l, 1, g, 1, g, s1);
        if (lf) {
          // write_unformatted_rank2_rank2: both levels in one record
          slice(a2, G2, G2, l, 1, g, 1, g, s2);
          s1.insert(s1.end(), s2.begin(), s2.end());
        }
        o.rec(s1.data(), s1.size());
      }
    }
    const bool ok = o.ok && std::fclose(o.f) == 0;
    PQ_CHECK_ARG(ok, "Failed to write output data file: %s.d", path.c_str());
  }
  if (!supp) return PQ_OK;
  // supplementary data: write_s_supp_header (qg3p_io.F90:474-511), data
  // (qg3.F90:858-885)
  std::vector<std::string> vars = {"model_time_day", "total_time_day", "dt"};
  const char *names[6] = {"ui", "uj", "zeta_fi", "zeta_fj", "res", "res_beta"};
  const char *shapes[6] = {"  flux_i_ii flux_i_jj", "  flux_j_ii flux_j_jj", "  flux_i_ii flux_i_jj",
                           "  flux_j_ii flux_j_jj", "  ii jj", "  ii jj"};
  for (int k = 0; k < 6; ++k)
    for (int l = 1; l <= L; ++l) vars.push_back(std::string(names[k]) + std::to_string(l) + shapes[k]);
  const std::string path = join(dir, base_name(L, g, "s_supp") + sfx);
  int rc = write_meta(path, g, L, l_scale, t_scale, vars);
  if (rc) return rc;
  Out o;
  o.f = std::fopen((path + ".d").c_str(), "wb");
  PQ_CHECK_ARG(o.f != nullptr, "Failed to open output data file: %s.d", path.c_str());
  o.rec(model_time_day);
  o.rec(total_time_day);
  o.rec(v.dt);
  // i-face arrays (G+3, G+2): data(3:G+1, 3:G); j-face arrays (G+2, G+3):
  // data(3:G, 3:G+1); residuals (G, G)
  const double *dev[6] = {v.ui, v.uj, v.zfi, v.zfj, v.res, v.res_beta};
  std::vector<double> a, s;
  for (int k = 0; k < 6 && o.ok; ++k) {
    const bool fi = k == 0 || k == 2, fj = k == 1 || k == 3;
    const int d1 = fi ? g + 3 : fj ? g + 2 : g, d2 = fi ? g + 2 : fj ? g + 3 : g;
    rc = d2h(a, dev[k], (size_t)d1 * d2 * L, v.es);
    if (rc) {
      std::fclose(o.f);
      return rc;
    }
    for (int l = 0; l < L; ++l) {
      if (fi) slice(a, d1, d2, l, 2, g - 1, 2, g - 2, s);
      else if (fj) slice(a, d1, d2, l, 2, g - 2, 2, g - 1, s);
      else slice(a, d1, d2, l, 0, g, 0, g, s);
      o.rec(s.data(), s.size());
    }
  }
  const bool ok = o.ok && std::fclose(o.f) == 0;
  PQ_CHECK_ARG(ok, "Failed to write output data file: %s.d", path.c_str());
  return PQ_OK;
}

int pq_model_write_mean(pq_model *m, const char *dir, double model_time_day, double total_time_day,
                        double l_scale, double t_scale) {
  PQ_CHECK_ARG(m != nullptr, "pq_model_write_mean: null model");
  const ModelView v = model_view(m);
  if (v.g == 0) return PQ_ERR_HIP;  // model_view set the error
  PQ_CHECK_ARG(v.psi_av && v.diag_time > 0.0, "pq_model_write_mean: no time-mean samples yet");
  const int g = v.g, L = v.L;
  PQ_HIP(hipStreamSynchronize(v.stream));
  // write_s_av_header (qg3p_io.F90:514-534), data (qg3.F90:886-898)
  std::vector<std::string> vars = {"model_time_day", "total_time_day"};
  for (int l = 1; l <= L; ++l) vars.push_back("psi" + std::to_string(l) + "_av  ii jj");
  const std::string path = join(dir, base_name(L, g, "s_av"));
  int rc = write_meta(path, g, L, l_scale, t_scale, vars);
  if (rc) return rc;
  std::vector<double> a;
  rc = d2h(a, v.psi_av, (size_t)g * g * L);
  if (rc) return rc;
  const double cff = 1.0 / v.diag_time;
  for (double &x : a) x = x * cff;
  Out o;
  o.f = std::fopen((path + ".d").c_str(), "wb");
  PQ_CHECK_ARG(o.f != nullptr, "Failed to open output data file: %s.d", path.c_str());
  o.rec(model_time_day);
  o.rec(total_time_day);
  for (int l = 0; l < L; ++l) o.rec(a.data() + (size_t)g * g * l, (size_t)g * g);
  const bool ok = o.ok && std::fclose(o.f) == 0;
  PQ_CHECK_ARG(ok, "Failed to write output data file: %s.d", path.c_str());
  return PQ_OK;
}

int pq_model_read_checkpoint(pq_model *m, const char *dir, double *model_time_day,
                             double *total_time_day) {
  PQ_CHECK_ARG(m != nullptr, "pq_model_read_checkpoint: null model");
  const ModelView v = model_view(m);
  if (v.g == 0) return PQ_ERR_HIP;  // model_view set the error
  const int g = v.g, L = v.L, G2 = g + 2;
  const bool lf = v.scheme == PQ_SCHEME_LEAPFROG;
  PQ_HIP(hipStreamSynchronize(v.stream));
  // read_parallel_checkpoint (qg3p_io.F90:202-317)
  In in;
  in.path = join(dir, base_name(L, g, "s") + ".d");
  in.f = std::fopen(in.path.c_str(), "rb");
  PQ_CHECK_ARG(in.f != nullptr, "File %s not found", in.path.c_str());
  double mtd = 0.0, ttd = 0.0;
  int rc = in.rec(&mtd, 1);
  if (!rc) rc = in.rec(&ttd, 1);
  const size_t n = (size_t)G2 * G2 * L;
  std::vector<double> f[4];  // psi, zeta (levels 1 and 2)
  for (auto &x : f) x.assign(n, 0.0);
  std::vector<double> r((size_t)g * g * (lf ? 2 : 1));
  for (int which = 0; which < 2 && !rc; ++which)
    for (int l = 0; l < L && !rc; ++l) {
      rc = in.rec(r.data(), r.size());
      if (rc) break;
      unslice(f[which], G2, G2, l, 1, g, 1, g, r.data());
      if (lf) unslice(f[2 + which], G2, G2, l, 1, g, 1, g, r.data() + (size_t)g * g);
    }
  std::fclose(in.f);
  if (rc) return rc;
  double dt = v.dt;
  std::vector<double> sup[6];
  if (!lf) {
    In is;
    is.path = join(dir, base_name(L, g, "s_supp") + ".d");
    is.f = std::fopen(is.path.c_str(), "rb");
    PQ_CHECK_ARG(is.f != nullptr, "File %s not found", is.path.c_str());
    double mts = 0.0, tts = 0.0;
    rc = is.rec(&mts, 1);
    if (!rc) rc = is.rec(&tts, 1);
    if (!rc && (fne(mts, mtd) || fne(tts, ttd))) {
      pq_set_error("Input times do not match in CABARET checkpoint data");
      rc = PQ_ERR_ARG;
    }
    if (!rc) rc = is.rec(&dt, 1);
    std::vector<double> s;
    for (int k = 0; k < 6 && !rc; ++k) {
      const bool fi = k == 0 || k == 2, fj = k == 1 || k == 3;
      const int d1 = fi ? g + 3 : fj ? g + 2 : g, d2 = fi ? g + 2 : fj ? g + 3 : g;
      const int ni = fi ? g - 1 : fj ? g - 2 : g, nj = fi ? g - 2 : fj ? g - 1 : g;
      sup[k].assign((size_t)d1 * d2 * L, 0.0);
      s.resize((size_t)ni * nj);
      for (int l = 0; l < L && !rc; ++l) {
        rc = is.rec(s.data(), s.size());
        if (!rc) unslice(sup[k], d1, d2, l, (fi || fj) ? 2 : 0, ni, (fi || fj) ? 2 : 0, nj, s.data());
      }
    }
    std::fclose(is.f);
    if (rc) return rc;
    PQ_CHECK_ARG(dt > 0.0 && std::isfinite(dt), "pq_model_read_checkpoint: bad dt %g", dt);
  }
  // device state: the fields read, zero elsewhere
  double *dst[4] = {v.psi, v.zeta, v.psi2, v.zeta2};
  for (int k = 0; k < (lf ? 4 : 2) && !rc; ++k) rc = h2d(dst[k], f[k], v.es);
  if (!lf) {
    double *sd[6] = {v.ui, v.uj, v.zfi, v.zfj, v.res, v.res_beta};
    for (int k = 0; k < 6 && !rc; ++k) rc = h2d(sd[k], sup[k], v.es);
    if (!rc) PQ_HIP(hipMemset(v.rel, 0, n * v.es));
  }
  if (rc) return rc;
  rc = model_pickup(m, dt);
  if (rc) return rc;
  if (model_time_day) *model_time_day = mtd;
  if (total_time_day) *total_time_day = ttd;
  return PQ_OK;
}

}  // extern "C"
