// pq_ensemble.hip -- ensembles of independent members of mixed grid sizes,
// layer counts and stratifications on one GPU.
//
// Members are grouped into size classes: same grid, array leading
// dimensions and transform precision, and a layer bucket: up to 4, 5-6 or
// 7-8 layers (the column pass's geometry and registers follow a class's
// largest layer count, so few-layer members do not run the many-layer
// kernel).
// Each class owns one PV-inversion handle for its largest layer count whose
// member slots carry the members' own layer counts and stratifications
// (pq_inv_set_member_layers), so one inversion of the whole ensemble is one
// ragged batched launch sequence per class.
// Classes run on their own streams, forked from and joined back to the
// ensemble's stream, so small classes overlap on the device.  The sequence
// (and, for members that are CABARET models, the whole time step: the first
// half of every member, the batched inversions, the second half) can be
// captured once into a HIP graph and replayed, which removes the per-kernel
// launch cost of the O(100) small launches of a mixed ensemble.
#include "pq_common.h"
#include "pq_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>


using namespace pq;

struct pq_ens {
  int prec = 0;
  bool finalized = false;
  struct Member {
    int g, L;
    std::vector<double> ss, m2z, z2m;
    const void *zeta;
    void *psi;
    long ldz1, ldz2, ldp1, ldp2;
    pq_model *model;
    int cls = -1, slot = -1;
    int step_kind = 0;  // model_step_kind at the capture of the step
  };
  std::vector<Member> mem;
  struct Class {
    int g, L;
    long ldz1, ldz2, ldp1, ldp2;
    pq_inv *inv = nullptr;
    int prec = 0;
    hipStream_t stream = nullptr;
    std::vector<int> members;
    std::vector<const void *> z;
    std::vector<void *> p;
  };
  std::vector<Class> cls;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> ev;  // fork/join events: 1 + one per class and model
  hipGraph_t graph_inv = nullptr, graph_step = nullptr;
  hipGraphExec_t exec_inv = nullptr, exec_step = nullptr;
  long long steps = 0;
};

static int ens_check(const pq_ens *e, const char *what) {
  if (!e) {
    pq_set_error("%s: null ensemble", what);
    return PQ_ERR_ARG;
  }
  if (!e->finalized) {
    pq_set_error("%s: call pq_ens_finalize first", what);
    return PQ_ERR_ARG;
  }
  return PQ_OK;
}

// Inversion of every class, forked from s and joined back to it.
static int ens_invert_on(pq_ens *e, hipStream_t s) {
  PQ_HIP(hipEventRecord(e->ev[0], s));
  for (size_t k = 0; k < e->cls.size(); ++k) {
    auto &c = e->cls[k];
    PQ_HIP(hipStreamWaitEvent(c.stream, e->ev[0], 0));
    int rc = pq_inv_run_dev(c.inv, (int)c.members.size(), c.z.data(), c.ldz1, c.ldz2, c.p.data(), c.ldp1,
                            c.ldp2, c.stream);
    if (rc) return rc;
    PQ_HIP(hipEventRecord(e->ev[1 + k], c.stream));
  }
  for (size_t k = 0; k < e->cls.size(); ++k) PQ_HIP(hipStreamWaitEvent(s, e->ev[1 + k], 0));
  return PQ_OK;
}

// One time step of every model member: first halves, inversions, second halves.
static int ens_step_on(pq_ens *e, hipStream_t s) {
  const size_t nc = e->cls.size();
  auto fork_models = [&](bool pre) -> int {
    PQ_HIP(hipEventRecord(e->ev[0], s));
    size_t k = 0;
    for (auto &m : e->mem) {
      if (!m.model) continue;
      hipStream_t ms = model_stream(m.model);
      PQ_HIP(hipStreamWaitEvent(ms, e->ev[0], 0));
      int rc = pre ? model_step_pre(m.model, ms) : model_step_post(m.model, ms, false);
      if (rc) return rc;
      PQ_HIP(hipEventRecord(e->ev[1 + nc + k], ms));
      ++k;
    }
    for (size_t q = 0; q < k; ++q) PQ_HIP(hipStreamWaitEvent(s, e->ev[1 + nc + q], 0));
    return PQ_OK;
  };
  int rc = fork_models(true);
  if (rc) return rc;
  rc = ens_invert_on(e, s);
  if (rc) return rc;
  return fork_models(false);
}

extern "C" {

int pq_ens_create(pq_ens **out, int precision) {
  PQ_CHECK_ARG(out != nullptr, "pq_ens_create: null pointer");
  PQ_CHECK_ARG(precision >= 0 && precision <= PQ_PREC_FP32_MFMA, "pq_ens_create: bad precision %d",
               precision);
  pq_ens *e = new pq_ens();
  e->prec = precision;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    pq_set_error("pq_ens_create: cannot create a stream");
    return PQ_ERR_HIP;
  }
  *out = e;
  return PQ_OK;
}

int pq_ens_add(pq_ens *e, int g, int layers, const double *ss, const double *modes_to_z,
               const double *z_to_modes, const void *zeta, long ldz1, long ldz2, void *psi, long ldp1,
               long ldp2, int *index) {
  PQ_CHECK_ARG(e && !e->finalized, "pq_ens_add: null or finalized ensemble");
  PQ_CHECK_ARG(ss && modes_to_z && z_to_modes && zeta && psi, "pq_ens_add: null argument");
  PQ_CHECK_ARG(layers >= 1 && layers <= 8, "pq_ens_add: layers must be in 1..8");
  PQ_CHECK_ARG(ldz1 >= g && ldz2 >= g && ldp1 >= g && ldp2 >= g, "pq_ens_add: leading dimensions < g");
  pq_ens::Member m;
  m.g = g;
  m.L = layers;
  m.ss.assign(ss, ss + layers);
  m.m2z.assign(modes_to_z, modes_to_z + layers * layers);
  m.z2m.assign(z_to_modes, z_to_modes + layers * layers);
  m.zeta = zeta;
  m.psi = psi;
  m.ldz1 = ldz1;
  m.ldz2 = ldz2;
  m.ldp1 = ldp1;
  m.ldp2 = ldp2;
  m.model = nullptr;
  e->mem.push_back(m);
  if (index) *index = (int)e->mem.size() - 1;
  return PQ_OK;
}

int pq_ens_add_model(pq_ens *e, pq_model *model, const double *ss, const double *modes_to_z,
                     const double *z_to_modes, int *index) {
  PQ_CHECK_ARG(e && !e->finalized && model, "pq_ens_add_model: null argument or finalized ensemble");
  PQ_CHECK_ARG(!model_adaptive(model), "pq_ens_add_model: models with adapt_dt are stepped alone");
  PQ_CHECK_ARG(model_scheme(model) == PQ_SCHEME_CABARET,
               "pq_ens_add_model: centred-scheme models are stepped alone");
  int g, L, es;
  long ld;
  void *z, *p;
  model_views(model, &g, &L, &z, &p, &ld, &es);
  const int ens_es = (e->prec == PQ_PREC_FP64 || e->prec == PQ_PREC_MIXED) ? 8 : 4;
  PQ_CHECK_ARG(es == ens_es,
               "pq_ens_add_model: the model stores %s fields and the ensemble's precision takes %s "
               "(FP64: PQ_PREC_FP64 or PQ_PREC_MIXED; FP32: the FP32-storage precisions)",
               es == 8 ? "FP64" : "FP32", ens_es == 8 ? "FP64" : "FP32");
  int rc = pq_ens_add(e, g, L, ss, modes_to_z, z_to_modes, z, ld, ld, p, ld, ld, index);
  if (rc) return rc;
  e->mem.back().model = model;
  return PQ_OK;
}

int pq_ens_finalize(pq_ens *e) {
  PQ_CHECK_ARG(e && !e->finalized, "pq_ens_finalize: null or finalized ensemble");
  PQ_CHECK_ARG(!e->mem.empty(), "pq_ens_finalize: no members");
  std::map<std::tuple<int, long, long, long, long, int, int>, int> key2cls;
  for (size_t i = 0; i < e->mem.size(); ++i) {
    auto &m = e->mem[i];
    // FP32_AUTO resolves per member (pq_prec_fp32_auto)
    const int prec = e->prec == PQ_PREC_FP32_AUTO
                         ? pq_prec_fp32_auto(m.g, m.L, m.m2z.data(), m.z2m.data())
                         : e->prec;
    auto key = std::make_tuple(m.g, m.ldz1, m.ldz2, m.ldp1, m.ldp2, prec, m.L > 4 ? (m.L > 6 ? 2 : 1) : 0);
    auto it = key2cls.find(key);
    int k;
    if (it == key2cls.end()) {
      k = (int)e->cls.size();
      key2cls[key] = k;
      pq_ens::Class c;
      c.g = m.g;
      c.L = 0;
      c.ldz1 = m.ldz1;
      c.ldz2 = m.ldz2;
      c.ldp1 = m.ldp1;
      c.ldp2 = m.ldp2;
      c.prec = prec;
      e->cls.push_back(c);
    } else {
      k = it->second;
    }
    auto &c = e->cls[k];
    m.cls = k;
    m.slot = (int)c.members.size();
    c.L = std::max(c.L, m.L);
    c.members.push_back((int)i);
    c.z.push_back(m.zeta);
    c.p.push_back(m.psi);
  }
  for (auto &c : e->cls) {
    // created with a member of the largest layer count; every slot then
    // gets its member's own layers and stratification
    int top = c.members[0];
    for (int q : c.members)
      if (e->mem[q].L == c.L) {
        top = q;
        break;
      }
    const auto &m0 = e->mem[top];
    int rc = pq_inv_create(&c.inv, c.g, c.L, m0.ss.data(), m0.m2z.data(), m0.z2m.data(), c.prec,
                           (int)c.members.size());
    if (rc) return rc;
    for (size_t q = 0; q < c.members.size(); ++q) {
      const auto &m = e->mem[c.members[q]];
      rc = pq_inv_set_member_layers(c.inv, (int)q, m.L, m.ss.data(), m.m2z.data(), m.z2m.data());
      if (rc) return rc;
    }
    c.stream = (hipStream_t)pq_inv_stream(c.inv);
  }
  e->ev.resize(1 + e->cls.size() + e->mem.size());
  for (auto &ev : e->ev) PQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  e->finalized = true;
  return PQ_OK;
}

int pq_ens_class_count(pq_ens *e) { return e && e->finalized ? (int)e->cls.size() : 0; }

pq_inv *pq_ens_class_handle(pq_ens *e, int k) {
  if (!e || !e->finalized || k < 0 || k >= (int)e->cls.size()) return nullptr;
  return e->cls[k].inv;
}

int pq_ens_class_precision(pq_ens *e, int k) {
  if (!e || !e->finalized || k < 0 || k >= (int)e->cls.size()) return -1;
  return e->cls[k].prec;
}

int pq_ens_member_class(pq_ens *e, int member, int *cls, int *slot) {
  int rc = ens_check(e, "pq_ens_member_class");
  if (rc) return rc;
  PQ_CHECK_ARG(member >= 0 && member < (int)e->mem.size() && cls && slot, "pq_ens_member_class: bad argument");
  *cls = e->mem[member].cls;
  *slot = e->mem[member].slot;
  return PQ_OK;
}

void *pq_ens_stream(pq_ens *e) { return e ? (void *)e->stream : nullptr; }

int pq_ens_invert(pq_ens *e) {
  int rc = ens_check(e, "pq_ens_invert");
  if (rc) return rc;
  if (e->exec_inv) {
    PQ_HIP(hipGraphLaunch(e->exec_inv, e->stream));
    return PQ_OK;
  }
  return ens_invert_on(e, e->stream);
}

int pq_ens_step(pq_ens *e, int nsteps) {
  int rc = ens_check(e, "pq_ens_step");
  if (rc) return rc;
  PQ_CHECK_ARG(nsteps >= 0, "pq_ens_step: nsteps < 0");
  for (int it = 0; it < nsteps; ++it) {
    if (e->exec_step) {
      for (auto &m : e->mem)
        PQ_CHECK_ARG(!m.model || model_step_kind(m.model) == m.step_kind,
                     "pq_ens_step: a model's time-mean setting changed since the capture; capture again");
      PQ_HIP(hipGraphLaunch(e->exec_step, e->stream));
      for (auto &m : e->mem)
        if (m.model) model_account_step(m.model);
    } else {
      rc = ens_step_on(e, e->stream);
      if (rc) return rc;
    }
    ++e->steps;
  }
  return PQ_OK;
}

// what: 0 the ensemble inversion, 1 one time step of the model members.
// Tile shapes and member stratifications are fixed from here on (a later
// pq_inv_set_tiles / pq_inv_set_member needs a new capture).
int pq_ens_capture(pq_ens *e, int what) {
  int rc = ens_check(e, "pq_ens_capture");
  if (rc) return rc;
  PQ_CHECK_ARG(what == 0 || what == 1, "pq_ens_capture: what must be 0 or 1");
  // a captured step holds the time-mean accumulation or not for good
  if (what == 1)
    for (auto &m : e->mem) {
      PQ_CHECK_ARG(!m.model || model_step_kind(m.model) >= 0,
                   "pq_ens_capture: a model's time mean starts within the coming steps (spin-up not "
                   "over); step it eagerly until then");
      if (m.model) m.step_kind = model_step_kind(m.model);
    }
  // one eager inversion first (it changes only psi, which every step
  // recomputes): kernels get their LDS limits outside the capture
  rc = ens_invert_on(e, e->stream);
  if (rc) return rc;
  PQ_HIP(hipStreamSynchronize(e->stream));
  hipGraph_t &graph = what == 0 ? e->graph_inv : e->graph_step;
  hipGraphExec_t &exec = what == 0 ? e->exec_inv : e->exec_step;
  if (exec) {
    (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  if (graph) {
    (void)hipGraphDestroy(graph);
    graph = nullptr;
  }
  // host-side model clocks advance as the graph replays, not at capture
  std::vector<ModelClock> times;
  for (auto &m : e->mem)
    if (m.model) times.push_back(model_clock(m.model));
  PQ_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeGlobal));
  rc = what == 0 ? ens_invert_on(e, e->stream) : ens_step_on(e, e->stream);
  {
    size_t k = 0;
    for (auto &m : e->mem)
      if (m.model) model_set_clock(m.model, times[k++]);
  }
  hipGraph_t g = nullptr;
  hipError_t ee = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ee != hipSuccess) {
    pq_set_error("pq_ens_capture: capture failed: %s", hipGetErrorString(ee));
    return PQ_ERR_HIP;
  }
  graph = g;
  PQ_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  return PQ_OK;
}

int pq_ens_synchronize(pq_ens *e) {
  PQ_CHECK_ARG(e != nullptr, "pq_ens_synchronize: null ensemble");
  PQ_HIP(hipStreamSynchronize(e->stream));
  return PQ_OK;
}

// Worst status (PQ_ERR_NAN before PQ_WARN_LARGE) over the classes' last runs.
int pq_ens_status(pq_ens *e) {
  int rc = ens_check(e, "pq_ens_status");
  if (rc) return rc;
  PQ_HIP(hipStreamSynchronize(e->stream));
  int worst = PQ_OK;
  for (auto &c : e->cls) {
    int r = pq_inv_status(c.inv, nullptr);
    if (r == PQ_ERR_NAN || (r != PQ_OK && worst == PQ_OK)) worst = r;
  }
  return worst;
}

int pq_ens_destroy(pq_ens *e) {
  if (!e) return PQ_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->exec_inv) (void)hipGraphExecDestroy(e->exec_inv);
  if (e->exec_step) (void)hipGraphExecDestroy(e->exec_step);
  if (e->graph_inv) (void)hipGraphDestroy(e->graph_inv);
  if (e->graph_step) (void)hipGraphDestroy(e->graph_step);
  for (auto &ev : e->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &c : e->cls) pq_inv_destroy(c.inv);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PQ_OK;
}

}  // extern "C"
