// pq_serve.cpp -- request-serving engine: members resident in HBM, requests
// of two kinds (a Helmholtz / PV inversion of a field member, an advance of
// a CABARET model member by n steps) arriving at any time, continuous
// batching, admission by HBM footprint and latency percentiles.
//
// Field members of one grid size form a class served by a small pool of
// ragged inversion handles (largest layer count 8); every member brings its
// own coefficient block (its layers and stratification), so any set of
// waiting members of a class can form a batch.  A scheduler thread loops:
//   * completed work (HIP events) -> requests done, members and handles free;
//   * per class with a free handle: the waiting members' oldest inversion
//     requests (up to max_batch, oldest first) are launched as one batch;
//   * every idle model member with a waiting advance replays its captured
//     one-step graph n times on its own stream.
// Requests that arrive while others run are admitted at the next pass of the
// loop, so work of different requests overlaps on the device.  One request
// per member is in flight at a time (requests of a member run in order).
// Admission by footprint: opening a member reserves its device bytes (and,
// for the first member of a class, the class's handles); an open that would
// exceed the cap fails with PQ_ERR_NOMEM and leaves the engine unchanged.
#include "pq_common.h"
#include "pq_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

using namespace pq;

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
  return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

constexpr int kHandlesPerClass = 2;
constexpr int kClassLayers = 8;

struct Req {
  int member = -1, kind = 0, steps = 0;
  double t_sub = 0, t_start = 0, t_done = 0;
  std::atomic<int> state{0};  // 0 queued, 1 running, 2 done, 3 failed
  // why a request failed (PQ_ERR_NAN: its inversion gave a NaN psiC,
  // PQ_ERR_HIP: a launch or the device failed), or PQ_WARN_LARGE for a done
  // request whose inversion gave |psiC| > 1e100
  int code = PQ_OK;
};

struct Member {
  int kind = 0;  // 0 field, 1 model
  int g = 0, L = 0, cls = -1;
  size_t bytes = 0;
  bool open = true, busy = false;
  std::deque<long long> q;  // waiting request ids, in submission order
  // field members
  const void *zeta = nullptr;
  void *psi = nullptr;
  long ldz1 = 0, ldz2 = 0, ldp1 = 0, ldp2 = 0;
  double *coef = nullptr;
  // model members
  pq_model *model = nullptr;
  hipGraph_t graph = nullptr, graph_k = nullptr;
  hipGraphExec_t exec = nullptr, exec_k = nullptr;  // 1 step, kStepsPerGraph steps
  int step_kind = 0;  // model_step_kind at the capture
  hipEvent_t ev = nullptr;
  long long running = -1;
  int left = 0;  // steps of the running advance not yet launched
  int *st_host = nullptr;  // pinned: the model inversion's status after an advance
};

// An advance of n steps replays a graph of kStepsPerGraph steps n / k times
// and the one-step graph for the rest: one launch per step would make the
// scheduler thread's launch rate the engine's limit.  The chunks of one
// advance are launched one at a time, each after the previous completes, so
// inversions that arrive meanwhile are not queued behind the whole advance.
constexpr int kStepsPerGraph = 10;

struct Class {
  int g = 0;
  long ldz1 = 0, ldz2 = 0, ldp1 = 0, ldp2 = 0;
  pq_inv *h[kHandlesPerClass] = {};
  hipEvent_t ev[kHandlesPerClass] = {};
  // pinned: the handle's per-slot status words of the running batch, copied
  // back (and cleared on the device) behind the batch, before its event
  int *st_host[kHandlesPerClass] = {};
  std::vector<long long> running[kHandlesPerClass];
  size_t bytes = 0;
};

// Status bits of an inversion (bit 0 NaN psiC, bit 1 |psiC| > 1e100) as a
// request's code.
int status_code(int st) { return (st & 1) ? PQ_ERR_NAN : (st & 2) ? PQ_WARN_LARGE : PQ_OK; }

// Enqueues, behind the work on s, the copy of n status words to the pinned
// host array and their clearing on the device.
int status_take(int *dev, int *host, int n, hipStream_t s) {
  if (hipMemcpyAsync(host, dev, sizeof(int) * n, hipMemcpyDeviceToHost, s) != hipSuccess) return PQ_ERR_HIP;
  if (hipMemsetAsync(dev, 0, sizeof(int) * n, s) != hipSuccess) return PQ_ERR_HIP;
  return PQ_OK;
}

}  // namespace

struct pq_srv {
  int prec = PQ_PREC_FP32;
  size_t cap = 0, used = 0;
  int max_batch = 16;
  std::mutex mu;  // guards members/classes layout, submissions and the request table
  std::vector<std::unique_ptr<Member>> mem;
  std::vector<std::unique_ptr<Class>> cls;
  std::map<std::tuple<int, long, long, long, long>, int> key2cls;
  std::deque<Req> req;              // stable addresses
  std::vector<long long> incoming;  // submitted, not yet queued at their member
  std::atomic<long long> done{0}, failed{0};
  std::atomic<bool> stop{false};
  std::atomic<int> error{PQ_OK};
  std::thread th;

  void loop();
  bool pass();
};

// One scheduler pass; true if anything completed or was launched.
bool pq_srv::pass() {
  std::vector<long long> fresh;
  std::unique_lock<std::mutex> lk(mu);
  fresh.swap(incoming);
  for (long long id : fresh) mem[req[id].member]->q.push_back(id);
  bool progress = false;
  const double t = now_s();
  // code: PQ_OK or PQ_WARN_LARGE (done), else the failure
  auto finish = [&](long long id, int code) {
    Req &r = req[id];
    const bool ok = code == PQ_OK || code == PQ_WARN_LARGE;
    r.t_done = t;
    r.code = code;
    r.state.store(ok ? 2 : 3, std::memory_order_release);
    (ok ? done : failed).fetch_add(1);
  };
  // completions
  for (auto &cp : cls) {
    Class &c = *cp;
    for (int k = 0; k < kHandlesPerClass; ++k) {
      if (c.running[k].empty()) continue;
      const hipError_t q = hipEventQuery(c.ev[k]);
      if (q == hipErrorNotReady) continue;
      // slot b of the batch is the b-th request launched
      for (size_t b = 0; b < c.running[k].size(); ++b) {
        const long long id = c.running[k][b];
        finish(id, q == hipSuccess ? status_code(c.st_host[k][b]) : PQ_ERR_HIP);
        mem[req[id].member]->busy = false;
      }
      c.running[k].clear();
      progress = true;
    }
  }
  // one chunk of a model member's advance (kStepsPerGraph steps, or the
  // remaining single steps), then its completion event
  auto launch_chunk = [&](Member &m) -> int {
    hipStream_t s = model_stream(m.model);
    // a captured step goes stale when pq_model_average changed what a step does
    if (model_step_kind(m.model) != m.step_kind) return PQ_ERR_ARG;
    if (m.left >= kStepsPerGraph) {
      if (hipGraphLaunch(m.exec_k, s) != hipSuccess) return PQ_ERR_HIP;
      for (int q = 0; q < kStepsPerGraph; ++q) model_account_step(m.model);
      m.left -= kStepsPerGraph;
    } else {
      for (; m.left > 0; --m.left) {
        if (hipGraphLaunch(m.exec, s) != hipSuccess) return PQ_ERR_HIP;
        model_account_step(m.model);
      }
    }
    // after the advance's last chunk, the status of its inversions
    if (m.left == 0) {
      const int rc = status_take(model_status_dev(m.model), m.st_host, 1, s);
      if (rc) return rc;
    }
    return hipEventRecord(m.ev, s) == hipSuccess ? PQ_OK : PQ_ERR_HIP;
  };
  for (auto &mp : mem) {
    Member &m = *mp;
    if (m.kind != 1 || m.running < 0) continue;
    const hipError_t q = hipEventQuery(m.ev);
    if (q == hipErrorNotReady) continue;
    progress = true;
    if (q == hipSuccess && m.left > 0) {
      const int rc = launch_chunk(m);
      if (rc == PQ_OK) continue;
      error.store(rc);
      finish(m.running, rc);
    } else {
      finish(m.running, q == hipSuccess ? status_code(*m.st_host) : PQ_ERR_HIP);
    }
    m.running = -1;
    m.left = 0;
    m.busy = false;
  }
  // inversion batches: per class and free handle, the oldest waiting requests
  for (size_t ci = 0; ci < cls.size(); ++ci) {
    Class &c = *cls[ci];
    for (int k = 0; k < kHandlesPerClass; ++k) {
      if (!c.running[k].empty()) continue;
      std::vector<int> cand;
      for (size_t mi = 0; mi < mem.size(); ++mi) {
        Member &m = *mem[mi];
        if (m.kind == 0 && m.cls == (int)ci && !m.busy && !m.q.empty()) cand.push_back((int)mi);
      }
      if (cand.empty()) break;
      std::sort(cand.begin(), cand.end(), [&](int a, int b) {
        return req[mem[a]->q.front()].t_sub < req[mem[b]->q.front()].t_sub;
      });
      if ((int)cand.size() > max_batch) cand.resize(max_batch);
      std::vector<const void *> z;
      std::vector<void *> p;
      std::vector<const double *> cf;
      std::vector<int> lay;
      for (int mi : cand) {
        Member &m = *mem[mi];
        z.push_back(m.zeta);
        p.push_back(m.psi);
        cf.push_back(m.coef);
        lay.push_back(m.L);
      }
      hipStream_t s = (hipStream_t)pq_inv_stream(c.h[k]);
      int rc = inv_run_members(c.h[k], (int)cand.size(), z.data(), c.ldz1, c.ldz2, p.data(), c.ldp1,
                               c.ldp2, cf.data(), lay.data(), s);
      if (rc == PQ_OK) rc = status_take(inv_status_dev(c.h[k]), c.st_host[k], (int)cand.size(), s);
      if (rc == PQ_OK && hipEventRecord(c.ev[k], s) != hipSuccess) rc = PQ_ERR_HIP;
      for (int mi : cand) {
        Member &m = *mem[mi];
        const long long id = m.q.front();
        m.q.pop_front();
        if (rc != PQ_OK) {
          finish(id, rc);
          continue;
        }
        req[id].t_start = t;
        req[id].state.store(1, std::memory_order_release);
        m.busy = true;
        c.running[k].push_back(id);
      }
      if (rc != PQ_OK) error.store(rc);
      progress = true;
    }
  }
  // model advances
  for (auto &mp : mem) {
    Member &m = *mp;
    if (m.kind != 1 || m.busy || m.q.empty()) continue;
    const long long id = m.q.front();
    m.q.pop_front();
    Req &r = req[id];
    m.left = r.steps;
    const int rc = launch_chunk(m);
    if (rc != PQ_OK) {
      m.left = 0;
      error.store(rc);
      finish(id, rc);
    } else {
      r.t_start = t;
      r.state.store(1, std::memory_order_release);
      m.busy = true;
      m.running = id;
    }
    progress = true;
  }
  return progress;
}

void pq_srv::loop() {
  while (!stop.load()) {
    if (!pass()) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

namespace {

void destroy_class(Class &c) {
  for (int k = 0; k < kHandlesPerClass; ++k) {
    pq_inv_destroy(c.h[k]);
    if (c.ev[k]) (void)hipEventDestroy(c.ev[k]);
    if (c.st_host[k]) (void)hipHostFree(c.st_host[k]);
    c.h[k] = nullptr;
    c.ev[k] = nullptr;
    c.st_host[k] = nullptr;
  }
}

// Destroys a member's device objects (not its footprint accounting).
void destroy_member(Member &m) {
  if (m.coef) (void)hipFree(m.coef);
  if (m.exec) (void)hipGraphExecDestroy(m.exec);
  if (m.graph) (void)hipGraphDestroy(m.graph);
  if (m.exec_k) (void)hipGraphExecDestroy(m.exec_k);
  if (m.graph_k) (void)hipGraphDestroy(m.graph_k);
  if (m.ev) (void)hipEventDestroy(m.ev);
  if (m.st_host) (void)hipHostFree(m.st_host);
  if (m.model) pq_model_destroy(m.model);
  m.coef = nullptr;
  m.exec = nullptr;
  m.graph = nullptr;
  m.exec_k = nullptr;
  m.graph_k = nullptr;
  m.ev = nullptr;
  m.st_host = nullptr;
  m.model = nullptr;
}

int class_for(pq_srv *s, const Member &m, size_t *new_bytes) {
  auto key = std::make_tuple(m.g, m.ldz1, m.ldz2, m.ldp1, m.ldp2);
  auto it = s->key2cls.find(key);
  if (it != s->key2cls.end()) {
    *new_bytes = 0;
    return it->second;
  }
  auto c = std::make_unique<Class>();
  c->g = m.g;
  c->ldz1 = m.ldz1;
  c->ldz2 = m.ldz2;
  c->ldp1 = m.ldp1;
  c->ldp2 = m.ldp2;
  // creation stratification: barotropic-only identity of 8 layers (every
  // member brings its own coefficient block)
  std::vector<double> ss(kClassLayers, 0.0), eye(kClassLayers * kClassLayers, 0.0);
  for (int l = 0; l < kClassLayers; ++l) {
    eye[l * (kClassLayers + 1)] = 1.0;
    ss[l] = l == 0 ? 0.0 : 1.0;
  }
  std::vector<double> m2z(eye);
  for (int l = 0; l < kClassLayers; ++l) m2z[l] = 1.0;  // modes_to_z(:,1) = 1
  for (int k = 0; k < kHandlesPerClass; ++k) {
    int rc = pq_inv_create(&c->h[k], m.g, kClassLayers, ss.data(), m2z.data(), eye.data(), s->prec,
                           s->max_batch);
    if (rc == PQ_OK && hipEventCreateWithFlags(&c->ev[k], hipEventDisableTiming) != hipSuccess)
      rc = PQ_ERR_HIP;
    if (rc == PQ_OK && hipHostMalloc((void **)&c->st_host[k], sizeof(int) * s->max_batch, 0) != hipSuccess)
      rc = PQ_ERR_NOMEM;
    if (rc) {
      destroy_class(*c);
      return -rc;
    }
    c->bytes += inv_device_bytes(c->h[k]);
  }
  *new_bytes = c->bytes;
  s->cls.push_back(std::move(c));
  const int k = (int)s->cls.size() - 1;
  s->key2cls[key] = k;
  return k;
}

}  // namespace

extern "C" {

int pq_srv_create(pq_srv **out, int precision, size_t hbm_cap_bytes, int max_batch) {
  PQ_CHECK_ARG(out != nullptr, "pq_srv_create: null pointer");
  PQ_CHECK_ARG(precision == PQ_PREC_FP32 || precision == PQ_PREC_FP64 || precision == PQ_PREC_MIXED ||
                   precision == PQ_PREC_FP32_STORAGE,
               "pq_srv_create: precision %d not served (FP32, FP32 storage, mixed or FP64)", precision);
  PQ_CHECK_ARG(max_batch >= 1 && max_batch <= 64, "pq_srv_create: max_batch must be in 1..64");
  pq_srv *s = new pq_srv();
  s->prec = precision;
  s->cap = hbm_cap_bytes;
  s->max_batch = max_batch;
  s->th = std::thread([s] { s->loop(); });
  *out = s;
  return PQ_OK;
}

int pq_srv_open_field(pq_srv *s, int g, int layers, const double *ss, const double *modes_to_z,
                      const double *z_to_modes, const void *zeta, long ldz1, long ldz2, void *psi,
                      long ldp1, long ldp2, int *member) {
  PQ_CHECK_ARG(s && ss && modes_to_z && z_to_modes && zeta && psi && member,
               "pq_srv_open_field: null argument");
  PQ_CHECK_ARG(layers >= 1 && layers <= kClassLayers, "pq_srv_open_field: layers must be in 1..8");
  PQ_CHECK_ARG(ldz1 >= g && ldz2 >= g && ldp1 >= g && ldp2 >= g, "pq_srv_open_field: leading dimensions < g");
  auto m = std::make_unique<Member>();
  m->kind = 0;
  m->g = g;
  m->L = layers;
  m->zeta = zeta;
  m->psi = psi;
  m->ldz1 = ldz1;
  m->ldz2 = ldz2;
  m->ldp1 = ldp1;
  m->ldp2 = ldp2;
  std::lock_guard<std::mutex> lk(s->mu);
  size_t cbytes = 0;
  const bool had = s->key2cls.count(std::make_tuple(g, ldz1, ldz2, ldp1, ldp2)) > 0;
  const int k = class_for(s, *m, &cbytes);
  if (k < 0) return -k;
  Class &c = *s->cls[k];
  const size_t el = (s->prec == PQ_PREC_FP64 || s->prec == PQ_PREC_MIXED) ? 8 : 4;
  const size_t cb = inv_coef_size(c.h[0]) * sizeof(double);
  // the member's zeta / psi views (the caller's memory, resident) and its
  // coefficient block count against the cap
  m->bytes = cb + 2 * el * (size_t)g * g * layers;
  // a failure below leaves the engine as it was: a class created for this
  // member goes again
  auto undo = [&](int code) {
    if (m->coef) (void)hipFree(m->coef);
    m->coef = nullptr;
    if (!had) {
      destroy_class(c);
      s->cls.pop_back();
      s->key2cls.erase(std::make_tuple(g, ldz1, ldz2, ldp1, ldp2));
    }
    return code;
  };
  if (s->used + cbytes + m->bytes > s->cap) {
    pq_set_error("pq_srv_open_field: %zu bytes would exceed the HBM footprint cap (%zu of %zu in use)",
                 cbytes + m->bytes, s->used, s->cap);
    return undo(PQ_ERR_NOMEM);
  }
  if (hipMalloc(&m->coef, cb) != hipSuccess) {
    m->coef = nullptr;
    pq_set_error("pq_srv_open_field: hipMalloc of %zu bytes failed", cb);
    return undo(PQ_ERR_NOMEM);
  }
  int rc = inv_make_coefs(c.h[0], layers, ss, modes_to_z, z_to_modes, m->coef);
  if (rc) return undo(rc);
  m->cls = k;
  s->used += cbytes + m->bytes;
  s->mem.push_back(std::move(m));
  *member = (int)s->mem.size() - 1;
  return PQ_OK;
}

int pq_srv_open_model(pq_srv *s, const pq_model_params *p, const double *zeta_sdata_host, int *member) {
  PQ_CHECK_ARG(s && p && zeta_sdata_host && member, "pq_srv_open_model: null argument");
  PQ_CHECK_ARG(!p->adapt_dt, "pq_srv_open_model: models with adapt_dt are stepped alone");
  PQ_CHECK_ARG(p->scheme == PQ_SCHEME_CABARET, "pq_srv_open_model: centred-scheme models are stepped alone");
  auto m = std::make_unique<Member>();
  m->kind = 1;
  m->g = p->g;
  m->L = p->layers;
  int rc = pq_model_create(&m->model, p);
  if (rc) return rc;
  auto fail = [&](int code) {
    destroy_member(*m);
    return code;
  };
  m->bytes = model_device_bytes(m->model);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->used + m->bytes > s->cap) {
      pq_set_error("pq_srv_open_model: %zu bytes would exceed the HBM footprint cap (%zu of %zu in use)",
                   m->bytes, s->used, s->cap);
      return fail(PQ_ERR_NOMEM);
    }
    s->used += m->bytes;  // reserved now, released below on failure
  }
  auto unreserve = [&](int code) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= m->bytes;
    return fail(code);
  };
  rc = pq_model_init(m->model, zeta_sdata_host);
  if (rc) return unreserve(rc);
  // one full time step, and kStepsPerGraph of them, captured once and
  // replayed (the inversion's kernels were first launched by
  // pq_model_init, so no attribute call happens in a capture)
  hipStream_t st = model_stream(m->model);
  const ModelClock t0 = model_clock(m->model);
  m->step_kind = model_step_kind(m->model);
  for (int k : {1, kStepsPerGraph}) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return unreserve(PQ_ERR_HIP);
    // the three parts of pq_model_step, k times
    for (int q = 0; q < k && !rc; ++q) {
      rc = model_step_pre(m->model, st);
      if (!rc) rc = model_step_invert(m->model);
      if (!rc) rc = model_step_post(m->model, st, false);
    }
    hipGraph_t gr = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &gr);
    model_set_clock(m->model, t0);
    if (rc || ee != hipSuccess) {
      if (gr) (void)hipGraphDestroy(gr);
      if (!rc) pq_set_error("pq_srv_open_model: capture failed: %s", hipGetErrorString(ee));
      return unreserve(rc ? rc : PQ_ERR_HIP);
    }
    (k == 1 ? m->graph : m->graph_k) = gr;
    if (hipGraphInstantiate(k == 1 ? &m->exec : &m->exec_k, gr, nullptr, nullptr, 0) != hipSuccess)
      return unreserve(PQ_ERR_HIP);
  }
  if (hipEventCreateWithFlags(&m->ev, hipEventDisableTiming) != hipSuccess) return unreserve(PQ_ERR_HIP);
  if (hipHostMalloc((void **)&m->st_host, sizeof(int), 0) != hipSuccess) return unreserve(PQ_ERR_NOMEM);
  *m->st_host = 0;
  std::lock_guard<std::mutex> lk(s->mu);
  s->mem.push_back(std::move(m));
  *member = (int)s->mem.size() - 1;
  return PQ_OK;
}

pq_model *pq_srv_model(pq_srv *s, int member) {
  if (!s) return nullptr;
  std::lock_guard<std::mutex> lk(s->mu);
  if (member < 0 || member >= (int)s->mem.size() || s->mem[member]->kind != 1) return nullptr;
  return s->mem[member]->model;
}

// kind 0: inversion of a field member; 1: `steps` time steps of a model member.
int pq_srv_submit(pq_srv *s, int member, int kind, int steps, long long *request) {
  PQ_CHECK_ARG(s && request, "pq_srv_submit: null argument");
  const double t = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  PQ_CHECK_ARG(member >= 0 && member < (int)s->mem.size() && s->mem[member]->open,
               "pq_srv_submit: no open member %d", member);
  const Member &m = *s->mem[member];
  PQ_CHECK_ARG(kind == m.kind, "pq_srv_submit: request kind %d does not match member kind %d", kind, m.kind);
  PQ_CHECK_ARG(kind == 0 || steps >= 1, "pq_srv_submit: steps must be >= 1");
  s->req.emplace_back();
  Req &r = s->req.back();
  r.member = member;
  r.kind = kind;
  r.steps = steps;
  r.t_sub = t;
  *request = (long long)s->req.size() - 1;
  s->incoming.push_back(*request);
  return PQ_OK;
}

// 1 done, 0 pending, negative error; latency_s (may be NULL) receives
// completion - submission time of a done request.
int pq_srv_poll(pq_srv *s, long long request, double *latency_s) {
  PQ_CHECK_ARG(s != nullptr, "pq_srv_poll: null engine");
  std::lock_guard<std::mutex> lk(s->mu);
  PQ_CHECK_ARG(request >= 0 && request < (long long)s->req.size(), "pq_srv_poll: no request %lld", request);
  const Req &r = s->req[request];
  const int st = r.state.load(std::memory_order_acquire);
  if (st == 3) {
    if (r.code == PQ_ERR_NAN)
      pq_set_error("pq_srv: request %lld failed: NaN psiC encountered", request);
    else
      pq_set_error("pq_srv: request %lld failed (status %d)", request, r.code);
    return -r.code;
  }
  if (st != 2) return 0;
  if (r.code == PQ_WARN_LARGE)
    pq_set_error("pq_srv: request %lld: WARNING: Large psiC encountered; simulation may be blowing up",
                 request);
  if (latency_s) *latency_s = r.t_done - r.t_sub;
  return 1;
}

// Waits until every submitted request has completed (or timeout_s passes:
// PQ_ERR_ARG).  Returns the engine's first launch error, if any.
int pq_srv_drain(pq_srv *s, double timeout_s) {
  PQ_CHECK_ARG(s != nullptr, "pq_srv_drain: null engine");
  const double t0 = now_s();
  for (;;) {
    long long n;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      n = (long long)s->req.size();
    }
    if (s->done.load() + s->failed.load() >= n) break;
    if (now_s() - t0 > timeout_s) {
      pq_set_error("pq_srv_drain: %lld of %lld requests still pending after %.1f s",
                   n - s->done.load() - s->failed.load(), n, timeout_s);
      return PQ_ERR_ARG;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(100));
  }
  const int e = s->error.load();
  if (e) pq_set_error("pq_srv: a launch failed (status %d)", e);
  return e;
}

// Latency statistics of the completed requests of `kind` (-1: all) numbered
// `first` or later: out[0] count, [1] mean, [2] p50, [3] p95, [4] p99,
// [5] max latency (s), [6] completed requests per second from the first
// submission to the last completion, [7] requests that failed.
int pq_srv_stats(pq_srv *s, int kind, long long first, double *out8) {
  PQ_CHECK_ARG(s && out8, "pq_srv_stats: null argument");
  std::vector<double> lat;
  double t0 = INFINITY, t1 = -INFINITY;
  long long nf = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (long long id = std::max(first, 0LL); id < (long long)s->req.size(); ++id) {
      const Req &r = s->req[id];
      if (kind >= 0 && r.kind != kind) continue;
      const int st = r.state.load(std::memory_order_acquire);
      if (st == 3) ++nf;
      if (st != 2) continue;
      lat.push_back(r.t_done - r.t_sub);
      t0 = std::min(t0, r.t_sub);
      t1 = std::max(t1, r.t_done);
    }
  }
  for (int i = 0; i < 8; ++i) out8[i] = 0.0;
  out8[7] = (double)nf;
  if (lat.empty()) return PQ_OK;
  std::sort(lat.begin(), lat.end());
  auto pct = [&](double q) {
    const double x = q * (double)(lat.size() - 1);
    const size_t i = (size_t)x;
    const double f = x - (double)i;
    return i + 1 < lat.size() ? lat[i] * (1.0 - f) + lat[i + 1] * f : lat[i];
  };
  double sum = 0.0;
  for (double v : lat) sum += v;
  out8[0] = (double)lat.size();
  out8[1] = sum / (double)lat.size();
  out8[2] = pct(0.50);
  out8[3] = pct(0.95);
  out8[4] = pct(0.99);
  out8[5] = lat.back();
  out8[6] = t1 > t0 ? (double)lat.size() / (t1 - t0) : 0.0;
  return PQ_OK;
}

// Bytes reserved against the cap (members and class handles).
size_t pq_srv_used_bytes(pq_srv *s) {
  if (!s) return 0;
  std::lock_guard<std::mutex> lk(s->mu);
  return s->used;
}

// Closes an idle member (no waiting or running request) and releases its
// footprint (a class's handles stay while the engine lives).
int pq_srv_close(pq_srv *s, int member) {
  PQ_CHECK_ARG(s != nullptr, "pq_srv_close: null engine");
  std::lock_guard<std::mutex> lk(s->mu);
  PQ_CHECK_ARG(member >= 0 && member < (int)s->mem.size() && s->mem[member]->open,
               "pq_srv_close: no open member %d", member);
  Member &m = *s->mem[member];
  bool pending = m.busy || !m.q.empty();
  for (long long id : s->incoming)
    if (s->req[id].member == member) pending = true;
  PQ_CHECK_ARG(!pending, "pq_srv_close: member %d has requests in flight", member);
  destroy_member(m);
  m.open = false;
  s->used -= m.bytes;
  return PQ_OK;
}

int pq_srv_destroy(pq_srv *s) {
  if (!s) return PQ_OK;
  (void)pq_srv_drain(s, 600.0);
  s->stop.store(true);
  if (s->th.joinable()) s->th.join();
  for (auto &mp : s->mem)
    if (mp->open) destroy_member(*mp);
  for (auto &cp : s->cls) destroy_class(*cp);
  delete s;
  return PQ_OK;
}

}  // extern "C"
