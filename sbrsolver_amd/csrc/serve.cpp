// Serving loop behind the C-ABI (sbr_server_*, sbr_submit, sbr_poll; declared in sbr_capi.h).
//
// A server owns `slots` engines.  Each engine holds its own device copy of the scene and materials,
// its own stream and its own per-trace buffers, and is driven by one worker thread, so the queries
// of different slots overlap on the device.  Admission is first come, first served: the query at
// the head of the queue starts on the first free slot once the device memory of all slots fits the
// HBM budget with it.  A slot's memory is what its engine holds between queries (per-trace buffers
// kept at their largest size, cached launch directions), or while it runs a query the larger of
// that and the query's estimate; after a query, a slot holding more than its share of the budget
// releases its per-trace buffers.  The scene copies are outside the budget.  A query is traced exactly as sbr_trace
// traces it (one transmitter, its own receivers), so its result does not depend on its neighbours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "engine.h"
#include "sbr_capi.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

enum : int32_t { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
  int64_t id = 0;
  sbr_trace_params p{};
  std::vector<double> rx;
  int32_t nrx = 0;
  int32_t state = kQueued, status = SBR_OK, slot = -1;
  int64_t launches = 0, bounces = 0, paths = 0, est = 0;
  Clock::time_point tSubmit, tStart, tEnd;
  std::vector<double> fields;
  std::string err;
};

// Tree nodes per launch observed for a depth class (maxR, maxT, maxTotal): the admission estimate.
struct NplStat {
  double sum = 0.0;
  int64_t n = 0;
};

int64_t launches_of(const sbr_trace_params& p) {
  const int64_t nl = 10 * (int64_t(1) << (2 * p.tess)) + 2;
  const int stride = p.launchStride > 0 ? p.launchStride : 1;
  return nl > p.launchOffset ? (nl - p.launchOffset + stride - 1) / stride : 0;
}

}  // namespace

struct sbr_server {
  std::mutex mu;
  std::condition_variable cvWork;   // workers: queue, budget or stop changed
  std::condition_variable cvDone;   // pollers: a query finished
  std::vector<std::unique_ptr<sbr::Engine>> eng;
  std::vector<std::thread> workers;
  std::deque<int64_t> queue;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::map<std::tuple<int, int, int>, NplStat> npl;
  int64_t nextId = 1;
  bool stop = false;
  int device = -1;
  int32_t maxQueued = 4096;
  int64_t budget = 0, inflight = 0, peakInflight = 0;   // inflight: the sum of slotUse
  std::vector<int64_t> slotHeld, slotUse;   // per slot: bytes held between queries; bytes counted now
  int64_t submitted = 0, completed = 0, failed = 0, refused = 0, running = 0, bounces = 0;
  std::vector<double> latency;   // ms, completed queries
  bool anySubmit = false;
  Clock::time_point tFirst, tLast;
  std::string err;

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }

  // Device bytes a query needs while it runs: interaction trees, level queues, the capture's node
  // list, per-node keys and heavy list (and image sources), the launch directions, receivers and
  // their grid, line of sight, fields and the duplicate-removal table.
  int64_t estimate(const sbr_trace_params& p, int32_t nrx) const {
    const int64_t launches = launches_of(p);
    double perLaunch = std::min(p.maxR + p.maxT + 1, 8);
    auto it = npl.find(std::make_tuple(p.maxR, p.maxT, p.maxTotal));
    if (it != npl.end() && it->second.n > 0) perLaunch = 1.25 * it->second.sum / (double)it->second.n;
    perLaunch = std::max(perLaunch, 1.0);
    const double nodes = (double)launches * perLaunch + 1024.0;
    int64_t tab = 1;
    while (tab < std::max<int64_t>(1 << 16, 4 * (int64_t)nrx)) tab <<= 1;
    // reference-sphere queries also hold an image source per node (the correction's pre-test)
    const double perNode = (double)(sizeof(sbr::TNode) + 4 * sizeof(int32_t) + 3 * sizeof(uint64_t)) +
                           (p.receptionModel == 0 ? 3.0 * sizeof(double) : 0.0);
    const double bytes = nodes * perNode + 24.0 * (double)launches +
                         (double)nrx * (24.0 + 8.0 + 1.0 + 48.0) + (double)tab * 28.0;
    return (int64_t)bytes;
  }

  // Whether the head of the queue may start on idle slot `slot`: always when nothing runs (so that
  // a query that alone exceeds what the idle slots leave free still runs), otherwise when all slots'
  // memory with the head's on this slot fits the budget.
  bool head_fits(int slot) const {
    if (queue.empty()) return false;
    const Request& r = *reqs.at(queue.front());
    return running == 0 || inflight - slotUse[slot] + std::max(slotHeld[slot], r.est) <= budget;
  }

  void worker(int slot) {
    sbr::Engine& e = *eng[slot];
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cvWork.wait(lk, [&] { return stop || head_fits(slot); });
      if (stop) return;
      const int64_t id = queue.front();
      queue.pop_front();
      Request* r = reqs.at(id).get();   // not released while running
      r->state = kRunning;
      r->slot = slot;
      r->tStart = Clock::now();
      const int64_t use = std::max(slotHeld[slot], r->est);
      inflight += use - slotUse[slot];
      slotUse[slot] = use;
      peakInflight = std::max(peakInflight, inflight);
      running++;
      lk.unlock();
      cvWork.notify_all();   // another slot may admit the next query

      int rc;
      std::vector<double> f;
      std::string msg;
      int64_t launches = 0, nb = 0, np = 0, nodes = 0;
      try {
        rc = e.trace(&r->p, 1, r->nrx > 0 ? r->rx.data() : nullptr, r->nrx);
        if (rc == SBR_OK) {
          f.assign(e.field_host(), e.field_host() + e.hFieldN);
          launches = e.stats.launches;
          nb = e.stats.bounces;
          np = e.stats.winners;
          nodes = e.nodes;
        } else {
          msg = e.err;
        }
      } catch (const std::bad_alloc&) {
        rc = SBR_E_NOMEM;
        msg = "host allocation failed";
      } catch (const std::exception& x) {
        rc = SBR_E_INTERNAL;
        msg = x.what();
      }
      // only this worker touches its engine: its buffers are measured (and trimmed) unlocked
      const int64_t share = budget / (int64_t)eng.size();
      if ((int64_t)e.held_bytes() > share) e.trim();
      const int64_t held = (int64_t)e.held_bytes();

      lk.lock();
      r->tEnd = Clock::now();
      inflight += held - slotUse[slot];
      slotHeld[slot] = slotUse[slot] = held;
      running--;
      if (rc == SBR_OK) {
        r->state = kDone;
        r->fields.swap(f);
        r->launches = launches;
        r->bounces = nb;
        r->paths = np;
        completed++;
        bounces += nb;
        latency.push_back(ms_between(r->tSubmit, r->tEnd));
        tLast = r->tEnd;
        if (launches > 0) {
          NplStat& st = npl[std::make_tuple(r->p.maxR, r->p.maxT, r->p.maxTotal)];
          st.sum += (double)nodes / (double)launches;
          st.n++;
        }
      } else {
        r->state = kFailed;
        r->status = rc;
        r->err = msg;
        failed++;
      }
      cvDone.notify_all();
      cvWork.notify_all();   // budget freed
    }
  }
};

namespace {

template <typename F>
int guarded_srv(sbr_server* s, F&& f) {
  if (!s) return SBR_E_ARG;
  try {
    return f();
  } catch (const std::bad_alloc&) {
    std::lock_guard<std::mutex> g(s->mu);
    return s->fail(SBR_E_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    std::lock_guard<std::mutex> g(s->mu);
    return s->fail(SBR_E_INTERNAL, e.what());
  }
}

void fill_info(const Request& r, sbr_request_info* info) {
  if (!info) return;
  info->state = r.state;
  info->status = r.status;
  info->n_rx = r.nrx;
  info->slot = r.slot;
  info->launches = r.launches;
  info->bounces = r.bounces;
  info->paths = r.paths;
  info->est_bytes = r.est;
  const Clock::time_point now = Clock::now();
  const Clock::time_point start = r.state == kQueued ? now : r.tStart;
  const Clock::time_point end = r.state >= kDone ? r.tEnd : now;
  info->ms_queued = ms_between(r.tSubmit, start);
  info->ms_run = r.state == kQueued ? 0.0 : ms_between(r.tStart, end);
  info->ms_total = ms_between(r.tSubmit, end);
}

}  // namespace

extern "C" {

int sbr_server_create(sbr_ctx* src, const sbr_server_config* cfg, sbr_server** out) {
  if (!src || !out) return SBR_E_ARG;
  *out = nullptr;
  if (!src->eng.haveScene) return src->eng.fail(SBR_E_STATE, "server: the source context has no scene");
  if (src->eng.mats.empty()) return src->eng.fail(SBR_E_STATE, "server: the source context has no materials");
  std::unique_ptr<sbr_server> s;
  try {
    s.reset(new sbr_server());
    const int slots = cfg && cfg->slots > 0 ? std::min(cfg->slots, 16) : 3;
    s->maxQueued = cfg && cfg->max_queued > 0 ? cfg->max_queued : 4096;
    s->device = src->eng.device;
    for (int i = 0; i < slots; i++) {
      std::unique_ptr<sbr::Engine> e(new sbr::Engine());
      int rc = e->init(s->device);
      if (rc == SBR_OK) rc = e->clone_scene(src->eng);
      if (rc != SBR_OK) return src->eng.fail(rc, "server: " + e->err);
      e->tune = src->eng.tune;
      s->eng.push_back(std::move(e));
    }
    if (cfg && cfg->hbm_budget_bytes > 0) {
      s->budget = cfg->hbm_budget_bytes;
    } else if (s->device >= 0) {
      size_t freeB = 0, totalB = 0;
      if (hipSetDevice(s->device) != hipSuccess || hipMemGetInfo(&freeB, &totalB) != hipSuccess)
        return src->eng.fail(SBR_E_HIP, "server: hipMemGetInfo failed");
      s->budget = (int64_t)(freeB / 4 * 3);
    } else {
      s->budget = int64_t(1) << 34;
    }
    s->slotHeld.assign(slots, 0);
    s->slotUse.assign(slots, 0);
    for (int i = 0; i < slots; i++) s->workers.emplace_back([p = s.get(), i] { p->worker(i); });
  } catch (const std::bad_alloc&) {
    return src->eng.fail(SBR_E_NOMEM, "server: host allocation failed");
  } catch (const std::exception& e) {
    return src->eng.fail(SBR_E_INTERNAL, std::string("server: ") + e.what());
  }
  *out = s.release();
  return SBR_OK;
}

int sbr_server_destroy(sbr_server* s) {
  if (!s) return SBR_E_ARG;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    for (int64_t id : s->queue) {
      Request& r = *s->reqs.at(id);
      r.state = kFailed;
      r.status = SBR_E_STATE;
      r.err = "server stopped before the query was admitted";
      r.tEnd = Clock::now();
      s->failed++;
    }
    s->queue.clear();
  }
  s->cvWork.notify_all();
  s->cvDone.notify_all();
  for (std::thread& t : s->workers) t.join();   // running queries finish first
  if (s->device >= 0) (void)hipSetDevice(s->device);
  delete s;
  return SBR_OK;
}

// The message is copied under the lock into storage of the calling thread, valid until that
// thread's next call of sbr_server_last_error or sbr_request_error.
static thread_local std::string tlMessage;

const char* sbr_server_last_error(const sbr_server* s) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> g(const_cast<sbr_server*>(s)->mu);
  tlMessage = s->err;
  return tlMessage.c_str();
}

int sbr_submit(sbr_server* s, const sbr_trace_params* params, const double* rx, int32_t n_rx, int64_t* id) {
  return guarded_srv(s, [&] {
    std::unique_lock<std::mutex> lk(s->mu);
    if (!params || !id || n_rx < 0 || (n_rx > 0 && !rx)) return s->fail(SBR_E_ARG, "submit: bad arguments");
    if (params->tess < 0 || params->tess > sbr::kMaxTess)
      return s->fail(SBR_E_ARG, "submit: tess must be in [0, " + std::to_string(sbr::kMaxTess) + "]");
    if (s->stop) return s->fail(SBR_E_STATE, "submit: server stopping");
    if ((int32_t)s->queue.size() >= s->maxQueued) {
      s->refused++;
      return s->fail(SBR_E_LIMIT, "submit: queue full");
    }
    std::unique_ptr<Request> r(new Request());
    r->p = *params;
    r->nrx = n_rx;
    r->rx.assign(rx, rx + 3 * (size_t)n_rx);
    r->est = s->estimate(*params, n_rx);
    if (r->est > s->budget) {
      s->refused++;
      return s->fail(SBR_E_NOMEM, "submit: the query alone exceeds the HBM budget (estimate " +
                                      std::to_string(r->est) + " B, budget " + std::to_string(s->budget) + " B)");
    }
    r->id = s->nextId++;
    r->tSubmit = Clock::now();
    if (!s->anySubmit) {
      s->anySubmit = true;
      s->tFirst = r->tSubmit;
    }
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->submitted++;
    lk.unlock();
    s->cvWork.notify_all();
    return SBR_OK;
  });
}

int sbr_poll(sbr_server* s, int64_t id, int32_t wait_ms, sbr_request_info* info) {
  return guarded_srv(s, [&] {
    std::unique_lock<std::mutex> lk(s->mu);
    // Looked up again after every wake-up: another thread may release the query meanwhile.
    auto finished = [&] {
      auto it = s->reqs.find(id);
      return it == s->reqs.end() || it->second->state >= kDone;
    };
    if (!finished()) {
      if (wait_ms < 0) s->cvDone.wait(lk, finished);
      else if (wait_ms > 0) s->cvDone.wait_for(lk, std::chrono::milliseconds(wait_ms), finished);
    }
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return s->fail(SBR_E_ARG, "poll: unknown query id");
    const Request& r = *it->second;
    fill_info(r, info);
    if (r.state == kDone) return SBR_OK;
    if (r.state == kFailed) return r.status;
    return SBR_PENDING;
  });
}

const char* sbr_request_error(sbr_server* s, int64_t id) {
  if (!s) return "null server";
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(id);
  tlMessage = it == s->reqs.end() ? std::string("unknown query id") : it->second->err;
  return tlMessage.c_str();
}

int sbr_request_fields(sbr_server* s, int64_t id, double* out, int32_t n_rx) {
  return guarded_srv(s, [&] {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return s->fail(SBR_E_ARG, "fields: unknown query id");
    const Request& r = *it->second;
    if (r.state != kDone) return s->fail(SBR_E_STATE, "fields: query not finished");
    if (!out || n_rx != r.nrx) return s->fail(SBR_E_ARG, "fields: n_rx must equal the query's receiver count");
    std::copy(r.fields.begin(), r.fields.end(), out);
    return SBR_OK;
  });
}

int sbr_request_release(sbr_server* s, int64_t id) {
  return guarded_srv(s, [&] {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return s->fail(SBR_E_ARG, "release: unknown query id");
    if (it->second->state < kDone) return s->fail(SBR_E_STATE, "release: query not finished");
    s->reqs.erase(it);
    return SBR_OK;
  });
}

int sbr_server_get_stats(sbr_server* s, sbr_server_stats* out) {
  return guarded_srv(s, [&] {
    if (!out) return SBR_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    *out = sbr_server_stats{};
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->refused = s->refused;
    out->queued = (int64_t)s->queue.size();
    out->running = s->running;
    out->bounces = s->bounces;
    out->budget_bytes = s->budget;
    out->peak_inflight_bytes = s->peakInflight;
    if (!s->latency.empty()) {
      std::vector<double> l = s->latency;
      std::sort(l.begin(), l.end());
      auto rank = [&](double q) {   // nearest rank
        size_t k = (size_t)std::ceil(q * (double)l.size());
        return l[std::min(l.size(), std::max<size_t>(k, 1)) - 1];
      };
      out->p50_ms = rank(0.50);
      out->p99_ms = rank(0.99);
      out->max_ms = l.back();
      double sum = 0.0;
      for (double v : l) sum += v;
      out->mean_ms = sum / (double)l.size();
      out->window_s = std::chrono::duration<double>(s->tLast - s->tFirst).count();
      if (out->window_s > 0) {
        out->req_per_s = (double)s->completed / out->window_s;
        out->mray_bounces_per_s = (double)s->bounces / out->window_s / 1e6;
      }
    }
    return SBR_OK;
  });
}

}  // extern "C"
