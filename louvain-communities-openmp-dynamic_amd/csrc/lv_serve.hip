// Serving loop of the MI355X Louvain library: the lv_server_* functions of
// include/louvain_hip.h.  It is written on the library's own C ABI (device
// graphs, states and batches, lv_dgraph_update, lv_run): resident graphs and
// states, one worker thread and lv_context per slot, requests of one graph
// in submission order, requests of different graphs overlapping on the
// device, admission by device memory.  The per-request step is the
// reference driver's batch loop body (main.cxx:101-137): update the graph,
// run Dynamic Frontier from the carried state (main.cxx:127-135).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "louvain_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes a request holds while it runs, per raw batch entry: its tidy
// lists (u, v, w and the stored deletion weight, 16 B) and what the tidy's
// scratch may grow by if no earlier request on the slot was as large (the
// uploaded lists, sort keys and indices, flags, selection: 41 B).
constexpr uint64_t kRequestBytesPerEntry = 16 + 41;

struct Graph {
  lv_dgraph* g = nullptr;
  lv_dstate* s = nullptr;
  uint64_t span = 0;
  uint64_t bytes = 0;  // graph + state, as of the last request on it
  bool busy = false;
  lv_tune_choice tuning{};  // its DF schedule (candidates 0: the defaults)
};

struct Request {
  uint64_t ticket = 0;
  uint32_t graph = 0;
  int warm_slot = -1;  // a warm-up run of this slot on the graph (no batch, not reported)
  std::vector<uint32_t> du, dv, iu, iv;
  std::vector<float> dw, iw;
  double submitted_ms = 0;
  uint64_t bytes = 0;
};

struct Finished {
  lv_completion c{};
  std::vector<uint32_t> memb;
  std::vector<double> vtot, ctot;
};

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// A raw probe batch for graph g: b undirected updates (80% insertions of
// random id pairs, 20% deletions of random existing edges), both directions
// of each, as the requests bring them.
void probe_batch(const lv_graph* g, double fraction, uint64_t seed, std::vector<uint32_t>& du,
                 std::vector<uint32_t>& dv, std::vector<float>& dw, std::vector<uint32_t>& iu,
                 std::vector<uint32_t>& iv, std::vector<float>& iw) {
  const uint64_t S = g->span, E = S ? g->offsets[S] : 0;
  if (!S || !E) return;
  const uint64_t b = std::max<uint64_t>(1, (uint64_t)std::llround(fraction * (double)(E / 2)));
  const uint64_t ni = (uint64_t)std::llround(0.8 * (double)b), nd = b - ni;
  uint64_t x = seed * 0x2545F4914F6CDD1Dull + 1;
  for (uint64_t k = 0, tries = 0; k < nd && tries < 16 * nd; ++tries) {
    const uint64_t i = splitmix64(x) % E;
    const uint64_t u = (uint64_t)(std::upper_bound(g->offsets, g->offsets + S + 1, i) - g->offsets) - 1;
    if (u >= S || (g->degrees && i - g->offsets[u] >= g->degrees[u]) || (g->exists && !g->exists[u])) continue;
    const uint32_t v = g->keys[i];
    const float w = g->weights ? g->weights[i] : (float)g->weights_f64[i];
    du.push_back((uint32_t)u);
    dv.push_back(v);
    dw.push_back(w);
    du.push_back(v);
    dv.push_back((uint32_t)u);
    dw.push_back(w);
    ++k;
  }
  for (uint64_t k = 0, tries = 0; k < ni && tries < 16 * ni; ++tries) {
    const uint32_t u = (uint32_t)(splitmix64(x) % S), v = (uint32_t)(splitmix64(x) % S);
    if (g->exists && (!g->exists[u] || !g->exists[v])) continue;
    iu.push_back(u);
    iv.push_back(v);
    iw.push_back(1.0f);
    iu.push_back(v);
    iv.push_back(u);
    iw.push_back(1.0f);
    ++k;
  }
}

lv_batch as_batch(const std::vector<uint32_t>& u, const std::vector<uint32_t>& v, const std::vector<float>& w) {
  lv_batch b;
  b.u = u.data();
  b.v = v.data();
  b.w = w.data();
  b.size = u.size();
  return b;
}

}  // namespace

struct lv_server {
  lv_server_options o{};
  lv_context* ctrl = nullptr;         // the calling threads' context (uploads, Static, downloads)
  std::vector<lv_context*> slot;      // one per worker
  std::vector<uint64_t> slot_ws;      // each slot's workspace bytes after its last request
  uint64_t ctrl_ws = 0;               // ctrl's workspace (Static runs of added graphs)
  std::vector<std::thread> workers;
  std::deque<Graph> graphs;           // (a deque: workers keep references while graphs are added)
  std::deque<Request*> pending;       // submission order
  std::set<uint64_t> outstanding;     // submitted, not finished
  std::map<uint64_t, Finished> finished;
  uint64_t next_ticket = 1;
  uint64_t inflight = 0, peak = 0;
  uint64_t submitted = 0, completed = 0, failed = 0, running = 0;
  int warm_left = 0;
  bool stopping = false;
  std::mutex mu;
  std::mutex ctrl_mu;  // serialises the callers' use of ctrl
  std::condition_variable cv_work, cv_done;
  std::string err;
  Clock::time_point t0 = Clock::now();

  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
  uint64_t resident() const {
    uint64_t b = 0;
    for (const Graph& g : graphs) b += g.bytes;
    return b;
  }
  uint64_t workspace() const {
    uint64_t b = ctrl_ws;
    for (uint64_t w : slot_ws) b += w;
    return b;
  }
  // The next request worker k may start (under mu): its own warm-up first,
  // else the oldest request whose graph is idle and has no older request
  // waiting, if it fits the memory budget (or nothing runs).  A request that
  // does not fit stops the scan, so later ones do not overtake it.
  std::deque<Request*>::iterator pick(int k) {
    std::vector<char> seen(graphs.size(), 0);
    for (auto it = pending.begin(); it != pending.end(); ++it) {
      Request* r = *it;
      if (r->warm_slot >= 0) {
        if (r->warm_slot == k) return it;
        continue;
      }
      if (seen[r->graph] || graphs[r->graph].busy) {
        seen[r->graph] = 1;
        continue;
      }
      seen[r->graph] = 1;
      const uint64_t use = resident() + workspace() + inflight;
      if (running && o.memory_budget && use + r->bytes > o.memory_budget) return pending.end();
      return it;
    }
    return pending.end();
  }
  void work(int k);
};

namespace {

// One request on slot k's context (no lock held; g was looked up under it,
// since adding a graph may move the deque's index while this runs).
void serve_one(lv_server* s, int k, Request* r, Graph& g, Finished& f) {
  lv_context* c = s->slot[k];
  f.c.ticket = r->ticket;
  f.c.graph_id = r->graph;
  f.c.submitted_ms = r->submitted_ms;
  f.c.started_ms = s->now_ms();
  f.c.modularity = std::numeric_limits<double>::quiet_NaN();
  int rc = LV_OK;
  lv_result res;
  std::memset(&res, 0, sizeof res);
  if (r->warm_slot >= 0) {
    // the graph's DF schedule on this slot, then its workspace sized for
    // the graph: a run with no batch on a copy of the state (the graph's
    // state keeps its own run count)
    lv_dstate* tmp = nullptr;
    lv_dbatch* none = nullptr;
    if (g.tuning.candidates) rc = lv_context_add_tuning(c, LV_DYNAMIC_FRONTIER, g.g, &g.tuning);
    if (rc == LV_OK) rc = lv_dstate_create(c, g.span, &tmp);
    if (rc == LV_OK) rc = lv_dstate_copy(c, tmp, g.s);
    if (rc == LV_OK) rc = lv_dbatch_create(c, nullptr, nullptr, &none);
    if (rc == LV_OK) rc = lv_run(c, LV_DYNAMIC_FRONTIER, g.g, none, tmp, &s->o.options, &res);
    lv_dbatch_destroy(none);
    lv_dstate_destroy(tmp);
  } else {
    const lv_batch del = as_batch(r->du, r->dv, r->dw), ins = as_batch(r->iu, r->iv, r->iw);
    lv_dbatch* tb = nullptr;
    rc = lv_dgraph_update(c, g.g, &del, &ins, &tb);
    if (rc == LV_OK) rc = lv_dbatch_sizes(tb, &f.c.deletions, &f.c.insertions);
    if (rc == LV_OK) rc = lv_run(c, LV_DYNAMIC_FRONTIER, g.g, tb, g.s, &s->o.options, &res);
    lv_dbatch_destroy(tb);
    if (rc == LV_OK && s->o.modularity) rc = lv_dmodularity(c, g.g, g.s, s->o.options.resolution, &f.c.modularity);
    if (rc == LV_OK && s->o.keep_states) {
      f.memb.resize(g.span);
      f.vtot.resize(g.span);
      f.ctot.resize(g.span);
      rc = lv_dstate_download(c, g.s, f.memb.data(), f.vtot.data(), f.ctot.data());
    }
  }
  f.c.status = rc;
  f.c.result = res;
  f.c.done_ms = s->now_ms();
  if (rc != LV_OK) {
    std::lock_guard<std::mutex> lock(s->mu);
    s->err = std::string("request ") + std::to_string(r->ticket) + ": " + lv_last_error(c);
  }
}

// The DF schedule of a graph being added (ctrl held): a copy of the graph
// and its state, two runs with no batch (the second builds the hub tables,
// as a graph's second request does), the probe batch applied by
// lv_dgraph_update, then lv_autotune on the copy.
int tune_graph(lv_server* s, const lv_graph* host, Graph& G, uint64_t seed) {
  lv_context* c = s->ctrl;
  lv_dgraph* gc = nullptr;
  lv_dstate* sc = nullptr;
  lv_dbatch *none = nullptr, *tb = nullptr;
  std::vector<uint32_t> du, dv, iu, iv;
  std::vector<float> dw, iw;
  probe_batch(host, s->o.probe_fraction, seed, du, dv, dw, iu, iv, iw);
  lv_result res;
  std::memset(&res, 0, sizeof res);
  int rc = lv_dgraph_create(c, host, &gc);
  if (rc == LV_OK) rc = lv_dstate_create(c, G.span, &sc);
  if (rc == LV_OK) rc = lv_dstate_copy(c, sc, G.s);
  if (rc == LV_OK) rc = lv_dbatch_create(c, nullptr, nullptr, &none);
  for (int k = 0; k < 2 && rc == LV_OK; ++k) rc = lv_run(c, LV_DYNAMIC_FRONTIER, gc, none, sc, &s->o.options, &res);
  const lv_batch del = as_batch(du, dv, dw), ins = as_batch(iu, iv, iw);
  if (rc == LV_OK) rc = lv_dgraph_update(c, gc, &del, &ins, &tb);
  lv_tune_choice k{};
  if (rc == LV_OK && (!du.empty() || !iu.empty()))
    rc = lv_autotune(c, LV_DYNAMIC_FRONTIER, gc, tb, sc, &s->o.options, 3, &k);
  if (rc == LV_OK) G.tuning = k;
  lv_dbatch_destroy(tb);
  lv_dbatch_destroy(none);
  lv_dstate_destroy(sc);
  lv_dgraph_destroy(gc);
  return rc;
}

}  // namespace

void lv_server::work(int k) {
  std::unique_lock<std::mutex> lock(mu);
  for (;;) {
    std::deque<Request*>::iterator it;
    cv_work.wait(lock, [&] { return (stopping && pending.empty()) || (it = pick(k)) != pending.end(); });
    if (stopping && pending.empty()) return;
    Request* r = *it;
    pending.erase(it);
    const bool warm = r->warm_slot >= 0;
    Graph& g = graphs[r->graph];
    if (!warm) {
      g.busy = true;
      inflight += r->bytes;
      ++running;
      peak = std::max(peak, resident() + workspace() + inflight);
    }
    lock.unlock();
    Finished f;
    serve_one(this, k, r, g, f);
    const uint64_t ws = lv_context_workspace_bytes(slot[k]);
    const uint64_t gb = lv_dgraph_device_bytes(g.g) + lv_dstate_device_bytes(g.s);
    lock.lock();
    slot_ws[k] = ws;
    if (warm) {
      --warm_left;
    } else {
      g.busy = false;
      g.bytes = gb;
      inflight -= r->bytes;
      --running;
      ++completed;
      if (f.c.status != LV_OK) ++failed;
      outstanding.erase(r->ticket);
      finished[r->ticket] = std::move(f);
    }
    delete r;
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" {

void lv_server_options_default(lv_server_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->device = 0;
  o->slots = 4;
  o->memory_budget = 0;
  lv_options_default(&o->options);
  o->path = 2;
  o->modularity = 0;
  o->keep_states = 0;
  o->autotune = 1;
  o->probe_fraction = 1e-4;
}

int lv_server_create(const lv_server_options* o, lv_server** out) {
  if (!o || !out || o->slots < 1 || o->slots > 64 || o->path < 0 || o->path > 2 || o->autotune < 0 ||
      o->autotune > 1 || !(o->probe_fraction > 0) || o->probe_fraction > 1)
    return LV_ERR_INVALID;
  *out = nullptr;
  auto* s = new lv_server();
  s->o = *o;
  int rc = lv_context_create(o->device, &s->ctrl);
  if (rc == LV_OK) rc = lv_context_set_path(s->ctrl, o->path);
  for (int k = 0; rc == LV_OK && k < o->slots; ++k) {
    lv_context* c = nullptr;
    rc = lv_context_create(o->device, &c);
    if (rc == LV_OK) rc = lv_context_set_path(c, o->path);
    if (c) s->slot.push_back(c);
  }
  if (rc != LV_OK) {
    for (lv_context* c : s->slot) lv_context_destroy(c);
    lv_context_destroy(s->ctrl);
    delete s;
    return rc;
  }
  s->slot_ws.assign(o->slots, 0);
  for (int k = 0; k < o->slots; ++k) s->workers.emplace_back([s, k] { s->work(k); });
  *out = s;
  return LV_OK;
}

void lv_server_destroy(lv_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (std::thread& t : s->workers) t.join();
  for (Graph& g : s->graphs) {
    lv_dgraph_destroy(g.g);
    lv_dstate_destroy(g.s);
  }
  for (lv_context* c : s->slot) lv_context_destroy(c);
  lv_context_destroy(s->ctrl);
  delete s;
}

const char* lv_server_last_error(const lv_server* s) { return s ? s->err.c_str() : "no server"; }

double lv_server_now_ms(const lv_server* s) { return s ? s->now_ms() : 0.0; }

int lv_server_add_graph(lv_server* s, const lv_graph* g, const uint32_t* q, const double* qvtot, const double* qctot,
                        uint32_t* graph_id) {
  if (!s || !g || !graph_id || (q && (!qvtot || !qctot))) return LV_ERR_INVALID;
  std::lock_guard<std::mutex> ctrl_lock(s->ctrl_mu);
  Graph G;
  int rc = lv_dgraph_create(s->ctrl, g, &G.g);
  if (rc == LV_OK) {
    G.span = lv_dgraph_span(G.g);
    rc = lv_dstate_create(s->ctrl, G.span, &G.s);
  }
  if (rc == LV_OK) {
    if (q) {
      rc = lv_dstate_upload(s->ctrl, G.s, q, qvtot, qctot);
    } else {
      lv_result res;
      std::memset(&res, 0, sizeof res);
      rc = lv_run(s->ctrl, LV_STATIC, G.g, nullptr, G.s, &s->o.options, &res);
    }
  }
  if (rc == LV_OK && s->o.autotune) rc = tune_graph(s, g, G, (uint64_t)s->graphs.size() + 1);
  if (rc != LV_OK) {
    std::lock_guard<std::mutex> lock(s->mu);
    s->err = lv_last_error(s->ctrl);
    lv_dgraph_destroy(G.g);
    lv_dstate_destroy(G.s);
    return rc;
  }
  G.bytes = lv_dgraph_device_bytes(G.g) + lv_dstate_device_bytes(G.s);
  std::unique_lock<std::mutex> lock(s->mu);
  s->ctrl_ws = lv_context_workspace_bytes(s->ctrl);
  *graph_id = (uint32_t)s->graphs.size();
  s->graphs.push_back(G);
  // a warm-up run on every slot, ahead of any request queued after this
  for (int k = 0; k < (int)s->slot.size(); ++k) {
    auto* r = new Request();
    r->graph = *graph_id;
    r->warm_slot = k;
    s->pending.push_front(r);
    ++s->warm_left;
  }
  s->cv_work.notify_all();
  s->cv_done.wait(lock, [&] { return s->warm_left == 0; });
  return LV_OK;
}

int lv_server_submit(lv_server* s, uint32_t graph_id, const lv_batch* del, const lv_batch* ins, uint64_t* ticket) {
  if (!s || !ticket) return LV_ERR_INVALID;
  const uint64_t nd = del ? del->size : 0, ni = ins ? ins->size : 0;
  if ((nd && (!del->u || !del->v || !del->w)) || (ni && (!ins->u || !ins->v || !ins->w))) return LV_ERR_INVALID;
  auto* r = new Request();
  r->du.assign(del ? del->u : nullptr, del ? del->u + nd : nullptr);
  r->dv.assign(del ? del->v : nullptr, del ? del->v + nd : nullptr);
  r->dw.assign(del ? del->w : nullptr, del ? del->w + nd : nullptr);
  r->iu.assign(ins ? ins->u : nullptr, ins ? ins->u + ni : nullptr);
  r->iv.assign(ins ? ins->v : nullptr, ins ? ins->v + ni : nullptr);
  r->iw.assign(ins ? ins->w : nullptr, ins ? ins->w + ni : nullptr);
  r->bytes = (nd + ni) * kRequestBytesPerEntry;
  std::lock_guard<std::mutex> lock(s->mu);
  if (graph_id >= s->graphs.size()) {
    delete r;
    return LV_ERR_INVALID;
  }
  r->graph = graph_id;
  r->ticket = s->next_ticket++;
  r->submitted_ms = s->now_ms();
  *ticket = r->ticket;
  s->pending.push_back(r);
  s->outstanding.insert(r->ticket);
  ++s->submitted;
  s->cv_work.notify_all();
  return LV_OK;
}

int lv_server_poll(lv_server* s, uint64_t ticket, int timeout_ms, lv_completion* out) {
  if (!s || !out) return LV_ERR_INVALID;
  std::unique_lock<std::mutex> lock(s->mu);
  if (ticket == 0 || ticket >= s->next_ticket) return LV_ERR_INVALID;
  auto ready = [&] { return s->finished.count(ticket) != 0; };
  if (!ready() && !s->outstanding.count(ticket)) return LV_ERR_INVALID;  // reported already
  if (!ready()) {
    if (timeout_ms == 0) return LV_ERR_PENDING;
    if (timeout_ms < 0) s->cv_done.wait(lock, ready);
    else if (!s->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready)) return LV_ERR_PENDING;
  }
  Finished& f = s->finished[ticket];
  uint32_t* m = out->result.membership;
  double* vt = out->result.vertex_weight;
  double* ct = out->result.community_weight;
  *out = f.c;
  out->result.membership = m;
  out->result.vertex_weight = vt;
  out->result.community_weight = ct;
  if (!f.memb.empty()) {
    if (m) std::memcpy(m, f.memb.data(), f.memb.size() * sizeof(uint32_t));
    if (vt) std::memcpy(vt, f.vtot.data(), f.vtot.size() * sizeof(double));
    if (ct) std::memcpy(ct, f.ctot.data(), f.ctot.size() * sizeof(double));
  }
  s->finished.erase(ticket);
  return LV_OK;
}

int lv_server_drain(lv_server* s) {
  if (!s) return LV_ERR_INVALID;
  std::unique_lock<std::mutex> lock(s->mu);
  s->cv_done.wait(lock, [&] { return s->completed == s->submitted; });
  return LV_OK;
}

int lv_server_graph_state(lv_server* s, uint32_t graph_id, uint32_t* q, double* qvtot, double* qctot) {
  if (!s) return LV_ERR_INVALID;
  std::lock_guard<std::mutex> ctrl_lock(s->ctrl_mu);
  std::unique_lock<std::mutex> lock(s->mu);
  if (graph_id >= s->graphs.size()) return LV_ERR_INVALID;
  auto idle = [&] {
    if (s->graphs[graph_id].busy) return false;
    for (const Request* r : s->pending)
      if (r->graph == graph_id) return false;
    return true;
  };
  s->cv_done.wait(lock, idle);
  Graph& g = s->graphs[graph_id];
  g.busy = true;  // (no request of it can start meanwhile)
  lock.unlock();
  const int rc = lv_dstate_download(s->ctrl, g.s, q, qvtot, qctot);
  lock.lock();
  g.busy = false;
  if (rc != LV_OK) s->err = lv_last_error(s->ctrl);
  s->cv_work.notify_all();
  return rc;
}

int lv_server_graph_tuning(lv_server* s, uint32_t graph_id, lv_tune_choice* out) {
  if (!s || !out) return LV_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  if (graph_id >= s->graphs.size()) return LV_ERR_INVALID;
  *out = s->graphs[graph_id].tuning;
  return LV_OK;
}

int lv_server_get_stats(lv_server* s, lv_server_stats* out) {
  if (!s || !out) return LV_ERR_INVALID;
  std::lock_guard<std::mutex> lock(s->mu);
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->failed = s->failed;
  out->resident_bytes = s->resident();
  out->workspace_bytes = s->workspace();
  out->inflight_bytes = s->inflight;
  out->peak_bytes = s->peak;
  return LV_OK;
}

}  // extern "C"
