// Serving loop: requests (one cell each, a number of steps) arrive at any
// time; a scheduler thread admits them into one running batch while the
// reserved slots, instance slots and cell keys allow (first come, first
// served), steps the batch, and retires each request when it has run its
// steps -- its final state is copied out and its slots are dropped by the
// next sort.  Because every instance of a batch evolves exactly as alone,
// a request's result does not depend on what else was running.
// No reference counterpart: the reference runs one cell per process.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/spatial_cell.h"
#include "sc_batch.h"
#include "sc_json.h"
#include "sc_setup.h"

namespace {

struct Request {
  int64_t id = 0;
  scs::HostSetup h;
  int steps = 0;
  int state = SC_REQ_QUEUED;
  int slot = -1;
  int steps_done = 0;
  double t_submit = -1.0, t_admit = -1.0, t_done = -1.0;
  size_t need_plain = 0, need_metabolic = 0;   // reserved slots: capacities of the cell
  sc_counts counts{};
  int stage = -1;          // retire stage holding the final state (from retirement until release)
  scb::Retired ret;
  std::string error;
};

thread_local std::string g_start_error;

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct sc_server {
  sc_sim* sim = nullptr;
  scs::HostSetup templ;
  scb::ServerLimits lim{};
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::thread worker;
  bool stopping = false;
  std::deque<int64_t> queue;
  std::map<int64_t, Request> req;
  std::vector<int64_t> running;
  std::vector<int> free_slots;
  std::vector<std::pair<int, std::pair<size_t, size_t>>> leaving;   // slot, reserved slots: free after a sort
  std::vector<int64_t> copying;                 // retired, final state on its way to the host
  std::vector<scb::RetireStage> stages;         // retire stages; free ones listed in free_stages
  std::vector<int> free_stages;
  size_t reserved_plain = 0, reserved_metabolic = 0;
  size_t stage_limit = 0;                       // largest retire stage to allocate (0: no limit)
  int64_t next_id = 1;
  std::string err;
  uint64_t steps_run = 0;
  sc_server_stats stats{};   // guarded by mu
};

// seconds between two points of the scheduler's loop
struct Lap {
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  double next() {
    const auto n = std::chrono::steady_clock::now();
    const double d = std::chrono::duration<double>(n - t).count();
    t = n;
    return d;
  }
};

// A free retire stage of at least `bytes` (the smallest that fits), or a new
// one of the next power of two; -1 when allocation fails.  Called by the
// scheduler thread with S->mu held.
static int take_stage(sc_server* S, size_t bytes) {
  int best = -1;
  for (size_t q = 0; q < S->free_stages.size(); q++) {
    const int k = S->free_stages[q];
    if (S->stages[k].bytes >= bytes && (best < 0 || S->stages[k].bytes < S->stages[S->free_stages[best]].bytes))
      best = (int)q;
  }
  if (best >= 0) {
    const int k = S->free_stages[best];
    S->free_stages.erase(S->free_stages.begin() + best);
    return k;
  }
  size_t b = (size_t)1 << 20;
  while (b < bytes) b *= 2;
  if (S->stage_limit && b > S->stage_limit) return -1;
  scb::RetireStage st;
  if (scb::server_stage_alloc(S->sim, b, st) != SC_OK) return -1;
  S->stages.push_back(st);
  return (int)S->stages.size() - 1;
}

// Requests whose final state has reached the host are done.
static void finish_copies(sc_server* S) {
  std::lock_guard<std::mutex> lk(S->mu);
  for (size_t q = 0; q < S->copying.size();) {
    Request& r = S->req[S->copying[q]];
    if (!scb::server_retire_ready(S->stages[r.stage])) { q++; continue; }
    scb::server_retire_counts(S->stages[r.stage], r.ret, &r.counts);
    r.state = SC_REQ_DONE;
    r.t_done = seconds_since(S->t0);
    S->copying.erase(S->copying.begin() + q);
    S->cv.notify_all();
  }
}

static void serve(sc_server* S) {
  std::vector<int64_t> done;
  while (true) {
    finish_copies(S);
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (S->stopping) return;
      // admit, first come first served
      while (!S->queue.empty()) {
        Request& r = S->req[S->queue.front()];
        if (S->free_slots.empty() || S->reserved_plain + r.need_plain > (size_t)S->lim.cap_plain ||
            S->reserved_metabolic + r.need_metabolic > (size_t)S->lim.cap_metabolic)
          break;
        const int slot = S->free_slots.back();
        // kernel grids must cover every slot in use, the newcomer's included
        scb::server_set_launch(S->sim, S->reserved_plain + r.need_plain, S->reserved_metabolic + r.need_metabolic);
        lk.unlock();
        Lap lap;
        sc_status st = scb::server_admit(S->sim, slot, r.h);
        lk.lock();
        S->stats.t_admit += lap.next();
        if (st == SC_OK) S->stats.admitted++;
        if (st != SC_OK) scb::server_set_launch(S->sim, S->reserved_plain, S->reserved_metabolic);
        if (st == SC_ERR_BUDGET && !(S->running.empty() && S->leaving.empty())) break;   // wait for room
        S->queue.pop_front();
        if (st != SC_OK) {
          r.state = SC_REQ_FAILED;
          r.error = scb::sim_error(S->sim);
          r.t_done = seconds_since(S->t0);
          continue;
        }
        S->free_slots.pop_back();
        r.slot = slot;
        r.state = SC_REQ_RUNNING;
        r.t_admit = seconds_since(S->t0);
        S->reserved_plain += r.need_plain;
        S->reserved_metabolic += r.need_metabolic;
        S->running.push_back(r.id);
        r.h.plain.clear(); r.h.plain.shrink_to_fit();
        r.h.metabolic.clear(); r.h.metabolic.shrink_to_fit();
        if (r.steps == 0) done.push_back(r.id);
      }
      if (S->running.empty() && S->leaving.empty()) {
        Lap lap;
        S->cv.wait_for(lk, std::chrono::milliseconds(S->copying.empty() ? 2 : 1));
        S->stats.t_idle += lap.next();
        continue;
      }
    }
    // one step of every resident cell (outside the lock: submit and poll stay responsive)
    if (done.empty()) {
      Lap lap;
      sc_status st = sc_step(S->sim, 1);
      if (st == SC_OK) st = sc_synchronize(S->sim);
      std::lock_guard<std::mutex> lk(S->mu);
      S->stats.t_step += lap.next();
      S->stats.steps++;
      S->stats.slot_steps += (double)(S->reserved_plain + S->reserved_metabolic);
      if (st != SC_OK) {
        S->err = sc_last_error(S->sim);
        for (int64_t id : S->running) { S->req[id].state = SC_REQ_FAILED; S->req[id].error = S->err; }
        S->running.clear();
        S->stopping = true;
        S->cv.notify_all();
        return;
      }
      S->steps_run++;
      // the sort of this step dropped the cells that left before it
      for (auto& l : S->leaving) {
        S->free_slots.push_back(l.first);
        S->reserved_plain -= l.second.first;
        S->reserved_metabolic -= l.second.second;
      }
      S->leaving.clear();
      scb::server_set_launch(S->sim, S->reserved_plain, S->reserved_metabolic);
      for (int64_t id : S->running) {
        Request& r = S->req[id];
        if (++r.steps_done >= r.steps) done.push_back(id);
      }
    }
    // retire: gather the final state and start its copy to the host (it
    // completes beside the next steps), then leave the batch
    Lap retire_lap;
    for (int64_t id : done) {
      Request* r;
      int stage;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        r = &S->req[id];
        r->ret.cap[0] = r->need_plain;
        r->ret.cap[1] = r->need_metabolic;
        stage = take_stage(S, scb::server_retire_bytes(r->need_plain, r->need_metabolic));
      }
      sc_status st = stage < 0 ? SC_ERR_OOM : scb::server_retire_begin(S->sim, r->slot, S->stages[stage], r->ret);
      std::string why;
      if (st != SC_OK) why = stage < 0 ? std::string("no memory for the final state") : scb::sim_error(S->sim);
      // the cell leaves the batch whether or not its final state could be
      // copied out; if it cannot even be evicted, its slot and reservation
      // stay taken (its particles keep stepping, so the launch bounds must
      // keep covering them)
      const sc_status ev = scb::server_evict(S->sim, r->slot);
      if (st != SC_OK && stage >= 0) scb::server_drain(S->sim);   // no copy into the stage is in flight
      std::lock_guard<std::mutex> lk(S->mu);
      S->running.erase(std::remove(S->running.begin(), S->running.end(), id), S->running.end());
      if (st == SC_OK && ev != SC_OK) {
        st = ev;
        why = scb::sim_error(S->sim);
        scb::server_drain(S->sim);
      }
      if (st != SC_OK) {
        r->state = SC_REQ_FAILED;
        r->error = why;
        r->t_done = seconds_since(S->t0);
        if (stage >= 0) S->free_stages.push_back(stage);
      } else {
        r->stage = stage;
        S->copying.push_back(id);
      }
      if (ev == SC_OK) S->leaving.push_back({r->slot, {r->need_plain, r->need_metabolic}});
      S->stats.retired++;
      S->cv.notify_all();
    }
    if (!done.empty()) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stats.t_retire += retire_lap.next();
    }
    done.clear();
  }
}

extern "C" {

sc_status sc_server_start(const char* config_json, const char* setup_json, const char* server_json, int device,
                          sc_server** out) {
  if (!out) return SC_ERR_INVALID;
  *out = nullptr;
  sc_server* S = new sc_server();
  std::string err;
  if (!scs::build_setup(config_json, setup_json, S->templ, err)) {
    g_start_error = "template: " + err;
    delete S;
    return SC_ERR_INVALID;
  }
  scj::Value v;
  if (!server_json || !scj::parse(server_json, v, err) || v.kind != scj::Value::Object) {
    g_start_error = "server document: " + (err.empty() ? std::string("not an object") : err);
    delete S;
    return SC_ERR_INVALID;
  }
  S->lim.max_instances = v.get_int("max_instances", 256);
  // largest retire stage (device + pinned host copy of a finished cell) the
  // server allocates; a cell needing more fails.  0 = no limit.
  S->stage_limit = (size_t)std::max(0.0, v.get_double("max_retire_stage_bytes", 0.0));
  S->lim.max_cells = (uint32_t)v.get_double("max_cells", (double)(1u << 28));
  if (v.has("hbm_budget_bytes") && !v.has("max_particles")) {
    // slots from a device-memory budget: per plain and per metabolic slot
    // the bytes the batch allocates, split by metabolic_fraction of slots
    const double budget = v.get_double("hbm_budget_bytes", 0.0);
    const double f = v.get_double("metabolic_fraction", 0.55);
    const double cells_bytes = 2.0 * 4.0 * (double)S->lim.max_cells;
    const double per = (1.0 - f) * scb::bytes_per_slot(false) + f * scb::bytes_per_slot(true);
    const double slots = std::max(0.0, (budget - cells_bytes) / per);
    S->lim.cap_plain = (int)std::min(slots * (1.0 - f), (double)((1 << 30) - 1));
    S->lim.cap_metabolic = (int)std::min(slots * f, (double)((1 << 30) - 1));
  } else {
    S->lim.cap_plain = v.get_int("max_particles", 1 << 20);
    S->lim.cap_metabolic = v.get_int("max_metabolic_particles", 1 << 20);
  }
  sc_status st = scb::server_sim_create(S->templ, S->lim, device, &S->sim, err);
  if (st != SC_OK) {
    g_start_error = err;
    delete S;
    return st;
  }
  for (int i = S->lim.max_instances - 1; i >= 0; i--) S->free_slots.push_back(i);
  S->t0 = std::chrono::steady_clock::now();
  S->worker = std::thread(serve, S);
  *out = S;
  return SC_OK;
}

sc_status sc_server_submit(sc_server* S, const char* config_json, const char* setup_json, int steps, int64_t* id) {
  if (!S || !id || steps < 0) return SC_ERR_INVALID;
  Request r;
  std::string err, why;
  if (!scs::build_setup(config_json, setup_json, r.h, err)) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = err;
    return SC_ERR_INVALID;
  }
  if (!scb::server_compatible(S->templ, r.h, why)) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = "request differs from the server's template in " + why;
    return SC_ERR_INVALID;
  }
  r.need_plain = (size_t)std::max(r.h.P.num_particles, 0);
  r.need_metabolic = (size_t)std::max(r.h.P.num_metabolic_particles, 0);
  if (r.need_plain > (size_t)S->lim.cap_plain || r.need_metabolic > (size_t)S->lim.cap_metabolic) {
    std::lock_guard<std::mutex> lk(S->mu);
    S->err = "request larger than the server's capacity";
    return SC_ERR_BUDGET;
  }
  r.steps = steps;
  std::lock_guard<std::mutex> lk(S->mu);
  r.id = S->next_id++;
  r.t_submit = seconds_since(S->t0);
  *id = r.id;
  S->queue.push_back(r.id);
  S->req.emplace(r.id, std::move(r));
  S->cv.notify_all();
  return SC_OK;
}

sc_status sc_server_poll(sc_server* S, int64_t id, sc_request_info* info) {
  if (!S || !info) return SC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->req.find(id);
  if (it == S->req.end()) return SC_ERR_INVALID;
  const Request& r = it->second;
  info->id = r.id;
  info->state = r.state;
  info->steps = r.steps;
  info->steps_done = r.steps_done;
  info->t_submit = r.t_submit;
  info->t_admit = r.t_admit;
  info->t_done = r.t_done;
  info->counts = r.counts;
  return SC_OK;
}

sc_status sc_server_wait(sc_server* S, int64_t id, double timeout_s) {
  if (!S) return SC_ERR_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->req.find(id);
  if (it == S->req.end()) return SC_ERR_INVALID;
  const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double>(std::max(timeout_s, 0.0));
  while (it->second.state == SC_REQ_QUEUED || it->second.state == SC_REQ_RUNNING) {
    if (S->stopping) return SC_ERR_DEVICE;
    if (S->cv.wait_until(lk, until) == std::cv_status::timeout) return SC_ERR_BUDGET;
  }
  return it->second.state == SC_REQ_DONE ? SC_OK : SC_ERR_DEVICE;
}

sc_status sc_server_result(sc_server* S, int64_t id, sc_particle* plain, size_t plain_cap, sc_particle* metabolic,
                           size_t metabolic_cap) {
  if (!S) return SC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->req.find(id);
  if (it == S->req.end() || it->second.state != SC_REQ_DONE) return SC_ERR_INVALID;
  const Request& r = it->second;
  if (r.stage < 0) return SC_ERR_INVALID;
  scb::server_retire_decode(S->stages[r.stage], r.ret, plain, plain_cap, metabolic, metabolic_cap);
  return SC_OK;
}

sc_status sc_server_release(sc_server* S, int64_t id) {
  if (!S) return SC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->req.find(id);
  if (it == S->req.end() || (it->second.state != SC_REQ_DONE && it->second.state != SC_REQ_FAILED))
    return SC_ERR_INVALID;
  if (it->second.stage >= 0) S->free_stages.push_back(it->second.stage);
  S->req.erase(it);
  return SC_OK;
}

sc_status sc_server_stop(sc_server* S) {
  if (!S) return SC_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
    S->cv.notify_all();
  }
  if (S->worker.joinable()) S->worker.join();
  for (scb::RetireStage& st : S->stages) scb::server_stage_free(S->sim, st);
  sc_destroy(S->sim);
  delete S;
  return SC_OK;
}

sc_status sc_server_get_stats(sc_server* S, sc_server_stats* out) {
  if (!S || !out) return SC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  *out = S->stats;
  return SC_OK;
}

const char* sc_server_last_error(sc_server* S) {
  if (!S) return g_start_error.c_str();
  std::lock_guard<std::mutex> lk(S->mu);
  return S->err.c_str();
}

}  // extern "C"
