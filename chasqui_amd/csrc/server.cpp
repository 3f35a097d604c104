// Continuous-batching serving loop (include/chasqui_amd.h, chq_server_*): runs arrive as decks,
// are admitted against an HBM budget and a per-batch slot limit, join the batch of their scheme
// between steps (their work overlaps on the device with everything already running: one
// batch per scheme, each on its own stream), and leave it when they reach tlim / nlim, which
// frees their slot for the next arrival. The result stays on the device until the client
// releases the run, which returns its memory to the budget.
//
// The loop keeps the device one chunk ahead of the host: every pump collects the chunk the
// previous pump launched, launches the next one, and does the admission work while it runs.
//
// Admission is first-fit over the queue in arrival order, with the head's memory reserved: a
// later run may pass a head that does not fit only if the head would still fit beside it, so
// the head is admitted as soon as the running work ahead of it drains.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <functional>
#include <future>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "ctx.hpp"

using namespace chq;

namespace {

double now_s() {
  using clk = std::chrono::steady_clock;
  static const clk::time_point t0 = clk::now();
  return std::chrono::duration<double>(clk::now() - t0).count();
}

enum : int { JOB_QUEUED = 0, JOB_RUNNING = 1, JOB_DONE = 2, JOB_FAILED = 3, JOB_RELEASED = 4,
             JOB_STARTING = 5 /* admitted, device setup on a worker thread; reported as running */ };

struct Job {
  int id = 0;
  std::string text, overrides;   // kept to set the run up again (FP64 fallback)
  chq_deck *deck = nullptr;
  std::shared_future<int> prep;   // host problem setup running ahead of admission
  std::shared_future<int> start;  // device setup (context, upload, con2prim) after admission
  chq_ctx *started = nullptr;     // the context the start task built
  chq_mesh_desc mesh{};
  chq_config cfg{};
  double tlim = 0.0;
  int nlim = -1;
  size_t bytes = 0;
  int state = JOB_QUEUED;
  int fell_back = 0;
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
  chq_ctx *ctx = nullptr;
  std::string error;
};

using Key = std::tuple<int, int, int, int>;   // reconstruction, Riemann solver, stages, precision
Key key_of(const chq_config &c) { return Key{c.recon, c.rsolver, c.nstage, c.precision}; }

}  // namespace

struct chq_server {
  size_t budget = 0, in_use = 0;
  int slots = 0;        // runs per batch
  int chunk = 1;        // steps per pump between polls
  std::map<int, Job> jobs;
  std::deque<int> queue;
  std::map<Key, chq_batch *> batches;
  std::map<chq_ctx *, int> running;   // context -> job
  std::vector<int> failed;            // runs that could not be created, not yet reported
  std::vector<chq_ctx *> to_free;     // released runs, destroyed when the device is idle
  std::vector<int> starting;          // admitted runs whose device setup is in progress
  std::map<Key, int> starting_per_key;
  bool inflight = false;              // a chunk launched by the last pump is running
  int next_id = 1;
  long long steps_issued = 0;
  double t_launch = 0.0, t_admit = 0.0, t_poll = 0.0, t_retire = 0.0;   // host seconds per phase
};

namespace {

// a mixed-precision run whose state is no longer finite is re-run in FP64
bool state_ok(chq_ctx *c) {
  chq_status st{};
  if (chq_get_status(c, &st)) return false;
  return std::isfinite(st.t) && std::isfinite(st.min_dt) && st.min_dt > 0.0;
}

// start the host problem setup of the next queued runs on worker threads, at most `ahead`
// prepared and not yet admitted at a time (each holds its initial state in host memory)
void prepare_ahead(chq_server *s, int ahead) {
  int live = 0;
  for (int id : s->queue) {
    Job &j = s->jobs[id];
    if (j.prep.valid()) { ++live; continue; }
    if (live >= ahead) break;
    chq_deck *d = j.deck;
    j.prep = std::async(std::launch::async, [d] { return chq_deck_prepare(d); }).share();
    ++live;
  }
}

// Admission reserves the run's memory and slot and hands its device setup (context, upload of
// the initial state, con2prim) to a worker thread on the context's own stream; the run joins
// its batch at the first pump after the setup is done, so the host never waits for uploads
// while the device steps.
int admit_one(chq_server *s, Job &j) {
  const Key key = key_of(j.cfg);
  chq_batch *&b = s->batches[key];
  if (!b) {
    if (int e = chq_batch_create(s->slots, &b)) { s->batches.erase(key); return e; }
  }
  if (chq_batch_size(b) + s->starting_per_key[key] >= s->slots) return 1;   // no free slot
  if (s->starting.size() >= 4) return 1;                                   // setups in flight
  Job *jp = &j;   // map nodes do not move
  std::shared_future<int> prep = j.prep;
  j.start = std::async(std::launch::async, [jp, prep]() -> int {
              if (prep.valid()) prep.wait();
              chq_ctx *c = nullptr;
              if (int e = chq_create(&jp->mesh, &jp->cfg, &c)) { jp->error = chq_last_error(); return e; }
              int e = chq_deck_upload(jp->deck, c);
              if (e == CHQ_EINVAL) jp->error = std::string("problem setup: ") + chq_deck_last_error();
              else if (e) jp->error = chq_last_error();
              if (!e && (e = chq_init(c, 0.0, jp->tlim, jp->nlim))) jp->error = chq_last_error();
              if (e) { chq_destroy(c); return e; }
              jp->started = c;
              return CHQ_OK;
            }).share();
  j.state = JOB_STARTING;
  s->in_use += j.bytes;
  s->starting.push_back(j.id);
  s->starting_per_key[key] += 1;
  return CHQ_OK;
}

// runs whose device setup finished join their batch (or fail); wait = block until all have
int join_started(chq_server *s, bool wait, const std::function<void(int)> &report) {
  for (auto it = s->starting.begin(); it != s->starting.end();) {
    Job &j = s->jobs[*it];
    if (!wait && j.start.wait_for(std::chrono::seconds(0)) != std::future_status::ready) { ++it; continue; }
    const int e = j.start.get();
    s->starting_per_key[key_of(j.cfg)] -= 1;
    it = s->starting.erase(it);
    if (j.deck) { chq_deck_close(j.deck); j.deck = nullptr; }   // the initial state is on the device
    if (e) {
      j.state = JOB_FAILED;
      j.t_done = now_s();
      s->in_use -= j.bytes;
      report(j.id);
      continue;
    }
    chq_batch *b = s->batches[key_of(j.cfg)];
    if (int e2 = chq_batch_add(b, j.started)) {
      chq_destroy(j.started);
      j.started = nullptr;
      j.state = JOB_FAILED;
      j.error = chq_last_error();
      j.t_done = now_s();
      s->in_use -= j.bytes;
      report(j.id);
      (void)e2;
      continue;
    }
    j.ctx = j.started;
    j.started = nullptr;
    j.state = JOB_RUNNING;
    j.t_admit = now_s();
    s->running[j.ctx] = j.id;
  }
  return CHQ_OK;
}

int admit(chq_server *s) {
  size_t reserve = 0;   // memory held for the first run that did not fit
  bool head_blocked = false;
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    Job &j = s->jobs[*it];
    const bool fits = s->in_use + j.bytes + reserve <= s->budget;
    if (fits) {
      const int e = admit_one(s, j);
      if (e == CHQ_OK) { it = s->queue.erase(it); continue; }
      if (e < 0) {   // this run cannot be created at all: fail it, keep serving
        j.state = JOB_FAILED;
        j.error = chq_last_error();
        j.t_done = now_s();
        s->failed.push_back(j.id);
        it = s->queue.erase(it);
        continue;
      }
      // no slot in its batch: it waits without reserving memory
    } else if (!head_blocked) {
      head_blocked = true;
      reserve = j.bytes;
    }
    ++it;
  }
  return CHQ_OK;
}

}  // namespace

extern "C" {

int chq_server_create(size_t hbm_budget, int slots_per_batch, int chunk_steps, chq_server **out) {
  if (!out || hbm_budget == 0 || slots_per_batch < 1 || chunk_steps < 1)
    return fail(CHQ_EINVAL, "budget, slots and chunk must be positive");
  auto *s = new (std::nothrow) chq_server();
  if (!s) return fail(CHQ_ENOMEM, "server");
  s->budget = hbm_budget;
  s->slots = slots_per_batch;
  s->chunk = chunk_steps;
  *out = s;
  return CHQ_OK;
}

int chq_server_destroy(chq_server *s) {
  if (!s) return CHQ_OK;
  for (auto &kv : s->batches) chq_batch_destroy(kv.second);
  for (chq_ctx *c : s->to_free) chq_destroy(c);
  for (auto &kv : s->jobs) {
    if (kv.second.start.valid()) {
      kv.second.start.wait();
      if (kv.second.started) chq_destroy(kv.second.started);
    }
    if (kv.second.prep.valid()) kv.second.prep.wait();
    if (kv.second.ctx) chq_destroy(kv.second.ctx);
    if (kv.second.deck) chq_deck_close(kv.second.deck);
  }
  delete s;
  return CHQ_OK;
}

// precision: -1 = the deck's, 0 = FP64, 1 = mixed (FP64 when the deck needs dual energy)
int chq_server_submit(chq_server *s, const char *deck_text, const char *overrides, int precision, int *job_id) {
  if (!s || !deck_text || !job_id) return fail(CHQ_EINVAL, "bad argument");
  Job j;
  // parse now; the problem setup runs at admission, on the host while batches step
  if (chq_deck_open_lazy(deck_text, overrides, &j.deck))
    return fail(CHQ_EINVAL, std::string("deck: ") + chq_deck_last_error());
  int ndiag = 0;
  chq_deck_mesh(j.deck, &j.mesh);
  chq_deck_config(j.deck, &j.cfg);
  chq_deck_run(j.deck, &j.tlim, &j.nlim, &ndiag);
  if (precision >= 0) j.cfg.precision = (precision == 1 && !j.cfg.dual_energy) ? 1 : 0;
  if (j.cfg.dual_energy) { chq_deck_close(j.deck); return fail(CHQ_EUNSUP, "served runs need dual energy off"); }
  if (j.cfg.characteristic && j.cfg.recon != 0) {
    chq_deck_close(j.deck);
    return fail(CHQ_EUNSUP, "served runs need primitive-variable reconstruction");
  }
  if (j.cfg.nu > 0.0 || j.cfg.kappa > 0.0 || j.cfg.eta > 0.0 || j.cfg.sts) {
    chq_deck_close(j.deck);
    return fail(CHQ_EUNSUP, "served runs need diffusion off");
  }
  if (j.cfg.flux_limiter || j.cfg.grav_uniform || j.cfg.grav_point || j.cfg.cooling) {
    chq_deck_close(j.deck);
    return fail(CHQ_EUNSUP, "served runs need the flux limiter and source terms off");
  }
  j.bytes = chq_bytes_for(&j.mesh, &j.cfg);
  if (j.bytes == 0) { chq_deck_close(j.deck); return fail(CHQ_EINVAL, "invalid mesh or configuration"); }
  if (j.bytes > s->budget) { chq_deck_close(j.deck); return fail(CHQ_ENOMEM, "run larger than the HBM budget"); }
  j.text = deck_text;
  j.overrides = overrides ? overrides : "";
  j.id = s->next_id++;
  j.t_submit = now_s();
  s->queue.push_back(j.id);
  *job_id = j.id;
  s->jobs.emplace(j.id, j);
  return CHQ_OK;
}

// One iteration of the loop. The device runs one chunk (`chunk` steps of every batch, batches
// concurrently on their own streams) per call, ahead of the host: a call first collects the
// chunk launched by the previous call (retiring the runs that finished), frees released runs
// while the device is idle, launches the next chunk, and then does the admission work (deck
// setup, uploads, con2prim of new runs on their own streams) while that chunk runs, returning
// without waiting for it. done_ids receives up to max_done finished (or failed) job ids.
int chq_server_pump(chq_server *s, int max_done, int *done_ids, int *n_done) {
  if (!s || !n_done) return fail(CHQ_EINVAL, "bad argument");
  *n_done = 0;
  auto report = [&](int id) {
    if (done_ids && *n_done < max_done) done_ids[*n_done] = id;
    ++*n_done;
  };
  const double t0 = now_s();
  const double poll0 = s->t_poll;   // the retire phase below excludes the time spent waiting on the device
  if (s->inflight) {
    std::vector<int> flags;
    for (auto &kv : s->batches) {
      chq_batch *b = kv.second;
      const int n = chq_batch_size(b);
      if (!n) continue;
      flags.assign(size_t(n), 1);
      int live = 0;
      const double tp = now_s();
      if (int e = chq_batch_unfinished(b, &live, flags.data())) return e;
      s->t_poll += now_s() - tp;
      if (live == n) continue;
      std::vector<chq_ctx *> gone;
      for (int r = 0; r < n; ++r)
        if (!flags[size_t(r)]) gone.push_back(chq_batch_run(b, r));
      const double t = now_s();
      for (chq_ctx *c : gone) {
        if (int e = chq_batch_remove(b, c)) return e;
        auto rj = s->running.find(c);
        if (rj == s->running.end()) continue;
        Job &j = s->jobs[rj->second];
        s->running.erase(rj);
        if (j.cfg.precision == 1 && !state_ok(c)) {   // FP64 fallback: run again from the start
          chq_destroy(c);
          j.ctx = nullptr;
          s->in_use -= j.bytes;
          j.cfg.precision = 0;
          j.bytes = chq_bytes_for(&j.mesh, &j.cfg);
          j.fell_back = 1;
          j.prep = std::shared_future<int>();
        if (chq_deck_open_lazy(j.text.c_str(), j.overrides.c_str(), &j.deck)) {
            j.state = JOB_FAILED;
            j.error = std::string("deck: ") + chq_deck_last_error();
            j.t_done = t;
            report(j.id);
            continue;
          }
          j.state = JOB_QUEUED;
          s->queue.push_front(j.id);
        } else {
          j.state = JOB_DONE;
          j.t_done = t;
          report(j.id);
        }
      }
    }
    s->inflight = false;
  }
  for (chq_ctx *c : s->to_free) chq_destroy(c);   // the device is idle here
  s->to_free.clear();
  if (int e = join_started(s, false, report)) return e;   // finished setups join this chunk
  const double t1 = now_s();
  s->t_retire += (t1 - t0) - (s->t_poll - poll0);
  for (auto &kv : s->batches)
    if (chq_batch_size(kv.second) > 0) {
      if (int e = chq_batch_step(kv.second, s->chunk)) return e;
      s->steps_issued += s->chunk;
      s->inflight = true;
    }
  const double t2 = now_s();
  s->t_launch += t2 - t1;
  if (int e = admit(s)) return e;
  for (int id : s->failed) report(id);
  s->failed.clear();
  prepare_ahead(s, 4);
  s->t_admit += now_s() - t2;
  return CHQ_OK;
}

// info[0..9] = state, precision, fell_back, cycle, t, t_submit, t_admit, t_done, bytes, fofc
int chq_server_job(chq_server *s, int job_id, double info[10]) {
  if (!s || !info) return fail(CHQ_EINVAL, "bad argument");
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(CHQ_EINVAL, "unknown job");
  const Job &j = it->second;
  chq_status st{};
  if (j.ctx && chq_get_status(j.ctx, &st)) return CHQ_EHIP;
  info[0] = j.state == JOB_STARTING ? JOB_RUNNING : j.state;
  info[1] = j.cfg.precision;
  info[2] = j.fell_back;
  info[3] = st.cycle;
  info[4] = st.t;
  info[5] = j.t_submit;
  info[6] = j.t_admit;
  info[7] = j.t_done;
  info[8] = double(j.bytes);
  info[9] = double(st.fofc_total);
  return CHQ_OK;
}

chq_ctx *chq_server_job_ctx(chq_server *s, int job_id) {
  if (!s) return nullptr;
  auto it = s->jobs.find(job_id);
  return it == s->jobs.end() ? nullptr : it->second.ctx;
}

const char *chq_server_job_error(chq_server *s, int job_id) {
  if (!s) return "";
  auto it = s->jobs.find(job_id);
  return it == s->jobs.end() ? "" : it->second.error.c_str();
}

// free a finished run's device memory (its result is gone after this)
int chq_server_release(chq_server *s, int job_id) {
  if (!s) return fail(CHQ_EINVAL, "null server");
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return fail(CHQ_EINVAL, "unknown job");
  Job &j = it->second;
  if (j.state != JOB_DONE && j.state != JOB_FAILED) return fail(CHQ_EINVAL, "job has not finished");
  if (j.ctx) {   // freeing would wait for the running chunk: defer to the next pump
    if (s->inflight) s->to_free.push_back(j.ctx);
    else chq_destroy(j.ctx);
    j.ctx = nullptr;
    s->in_use -= j.bytes;
  }
  if (j.prep.valid()) j.prep.wait();
  if (j.deck) { chq_deck_close(j.deck); j.deck = nullptr; }
  j.state = JOB_RELEASED;
  return CHQ_OK;
}

// stats[0..10] = queued, running, finished (not released), in_use bytes, budget, batches,
// batch steps, host seconds launching steps / admitting / waiting for polls / retiring
int chq_server_stats(chq_server *s, double stats[11]) {
  if (!s || !stats) return fail(CHQ_EINVAL, "bad argument");
  int q = 0, r = 0, d = 0;
  for (auto &kv : s->jobs) {
    q += kv.second.state == JOB_QUEUED;
    r += kv.second.state == JOB_RUNNING || kv.second.state == JOB_STARTING;
    d += kv.second.state == JOB_DONE || kv.second.state == JOB_FAILED;
  }
  stats[0] = q;
  stats[1] = r;
  stats[2] = d;
  stats[3] = double(s->in_use);
  stats[4] = double(s->budget);
  stats[5] = double(s->batches.size());
  stats[6] = double(s->steps_issued);
  stats[7] = s->t_launch;
  stats[8] = s->t_admit;
  stats[9] = s->t_poll;
  stats[10] = s->t_retire;
  return CHQ_OK;
}

double chq_server_clock(void) { return now_s(); }

}  // extern "C"
