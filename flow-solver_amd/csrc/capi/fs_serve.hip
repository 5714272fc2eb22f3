// Serving: a C++ request loop over resident steady solves.
//
// Threads: `prep_threads` host workers run the mesh preprocessing of queued
// requests; one device thread owns every engine.  At each pseudo-time-step
// boundary the device thread admits prepared requests (first come, first
// served) while their footprint fits the device budget and packs the admitted
// ones into one new batched engine.  Every engine runs on its own stream and
// advances on its own: whenever its previous step has finished, the device
// thread retires the solves that met the stopping test and enqueues the next
// step at once, so the steps of all resident solves overlap on the device and
// a small solve is not held to the pace of a large one.  An engine whose
// solves have all finished is destroyed and its memory returns to the
// budget.  A solve's arithmetic depends only on its own mesh and controls,
// so its result does not depend on what it shares the device with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fs_capi.h"
#include "fs_capi_internal.hpp"
#include "fs_engine.hpp"

namespace {

struct Request {
  int64_t id = 0;
  std::unique_ptr<fs_mesh> mesh;
  fs_params prm{};
  double t_submit = 0, t_admit = -1, t_done = -1;
  int status = FS_INST_RUNNING;
  int steps = 0;            // steps taken, those a recovery rolled back included
  int limit = 0;            // step limit
  fse::RecoveryTrack rt;
  double r0 = -1.0, drop = 1.0;
  int64_t bytes = 0;
  int nn = 0, ng = 0;
  bool cancel = false;
  bool finished = false;
  std::string error;
  std::vector<double> Q;
  double forces[2] = {0., 0.};
};

struct Group {
  std::unique_ptr<fse::Engine> eng;
  std::vector<Request*> reqs;                    // instance i -> request, null once finished
  std::vector<std::unique_ptr<fs_mesh>> meshes;  // the engine's host meshes
  int64_t bytes = 0;
  int live = 0;
};

int ts_ref(const fs_params& p) { return (p.order == 3) ? p.foits + p.soits + 1 : p.foits + 1; }

}  // namespace

struct fs_server {
  fs_server_config cfg{};
  int device = 0;
  int64_t budget = 0;
  std::chrono::steady_clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_prep, cv_gpu, cv_done;
  bool stop = false;
  int64_t next_id = 1;
  std::map<int64_t, std::unique_ptr<Request>> all;
  std::deque<Request*> to_prepare, ready;
  std::deque<Request*> done;
  int running = 0;
  int64_t bytes_in_use = 0;

  std::vector<std::unique_ptr<Group>> groups;  // device thread only
  std::vector<std::thread> workers;
  std::thread gpu;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  // caller holds mu
  void finish_locked(Request* r, int status, const std::string& err) {
    r->status = status;
    r->error = err;
    r->t_done = now();
    r->finished = true;
    done.push_back(r);
    cv_done.notify_all();
  }

  void prep_loop() {
    for (;;) {
      Request* r = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_prep.wait(lk, [&] { return stop || !to_prepare.empty(); });
        if (stop) return;
        r = to_prepare.front();
        to_prepare.pop_front();
      }
      std::string err;
      std::string why;
      bool ok = true;
      if (!fse::supported(r->prm, &why)) {
        ok = false;
        err = why;
      } else if (!r->mesh->prepared) {
        if (!fs::prepare(&r->mesh->m, fse::to_params(&r->prm), &err)) ok = false;
        else {
          r->mesh->prepared = true;
          r->mesh->order_prepared = r->prm.order;
        }
      } else if (r->prm.order == 3 && r->mesh->order_prepared != 3) {
        ok = false;
        err = "order 3 needs a mesh prepared with order 3";
      }
      std::lock_guard<std::mutex> lk(mu);
      if (!ok || r->cancel) {
        finish_locked(r, r->cancel ? FS_INST_CANCELLED : FS_INST_FAILED, r->cancel ? "cancelled" : err);
        continue;
      }
      r->bytes = (int64_t)fse::footprint_bytes(r->mesh->m, cfg.precision, cfg.max_recoveries > 0);
      r->nn = r->mesh->m.nn;
      r->ng = r->mesh->m.ng;
      r->limit = cfg.max_steps;
      if (cfg.max_steps_per_sqrt_node > 0)
        r->limit = std::max(cfg.max_steps, (int)std::ceil(cfg.max_steps_per_sqrt_node * std::sqrt((double)r->nn)));
      ready.push_back(r);
      cv_gpu.notify_all();
    }
  }

  void fail_group(Group* g, const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu);
    for (size_t i = 0; i < g->reqs.size(); ++i) {
      Request* r = g->reqs[i];
      if (!r) continue;
      g->reqs[i] = nullptr;
      --running;
      finish_locked(r, FS_INST_FAILED, msg);
    }
    g->live = 0;
  }

  void admit() {
    std::vector<Request*> take;
    int64_t bytes = 0;
    {
      std::lock_guard<std::mutex> lk(mu);
      while (!ready.empty()) {
        Request* r = ready.front();
        if (r->cancel) {
          ready.pop_front();
          finish_locked(r, FS_INST_CANCELLED, "cancelled");
          continue;
        }
        if (r->bytes > budget) {
          ready.pop_front();
          finish_locked(r, FS_INST_FAILED, "device footprint exceeds the server budget");
          continue;
        }
        if (bytes_in_use + bytes + r->bytes > budget) break;  // first come, first served
        ready.pop_front();
        take.push_back(r);
        bytes += r->bytes;
      }
      bytes_in_use += bytes;
      running += (int)take.size();
    }
    if (take.empty()) return;
    auto g = std::make_unique<Group>();
    g->reqs = take;
    g->bytes = bytes;
    g->live = (int)take.size();
    std::vector<const fs::Mesh*> ms;
    std::vector<fs_params> ps;
    for (Request* r : take) {
      ms.push_back(&r->mesh->m);
      ps.push_back(r->prm);
      g->meshes.push_back(std::move(r->mesh));
    }
    g->eng = std::make_unique<fse::Engine>();
    int rc = g->eng->create(ms, ps, cfg.precision);
    if (!rc) rc = g->eng->init_freestream(-1);
    if (!rc && cfg.max_recoveries > 0) rc = g->eng->enable_checkpoints();
    for (auto& I : g->eng->inst) I.freeze = cfg.limiter_freeze;
    // the engine's step counters start at 0: each solve runs its own schedule
    if (!rc) rc = g->eng->sync();
    double t = now();
    if (rc) {
      fail_group(g.get(), std::string("admission failed: ") + fse::last_error_text());
      std::lock_guard<std::mutex> lk(mu);
      bytes_in_use -= g->bytes;
      return;
    }
    for (Request* r : take) r->t_admit = t;
    groups.push_back(std::move(g));
  }

  // Bookkeeping after a finished step of engine g: the recovery policy's
  // restores and checkpoints, the stopping test per solve (as
  // fs_solve_steady), final states copied out and the body forces of the
  // finished solves from one pass over the engine, finished solves masked off.
  void retire_finished(Group* g) {
    fse::Engine& E = *g->eng;
    const fse::RecoveryPolicy pol{cfg.max_recoveries, cfg.checkpoint_every, cfg.limiter_fallback,
                                  cfg.first_order_fallback != 0, cfg.stall_steps_per_sqrt_node};
    std::vector<int> mode(E.ninst(), 0);
    std::vector<int> fin_inst, fin_status;
    std::vector<std::string> fin_err;
    for (int i = 0; i < E.ninst(); ++i) {
      fse::Instance& I = E.inst[i];
      Request* r = g->reqs[i];
      if (!I.active || !r) continue;
      ++r->steps;
      const double rn = std::sqrt(I.r2);
      const bool failed = I.err != 0 || !(rn == rn);
      if (E.recover_or_checkpoint(i, failed, rn, pol, r->rt, mode)) continue;
      int status = FS_INST_RUNNING;
      std::string err;
      if (failed) {
        status = FS_INST_FAILED;
        err = I.err ? fse::device_error_message(I.err) : "non-finite residual";
      } else {
        int tr = ts_ref(r->prm);  // as submitted (a fallback may change the engine's copy)
        if (I.ts == tr || (I.ts > tr && r->r0 < 0)) r->r0 = rn;
        if (r->r0 > 0) r->drop = rn / r->r0;
        if (r->r0 > 0 && r->drop <= cfg.tol) status = FS_INST_CONVERGED;
        else if (r->steps >= r->limit) status = FS_INST_MAX_STEPS;
      }
      if (status == FS_INST_RUNNING) continue;
      fin_inst.push_back(i);
      fin_status.push_back(status);
      fin_err.push_back(err);
    }
    if (E.checkpoints_enabled() && E.checkpoint_pass(mode)) {
      fail_group(g, fse::last_error_text());
      return;
    }
    if (fin_inst.empty()) return;
    // main.C:874-876: the body forces of the final states, one pass for all
    std::vector<double> f;
    bool forces_ok = false;
    for (int st : fin_status) forces_ok = forces_ok || st == FS_INST_CONVERGED || st == FS_INST_MAX_STEPS;
    if (forces_ok) {
      std::vector<int> citer(E.ninst());
      std::vector<char> want(E.ninst(), 0);
      f.assign(2 * (size_t)E.ninst(), 0.0);
      for (int k = 0; k < E.ninst(); ++k) citer[k] = E.inst[k].ts;
      for (int k : fin_inst) want[k] = 1;
      forces_ok = E.body_forces(citer.data(), f.data(), &want) == FS_OK;
    }
    for (size_t k = 0; k < fin_inst.size(); ++k) {
      const int i = fin_inst[k];
      fse::Instance& I = E.inst[i];
      Request* r = g->reqs[i];
      int status = fin_status[k];
      std::string err = fin_err[k];
      int64_t n = 4 * ((int64_t)I.nn + I.nbe);
      r->Q.assign((size_t)n, 0.0);
      if (E.get_field(i, FS_F_Q, r->Q.data(), n)) {
        status = FS_INST_FAILED;
        err = fse::last_error_text();
      }
      if (forces_ok && (status == FS_INST_CONVERGED || status == FS_INST_MAX_STEPS)) {
        r->forces[0] = f[2 * (size_t)i];
        r->forces[1] = f[2 * (size_t)i + 1];
      }
      I.active = false;
      --g->live;
      g->reqs[i] = nullptr;
      std::lock_guard<std::mutex> lk(mu);
      --running;
      finish_locked(r, status, err);
    }
  }

  // Advances engine g if its previous step has finished: bookkeeping, then
  // the next step is enqueued at once.  Engines progress independently, so a
  // small solve is not held to the step rate of a large one.  Returns whether
  // anything happened.
  bool service(Group* g) {
    if (g->live == 0) return false;
    fse::Engine& E = *g->eng;
    if (E.step_pending()) {
      if (!E.step_ready()) return false;
      if (E.step_end()) {
        fail_group(g, fse::last_error_text());
        return true;
      }
      retire_finished(g);
    }
    {
      // cancellations take effect at the step boundary
      std::lock_guard<std::mutex> lk(mu);
      for (size_t i = 0; i < g->reqs.size(); ++i) {
        Request* r = g->reqs[i];
        if (r && r->cancel) {
          E.inst[i].active = false;
          g->reqs[i] = nullptr;
          --g->live;
          --running;
          finish_locked(r, FS_INST_CANCELLED, "cancelled");
        }
      }
    }
    if (g->live > 0 && E.step_begin(true, cfg.use_graph != 0)) fail_group(g, fse::last_error_text());
    return true;
  }

  void retire_groups() {
    for (auto it = groups.begin(); it != groups.end();) {
      if ((*it)->live > 0) {
        ++it;
        continue;
      }
      int64_t b = (*it)->bytes;
      it = groups.erase(it);  // the engine's destructor waits for its stream
      std::lock_guard<std::mutex> lk(mu);
      bytes_in_use -= b;
    }
  }

  void gpu_loop() {
    (void)hipSetDevice(device);
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_gpu.wait(lk, [&] { return stop || !groups.empty() || !ready.empty(); });
        if (stop) break;
      }
      admit();
      bool any = false;
      for (auto& g : groups) any = service(g.get()) || any;
      retire_groups();
      if (!any) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    groups.clear();
  }
};

namespace {
int fail(int code, const std::string& msg) { return fse::set_error(code, msg); }
}  // namespace

extern "C" {

void fs_server_config_default(fs_server_config* c) {
  if (!c) return;
  c->device_budget = 0;
  c->precision = FS_PREC_MIXED;
  c->max_steps = 2000;
  c->tol = 1e-10;
  c->prep_threads = 8;
  c->use_graph = 1;
  c->max_recoveries = 0;
  c->checkpoint_every = 10;
  c->max_steps_per_sqrt_node = 0.0;
  c->limiter_freeze = 0;
  c->limiter_fallback = 0;
  c->first_order_fallback = 0;
  c->stall_steps_per_sqrt_node = 0.0;
}

int fs_server_create(const fs_server_config* c, fs_server** out) {
  if (!c || !out) return fail(FS_ERR_ARG, "null argument");
  if (c->precision != FS_PREC_F64 && c->precision != FS_PREC_MIXED && c->precision != FS_PREC_MIXED16)
    return fail(FS_ERR_ARG, "unknown precision");
  if (c->max_steps < 1 || !(c->tol > 0) || c->prep_threads < 1 || c->device_budget < 0 || c->max_recoveries < 0 ||
      c->checkpoint_every < 1 || !(c->max_steps_per_sqrt_node >= 0) || c->limiter_freeze < 0 ||
      c->limiter_fallback < 0 || !(c->stall_steps_per_sqrt_node >= 0))
    return fail(FS_ERR_ARG, "bad server configuration");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(FS_ERR_HIP, "no HIP device available");
  auto s = std::make_unique<fs_server>();
  s->cfg = *c;
  FSE_HIP(hipGetDevice(&s->device));
  if (c->device_budget > 0) {
    s->budget = c->device_budget;
  } else {
    size_t fr = 0, tot = 0;
    FSE_HIP(hipMemGetInfo(&fr, &tot));
    s->budget = (int64_t)(0.8 * (double)fr);
  }
  s->t0 = std::chrono::steady_clock::now();
  fs_server* sp = s.get();
  for (int i = 0; i < c->prep_threads; ++i) s->workers.emplace_back([sp] { sp->prep_loop(); });
  s->gpu = std::thread([sp] { sp->gpu_loop(); });
  *out = s.release();
  return FS_OK;
}

void fs_server_destroy(fs_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_prep.notify_all();
  s->cv_gpu.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  if (s->gpu.joinable()) s->gpu.join();
  delete s;
}

int fs_server_submit(fs_server* s, fs_mesh* mesh, const fs_params* p, int64_t* id) {
  if (!s || !mesh || !p || !id) return fail(FS_ERR_ARG, "null argument");
  auto r = std::make_unique<Request>();
  r->mesh.reset(mesh);
  r->prm = *p;
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  r->t_submit = s->now();
  *id = r->id;
  s->to_prepare.push_back(r.get());
  s->all[r->id] = std::move(r);
  s->cv_prep.notify_one();
  return FS_OK;
}

int fs_server_wait(fs_server* s, double timeout_s, fs_request_info* out, int* got) {
  if (!s || !out || !got) return fail(FS_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  *got = 0;
  auto pred = [&] { return !s->done.empty() || s->stop; };
  if (timeout_s < 0)
    s->cv_done.wait(lk, pred);
  else
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  if (s->done.empty()) return FS_OK;
  Request* r = s->done.front();
  s->done.pop_front();
  out->id = r->id;
  out->status = r->status;
  out->steps = r->steps;
  out->drop = r->drop;
  out->t_submit = r->t_submit;
  out->t_admit = r->t_admit;
  out->t_done = r->t_done;
  out->device_bytes = r->bytes;
  out->nn = r->nn;
  out->ng = r->ng;
  out->forces[0] = r->forces[0];
  out->forces[1] = r->forces[1];
  out->recoveries = r->rt.recoveries;
  out->fallbacks = r->rt.fallbacks;
  out->max_steps = r->limit;
  out->r0 = r->r0;
  *got = 1;
  if (!r->error.empty()) fse::set_error(FS_ERR_STATE, r->error);
  return FS_OK;
}

int fs_server_get_state(fs_server* s, int64_t id, double* Q, int64_t n) {
  if (!s || !Q) return fail(FS_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return fail(FS_ERR_ARG, "unknown request id");
  Request* r = it->second.get();
  if (!r->finished) return fail(FS_ERR_ARG, "request not finished");
  if (r->Q.empty()) return fail(FS_ERR_STATE, "request has no state: " + r->error);
  if ((int64_t)r->Q.size() != n) return fail(FS_ERR_ARG, "size mismatch: expected " + std::to_string(r->Q.size()));
  std::copy(r->Q.begin(), r->Q.end(), Q);
  return FS_OK;
}

int fs_server_release(fs_server* s, int64_t id) {
  if (!s) return fail(FS_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return fail(FS_ERR_ARG, "unknown request id");
  if (!it->second->finished) return fail(FS_ERR_ARG, "request not finished (cancel it first)");
  for (auto d = s->done.begin(); d != s->done.end(); ++d)
    if (*d == it->second.get()) {
      s->done.erase(d);
      break;
    }
  s->all.erase(it);
  return FS_OK;
}

int fs_server_cancel(fs_server* s, int64_t id) {
  if (!s) return fail(FS_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return fail(FS_ERR_ARG, "unknown request id");
  it->second->cancel = true;
  s->cv_gpu.notify_all();
  return FS_OK;
}

int fs_server_counts(fs_server* s, int* queued, int* running, int* finished) {
  if (!s) return fail(FS_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = (int)(s->to_prepare.size() + s->ready.size());
  if (running) *running = s->running;
  if (finished) *finished = (int)s->done.size();
  return FS_OK;
}

double fs_server_clock(fs_server* s) { return s ? s->now() : 0.0; }

}  // extern "C"
