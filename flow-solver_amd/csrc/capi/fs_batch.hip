// C-ABI of batched steady solves: many meshes in one device engine, one
// launch per kernel per pseudo-time step, per-instance convergence masks.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "fs_capi.h"
#include "fs_capi_internal.hpp"
#include "fs_engine.hpp"

// A batch starts as one engine over all instances.  With re-packing on
// (fs_batch_set_compaction), during steady solves, once half of the newest
// engine's instances have stopped, the running ones are re-packed into a
// smaller engine (their state and step counters copied on the device), so
// the launches of later steps cover only live work.  The
// engine an instance stopped in keeps its final fields; `where` maps every
// instance to the engine holding its current state.  An instance's
// arithmetic does not depend on the other instances of its engine, so
// re-packing changes no result.
struct fs_batch {
  std::vector<std::unique_ptr<fse::Engine>> eng;
  std::vector<std::pair<int, int>> where;   // instance -> (engine, index in it)
  std::vector<const fs::Mesh*> meshes;
  std::vector<fs_params> params;
  int precision = FS_PREC_F64;
  bool graph = false;
  bool compact = false;
  fse::RecoveryPolicy rec;
  struct Track {
    int status = FS_INST_RUNNING;
    int steps = 0;            // steps taken, those a recovery rolled back included
    double r0 = -1.0, drop = 1.0;
    int err = 0;
    fse::RecoveryTrack rt;
  };
  std::vector<Track> track;

  fse::Engine& engine_of(int i) { return *eng[where[i].first]; }
  fse::Instance& inst(int i) { return eng[where[i].first]->inst[where[i].second]; }
  int ninst() const { return (int)where.size(); }
};

namespace {
int fail(int code, const std::string& msg) { return fse::set_error(code, msg); }

int ts_ref(const fs_params& p) { return (p.order == 3) ? p.foits + p.soits + 1 : p.foits + 1; }

// Re-packs the running instances of the newest engine into a new engine
// once at most half of its instances still run (and it holds at least 8).
// Called only before a step: every moved instance then steps in the new
// engine before any of its fields is read there.
int maybe_compact(fs_batch* b) {
  fse::Engine& E = *b->eng.back();
  const int ne = (int)b->eng.size() - 1;
  std::vector<int> live;
  for (int i = 0; i < b->ninst(); ++i)
    if (b->where[i].first == ne && b->inst(i).active) live.push_back(i);
  if (!b->compact || E.ninst() < 8 || live.empty() || 2 * (int)live.size() > E.ninst()) return FS_OK;
  std::vector<const fs::Mesh*> ms;
  std::vector<fs_params> ps;
  for (int i : live) {
    ms.push_back(b->meshes[i]);
    ps.push_back(b->params[i]);
  }
  auto N = std::make_unique<fse::Engine>();
  int rc = N->create(ms, ps, b->precision);
  if (!rc && E.checkpoints_enabled()) rc = N->enable_checkpoints();
  if (rc) return rc;
  N->dm.tune = E.dm.tune;
  N->defer_k2 = E.defer_k2;
  FSE_HIP(hipStreamSynchronize(E.stream));
  for (size_t k = 0; k < live.size(); ++k) {
    fse::Instance& from = b->inst(live[k]);
    fse::Instance& to = N->inst[k];
    // state with the ghosts' memory: real nodes, then the instance's ghosts
    FSE_HIP(hipMemcpyAsync(N->ds.Q + 4 * (size_t)to.noff, E.ds.Q + 4 * (size_t)from.noff,
                           4 * (size_t)from.nn * sizeof(double), hipMemcpyDeviceToDevice, N->stream));
    if (from.nbe > 0)
      FSE_HIP(hipMemcpyAsync(N->ds.Q + 4 * ((size_t)N->dm.nn + to.boff), E.ds.Q + 4 * ((size_t)E.dm.nn + from.boff),
                             4 * (size_t)from.nbe * sizeof(double), hipMemcpyDeviceToDevice, N->stream));
    // limiter values, which a frozen limiter keeps from step to step
    FSE_HIP(hipMemcpyAsync(N->ds.phi + to.noff, E.ds.phi + from.noff, (size_t)from.nn * sizeof(double),
                           hipMemcpyDeviceToDevice, N->stream));
    N->phi_written((int)k);
    to.ts = from.ts;
    to.freeze = from.freeze;
    to.active = true;
    to.r2 = from.r2;
    to.dq2 = from.dq2;
    if (E.checkpoints_enabled()) {
      rc = N->adopt_checkpoint((int)k, E, b->where[live[k]].second);
      if (rc) return rc;
    }
    from.active = false;
    b->where[live[k]] = {ne + 1, (int)k};
  }
  FSE_HIP(hipStreamSynchronize(N->stream));
  b->eng.push_back(std::move(N));
  return FS_OK;
}

// one step of the running instances, then the per-instance convergence
// bookkeeping of fs_solve_steady (and, with the recovery policy on, the
// restores and checkpoints it asks for)
int batch_step(fs_batch* b, double tol, bool stop_on_tol, bool record = false) {
  for (auto& E : b->eng) {
    bool any = false;
    for (const auto& I : E->inst) any = any || I.active;
    if (!any) continue;
    int rc = E->step(true, b->graph && !record, record);
    if (rc) return rc;
  }
  std::vector<std::vector<int>> mode(b->eng.size());
  for (size_t e = 0; e < b->eng.size(); ++e) mode[e].assign(b->eng[e]->ninst(), 0);
  for (int i = 0; i < b->ninst(); ++i) {
    fse::Instance& I = b->inst(i);
    fs_batch::Track& T = b->track[i];
    if (!I.active) continue;
    ++T.steps;
    const double rn = std::sqrt(I.r2);
    // a non-finite residual without a device flag fails too, as in serving
    const bool failed = I.err != 0 || !(rn == rn);
    const auto& w = b->where[i];
    if (b->eng[w.first]->recover_or_checkpoint(w.second, failed, rn, b->rec, T.rt, mode[w.first])) continue;
    if (failed) {
      T.status = FS_INST_FAILED;
      T.err = I.err;
      I.active = false;
      continue;
    }
    // the reference step of the solve as submitted (a fallback may have
    // changed the engine's copy of its controls)
    int tr = ts_ref(b->params[i]);
    if (I.ts == tr || (I.ts > tr && T.r0 < 0)) T.r0 = rn;
    if (T.r0 > 0) {
      T.drop = rn / T.r0;
      if (stop_on_tol && T.drop <= tol) {
        T.status = FS_INST_CONVERGED;
        I.active = false;
      }
    }
  }
  for (size_t e = 0; e < b->eng.size(); ++e) {
    if (!b->eng[e]->checkpoints_enabled()) continue;
    int rc = b->eng[e]->checkpoint_pass(mode[e]);
    if (rc) return rc;
  }
  return FS_OK;
}
}  // namespace

extern "C" {

int fs_batch_create(int n, const fs_mesh* const* meshes, const fs_params* params, int precision, fs_batch** out) {
  if (n < 1 || !meshes || !params || !out) return fail(FS_ERR_ARG, "bad arguments");
  std::vector<const fs::Mesh*> ms(n);
  std::vector<fs_params> ps(params, params + n);
  for (int i = 0; i < n; ++i) {
    if (!meshes[i] || !meshes[i]->prepared) return fail(FS_ERR_ARG, "mesh " + std::to_string(i) + " not prepared");
    if (params[i].order == 3 && meshes[i]->order_prepared != 3)
      return fail(FS_ERR_ARG, "mesh " + std::to_string(i) + ": order 3 needs a mesh prepared with order 3");
    ms[i] = &meshes[i]->m;
  }
  auto b = std::make_unique<fs_batch>();
  b->eng.push_back(std::make_unique<fse::Engine>());
  int rc = b->eng[0]->create(ms, ps, precision);
  if (rc) return rc;
  b->meshes = ms;
  b->params = ps;
  b->precision = precision;
  b->where.resize(n);
  for (int i = 0; i < n; ++i) b->where[i] = {0, i};
  b->track.assign(n, fs_batch::Track{});
  *out = b.release();
  return FS_OK;
}

void fs_batch_destroy(fs_batch* b) { delete b; }

int fs_batch_get_info(const fs_batch* b, fs_batch_info* info) {
  if (!b || !info) return fail(FS_ERR_ARG, "null argument");
  const fse::Engine& E = *b->eng[0];
  info->ninst = b->ninst();
  info->nn = E.dm.nn; info->ne = E.dm.ne; info->nsub = E.dm.nsub; info->nbe = E.dm.nbe; info->nnz = E.dm.nnz;
  info->ncolor = E.dm.ncolor;
  int64_t bytes = 0;
  for (const auto& e : b->eng) bytes += (int64_t)e->device_bytes();
  info->device_bytes = bytes;
  return FS_OK;
}

int fs_batch_init_freestream(fs_batch* b) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  for (int i = 0; i < b->ninst(); ++i) {
    int rc = b->engine_of(i).init_freestream(b->where[i].second);
    if (rc) return rc;
  }
  return FS_OK;
}

int fs_batch_use_graph(fs_batch* b, int on) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  b->graph = on != 0;
  return FS_OK;
}

int fs_batch_set_compaction(fs_batch* b, int on) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  b->compact = on != 0;
  return FS_OK;
}

int fs_batch_set_recovery(fs_batch* b, int max_recoveries, int checkpoint_every, int limiter_fallback,
                          int first_order_fallback) {
  if (!b || max_recoveries < 0 || checkpoint_every < 1 || limiter_fallback < 0) return fail(FS_ERR_ARG, "bad arguments");
  b->rec.max_recoveries = max_recoveries;
  b->rec.every = checkpoint_every;
  b->rec.limiter_fallback = limiter_fallback;
  b->rec.first_order_fallback = first_order_fallback != 0;
  if (max_recoveries > 0)
    for (auto& e : b->eng) {
      int rc = e->enable_checkpoints();
      if (rc) return rc;
    }
  return FS_OK;
}

int fs_batch_set_stall(fs_batch* b, double steps_per_sqrt_node) {
  if (!b || !(steps_per_sqrt_node >= 0)) return fail(FS_ERR_ARG, "bad arguments");
  b->rec.stall_steps_per_sqrt_node = steps_per_sqrt_node;
  return FS_OK;
}

int fs_batch_defer_k2(fs_batch* b, int on) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  for (auto& e : b->eng) e->defer_k2 = on != 0;
  return FS_OK;
}

int fs_batch_set_limiter_freeze(fs_batch* b, int steps) {
  if (!b || steps < 0) return fail(FS_ERR_ARG, "bad arguments");
  // a freeze changes which limiter values a deferred evaluation would form
  for (auto& e : b->eng)
    if (int rc = e->ensure_k2_fields()) return rc;
  for (int i = 0; i < b->ninst(); ++i) b->inst(i).freeze = steps;
  return FS_OK;
}

int fs_batch_inst_recovery(const fs_batch* b, int i, int* restores, int* fallbacks) {
  if (!b || i < 0 || i >= b->ninst()) return fail(FS_ERR_ARG, "instance out of range");
  if (restores) *restores = b->track[i].rt.recoveries;
  if (fallbacks) *fallbacks = b->track[i].rt.fallbacks;
  return FS_OK;
}

int fs_batch_autotune(fs_batch* b, double* seconds) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  return b->eng.back()->autotune(seconds);
}

int fs_batch_get_tuning(const fs_batch* b, int* values, int* from_cache) {
  if (!b || !values) return fail(FS_ERR_ARG, "null argument");
  b->eng.back()->tuning_values(values);
  if (from_cache) *from_cache = b->eng.back()->tuned_from_cache() ? 1 : 0;
  return FS_OK;
}

int fs_batch_set_tuning(fs_batch* b, const int* values) {
  if (!b || !values) return fail(FS_ERR_ARG, "null argument");
  for (auto& e : b->eng) {
    int rc = e->set_tuning(values);
    if (rc) return rc;
  }
  return FS_OK;
}

int fs_batch_set_active(fs_batch* b, int i, int active) {
  if (!b || i < 0 || i >= b->ninst()) return fail(FS_ERR_ARG, "instance out of range");
  fs_batch::Track& T = b->track[i];
  if (T.status == FS_INST_FAILED) return fail(FS_ERR_STATE, "instance failed: " + std::string(fse::device_error_message(T.err)));
  b->inst(i).active = active != 0;
  T.status = active ? FS_INST_RUNNING : FS_INST_PAUSED;
  return FS_OK;
}

int fs_batch_step(fs_batch* b, double* r2, double* dq) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  int rc = batch_step(b, 0.0, false);
  if (rc) return rc;
  for (int i = 0; i < b->ninst(); ++i) {
    if (r2) r2[i] = b->inst(i).r2;
    if (dq) dq[i] = std::sqrt(b->inst(i).dq2);
  }
  return FS_OK;
}

int fs_batch_solve_steady(fs_batch* b, int max_steps, double tol, int* steps, double* drop, int* status) {
  if (!b || max_steps < 1) return fail(FS_ERR_ARG, "bad arguments");
  for (int k = 0; k < max_steps; ++k) {
    bool any = false;
    for (int i = 0; i < b->ninst(); ++i) any = any || b->inst(i).active;
    if (!any) break;
    int rc = maybe_compact(b);
    if (!rc) rc = batch_step(b, tol, true);
    if (rc) return rc;
  }
  for (int i = 0; i < b->ninst(); ++i) {
    fs_batch::Track& T = b->track[i];
    if (T.status == FS_INST_RUNNING) T.status = FS_INST_MAX_STEPS;
    if (steps) steps[i] = T.steps;
    if (drop) drop[i] = T.drop;
    if (status) status[i] = T.status;
  }
  return FS_OK;
}

int fs_batch_inst_status(const fs_batch* b, int i, int* status, int* steps, double* drop) {
  if (!b || i < 0 || i >= b->ninst()) return fail(FS_ERR_ARG, "instance out of range");
  const fs_batch::Track& T = b->track[i];
  if (status) *status = T.status;
  if (steps) *steps = T.steps;
  if (drop) *drop = T.drop;
  return FS_OK;
}

int fs_batch_get(fs_batch* b, int i, int field, double* host, int64_t n) {
  if (!b || !host) return fail(FS_ERR_ARG, "null argument");
  if (i < 0 || i >= b->ninst()) return fail(FS_ERR_ARG, "instance out of range");
  return b->engine_of(i).get_field(b->where[i].second, field, host, n);
}

int fs_batch_set(fs_batch* b, int i, int field, const double* host, int64_t n) {
  if (!b || !host) return fail(FS_ERR_ARG, "null argument");
  if (i < 0 || i >= b->ninst()) return fail(FS_ERR_ARG, "instance out of range");
  return b->engine_of(i).set_field(b->where[i].second, field, host, n);
}

int fs_batch_time_steps(fs_batch* b, int nsteps, fs_timing* t) {
  if (!b || !t || nsteps < 1) return fail(FS_ERR_ARG, "bad arguments");
  for (auto& e : b->eng) {
    int rc = e->sync();
    if (rc) return rc;
    e->t_edge = e->t_node = e->t_gs = e->t_lu = 0;
  }
  // without graphs the kernel groups are timed with events on each engine's
  // stream (fs_timing's kernel fields); with graphs only the step
  const bool record = !b->graph;
  int64_t evals = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < nsteps; ++k) {
    for (int i = 0; i < b->ninst(); ++i) {
      const fse::Instance& I = b->inst(i);
      if (I.active) evals += (int64_t)I.nsub + I.nbe;
    }
    int rc = batch_step(b, 0.0, false, record);
    if (rc) return rc;
  }
  for (auto& e : b->eng) {
    int rc = e->sync();
    if (rc) return rc;
  }
  auto t1 = std::chrono::steady_clock::now();
  *t = fs_timing{};
  t->step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / nsteps;
  t->flux_evals = evals / nsteps;
  if (record) {
    for (auto& e : b->eng) {
      t->edge_ms += e->t_edge / nsteps;
      t->node_ms += e->t_node / nsteps;
      t->gs_ms += e->t_gs / nsteps;
      t->lu_ms += e->t_lu / nsteps;
    }
    t->other_ms = t->step_ms - t->edge_ms - t->node_ms - t->gs_ms - t->lu_ms;
  }
  return FS_OK;
}

int fs_batch_body_forces(fs_batch* b, double* forces) {
  if (!b || !forces) return fail(FS_ERR_ARG, "null argument");
  for (int e = 0; e < (int)b->eng.size(); ++e) {
    fse::Engine& E = *b->eng[e];
    std::vector<int> citer(E.ninst());
    std::vector<double> f(2 * (size_t)E.ninst());
    for (int k = 0; k < E.ninst(); ++k) citer[k] = E.inst[k].ts;
    int rc = E.body_forces(citer.data(), f.data());
    if (rc) return rc;
    for (int i = 0; i < b->ninst(); ++i)
      if (b->where[i].first == e) {
        forces[2 * (size_t)i] = f[2 * (size_t)b->where[i].second];
        forces[2 * (size_t)i + 1] = f[2 * (size_t)b->where[i].second + 1];
      }
  }
  return FS_OK;
}

int fs_mesh_footprint(const fs_mesh* m, int precision, int64_t* bytes) {
  if (!m || !bytes) return fail(FS_ERR_ARG, "null argument");
  if (!m->prepared) return fail(FS_ERR_ARG, "mesh not prepared");
  *bytes = (int64_t)fse::footprint_bytes(m->m, precision);
  return FS_OK;
}

}  // extern "C"
