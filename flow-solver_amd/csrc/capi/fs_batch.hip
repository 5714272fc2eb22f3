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

struct fs_batch {
  fse::Engine eng;
  bool graph = false;
  struct Track {
    int status = FS_INST_RUNNING;
    int steps = 0;
    double r0 = -1.0, drop = 1.0;
    int err = 0;
  };
  std::vector<Track> track;
};

namespace {
int fail(int code, const std::string& msg) { return fse::set_error(code, msg); }

int ts_ref(const fs_params& p) { return (p.order == 3) ? p.foits + p.soits + 1 : p.foits + 1; }

// one step of the running instances, then the per-instance convergence
// bookkeeping of fs_solve_steady
int batch_step(fs_batch* b, double tol, bool stop_on_tol) {
  fse::Engine& E = b->eng;
  int rc = E.step(true, b->graph, false);
  if (rc) return rc;
  for (int i = 0; i < E.ninst(); ++i) {
    fse::Instance& I = E.inst[i];
    fs_batch::Track& T = b->track[i];
    if (!I.active) continue;
    T.steps = I.ts;
    if (I.err) {
      T.status = FS_INST_FAILED;
      T.err = I.err;
      I.active = false;
      continue;
    }
    double rn = std::sqrt(I.r2);
    int tr = ts_ref(I.prm);
    if (I.ts == tr || (I.ts > tr && T.r0 < 0)) T.r0 = rn;
    if (T.r0 > 0) {
      T.drop = rn / T.r0;
      if (stop_on_tol && T.drop <= tol) {
        T.status = FS_INST_CONVERGED;
        I.active = false;
      }
    }
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
  int rc = b->eng.create(ms, ps, precision);
  if (rc) return rc;
  b->track.assign(n, fs_batch::Track{});
  *out = b.release();
  return FS_OK;
}

void fs_batch_destroy(fs_batch* b) { delete b; }

int fs_batch_get_info(const fs_batch* b, fs_batch_info* info) {
  if (!b || !info) return fail(FS_ERR_ARG, "null argument");
  const fse::Engine& E = b->eng;
  info->ninst = E.ninst();
  info->nn = E.dm.nn; info->ne = E.dm.ne; info->nsub = E.dm.nsub; info->nbe = E.dm.nbe; info->nnz = E.dm.nnz;
  info->ncolor = E.dm.ncolor;
  info->device_bytes = (int64_t)E.device_bytes();
  return FS_OK;
}

int fs_batch_init_freestream(fs_batch* b) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  return b->eng.init_freestream(-1);
}

int fs_batch_use_graph(fs_batch* b, int on) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  b->graph = on != 0;
  return FS_OK;
}

int fs_batch_autotune(fs_batch* b, double* seconds) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  return b->eng.autotune(seconds);
}

int fs_batch_get_tuning(const fs_batch* b, int* values, int* from_cache) {
  if (!b || !values) return fail(FS_ERR_ARG, "null argument");
  b->eng.tuning_values(values);
  if (from_cache) *from_cache = b->eng.tuned_from_cache() ? 1 : 0;
  return FS_OK;
}

int fs_batch_set_active(fs_batch* b, int i, int active) {
  if (!b || i < 0 || i >= b->eng.ninst()) return fail(FS_ERR_ARG, "instance out of range");
  fs_batch::Track& T = b->track[i];
  if (T.status == FS_INST_FAILED) return fail(FS_ERR_STATE, "instance failed: " + std::string(fse::device_error_message(T.err)));
  b->eng.inst[i].active = active != 0;
  T.status = active ? FS_INST_RUNNING : FS_INST_PAUSED;
  return FS_OK;
}

int fs_batch_step(fs_batch* b, double* r2, double* dq) {
  if (!b) return fail(FS_ERR_ARG, "null batch");
  int rc = batch_step(b, 0.0, false);
  if (rc) return rc;
  for (int i = 0; i < b->eng.ninst(); ++i) {
    if (r2) r2[i] = b->eng.inst[i].r2;
    if (dq) dq[i] = std::sqrt(b->eng.inst[i].dq2);
  }
  return FS_OK;
}

int fs_batch_solve_steady(fs_batch* b, int max_steps, double tol, int* steps, double* drop, int* status) {
  if (!b || max_steps < 1) return fail(FS_ERR_ARG, "bad arguments");
  fse::Engine& E = b->eng;
  for (int k = 0; k < max_steps; ++k) {
    bool any = false;
    for (const auto& I : E.inst) any = any || I.active;
    if (!any) break;
    int rc = batch_step(b, tol, true);
    if (rc) return rc;
  }
  for (int i = 0; i < E.ninst(); ++i) {
    fs_batch::Track& T = b->track[i];
    if (T.status == FS_INST_RUNNING) T.status = FS_INST_MAX_STEPS;
    if (steps) steps[i] = T.steps;
    if (drop) drop[i] = T.drop;
    if (status) status[i] = T.status;
  }
  return FS_OK;
}

int fs_batch_inst_status(const fs_batch* b, int i, int* status, int* steps, double* drop) {
  if (!b || i < 0 || i >= b->eng.ninst()) return fail(FS_ERR_ARG, "instance out of range");
  const fs_batch::Track& T = b->track[i];
  if (status) *status = T.status;
  if (steps) *steps = T.steps;
  if (drop) *drop = T.drop;
  return FS_OK;
}

int fs_batch_get(fs_batch* b, int i, int field, double* host, int64_t n) {
  if (!b || !host) return fail(FS_ERR_ARG, "null argument");
  return b->eng.get_field(i, field, host, n);
}

int fs_batch_set(fs_batch* b, int i, int field, const double* host, int64_t n) {
  if (!b || !host) return fail(FS_ERR_ARG, "null argument");
  return b->eng.set_field(i, field, host, n);
}

int fs_batch_time_steps(fs_batch* b, int nsteps, fs_timing* t) {
  if (!b || !t || nsteps < 1) return fail(FS_ERR_ARG, "bad arguments");
  fse::Engine& E = b->eng;
  int rc = E.sync();
  if (rc) return rc;
  int64_t evals = 0;
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < nsteps; ++k) {
    for (const auto& I : E.inst)
      if (I.active) evals += (int64_t)I.nsub + I.nbe;
    rc = batch_step(b, 0.0, false);
    if (rc) return rc;
  }
  rc = E.sync();
  if (rc) return rc;
  auto t1 = std::chrono::steady_clock::now();
  *t = fs_timing{};
  t->step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / nsteps;
  t->flux_evals = evals / nsteps;
  return FS_OK;
}

int fs_batch_body_forces(fs_batch* b, double* forces) {
  if (!b || !forces) return fail(FS_ERR_ARG, "null argument");
  fse::Engine& E = b->eng;
  std::vector<int> citer(E.ninst());
  for (int i = 0; i < E.ninst(); ++i) citer[i] = E.inst[i].ts;
  return E.body_forces(citer.data(), forces);
}

int fs_mesh_footprint(const fs_mesh* m, int precision, int64_t* bytes) {
  if (!m || !bytes) return fail(FS_ERR_ARG, "null argument");
  if (!m->prepared) return fail(FS_ERR_ARG, "mesh not prepared");
  *bytes = (int64_t)fse::footprint_bytes(m->m, precision);
  return FS_OK;
}

}  // extern "C"
