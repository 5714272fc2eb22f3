// The device engine behind both the single-mesh solver and the batch: a set
// of prepared meshes packed into one set of device arrays, their per-instance
// controls, and the pseudo-time step (solve.C:44-931) over all of them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../host/fs_setup.hpp"
#include "../kernels/fs_kernels.hpp"
#include "fs_capi.h"

namespace fse {

// records `msg` as the calling thread's last error and returns `code`
int set_error(int code, const std::string& msg);

#define FSE_HIP(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return ::fse::set_error(FS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

fs::Params to_params(const fs_params* p);
void from_params(const fs::Params& q, fs_params* p);

// Which reconstruction a pseudo-time step uses (solve.C:389-473 with the
// order/foits/soits schedule of solve.C:57-114).
struct Phase {
  int recon_flux;  // 0 nodal face states, 1 linear limited
  bool gradient;
  bool k2;
  int lim;         // fsk::LimMode bits, 0 none (phi = 1)
};
Phase phase_for(const fs_params& p, int ts);
double cfl_for(const fs_params& p, int citer);   // init_memory.C:200-216
bool supported(const fs_params& p, std::string* why);

struct Instance {
  const fs::Mesh* mesh = nullptr;
  fs_params prm{};
  double uinf = 0, vinf = 0, pinf = 0;
  int ts = 0;          // pseudo-time steps taken (grid->citer)
  bool active = true;
  int noff = 0, nn = 0, eoff = 0, ne = 0, soff = 0, nsub = 0, boff = 0, nbe = 0;
  int nnz = 0;
  int ncolor = 0;
  std::vector<int> slot_dev;   // reference block slot -> packed device slot
  int err = 0;                 // device error flags of the last step
  double r2 = 0, dq2 = 0;      // sum R^2 and sum dQ^2 of the last step
  // recovery (never touched unless the caller enables it): the CFL target is
  // scaled by cfl_scale and its ramp restarts after step ramp_from; the
  // latest checkpoint holds the state after step saved_ts (-1: none)
  double cfl_scale = 1.0;
  int ramp_from = 0;
  int saved_ts = -1;
  int phase_ts = -1;   // the phase checkpoint holds the state after this step
  // limiter freeze (opt-in, not in the reference): once the limited phase
  // has run more than this many steps the limiter values are kept (0: the
  // limiter runs every step, as the reference's)
  int freeze = 0;
};

// CFL of instance I's next step: the reference's ramp (init_memory.C:200-216)
// when no recovery has touched I, else the ramp restarted after step
// ramp_from towards cfl_scale * CFL (at least 1)
double cfl_for_inst(const Instance& I);

// Recovery of steady solves from non-physical states.  It is not part of
// the reference (which exits, cvbc.C:330-382, init_memory.C:257-268) and
// changes the iterate path of a solve it acts on (never its discretization
// or stopping test): off unless max_recoveries > 0.
struct RecoveryPolicy {
  int max_recoveries = 0;   // restores of one solve before it counts as failed
  int every = 10;           // steps between checkpoints of a healthy solve
  // > 0: the first failure of a solve running without a limiter in a phase
  // that would limit returns it to the start of that phase (the phase
  // checkpoint) with the Barth limiter switched on, frozen after this many
  // limited steps
  int limiter_fallback = 0;
  // the next failure of a solve in a reconstructing phase (linear or k=2)
  // returns it to the start of that phase with first-order face states for
  // the rest of the solve
  bool first_order_fallback = false;
  // > 0: a solve whose residual norm has not fallen below half its lowest
  // mark of the phase for ceil(this * sqrt(nodes)) steps (at least
  // kStallMinSteps) has stalled; a stall hands it to the next fallback that
  // applies, at the CFL it runs with (no fallback left: it runs on)
  double stall_steps_per_sqrt_node = 0.0;
};
constexpr int kStallMinSteps = 100;
// fallback bits of RecoveryTrack::fallbacks (fs_capi.h FS_FALLBACK_*)
constexpr int kFallbackLimiter = 1, kFallbackFirstOrder = 2, kFallbackStall = 4;
// per-solve bookkeeping of the policy
struct RecoveryTrack {
  int recoveries = 0;
  int fallbacks = 0;            // kFallback* bits of the fallbacks that acted
  int since_save = 0;
  int phase = -1;               // phase of the last step (a change resets r_best)
  double r_best = 0.0;          // smallest residual norm in the current phase
  double r_mark = 0.0;          // stall test: the last mark of the phase
  int since_mark = 0;           // steps since r_mark was set
};

// Device bytes that packing `m` into an engine of precision `prec` costs.
size_t footprint_bytes(const fs::Mesh& m, int prec, bool checkpoints = false);

class Engine {
 public:
  Engine() = default;
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;
  ~Engine();

  // Packs the meshes (prepared, kept alive by the caller) with one parameter
  // set each.  All instances share subits, iter and jufreq.
  int create(const std::vector<const fs::Mesh*>& meshes, const std::vector<fs_params>& prm, int precision);

  int ninst() const { return (int)inst.size(); }

  // Freestream initial state (ic.C:808-840, ghosts ic.C:77-86) for instance i
  // (or all when i < 0).
  int init_freestream(int i);

  // Writes the controls of Newton iteration `it` of every instance (from its
  // step counter and precision) to the device table on the stream.
  int upload_params(int it);
  // Controls with a forced limiter/gradient mode for the single-function
  // mirrors (instance 0 of a one-instance engine).
  int upload_params_custom(const fsk::InstParams& p);
  fsk::InstParams params_for(int i, int it) const;

  // One pseudo-time step of every active instance.  With `advance` each
  // active instance's step counter increments first.  With `graph` the step
  // runs as one hipGraph launch (captured per phase combination); with
  // `record` the kernel groups are timed with events (no graph).
  int step(bool advance, bool graph, bool record);
  // The same step split in two so several engines can overlap on the device:
  // step_begin enqueues the step and the norm/error readback and returns;
  // step_end waits for them and updates Instance::r2/dq2/err.
  int step_begin(bool advance, bool graph);
  int step_end();
  bool step_pending() const { return pending_; }
  // the enqueued step has completed or failed (step_end will not block and
  // reports a failure)
  bool step_ready() const { return !pending_ || hipStreamQuery(stream) != hipErrorNotReady; }

  // Compute_Body_Forces of every instance from its resident state (lift,
  // drag: 2 per instance); citer[i] is the step the state comes from.  With
  // `want`, only the instances it marks are evaluated (the others' values
  // are not meaningful).
  int body_forces(const int* citer, double* forces, const std::vector<char>* want = nullptr);

  // Deferred k=2 fields.  In the k=2 phase neither the residual nor the
  // Jacobian reads the k=2 derivatives, point values or limiter values (the
  // Gauss-point interior flux is overwritten by the nodal one,
  // residual.C:208-290, and the boundary Gauss branch is disabled,
  // residual.C:446); the reference still evaluates them every step
  // (solve.C:396-472), and its body forces, Cp and field reads see the values
  // of the last step.  A step of an instance in that phase (one Newton
  // iteration, no limiter freeze) therefore skips them and keeps the state it
  // started from; ensure_k2_fields evaluates them from that state for the
  // marked (or all) instances whose last step skipped them, giving the
  // values the step would have written.  Every output that reads them calls
  // it first.  defer_k2 = false evaluates them every step instead.
  int ensure_k2_fields(const std::vector<char>* want = nullptr);
  bool defer_k2 = true;
  // instance i's phi was written from outside the engine's passes (re-packing)
  void phi_written(int i) { phi_one_[i] = 0; }

  // copies the per-instance device error flags into Instance::err (and clears
  // them on the device)
  int fetch_errors();
  int sync();

  // per-instance field access (host <-> device), field ids of fs_capi.h
  int field_count(int i, int field, int64_t* n) const;
  int get_field(int i, int field, double* host, int64_t n);
  int set_field(int i, int field, const double* host, int64_t n);

  size_t device_bytes() const { return bytes_; }

  // Checkpoints for recovery: allocates two saved states (4 doubles per real
  // node and ghost each: the latest checkpoint and the start of the current
  // phase); each instance's first checkpoint is taken after its first
  // healthy step, its phase checkpoint after the last step of each phase.
  int enable_checkpoints();
  bool checkpoints_enabled() const { return qsave_ != nullptr; }
  // After a finished step of instance i (failed: device error flag or a
  // non-finite residual; rn: its residual norm): with the policy on, a failed
  // step returns the instance to a checkpoint (step counter included) with
  // half the CFL target and a restarted ramp: its latest checkpoint, or, on
  // the limiter fallback, the start of its phase.  A healthy step is
  // checkpointed after the first healthy step and then every `every` steps,
  // unless its residual is above ten times the best of its phase, and the
  // last step of a phase also saves the phase checkpoint.  A healthy step
  // that completes a stall (RecoveryPolicy::stall_steps_per_sqrt_node)
  // starts the next fallback that applies, at the same CFL.  Marks the copies
  // in mode[i] (fsk::CheckpointMode bits); returns whether the instance was
  // restored.  A failed step that is not restored (policy off, or out of
  // recoveries) is left to the caller.
  bool recover_or_checkpoint(int i, bool failed, double rn, const RecoveryPolicy& pol, RecoveryTrack& t,
                             std::vector<int>& mode);
  // enqueues the copies that mode marks (one launch over the engine)
  int checkpoint_pass(const std::vector<int>& mode);
  // copies instance i's checkpoint and recovery controls (and the limiter
  // switch of the fallback) from engine `from`'s instance j (re-packing)
  int adopt_checkpoint(int i, Engine& from, int j);

  // Times every candidate launch shape of the tunable kernels on this
  // engine's own data and keeps the fastest (coordinate search; the engine
  // must have run a step, and the next step recomputes everything the trial
  // launches overwrite).  The choice is stored in a cache file keyed by
  // (arch, precision, k=2 present, log2 of the node count) and applied by
  // later engines with the same key at creation.
  int autotune(double* seconds);
  bool tuned_from_cache() const { return tuned_from_cache_; }
  int steps_run() const { return steps_run_; }
  // sets the launch shapes (fs_capi.h's tuning value order); rejects
  // values outside the candidates
  int set_tuning(const int* v);
  void tuning_values(int* v) const {
    const fsk::Tuning& t = dm.tune;
    v[0] = t.gs_block; v[1] = t.gs_lanes; v[2] = t.gs32_block; v[3] = t.gs32_pre;
    v[4] = t.node_block; v[5] = t.face_block; v[6] = t.cell_block; v[7] = t.deriv_lds; v[8] = t.gs32_scaled;
    v[9] = t.node_fused; v[10] = t.gs32_dpp;
  }

  // state
  int precision = FS_PREC_F64;
  int subits = 0, iter = 1, jufreq = 1;
  std::vector<Instance> inst;
  fsk::DevMesh dm{};
  fsk::DevState ds{};
  fsk::NormChunks nc{};
  std::vector<int> color_off, color_nback;
  hipStream_t stream = nullptr;
  int device = 0;

  // timing of the last recorded steps (ms, accumulated)
  double t_edge = 0, t_node = 0, t_gs = 0, t_lu = 0;

 private:
  struct Buf {
    void* p = nullptr;
  };
  std::vector<Buf> bufs_;
  size_t bytes_ = 0;
  fsk::InstParams* d_ip_ = nullptr;   // iter * ninst
  fsk::InstParams* h_ip_ = nullptr;   // pinned staging, iter * ninst
  double* h_norms_ = nullptr;         // pinned, 2 * ninst
  int* h_err_ = nullptr;              // pinned, ninst
  fsk::ForceParams* d_fp_ = nullptr;  // ninst
  fsk::ForceParams* h_fp_ = nullptr;  // pinned staging, ninst
  double* h_bforce_ = nullptr;        // pinned, 2 * nbe
  double* qsave_ = nullptr;           // checkpointed Q with ghosts, 4 * (nn + ng)
  double* qphase_ = nullptr;          // Q with ghosts at the start of the phase, 4 * (nn + ng)
  int* d_rmode_ = nullptr;            // ninst
  int* h_rmode_ = nullptr;            // pinned staging, ninst
  double* qpre_ = nullptr;            // state (with ghosts) the last step of a deferring instance started from
  std::vector<char> k2_stale_;        // per instance: its last step deferred the k=2 fields
  bool k2_trial_ = false;             // autotune: the k=2 passes run as they would without deferral
  bool k2_deferred(const Instance& I) const;
  void mark_k2_stale();
  // per instance: every phi is 1 (a limiter pass without a limiter would
  // write nothing new, so it is skipped and, with no other work, not launched)
  std::vector<char> phi_one_;
  void note_phi(int i, int lim);
  // the limiter pass's node list (fsk::DevMesh::lim_chunks, batches only):
  // the chunks holding nodes of the instances whose pass has work, when they
  // hold at most three quarters of the nodes; rebuilt when that set changes
  // (the graphs are dropped then)
  int* d_lim_chunks_ = nullptr;
  std::vector<char> lim_set_;
  std::vector<int> lim_h_;
  int refresh_limiter_list();
  hipEvent_t ev_[8] = {};
  std::map<int, hipGraphExec_t> graphs_;
  // Sweep work lists (fsk::ColorRanges::chunks): once at most three quarters
  // of the nodes belong to active instances, each sweep launch covers only
  // the 16-row chunks of its range that hold an active row.  Rebuilt when an
  // instance outside the lists becomes active or the active nodes fall below
  // three quarters of those at the last build; every rebuild drops the
  // captured graphs (their launch grids change).
  // The node, edge and position kernels get lists of the 64-element chunks
  // holding elements of running instances, the LDS derivative pass a list of
  // its tiles (fsk::DevMesh::*_chunks, tile_list); the workgroups are dealt
  // to the XCDs in runs meanwhile (fsk::DevMesh::xcd_run).
  struct Seg {
    int t0, t1, inst;   // color positions [t0, t1) of one instance
  };
  std::vector<std::vector<Seg>> color_segs_;   // per color, in position order
  std::vector<int> dt_off_h_;         // first node of each LDS derivative tile
  std::vector<int> chunk_off_h_;      // 2 ncolor + 1
  std::vector<int> chunk_h_;          // sweep chunks' first positions, then the other lists
  std::vector<char> chunk_inst_;      // instances the lists cover
  int* d_chunks_ = nullptr;
  bool chunks_on_ = false;
  long long chunk_nodes_ = 0;         // active nodes when the lists were built
  int refresh_work_lists();
  fsk::ColorRanges color_ranges() const;
  void drop_graphs();

  template <typename T>
  int alloc(T** out, size_t count);
  template <typename T>
  int upload(const T** out, const std::vector<T>& v);
  int enqueue_step(int key, bool record);
  int newton_iteration(int it, int key, bool record);
  int step_key() const;
  int check_errors(bool norms);
  int enqueue_readback();
  bool pending_ = false;
  bool tuned_from_cache_ = false;
  int steps_run_ = 0;
  std::string tune_key() const;
};

}  // namespace fse
