// Device data layout and kernel launchers.
//
// One launch serves a batch of independent mesh instances packed CSR-style:
// every per-node / per-edge / per-subedge / per-bedge array is the
// concatenation of the instances' arrays with their indices shifted by the
// instance offsets, and every element looks up its instance's controls in a
// small device table (InstParams).  A single mesh is a batch of one.
//
// HBM layout (all arrays 0-based, structure-of-arrays where a kernel streams
// one field per face):
//   per node       nQ/Q (4 x f64, interleaved rho,rhou,rhov,E), grad (8),
//                  phi (1), hess (20), pv (4), R/RHS/dQ (4), cv_area, x (2)
//   per ghost      nQ/Q after all real nodes: ghost of packed bedge b at nn+b
//   per edge       (L,R), its 1-2 subedges, the two off-diagonal block slots
//   per subedge    (nx,ny,len) and, at each endpoint's incidence slot, the
//                  face point (2), Gauss points (6) and the signed flux (4)
//   per bedge      real node, bc code, (nx,ny,len), flux (4), block (16)
//   per node CSR   incident subedges / edges / bedges in ascending order
//   block CSR      color-major: all rows of color c (of every instance) form
//                  one contiguous run of rows and of blocks
#pragma once
#include <hip/hip_runtime.h>

namespace fsk {

// Per-instance controls of one Newton iteration, written by the host before
// the launches and read by every element of the instance.
struct InstParams {
  double gamma, uinf, vinf, pinf, cfl;
  int recon_flux;  // face states: 0 nodal, 1 linear limited
  int grad;        // LSQ gradient this iteration
  int k2;          // k=2 derivatives this iteration
  int lim;         // LimMode bits; 0: phi = 1
  int jac;         // Jacobian update this iteration
  int active;      // instance takes part in this step
  // k=2 phase with the derivatives and limiter deferred (k2 0 and lim
  // LIM_FROZEN then): no residual or Jacobian of that phase reads them, so
  // they are evaluated when an output does, from the state the step started
  // from, which the step's first copy keeps for the instance
  int k2_deferred;
  int pad;
};

// LIM_FROZEN: the limiter pass leaves phi as the last pass wrote it (the
// opt-in limiter freeze of batches and serving; never set on the parity path)
enum LimMode { LIM_ON = 1, LIM_GAUSS = 2, LIM_VENKAT = 4, LIM_FROZEN = 8 };

// Launch shapes chosen by the autotuner.  None of them changes what a
// thread computes or the order it sums in, so every choice gives the same
// results bit for bit.
constexpr int kMaxBlock = 512;
struct Tuning {
  int gs_block = 256;     // FP64 sweep workgroup size
  int gs_lanes = 4;       // FP64 sweep: 4 lanes per row (one per component) or 1 thread per row
  int gs32_block = 256;   // FP32 sweep workgroup size
  int gs32_pre = 6;       // the engine's sweep (FP32 or FP16 operands; FP64 with 4 lanes per row): row blocks
                          // whose loads are issued before the arithmetic (4, 6, 8; 0: the FP64 sweep loads
                          // block by block, the mixed sweeps take it as 4)
  int node_block = 256;   // node assembly
  int face_block = 256;   // subedge / edge kernels
  int cell_block = 256;   // gradient, limiter, derivatives, diagonal factorisation
  int deriv_lds = 1;      // k=2 derivatives: 0 neighbour states gathered through L2, 1 staged per tile in LDS
  int gs32_scaled = 1;    // FP32 sweeps: 0 apply the stored diagonal inverse per row, 1 operands pre-multiplied by it
  int node_fused = 1;     // pre-multiplied operands: 1 formed by the mixed node pass itself (the row's blocks
                          // staged in LDS), 0 by a separate diagonal pass over the stored blocks
  int gs32_dpp = 0;       // the engine's 4-lane sweep: 0 every lane loads a neighbour's whole dQ, 1 one
                          // component per lane shared within the quad by DPP broadcasts
};

// LDS tiles of the k=2 derivative pass: runs of kDerivTileNodes consecutive
// nodes; a tile's distinct stencil neighbours (at most kDerivTileMaxU) are
// staged in LDS once and its stencil entries address them by slot
constexpr int kDerivTileNodes = 32;

// block-column padding: the mixed sweeps read up to 8 column entries of a
// row without testing its degree
constexpr int kJaPad = 8;
constexpr int kDerivTileMaxU = 576;

struct DevMesh {
  int nn, ng, ne, nsub, nbe, nnz, ncolor, ninst;
  int maxdeg;               // largest number of edges at one node (off-diagonal blocks of a row)
  int slice_w;              // 0, or the uniform slice width: sbase[s] = 16 slice_w s (mixed sweeps compute it)
  int xcd_run;              // xcd_block's run length (0: one contiguous range per XCD)
  // Work lists of a batch most of whose instances have stopped (null: every
  // element): kWorkChunk-element chunks holding elements of running
  // instances, as first elements, and the LDS derivative tiles holding nodes
  // of running instances.
  const int* node_chunks;  int n_node_chunks;
  const int* edge_chunks;  int n_edge_chunks;
  const int* pos_chunks;   int n_pos_chunks;    // color positions
  const int* tile_list;    int n_tile_list;
  // the limiter pass's own list (null: node_chunks): chunks holding nodes of
  // the instances whose limiter pass has work this step
  const int* lim_chunks;   int n_lim_chunks;
  Tuning tune;
  const int* inst_node;     // nn
  const int* inst_edge;     // ne
  const int* inst_bedge;    // nbe
  const int* pos_inst;      // nn: instance of each color-order position
  const double* x;          // 2*nn
  const double* cv_area;    // nn
  const int* edges;         // 2*ne
  const int* edge_sub;      // 2*ne (second = -1 for a boundary edge)
  const int* edge_slot;     // 2*ne (slot of block (L,R), slot of block (R,L))
  const double* sub_n;      // 3*nsub
  const int* sub_edge;      // nsub: the subedge's edge
  // Node-incidence order: node n's incident subedges (ascending subedge
  // index) occupy slots nsub_off[n]..nsub_off[n+1]; its incident edges
  // (ascending edge index) occupy nedge_off[n]..nedge_off[n+1].
  const int* nsub_off;      // nn+1
  const int* nedge_off;     // nn+1
  const int* sub_inc;       // 2*nsub: incidence slot of each subedge at L and at R
  const double* inc_n;      // 4 per subedge incidence: (nx, ny, len, 0)
  const double* inc_fc;     // 2 per subedge incidence: face point (k=1)
  const double* inc_xg;     // 6 per subedge incidence: Gauss points (k=2)
  const int* inc_sother;    // per subedge incidence: other node | (1<<31 if this node is R)
  const int* inc_other;     // per edge incidence: other node | (1<<31 if this node is R)
  const double* inc_gw;     // 2 per edge incidence: LSQ weights of this side
  // Mixed path only (null otherwise), per edge incidence: the edge's two
  // subedges' (nx, ny, len) in the edge's L->R orientation (len = 0 where a
  // boundary edge has one), their face points, and the block slot of
  // (this node, other node).
  const double* einc_geo;   // 6 per edge incidence
  const double* einc_fc;    // 4 per edge incidence
  const int* einc_slot;     // 1 per edge incidence
  const int* nbe_off;       // nn+1
  const int* nbe_list;
  const int* bedge_real;    // nbe
  const int* bedge_bc;      // nbe
  const double* bedge_n;    // 3*nbe
  const double* bedge_xg;   // 6*nbe: Gauss points of the bedge's half segment (half arc on a curved wall)
  const double* bedge_gn;   // 6*nbe: unit normal at each Gauss point
  // Block storage.  Rows are visited in color order (position t = index in
  // color_rows).  Off-diagonal blocks are sliced: positions are cut into
  // slices of 16, and off-diagonal block e (the reference's nsn order) of the
  // row at position t lives in slot sbase[t/16] + 16 e + t%16, so a sweep
  // wave reading block e of 16 consecutive rows reads one contiguous run.
  // The diagonal block of node n is slot iau[n] (all diagonals after the
  // off-diagonal slots, in position order).
  const int* roff;          // nn+1: node n's off-diagonal entries roff[n]..roff[n+1]
  const int* tslot;         // per entry: slot of the transposed block (m,n) of entry (n,m)
  const int* iau;           // nn: slot of the diagonal block
  const int* ja;            // per off-diagonal slot: column node (padding slots: 0); on the mixed path the
                            // column's color position; 16 kJaPad padding entries follow the last slot
  const int* pdeg;          // nn, by position: off-diagonal blocks of the row
  const int* sbase;         // nn/16+1: first slot of each slice
  const int* color_rows;    // nn (rows grouped by color)
  const int* st_off;        // nn+1 (k=2; empty stencils for k=1 instances)
  const int* st_nodes;
  const double* st_Wt;      // [entry][5] pseudo-inverse columns, weights folded
  int ndtile;               // LDS tiles of the k=2 derivative pass (0: none)
  const int* dt_off;        // ndtile+1: first node of each tile
  const int* dt_uoff;       // ndtile+1: the tile's distinct neighbours in dt_unodes
  const int* dt_unodes;
  const unsigned short* st_loc;  // per stencil entry: slot of its node in the tile's list
  const double* moments;    // 9*nn
};

struct DevState {
  double* Q;      // (nn+ng)*4
  double* nQ;     // (nn+ng)*4
  double* dQ;     // nn*4
  double* R;      // nn*4
  double* RHS;    // nn*4
  double* grad;   // nn*8
  double* phi;    // nn
  double* hess;   // nn*20
  double* pv;     // nn*4
  double* ev_con; // nn
  double* dt;     // nn
  double* LHS;    // nnz*16, nnz = block slots (padding included)
  double* LU;     // nn*16, indexed by color position
  double* inc_flux;  // 2*nsub*4: subedge flux at each endpoint's incidence slot, signed
  double* bflux;     // nbe*4
  double* bjac;      // nbe*16
  double* bforce;    // nbe*2: per-bedge body-force terms
  int* err;          // per-instance device error flags
  // mixed precision (FP32 Jacobian blocks and sweeps)
  float* LHS32;      // nnz*16, color-major slots
  float* LU32;       // nn*16, color position: inverse of the diagonal block (pivoted, FP64, stored FP32)
  float* RHS32;      // nn*4 by color position: the FP32 right-hand side; with pre-multiplied sweep
                     // operands its product with the row's diagonal inverse (and the off-diagonal
                     // blocks of LHS32 likewise)
  float* dQ32;       // nn*4 by color position
  unsigned char* pos_act;  // nn by color position: the row's instance is active (written before the sweeps)
  float* D32;        // nn*16 by color position: the FP32 diagonal blocks (the tail of LHS32, slots
                     // nnz-nn.., when LHS32 exists)
  unsigned short* LHS16;  // FS_PREC_MIXED16 only: (nnz-nn)*16 FP16 bit patterns, the off-diagonal slots
                          // of LHS32's layout holding the pre-multiplied blocks (LHS32 is not allocated)
  const InstParams* ip;   // ninst, controls of the current Newton iteration
  double* inst_norms;     // 2 per instance: sum R^2, sum dQ^2
  double* chunk_part;     // 2 per reduction chunk
};

// Workgroups are dispatched to the 8 XCDs round-robin.  Renumbered this way
// (run 0), XCD x runs one contiguous range of the grid, so the neighbour
// gathers of adjacent rows and faces are served by the same L2 instead of
// being fetched into eight.  With run R > 0 the grid is dealt to the XCDs in
// runs of R consecutive workgroups instead (the blocks past the last full
// round of 8 R split contiguously): work that sits in one part of the grid,
// the instances still running in a batch, then spreads over all eight.  A
// bijection of [0, gridDim.x) for any grid size; which XCD runs a workgroup
// changes no result.
__device__ __forceinline__ int xcd_block(int run) {
  const int G = gridDim.x, b = blockIdx.x, x = b & 7;
  int base = 0, n = G, bb = b;
  if (run > 0) {
    const int S = 8 * run, full = G / S, k = b >> 3;
    if (k < full * run) return (k / run) * S + x * run + (k % run);
    base = full * S;
    n = G - base;
    bb = b - base;
  }
  const int q = n >> 3, r = n & 7;
  return base + x * q + (x < r ? x : r) + (bb >> 3);
}
// the run length used once a batch's work lists are on
constexpr int kXcdRun = 16;

// The element of virtual index vn (a thread index divided by its lanes per
// element) under a work list of 64-element chunks: element vn % 64 of the
// (vn / 64)-th listed chunk, or `none` past the list; without a list, vn.  A
// wave's virtual indices lie in one chunk for any lanes per element, so the
// chunk's first element is a scalar load.
constexpr int kWorkChunk = 64;
__device__ __forceinline__ int work_elem(const int* __restrict__ chunks, int nchunks, int vn, int none) {
  if (!chunks) return vn;
  const int c = __builtin_amdgcn_readfirstlane(vn / kWorkChunk);
  return c < nchunks ? chunks[c] + (vn % kWorkChunk) : none;
}
// elements a launch covers: every one, or the listed chunks
inline int work_count(const int* chunks, int nchunks, int n) { return chunks ? kWorkChunk * nchunks : n; }

// Instance of element `idx` (a single-instance launch skips the lookup).
__device__ __forceinline__ int inst_of(const DevMesh& m, const int* map, int idx) {
  return m.ninst == 1 ? 0 : map[idx];
}

// NODE_EVONLY with NODE_EVCON: ev_con alone (cv_calc_eigenvalue_contribution),
// neither dt nor the V/dt diagonal
enum NodeFlags { NODE_EVCON = 1, NODE_INTERIOR = 2, NODE_BOUNDARY = 4, NODE_JAC = 8, NODE_RHS = 16, NODE_EVONLY = 32 };

// Deterministic per-instance norms: each instance's nodes are cut into
// fixed-size chunks counted from the instance's first node, so the sums do
// not depend on what else shares the launch.
constexpr int kNormChunkNodes = 1024;
struct NormChunks {
  int nchunk;
  const int* chunk_inst;   // nchunk
  const int* chunk_begin;  // nchunk, packed node index
  const int* chunk_end;    // nchunk
  const int* inst_chunk0;  // ninst+1: first chunk of each instance
};

// One color launch of the symmetric sweeps.  Within a color, the rows of
// instances whose last color it is are stored last; a backward launch skips
// them, so every instance sees exactly the schedule of its own color count.
struct GsLaunch {
  int color;
  int backward;
};
// color visiting order of `subits` symmetric sweeps over `ncolor` colors;
// returns its length (out holds at least subits*ncolor entries)
int gs_schedule(int ncolor, int subits, GsLaunch* out);
struct ColorRanges {
  const int* off;     // ncolor+1, host
  const int* nback;   // ncolor, host: rows of a color visited by backward launches
  // Optional work lists (both null: every row of a launch's range): range
  // 2c + b (color c forward, b = 0, or backward, b = 1) launches only the
  // kGsChunk-row chunks chunks[chunk_off[2c + b] ..], each a first position,
  // so rows of stopped instances cost nothing once most have stopped.
  const int* chunk_off = nullptr;  // 2 ncolor + 1, host
  const int* chunks = nullptr;     // device
};
// rows per work-list chunk: one wave of the 4-lane sweeps
constexpr int kGsChunk = 16;

// Per-instance modes come from DevState::ip.
void launch_copy(double* dst, const double* src, size_t n, hipStream_t st);
// dst = src for the state with ghosts (4 words per node and ghost): the
// nodes of the work list's chunks (every node without one) and every ghost
void launch_copy_state(const DevMesh& m, double* dst, const double* src, hipStream_t st);
// the same copy, which also writes the real nodes and ghosts of instances
// with ip[i].k2_deferred (and active) to `save`
void launch_copy_state_save(const DevMesh& m, const InstParams* ip, double* dst, const double* src, double* save,
                            hipStream_t st);
void launch_gradient(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_derivatives(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_limiter(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_edge_flux(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_bedge(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_node(const DevMesh& m, const DevState& s, int flags, hipStream_t st);
void launch_zero_offdiag(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_diag_lu(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_gs_sweeps(const DevMesh& m, const DevState& s, const ColorRanges& cr, int subits, hipStream_t st);
void launch_update(const DevMesh& m, const DevState& s, hipStream_t st);
// Per-instance checkpoints of the state with its ghosts (nn real nodes, then
// the ghost of packed boundary edge b at nn + b), two slots: the latest
// (Qsave) and the start of the current phase (Qphase).  Bits of mode[i]:
// CKPT_SAVE / CKPT_SAVE_PHASE copy instance i's Q into the slot,
// CKPT_RESTORE / CKPT_RESTORE_PHASE copy the slot back into Q (a restore
// excludes saves in the same pass).
enum CheckpointMode { CKPT_SAVE = 1, CKPT_RESTORE = 2, CKPT_SAVE_PHASE = 4, CKPT_RESTORE_PHASE = 8 };
void launch_checkpoint(const DevMesh& m, double* Q, double* Qsave, double* Qphase, const int* mode, hipStream_t st);
void launch_norms(const DevMesh& m, const DevState& s, const NormChunks& c, hipStream_t st);
// mixed precision
// FP64 residual, ev_con/dt, FP32 analytic Jacobian blocks (off-diagonal and
// diagonal) and right-hand sides in one node pass
void launch_resid_jac32(const DevMesh& m, const DevState& s, hipStream_t st);
// whether launch_resid_jac32 also forms the sweep operands (diagonal
// inverse, pre-multiplied blocks and right-hand sides, zero dQ), so no
// diagonal pass follows it: FP16 blocks always; FP32 pre-multiplied operands
// when the tuning asks for it and a workgroup's rows fit in LDS
bool resid_jac32_fused(const DevMesh& m, const DevState& s);
// whether a workgroup of the fused node pass can stage its rows' blocks in
// LDS at the mesh's largest node degree (FS_PREC_MIXED16 requires it)
bool resid_jac32_fused_fits(const DevMesh& m);
void launch_diag_lu32(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_gs32_sweeps(const DevMesh& m, const DevState& s, const ColorRanges& cr, int subits, hipStream_t st);
void launch_update_mixed(const DevMesh& m, const DevState& s, hipStream_t st);
// form 0: reference operation order (bit-exact); 1: characteristic-wave form
// Per-instance controls of the body-force pass: the reconstruction of the
// phase of the step the state comes from (0 nodal, 1 limited linear, 3 k=2
// point value + moments), gamma, and cos/sin of the angle of attack.
struct ForceParams {
  int rorder;
  double gamma, ca, sa;
};
void launch_body_terms(const DevMesh& m, const DevState& s, const ForceParams* fp, hipStream_t st);
// streaming read / copy of n16 16-byte words (achievable bandwidth probes)
void launch_probe_read(const void* a, size_t n16, float* out, hipStream_t st);
void launch_probe_copy(const void* a, void* b, size_t n16, hipStream_t st);
void launch_roe_batch(int n, int form, double gamma, const double* normals, const double* qL, const double* qR,
                      double* flux, hipStream_t st);

}  // namespace fsk
