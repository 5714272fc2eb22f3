// Device data layout and kernel launchers for one mesh instance.
//
// HBM layout (all arrays 0-based, structure-of-arrays where a kernel streams
// one field per face):
//   per node       nQ/Q (4 x f64, interleaved rho,rhou,rhov,E), grad (8),
//                  phi (1), hess (20), pv (4), R/RHS/dQ (4), cv_area, x (2)
//   per edge       (L,R), its 1-2 subedges, the two off-diagonal block slots,
//                  4 LSQ weights
//   per subedge    (nx,ny,len), face point (2), Gauss points (6), flux (4)
//   per bedge      real node, bc code, (nx,ny,len), flux (4), block (16)
//   per node CSR   incident subedges / edges / bedges in ascending order
//   block CSR      ia/ja/iau + transposed-slot map, LHS 16 x f64 per block
#pragma once
#include <hip/hip_runtime.h>

namespace fsk {

struct DevMesh {
  int nn, ng, ne, nsub, nbe, nnz, ncolor;
  const double* x;          // 2*nn
  const double* cv_area;    // nn
  const int* edges;         // 2*ne
  const int* edge_sub;      // 2*ne (second = -1 for a boundary edge)
  const int* edge_slot;     // 2*ne (slot of block (L,R), slot of block (R,L))
  const double* edge_w;     // 4*ne
  const double* sub_n;      // 3*nsub
  const double* sub_fc;     // 2*nsub
  const double* sub_xg;     // 6*nsub
  // Node-incidence order: node n's incident subedges (ascending subedge
  // index) occupy slots nsub_off[n]..nsub_off[n+1]; its incident edges
  // (ascending edge index) occupy nedge_off[n]..nedge_off[n+1].
  const int* nsub_off;      // nn+1
  const int* nedge_off;     // nn+1
  const int* sub_inc;       // 2*nsub: incidence slot of each subedge at L and at R
  const double* inc_n;      // 4 per subedge incidence: (nx, ny, len, 0)
  const double* inc_fc;     // 2 per subedge incidence: face point (k=1)
  const double* inc_xg;     // 6 per subedge incidence: Gauss points (k=2)
  const int* inc_other;     // per edge incidence: other node | (1<<31 if this node is R)
  const double* inc_gw;     // 2 per edge incidence: LSQ weights of this side
  const int* nbe_off;       // nn+1
  const int* nbe_list;
  const int* bedge_real;    // nbe
  const int* bedge_bc;      // nbe
  const double* bedge_n;    // 3*nbe
  // Block CSR in color-major device order: the rows of one color are one
  // contiguous run of blocks.  Row n owns device slots rstart[n]..iau[n]
  // (off-diagonals in the reference's nsn order, diagonal last).
  const int* rstart;        // nn
  const int* ja;            // nnz, device slot order
  const int* iau;           // nn
  const int* slot_t;        // nnz, device slot of the transposed block
  const int* color_rows;    // nn (rows grouped by color)
  const int* st_off;        // nn+1 (k=2)
  const int* st_nodes;
  const double* st_Wt;      // [entry][5] pseudo-inverse columns, weights folded
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
  double* LHS;    // nnz*16
  double* LU;     // nn*16, indexed by color position
  double* inc_flux;  // 2*nsub*4: subedge flux at each endpoint's incidence slot, signed
  double* bflux;     // nbe*4
  double* bjac;      // nbe*16
  double* partial;   // reduction scratch
  int* err;          // device error flag
  // mixed precision (FP32 Jacobian blocks and sweeps)
  float* LHS32;      // nnz*16, color-major slots
  float* LU32;       // nn*16, color position; reciprocal pivots on the diagonal
  float* RHS32;      // nn*4
  float* dQ32;       // nn*4
};

struct PhysParams {
  double gamma, uinf, vinf, pinf, cfl;
};

enum NodeFlags { NODE_EVCON = 1, NODE_INTERIOR = 2, NODE_BOUNDARY = 4, NODE_JAC = 8, NODE_RHS = 16 };

// recon: 0 nodal, 1 linear limited (k=1), 3 quadratic (k=2 phase: nodal flux)
void launch_gradient(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_derivatives(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_limiter(const DevMesh& m, const DevState& s, int recon, hipStream_t st);
void launch_edge_flux(const DevMesh& m, const DevState& s, const PhysParams& p, int recon, int jupdate,
                      hipStream_t st);
void launch_bedge(const DevMesh& m, const DevState& s, const PhysParams& p, int jupdate, hipStream_t st);
void launch_node(const DevMesh& m, const DevState& s, const PhysParams& p, int flags, hipStream_t st);
void launch_zero_offdiag(const DevMesh& m, const DevState& s, hipStream_t st);
// color visiting order of `subits` symmetric multicolor sweeps; returns its length
int gs_schedule(int ncolor, int subits, int* out);
void launch_diag_lu(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_gs_sweeps(const DevMesh& m, const DevState& s, const int* color_off_host, int subits, hipStream_t st);
inline void launch_gs(const DevMesh& m, const DevState& s, const int* color_off_host, int subits, hipStream_t st) {
  launch_diag_lu(m, s, st);
  launch_gs_sweeps(m, s, color_off_host, subits, st);
}
void launch_update(const DevMesh& m, const DevState& s, hipStream_t st);
// deterministic sums over nn*4 values: out[0]=sum R^2, out[1]=sum dQ^2 (device)
void launch_norms(const DevMesh& m, const DevState& s, double* d_out, hipStream_t st);
// mixed precision
void launch_edge_jac32(const DevMesh& m, const DevState& s, float gamma, int recon, int jupdate, hipStream_t st);
void launch_node_mixed(const DevMesh& m, const DevState& s, const PhysParams& p, int jac, hipStream_t st);
void launch_diag_lu32(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_gs32_sweeps(const DevMesh& m, const DevState& s, const int* color_off, int subits, hipStream_t st);
void launch_update_mixed(const DevMesh& m, const DevState& s, hipStream_t st);
void launch_roe_batch(int n, double gamma, const double* normals, const double* qL, const double* qR, double* flux,
                      hipStream_t st);

}  // namespace fsk
