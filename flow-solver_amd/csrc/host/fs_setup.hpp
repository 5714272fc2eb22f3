// Host-side setup for the median-dual Euler residual/implicit-update path.
//
// Everything here is one-time FP64/integer preprocessing that the device
// kernels consume: the reference grid format, seeded mesh generators,
// Cuthill-McKee renumbering, node/edge/subedge maps, median-dual geometry,
// control-volume moments, least-squares edge weights, block-CSR pattern,
// multicolor row coloring and the k=2 pseudo-inverse stencils.
//
// Internal indexing is 0-based.  The reference is 1-based; node n here is
// node n+1 there, edge e here is edge e+1 there, and so on.  Every ordering
// (node numbering, edge numbering, nsn/esn order, subedge order, stencil
// order) reproduces the reference's exactly so that summation orders and
// integer maps can be compared bit for bit.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace fs {

// Run controls, same fields and meaning as the reference's PARAMS
// (include/params.h:19-76), read in the order of params.C:44-217.
struct Params {
  int nsteps = 0, iter = 1, jufreq = 1;
  double CFL = 1.0;
  int ramp = 0;
  double gamma = 1.4, alpha = 0.0, mach = 0.0;
  int subits = 20, unsteady = 0, tacc = 0;
  double mintime = 0.0;
  int method = 1, order = 1, foits = 0, soits = 0, toits = 0;
  int isrestart = 0, limit = 1, reorder = 1, ic = 0, p_or_cv = 0;
  int start_post_pro = 0, post_pro_freq = 0;
};

bool read_params(const std::string& path, Params* p, std::string* err);
bool write_params(const std::string& path, const Params& p);

struct Mesh {
  // ---- raw grid (after optional reordering) ----
  int nn = 0;      // real nodes
  int ng = 0;      // ghost nodes (= nbedges)
  int ntri = 0, nquad = 0;
  int nb = 0;      // boundaries
  std::vector<double> x;      // 2*(nn+ng)
  std::vector<int> tri;       // 3*ntri
  std::vector<int> quad;      // 4*nquad
  std::vector<int> bseg_off;  // nb+1, offsets into bseg (segments)
  std::vector<int> bseg;      // 2*nseg: (n0,n1) boundary-major
  std::vector<int> bbc;       // nb boundary condition codes
  std::vector<int> nn_map;    // file order -> new id (empty if no reorder)
  // restart payload of the grid file (mesh_io.C:217-263), if it has one
  bool has_restart = false;
  int restart_citer = 0;
  double restart_cfl = 0.0;
  std::vector<double> restart_Q;        // 4*nn
  std::vector<double> restart_ghost_Q;  // 4*ng when the file carries the ghost extension

  // ---- integer maps (maps.C) ----
  std::vector<int> ncsn, csn;   // cells surrounding node (global elem ids)
  std::vector<int> nnsn, nsn;   // nodes surrounding node
  std::vector<int> esn;         // edge per nsn slot
  int ne = 0;
  std::vector<int> edges;       // 2*ne, L<R
  std::vector<int> c2e;         // 3*ntri + 4*nquad

  // ---- median-dual geometry (geometry.C, cv_calc.C) ----
  std::vector<double> el_cent;  // 2*nelem
  int nsub = 0;
  std::vector<int> sub_edge, sub_elem;          // nsub
  std::vector<double> sub_n;                    // 3*nsub (nx,ny,len)
  std::vector<double> sub_xm;                   // 2*nsub edge midpoint
  std::vector<double> sub_xg;                   // 6*nsub Gauss points
  int nbe = 0;
  std::vector<int> bedge;                       // 5*nbe: real,ghost,gelem,b,s
  std::vector<double> bedge_n;                  // 3*nbe
  std::vector<double> bedge_xg;                 // 6*nbe Gauss points of the half segment
  std::vector<double> bedge_gn;                 // 6*nbe unit normal at each Gauss point
                                                //  (the bedge normal on a straight wall)
  std::vector<double> cv_area;                  // nn
  std::vector<double> moments;                  // 9*nn
  std::vector<double> edge_weight;              // 4*ne

  // ---- block CSR (init_memory.C Build_CRS) ----
  std::vector<int> ia, iau, ja;                 // ia: nn+1, iau: nn, ja: nnz
  // slot of block (L,R) and (R,L) for every edge
  std::vector<int> edge_slot_lr, edge_slot_rl;
  // transposed slot: for every off-diagonal slot k in row n (column m),
  // the slot holding block (m,n)
  std::vector<int> slot_t;

  // ---- node-centric incidence (for deterministic gathers) ----
  // incident subedges per node in ascending subedge order; sign +1 if the
  // node is the edge's L node, -1 if R.
  std::vector<int> nsub_off, nsub_list;         // nn+1, 2*nsub (sub | side<<31)
  std::vector<int> nedge_off, nedge_list;       // nn+1, 2*ne  (edge | side<<31)
  std::vector<int> nbe_off, nbe_list;           // nn+1, nbe ascending
  // per-edge subedge list (1 or 2 subedges, ascending)
  std::vector<int> esub_off, esub_list;         // ne+1, nsub

  // ---- multicolor Gauss-Seidel coloring ----
  int ncolor = 0;
  std::vector<int> color;                       // nn
  std::vector<int> color_off, color_rows;       // ncolor+1, nn

  // ---- k=2 stencils (maps.C Add_Layer_to_GNSN) + pseudo-inverse ----
  std::vector<int> st_off, st_nodes;            // nn+1, stencil w/o self
  std::vector<double> st_W;                     // 5*m per node, weights folded
};

// ---------------- generators / IO ----------------
// Bump channel [0,3]x[0,1] with a circular-arc bump of height 0.1 on the
// lower wall over x in [1,2]; nx*ny structured nodes split into CCW
// triangles; interior nodes jittered by U(-jitter,jitter) of local spacing.
void make_bump_channel(int nx, int ny, double jitter, uint64_t seed, Mesh* m);
// Symmetric 12%-thick analytic-profile O-mesh: nwall wall nodes,
// nrad radial layers, geometric stretching from first cell height h0 to
// radius rmax chords, seeded jitter of interior nodes.
void make_airfoil_omesh(int nwall, int nrad, double h0, double rmax,
                        double jitter, uint64_t seed, Mesh* m);
// The same O-mesh (nwall even) with the wall cut at the leading and trailing
// edges into an upper and a lower boundary with their own codes (1 for
// polyline walls, 17 and 19 for the analytic NACA 0012 surfaces); boundaries
// upper, lower, farfield.
void make_airfoil_omesh_split(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed,
                              int upper_bc, int lower_bc, Mesh* m);
// Annulus r_in <= r <= r_out around the origin: ntheta (even) nodes per
// ring on nrad geometrically spaced rings, nodes at angles 0 and pi exactly on
// the x axis, interior nodes jittered; boundaries inner (bc_in), outer
// (bc_out).  With the circle codes 11/12 (radius 1) and 21/22 (radius 300)
// the rings are walls or farfields on the analytic circles.
void make_annulus(int ntheta, int nrad, double r_in, double r_out, double jitter, uint64_t seed, int bc_in,
                  int bc_out, Mesh* m);

bool write_grid(const std::string& path, const Mesh& m);
bool read_grid(const std::string& path, Mesh* m, std::string* err);
bool write_bc(const std::string& path, const Mesh& m);
// The grid in its current numbering plus the restart payload of
// write_mesh (mesh_io.C:318-395): last step, last CFL and Q of the nn real
// nodes; with `ghosts`, the ng ghost states after it (an extension the
// reference's reader does not read, so its files stay readable there).
bool write_restart(const std::string& path, const Mesh& m, const double* Q, int ng, int citer, double cfl,
                   bool ghosts);
bool read_bc(const std::string& path, Mesh* m, std::string* err);

// ---------------- preprocessing ----------------
void cuthill_mckee(Mesh* m);                  // reorder.C:58-303
bool check_orientation(Mesh* m, std::string* err);  // metrics.C:31-294
void build_maps(Mesh* m);                     // maps.C:35-517
bool build_geometry(Mesh* m, std::string* err);    // geometry.C + cv_calc.C
void build_edge_weights(Mesh* m);             // gradient.C:923-1121
void build_crs(Mesh* m);                      // init_memory.C:36-138
void build_incidence(Mesh* m);
void build_coloring(Mesh* m);
// k=2 stencil: layer 1 and 2 for every node (solve.C:57-114 schedule),
// extra layers while fewer than 6 neighbours; then FP64 pseudo-inverse.
bool build_stencils(Mesh* m, std::string* err);
bool build_pseudo_inverse(Mesh* m, std::string* err);

// Full pipeline used by the solver: reorder (p.reorder), orientation check,
// maps, geometry, weights, CRS, incidence, coloring, stencils if order 3.
bool prepare(Mesh* m, const Params& p, std::string* err);

// One-sided Jacobi SVD pseudo-inverse of a column-major m x n matrix A
// (m >= n).  Writes W (n x m, row-major: W[i*m+j]).
bool pinv_jacobi(const double* A, int m, int n, double* W);

}  // namespace fs
