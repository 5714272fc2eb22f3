// Checks the reference-shaped entry points of include/fs_reference_shim.hpp.
//
//   shim_check roe <in> <out>     host Roe_flux_centered over a file of faces
//                                 (n, then n x [nx ny len qL(4) qR(4)]); writes n x 4 fluxes
//   shim_check solve nx ny seed order limit steps
//                                 builds a reference-layout GRID (1-based, after the
//                                 reordering) of a seeded bump channel, advances it with
//                                 fsref::solve and, in step, the C-ABI solver on the same
//                                 mesh; then one reference call sequence through the
//                                 loop-level wrappers.  Exit 0 when every compared array
//                                 is identical bit for bit.
//   shim_check curved bc xL yL xR yR
//                                 the curved_boundary_* functions on one segment: prints
//                                 the midpoint, the 3 Gauss points, the normal at (xL, yL)
//                                 and the arclength (%.17g)
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fs_reference_shim.hpp"

static int roe_mode(const char* in, const char* out) {
  FILE* f = std::fopen(in, "rb");
  if (!f) return 2;
  int64_t n = 0;
  if (std::fread(&n, sizeof n, 1, f) != 1) return 2;
  std::vector<double> a(11 * (size_t)n), r(4 * (size_t)n);
  if (std::fread(a.data(), sizeof(double), a.size(), f) != a.size()) return 2;
  std::fclose(f);
  for (int64_t i = 0; i < n; ++i) {
    double* p = &a[11 * i];
    fsref::Roe_flux_centered(p[0], p[1], p[2], 1.4, p + 3, p + 7, &r[4 * i]);
  }
  f = std::fopen(out, "wb");
  if (!f) return 2;
  std::fwrite(r.data(), sizeof(double), r.size(), f);
  std::fclose(f);
  return 0;
}

template <typename T>
static std::vector<T> arr(fs_mesh* m, int which) {
  const void* d = nullptr;
  int64_t n = 0;
  int dt = 0;
  if (fs_mesh_get_array(m, which, &d, &n, &dt)) { std::printf("get_array: %s\n", fs_last_error()); std::exit(2); }
  const T* p = static_cast<const T*>(d);
  return std::vector<T>(p, p + n);
}

static bool same(const char* what, const double* a, const double* b, size_t n) {
  size_t bad = 0;
  double mx = 0;
  for (size_t i = 0; i < n; ++i)
    if (std::memcmp(&a[i], &b[i], sizeof(double)) != 0) {
      ++bad;
      double d = a[i] - b[i];
      if (d < 0) d = -d;
      if (d > mx) mx = d;
    }
  std::printf("  %-8s %s (%zu of %zu differ, max |diff| %.3e)\n", what, bad ? "DIFFERS" : "identical", bad, n, mx);
  return bad == 0;
}

static int solve_mode(int nx, int ny, int seed, int order, int limit, int steps) {
  fs_params p;
  fs_params_default(&p);
  p.iter = 1; p.jufreq = 1; p.CFL = 20.0; p.ramp = 10; p.gamma = 1.4; p.alpha = 1.0; p.mach = 0.5;
  p.subits = 10; p.method = 1; p.order = order; p.foits = 2; p.soits = order == 3 ? 2 : 0; p.limit = limit;
  p.reorder = 1;
  // the reference's preprocessing order: read, reorder, then the solver sees the renumbered GRID
  fs_mesh* m = nullptr;
  if (fs_mesh_create_bump(nx, ny, 0.2, (uint64_t)seed, &m) || fs_mesh_prepare(m, &p)) {
    std::printf("mesh: %s\n", fs_last_error());
    return 2;
  }
  fs_mesh_info mi;
  fs_mesh_get_info(m, &mi);
  std::vector<double> x = arr<double>(m, FS_M_X);
  std::vector<int> tri = arr<int>(m, FS_M_TRI), boff = arr<int>(m, FS_M_BSEG_OFF), bseg = arr<int>(m, FS_M_BSEG);
  std::vector<int> bbc = arr<int>(m, FS_M_BBC);
  // GRID in the reference layout
  fsref::GRID g;
  g.nn = mi.nn;
  g.num_elem[fsref::kTri] = mi.ntri;
  std::vector<int> c2n_tri(3 * ((size_t)mi.ntri + 1));
  for (int t = 0; t < mi.ntri; ++t)
    for (int k = 0; k < 3; ++k) c2n_tri[3 * (t + 1) + k] = tri[3 * t + k] + 1;
  int* c2n[3] = {nullptr, c2n_tri.data(), nullptr};
  g.c2n = c2n;
  g.nb = mi.nb;
  std::vector<int> nbs(mi.nb + 1), bbc1(mi.nb + 1);
  std::vector<std::vector<int*>> bsb(mi.nb + 1);
  std::vector<int> segs(2 * bseg.size() + 2);
  std::vector<int**> bs(mi.nb + 1);
  for (int b = 0; b < mi.nb; ++b) {
    nbs[b + 1] = boff[b + 1] - boff[b];
    bbc1[b + 1] = bbc[b];
    bsb[b + 1].assign(nbs[b + 1] + 1, nullptr);
    for (int s = 0; s < nbs[b + 1]; ++s) {
      int* seg = &segs[2 * (boff[b] + s)];
      seg[0] = bseg[2 * (boff[b] + s)] + 1;
      seg[1] = bseg[2 * (boff[b] + s) + 1] + 1;
      bsb[b + 1][s + 1] = seg;
    }
    bs[b + 1] = bsb[b + 1].data();
  }
  g.nbs = nbs.data();
  g.bs = bs.data();
  g.bbc = bbc1.data();
  size_t ntot = (size_t)mi.nn + mi.ng;
  std::vector<double> x1(2 * (ntot + 1)), Q1(4 * (ntot + 1)), nQ1(4 * (ntot + 1));
  for (size_t i = 0; i < 2 * (size_t)mi.nn; ++i) x1[2 + i] = x[i];
  g.x = x1.data();
  g.Q = Q1.data();
  g.nQ = nQ1.data();
  fsref::SYS_MEM sm;
  std::vector<double> lhs(16 * (size_t)mi.nnz), rhs(4 * ((size_t)mi.nn + 1));
  sm.LHS = lhs.data();
  sm.RHS = rhs.data();

  // the C-ABI solver on the same prepared mesh
  fs_solver* s = nullptr;
  if (fs_solver_create(m, &p, FS_PREC_F64, &s) || fs_solver_init_freestream(s)) {
    std::printf("solver: %s\n", fs_last_error());
    return 2;
  }
  std::vector<double> Q(4 * ntot), buf;
  fs_solver_get(s, FS_F_Q, Q.data(), (int64_t)Q.size());
  std::memcpy(&Q1[4], Q.data(), Q.size() * sizeof(double));
  bool ok = true;
  for (int ts = 1; ts <= steps; ++ts) {
    double dq_shim = fsref::solve(&g, &sm, p, ts);
    double dq = 0, r2 = 0;
    fs_solve(s, ts, &dq, &r2);
    fs_solver_get(s, FS_F_Q, Q.data(), (int64_t)Q.size());
    std::printf("step %d |dQ| shim %.17g c-abi %.17g\n", ts, dq_shim, dq);
    ok = same("Q", &Q1[4], Q.data(), Q.size()) && ok;
    ok = (dq_shim == dq) && ok;
    buf.resize(4 * (size_t)mi.nn);
    fs_solver_get(s, FS_F_R, buf.data(), (int64_t)buf.size());
    ok = same("R", &g.R[4], buf.data(), buf.size()) && ok;
  }
  // one reference call sequence (solve.C:352-767) through the loop-level wrappers
  int ts = steps + 1;
  g.citer = ts;
  std::memcpy(&nQ1[4], &Q1[4], 4 * ntot * sizeof(double));
  fsref::build_matrix(&g, p, &sm);
  bool grad_phase = (order == 1 && ts > p.foits) || (order == 3 && ts > p.foits && ts <= p.foits + p.soits);
  if (grad_phase) fsref::Compute_Gradient_LeastSquares(&g);
  if (order == 3 && ts > p.foits + p.soits) fsref::Compute_Derivatives(&g, p, 3);
  if (limit == 1 && order == 1 && ts > p.foits) fsref::compute_limiter_barth(&g, p);
  if (limit == 2 && order == 1 && ts > p.foits) fsref::compute_limiter_venkatakrishnan(&g, p);
  fsref::build_interior_residuals(&g, &sm, p, 1, ts);
  fsref::build_boundary_residuals(&g, &sm, p, 1);
  for (int n = 1; n <= mi.nn; ++n)
    for (int v = 0; v < 4; ++v) rhs[4 * n + v] = -g.R[4 * n + v];
  double rms = fsref::linear_solve(&g, &sm, p);
  // the same sequence on the C-ABI solver
  fs_solver_set(s, FS_F_NQ, Q.data(), (int64_t)Q.size());
  fs_build_matrix(s, ts);
  if (grad_phase) fs_gradient(s);
  if (order == 3 && ts > p.foits + p.soits) fs_derivatives(s);
  if (limit > 0 && order == 1 && ts > p.foits) fs_limiter(s, ts);
  fs_interior_residual(s, 1, ts);
  fs_boundary_residual(s, 1);
  buf.resize(4 * (size_t)mi.nn);
  fs_solver_get(s, FS_F_R, buf.data(), (int64_t)buf.size());
  ok = same("seq R", &g.R[4], buf.data(), buf.size()) && ok;
  for (double& v : buf) v = -v;
  fs_solver_set(s, FS_F_RHS, buf.data(), (int64_t)buf.size());
  double rms2 = 0;
  fs_linear_solve(s, &rms2);
  fs_solver_get(s, FS_F_DQ, buf.data(), (int64_t)buf.size());
  ok = same("seq dQ", &g.dQ[4], buf.data(), buf.size()) && ok;
  std::vector<double> L(16 * (size_t)mi.nnz);
  fs_solver_get(s, FS_F_LHS, L.data(), (int64_t)L.size());
  ok = same("seq LHS", lhs.data(), L.data(), L.size()) && ok;
  std::printf("linear rms shim %.17g c-abi %.17g\n", rms, rms2);
  ok = (rms == rms2) && ok;
  fsref::detach(&g);
  fs_solver_destroy(s);
  fs_mesh_destroy(m);
  std::free(g.dQ); std::free(g.R); std::free(g.grad); std::free(g.phi); std::free(g.hess);
  std::free(g.point_value); std::free(g.ev_con); std::free(g.dt);
  std::printf("%s\n", ok ? "ALL IDENTICAL" : "MISMATCH");
  return ok ? 0 : 1;
}

struct NoGrid {};  // the curved functions take a GRID* they do not read

static int curved_mode(int bc, double xl, double yl, double xr, double yr) {
  double xL[2] = {xl, yl}, xR[2] = {xr, yr}, xM[2], GP[6], nx, ny, len;
  NoGrid* g = nullptr;
  fsref::curved_boundary_midpoint(g, bc, xL, xR, xM);
  fsref::curved_boundary_gauss_points(g, bc, xL, xR, GP);
  fsref::curved_boundary_normal_vector(g, bc, xL, &nx, &ny);
  fsref::curved_boundary_arclength(g, bc, xL, xR, &len);
  std::printf("%.17g %.17g\n", xM[0], xM[1]);
  for (int k = 0; k < 3; ++k) std::printf("%.17g %.17g\n", GP[2 * k], GP[2 * k + 1]);
  std::printf("%.17g %.17g\n%.17g\n", nx, ny, len);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && !std::strcmp(argv[1], "roe")) return roe_mode(argv[2], argv[3]);
  if (argc == 8 && !std::strcmp(argv[1], "solve"))
    return solve_mode(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                      std::atoi(argv[6]), std::atoi(argv[7]));
  if (argc == 7 && !std::strcmp(argv[1], "curved"))
    return curved_mode(std::atoi(argv[2]), std::atof(argv[3]), std::atof(argv[4]), std::atof(argv[5]),
                       std::atof(argv[6]));
  std::printf("usage: shim_check roe <in> <out> | solve nx ny seed order limit steps | curved bc xL yL xR yR\n");
  return 2;
}
