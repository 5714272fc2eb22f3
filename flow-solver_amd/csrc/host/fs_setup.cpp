// Host-side setup: generators, grid format, maps, renumbering, geometry.
// Orderings and floating-point operation order follow the reference files
// cited per function so that maps are bit-exact and FP64 geometry agrees to
// the last bit where the same operations are performed.
#include "fs_setup.hpp"

#include "fs_curved.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace fs {

// ------------------------------------------------------------------------
// Seeded PRNG: SplitMix64 seeding xoshiro256**.
// ------------------------------------------------------------------------
namespace {
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t t = z;
      t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
      t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
      s[i] = t ^ (t >> 31);
    }
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double sym(double a) { return (2.0 * uniform() - 1.0) * a; }
};

inline int nverts(int type) { return type == 0 ? 3 : 4; }
}  // namespace

// ------------------------------------------------------------------------
// Parameters file: 24 values, each preceded by one comment line
// (params.C:44-217).
// ------------------------------------------------------------------------
bool read_params(const std::string& path, Params* p, std::string* err) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) { if (err) *err = "cannot open " + path; return false; }
  char buf[256];
  auto nexti = [&](int* v) { if (!std::fgets(buf, sizeof buf, fp) || !std::fgets(buf, sizeof buf, fp)) return false; return std::sscanf(buf, "%d", v) == 1; };
  auto nextd = [&](double* v) { if (!std::fgets(buf, sizeof buf, fp) || !std::fgets(buf, sizeof buf, fp)) return false; return std::sscanf(buf, "%lf", v) == 1; };
  bool ok = nexti(&p->nsteps) && nexti(&p->iter) && nexti(&p->jufreq) && nextd(&p->CFL) &&
            nexti(&p->ramp) && nextd(&p->gamma) && nextd(&p->alpha) && nextd(&p->mach) &&
            nexti(&p->subits) && nexti(&p->unsteady) && nexti(&p->tacc) && nextd(&p->mintime) &&
            nexti(&p->method) && nexti(&p->order) && nexti(&p->foits) && nexti(&p->soits) &&
            nexti(&p->toits) && nexti(&p->isrestart) && nexti(&p->limit) && nexti(&p->reorder) &&
            nexti(&p->ic) && nexti(&p->p_or_cv) && nexti(&p->start_post_pro) && nexti(&p->post_pro_freq);
  std::fclose(fp);
  if (!ok && err) *err = "malformed parameter file " + path;
  return ok;
}

bool write_params(const std::string& path, const Params& p) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return false;
  std::fprintf(fp, "#nsteps\n%d\n#iter\n%d\n#jufreq\n%d\n#CFL\n%.17g\n#ramp\n%d\n", p.nsteps, p.iter, p.jufreq, p.CFL, p.ramp);
  std::fprintf(fp, "#gamma\n%.17g\n#alpha\n%.17g\n#mach\n%.17g\n#subits\n%d\n#unsteady\n%d\n", p.gamma, p.alpha, p.mach, p.subits, p.unsteady);
  std::fprintf(fp, "#tacc\n%d\n#mintime\n%.17g\n#method\n%d\n#order\n%d\n#foits\n%d\n#soits\n%d\n#toits\n%d\n", p.tacc, p.mintime, p.method, p.order, p.foits, p.soits, p.toits);
  std::fprintf(fp, "#isrestart\n%d\n#limit\n%d\n#reorder\n%d\n#ic\n%d\n#p_or_cv\n%d\n#start_post_pro\n%d\n#post_pro_freq\n%d\n",
               p.isrestart, p.limit, p.reorder, p.ic, p.p_or_cv, p.start_post_pro, p.post_pro_freq);
  std::fclose(fp);
  return true;
}

// ------------------------------------------------------------------------
// Grid format (mesh_io.C:53-214, 318-378).
// ------------------------------------------------------------------------
bool write_grid(const std::string& path, const Mesh& m) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return false;
  std::fprintf(fp, "#Number of grid points\n%d\n", m.nn);
  for (int n = 0; n < m.nn; ++n) std::fprintf(fp, "%.17g %.17g\n", m.x[2 * n], m.x[2 * n + 1]);
  std::fprintf(fp, "#Number of blocks\n1\n#Number of triangles\n%d\n", m.ntri);
  for (int t = 0; t < m.ntri; ++t) std::fprintf(fp, "%d %d %d\n", m.tri[3 * t] + 1, m.tri[3 * t + 1] + 1, m.tri[3 * t + 2] + 1);
  std::fprintf(fp, "#Number of quads\n%d\n", m.nquad);
  for (int t = 0; t < m.nquad; ++t)
    std::fprintf(fp, "%d %d %d %d\n", m.quad[4 * t] + 1, m.quad[4 * t + 1] + 1, m.quad[4 * t + 2] + 1, m.quad[4 * t + 3] + 1);
  std::fprintf(fp, "#Number of boundaries\n%d\n", m.nb);
  for (int b = 0; b < m.nb; ++b) {
    int s0 = m.bseg_off[b], s1 = m.bseg_off[b + 1];
    std::fprintf(fp, "#Number of edges in boundary %d\n%d\n", b + 1, s1 - s0);
    for (int s = s0; s < s1; ++s) std::fprintf(fp, "%d %d\n", m.bseg[2 * s] + 1, m.bseg[2 * s + 1] + 1);
  }
  std::fclose(fp);
  return true;
}

bool read_grid(const std::string& path, Mesh* m, std::string* err) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) { if (err) *err = "cannot open " + path; return false; }
  char buf[512];
  auto line = [&]() { return std::fgets(buf, sizeof buf, fp) != nullptr; };
  auto fail = [&](const char* what) { if (err) *err = std::string("malformed grid: ") + what; std::fclose(fp); return false; };
  int nblk = 0;
  if (!line() || !line() || std::sscanf(buf, "%d", &m->nn) != 1 || m->nn <= 0) return fail("node count");
  m->x.assign(2 * (size_t)m->nn, 0.0);
  for (int n = 0; n < m->nn; ++n)
    if (!line() || std::sscanf(buf, "%lf %lf", &m->x[2 * n], &m->x[2 * n + 1]) != 2) return fail("coordinates");
  if (!line() || !line() || std::sscanf(buf, "%d", &nblk) != 1 || nblk != 1) return fail("block count must be 1");
  if (!line() || !line() || std::sscanf(buf, "%d", &m->ntri) != 1 || m->ntri < 0) return fail("triangle count");
  m->tri.assign(3 * (size_t)m->ntri, 0);
  for (int t = 0; t < m->ntri; ++t) {
    int a, b, c;
    if (!line() || std::sscanf(buf, "%d %d %d", &a, &b, &c) != 3) return fail("triangles");
    m->tri[3 * t] = a - 1; m->tri[3 * t + 1] = b - 1; m->tri[3 * t + 2] = c - 1;
  }
  if (!line() || !line() || std::sscanf(buf, "%d", &m->nquad) != 1 || m->nquad < 0) return fail("quad count");
  m->quad.assign(4 * (size_t)m->nquad, 0);
  for (int t = 0; t < m->nquad; ++t) {
    int a, b, c, d;
    if (!line() || std::sscanf(buf, "%d %d %d %d", &a, &b, &c, &d) != 4) return fail("quads");
    m->quad[4 * t] = a - 1; m->quad[4 * t + 1] = b - 1; m->quad[4 * t + 2] = c - 1; m->quad[4 * t + 3] = d - 1;
  }
  if (!line() || !line() || std::sscanf(buf, "%d", &m->nb) != 1 || m->nb <= 0) return fail("boundary count");
  m->bseg_off.assign(m->nb + 1, 0);
  m->bseg.clear();
  for (int b = 0; b < m->nb; ++b) {
    int ns;
    if (!line() || !line() || std::sscanf(buf, "%d", &ns) != 1 || ns < 0) return fail("segment count");
    for (int s = 0; s < ns; ++s) {
      int a, c;
      if (!line() || std::sscanf(buf, "%d %d", &a, &c) != 2) return fail("segments");
      m->bseg.push_back(a - 1); m->bseg.push_back(c - 1);
    }
    m->bseg_off[b + 1] = m->bseg_off[b] + ns;
  }
  // optional restart payload (mesh_io.C:217-263): constants (last step, last
  // CFL) and the state of every real node; then, as an extension the
  // reference's reader skips, the ghost states
  m->has_restart = false;
  m->restart_Q.clear();
  m->restart_ghost_Q.clear();
  bool more = line();
  while (more && buf[0] != '#' && std::strspn(buf, " \t\r\n") == std::strlen(buf)) more = line();  // blank lines
  if (more && buf[0] == '#' && std::strstr(buf, "Constants")) {
    int nconst = 0, nvar = 0;
    if (!line() || std::sscanf(buf, "%d", &nconst) != 1 || nconst != 2) return fail("restart constants");
    if (!line() || std::sscanf(buf, "%d", &m->restart_citer) != 1) return fail("restart step");
    if (!line() || std::sscanf(buf, "%lf", &m->restart_cfl) != 1) return fail("restart CFL");
    if (!line() || !line() || std::sscanf(buf, "%d", &nvar) != 1 || nvar != 4) return fail("restart variable count");
    m->restart_Q.assign(4 * (size_t)m->nn, 0.0);
    for (int n = 0; n < m->nn; ++n) {
      double* q = &m->restart_Q[4 * (size_t)n];
      if (!line() || std::sscanf(buf, "%lf %lf %lf %lf", q, q + 1, q + 2, q + 3) != 4) return fail("restart state");
    }
    m->has_restart = true;
    int ng = 0;
    if (line() && line() && std::sscanf(buf, "%d", &ng) == 1 && ng > 0) {
      m->restart_ghost_Q.assign(4 * (size_t)ng, 0.0);
      for (int g = 0; g < ng; ++g) {
        double* q = &m->restart_ghost_Q[4 * (size_t)g];
        if (!line() || std::sscanf(buf, "%lf %lf %lf %lf", q, q + 1, q + 2, q + 3) != 4) return fail("restart ghost state");
      }
    }
  }
  std::fclose(fp);
  for (int v : m->tri) if (v < 0 || v >= m->nn) { if (err) *err = "triangle node out of range"; return false; }
  for (int v : m->quad) if (v < 0 || v >= m->nn) { if (err) *err = "quad node out of range"; return false; }
  for (int v : m->bseg) if (v < 0 || v >= m->nn) { if (err) *err = "segment node out of range"; return false; }
  if ((int)m->bbc.size() != m->nb) m->bbc.assign(m->nb, 0);
  return true;
}

bool write_restart(const std::string& path, const Mesh& m, const double* Q, int ng, int citer, double cfl,
                   bool ghosts) {
  if (!write_grid(path, m)) return false;
  FILE* fp = std::fopen(path.c_str(), "a");
  if (!fp) return false;
  // mesh_io.C:380-389 layout of the payload
  std::fprintf(fp, "#Number of Constants\n%d\n%d\n%.17E\n", 2, citer, cfl);
  std::fprintf(fp, "#Number of Variables\n%d\n", 4);
  for (int n = 0; n < m.nn; ++n)
    std::fprintf(fp, "%.17e %.17e %.17e %.17e\n", Q[4 * n], Q[4 * n + 1], Q[4 * n + 2], Q[4 * n + 3]);
  if (ghosts) {
    std::fprintf(fp, "#Number of ghost states (bedge order)\n%d\n", ng);
    for (int g = 0; g < ng; ++g) {
      const double* q = Q + 4 * ((size_t)m.nn + g);
      std::fprintf(fp, "%.17e %.17e %.17e %.17e\n", q[0], q[1], q[2], q[3]);
    }
  }
  std::fclose(fp);
  return true;
}

bool write_bc(const std::string& path, const Mesh& m) {
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return false;
  for (int b = 0; b < m.nb; ++b) std::fprintf(fp, "%d\n", m.bbc[b]);
  std::fclose(fp);
  return true;
}

bool read_bc(const std::string& path, Mesh* m, std::string* err) {
  FILE* fp = std::fopen(path.c_str(), "r");
  if (!fp) { if (err) *err = "cannot open " + path; return false; }
  char buf[256];
  m->bbc.assign(m->nb, 0);
  for (int b = 0; b < m->nb; ++b) {
    if (!std::fgets(buf, sizeof buf, fp) || std::sscanf(buf, "%d", &m->bbc[b]) != 1) {
      std::fclose(fp);
      if (err) *err = "bc file has fewer entries than boundaries";
      return false;
    }
  }
  std::fclose(fp);
  return true;
}

// ------------------------------------------------------------------------
// Generators.
// ------------------------------------------------------------------------
void make_bump_channel(int nx, int ny, double jitter, uint64_t seed, Mesh* m) {
  *m = Mesh();
  m->nn = nx * ny;
  m->x.assign(2 * (size_t)m->nn, 0.0);
  auto yb = [](double x) { return (x >= 1.0 && x <= 2.0) ? std::sqrt(1.3 * 1.3 - (x - 1.5) * (x - 1.5)) - 1.2 : 0.0; };
  const double dx = 3.0 / (nx - 1);
  Rng rng(seed);
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      double xi = 3.0 * i / (nx - 1);
      double ybi = yb(xi);
      double dy = (1.0 - ybi) / (ny - 1);
      double px = xi, py = ybi + (1.0 - ybi) * j / (double)(ny - 1);
      if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && jitter > 0.0) {
        double ux = rng.sym(jitter), uy = rng.sym(jitter);
        px += ux * dx;
        py += uy * dy;
      }
      int n = j * nx + i;
      m->x[2 * n] = px;
      m->x[2 * n + 1] = py;
    }
  }
  auto id = [nx](int i, int j) { return j * nx + i; };
  m->ntri = 2 * (nx - 1) * (ny - 1);
  m->tri.reserve(3 * (size_t)m->ntri);
  for (int j = 0; j < ny - 1; ++j)
    for (int i = 0; i < nx - 1; ++i) {
      int a = id(i, j), b = id(i + 1, j), c = id(i + 1, j + 1), d = id(i, j + 1);
      m->tri.insert(m->tri.end(), {a, b, c});
      m->tri.insert(m->tri.end(), {a, c, d});
    }
  m->nb = 4;
  m->bseg_off = {0};
  for (int i = 0; i < nx - 1; ++i) m->bseg.insert(m->bseg.end(), {id(i, 0), id(i + 1, 0)});
  m->bseg_off.push_back((int)m->bseg.size() / 2);
  for (int j = 0; j < ny - 1; ++j) m->bseg.insert(m->bseg.end(), {id(nx - 1, j), id(nx - 1, j + 1)});
  m->bseg_off.push_back((int)m->bseg.size() / 2);
  for (int i = nx - 1; i > 0; --i) m->bseg.insert(m->bseg.end(), {id(i, ny - 1), id(i - 1, ny - 1)});
  m->bseg_off.push_back((int)m->bseg.size() / 2);
  for (int j = ny - 1; j > 0; --j) m->bseg.insert(m->bseg.end(), {id(0, j), id(0, j - 1)});
  m->bseg_off.push_back((int)m->bseg.size() / 2);
  m->bbc = {1, 0, 1, 0};  // lower wall, outlet, upper wall, inlet
}

namespace {
// Triangles and boundary segments of a ring mesh with nw nodes per ring and
// nrad rings (node r*nw + k): each quad split into two counter-clockwise
// triangles; boundary 0 the inner ring traversed clockwise, boundary 1 the
// outer ring counter-clockwise, so the domain lies left of every segment.
void ring_connect(int nw, int nrad, Mesh* m) {
  auto id = [nw](int r, int k) { return r * nw + (k % nw); };
  m->ntri = 2 * nw * (nrad - 1);
  m->tri.reserve(3 * (size_t)m->ntri);
  for (int r = 0; r < nrad - 1; ++r)
    for (int k = 0; k < nw; ++k) {
      int a = id(r, k), b = id(r, k + 1), c = id(r + 1, k + 1), d = id(r + 1, k);
      m->tri.insert(m->tri.end(), {a, d, c});
      m->tri.insert(m->tri.end(), {a, c, b});
    }
  m->nb = 2;
  m->bseg_off = {0};
  for (int k = 0; k < nw; ++k) m->bseg.insert(m->bseg.end(), {id(0, k + 1), id(0, k)});
  m->bseg_off.push_back((int)m->bseg.size() / 2);
  for (int k = 0; k < nw; ++k) m->bseg.insert(m->bseg.end(), {id(nrad - 1, k), id(nrad - 1, k + 1)});
  m->bseg_off.push_back((int)m->bseg.size() / 2);
}
}  // namespace

void make_airfoil_omesh(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed, Mesh* m) {
  *m = Mesh();
  const double PI = 3.14159265358979323846;
  auto yt = [](double x) {
    x = std::min(std::max(x, 0.0), 1.0);
    return 0.6 * (0.2969 * std::sqrt(x) - 0.126 * x - 0.3516 * x * x + 0.2843 * x * x * x - 0.1036 * x * x * x * x);
  };
  // stretching ratio q such that the first normalised layer has height h0/rmax
  double target = h0 / rmax, lo = 1.0 + 1e-12, hi = 2.0;
  int N = nrad - 1;
  for (int it = 0; it < 200; ++it) {
    double q = 0.5 * (lo + hi);
    double s1 = (q - 1.0) / (std::pow(q, N) - 1.0);
    if (s1 > target) lo = q; else hi = q;
  }
  double q = 0.5 * (lo + hi);
  std::vector<double> s(nrad);
  for (int r = 0; r < nrad; ++r) s[r] = (std::pow(q, r) - 1.0) / (std::pow(q, N) - 1.0);
  m->nn = nwall * nrad;
  m->x.assign(2 * (size_t)m->nn, 0.0);
  Rng rng(seed);
  const double dbeta = 2.0 * PI / nwall;
  for (int r = 0; r < nrad; ++r) {
    for (int k = 0; k < nwall; ++k) {
      double beta = k * dbeta, sr = s[r];
      if (r > 0 && r < nrad - 1 && jitter > 0.0) {
        double ub = rng.sym(jitter), us = rng.sym(jitter);
        beta += ub * dbeta;
        double ds = std::min(s[r + 1] - s[r], s[r] - s[r - 1]);
        sr += us * ds;
      }
      double cb = std::cos(beta), sb = std::sin(beta);
      double xw = 0.5 * (1.0 + cb);
      double yw = std::sin(beta) >= 0.0 ? yt(xw) : -yt(xw);
      if (k == 0 && r == 0) yw = 0.0;
      double xo = 0.5 + rmax * cb, yo = rmax * sb;
      int n = r * nwall + k;
      m->x[2 * n] = (1.0 - sr) * xw + sr * xo;
      m->x[2 * n + 1] = (1.0 - sr) * yw + sr * yo;
    }
  }
  ring_connect(nwall, nrad, m);
  m->bbc = {1, 0};  // wall (polyline), farfield
}

void make_airfoil_omesh_split(int nwall, int nrad, double h0, double rmax, double jitter, uint64_t seed,
                              int upper_bc, int lower_bc, Mesh* m) {
  make_airfoil_omesh(nwall, nrad, h0, rmax, jitter, seed, m);
  // wall segment k runs from wall node k+1 to k; nodes 0..nwall/2 are the
  // upper surface (trailing edge to leading edge)
  m->bseg_off = {0, nwall / 2, nwall, 2 * nwall};
  m->nb = 3;
  m->bbc = {upper_bc, lower_bc, 0};
}

void make_annulus(int ntheta, int nrad, double r_in, double r_out, double jitter, uint64_t seed, int bc_in,
                  int bc_out, Mesh* m) {
  *m = Mesh();
  const double PI = 3.14159265358979323846;
  m->nn = ntheta * nrad;
  m->x.assign(2 * (size_t)m->nn, 0.0);
  Rng rng(seed);
  const double dth = 2.0 * PI / ntheta;
  auto ring = [&](int r) { return r == 0 ? r_in : r == nrad - 1 ? r_out : r_in * std::pow(r_out / r_in, r / (nrad - 1.0)); };
  for (int r = 0; r < nrad; ++r)
    for (int k = 0; k < ntheta; ++k) {
      double th = k * dth, rr = ring(r);
      if (r > 0 && r < nrad - 1 && jitter > 0.0) {
        const double dr = std::min(ring(r + 1) - rr, rr - ring(r - 1));
        th += rng.sym(jitter) * dth;
        rr += rng.sym(jitter) * dr;
      }
      double px = rr * std::cos(th), py = rr * std::sin(th);
      // the two nodes on the x axis lie exactly on it
      if (k == 0) { px = rr; py = 0.0; }
      if (2 * k == ntheta) { px = -rr; py = 0.0; }
      int n = r * ntheta + k;
      m->x[2 * n] = px;
      m->x[2 * n + 1] = py;
    }
  ring_connect(ntheta, nrad, m);
  m->bbc = {bc_in, bc_out};
}

// ------------------------------------------------------------------------
// Maps (maps.C:35-517, mesh_utils.C:285-379).
// ------------------------------------------------------------------------
namespace {
// global element id g (0-based): triangles first, then quads
inline void elem_nodes(const Mesh& m, int g, const int** nodes, int* nv) {
  if (g < m.ntri) { *nodes = &m.tri[3 * (size_t)g]; *nv = 3; }
  else { *nodes = &m.quad[4 * (size_t)(g - m.ntri)]; *nv = 4; }
}

// nodes adjacent to vertex `node` inside element g, in the reference's
// vertex-dependent order (mesh_utils.C:285-379)
inline void elem_nodes_around(const Mesh& m, int g, int node, int out[2]) {
  const int* c; int nv;
  elem_nodes(m, g, &c, &nv);
  int v = 0;
  for (int i = 0; i < nv; ++i) if (c[i] == node) { v = i; break; }
  if (nv == 3) {
    static const int a[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    out[0] = c[a[v][0]]; out[1] = c[a[v][1]];
  } else {
    static const int a[4][2] = {{1, 3}, {0, 2}, {1, 3}, {0, 2}};
    out[0] = c[a[v][0]]; out[1] = c[a[v][1]];
  }
}

void cells_surrounding_node(const Mesh& m, std::vector<int>& ncsn, std::vector<int>& csn) {
  ncsn.assign(m.nn + 1, 0);
  int nel = m.ntri + m.nquad;
  for (int g = 0; g < nel; ++g) {
    const int* c; int nv; elem_nodes(m, g, &c, &nv);
    for (int k = 0; k < nv; ++k) ncsn[c[k] + 1]++;
  }
  for (int i = 1; i <= m.nn; ++i) ncsn[i] += ncsn[i - 1];
  csn.assign(ncsn[m.nn], 0);
  std::vector<int> pos(ncsn.begin(), ncsn.end() - 1);
  for (int g = 0; g < nel; ++g) {
    const int* c; int nv; elem_nodes(m, g, &c, &nv);
    for (int k = 0; k < nv; ++k) csn[pos[c[k]]++] = g;
  }
}

void nodes_surrounding_node(const Mesh& m, const std::vector<int>& ncsn, const std::vector<int>& csn,
                            std::vector<int>& nnsn, std::vector<int>& nsn) {
  nnsn.assign(m.nn + 1, 0);
  nsn.clear();
  nsn.reserve(6 * (size_t)m.nn + 16);
  for (int i = 0; i < m.nn; ++i) {
    int base = (int)nsn.size();
    for (int j = ncsn[i]; j < ncsn[i + 1]; ++j) {
      int list[2];
      elem_nodes_around(m, csn[j], i, list);
      for (int k = 0; k < 2; ++k) {
        bool dup = false;
        if (j != ncsn[i])
          for (int l = base; l < (int)nsn.size(); ++l) if (nsn[l] == list[k]) { dup = true; break; }
        if (!dup) nsn.push_back(list[k]);
      }
    }
    nnsn[i + 1] = (int)nsn.size();
  }
}
}  // namespace

// Cuthill-McKee renumbering with a stable descending-degree sort of each
// newly reached neighbour set (reorder.C:58-303).  glibc's qsort sorts
// these arrays with a merge sort, so equal degrees keep their nsn order.
void cuthill_mckee(Mesh* m) {
  std::vector<int> ncsn, csn, nnsn, nsn;
  cells_surrounding_node(*m, ncsn, csn);
  nodes_surrounding_node(*m, ncsn, csn, nnsn, nsn);
  const int nn = m->nn;
  std::vector<int> map(nn, -1), queued(nn, 0), S;
  S.reserve(nn);
  auto deg = [&](int i) { return nnsn[i + 1] - nnsn[i]; };
  int seed = 0, maxdeg = deg(0);
  for (int i = 0; i < nn; ++i) if (deg(i) > maxdeg) { maxdeg = deg(i); seed = i; }
  int inode = 0, startS = 0;
  map[seed] = inode++;
  queued[seed] = 1;
  std::vector<int> R;
  while (inode < nn) {
    R.clear();
    for (int j = nnsn[seed]; j < nnsn[seed + 1]; ++j)
      if (!queued[nsn[j]]) { R.push_back(nsn[j]); queued[nsn[j]] = 1; }
    std::stable_sort(R.begin(), R.end(), [&](int a, int b) { return deg(a) > deg(b); });
    for (int r : R) { map[r] = inode++; S.push_back(r); }
    if (startS >= (int)S.size()) break;  // disconnected remainder: keep reference behaviour bounded
    seed = S[startS++];
  }
  // any node never reached keeps a fresh number in file order
  for (int i = 0; i < nn; ++i) if (map[i] < 0) map[i] = inode++;
  std::vector<double> tx(m->x.size());
  for (int i = 0; i < nn; ++i) { tx[2 * map[i]] = m->x[2 * i]; tx[2 * map[i] + 1] = m->x[2 * i + 1]; }
  std::copy(tx.begin(), tx.begin() + 2 * (size_t)nn, m->x.begin());
  for (int& v : m->tri) v = map[v];
  for (int& v : m->quad) v = map[v];
  for (int& v : m->bseg) v = map[v];
  m->nn_map = map;
}

bool check_orientation(Mesh* m, std::string* err) {
  // metrics.C:31-294: every element must have the same sign; an all-CW mesh
  // is rewound to CCW.
  int neg = 0, tot = m->ntri + m->nquad;
  for (int t = 0; t < m->ntri; ++t) {
    const int* c = &m->tri[3 * (size_t)t];
    double v1x = m->x[2 * c[1]] - m->x[2 * c[0]], v1y = m->x[2 * c[1] + 1] - m->x[2 * c[0] + 1];
    double v2x = m->x[2 * c[2]] - m->x[2 * c[0]], v2y = m->x[2 * c[2] + 1] - m->x[2 * c[0] + 1];
    double area = 0.5 * (v1x * v2y - v1y * v2x);
    if (std::fabs(area) < 1e-15) { if (err) *err = "degenerate triangle"; return false; }
    if (area < 0) ++neg;
  }
  for (int t = 0; t < m->nquad; ++t) {
    const int* c = &m->quad[4 * (size_t)t];
    double v1x = m->x[2 * c[2]] - m->x[2 * c[0]], v1y = m->x[2 * c[2] + 1] - m->x[2 * c[0] + 1];
    double v2x = m->x[2 * c[3]] - m->x[2 * c[1]], v2y = m->x[2 * c[3] + 1] - m->x[2 * c[1] + 1];
    double area = 0.5 * (v1x * v2y - v1y * v2x);
    if (std::fabs(area) < 1e-15) { if (err) *err = "degenerate quad"; return false; }
    if (area < 0) ++neg;
  }
  if (neg == 0) return true;
  if (neg != tot) { if (err) *err = "mixed element orientation"; return false; }
  for (int t = 0; t < m->ntri; ++t) std::swap(m->tri[3 * (size_t)t + 1], m->tri[3 * (size_t)t + 2]);
  for (int t = 0; t < m->nquad; ++t) std::swap(m->quad[4 * (size_t)t + 1], m->quad[4 * (size_t)t + 3]);
  return true;
}

void build_maps(Mesh* m) {
  cells_surrounding_node(*m, m->ncsn, m->csn);
  nodes_surrounding_node(*m, m->ncsn, m->csn, m->nnsn, m->nsn);
  // edges (maps.C:375-467): numbered by lower node, then nsn order
  const int nn = m->nn;
  m->esn.assign(m->nsn.size(), -1);
  m->edges.clear();
  m->ne = 0;
  for (int i = 0; i < nn; ++i) {
    for (int j = m->nnsn[i]; j < m->nnsn[i + 1]; ++j) {
      int r = m->nsn[j];
      if (i > r) continue;
      int e = m->ne++;
      m->edges.push_back(i);
      m->edges.push_back(r);
      m->esn[j] = e;
      for (int k = m->nnsn[r]; k < m->nnsn[r + 1]; ++k)
        if (m->nsn[k] == i) { m->esn[k] = e; break; }
    }
  }
  // cell to edge (maps.C:517-600)
  m->c2e.clear();
  int nel = m->ntri + m->nquad;
  for (int g = 0; g < nel; ++g) {
    const int* c; int nv; elem_nodes(*m, g, &c, &nv);
    for (int k = 0; k < nv; ++k) {
      int l = c[k], r = c[(k + 1) % nv];
      if (l > r) std::swap(l, r);
      int e = -1;
      for (int s = m->nnsn[l]; s < m->nnsn[l + 1]; ++s) if (m->nsn[s] == r) { e = m->esn[s]; break; }
      m->c2e.push_back(e);
    }
  }
}

// ------------------------------------------------------------------------
// Geometry (geometry.C:31-391, 407-835, 993-1197; cv_calc.C:34-225,
// 981-2148).  Operation order follows the reference expression by
// expression.
// ------------------------------------------------------------------------
bool build_geometry(Mesh* m, std::string* err) {
  const int nn = m->nn;
  const int nel = m->ntri + m->nquad;
  const double* X = nullptr;
  // element centroids
  m->el_cent.assign(2 * (size_t)nel, 0.0);
  for (int g = 0; g < nel; ++g) {
    const int* c; int nv; elem_nodes(*m, g, &c, &nv);
    double xc = 0., yc = 0.;
    for (int k = 0; k < nv; ++k) { xc += m->x[2 * c[k]]; yc += m->x[2 * c[k] + 1]; }
    xc /= (nv * 1.);
    yc /= (nv * 1.);
    m->el_cent[2 * g] = xc;
    m->el_cent[2 * g + 1] = yc;
  }
  // subedges, element-major
  m->nsub = 3 * m->ntri + 4 * m->nquad;
  m->sub_edge.assign(m->nsub, 0);
  m->sub_elem.assign(m->nsub, 0);
  m->sub_n.assign(3 * (size_t)m->nsub, 0.0);
  m->sub_xm.assign(2 * (size_t)m->nsub, 0.0);
  std::vector<int> hits(m->ne, 0);
  X = m->x.data();
  int is = 0;
  for (int g = 0; g < nel; ++g) {
    const int* c; int nv; elem_nodes(*m, g, &c, &nv);
    double xc = m->el_cent[2 * g], yc = m->el_cent[2 * g + 1];
    for (int k = 0; k < nv; ++k, ++is) {
      int L = c[k], R = c[(k + 1) % nv];
      if (L > R) std::swap(L, R);
      double xmid = 0.5 * (X[2 * L] + X[2 * R]);
      double ymid = 0.5 * (X[2 * L + 1] + X[2 * R + 1]);
      int e = -1;
      for (int s = m->nnsn[L]; s < m->nnsn[L + 1]; ++s) {
        int ee = m->esn[s];
        if (m->edges[2 * ee + 1] == R) { e = ee; break; }
      }
      if (e < 0) { if (err) *err = "subedge without edge"; return false; }
      double nx = yc - ymid, ny = xmid - xc;
      if (R == c[k]) { nx = -nx; ny = -ny; }
      m->sub_edge[is] = e;
      m->sub_elem[is] = g;
      m->sub_n[3 * is] = nx;
      m->sub_n[3 * is + 1] = ny;
      m->sub_xm[2 * is] = xmid;
      m->sub_xm[2 * is + 1] = ymid;
      hits[e]++;
    }
  }
  for (int e = 0; e < m->ne; ++e)
    if (hits[e] < 1 || hits[e] > 2) { if (err) *err = "edge shared by more than two elements"; return false; }
  for (int s = 0; s < m->nsub; ++s) {
    double nx = m->sub_n[3 * s], ny = m->sub_n[3 * s + 1];
    double mag = std::sqrt(nx * nx + ny * ny);
    if (mag > 1.0e-20) { nx /= mag; ny /= mag; }
    m->sub_n[3 * s] = nx; m->sub_n[3 * s + 1] = ny; m->sub_n[3 * s + 2] = mag;
  }
  // boundary edges: two per segment, each with its own ghost node
  int nseg = m->bseg_off[m->nb];
  m->nbe = 2 * nseg;
  m->ng = m->nbe;
  m->x.resize(2 * (size_t)(nn + m->ng), 0.0);
  X = m->x.data();
  m->bedge.assign(5 * (size_t)m->nbe, 0);
  m->bedge_n.assign(3 * (size_t)m->nbe, 0.0);
  int ib = 0;
  for (int b = 0; b < m->nb; ++b) {
    const int bc = m->bbc[b];
    if (curved_bc(bc) && !curved_bc_known(bc)) {
      if (err) *err = "unknown curved boundary code " + std::to_string(bc) + " (codes 11-22 are defined)";
      return false;
    }
    for (int s = m->bseg_off[b]; s < m->bseg_off[b + 1]; ++s) {
      int n0 = m->bseg[2 * s], n1 = m->bseg[2 * s + 1];
      if (curved_bc(bc)) {
        // geometry.C:636-690: both ghosts at the curved midpoint, the
        // normals of the curve at the segment's own nodes
        const double* xL = &m->x[2 * (size_t)n0];
        const double* xR = &m->x[2 * (size_t)n1];
        if (bc >= 17 && bc <= 20 && (xL[0] < 0.0 || xR[0] < 0.0)) {
          if (err) *err = "NACA 0012 boundary node with x < 0";
          return false;
        }
        double xM[2];
        curved_midpoint(bc, xL, xR, xM);
        for (int h = 0; h < 2; ++h) {
          int real = h == 0 ? n0 : n1;
          int* be = &m->bedge[5 * (size_t)ib];
          be[0] = real; be[1] = ib; be[3] = b; be[4] = s - m->bseg_off[b];
          curved_normal(bc, h == 0 ? xL : xR, &m->bedge_n[3 * ib], &m->bedge_n[3 * ib + 1]);
          m->x[2 * (nn + ib)] = xM[0];
          m->x[2 * (nn + ib) + 1] = xM[1];
          ++ib;
        }
        continue;
      }
      double xmid = (m->x[2 * n0] + m->x[2 * n1]) / 2.;
      double ymid = (m->x[2 * n0 + 1] + m->x[2 * n1 + 1]) / 2.;
      double nx = m->x[2 * n1 + 1] - ymid;
      double ny = xmid - m->x[2 * n1];
      for (int h = 0; h < 2; ++h) {
        int real = h == 0 ? n0 : n1;
        int* be = &m->bedge[5 * (size_t)ib];
        be[0] = real; be[1] = ib; be[3] = b; be[4] = s - m->bseg_off[b];
        m->bedge_n[3 * ib] = nx;
        m->bedge_n[3 * ib + 1] = ny;
        m->x[2 * (nn + ib)] = (m->x[2 * real] + xmid) / 2.;
        m->x[2 * (nn + ib) + 1] = (m->x[2 * real + 1] + ymid) / 2.;
        ++ib;
      }
    }
  }
  // element owning each segment: first element (triangles, then quads, in
  // order) containing n0 followed by n1 in its cyclic vertex order
  {
    // directed-edge lookup through the nodes' cell lists instead of a full scan;
    // the lowest global element id matching is the reference's first match
    for (int b = 0; b < m->nb; ++b)
      for (int s = m->bseg_off[b]; s < m->bseg_off[b + 1]; ++s) {
        int n0 = m->bseg[2 * s], n1 = m->bseg[2 * s + 1];
        int best = -1;
        for (int j = m->ncsn[n0]; j < m->ncsn[n0 + 1]; ++j) {
          int g = m->csn[j];
          const int* c; int nv; elem_nodes(*m, g, &c, &nv);
          int k = 0;
          for (; k < nv; ++k) if (c[k] == n0) break;
          if (k == nv) continue;
          if (c[(k + 1) % nv] != n1) continue;
          if (best < 0 || g < best) best = g;
        }
        if (best < 0) { if (err) *err = "boundary segment not found in any element (orientation?)"; return false; }
        int seg_local = s - m->bseg_off[b];
        for (int i = 0; i < m->nbe; ++i)
          if (m->bedge[5 * (size_t)i + 3] == b && m->bedge[5 * (size_t)i + 4] == seg_local) m->bedge[5 * (size_t)i + 2] = best;
      }
  }
  for (int i = 0; i < m->nbe; ++i) {
    const int bc = m->bbc[m->bedge[5 * (size_t)i + 3]];
    if (curved_bc(bc)) {
      // geometry.C:773-784: unit normal kept, length along the curve from
      // the node to the curved midpoint
      const int node = m->bedge[5 * (size_t)i];
      m->bedge_n[3 * i + 2] = curved_arclength(bc, &m->x[2 * (size_t)node], &m->x[2 * (size_t)(nn + i)]);
      continue;
    }
    double nx = m->bedge_n[3 * i], ny = m->bedge_n[3 * i + 1];
    double mag = std::sqrt(nx * nx + ny * ny);
    if (mag > 1.0E-20) { nx /= mag; ny /= mag; }
    m->bedge_n[3 * i] = nx; m->bedge_n[3 * i + 1] = ny; m->bedge_n[3 * i + 2] = mag;
  }
  // Fix_Subedge_Data (geometry.C:852-981): the subedge of a curved segment's
  // edge runs from the element centroid to the curved midpoint
  {
    std::vector<int> first_sub(m->ne, -1);
    for (int s = m->nsub - 1; s >= 0; --s) first_sub[m->sub_edge[s]] = s;
    for (int i = 0; i < m->nbe; ++i) {
      const int* be = &m->bedge[5 * (size_t)i];
      const int b = be[3], bc = m->bbc[b];
      if (!curved_bc(bc)) continue;
      const int sg = m->bseg_off[b] + be[4];
      const int nL = m->bseg[2 * sg], nR = m->bseg[2 * sg + 1];
      const int lo = std::min(nL, nR), hi = std::max(nL, nR);
      int e = -1;
      for (int k = m->nnsn[lo]; k < m->nnsn[lo + 1]; ++k)
        if (m->edges[2 * m->esn[k] + 1] == hi) { e = m->esn[k]; break; }
      const int j = e < 0 ? -1 : first_sub[e];
      if (j < 0) { if (err) *err = "curved segment without an edge"; return false; }
      const int g = be[2];
      if (m->sub_elem[j] != g) { if (err) *err = "curved segment and its subedge lie in different elements"; return false; }
      const double XC0 = m->el_cent[2 * (size_t)g], XC1 = m->el_cent[2 * (size_t)g + 1];
      const double XM0 = m->x[2 * (size_t)(nn + be[1])], XM1 = m->x[2 * (size_t)(nn + be[1]) + 1];
      double nx = XC1 - XM1, ny = XM0 - XC0;
      const double len = std::sqrt(nx * nx + ny * ny);
      nx /= len;
      ny /= len;
      if (nL > nR) { nx = -nx; ny = -ny; }
      m->sub_n[3 * (size_t)j] = nx;
      m->sub_n[3 * (size_t)j + 1] = ny;
      m->sub_n[3 * (size_t)j + 2] = len;
      m->sub_xm[2 * (size_t)j] = XM0;
      m->sub_xm[2 * (size_t)j + 1] = XM1;
    }
  }
  // Gauss points (geometry.C:993-1197)
  const double t1 = std::sqrt(15.0) / 5.0, t2 = 0., t3 = std::sqrt(15.0) / (-5.0);
  m->sub_xg.assign(6 * (size_t)m->nsub, 0.0);
  for (int s = 0; s < m->nsub; ++s) {
    double xmid = m->sub_xm[2 * s], ymid = m->sub_xm[2 * s + 1];
    int g = m->sub_elem[s];
    double xc = m->el_cent[2 * g], yc = m->el_cent[2 * g + 1];
    double* o = &m->sub_xg[6 * (size_t)s];
    o[0] = (xmid + xc) * 0.5 + (xc - xmid) * 0.5 * t1;
    o[1] = (ymid + yc) * 0.5 + (yc - ymid) * 0.5 * t1;
    o[2] = (xmid + xc) * 0.5 + (xc - xmid) * 0.5 * t2;
    o[3] = (ymid + yc) * 0.5 + (yc - ymid) * 0.5 * t2;
    o[4] = (xmid + xc) * 0.5 + (xc - xmid) * 0.5 * t3;
    o[5] = (ymid + yc) * 0.5 + (yc - ymid) * 0.5 * t3;
  }
  m->bedge_xg.assign(6 * (size_t)m->nbe, 0.0);
  m->bedge_gn.assign(6 * (size_t)m->nbe, 0.0);
  for (int i = 0; i < m->nbe; ++i) {
    const int* be = &m->bedge[5 * (size_t)i];
    int real = be[0], b = be[3], s = m->bseg_off[b] + be[4];
    int nL = m->bseg[2 * s], nR = m->bseg[2 * s + 1];
    const int bc = m->bbc[b];
    double* o = &m->bedge_xg[6 * (size_t)i];
    double* on = &m->bedge_gn[6 * (size_t)i];
    if (curved_bc(bc)) {
      // geometry.C:1160-1194: Gauss points of the arc between the node and
      // the curved midpoint, each with the curve's normal there
      const double* xn = &m->x[2 * (size_t)real];
      const double* xg = &m->x[2 * (size_t)(nn + be[1])];
      if (!curved_gauss_points(bc, real == nL ? xn : xg, real == nL ? xg : xn, o, err)) return false;
      for (int k = 0; k < 3; ++k) curved_normal(bc, &o[2 * k], &on[2 * k], &on[2 * k + 1]);
      continue;
    }
    for (int k = 0; k < 3; ++k) {
      on[2 * k] = m->bedge_n[3 * (size_t)i];
      on[2 * k + 1] = m->bedge_n[3 * (size_t)i + 1];
    }
    double xi = m->x[2 * real], yi = m->x[2 * real + 1];
    double xmid = 0.5 * (m->x[2 * nL] + m->x[2 * nR]);
    double ymid = 0.5 * (m->x[2 * nL + 1] + m->x[2 * nR + 1]);
    double xL, xR, yL, yR;
    if (real == nL) { xL = xi; xR = xmid; yL = yi; yR = ymid; }
    else { xL = xmid; xR = xi; yL = ymid; yR = yi; }
    o[0] = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t1;
    o[1] = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t1;
    o[2] = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t2;
    o[3] = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t2;
    o[4] = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t3;
    o[5] = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t3;
  }
  // control-volume areas by Green's theorem (cv_calc.C:34-225)
  m->cv_area.assign(nn, 0.0);
  for (int s = 0; s < m->nsub; ++s) {
    int e = m->sub_edge[s], g = m->sub_elem[s];
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double qc = m->el_cent[2 * g];
    double nx = m->sub_n[3 * s], len = m->sub_n[3 * s + 2];
    double qmid = m->sub_xm[2 * s];
    m->cv_area[L] += (qc + qmid) / 2. * nx * len;
    m->cv_area[R] -= (qc + qmid) / 2. * nx * len;
  }
  const double wc[3] = {8. / 9., 5. / 9., 5. / 9.};  // weights of the curved Gauss points
  for (int b = 0; b < m->nb; ++b) {
    const int bc = m->bbc[b];
    for (int s = m->bseg_off[b]; s < m->bseg_off[b + 1]; ++s) {
      int n0 = m->bseg[2 * s], n1 = m->bseg[2 * s + 1];
      if (curved_bc(bc)) {
        // cv_calc.C:130-196: x nx integrated along each half arc
        const double* xL = &m->x[2 * (size_t)n0];
        const double* xR = &m->x[2 * (size_t)n1];
        double xM[2], GP[6];
        curved_midpoint(bc, xL, xR, xM);
        for (int h = 0; h < 2; ++h) {
          const double* a = h == 0 ? xL : xM;
          const double* c = h == 0 ? xM : xR;
          if (!curved_gauss_points(bc, a, c, GP, err)) return false;
          double Ftemp = 0.;
          for (int k = 0; k < 3; ++k) {
            double nx, ny;
            curved_normal(bc, &GP[2 * k], &nx, &ny);
            Ftemp += (wc[k] * GP[2 * k] * nx);
          }
          const double len = curved_arclength(bc, a, c);
          Ftemp = Ftemp * len * 0.5;
          m->cv_area[h == 0 ? n0 : n1] += Ftemp;
        }
        continue;
      }
      double ymid = (m->x[2 * n0 + 1] + m->x[2 * n1 + 1]) / 2.;
      double nx = m->x[2 * n1 + 1] - ymid;
      m->cv_area[n0] += (0.75 * m->x[2 * n0] + 0.25 * m->x[2 * n1]) * nx;
      m->cv_area[n1] += (0.75 * m->x[2 * n1] + 0.25 * m->x[2 * n0]) * nx;
    }
  }
  for (int n = 0; n < nn; ++n)
    if (!(m->cv_area[n] > 0.0)) {
      if (err) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "non-positive control volume area %.6e at node %d (%.6f, %.6f)", m->cv_area[n],
                      n, m->x[2 * (size_t)n], m->x[2 * (size_t)n + 1]);
        *err = buf;
      }
      return false;
    }
  // control-volume moments about each node (cv_calc.C:981-1518), boundary
  // integrals at 3 Gauss points, divided by the area
  m->moments.assign(9 * (size_t)nn, 0.0);
  auto face_moments = [&](double xi, double yi, double x1, double y1, double x2, double y2, double x3, double y3,
                          double nx, double len, double* I) {
    const double w1 = 5. / 9., w2 = 8. / 9., w3 = w1;
    double Ix = w1 * (x1 - xi) * (x1 - xi) + w2 * (x2 - xi) * (x2 - xi) + w3 * (x3 - xi) * (x3 - xi);
    I[0] = 0.5 * Ix * nx * len / 2.;
    double Iy = w1 * (x1 - xi) * (y1 - yi) + w2 * (x2 - xi) * (y2 - yi) + w3 * (x3 - xi) * (y3 - yi);
    I[1] = Iy * nx * len / 2.;
    double Ixx = w1 * (x1 - xi) * (x1 - xi) * (x1 - xi) + w2 * (x2 - xi) * (x2 - xi) * (x2 - xi) +
                 w3 * (x3 - xi) * (x3 - xi) * (x3 - xi);
    I[2] = (1. / 3.) * Ixx * nx * len / 2.;
    double Iyy = w1 * (x1 - xi) * (y1 - yi) * (y1 - yi) + w2 * (x2 - xi) * (y2 - yi) * (y2 - yi) +
                 w3 * (x3 - xi) * (y3 - yi) * (y3 - yi);
    I[3] = Iyy * nx * len / 2.;
    double Ixy = w1 * (x1 - xi) * (x1 - xi) * (y1 - yi) + w2 * (x2 - xi) * (x2 - xi) * (y2 - yi) +
                 w3 * (x3 - xi) * (x3 - xi) * (y3 - yi);
    I[4] = 0.5 * Ixy * nx * len / 2.;
    double Ixxx = w1 * (x1 - xi) * (x1 - xi) * (x1 - xi) * (x1 - xi) + w2 * (x2 - xi) * (x2 - xi) * (x2 - xi) * (x2 - xi) +
                  w3 * (x3 - xi) * (x3 - xi) * (x3 - xi) * (x3 - xi);
    I[5] = (0.25) * Ixxx * nx * len / 2.;
    double Iyyy = w1 * (x1 - xi) * (y1 - yi) * (y1 - yi) * (y1 - yi) + w2 * (x2 - xi) * (y2 - yi) * (y2 - yi) * (y2 - yi) +
                  w3 * (x3 - xi) * (y3 - yi) * (y3 - yi) * (y3 - yi);
    I[6] = Iyyy * nx * len / 2.;
    double Ixxy = w1 * (x1 - xi) * (x1 - xi) * (x1 - xi) * (y1 - yi) + w2 * (x2 - xi) * (x2 - xi) * (x2 - xi) * (y2 - yi) +
                  w3 * (x3 - xi) * (x3 - xi) * (x3 - xi) * (y3 - yi);
    I[7] = (1. / 3.) * Ixxy * nx * len / 2.;
    double Ixyy = w1 * (x1 - xi) * (x1 - xi) * (y1 - yi) * (y1 - yi) + w2 * (x2 - xi) * (x2 - xi) * (y2 - yi) * (y2 - yi) +
                  w3 * (x3 - xi) * (x3 - xi) * (y3 - yi) * (y3 - yi);
    I[8] = 0.5 * Ixyy * nx * len / 2.;
  };
  for (int s = 0; s < m->nsub; ++s) {
    int e = m->sub_edge[s], g = m->sub_elem[s];
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double xc = m->el_cent[2 * g], yc = m->el_cent[2 * g + 1];
    double nx = m->sub_n[3 * s], len = m->sub_n[3 * s + 2];
    double xmid = m->sub_xm[2 * s], ymid = m->sub_xm[2 * s + 1];
    double x1 = (xmid + xc) * 0.5 + (xc - xmid) * 0.5 * t1, y1 = (ymid + yc) * 0.5 + (yc - ymid) * 0.5 * t1;
    double x2 = (xmid + xc) * 0.5 + (xc - xmid) * 0.5 * t2, y2 = (ymid + yc) * 0.5 + (yc - ymid) * 0.5 * t2;
    double x3 = (xmid + xc) * 0.5 + (xc - xmid) * 0.5 * t3, y3 = (ymid + yc) * 0.5 + (yc - ymid) * 0.5 * t3;
    double I[9];
    face_moments(m->x[2 * L], m->x[2 * L + 1], x1, y1, x2, y2, x3, y3, nx, len, I);
    for (int k = 0; k < 9; ++k) m->moments[9 * (size_t)L + k] += I[k];
    face_moments(m->x[2 * R], m->x[2 * R + 1], x1, y1, x2, y2, x3, y3, nx, len, I);
    for (int k = 0; k < 9; ++k) m->moments[9 * (size_t)R + k] -= I[k];
  }
  for (int i = 0; i < m->nbe; ++i) {
    const int* be = &m->bedge[5 * (size_t)i];
    int node = be[0], b = be[3], s = m->bseg_off[b] + be[4];
    int nL = m->bseg[2 * s], nR = m->bseg[2 * s + 1];
    double xi = m->x[2 * node], yi = m->x[2 * node + 1];
    const int bc = m->bbc[b];
    if (curved_bc(bc)) {
      // cv_calc.C:1428-1568: the moments' boundary integrals along the half
      // arc between the node and the curved midpoint
      double XM[2], XL[2], XR[2], GP[6];
      curved_midpoint(bc, &m->x[2 * (size_t)nL], &m->x[2 * (size_t)nR], XM);
      if (node == nL) { XL[0] = xi; XL[1] = yi; XR[0] = XM[0]; XR[1] = XM[1]; }
      else { XL[0] = XM[0]; XL[1] = XM[1]; XR[0] = xi; XR[1] = yi; }
      if (!curved_gauss_points(bc, XL, XR, GP, err)) return false;
      const double len = curved_arclength(bc, XL, XR);
      double I[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
      for (int k = 0; k < 3; ++k) {
        double nx, ny;
        curved_normal(bc, &GP[2 * k], &nx, &ny);
        const double dx = GP[2 * k] - xi, dy = GP[2 * k + 1] - yi;
        I[0] += (wc[k] * dx * dx * nx * len);
        I[1] += (wc[k] * dx * dy * nx * len);
        I[2] += (wc[k] * dx * dx * dx * nx * len);
        I[3] += (wc[k] * dx * dy * dy * nx * len);
        I[4] += (wc[k] * dx * dx * dy * nx * len);
        I[5] += (wc[k] * dx * dx * dx * dx * nx * len);
        I[6] += (wc[k] * dx * dy * dy * dy * nx * len);
        I[7] += (wc[k] * dx * dx * dx * dy * nx * len);
        I[8] += (wc[k] * dx * dx * dy * dy * nx * len);
      }
      I[0] = 0.5 * I[0] * 0.5;
      I[1] = I[1] * 0.5;
      I[2] = (1. / 3.) * I[2] * 0.5;
      I[3] = I[3] * 0.5;
      I[4] = 0.5 * I[4] * 0.5;
      I[5] = (0.25) * I[5] * 0.5;
      I[6] = I[6] * 0.5;
      I[7] = (1. / 3.) * I[7] * 0.5;
      I[8] = 0.5 * I[8] * 0.5;
      for (int k = 0; k < 9; ++k) m->moments[9 * (size_t)node + k] += I[k];
      continue;
    }
    double xmid = 0.5 * (m->x[2 * nL] + m->x[2 * nR]);
    double ymid = 0.5 * (m->x[2 * nL + 1] + m->x[2 * nR + 1]);
    double len = m->bedge_n[3 * i + 2], nx = m->bedge_n[3 * i];
    double xL, xR, yL, yR;
    if (node == nL) { xL = xi; xR = xmid; yL = yi; yR = ymid; }
    else { xL = xmid; xR = xi; yL = ymid; yR = yi; }
    double x1 = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t1, y1 = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t1;
    double x2 = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t2, y2 = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t2;
    double x3 = (xL + xR) * 0.5 + (xR - xL) * 0.5 * t3, y3 = (yL + yR) * 0.5 + (yR - yL) * 0.5 * t3;
    double I[9];
    face_moments(xi, yi, x1, y1, x2, y2, x3, y3, nx, len, I);
    for (int k = 0; k < 9; ++k) m->moments[9 * (size_t)node + k] += I[k];
  }
  for (int n = 0; n < nn; ++n)
    for (int k = 0; k < 9; ++k) m->moments[9 * (size_t)n + k] /= m->cv_area[n];
  return true;
}

// LSQ edge weights: closed-form 2x2 normal equations (gradient.C:923-1121)
void build_edge_weights(Mesh* m) {
  const int nn = m->nn;
  std::vector<double> r11(nn, 0.), r12(nn, 0.), s22(nn, 0.), r22o(nn, 0.);
  for (int e = 0; e < m->ne; ++e) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double dx0 = m->x[2 * R] - m->x[2 * L], dx1 = m->x[2 * R + 1] - m->x[2 * L + 1];
    double t = dx0 * dx0;
    r11[L] += t; r11[R] += t;
    t = dx0 * dx1; r12[L] += t;
    t = (-dx0) * (-dx1); r12[R] += t;
    t = dx1 * dx1;
    s22[L] += t; s22[R] += t;
  }
  for (int i = 0; i < nn; ++i) r22o[i] = s22[i] - (r12[i] * r12[i]) / r11[i];
  m->edge_weight.assign(4 * (size_t)m->ne, 0.0);
  for (int e = 0; e < m->ne; ++e) {
    int L = m->edges[2 * e], R = m->edges[2 * e + 1];
    double dx0 = m->x[2 * R] - m->x[2 * L], dx1 = m->x[2 * R + 1] - m->x[2 * L + 1];
    double* w = &m->edge_weight[4 * (size_t)e];
    w[1] = (+dx1 - r12[L] / r11[L] * (+dx0)) / r22o[L];
    w[0] = (+dx0 - r12[L] * w[1]) / r11[L];
    w[3] = (-dx1 - r12[R] / r11[R] * (-dx0)) / r22o[R];
    w[2] = (-dx0 - r12[R] * w[3]) / r11[R];
  }
}

// Block-CSR pattern: each row lists its nsn neighbours in order, then the
// diagonal last (init_memory.C:36-138).
void build_crs(Mesh* m) {
  const int nn = m->nn;
  m->ia.assign(nn + 1, 0);
  m->iau.assign(nn, 0);
  m->ja.clear();
  for (int n = 0; n < nn; ++n) {
    m->ia[n] = m->nnsn[n] + n;
    for (int j = m->nnsn[n]; j < m->nnsn[n + 1]; ++j) m->ja.push_back(m->nsn[j]);
    m->iau[n] = (int)m->ja.size();
    m->ja.push_back(n);
  }
  m->ia[nn] = (int)m->ja.size();
  // slots of the two off-diagonal blocks owned by each edge
  m->edge_slot_lr.assign(m->ne, -1);
  m->edge_slot_rl.assign(m->ne, -1);
  for (int n = 0; n < nn; ++n)
    for (int j = m->nnsn[n]; j < m->nnsn[n + 1]; ++j) {
      int e = m->esn[j], slot = m->ia[n] + (j - m->nnsn[n]);
      if (m->edges[2 * e] == n) m->edge_slot_lr[e] = slot; else m->edge_slot_rl[e] = slot;
    }
  m->slot_t.assign(m->ja.size(), -1);
  for (int e = 0; e < m->ne; ++e) {
    m->slot_t[m->edge_slot_lr[e]] = m->edge_slot_rl[e];
    m->slot_t[m->edge_slot_rl[e]] = m->edge_slot_lr[e];
  }
  for (int n = 0; n < nn; ++n) m->slot_t[m->iau[n]] = m->iau[n];
}

void build_incidence(Mesh* m) {
  const int nn = m->nn;
  m->nsub_off.assign(nn + 1, 0);
  for (int s = 0; s < m->nsub; ++s) {
    int e = m->sub_edge[s];
    m->nsub_off[m->edges[2 * e] + 1]++;
    m->nsub_off[m->edges[2 * e + 1] + 1]++;
  }
  for (int n = 0; n < nn; ++n) m->nsub_off[n + 1] += m->nsub_off[n];
  m->nsub_list.assign(m->nsub_off[nn], 0);
  {
    std::vector<int> pos(m->nsub_off.begin(), m->nsub_off.end() - 1);
    for (int s = 0; s < m->nsub; ++s) {
      int e = m->sub_edge[s];
      m->nsub_list[pos[m->edges[2 * e]]++] = s;
      m->nsub_list[pos[m->edges[2 * e + 1]]++] = s | (int)0x80000000u;
    }
  }
  m->nedge_off.assign(nn + 1, 0);
  for (int e = 0; e < m->ne; ++e) { m->nedge_off[m->edges[2 * e] + 1]++; m->nedge_off[m->edges[2 * e + 1] + 1]++; }
  for (int n = 0; n < nn; ++n) m->nedge_off[n + 1] += m->nedge_off[n];
  m->nedge_list.assign(m->nedge_off[nn], 0);
  {
    std::vector<int> pos(m->nedge_off.begin(), m->nedge_off.end() - 1);
    for (int e = 0; e < m->ne; ++e) {
      m->nedge_list[pos[m->edges[2 * e]]++] = e;
      m->nedge_list[pos[m->edges[2 * e + 1]]++] = e | (int)0x80000000u;
    }
  }
  m->nbe_off.assign(nn + 1, 0);
  for (int i = 0; i < m->nbe; ++i) m->nbe_off[m->bedge[5 * (size_t)i] + 1]++;
  for (int n = 0; n < nn; ++n) m->nbe_off[n + 1] += m->nbe_off[n];
  m->nbe_list.assign(m->nbe, 0);
  {
    std::vector<int> pos(m->nbe_off.begin(), m->nbe_off.end() - 1);
    for (int i = 0; i < m->nbe; ++i) m->nbe_list[pos[m->bedge[5 * (size_t)i]]++] = i;
  }
  m->esub_off.assign(m->ne + 1, 0);
  for (int s = 0; s < m->nsub; ++s) m->esub_off[m->sub_edge[s] + 1]++;
  for (int e = 0; e < m->ne; ++e) m->esub_off[e + 1] += m->esub_off[e];
  m->esub_list.assign(m->nsub, 0);
  {
    std::vector<int> pos(m->esub_off.begin(), m->esub_off.end() - 1);
    for (int s = 0; s < m->nsub; ++s) m->esub_list[pos[m->sub_edge[s]]++] = s;
  }
}

// Greedy distance-1 coloring of the matrix graph in node order: every node
// takes the smallest color none of its already-colored neighbours holds.
// Deterministic, so the host result is the coloring every device uses.
void build_coloring(Mesh* m) {
  const int nn = m->nn;
  m->color.assign(nn, -1);
  std::vector<int> mark;
  int nc = 0;
  for (int n = 0; n < nn; ++n) {
    mark.assign(nc + 1, 0);
    for (int j = m->nnsn[n]; j < m->nnsn[n + 1]; ++j) {
      int c = m->color[m->nsn[j]];
      if (c >= 0) mark[c] = 1;
    }
    int c = 0;
    while (c < nc && mark[c]) ++c;
    m->color[n] = c;
    if (c == nc) ++nc;
  }
  m->ncolor = nc;
  m->color_off.assign(nc + 1, 0);
  for (int n = 0; n < nn; ++n) m->color_off[m->color[n] + 1]++;
  for (int c = 0; c < nc; ++c) m->color_off[c + 1] += m->color_off[c];
  m->color_rows.assign(nn, 0);
  std::vector<int> pos(m->color_off.begin(), m->color_off.end() - 1);
  for (int n = 0; n < nn; ++n) m->color_rows[pos[m->color[n]]++] = n;
}

// Generic stencil: the node's layer-1 neighbours in nsn order, then layer 2
// built by visiting the previous layer in order and each member's nsn slots
// in order, skipping nodes already present (maps.C:1316-1551).  Layers are
// added as solve.C:57-114 does for order 3: two layers for every node, then
// up to three more while the stencil has fewer than 6 neighbours.
bool build_stencils(Mesh* m, std::string* err) {
  const int nn = m->nn;
  m->st_off.assign(nn + 1, 0);
  m->st_nodes.clear();
  std::vector<int> st, seen_stamp(nn, -1);
  for (int i = 0; i < nn; ++i) {
    st.clear();
    st.push_back(i);
    seen_stamp[i] = i;
    int layer_begin = 0, layer_end = 1;
    auto add_layer = [&]() {
      int nb = (int)st.size();
      for (int p = layer_begin; p < layer_end; ++p) {
        int nb_node = st[p];
        for (int j = m->nnsn[nb_node]; j < m->nnsn[nb_node + 1]; ++j) {
          int e = m->esn[j];
          int a = m->edges[2 * e], b = m->edges[2 * e + 1];
          int cand = (a != nb_node) ? a : b;
          if (seen_stamp[cand] == i) continue;
          seen_stamp[cand] = i;
          st.push_back(cand);
        }
      }
      layer_begin = layer_end;
      layer_end = (int)st.size();
      (void)nb;
    };
    add_layer();
    add_layer();
    int k = 0;
    for (; (int)st.size() - 1 < 6 && k < 3; ++k) add_layer();
    if ((int)st.size() - 1 < 6) { if (err) *err = "stencil smaller than 6 after 5 layers"; return false; }
    m->st_nodes.insert(m->st_nodes.end(), st.begin() + 1, st.end());
    m->st_off[i + 1] = (int)m->st_nodes.size();
  }
  return true;
}

// One-sided Jacobi (Hestenes) SVD; W = V diag(1/s^2) (A V)^T = A^+.
bool pinv_jacobi(const double* A, int m, int n, double* W) {
  std::vector<double> U(A, A + (size_t)m * n);  // column-major copy
  std::vector<double> V((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double alpha = 0, beta = 0, gamma = 0;
        for (int i = 0; i < m; ++i) {
          double up = U[(size_t)p * m + i], uq = U[(size_t)q * m + i];
          alpha += up * up; beta += uq * uq; gamma += up * uq;
        }
        if (alpha == 0.0 || beta == 0.0) continue;
        double c = std::fabs(gamma) / std::sqrt(alpha * beta);
        off = std::max(off, c);
        if (c < 1e-15) continue;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
        for (int i = 0; i < m; ++i) {
          double up = U[(size_t)p * m + i], uq = U[(size_t)q * m + i];
          U[(size_t)p * m + i] = cs * up - sn * uq;
          U[(size_t)q * m + i] = sn * up + cs * uq;
        }
        for (int i = 0; i < n; ++i) {
          double vp = V[(size_t)p * n + i], vq = V[(size_t)q * n + i];
          V[(size_t)p * n + i] = cs * vp - sn * vq;
          V[(size_t)q * n + i] = sn * vp + cs * vq;
        }
      }
    if (off < 1e-15) break;
  }
  std::vector<double> s2(n, 0.0);
  for (int k = 0; k < n; ++k) {
    for (int i = 0; i < m; ++i) s2[k] += U[(size_t)k * m + i] * U[(size_t)k * m + i];
    if (!(s2[k] > 0.0)) return false;
  }
  for (int r = 0; r < n; ++r)
    for (int j = 0; j < m; ++j) {
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += V[(size_t)k * n + r] * U[(size_t)k * m + j] / s2[k];
      W[(size_t)r * m + j] = acc;
    }
  return true;
}

// Weighted k=2 least-squares matrix per node (compute_derivatives.C:859-913:
// rows w_j*[dx+Mx_j-Mx_i, dy+My_j-My_i, quadratic moment columns], w=1/dist)
// and its pseudo-inverse with the row weights folded into the columns.
// Nodes are independent and each writes only its own stencil's entries, so
// contiguous node ranges run on host threads; the result does not depend on
// the thread count.
namespace {
bool pinv_nodes(Mesh* m, int i0, int i1) {
  std::vector<double> A, W, wv;
  const double* mo = m->moments.data();
  for (int i = i0; i < i1; ++i) {
    int o = m->st_off[i], mm = m->st_off[i + 1] - o;
    A.assign((size_t)mm * 5, 0.0);
    W.assign((size_t)mm * 5, 0.0);
    wv.assign(mm, 0.0);
    double xi = m->x[2 * i], yi = m->x[2 * i + 1];
    for (int r = 0; r < mm; ++r) {
      int node = m->st_nodes[o + r];
      double xn = m->x[2 * node], yn = m->x[2 * node + 1];
      double dx0 = xn - xi, dx1 = yn - yi;
      double dist = std::sqrt(dx0 * dx0 + dx1 * dx1);
      double w = 1.0 / std::pow(dist, 1.0);
      wv[r] = w;
      const double* Mn = mo + 9 * (size_t)node;
      const double* Mi = mo + 9 * (size_t)i;
      A[r + 0 * mm] = w * (xn - xi + Mn[0] - Mi[0]);
      A[r + 1 * mm] = w * (yn - yi + Mn[1] - Mi[1]);
      A[r + 2 * mm] = 0.5 * w * ((xn - xi) * (xn - xi) + 2.0 * Mn[0] * (xn - xi) + Mn[2] - Mi[2]);
      A[r + 3 * mm] = 0.5 * w * ((yn - yi) * (yn - yi) + 2.0 * Mn[1] * (yn - yi) + Mn[3] - Mi[3]);
      A[r + 4 * mm] = w * (Mn[4] + Mn[0] * (yn - yi) + Mn[1] * (xn - xi) + (yn - yi) * (xn - xi) - Mi[4]);
    }
    if (!pinv_jacobi(A.data(), mm, 5, W.data())) return false;
    for (int r = 0; r < 5; ++r)
      for (int j = 0; j < mm; ++j) m->st_W[5 * (size_t)o + (size_t)r * mm + j] = W[(size_t)r * mm + j] * wv[j];
  }
  return true;
}
}  // namespace

bool build_pseudo_inverse(Mesh* m, std::string* err) {
  const int nn = m->nn;
  m->st_W.assign(5 * m->st_nodes.size(), 0.0);
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nth = nn < 8192 ? 1 : (int)std::min(16u, hw);
  bool ok = true;
  if (nth == 1) {
    ok = pinv_nodes(m, 0, nn);
  } else {
    std::vector<std::thread> th;
    std::vector<char> res(nth, 1);
    for (int t = 0; t < nth; ++t) {
      const int i0 = (int)((int64_t)nn * t / nth), i1 = (int)((int64_t)nn * (t + 1) / nth);
      th.emplace_back([m, i0, i1, t, &res] { res[t] = pinv_nodes(m, i0, i1) ? 1 : 0; });
    }
    for (auto& x : th) x.join();
    for (char r : res) ok = ok && r;
  }
  if (!ok && err) *err = "rank-deficient k=2 stencil";
  return ok;
}

bool prepare(Mesh* m, const Params& p, std::string* err) {
  if (p.isrestart == 0 && p.reorder == 1) cuthill_mckee(m);
  if (!check_orientation(m, err)) return false;
  build_maps(m);
  if (!build_geometry(m, err)) return false;
  build_edge_weights(m);
  build_crs(m);
  build_incidence(m);
  build_coloring(m);
  if (p.order == 3) {
    if (!build_stencils(m, err)) return false;
    if (!build_pseudo_inverse(m, err)) return false;
  }
  return true;
}

}  // namespace fs
