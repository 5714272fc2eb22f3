// Test-infrastructure harness: the reference-shaped entry points of
// include/fs_reference_shim.hpp against the reference's own functions, on one
// GRID prepared by the reference's own setup sequence (main.C:105-648).
//
// Compiled against the reference sources in place and linked with their
// objects and libflowsolver.so (oracle/Makefile target ref); never part of the
// product.  For every face / boundary edge of the mesh it calls the
// reference's function and the shim's with the same inputs and counts the
// outputs that differ in any bit:
//   host:   Reconstruct, Reconstruct_Gauss, Reconstruct_Gauss_Boundary in the
//           first-order, linear and k=2 phases; Get_Boundary_Value at the
//           node and the Gauss points of every boundary edge;
//           compute_numerical_jacobian_roe, compute_approximate_jacobian_roe,
//           compute_boundary_approximate_jacobian_roe and
//           compute_boundary_numerical_jacobian_roe accumulated over every
//           face into SYS_MEM's LHS; Build_CRS; Compute_Edge_Weights
//   device: Set_Boundary_Values and cv_calc_eigenvalue_contribution (device
//           passes over the GRID's twin)
// and, in jac mode, writes the reference's compute_approximate_jacobian_roe
// accumulated over every face at the linear phase's reconstructed (unequal)
// face states, with the state and limiter values it started from, for the
// oracle's restatement of that Jacobian to be compared with:
//   jac_nQ.bin (nn+ng+1)*4, jac_phi.bin (nn+1)*8, jac_lhs.bin nnz*16 doubles
//   (the reference's 1-based arrays as they are), jac_meta.txt "nn ng nnz ts"
//
// Usage: ref_pointwise <case_dir> host|device|jac
//   case_dir holds euler.params (order 3), bc.dat and grid.mesh.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unistd.h>

#include "Defined_Variables.h"
#include "grid.h"
#include "params.h"
#include "sys_mem.h"
#include "cv_calc.h"
#include "cvbc.h"
#include "flux.h"
#include "geometry.h"
#include "gradient.h"
#include "ic.h"
#include "init_memory.h"
#include "jacobian.h"
#include "maps.h"
#include "mesh_io.h"
#include "metrics.h"
#include "reconstruct.h"
#include "reorder.h"

#include "fs_reference_shim.hpp"

static int g_fail = 0;

static void report(const char* name, long bad, long total) {
  std::printf("check %-44s mismatches %ld of %ld\n", name, bad, total);
  if (bad || !total) g_fail = 1;
}

static long diff(const double* a, const double* b, long n) {
  long bad = 0;
  for (long i = 0; i < n; ++i)
    if (std::memcmp(&a[i], &b[i], sizeof(double)) != 0) ++bad;
  return bad;
}

// deterministic values in [0, 1)
static unsigned long long g_lcg = 12345;
static double urand() {
  g_lcg = g_lcg * 6364136223846793005ULL + 1442695040888963407ULL;
  return (double)(g_lcg >> 11) * (1.0 / 9007199254740992.0);
}

static double* zeros(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: ref_pointwise <case_dir> host|device|jac\n");
    return 2;
  }
  const bool device = std::strcmp(argv[2], "device") == 0;
  const bool jac = std::strcmp(argv[2], "jac") == 0;
  if (chdir(argv[1]) != 0) {
    std::fprintf(stderr, "cannot chdir to %s\n", argv[1]);
    return 2;
  }
  GRID grid;
  SYS_MEM smem;
  PARAMS p;
  int i, n, gn;
  char meshname[] = "grid.mesh";
  // main.C's setup sequence
  read_parameters(&p);
  read_mesh(meshname, &grid, p);
  if (p.isrestart == 0 && p.reorder == 1) Cuthill_McKee(&grid);
  metrics2D(&grid);
  Cells_Surrounding_Node(grid.nn, grid.num_elem, grid.c2n, &(grid.csn), &(grid.ncsn));
  Nodes_Surrounding_Node(grid.nn, grid.num_elem, grid.c2n, grid.csn, grid.ncsn, &(grid.nsn), &(grid.nnsn));
  Edges_Surrounding_Node(grid.nn, grid.nsn, grid.nnsn, &(grid.nedges), &(grid.edges), &(grid.esn));
  Build_Cell_to_Edge_Map(grid.c2n, grid.num_elem, grid.nsn, grid.nnsn, grid.esn, &(grid.c2e));
  Elements_Surrounding_Element(grid.num_elem, grid.c2n, grid.csn, grid.ncsn, &(grid.ese), &(grid.nese));
  Nodes_Surrounding_Node2(grid.nn, grid.num_elem, grid.c2n, grid.esn, grid.nnsn, grid.edges, &(grid.nsn2), &(grid.nnsn2));
  Initialize_Generic_Nodes_Surrounding_Node(grid.nn, grid.num_elem, grid.c2n, grid.esn, grid.nnsn, grid.edges,
                                            &(grid.gnsn), &(grid.gnnsn), &(grid.gnnsn1), &(grid.node_fill_level));
  Find_Element_Centroids(&grid);
  Build_Edge_Data(&grid);
  Build_Boundary_Edge_Data(&grid);
  Fix_Subedge_Data(&grid);
  Generate_Gaussian_Data_For_Edge(&grid);
  cv_calc_area_Green(&grid);
  cv_calc_check_closure(&grid);
  cv_calc_length_scale(&grid);
  cv_calc_Moments(&grid);
  Build_CRS(&grid, p, &smem);
  for (i = 1; i <= grid.nn; i++) grid.node_order[i] = p.order;
  for (i = 1; i <= grid.nbedges; i++) {
    n = grid.bedges[i * 5 + 0];
    gn = grid.bedges[i * 5 + 1] + grid.nn;
    grid.node_order[gn] = grid.node_order[n];
  }
  init_pv(&grid, p);
  Compute_Edge_Weights(&grid);
  if (p.order != 3) {
    std::fprintf(stderr, "the case must have order 3 (all three reconstruction phases)\n");
    return 2;
  }

  const int nn = grid.nn, ng = grid.nn_ghost;
  const long ntot = (long)(nn + ng + 1) * 4;
  // a perturbed state (ghosts included) and arbitrary reconstruction fields
  if (!grid.nQ) grid.nQ = zeros(ntot);
  for (long k = 4; k < ntot; ++k) grid.nQ[k] = grid.Q[k] * (1.0 + 0.05 * (2.0 * urand() - 1.0));
  if (!grid.grad) grid.grad = zeros((long)(nn + 1) * 8);
  Compute_Gradient_LeastSquares(&grid);
  if (!grid.phi) grid.phi = zeros((long)(nn + 1) * 8);
  if (!grid.hess) grid.hess = zeros((long)(nn + 1) * 4 * NUM_MOM);
  if (!grid.point_value) grid.point_value = zeros((long)(nn + 1) * 4);
  for (n = 1; n <= nn; ++n) {
    grid.phi[n * 8 + 0] = urand();
    grid.phi[n * 8 + 1] = urand();
    for (int j = 0; j < 4; ++j) {
      grid.point_value[n * 4 + j] = grid.nQ[n * 4 + j] * (1.0 + 0.01 * (2.0 * urand() - 1.0));
      for (int m = 0; m < 5; ++m) grid.hess[n * 4 * NUM_MOM + j * NUM_MOM + m] = 0.1 * (2.0 * urand() - 1.0);
    }
  }
  const int phase_ts[3] = {p.foits, p.foits + 1, p.foits + p.soits + 1};

  if (jac) {
    // the reference's approximate Jacobian over every face, linear phase
    const int nnz = smem.ia[nn + 1];
    if (!smem.LHS) smem.LHS = zeros((long)nnz * 16);
    std::memset(smem.LHS, 0, (size_t)nnz * 16 * sizeof(double));
    const int ts = phase_ts[1];
    grid.citer = ts;
    for (int s = 1; s <= grid.nsubedges; ++s) {
      const int e = grid.subedges[s * 2];
      const int L = grid.edges[e * 2], R = grid.edges[e * 2 + 1];
      const double nx = grid.xn_subedges[s * 3], ny = grid.xn_subedges[s * 3 + 1], len = grid.xn_subedges[s * 3 + 2];
      double QL[4], QR[4], f[4];
      Reconstruct(s, L, R, QL, QR, &grid, p, ts);
      Roe_flux_centered(nx, ny, len, p.gamma, QL, QR, f);
      compute_approximate_jacobian_roe(&grid, &smem, p, L, R, QL, QR, f, nx, ny, len);
    }
    struct {
      const char* name;
      const double* a;
      long n;
    } outs[3] = {{"jac_nQ.bin", grid.nQ, ntot}, {"jac_phi.bin", grid.phi, (long)(nn + 1) * 8},
                 {"jac_lhs.bin", smem.LHS, (long)nnz * 16}};
    for (const auto& o : outs) {
      FILE* f = std::fopen(o.name, "wb");
      if (!f || std::fwrite(o.a, sizeof(double), (size_t)o.n, f) != (size_t)o.n) {
        std::fprintf(stderr, "cannot write %s\n", o.name);
        return 2;
      }
      std::fclose(f);
    }
    FILE* f = std::fopen("jac_meta.txt", "w");
    if (!f) return 2;
    std::fprintf(f, "%d %d %d %d\n", nn, ng, nnz, ts);
    std::fclose(f);
    std::printf("wrote the approximate Jacobian of %d faces\n", grid.nsubedges);
    return 0;
  }

  if (!device) {
    // ---- reconstruction at face points and Gauss points ----
    long bad = 0, tot = 0, badg = 0, totg = 0, badb = 0, totb = 0;
    double a[4], b[4], c[4], d[4];
    for (int ph = 0; ph < 3; ++ph) {
      const int ts = phase_ts[ph];
      grid.citer = ts;
      for (int s = 1; s <= grid.nsubedges; ++s) {
        const int e = grid.subedges[s * 2];
        const int L = grid.edges[e * 2], R = grid.edges[e * 2 + 1];
        Reconstruct(s, L, R, a, b, &grid, p, ts);
        fsref::Reconstruct(s, L, R, c, d, &grid, p, ts);
        bad += diff(a, c, 4) + diff(b, d, 4);
        tot += 8;
        for (int q = 0; q < 3; ++q) {
          double GP[2] = {grid.xg_subedges[s * 6 + 2 * q], grid.xg_subedges[s * 6 + 2 * q + 1]};
          Reconstruct_Gauss(s, L, R, GP, a, b, &grid, p, ts);
          fsref::Reconstruct_Gauss(s, L, R, GP, c, d, &grid, p, ts);
          badg += diff(a, c, 4) + diff(b, d, 4);
          totg += 8;
        }
      }
      for (int ib = 1; ib <= grid.nbedges; ++ib) {
        const int L = grid.bedges[ib * 5];
        for (int q = 0; q < 3; ++q) {
          double GP[2] = {grid.xg_bedges[ib * 6 + 2 * q], grid.xg_bedges[ib * 6 + 2 * q + 1]};
          Reconstruct_Gauss_Boundary(L, GP, a, &grid, p);
          fsref::Reconstruct_Gauss_Boundary(L, GP, c, &grid, p);
          badb += diff(a, c, 4);
          totb += 4;
        }
      }
    }
    report("Reconstruct (3 phases)", bad, tot);
    report("Reconstruct_Gauss (3 phases, 3 points)", badg, totg);
    report("Reconstruct_Gauss_Boundary (3 phases, 3 points)", badb, totb);

    // ---- ghost states at the node and at the three Gauss points ----
    long badv = 0, totv = 0;
    grid.citer = phase_ts[2];
    for (int ib = 1; ib <= grid.nbedges; ++ib) {
      const int L = grid.bedges[ib * 5];
      for (int q = 0; q < 4; ++q) {
        double QL[4];
        if (q == 0) {
          for (int j = 0; j < 4; ++j) QL[j] = grid.nQ[L * 4 + j];
        } else {
          double GP[2] = {grid.xg_bedges[ib * 6 + 2 * (q - 1)], grid.xg_bedges[ib * 6 + 2 * (q - 1) + 1]};
          Reconstruct_Gauss_Boundary(L, GP, QL, &grid, p);
        }
        Get_Boundary_Value(&grid, p, ib, QL, a);
        fsref::Get_Boundary_Value(&grid, p, ib, QL, c);
        badv += diff(a, c, 4);
        totv += 4;
      }
    }
    report("Get_Boundary_Value (node, 3 Gauss points)", badv, totv);

    // ---- Jacobians accumulated over every face into two copies of LHS ----
    const int nnz = smem.ia[nn + 1];
    SYS_MEM s2;
    s2.ia = (int*)malloc((size_t)(nn + 2) * sizeof(int));
    s2.iau = (int*)malloc((size_t)(nn + 1) * sizeof(int));
    s2.ja = (int*)malloc((size_t)nnz * sizeof(int));
    std::memcpy(s2.ia, smem.ia, (size_t)(nn + 2) * sizeof(int));
    std::memcpy(s2.iau, smem.iau, (size_t)(nn + 1) * sizeof(int));
    std::memcpy(s2.ja, smem.ja, (size_t)nnz * sizeof(int));
    if (!smem.LHS) smem.LHS = zeros((long)nnz * 16);
    s2.LHS = zeros((long)nnz * 16);
    const int ts = phase_ts[1];
    for (int kind = 0; kind < 4; ++kind) {
      std::memset(smem.LHS, 0, (size_t)nnz * 16 * sizeof(double));
      std::memset(s2.LHS, 0, (size_t)nnz * 16 * sizeof(double));
      if (kind < 2) {
        for (int s = 1; s <= grid.nsubedges; ++s) {
          const int e = grid.subedges[s * 2];
          const int L = grid.edges[e * 2], R = grid.edges[e * 2 + 1];
          const double nx = grid.xn_subedges[s * 3], ny = grid.xn_subedges[s * 3 + 1],
                       len = grid.xn_subedges[s * 3 + 2];
          double QL[4], QR[4], f[4];
          Reconstruct(s, L, R, QL, QR, &grid, p, ts);
          Roe_flux_centered(nx, ny, len, p.gamma, QL, QR, f);
          if (kind == 0) {
            compute_numerical_jacobian_roe(&grid, &smem, p, L, R, QL, QR, f, nx, ny, len);
            fsref::compute_numerical_jacobian_roe(&grid, &s2, p, L, R, QL, QR, f, nx, ny, len);
          } else {
            compute_approximate_jacobian_roe(&grid, &smem, p, L, R, QL, QR, f, nx, ny, len);
            fsref::compute_approximate_jacobian_roe(&grid, &s2, p, L, R, QL, QR, f, nx, ny, len);
          }
        }
      } else {
        for (int ib = 1; ib <= grid.nbedges; ++ib) {
          const int L = grid.bedges[ib * 5], G = grid.bedges[ib * 5 + 1] + nn;
          const double nx = grid.xn_bedges[ib * 3], ny = grid.xn_bedges[ib * 3 + 1], len = grid.xn_bedges[ib * 3 + 2];
          double QL[4], QR[4], f[4];
          for (int j = 0; j < 4; ++j) {
            QL[j] = grid.nQ[L * 4 + j];
            QR[j] = grid.nQ[G * 4 + j];
          }
          if (kind == 2) {
            Roe_flux_centered(nx, ny, len, p.gamma, QL, QR, f);
            compute_boundary_approximate_jacobian_roe(&grid, &smem, p, ib, QL, QR, f, nx, ny, len);
            fsref::compute_boundary_approximate_jacobian_roe(&grid, &s2, p, ib, QL, QR, f, nx, ny, len);
          } else {
            // the flux against the ghost the unperturbed state implies
            Get_Boundary_Value(&grid, p, ib, QL, QR);
            Roe_flux_centered(nx, ny, len, p.gamma, QL, QR, f);
            compute_boundary_numerical_jacobian_roe(&grid, &smem, p, ib, QL, f, nx, ny, len);
            fsref::compute_boundary_numerical_jacobian_roe(&grid, &s2, p, ib, QL, f, nx, ny, len);
          }
        }
      }
      long nz = 0;
      for (long k = 0; k < (long)nnz * 16; ++k) nz += smem.LHS[k] != 0.0;
      const char* names[4] = {"compute_numerical_jacobian_roe (all faces)", "compute_approximate_jacobian_roe (all faces)",
                              "compute_boundary_approximate_jacobian_roe", "compute_boundary_numerical_jacobian_roe"};
      report(names[kind], diff(smem.LHS, s2.LHS, (long)nnz * 16) + (nz == 0), (long)nnz * 16);
    }

    // ---- setup copies ----
    SYS_MEM s3;
    fsref::Build_CRS(&grid, p, &s3);
    long bc = 0;
    for (i = 0; i <= nn + 1; ++i) bc += s3.ia[i] != smem.ia[i];
    for (i = 0; i <= nn; ++i) bc += s3.iau[i] != smem.iau[i];
    for (i = 0; i < nnz; ++i) bc += s3.ja[i] != smem.ja[i];
    report("Build_CRS (ia, iau, ja)", bc, (long)(nn + 2) + (nn + 1) + nnz);
    double* ref_w = grid.edge_weight;
    grid.edge_weight = NULL;
    fsref::Compute_Edge_Weights(&grid);
    report("Compute_Edge_Weights", diff(ref_w + 4, grid.edge_weight + 4, (long)grid.nedges * 4), (long)grid.nedges * 4);
    free(grid.edge_weight);
    grid.edge_weight = ref_w;
  } else {
    // ---- device passes over the twin ----
    double* start = zeros(ntot);
    std::memcpy(start, grid.nQ, (size_t)ntot * sizeof(double));
    Set_Boundary_Values(&grid, p);
    double* ref = zeros(ntot);
    std::memcpy(ref, grid.nQ, (size_t)ntot * sizeof(double));
    std::memcpy(grid.nQ, start, (size_t)ntot * sizeof(double));
    fsref::Set_Boundary_Values(&grid, p);
    long changed = 0;
    for (long k = 4 * (long)(nn + 1); k < ntot; ++k) changed += ref[k] != start[k];
    report("Set_Boundary_Values (ghost states)", diff(ref + 4, grid.nQ + 4, ntot - 4) + (changed == 0), ntot - 4);
    // a second update from the updated ghosts: the memory term
    Set_Boundary_Values(&grid, p);
    std::memcpy(ref, grid.nQ, (size_t)ntot * sizeof(double));
    std::memcpy(grid.nQ, start, (size_t)ntot * sizeof(double));
    fsref::Set_Boundary_Values(&grid, p);
    fsref::Set_Boundary_Values(&grid, p);
    report("Set_Boundary_Values (twice, with memory)", diff(ref + 4, grid.nQ + 4, ntot - 4), ntot - 4);
    cv_calc_eigenvalue_contribution(&grid, p);
    double* ev = zeros(nn + 1);
    std::memcpy(ev, grid.ev_con, (size_t)(nn + 1) * sizeof(double));
    std::memset(grid.ev_con, 0, (size_t)(nn + 1) * sizeof(double));
    fsref::cv_calc_eigenvalue_contribution(&grid, p);
    report("cv_calc_eigenvalue_contribution", diff(ev + 1, grid.ev_con + 1, nn), nn);
    free(start);
    free(ref);
    free(ev);
    fsref::detach(&grid);
  }
  std::printf("%s\n", g_fail ? "FAILED" : "ALL BIT-EXACT");
  return g_fail;
}
