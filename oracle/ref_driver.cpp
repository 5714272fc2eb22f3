// Test-infrastructure harness around the reference's own translation units.
//
// This program is compiled against the reference sources in place (see
// oracle/Makefile) and is used only to write golden vectors and as an
// optional CPU baseline.  It replays the reference driver's setup sequence
// (main.C:105-648) and pseudo-time loop (main.C:789-876) and dumps every
// array the parity tests compare.  It never enters the product.
//
// Usage: ref_driver <case_dir> <out.bin> <nsteps> [stencil] [dump_every]
//   case_dir holds euler.params, bc.dat and grid.mesh (the reference reads
//   the first two from the working directory, params.C:34, mesh_io.C:282).
//   stencil=1 also builds the order-3 generic stencils the way solve.C:57-114
//   does and dumps them, without entering the k=2 phase.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <unistd.h>

#include "Defined_Variables.h"
#include "grid.h"
#include "params.h"
#include "sys_mem.h"
#include "cv_calc.h"
#include "geometry.h"
#include "gradient.h"
#include "grid_io.h"
#include "grid.h"
#include "ic.h"
#include "init_memory.h"
#include "maps.h"
#include "mesh_io.h"
#include "metrics.h"
#include "params.h"
#include "reorder.h"
#include "solve.h"
#include "sys_mem.h"

static FILE* g_out = NULL;

static void dump(const char* name, int type, const void* data, long count) {
  char nm[32];
  memset(nm, 0, sizeof nm);
  strncpy(nm, name, 31);
  fwrite(nm, 1, 32, g_out);
  int t = type;
  fwrite(&t, sizeof(int), 1, g_out);
  long long c = count;
  fwrite(&c, sizeof(long long), 1, g_out);
  fwrite(data, type == 0 ? sizeof(int) : sizeof(double), count, g_out);
}
static void dumpi(const char* n, const int* d, long c) { dump(n, 0, d, c); }
static void dumpd(const char* n, const double* d, long c) { dump(n, 1, d, c); }

static double now_s() {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: ref_driver <case_dir> <out.bin> <nsteps> [stencil] [dump_every]\n");
    return 2;
  }
  const char* outpath = argv[2];
  int nsteps = atoi(argv[3]);
  int stencil = argc > 4 ? atoi(argv[4]) : 0;
  int dump_every = argc > 5 ? atoi(argv[5]) : 1;
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return 2;
  char abs_out[8192];
  if (outpath[0] == '/') snprintf(abs_out, sizeof abs_out, "%s", outpath);
  else snprintf(abs_out, sizeof abs_out, "%s/%s", cwd, outpath);
  if (chdir(argv[1]) != 0) { fprintf(stderr, "cannot chdir to %s\n", argv[1]); return 2; }
  g_out = fopen(abs_out, "wb");
  if (!g_out) { fprintf(stderr, "cannot open %s\n", abs_out); return 2; }

  GRID grid;
  SYS_MEM smem;
  PARAMS p;
  int i, n, gn;
  char meshname[] = "grid.mesh";

  double t0 = now_s();
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
  double t_setup = now_s() - t0;

  int nn = grid.nn, ng = grid.nn_ghost;
  int hdr[8] = {nn, ng, grid.nedges, grid.nsubedges, grid.nbedges, grid.num_elem[Tri], grid.num_elem[Quad], grid.nb};
  dumpi("hdr", hdr, 8);
  if (grid.nn_map) dumpi("nn_map", grid.nn_map, nn + 1);
  dumpd("x", grid.x, (long)(nn + ng + 1) * 2);
  dumpi("c2n_tri", grid.c2n[Tri], (long)(grid.num_elem[Tri] + 1) * 3);
  if (grid.num_elem[Quad] > 0) dumpi("c2n_quad", grid.c2n[Quad], (long)(grid.num_elem[Quad] + 1) * 4);
  dumpi("ncsn", grid.ncsn, nn + 2);
  dumpi("csn", grid.csn, grid.ncsn[nn + 1]);
  dumpi("nnsn", grid.nnsn, nn + 2);
  dumpi("nsn", grid.nsn, grid.nnsn[nn + 1]);
  dumpi("esn", grid.esn, grid.nnsn[nn + 1]);
  dumpi("edges", grid.edges, (long)(grid.nedges + 1) * 2);
  dumpi("c2e_tri", grid.c2e[Tri], (long)(grid.num_elem[Tri] + 1) * 3);
  if (grid.num_elem[Quad] > 0) dumpi("c2e_quad", grid.c2e[Quad], (long)(grid.num_elem[Quad] + 1) * 4);
  dumpi("subedges", grid.subedges, (long)(grid.nsubedges + 1) * 2);
  dumpd("xn_subedges", grid.xn_subedges, (long)(grid.nsubedges + 1) * 3);
  dumpd("xm_subedges", grid.xm_subedges, (long)(grid.nsubedges + 1) * 2);
  dumpd("xg_subedges", grid.xg_subedges, (long)(grid.nsubedges + 1) * 6);
  int nel = grid.num_elem[Tri] + grid.num_elem[Quad];
  dumpd("el_cent", grid.el_cent, (long)(nel + 1) * 2);
  dumpi("bedges", grid.bedges, (long)(grid.nbedges + 1) * 5);
  dumpd("xn_bedges", grid.xn_bedges, (long)(grid.nbedges + 1) * 3);
  dumpd("xg_bedges", grid.xg_bedges, (long)(grid.nbedges + 1) * 6);
  dumpd("cv_area", grid.cv_area, nn + 1);
  dumpd("Moments", grid.Moments, (long)(nn + 1) * NUM_MOM);
  dumpd("edge_weight", grid.edge_weight, (long)(grid.nedges + 1) * 4);
  dumpi("ia", smem.ia, nn + 2);
  dumpi("iau", smem.iau, nn + 1);
  dumpi("ja", smem.ja, smem.ia[nn + 1]);
  dumpd("Q0", grid.Q, (long)(nn + ng + 1) * 4);

  if (stencil) {
    // the order-3 stencil build of solve.C:57-114
    for (i = 1; i <= grid.nn; i++)
      Add_Layer_to_GNSN(grid.nn, grid.num_elem, grid.c2n, grid.esn, grid.nnsn, grid.edges, &(grid.gnsn), &(grid.gnnsn),
                        grid.gnnsn1, grid.node_fill_level, i);
    for (i = 1; i <= grid.nn; i++)
      if (grid.node_order[i] > 2)
        Add_Layer_to_GNSN(grid.nn, grid.num_elem, grid.c2n, grid.esn, grid.nnsn, grid.edges, &(grid.gnsn),
                          &(grid.gnnsn), grid.gnnsn1, grid.node_fill_level, i);
    for (i = 1; i <= grid.nn; i++) {
      if (grid.node_order[i] < 3) continue;
      int j = grid.gnnsn[grid.gnnsn1[i + 1]] - grid.gnnsn[(grid.gnnsn1[i] + 1)];
      for (int k = 0; j < 6 && k < 3; k++) {
        Add_Layer_to_GNSN(grid.nn, grid.num_elem, grid.c2n, grid.esn, grid.nnsn, grid.edges, &(grid.gnsn),
                          &(grid.gnnsn), grid.gnnsn1, grid.node_fill_level, i);
        j = grid.gnnsn[grid.gnnsn1[i + 1]] - grid.gnnsn[(grid.gnnsn1[i] + 1)];
      }
    }
    dumpi("gnnsn1", grid.gnnsn1, nn + 2);
    dumpi("gnnsn", grid.gnnsn, grid.gnnsn1[nn + 1] + 1);
    dumpi("gnsn", grid.gnsn, grid.gnnsn[grid.gnnsn1[nn + 1]]);
  }

  double t_solve = 0.0;
  char nm[64];
  for (i = grid.citer + 1; i <= nsteps; i++) {
    grid.eltime = i * p.mintime;
    grid.citer = i;
    grid.iter = i;
    double ts0 = now_s();
    double rms = solve(&grid, &smem, p, i);
    t_solve += now_s() - ts0;
    // main.C:874-876: body forces from the state the step left
    double forces[2] = {0., 0.};
    Compute_Body_Forces(&grid, p, forces);
    if (dump_every > 0 && (i % dump_every == 0 || i == nsteps)) {
      snprintf(nm, sizeof nm, "rms_%d", i);      dumpd(nm, &rms, 1);
      snprintf(nm, sizeof nm, "forces_%d", i);   dumpd(nm, forces, 2);
      snprintf(nm, sizeof nm, "R_%d", i);        dumpd(nm, grid.R, (long)(nn + 1) * 4);
      snprintf(nm, sizeof nm, "dQ_%d", i);       dumpd(nm, grid.dQ, (long)(nn + 1) * 4);
      snprintf(nm, sizeof nm, "Q_%d", i);        dumpd(nm, grid.Q, (long)(nn + ng + 1) * 4);
      snprintf(nm, sizeof nm, "LHS_%d", i);      dumpd(nm, smem.LHS, (long)smem.ia[nn + 1] * 16);
      snprintf(nm, sizeof nm, "phi_%d", i);      dumpd(nm, grid.phi, (long)(nn + 1) * 8);
      snprintf(nm, sizeof nm, "grad_%d", i);     dumpd(nm, grid.grad, (long)(nn + 1) * 8);
      snprintf(nm, sizeof nm, "ev_con_%d", i);   dumpd(nm, grid.ev_con, nn + 1);
      snprintf(nm, sizeof nm, "dt_%d", i);       dumpd(nm, grid.dt, nn + 1);
    }
  }
  double times[2] = {t_setup, t_solve};
  dumpd("times", times, 2);
  dumpd("Qfinal", grid.Q, (long)(nn + ng + 1) * 4);
  fclose(g_out);
  return 0;
}
