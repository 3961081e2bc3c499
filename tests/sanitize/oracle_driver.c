/* The FP64 oracle (oracle/lbm_oracle.c) built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitize_cpu.py): cavity, channel
 * and pipe lattices, SRT and MRT-LES, over several tile decompositions
 * including 1-node-thick boxes, a few steps each with every read-back.
 * Prints "oracle ok" when every lattice stayed finite. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "lbm_oracle.h"

static int run_case(int nx, int ny, int nz, int casetype, int mrt, int xd, int yd, int zd) {
  orc_params p;
  orc_default_params(&p);
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.casetype = casetype;
  p.enable_mrt = mrt;
  p.enable_les = mrt;
  p.smagorinsky = mrt ? 0.17 : 0.0;
  p.tau = 0.56;
  p.ref_vel[0] = 0.05;
  p.init_lid_velocity = 1;
  p.seed = 11;
  p.pipe_radius = ny / 3.0;
  p.pipe_cy = ny / 2.0;
  p.pipe_cz = nz / 2.0;
  int(*slice)[3][2] = malloc(sizeof(int[3][2]) * 64);
  const int nt = orc_decomposition(nx, ny, nz, xd, yd, zd, slice, 64);
  if (nt <= 0) return 1;
  orc_lattice* L = orc_create(&p, nt, (const int(*)[3][2])slice);
  orc_initialize(L);
  orc_run(L, 4);
  const size_t n = (size_t)nx * ny * nz, np = (size_t)(nx + 2) * (ny + 2) * (nz + 2) * ORC_Q;
  double* a = malloc(np * sizeof(double));
  double* m = malloc(4 * n * sizeof(double));
  int* t = malloc(2 * n * sizeof(int));
  orc_read_arF_padded(L, a);
  orc_write_arF_padded(L, a);
  orc_step(L);
  orc_read_macro(L, m, m + n, m + 2 * n, m + 3 * n);
  orc_read_node_type(L, t, t + n);
  double mass, mom[3], d;
  orc_diagnostics(L, &mass, mom, &d);
  int bad = !isfinite(mass) || !isfinite(d);
  for (size_t i = 0; i < 4 * n; ++i) bad |= !isfinite(m[i]);
  orc_destroy(L);
  free(a);
  free(m);
  free(t);
  free(slice);
  return bad;
}

int main(void) {
  const int boxes[][3] = {{9, 7, 5}, {12, 6, 5}, {3, 1, 4}, {1, 5, 1}, {6, 6, 6}};
  const int decs[][3] = {{1, 1, 1}, {2, 2, 1}, {2, 3, 2}, {3, 1, 2}};
  for (int b = 0; b < 5; ++b)
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c)
        for (int mrt = 0; mrt < 2; ++mrt) {
          if (c != ORC_CAVITY && boxes[b][0] < 3) continue; /* channel and pipe need nx >= 3 */
          if (run_case(boxes[b][0], boxes[b][1], boxes[b][2], c, mrt, decs[k][0], decs[k][1], decs[k][2])) {
            fprintf(stderr, "non-finite state: box %d dec %d case %d mrt %d\n", b, k, c, mrt);
            return 1;
          }
        }
  printf("oracle ok\n");
  return 0;
}
