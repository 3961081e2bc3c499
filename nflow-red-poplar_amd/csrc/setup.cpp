// setup.cpp -- host-side set-up calls of the C ABI: parameter defaults, the
// sub-block decomposition and the MRT matrices.  No device code.
#include <cmath>
#include <cstring>

#include "d3q19.hpp"
#include "errors.hpp"
#include "nflow.h"

using namespace nflow;

extern "C" void nflow_input_defaults(nflow_input* in) {
  // main.cpp:56-85: 10^3 cavity, Re = 1000, lid 0.1, tau = 3 U h / Re + 0.5,
  // MRT + LES with m_dSmagorinsky left at its value-initialised 0.
  std::memset(in, 0, sizeof(*in));
  in->h = 10;
  in->nx = in->ny = in->nz = in->h;
  in->maxitr = 199;
  in->nwrite = 10000;
  in->casetype = NFLOW_CAVITY;
  in->enable_mrt = 1;
  in->enable_les = 1;
  in->storage = NFLOW_STORAGE_F32;
  in->init_lid_velocity = 1;
  in->ref_den = 1.0;
  in->ref_vel[0] = 0.1;
  in->tau = 3.0 * 0.1 * in->h / 1000.0 + 0.5;
  in->smagorinsky = 0.0;
  in->w_eps = 3.0;
  in->w_epsj = -5.5;
  in->w_xx = -0.5;
  in->inlet_peak = 0.05;
  in->perturb = 0.01;
  in->seed = 0;
  in->pipe_radius = 20.0;  // kernel.cu:1090: (y-25)^2 + (z-25)^2 > 400
  in->pipe_cy = 25.0;
  in->pipe_cz = 25.0;
  in->stochastic_rounding = 1;  // 16-bit storage rounds stochastically; FP32 ignores it
}

// decomposition.cpp:98-161: `dec` equal blocks of n/dec nodes per axis, the
// last block taking the remainder; tiles enumerated x outer, y, then z inner.
static void axis_block(int i, int dec, int n, int* st, int* ed) {
  const int chunk = n / dec;
  int e = (i + 1) * chunk;
  if (i == dec - 1 || e > n) e = n;
  *st = i * chunk;
  *ed = e;
}

extern "C" int nflow_decomposition(const nflow_input* in, int xdecomp, int ydecomp, int zdecomp, int slice[][3][2],
                                   int max_tiles) {
  if (!in || !slice) return fail(NFLOW_EINVAL, "nflow_decomposition: null argument");
  if (xdecomp < 1 || ydecomp < 1 || zdecomp < 1)
    return fail(NFLOW_EINVAL, "nflow_decomposition: decomposition counts must be >= 1");
  if ((long long)xdecomp * ydecomp * zdecomp > max_tiles)
    return fail(NFLOW_EINVAL, "nflow_decomposition: %d tiles exceed max_tiles = %d", xdecomp * ydecomp * zdecomp,
                max_tiles);
  int t = 0;
  for (int i = 0; i < xdecomp; ++i)
    for (int j = 0; j < ydecomp; ++j)
      for (int k = 0; k < zdecomp; ++k, ++t) {
        axis_block(i, xdecomp, in->nx, &slice[t][0][0], &slice[t][0][1]);
        axis_block(j, ydecomp, in->ny, &slice[t][1][0], &slice[t][1][1]);
        axis_block(k, zdecomp, in->nz, &slice[t][2][0], &slice[t][2][1]);
      }
  return t;
}

// slice_chunk, decomposition.cpp:66-94
extern "C" int nflow_slice_chunk(unsigned tile, const int slice[][3][2], int* xs, int* xe, int* ys, int* ye, int* zs,
                                 int* ze, int* chunk_nx, int* chunk_ny, int* chunk_nz, int* chunk_len) {
  *xs = slice[tile][0][0];
  *xe = slice[tile][0][1];
  *ys = slice[tile][1][0];
  *ye = slice[tile][1][1];
  *zs = slice[tile][2][0];
  *ze = slice[tile][2][1];
  *chunk_nx = *xe - *xs;
  *chunk_ny = *ye - *ys;
  *chunk_nz = *ze - *zs;
  *chunk_len = *chunk_nx * *chunk_ny * *chunk_nz;
  return (*chunk_nx < 1 || *chunk_ny < 1 || *chunk_nz < 1) ? 1 : 0;
}

// set_move_decomp, decomposition.cpp:17-62: offset of node x + e_q inside the
// tile's ghost-padded sub-block, relative to the node's unpadded local index
// (index_ghostcell, decomposition.cpp:10-15).
extern "C" int nflow_set_move_decomp(int ntile, const int slice[][3][2], int move[][NFLOW_DIRECTION_SIZE]) {
  if (ntile < 0 || (ntile > 0 && (!slice || !move))) return fail(NFLOW_EINVAL, "nflow_set_move_decomp: bad argument");
  for (int t = 0; t < ntile; ++t) {
    const int pnx = slice[t][0][1] - slice[t][0][0] + 2;
    const int pnz = slice[t][2][1] - slice[t][2][0] + 2;
    for (int q = 0; q < d3q19::Q; ++q)
      move[t][q] = (d3q19::EX[q] + 1) + (d3q19::EY[q] + 1) * pnx * pnz + (d3q19::EZ[q] + 1) * pnx;
  }
  return 0;
}

extern "C" void nflow_mrt_matrix(float m[NFLOW_DIRECTION_SIZE][NFLOW_DIRECTION_SIZE]) {
  for (int k = 0; k < d3q19::Q; ++k)
    for (int q = 0; q < d3q19::Q; ++q) m[k][q] = (float)d3q19::mrow(k, q);
}

// Gauss-Jordan elimination with partial pivoting, in double.  The reference's
// inverseMRT (inverseMRT.cpp:8-92) eliminates in float with a row rotation on
// near-zero pivots and exits on a singular matrix; here a singular matrix is
// an error code.
extern "C" int nflow_inverse_mrt(const float m[NFLOW_DIRECTION_SIZE][NFLOW_DIRECTION_SIZE],
                                 float minv[NFLOW_DIRECTION_SIZE][NFLOW_DIRECTION_SIZE]) {
  constexpr int N = NFLOW_DIRECTION_SIZE;
  if (!m || !minv) return fail(NFLOW_EINVAL, "nflow_inverse_mrt: null argument");
  double a[N][2 * N];
  double scale = 0.0;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      a[i][j] = m[i][j];
      a[i][N + j] = (i == j) ? 1.0 : 0.0;
      scale = std::fmax(scale, std::fabs((double)m[i][j]));
    }
  if (scale == 0.0) return fail(NFLOW_EINVAL, "nflow_inverse_mrt: singular matrix");
  for (int c = 0; c < N; ++c) {
    int piv = c;
    for (int r = c + 1; r < N; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
    if (std::fabs(a[piv][c]) <= 1e-12 * scale) return fail(NFLOW_EINVAL, "nflow_inverse_mrt: singular matrix");
    if (piv != c)
      for (int j = 0; j < 2 * N; ++j) std::swap(a[c][j], a[piv][j]);
    const double inv = 1.0 / a[c][c];
    for (int j = 0; j < 2 * N; ++j) a[c][j] *= inv;
    for (int r = 0; r < N; ++r) {
      if (r == c || a[r][c] == 0.0) continue;
      const double fct = a[r][c];
      for (int j = 0; j < 2 * N; ++j) a[r][j] -= fct * a[c][j];
    }
  }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) minv[i][j] = (float)a[i][N + j];
  return 0;
}
