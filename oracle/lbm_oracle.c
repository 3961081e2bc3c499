/*
 * lbm_oracle.c -- FP64 CPU restatement of the reference D3Q19 LBM step.
 *
 * TEST INFRASTRUCTURE ONLY (the correctness oracle and the timed CPU
 * baseline).  Nothing in the product library calls, links or executes this
 * file.  Every function cites the reference file:line it restates; the maths
 * is the reference's, evaluated in double precision with the same operation
 * order where the reference fixes one (dense 19x19 mat-vecs for MRT, the
 * per-direction sums for LES and macroscopic quantities).
 *
 * Pins (tests/test_oracle_*.py):
 *   tables, moment matrix   closed forms + the reference's literal matrix
 *   decomposition, Move     the reference decomposition.cpp built under oracle/_ref
 *   MRT / SRT / LES cells   MRT == SRT when all rates are 1/tau and RefDen = rho,
 *                           conservation, rest-state fixed point, Cs = 0, brute force
 *   streaming + walls       pure-streaming permutation, mass conservation,
 *                           hand-computed moving-wall link, decomposition invariance
 *   channel inlet/outlet    analytic duct Poiseuille profile, hand-computed links
 */
#include "lbm_oracle.h"

#include <math.h>

/* Working precision of the lattice state and arithmetic: double for the
 * oracle; oracle/Makefile also builds liboracle_f32.so with ORC_REAL=float,
 * the same algorithm in single precision, which measures how far plain FP32
 * arithmetic drifts from FP64 (the yardstick for the GPU tolerances). */
#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL real;
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------------
 * Tables.  main.cpp:87-113 (velocity set and opposite map, [XDIM/YDIM/ZDIM][q]),
 * main.cpp:141-146 (weights).
 * ------------------------------------------------------------------------ */
static const int EX[ORC_Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[ORC_Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int EZ[ORC_Q] = {0, 0, 0, -1, 1, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, 1, -1, 1};
static const int OPP[ORC_Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

static double g_dir[3][ORC_Q];
static double g_w[ORC_Q];
static double g_M[ORC_Q][ORC_Q];
static double g_Minv[ORC_Q][ORC_Q];

void orc_tables(double dir[3][ORC_Q], int opp[ORC_Q], double w[ORC_Q]) {
  for (int q = 0; q < ORC_Q; ++q) {
    dir[ORC_XDIM][q] = EX[q];
    dir[ORC_YDIM][q] = EY[q];
    dir[ORC_ZDIM][q] = EZ[q];
    opp[q] = OPP[q];
    w[q] = (q == 0) ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
}

/* The d'Humieres D3Q19 moment basis for this velocity order, written as its
 * polynomial rows (equal entry-by-entry to the literal of main.cpp:152-170). */
void orc_mrt_matrix(double m[ORC_Q][ORC_Q]) {
  for (int q = 0; q < ORC_Q; ++q) {
    const double x = EX[q], y = EY[q], z = EZ[q];
    const double e2 = x * x + y * y + z * z;
    m[0][q] = 1.0;
    m[1][q] = 19.0 * e2 - 30.0;
    m[2][q] = (21.0 * e2 * e2 - 53.0 * e2 + 24.0) / 2.0;
    m[3][q] = x;
    m[4][q] = (5.0 * e2 - 9.0) * x;
    m[5][q] = y;
    m[6][q] = (5.0 * e2 - 9.0) * y;
    m[7][q] = z;
    m[8][q] = (5.0 * e2 - 9.0) * z;
    m[9][q] = 3.0 * x * x - e2;
    m[10][q] = (3.0 * e2 - 5.0) * (3.0 * x * x - e2);
    m[11][q] = y * y - z * z;
    m[12][q] = (3.0 * e2 - 5.0) * (y * y - z * z);
    m[13][q] = x * y;
    m[14][q] = y * z;
    m[15][q] = x * z;
    m[16][q] = x * (y * y - z * z);
    m[17][q] = y * (z * z - x * x);
    m[18][q] = z * (x * x - y * y);
  }
}

/* The rows of M are mutually orthogonal, so M^-1 = M^T diag(1/|row_k|^2).
 * Replaces the FP32 Gauss-Jordan of inverseMRT.cpp:8-92 with the exact inverse. */
int orc_mrt_inverse_exact(const double m[ORC_Q][ORC_Q], double minv[ORC_Q][ORC_Q]) {
  double norm2[ORC_Q];
  for (int k = 0; k < ORC_Q; ++k) {
    for (int l = 0; l < ORC_Q; ++l) {
      double s = 0.0;
      for (int q = 0; q < ORC_Q; ++q) s += m[k][q] * m[l][q];
      if (k == l)
        norm2[k] = s;
      else if (fabs(s) > 1e-12)
        return -1;
    }
    if (norm2[k] <= 0.0) return -1;
  }
  for (int q = 0; q < ORC_Q; ++q)
    for (int k = 0; k < ORC_Q; ++k) minv[q][k] = m[k][q] / norm2[k];
  return 0;
}

static void orc_init_tables(void) {
  int opp[ORC_Q];
  orc_tables(g_dir, opp, g_w);
  orc_mrt_matrix(g_M);
  orc_mrt_inverse_exact(g_M, g_Minv);
}

/* main.cpp:59-85 defaults (the shipped 10^3 run), with s_rates of
 * lbm_codelets.cpp:488-506. */
void orc_default_params(orc_params* p) {
  memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = 10;
  p->casetype = ORC_CAVITY;
  p->enable_mrt = 1;
  p->enable_les = 1;
  p->init_lid_velocity = 1;
  p->ref_den = 1.0;
  p->ref_vel[0] = 0.1;
  p->tau = 3.0 * 0.1 * 10 / 1000.0 + 0.5;
  p->smagorinsky = 0.0; /* never assigned in main.cpp, value-initialised to 0 */
  p->w_eps = 3.0;
  p->w_epsj = -5.5;
  p->w_xx = -0.5;
  p->inlet_peak = 0.05;
  p->perturb = 0.01;
  p->seed = 0;
  p->pipe_radius = 20.0; /* kernel.cu:1090: (y-25)^2 + (z-25)^2 > 400 */
  p->pipe_cy = 25.0;
  p->pipe_cz = 25.0;
  static const double S[ORC_Q] = {0.0, 1.19, 1.4, 0.0,  1.2, 0.0,  1.2, 0.0,  1.2, -1.0,
                                  1.4, -1.0, 1.4, -1.0, -1.0, -1.0, 1.98, 1.98, 1.98};
  memcpy(p->s_rates, S, sizeof(S));
}

/* ---------------------------------------------------------------------------
 * Decomposition.  decomposition.cpp:98-161 (decomposition), 66-94 (slice_chunk),
 * 10-62 (index_ghostcell, set_move_decomp).
 * ------------------------------------------------------------------------ */
static void block_range(int i, int ndec, int n, int* st, int* ed) {
  const int chunk = n / ndec;
  int start = i * chunk, end = (i + 1) * chunk;
  if (i == ndec - 1) end = n;
  if (end > n) end = n;
  *st = start;
  *ed = end;
}

int orc_decomposition(int nx, int ny, int nz, int xdec, int ydec, int zdec, int slice[][3][2], int max_tiles) {
  if (xdec < 1 || ydec < 1 || zdec < 1) return -1;
  if (xdec * ydec * zdec > max_tiles) return -1;
  int tiles = 0;
  for (int i = 0; i < xdec; ++i) /* x outer, then y, then z (decomposition.cpp:111-146) */
    for (int j = 0; j < ydec; ++j)
      for (int k = 0; k < zdec; ++k) {
        block_range(i, xdec, nx, &slice[tiles][0][0], &slice[tiles][0][1]);
        block_range(j, ydec, ny, &slice[tiles][1][0], &slice[tiles][1][1]);
        block_range(k, zdec, nz, &slice[tiles][2][0], &slice[tiles][2][1]);
        ++tiles;
      }
  return tiles;
}

int orc_slice_chunk(int tile, const int slice[][3][2], int out[10]) {
  const int xs = slice[tile][0][0], xe = slice[tile][0][1];
  const int ys = slice[tile][1][0], ye = slice[tile][1][1];
  const int zs = slice[tile][2][0], ze = slice[tile][2][1];
  const int dx = xe - xs, dy = ye - ys, dz = ze - zs;
  out[0] = xs; out[1] = xe; out[2] = ys; out[3] = ye; out[4] = zs; out[5] = ze;
  out[6] = dx; out[7] = dy; out[8] = dz; out[9] = dx * dy * dz;
  return (dx < 1 || dy < 1 || dz < 1) ? 1 : 0;
}

int orc_set_move_decomp(int ntile, const int slice[][3][2], int move[][ORC_Q]) {
  for (int t = 0; t < ntile; ++t) {
    const int nx2 = (slice[t][0][1] - slice[t][0][0]) + 2;
    const int nz2 = (slice[t][2][1] - slice[t][2][0]) + 2;
    const int nslave2 = nx2 * nz2;
    for (int q = 0; q < ORC_Q; ++q) /* index_ghostcell(nx, nSlave, ex, ey, ez) */
      move[t][q] = (EX[q] + 1) + (EY[q] + 1) * nslave2 + (EZ[q] + 1) * nx2;
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * Channel inlet profile (no reference counterpart: the reference inlet is the
 * uniform RefVel of kernel.cu:601-605; the duct inlet is parabolic with a
 * seeded, counter-based perturbation).
 * ------------------------------------------------------------------------ */
unsigned long long orc_splitmix64(unsigned long long x) {
  uint64_t z = (uint64_t)x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double orc_inlet_xi(unsigned long long seed, int y, int z) {
  const uint64_t key = (uint64_t)seed ^ (((uint64_t)(uint32_t)y << 32) | (uint64_t)(uint32_t)z);
  const uint64_t h = orc_splitmix64(key);
  return 2.0 * ((double)(h >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

void orc_inlet_profile(const orc_params* p, double* uin) {
  for (int y = 0; y < p->ny; ++y)
    for (int z = 0; z < p->nz; ++z) {
      const double eta = (y + 0.5) / p->ny, zeta = (z + 0.5) / p->nz;
      const double base = p->inlet_peak * 16.0 * eta * (1.0 - eta) * zeta * (1.0 - zeta);
      uin[(size_t)y * p->nz + z] = base * (1.0 + p->perturb * orc_inlet_xi(p->seed, y, z));
    }
}

/* ---------------------------------------------------------------------------
 * Single-cell maths, in the working precision `real`.
 * ------------------------------------------------------------------------ */
static real R_DIR[3][ORC_Q], R_W[ORC_Q], R_M[ORC_Q][ORC_Q], R_MINV[ORC_Q][ORC_Q];
/* the same two matrices stored column by column, [k][i] (padded to 20 rows),
 * so that the mat-vecs below can run over i in vector registers */
#define ORC_QP 20
static real R_MT[ORC_Q][ORC_QP], R_MINVT[ORC_Q][ORC_QP];

__attribute__((constructor)) static void orc_init_real_tables(void) {
  orc_init_tables();
  for (int q = 0; q < ORC_Q; ++q) {
    for (int a = 0; a < 3; ++a) R_DIR[a][q] = (real)g_dir[a][q];
    R_W[q] = (real)g_w[q];
    for (int k = 0; k < ORC_Q; ++k) {
      R_M[q][k] = (real)g_M[q][k];
      R_MINV[q][k] = (real)g_Minv[q][k];
      R_MT[k][q] = R_M[q][k];
      R_MINVT[k][q] = R_MINV[q][k];
    }
  }
}

/* feq, lbm_codelets.cpp:222-231 */
static void cell_feq(real rho, const real u[3], real feq[ORC_Q]) {
  const real usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  /* e_q from the integer tables (the values of R_DIR), unrolled so that the
   * compiler sees the 0 / +-1 factors */
#pragma GCC unroll 19
  for (int q = 0; q < ORC_Q; ++q) {
    const real eu = (real)EX[q] * u[0] + (real)EY[q] * u[1] + (real)EZ[q] * u[2];
    feq[q] = R_W[q] * rho * ((real)1 + (real)3 * eu + (real)4.5 * eu * eu - (real)1.5 * usq);
  }
}

/* ComputeTurbulence_LES::compute body, lbm_codelets.cpp:597-662 */
static real cell_tau_eddy(const orc_params* p, const real f[ORC_Q], real rho, const real u[3]) {
  real feq[ORC_Q];
  cell_feq(rho, u, feq);
  real phi[3][3] = {{0}};
#pragma GCC unroll 19
  for (int q = 0; q < ORC_Q; ++q) {
    const real e[3] = {(real)EX[q], (real)EY[q], (real)EZ[q]}; /* = R_DIR[XDIM/YDIM/ZDIM][q] */
    const real mul = f[q] - feq[q];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) phi[a][b] += e[a] * (e[b] * mul);
  }
  real qles = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) qles += phi[a][b] * phi[a][b];
  const real tau = (real)p->tau, cs = (real)p->smagorinsky;
  const real smag = (real)0.5 * sqrt(tau * tau + (real)18.0 * (real)1.41421356237 * cs * cs * sqrt(qles) / rho);
  return (real)-0.5 * tau + smag;
}

/* Collision_MRT::compute body, lbm_codelets.cpp:449-537 (dense M and M^-1) */
static void cell_mrt(const orc_params* p, const real f[ORC_Q], real rho, const real u[3], real tau_eddy,
                     real fout[ORC_Q]) {
  const real j[3] = {rho * u[0], rho * u[1], rho * u[2]};
  const real jj = j[0] * j[0] + j[1] * j[1] + j[2] * j[2];
  const real r0 = (real)p->ref_den;
  real meq[ORC_Q];
  meq[0] = rho;
  meq[1] = (real)-11.0 * rho + (real)19.0 / r0 * jj;
  meq[2] = (real)p->w_eps * rho + (real)p->w_epsj / r0 * jj;
  meq[3] = j[0];
  meq[4] = (real)(-2.0 / 3.0) * j[0];
  meq[5] = j[1];
  meq[6] = (real)(-2.0 / 3.0) * j[1];
  meq[7] = j[2];
  meq[8] = (real)(-2.0 / 3.0) * j[2];
  meq[9] = (real)1.0 / r0 * ((real)2.0 * j[0] * j[0] - (j[1] * j[1] + j[2] * j[2]));
  meq[10] = (real)p->w_xx * meq[9];
  meq[11] = (real)1.0 / r0 * (j[1] * j[1] - j[2] * j[2]);
  meq[12] = (real)p->w_xx * meq[11];
  meq[13] = j[0] * j[1] / r0;
  meq[14] = j[1] * j[2] / r0;
  meq[15] = j[2] * j[0] / r0;
  meq[16] = meq[17] = meq[18] = 0;
  const real snu = (real)1.0 / ((real)p->tau + tau_eddy);
  /* m1 = M f and s = M^-1 tmp as dense 19 x 19 products; each row sums its
   * 19 terms in the order k = 0..18, as the row loops of the reference do,
   * with the loop over rows innermost (the same arithmetic, vectorisable) */
  real m1[ORC_QP] = {0}, tmp[ORC_Q], s[ORC_QP] = {0};
  for (int k = 0; k < ORC_Q; ++k)
    for (int i = 0; i < ORC_QP; ++i) m1[i] += R_MT[k][i] * f[k];
  for (int i = 0; i < ORC_Q; ++i) {
    const real s_i = p->s_rates[i] < 0.0 ? snu : (real)p->s_rates[i];
    tmp[i] = s_i * (m1[i] - meq[i]);
  }
  for (int k = 0; k < ORC_Q; ++k)
    for (int i = 0; i < ORC_QP; ++i) s[i] += R_MINVT[k][i] * tmp[k];
  for (int i = 0; i < ORC_Q; ++i) fout[i] = f[i] - s[i];
}

/* keCollision_SRT body, kernel.cu:466-486 */
static void cell_srt(const orc_params* p, const real f[ORC_Q], real rho, const real u[3], real tau_eddy,
                     real fout[ORC_Q]) {
  real feq[ORC_Q];
  cell_feq(rho, u, feq);
  for (int q = 0; q < ORC_Q; ++q) fout[q] = f[q] - (f[q] - feq[q]) / ((real)p->tau + tau_eddy);
}

/* double-precision entry points of the cell functions (known-answer tests) */
void orc_cell_feq(const double dir[3][ORC_Q], const double w[ORC_Q], double rho, const double u[3], double feq[ORC_Q]) {
  const double usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int q = 0; q < ORC_Q; ++q) {
    const double eu = dir[ORC_XDIM][q] * u[0] + dir[ORC_YDIM][q] * u[1] + dir[ORC_ZDIM][q] * u[2];
    feq[q] = w[q] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
  }
}

static void to_real(const double* a, real* b, int n) {
  for (int i = 0; i < n; ++i) b[i] = (real)a[i];
}

double orc_cell_tau_eddy(const orc_params* p, const double f[ORC_Q], double rho, const double u[3]) {
  real fr[ORC_Q], ur[3];
  to_real(f, fr, ORC_Q);
  to_real(u, ur, 3);
  return (double)cell_tau_eddy(p, fr, (real)rho, ur);
}

void orc_cell_mrt(const orc_params* p, const double f[ORC_Q], double rho, const double u[3], double tau_eddy,
                  double fout[ORC_Q]) {
  real fr[ORC_Q], ur[3], o[ORC_Q];
  to_real(f, fr, ORC_Q);
  to_real(u, ur, 3);
  cell_mrt(p, fr, (real)rho, ur, (real)tau_eddy, o);
  for (int q = 0; q < ORC_Q; ++q) fout[q] = (double)o[q];
}

void orc_cell_srt(const orc_params* p, const double f[ORC_Q], double rho, const double u[3], double tau_eddy,
                  double fout[ORC_Q]) {
  real fr[ORC_Q], ur[3], o[ORC_Q];
  to_real(f, fr, ORC_Q);
  to_real(u, ur, 3);
  cell_srt(p, fr, (real)rho, ur, (real)tau_eddy, o);
  for (int q = 0; q < ORC_Q; ++q) fout[q] = (double)o[q];
}

/* ---------------------------------------------------------------------------
 * Lattice.
 * ------------------------------------------------------------------------ */
struct orc_lattice {
  orc_params p;
  int ntile;
  int (*slice)[3][2];
  int (*move)[ORC_Q];
  size_t n;      /* nx*ny*nz */
  size_t npad;   /* (nx+2)(ny+2)(nz+2) */
  real* arF;     /* [ny][nz][nx][19] */
  real* arFTemp; /* [ny+2][nz+2][nx+2][19] */
  real *rho, *prev_rho, *ux, *uy, *uz, *tau_eddy;
  int *type, *dirstate;
  real* uin; /* [ny][nz] */
  unsigned char* mask; /* [ny][nz][nx], non-zero = solid; NULL: none */
  long steps;
};

static inline size_t nid(const orc_lattice* L, int x, int y, int z) {
  return (size_t)x + (size_t)z * L->p.nx + (size_t)y * L->p.nx * L->p.nz;
}
static inline size_t pid(const orc_lattice* L, int X, int Y, int Z) { /* padded coords */
  return (size_t)X + (size_t)Z * (L->p.nx + 2) + (size_t)Y * (L->p.nx + 2) * (L->p.nz + 2);
}

orc_lattice* orc_create(const orc_params* p, int ntile, const int slice[][3][2]) {
  if (p->nx < 1 || p->ny < 1 || p->nz < 1 || ntile < 1) return NULL;
  if (p->casetype != ORC_CAVITY && p->nx < 3) return NULL;
  orc_lattice* L = (orc_lattice*)calloc(1, sizeof(orc_lattice));
  L->p = *p;
  L->ntile = ntile;
  L->slice = (int(*)[3][2])malloc(sizeof(int[3][2]) * ntile);
  memcpy(L->slice, slice, sizeof(int[3][2]) * ntile);
  L->move = (int(*)[ORC_Q])malloc(sizeof(int[ORC_Q]) * ntile);
  orc_set_move_decomp(ntile, (const int(*)[3][2])L->slice, L->move);
  L->n = (size_t)p->nx * p->ny * p->nz;
  L->npad = (size_t)(p->nx + 2) * (p->ny + 2) * (p->nz + 2);
  L->arF = (real*)calloc(L->n * ORC_Q, sizeof(real));
  L->arFTemp = (real*)calloc(L->npad * ORC_Q, sizeof(real));
  L->rho = (real*)calloc(L->n, sizeof(real));
  L->prev_rho = (real*)calloc(L->n, sizeof(real));
  L->ux = (real*)calloc(L->n, sizeof(real));
  L->uy = (real*)calloc(L->n, sizeof(real));
  L->uz = (real*)calloc(L->n, sizeof(real));
  L->tau_eddy = (real*)calloc(L->n, sizeof(real));
  L->type = (int*)calloc(L->n, sizeof(int));
  L->dirstate = (int*)calloc(L->n, sizeof(int));
  L->uin = (real*)calloc((size_t)p->ny * p->nz, sizeof(real));
  double* uin = (double*)calloc((size_t)p->ny * p->nz, sizeof(double));
  if (!L->arF || !L->arFTemp || !L->rho || !L->prev_rho || !L->ux || !L->uy || !L->uz || !L->tau_eddy ||
      !L->type || !L->dirstate || !L->uin || !uin) {
    free(uin);
    orc_destroy(L);
    return NULL;
  }
  orc_inlet_profile(p, uin);
  for (size_t i = 0; i < (size_t)p->ny * p->nz; ++i)
    L->uin[i] = p->casetype == ORC_PIPE ? (real)p->ref_vel[0] : (real)uin[i]; /* pipe: kernel.cu:599-608 */
  free(uin);
  orc_set_mask(L, NULL);
  return L;
}

static int pipe_solid(const orc_params* p, int y, int z) {
  const double dy = y - p->pipe_cy, dz = z - p->pipe_cz;
  return dy * dy + dz * dz > p->pipe_radius * p->pipe_radius;
}

void orc_set_mask(orc_lattice* L, const unsigned char* mask) {
  const orc_params* p = &L->p;
  free(L->mask);
  L->mask = NULL;
  if (!mask && p->casetype != ORC_PIPE) return;
  L->mask = (unsigned char*)malloc(L->n);
  if (mask) {
    memcpy(L->mask, mask, L->n);
    return;
  }
  for (int y = 0; y < p->ny; ++y) /* keSetNodeType, kernel.cu:1078-1094 */
    for (int z = 0; z < p->nz; ++z)
      for (int x = 0; x < p->nx; ++x) L->mask[nid(L, x, y, z)] = (unsigned char)pipe_solid(p, y, z);
}

void orc_destroy(orc_lattice* L) {
  if (!L) return;
  free(L->slice); free(L->move); free(L->arF); free(L->arFTemp); free(L->rho); free(L->prev_rho);
  free(L->ux); free(L->uy); free(L->uz); free(L->tau_eddy); free(L->type); free(L->dirstate); free(L->uin);
  free(L->mask);
  free(L);
}

long orc_steps(const orc_lattice* L) { return L->steps; }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* Node type for global coordinates.  Cavity: lbm_codelets.cpp:183-201.
 * Channel (square duct, no reference counterpart beyond the inlet/outlet
 * flags of kernel.cu:1081-1088): x = 0 inlet, x = nx-1 outlet, the four side
 * faces are fixed walls and override inlet/outlet.  The pipe uses the duct's
 * types under its solid mask (applied by the caller). */
static int node_type(const orc_params* p, int x, int y, int z) {
  int t = ORC_NODE_FLUID;
  if (p->casetype == ORC_CAVITY) {
    if (y == 0) t = ORC_NODE_FIXEDWALL;
    if (y == p->ny - 1) t = ORC_NODE_MOVINGWALL;
    if (x == 0 || x == p->nx - 1 || z == 0 || z == p->nz - 1) t = ORC_NODE_FIXEDWALL;
  } else {
    if (x == 0) t = ORC_NODE_INLET;
    if (x == p->nx - 1) t = ORC_NODE_OUTLET;
    if (y == 0 || y == p->ny - 1 || z == 0 || z == p->nz - 1) t = ORC_NODE_FIXEDWALL;
  }
  return t;
}

/* Initialize vertex, lbm_codelets.cpp:140-241, run per tile with global coordinates. */
void orc_initialize(orc_lattice* L) {
  const orc_params* p = &L->p;
  for (int t = 0; t < L->ntile; ++t) {
    int c[10];
    if (orc_slice_chunk(t, (const int(*)[3][2])L->slice, c)) continue;
#pragma omp parallel for
    for (int j = 0; j < c[7]; ++j) {
      const int y = j + c[2];
      for (int k = 0; k < c[8]; ++k) {
        const int z = k + c[4];
        for (int i = 0; i < c[6]; ++i) {
          const int x = i + c[0];
          const size_t n = nid(L, x, y, z);
          int ds = 0; /* keInitialNodeDirection: bit q set iff x + e_q is inside */
          for (int q = ORC_Q - 1; q > 0; --q) {
            const int x1 = x + EX[q], y1 = y + EY[q], z1 = z + EZ[q];
            const int in = x1 >= 0 && x1 < p->nx && y1 >= 0 && y1 < p->ny && z1 >= 0 && z1 < p->nz;
            ds = in ? (ds | 1) << 1 : ds << 1;
          }
          L->dirstate[n] = ds | 1;
          const int ty = (L->mask && L->mask[n]) ? ORC_NODE_STRUCT : node_type(p, x, y, z);
          L->type[n] = ty;
          if (ty == ORC_NODE_STRUCT) { /* keInitialize skips it (kernel.cu:294): zeros */
            memset(&L->arF[n * ORC_Q], 0, ORC_Q * sizeof(real));
            L->rho[n] = L->ux[n] = L->uy[n] = L->uz[n] = L->prev_rho[n] = L->tau_eddy[n] = 0;
            continue;
          }
          real u[3] = {0, 0, 0};
          if (ty == ORC_NODE_MOVINGWALL && p->init_lid_velocity) {
            u[0] = (real)p->ref_vel[0]; u[1] = (real)p->ref_vel[1]; u[2] = (real)p->ref_vel[2];
          }
          cell_feq((real)p->ref_den, u, &L->arF[n * ORC_Q]);
          L->rho[n] = (real)p->ref_den;
          L->ux[n] = u[0]; L->uy[n] = u[1]; L->uz[n] = u[2];
          L->prev_rho[n] = 0;
          L->tau_eddy[n] = 0;
        }
      }
    }
  }
  L->steps = 0;
}

/* ComputeTurbulence_LES (lbm_codelets.cpp:558-668); skipped when LES is off,
 * leaving tau_eddy = 0 as the memset of kernel.cu:143 does. */
static void phase_les(orc_lattice* L) {
  if (!L->p.enable_les) return;
#pragma omp parallel for
  for (long n = 0; n < (long)L->n; ++n) {
    if (L->type[n] == ORC_NODE_STRUCT) continue;
    const real u[3] = {L->ux[n], L->uy[n], L->uz[n]};
    L->tau_eddy[n] = cell_tau_eddy(&L->p, &L->arF[n * ORC_Q], L->rho[n], u);
  }
}

/* Collision_MRT (lbm_codelets.cpp:434-549) or keCollision_SRT (kernel.cu:458-487),
 * writing the interior of the ghost-padded arFTemp (runSimulation.cpp:70). */
static void phase_collide(orc_lattice* L) {
  const orc_params* p = &L->p;
#pragma omp parallel for
  for (int y = 0; y < p->ny; ++y)
    for (int z = 0; z < p->nz; ++z)
      for (int x = 0; x < p->nx; ++x) {
        const size_t n = nid(L, x, y, z);
        if (L->type[n] == ORC_NODE_STRUCT) continue;
        const real u[3] = {L->ux[n], L->uy[n], L->uz[n]};
        real* out = &L->arFTemp[pid(L, x + 1, y + 1, z + 1) * ORC_Q];
        if (p->enable_mrt)
          cell_mrt(p, &L->arF[n * ORC_Q], L->rho[n], u, L->tau_eddy[n], out);
        else
          cell_srt(p, &L->arF[n * ORC_Q], L->rho[n], u, L->tau_eddy[n], out);
      }
}

/* Streaming vertex (lbm_codelets.cpp:37-104): pull through the tile's padded
 * arFTemp slice with the tile's Move offsets (runSimulation.cpp:101).  A solid
 * donor sends nothing (keStreaming, kernel.cu:494, 511); the link bounces back
 * at the node instead, f_opp(q)(x) = f*_q(x) (keBoundaryStruct, kernel.cu:
 * 659-697).  The reference applies that rule to FLUID nodes only and leaves
 * such links of inlet/outlet/wall nodes holding the previous step's value;
 * here every non-solid node bounces back. */
static void phase_stream(orc_lattice* L) {
  for (int t = 0; t < L->ntile; ++t) {
    int c[10];
    if (orc_slice_chunk(t, (const int(*)[3][2])L->slice, c)) continue;
    const int nx2 = c[6] + 2, nz2 = c[8] + 2, nslave2 = nx2 * nz2;
    const int* mv = L->move[t];
    /* loc = x + z nx2 + y nslave2 + Move[q] is the donor's index inside the
     * tile's padded slice; Move[q] = ox + oz nx2 + oy nslave2 with offsets in
     * {0, 1, 2}, so the donor's padded coordinates are (x + ox, y + oy, z + oz)
     * (decoded once per tile instead of once per link) */
    int ox[ORC_Q], oy[ORC_Q], oz[ORC_Q];
    for (int q = 0; q < ORC_Q; ++q) {
      oy[q] = mv[q] / nslave2;
      oz[q] = (mv[q] % nslave2) / nx2;
      ox[q] = mv[q] % nx2;
    }
#pragma omp parallel for
    for (int y = 0; y < c[7]; ++y)
      for (int z = 0; z < c[8]; ++z)
        for (int x = 0; x < c[6]; ++x) {
          const size_t n = nid(L, x + c[0], y + c[2], z + c[4]);
          if (L->type[n] == ORC_NODE_STRUCT) continue;
          const int ds = L->dirstate[n];
          for (int q = 0; q < ORC_Q; ++q) {
            if (!((ds >> q) & 1)) continue;
            const int lx = x + ox[q], ly = y + oy[q], lz = z + oz[q];
            const size_t g = pid(L, c[0] + lx, c[2] + ly, c[4] + lz);
            if (L->mask && L->mask[nid(L, c[0] + lx - 1, c[2] + ly - 1, c[4] + lz - 1)]) {
              L->arF[n * ORC_Q + OPP[q]] = L->arFTemp[pid(L, x + c[0] + 1, y + c[2] + 1, z + c[4] + 1) * ORC_Q + q];
              continue;
            }
            L->arF[n * ORC_Q + OPP[q]] = L->arFTemp[g * ORC_Q + OPP[q]];
          }
        }
  }
}

/* BoundaryCavity (lbm_codelets.cpp:347-401), keBoundaryInlet (kernel.cu:566-613),
 * keBoundaryOutlet (kernel.cu:615-657): fill the links whose donor is outside. */
static void phase_boundary(orc_lattice* L) {
  const orc_params* p = &L->p;
  const real r0 = (real)p->ref_den;
  const real ul[3] = {(real)p->ref_vel[0], (real)p->ref_vel[1], (real)p->ref_vel[2]};
  for (int t = 0; t < L->ntile; ++t) {
    int c[10];
    if (orc_slice_chunk(t, (const int(*)[3][2])L->slice, c)) continue;
#pragma omp parallel for
    for (int j = 0; j < c[7]; ++j)
      for (int k = 0; k < c[8]; ++k)
        for (int i = 0; i < c[6]; ++i) {
          const int x = i + c[0], y = j + c[2], z = k + c[4];
          const size_t n = nid(L, x, y, z);
          const int ty = L->type[n];
          if (ty == ORC_NODE_FLUID || ty == ORC_NODE_STRUCT) continue;
          const int ds = L->dirstate[n];
          const real* ft = &L->arFTemp[pid(L, x + 1, y + 1, z + 1) * ORC_Q];
          real* f = &L->arF[n * ORC_Q];
          for (int q = 0; q < ORC_Q; ++q) {
            if ((ds >> q) & 1) continue;
            const int o = OPP[q];
            if (ty == ORC_NODE_FIXEDWALL) {
              f[o] = ft[q];
            } else if (ty == ORC_NODE_MOVINGWALL) {
              const real eu = EX[q] * ul[0] + EY[q] * ul[1] + EZ[q] * ul[2];
              f[o] = ft[q] - (real)6.0 * R_W[q] * r0 * eu;
            } else if (ty == ORC_NODE_INLET) {
              const real rin = L->rho[n + 1]; /* CurrDen[n + Move[1]], Move[1] = +1 */
              const real eu = EX[q] * L->uin[(size_t)y * p->nz + z];
              f[o] = ft[q] - (real)6.0 * R_W[q] * rin * eu;
            } else if (ty == ORC_NODE_OUTLET) {
              const size_t m = n - 1; /* Move[2] = -1 */
              const real ub[3] = {L->ux[n] + (real)0.5 * (L->ux[n] - L->ux[m]),
                                  L->uy[n] + (real)0.5 * (L->uy[n] - L->uy[m]),
                                  L->uz[n] + (real)0.5 * (L->uz[n] - L->uz[m])};
              const real usq = ub[0] * ub[0] + ub[1] * ub[1] + ub[2] * ub[2];
              const real eu = EX[q] * ub[0] + EY[q] * ub[1] + EZ[q] * ub[2];
              f[o] = -ft[q] + (real)2.0 * R_W[q] * r0 * ((real)1.0 + (real)4.5 * eu * eu - (real)1.5 * usq);
            }
          }
        }
  }
}

/* Macroscopic (lbm_codelets.cpp:266-317) */
static void phase_macro(orc_lattice* L) {
#pragma omp parallel for
  for (long n = 0; n < (long)L->n; ++n) {
    if (L->type[n] == ORC_NODE_STRUCT) continue;
    const real* f = &L->arF[n * ORC_Q];
    real den = 0, vx = 0, vy = 0, vz = 0;
    for (int q = 0; q < ORC_Q; ++q) {
      den += f[q];
      vx += f[q] * EX[q];
      vy += f[q] * EY[q];
      vz += f[q] * EZ[q];
    }
    L->prev_rho[n] = L->rho[n];
    L->rho[n] = den;
    L->ux[n] = vx / den;
    L->uy[n] = vy / den;
    L->uz[n] = vz / den;
  }
}

/* One step in the reference order: LES -> collide -> stream -> BC -> macroscopic
 * (runSimulation.cpp:164-165; kernel.cu:168-199). */
void orc_step(orc_lattice* L) {
  phase_les(L);
  phase_collide(L);
  phase_stream(L);
  phase_boundary(L);
  phase_macro(L);
  L->steps++;
}

void orc_run(orc_lattice* L, int nsteps) {
  for (int s = 0; s < nsteps; ++s) orc_step(L);
}

void orc_read_arF_padded(const orc_lattice* L, double* out) {
  memset(out, 0, L->npad * ORC_Q * sizeof(double));
  const orc_params* p = &L->p;
  for (int y = 0; y < p->ny; ++y)
    for (int z = 0; z < p->nz; ++z)
      for (int x = 0; x < p->nx; ++x) {
        const real* s = &L->arF[nid(L, x, y, z) * ORC_Q];
        double* d = &out[pid(L, x + 1, y + 1, z + 1) * ORC_Q];
        for (int q = 0; q < ORC_Q; ++q) d[q] = (double)s[q];
      }
}

void orc_write_arF_padded(orc_lattice* L, const double* in) {
  const orc_params* p = &L->p;
  for (int y = 0; y < p->ny; ++y)
    for (int z = 0; z < p->nz; ++z)
      for (int x = 0; x < p->nx; ++x) {
        real* d = &L->arF[nid(L, x, y, z) * ORC_Q];
        const double* s = &in[pid(L, x + 1, y + 1, z + 1) * ORC_Q];
        for (int q = 0; q < ORC_Q; ++q) d[q] = (real)s[q];
      }
  /* rho/u as Macroscopic computes them; PrevDen restarts at 0 as after init */
  phase_macro(L);
  memset(L->prev_rho, 0, L->n * sizeof(real));
  memset(L->tau_eddy, 0, L->n * sizeof(real));
  L->steps = 0;
}

void orc_read_macro(const orc_lattice* L, double* rho, double* ux, double* uy, double* uz) {
  for (size_t i = 0; i < L->n; ++i) {
    rho[i] = (double)L->rho[i];
    ux[i] = (double)L->ux[i];
    uy[i] = (double)L->uy[i];
    uz[i] = (double)L->uz[i];
  }
}

void orc_read_node_type(const orc_lattice* L, int* type, int* dirstate) {
  memcpy(type, L->type, L->n * sizeof(int));
  memcpy(dirstate, L->dirstate, L->n * sizeof(int));
}

/* Mass, momentum and the residual sum |PrevDen - CurrDen| of kernel.cu:241-245,
 * accumulated in double. */
void orc_diagnostics(const orc_lattice* L, double* mass, double mom[3], double* drho_l1) {
  double m = 0.0, mx = 0.0, my = 0.0, mz = 0.0, d = 0.0;
  for (size_t n = 0; n < L->n; ++n) {
    const double r = (double)L->rho[n];
    m += r;
    mx += r * (double)L->ux[n];
    my += r * (double)L->uy[n];
    mz += r * (double)L->uz[n];
    d += fabs((double)L->prev_rho[n] - r);
  }
  *mass = m;
  mom[0] = mx; mom[1] = my; mom[2] = mz;
  *drho_l1 = d;
}

int orc_precision_bytes(void) { return (int)sizeof(real); }
