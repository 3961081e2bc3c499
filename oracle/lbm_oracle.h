/*
 * lbm_oracle.h -- FP64 CPU restatement of the reference D3Q19 LBM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * It restates, in plain double-precision loops, the reference's per-node
 * maths and step order:
 *   lbm_codelets.cpp : Streaming 21-105, Initialize 111-242, Macroscopic
 *                      248-318, BoundaryCavity 325-402, Collision_MRT
 *                      409-550, ComputeTurbulence_LES 558-668
 *   kernel.cu        : keCollision_SRT 458-487, keBoundaryInlet 566-613,
 *                      keBoundaryOutlet 615-657, keSetNodeType 1046-1097,
 *                      step order 154-203, residual 232-248
 *   decomposition.cpp: index_ghostcell 10-15, set_move_decomp 17-62,
 *                      slice_chunk 66-94, decomposition 98-161
 *   main.cpp         : parameters 59-85, tables 87-146, moment matrix 149-171
 *   runSimulation.cpp: step order 164-165
 * Layout follows the Poplar path: arF is AoS [ny][nz][nx][19], arFTemp is the
 * ghost-padded AoS [ny+2][nz+2][nx+2][19] (main.cpp:216-218).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_Q 19
#define ORC_XDIM 0 /* common.h:13-15 */
#define ORC_ZDIM 1
#define ORC_YDIM 2

#define ORC_NODE_FLUID (1 << 1) /* common.h:32-37 */
#define ORC_NODE_FIXEDWALL (1 << 2)
#define ORC_NODE_MOVINGWALL (1 << 3)
#define ORC_NODE_INLET (1 << 4)
#define ORC_NODE_OUTLET (1 << 5)
#define ORC_NODE_STRUCT (1 << 6)

#define ORC_PIPE 0   /* common.h:27 POISEUILLE of the CUDA path: the duct with a circular solid mask */
#define ORC_CAVITY 1 /* common.h:26 */
#define ORC_CHANNEL 2 /* the square-duct channel (no reference value) */

typedef struct orc_params {
  int nx, ny, nz;
  int casetype;          /* ORC_CAVITY, ORC_CHANNEL or ORC_PIPE */
  int enable_mrt;        /* 1: MRT (Collision_MRT), 0: SRT (keCollision_SRT) */
  int enable_les;        /* Smagorinsky eddy viscosity */
  int init_lid_velocity; /* 1: lid nodes start at U_lid (lbm_codelets.cpp:215-220) */
  double tau, smagorinsky, ref_den;
  double ref_vel[3];     /* lid velocity (x, y, z) */
  double w_eps, w_epsj, w_xx;
  double inlet_peak, perturb; /* channel inlet: peak * 16 eta(1-eta) zeta(1-zeta) * (1 + perturb*xi) */
  unsigned long long seed;
  /* MRT relaxation rates; a negative entry means 1/(tau + tau_eddy). */
  double s_rates[ORC_Q];
  /* pipe: node (x, y, z) is solid when (y - cy)^2 + (z - cz)^2 > radius^2 (kernel.cu:1090) */
  double pipe_radius, pipe_cy, pipe_cz;
} orc_params;

typedef struct orc_lattice orc_lattice;

/* ---- tables (main.cpp:87-171) ------------------------------------------- */
void orc_tables(double dir[3][ORC_Q], int opp[ORC_Q], double w[ORC_Q]);
void orc_mrt_matrix(double m[ORC_Q][ORC_Q]);
/* exact inverse of an orthogonal-row matrix: M^T diag(1/|row|^2). Returns 0, or -1 if rows are not orthogonal */
int orc_mrt_inverse_exact(const double m[ORC_Q][ORC_Q], double minv[ORC_Q][ORC_Q]);
void orc_default_params(orc_params* p);

/* ---- decomposition (decomposition.cpp) ---------------------------------- */
int orc_decomposition(int nx, int ny, int nz, int xdec, int ydec, int zdec, int slice[][3][2], int max_tiles);
int orc_slice_chunk(int tile, const int slice[][3][2], int out[10]);
int orc_set_move_decomp(int ntile, const int slice[][3][2], int move[][ORC_Q]);

/* ---- splitmix64-based inlet perturbation (counter-based, platform independent) */
unsigned long long orc_splitmix64(unsigned long long x);
double orc_inlet_xi(unsigned long long seed, int y, int z);
void orc_inlet_profile(const orc_params* p, double* uin /* [ny][nz] */);

/* ---- single-cell kernels (known-answer tests) --------------------------- */
void orc_cell_feq(const double dir[3][ORC_Q], const double w[ORC_Q], double rho, const double u[3], double feq[ORC_Q]);
double orc_cell_tau_eddy(const orc_params* p, const double f[ORC_Q], double rho, const double u[3]);
void orc_cell_mrt(const orc_params* p, const double f[ORC_Q], double rho, const double u[3], double tau_eddy,
                  double fout[ORC_Q]);
void orc_cell_srt(const orc_params* p, const double f[ORC_Q], double rho, const double u[3], double tau_eddy,
                  double fout[ORC_Q]);

/* ---- lattice ------------------------------------------------------------ */
orc_lattice* orc_create(const orc_params* p, int ntile, const int slice[][3][2]);
void orc_destroy(orc_lattice* L);
/* Solid nodes (NODE_STRUCT), mask[ny][nz][nx] non-zero = solid; NULL restores the
 * case's own geometry.  Takes effect at the next orc_initialize. */
void orc_set_mask(orc_lattice* L, const unsigned char* mask);
void orc_initialize(orc_lattice* L);
void orc_step(orc_lattice* L);
void orc_run(orc_lattice* L, int nsteps);
long orc_steps(const orc_lattice* L);
/* export [ny+2][nz+2][nx+2][19] with ghosts = 0 */
void orc_read_arF_padded(const orc_lattice* L, double* out);
/* import the same layout (ghosts ignored); recomputes rho/u as Macroscopic would */
void orc_write_arF_padded(orc_lattice* L, const double* in);
void orc_read_macro(const orc_lattice* L, double* rho, double* ux, double* uy, double* uz);
void orc_read_node_type(const orc_lattice* L, int* type, int* dirstate);
void orc_diagnostics(const orc_lattice* L, double* mass, double mom[3], double* drho_l1);
int orc_num_threads(void);
void orc_set_num_threads(int n); /* OpenMP worker count for later calls */
/* sizeof the working precision: 8 for liboracle.so, 4 for liboracle_f32.so */
int orc_precision_bytes(void);

#ifdef __cplusplus
}
#endif
#endif
