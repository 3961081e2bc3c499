/*
 * nflow.h -- C ABI of the MI355X-native D3Q19 lattice-Boltzmann solver.
 *
 * Drop-in for the reference's time-step loop (eugenekim-mzc/NFLOW-RED-Poplar):
 * it keeps the reference's parameter block, its decomposition API and its
 * exported population layout, and runs the step on hand-written HIP kernels.
 * Plain C types only; no exceptions cross this boundary.  Every function
 * returns 0 on success and a negative NFLOW_E* code on failure (the reference
 * calls exit(0) instead, inverseMRT.cpp:89, kernel.cu:711/979);
 * nflow_last_error() gives the message for the calling thread.
 *
 * Ownership: the library owns all device memory; callers own host buffers.
 * Reads and writes of host buffers are synchronous.  A solver, batch or
 * server handle is used by one host thread at a time (the server's submit and
 * poll calls are the exception: they may come from any thread).
 *
 * Layouts (axis order [y][z][x], y slowest, as in main.cpp:1 and
 * lbm_codelets.cpp:52):
 *   arF   : float [ny+2][nz+2][nx+2][19], one ghost layer per side, ghosts = 0
 *           (the layout README:9-11 documents; the reference's own tensor is
 *           the unpadded [ny][nz][nx][19] of main.cpp:216)
 *   macro : float [ny][nz][nx] for rho, ux, uy, uz
 */
#ifndef NFLOW_H
#define NFLOW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NFLOW_DIMENSION 3       /* common.h:7  */
#define NFLOW_DIRECTION_SIZE 19 /* common.h:8  */
#define NFLOW_XDIM 0            /* common.h:13-15: axis ids of Direction[a][q] */
#define NFLOW_ZDIM 1
#define NFLOW_YDIM 2

/* Case numbers pass the reference's values through unchanged (common.h:26-27):
 * a caller may copy Inputval.casetype straight into nflow_input.casetype. */
#define NFLOW_PIPE 0    /* common.h:27 POISEUILLE: circular pipe of solid nodes (kernel.cu:1078-1095) */
#define NFLOW_CAVITY 1  /* common.h:26 */
#define NFLOW_CHANNEL 2 /* no reference value: square duct, parabolic velocity inlet, pressure outlet, walls */

#define NFLOW_NODE_FLUID (1 << 1) /* common.h:32-37 */
#define NFLOW_NODE_FIXEDWALL (1 << 2)
#define NFLOW_NODE_MOVINGWALL (1 << 3)
#define NFLOW_NODE_INLET (1 << 4)
#define NFLOW_NODE_OUTLET (1 << 5)
#define NFLOW_NODE_STRUCT (1 << 6)

#define NFLOW_STORAGE_F32 0  /* FP32 populations */
#define NFLOW_STORAGE_F16 1  /* FP16 of f - w_q, FP32 arithmetic */
#define NFLOW_STORAGE_BF16 2 /* BF16 of f - w_q, FP32 arithmetic */

#define NFLOW_OK 0
#define NFLOW_EINVAL (-1)  /* bad argument */
#define NFLOW_ENOMEM (-2)  /* device or host allocation failed */
#define NFLOW_EDEVICE (-3) /* HIP runtime error */
#define NFLOW_ESTATE (-4)  /* call out of order (e.g. run before initialize) */
#define NFLOW_EBUDGET (-5) /* request larger than the server's HBM budget */
#define NFLOW_ENOTFOUND (-6)
#define NFLOW_EDIVERGED (-7) /* the lattice went unstable: some node's density was not a positive finite number */

/* Parameter block.  Superset of Inputval (struct.h:78-100): the fields the
 * step uses, plus storage precision and the channel inlet description. */
typedef struct nflow_input {
  int h;                 /* Inputval.h: reference length (main.cpp:59) */
  int nx, ny, nz;        /* interior nodes per axis */
  int maxitr;            /* steps nflow_main runs after initialisation (main.cpp:291) */
  int nwrite;            /* output period of nflow_main (kernel.cu:205) */
  int casetype;          /* NFLOW_PIPE (0), NFLOW_CAVITY (1) or NFLOW_CHANNEL (2) */
  int enable_mrt;        /* 1: MRT collision, 0: SRT/BGK */
  int enable_les;        /* Smagorinsky eddy viscosity */
  int storage;           /* NFLOW_STORAGE_* */
  int init_lid_velocity; /* 1: lid nodes start at the lid velocity (lbm_codelets.cpp:215-220) */
  double tau;            /* m_dTau */
  double smagorinsky;    /* m_dSmagorinsky (Cs) */
  double ref_den;        /* m_dRefDen */
  double ref_vel[3];     /* m_dRefVelx/y/z: lid velocity */
  double w_eps, w_epsj, w_xx;   /* dWeps, dWepsj, dWXX */
  double inlet_peak;     /* channel: peak inlet velocity of the parabola */
  double perturb;        /* channel: relative amplitude of the seeded inlet perturbation */
  unsigned long long seed;
  /* pipe: nodes with (y - cy)^2 + (z - cz)^2 > r^2 are solid (kernel.cu:1090
   * uses cy = cz = 25, r = 20); the inlet velocity is ref_vel (kernel.cu:601) */
  double pipe_radius, pipe_cy, pipe_cz;
  /* FP16/BF16 storage: 1 rounds each stored f - w_q stochastically instead of
   * to nearest: it goes to the grid point above with probability equal to its
   * distance from the one below, in units of the grid spacing (exact for
   * normal values; gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32).  The random
   * bits come from a hash of the node id and the collision count, so rounding
   * errors do not correlate from step to step and average out over long runs
   * instead of accumulating; runs stay deterministic and a lattice rounds the
   * same way alone and in a batch.  1 is the default (nflow_input_defaults):
   * it costs ~7% of the 16-bit step rate and holds the velocity error of the
   * long benchmark runs under 1e-3, which rounding to nearest (0) does not.
   * Ignored for FP32. */
  int stochastic_rounding;
} nflow_input;

typedef struct nflow_solver nflow_solver;
typedef struct nflow_batch nflow_batch;
typedef struct nflow_server nflow_server;

/* Reference defaults of main.cpp:59-85 (10^3 cavity, 199 steps, MRT+LES, Cs = 0). */
void nflow_input_defaults(nflow_input* in);

/* ---- decomposition (replaces decomposition.cpp) --------------------------
 * decomposition.cpp:98-161 with the XDECOMP/YDECOMP/ZDECOMP macros of
 * common.h:17-19 as arguments.  Fills slice[t][axis][start,end) (axis 0 = x,
 * 1 = y, 2 = z) and returns the number of tiles, or a negative code. */
int nflow_decomposition(const nflow_input* in, int xdecomp, int ydecomp, int zdecomp, int slice[][3][2],
                        int max_tiles);
/* slice_chunk, decomposition.cpp:66-94: returns 1 when the tile is empty. */
int nflow_slice_chunk(unsigned tile, const int slice[][3][2], int* xs, int* xe, int* ys, int* ye, int* zs, int* ze,
                      int* chunk_nx, int* chunk_ny, int* chunk_nz, int* chunk_len);
/* set_move_decomp, decomposition.cpp:17-62: per-tile neighbour offsets into
 * the tile's ghost-padded sub-block. */
int nflow_set_move_decomp(int ntile, const int slice[][3][2], int move[][NFLOW_DIRECTION_SIZE]);
/* The moment basis M of main.cpp:149-171. */
void nflow_mrt_matrix(float m[NFLOW_DIRECTION_SIZE][NFLOW_DIRECTION_SIZE]);
/* inverseMRT, inverseMRT.cpp:8-92: Gauss-Jordan with partial pivoting in
 * double; returns NFLOW_EINVAL for a singular matrix instead of exit(0). */
int nflow_inverse_mrt(const float m[NFLOW_DIRECTION_SIZE][NFLOW_DIRECTION_SIZE],
                      float minv[NFLOW_DIRECTION_SIZE][NFLOW_DIRECTION_SIZE]);

/* ---- one lattice (replaces initialize.cpp + runSimulation.cpp + the
 *      Engine run of main.cpp:278-317) ------------------------------------
 * nflow_create: the tiles must partition the box exactly; they become the
 * work partition of the kernels (results do not depend on them). */
int nflow_create(const nflow_input* in, int ntile, const int slice[][3][2], nflow_solver** out);
/* Initialize compute set (initialize.cpp:23-98, lbm_codelets.cpp:111-242). */
int nflow_initialize(nflow_solver* s);
/* Repeat(nsteps, runSimulation step) (main.cpp:291, runSimulation.cpp:164-165). */
int nflow_run(nflow_solver* s, int nsteps);
int nflow_synchronize(nflow_solver* s);
long long nflow_steps_done(const nflow_solver* s);
/* readTensor("D2H-arF") (main.cpp:317): post-stream, post-boundary populations. */
int nflow_read_arF(nflow_solver* s, float* dst);
/* writeTensor("H2D-arF") (main.cpp:303); also the checkpoint/restart path.
 * rho/u restart from the moments of src; the step counter restarts at 0. */
int nflow_write_arF(nflow_solver* s, const float* src);
/* CurrDen / CurrVelx/y/z after the last step (Macroscopic, lbm_codelets.cpp:248-318). */
int nflow_read_macro(nflow_solver* s, float* rho, float* ux, float* uy, float* uz);
/* Node flags and direction masks (keSetNodeType / keInitialNodeDirection), [ny][nz][nx]. */
int nflow_read_node_type(nflow_solver* s, int* node_type, int* direction_state);
/* Solid obstacles (NODE_STRUCT, kernel.cu:659-697, 1090): mask[ny][nz][nx],
 * non-zero = solid.  Solid nodes are not updated and read back as zero; a
 * link whose donor is solid bounces back at the node.  Call before
 * nflow_initialize (the pipe case builds its own mask). */
int nflow_set_solid_mask(nflow_solver* s, const unsigned char* mask);
/* FP64 reductions on the device: total mass, momentum and the residual
 * sum |PrevDen - CurrDen| of kernel.cu:241-245.  Returns NFLOW_EDIVERGED
 * (outputs still filled) when the lattice has gone unstable: some node's
 * density is not a positive finite number (checked over every node by this
 * call), an earlier check since initialisation found so, or a sum is not
 * finite.  The reference only prints the residual (kernel.cu:232-248) and
 * runs on. */
int nflow_diagnostics(nflow_solver* s, double* mass, double mom[3], double* drho_l1);
/* nflow_diagnostics without allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The same kernels and
 * the same sums, bit for bit; the FP64 partial sums live in a static device
 * array of the library (5 MiB per device, held until the process ends), so
 * lattices of more than 2^25 = 33,554,432 nodes are refused with NFLOW_EINVAL.
 * Calls on one device share that array: calls that overlap on a device race,
 * so callers serialise them. */
int nflow_diagnostics_noalloc(nflow_solver* s, double* mass, double mom[3], double* drho_l1);
/* Device time of the last nflow_run, per step (HIP events on the solver's stream). */
int nflow_timing(nflow_solver* s, double* ms_per_step, int* steps_timed);
void nflow_destroy(nflow_solver* s);
const char* nflow_last_error(void);

/* ---- restart files (Inputval.irestart/irst, struct.h:84; kernel.cu:706) ----
 * A checkpoint is the exported arF, the step count, the storage format and a
 * hash of the parameters the step uses (case, collision model, tau, LES,
 * lid/inlet description, pipe geometry, rounding mode): loading it into a
 * solver with the same extents, storage and parameters continues the run
 * exactly where it was saved (bitwise for FP32 storage and for stochastic
 * rounding's step-seeded stores); any other solver is refused with
 * NFLOW_EINVAL. */
int nflow_save_checkpoint(nflow_solver* s, const char* path);
int nflow_load_checkpoint(nflow_solver* s, const char* path);
int nflow_checkpoint_info(const char* path, int* nx, int* ny, int* nz, long long* steps);

/* ---- output files (saveFile, kernel.cu:699-1018; host only) -----------------
 * VTK unstructured-grid file (binary appended): points, hexahedral cells,
 * Velocity, Density and, when node_type is given, the six node-flag fields.
 * Arrays are [ny][nz][nx]. */
int nflow_write_vtu(const char* path, int nx, int ny, int nz, const float* rho, const float* ux, const float* uy,
                    const float* uz, const int* node_type);
/* result.csv: "x,y,z,rho,u,v,w", one line per node. */
int nflow_write_csv(const char* path, int nx, int ny, int nz, const float* rho, const float* ux, const float* uy,
                    const float* uz);

/* ---- many lattices (no reference counterpart) ------------------------------
 * Independent lattices of any sizes advanced together: one launch per time
 * step covers every lattice; steps are replayed from a captured HIP graph;
 * the launch shape comes from nflow_batch_autotune. All lattices share one
 * storage precision (the first input's). */
typedef struct nflow_tuning {
  int block;       /* threads per workgroup: 64, 128, 256 or 512 */
  int xcd_order;   /* 1: each XCD receives a contiguous range of work items */
  int persistent;  /* 0: one workgroup per work item; k > 0: k workgroups per CU, grid-stride */
  int graph_steps; /* steps per captured graph (even), 0 = no graph */
  /* Two steps per launch for a batch of one FP32 cavity or channel lattice
   * without solid nodes (other batches ignore this): x-z tiles of 256 x 8 nodes swept along
   * y in chunks of H planes, with results bit for bit those of one step per
   * launch.  A run of odd length ends with a single step; one of even length
   * ends with a two-step launch, which a read-back of its state re-runs as two
   * single steps (launches only), unless the previous run's state was read
   * back: then it ends with two single steps.  H > 0: that chunk
   * height; -1 (the default): chosen by the library from an estimate of the
   * launch time at heights 8 to 64, on where the estimate beats one step per
   * launch (cavities from about 176^3 up, channels of 128 x 128 cross-section
   * from 256 long); 0: one step per launch.  The autotuner times it. */
  int temporal;
  double ms_per_step; /* filled by the autotuner */
} nflow_tuning;

int nflow_batch_create(int count, const nflow_input* inputs, nflow_batch** out);
int nflow_batch_initialize(nflow_batch* b);
int nflow_batch_autotune(nflow_batch* b, int steps_per_trial, nflow_tuning* chosen);
int nflow_batch_set_tuning(nflow_batch* b, const nflow_tuning* t);
int nflow_batch_get_tuning(nflow_batch* b, nflow_tuning* t);
/* launch shape of a single-lattice solver (it is a batch of one; no
 * reference counterpart): the same fields and rules as for a batch */
int nflow_set_tuning(nflow_solver* s, const nflow_tuning* t);
/* chunk height the batch's runs use for two steps per launch under its tuning
 * (temporal = -1 resolved), 0 when every launch is a single step */
int nflow_batch_step2_height(nflow_batch* b);
int nflow_batch_run(nflow_batch* b, int nsteps);
int nflow_batch_synchronize(nflow_batch* b);
int nflow_batch_read_arF(nflow_batch* b, int idx, float* dst);
int nflow_batch_read_macro(nflow_batch* b, int idx, float* rho, float* ux, float* uy, float* uz);
/* As nflow_diagnostics, for lattice idx (NFLOW_EDIVERGED likewise). */
int nflow_batch_diagnostics(nflow_batch* b, int idx, double* mass, double mom[3], double* drho_l1);
/* flags[i] = 1 for each lattice that has gone unstable, else 0: some node's
 * rest population is not finite (checked now over every node of every
 * lattice, one population per node read: a NaN or infinity in any population
 * reaches it through the next collision) or an earlier check since the
 * lattice's initialisation found it unstable (see nflow_diagnostics). */
int nflow_batch_diverged(nflow_batch* b, int* flags);
int nflow_batch_timing(nflow_batch* b, double* ms_per_step, int* steps_timed);
long long nflow_batch_nodes(nflow_batch* b);
size_t nflow_batch_device_bytes(nflow_batch* b);
void nflow_batch_destroy(nflow_batch* b);

/* ---- serving (no reference counterpart) ------------------------------------
 * A continuous-batching loop on a host thread: requests are admitted in
 * arrival order as their device footprint fits the HBM budget and a slot is
 * free (smaller later requests may backfill), stay resident until their step
 * count is reached, and share every step launch with the other resident
 * lattices.  Results are returned in completion order. */
typedef struct nflow_request {
  nflow_input input;        /* storage is overridden by the server's */
  int nsteps;
  int want_fields;          /* 1: keep rho/u on the host for nflow_server_fetch_macro */
  unsigned long long tag;   /* caller's id, echoed in the result */
} nflow_request;

typedef struct nflow_result {
  unsigned long long tag;
  int status;               /* 0 ok, negative NFLOW_E*; NFLOW_EDIVERGED: the request went
                               unstable; the check after every tick stops it there */
  int nx, ny, nz;
  long long steps;
  double submit_s, admit_s, done_s; /* seconds since the server started */
  double mass, mom[3], drho_l1;     /* FP64 diagnostics of the final state */
} nflow_result;

typedef struct nflow_server_stats {
  long long submitted, admitted, completed, rejected;
  long long node_updates;   /* sum over completed requests of nodes * steps */
  long long launches;
  double busy_s;            /* wall time with at least one resident lattice */
  size_t peak_bytes;        /* peak device bytes held by resident lattices */
  int peak_resident;
  long long ticks;          /* rounds of step launches between host visits */
  double step_s;            /* host time from the first launch of a tick to its completion, summed */
} nflow_server_stats;

int nflow_server_create(size_t hbm_budget_bytes, int storage, int max_resident, int steps_per_tick,
                        nflow_server** out);
/* Requests submitted afterwards whose relaxation time tau lies below tau_min
 * are kept in FP32 storage (a second batch, its own launch per step) on a
 * server whose storage is FP16/BF16: near tau = 0.5 the 16-bit rounding of
 * f - w_q is weakly damped and accumulates.  0 (the default) turns it off. */
int nflow_server_set_tau_floor(nflow_server* s, double tau_min);
int nflow_server_submit(nflow_server* s, const nflow_request* r);
int nflow_server_poll(nflow_server* s, nflow_result* out, int max_results, int wait_ms);
int nflow_server_fetch_macro(nflow_server* s, unsigned long long tag, float* rho, float* ux, float* uy, float* uz);
int nflow_server_drain(nflow_server* s);
int nflow_server_stats_get(nflow_server* s, nflow_server_stats* st);
void nflow_server_destroy(nflow_server* s);

/* ---- measurement helpers ---------------------------------------------------
 * Device copy of `bytes` bytes split over `streams` equal arrays (a streaming
 * ceiling for the population traffic), timed with HIP events.  streams = 0
 * copies in the FP32 population layout instead: blocks of 19 runs of 64
 * floats, one 4-byte element of every run per thread. */
int nflow_copy_bandwidth(size_t bytes, int streams, int reps, double* gbps, double* ms_per_copy);
/* VALU issue ceiling: independent FP32 FMA chains (no memory traffic) on
 * 16 workgroups of 256 lanes per CU, timed with HIP events over `reps`
 * launches.  *g_wave_fma_per_s = wave-level v_fma_f32 instructions per
 * second in units of 1e9 (each one 64 lane FMAs). */
int nflow_valu_rate(int reps, double* g_wave_fma_per_s, double* ms_per_launch);
/* The step kernel's 16-bit store conversion applied to n host floats
 * (storage NFLOW_STORAGE_F16 or _BF16): to nearest even, or, with stochastic
 * = 1, stochastically with seed[i] for src[i] (the rounding described at
 * nflow_input.stochastic_rounding).  dst receives the 16-bit patterns.  For
 * checking the rounding; n <= 2^30. */
int nflow_round16(const float* src, const unsigned* seeds, size_t n, int storage, int stochastic,
                  unsigned short* dst);
int nflow_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif
