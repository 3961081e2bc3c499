// lattice.hpp -- structures shared by the host runtime and the HIP kernels.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

namespace nflow {

enum : int { CASE_PIPE = 0, CASE_CAVITY = 1, CASE_CHANNEL = 2 };  // NFLOW_PIPE/CAVITY/CHANNEL
enum : int { MODEL_SRT = 0, MODEL_MRT = 1 };
enum : int { ST_F32 = 0, ST_F16 = 1, ST_BF16 = 2 };
enum : int {
  NODE_FLUID = 1 << 1,
  NODE_FIXEDWALL = 1 << 2,
  NODE_MOVINGWALL = 1 << 3,
  NODE_INLET = 1 << 4,
  NODE_OUTLET = 1 << 5,
  NODE_STRUCT = 1 << 6
};

// Inlet / outlet state kept between steps (channel and pipe), one record of
// SIDE_REC floats per (y, z), [ny][nz]: the density of the x = 1 node (used by
// the inlet, kernel.cu:596-597) next to the inlet velocity u_in, and the
// velocities of the x = nx-1 and x = nx-2 nodes (used by the outlet,
// kernel.cu:645-648), each in its own 16-byte slot, so that an inlet node
// reads its inputs with one load and an outlet node with two.
enum : int { SIDE_RHO = 0, SIDE_UIN = 1, SIDE_U0 = 4, SIDE_U1 = 8, SIDE_REC = 12 };

// Words of LatticeDev::ctr: the two collision counters of the stochastic
// rounding, then the divergence flag: set by the export when some node's
// density is not a positive finite number and by the probe when some node's
// rest population is not finite; cleared only by initialisation.
enum : int { CTR_DIVERGED = 2, CTR_WORDS = 4 };
// Two steps per launch (k_step2): x-z tiles of STEP2_TX x STEP2_TZ nodes,
// STEP2_NT threads per workgroup (two nodes each).
enum : int { STEP2_TX = 256, STEP2_TZ = 8, STEP2_NT = 1024 };

// Exact unsigned division by a runtime constant d for n < 2^31:
// floor(n / d) = (n * magic) >> shift with shift = 31 + ceil(log2 d) and
// magic = ceil(2^shift / d) < 2^32 (Granlund & Montgomery 1994, Thm 4.2).
struct FastDiv {
  uint32_t magic;
  uint32_t shift;
};

__host__ __device__ inline uint32_t fastdiv(uint32_t n, FastDiv d) {
  return (uint32_t)(((uint64_t)n * d.magic) >> d.shift);
}

inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint32_t shift = 31 + l;
  const uint64_t p = 1ull << shift;
  FastDiv r;
  r.magic = (uint32_t)((p + d - 1) / d);
  r.shift = shift;
  return r;
}

// One lattice as the kernels see it.  Populations of node n = x + nx*(z + nz*y)
// live in buf[i] at the element pidx(q, n) of kernels.hip (256-byte runs of
// consecutive nodes per population); the two buffers and the two side-record
// arrays alternate each step (parity ^ parity0 is the one read).
struct alignas(16) LatticeDev {
  void* buf[2];
  float* side[2];
  const uint8_t* mask;  // solid nodes (non-zero), or null
  // Collision counter of the stochastic rounding, ping-pong by parity like
  // buf: the launch reading buf[p] rounds with ctr[p]; k_advance then writes
  // ctr[p] + 1 to ctr[p ^ 1] (k_init writes the first value).
  uint32_t* ctr;
  long long pstride;  // elements between population planes
  int nx, ny, nz, nxz;
  uint32_t n;         // interior nodes
  FastDiv div_nx, div_nxz;
  int casetype, model, les, parity0;
  int init_lid, sr, pad1, pad2;  // sr: stochastic rounding of 16-bit stores
  float tau, les_c, refden, inv_refden;
  float weps, wepsj, wxx, pad3;
  float ulid[4];
};

// A run of consecutive node ids of one lattice, processed by one workgroup.
struct alignas(16) WorkItem {
  int lat;
  uint32_t start;
  uint32_t count;
  uint32_t pad;
};

// Kernel launchers (kernels.hip).
// masked: use the instantiation that honours LatticeDev::mask (any lattice of
// the launch has solid nodes); sr: the one that can round stochastically (any
// lattice of the launch has LatticeDev::sr set), followed by k_advance over the
// nlat lattices.
hipError_t launch_step(int storage, bool masked, bool sr, int block, int grid, hipStream_t st,
                       const LatticeDev* d_lats, int nlat, const WorkItem* d_items, int nitems, int parity);
// Two steps in one launch (k_step2) for an FP32 cavity or channel lattice
// without solid nodes: reads buf[parity ^ parity0] (and the side records of
// that parity), writes the state two steps later to the other buffer (and
// records); y planes are swept in chunks of H.
hipError_t launch_step2(int storage, hipStream_t st, const LatticeDev& L, int parity, int H);
// step0: the step count the initialised (or imported) state stands at; the
// stochastic rounding of its look-ahead collision uses it.
hipError_t launch_init(int storage, hipStream_t st, const LatticeDev& L, int cur, const float* src_aos, uint32_t step0);
hipError_t launch_export(int storage, hipStream_t st, const LatticeDev& L, int prev, int prev_streamed, float* aos,
                         float* rho, float* ux, float* uy, float* uz, double* partials, int* nblocks_out);
hipError_t launch_reduce(hipStream_t st, const double* partials, int nblocks, double* out5);
// rest-population probe of every node of a work list (items of at most 512
// nodes), newest state = the buffer the launch with `parity` would read
hipError_t launch_probe(int storage, hipStream_t st, const LatticeDev* d_lats, const WorkItem* d_items, int nitems,
                        int parity);
// out[i] = divergence flag of lattice i (i < nlat)
hipError_t launch_flags(hipStream_t st, const LatticeDev* d_lats, int nlat, uint32_t* out);
hipError_t launch_node_type(hipStream_t st, const LatticeDev& L, int* type, int* dirstate);
hipError_t launch_copy_streams(hipStream_t st, const float* src, float* dst, size_t per_stream, int streams);
// nblk blocks of 19 runs of 64 floats (the FP32 population layout)
hipError_t launch_copy_runs(hipStream_t st, const float* src, float* dst, size_t nblk);
// grid workgroups of 256 lanes, each lane 64 * iters independent FMAs (out:
// grid * 256 floats)
hipError_t launch_fma_chains(hipStream_t st, float* out, int grid, int iters);
// 16-bit store conversion of n floats (storage ST_F16 or ST_BF16), to
// nearest or stochastically with per-element seeds (n < 2^32 * 256)
hipError_t launch_round16(hipStream_t st, int storage, const float* src, const uint32_t* seeds, uint16_t* dst, size_t n,
                          int stochastic);
int export_block_size();

}  // namespace nflow
