// batch.hpp -- host runtime: lattices resident in device memory, the work
// list that maps workgroups onto them, graph-captured stepping, export.
#pragma once
#include <cstddef>
#include <vector>

#include "lattice.hpp"
#include "nflow.h"

namespace nflow {

// A sub-block [x0,x1) x [y0,y1) x [z0,z1) of one lattice (an IPU tile of the
// reference's decomposition); workgroups never straddle two tiles.
struct Tile {
  int x0, x1, y0, y1, z0, z1;
};

struct HostLattice {
  nflow_input in{};
  std::vector<Tile> tiles;
  size_t bytes = 0;      // device bytes of this lattice
  void* mem = nullptr;   // base of its device memory
  bool own_mem = false;  // allocated with hipMalloc by this lattice
  void* mask_mem = nullptr;  // solid mask allocated by nflow_set_solid_mask (pipe: inside mem)
  LatticeDev dev{};
  long long steps_done = 0;
  int cur = 0;  // buffer holding the newest post-collision state
  bool initialized = false;
  // buf[cur ^ 1] holds the streamed state written by init/import (no step
  // launched since), rather than the previous post-collision state
  bool prev_streamed = true;
  // the last launch was a two-step launch: buf[cur ^ 1] holds the state two
  // steps back (the one that launch read), not the previous step's
  bool pair_last = false;
};

int validate_input(const nflow_input& in, int storage);
size_t lattice_bytes(const nflow_input& in, int storage);
// Lay the lattice's buffers out in `mem` (lattice_bytes long) and fill dev.
void lattice_layout(HostLattice& h, const nflow_input& in, int storage, void* mem);
// Upload the inlet profile of a channel or pipe (host-computed, double ->
// float) and the pipe's solid mask.
int lattice_upload_inlet(HostLattice& h, hipStream_t st);
// Pipe geometry: node (x, y, z) is solid when (y - cy)^2 + (z - cz)^2 > r^2.
bool pipe_solid(const nflow_input& in, int y, int z);

class Batch {
 public:
  int storage = NFLOW_STORAGE_F32;
  std::vector<HostLattice> lats;
  hipStream_t stream = nullptr;
  int device = 0;  // the device current at open(), which owns the stream
  int cu_count = 256;
  nflow_tuning tuning{256, 0, 0, 10, -1, 0.0};

  // device tables
  LatticeDev* d_lats = nullptr;
  size_t d_lats_cap = 0;
  WorkItem* d_items = nullptr;
  size_t d_items_cap = 0;
  int nitems = 0;
  int grid = 0;
  bool tables_dirty = true;
  bool masked = false;  // some lattice has solid nodes: step with the masked kernel
  bool sr_any = false;  // some lattice rounds its 16-bit stores stochastically

  int gparity = 0;  // parity argument of the next step launch
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_steps_built = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int last_steps = 0;
  long long launches = 0;
  // the state of the last run was read back: the next run of an even number
  // of steps ends with single steps rather than a two-step launch
  bool read_back = false;

  // export scratch
  float* d_aos = nullptr;
  size_t d_aos_cap = 0;
  float* d_macro = nullptr;
  size_t d_macro_cap = 0;
  double* d_partials = nullptr;
  size_t d_partials_cap = 0;
  double* d_out5 = nullptr;
  int* d_types = nullptr;
  size_t d_types_cap = 0;
  uint32_t* d_flags = nullptr;  // divergence flags of every lattice, gathered
  size_t d_flags_cap = 0;
  std::vector<uint32_t> h_flags;

  Batch() = default;
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;
  ~Batch();

  int open();  // stream, events, device properties
  int add_lattice(const nflow_input& in, const std::vector<Tile>& tiles, void* arena_mem);
  void remove_lattice(int idx);  // frees owned memory; the caller syncs first
  int upload_tables();
  // step0: the step count the state stands at (0 for a new lattice)
  int init_lattice(int idx, const float* d_src_aos, long long step0 = 0);
  int init_all();
  int launch_steps(int nsteps);  // enqueue only
  // After a run that ended with a two-step launch, re-run that launch as two
  // single steps from the state it read (still in buf[cur ^ 1]), so that
  // buf[cur ^ 1] holds the previous step's state again; the same values, bit
  // for bit.  Kernel launches only: no allocation.  No-op otherwise.
  int settle_pair();
  int run(int nsteps);           // enqueue with timing events
  int synchronize();
  int timing(double* ms_per_step, int* steps);
  // diverged (nullable): the lattice's divergence flag.  consume: the lattice
  // is not stepped again, so rho/u and the sums may be formed in its idle
  // population buffer instead of allocated scratch (no aos).
  int export_one(int idx, float* aos, float* rho, float* ux, float* uy, float* uz, double* diag5,
                 int* diverged = nullptr, bool consume = false);
  // flags of all lattices in order (tables must be current)
  int diverged_flags(std::vector<int>& out);
  int import_one(int idx, const float* aos_host, long long step0);
  int node_types(int idx, int* type, int* dirstate);
  int autotune(int trial_steps, nflow_tuning* out);
  int set_tuning(const nflow_tuning& t);
  // y-chunk height of two-step launches for this batch under `t` (tuning by
  // default), resolving temporal = -1; 0: one step per launch
  int step2_height(const nflow_tuning* t = nullptr) const;
  void destroy_graphs();
  long long total_nodes() const;
  size_t device_bytes() const;

 private:
  int ensure_scratch(size_t aos, size_t macro, size_t partials, size_t types);
  int build_graph(int parity);
};

}  // namespace nflow
