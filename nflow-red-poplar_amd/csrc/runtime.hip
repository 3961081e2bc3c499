// runtime.hip -- host runtime behind the C ABI: device memory of each
// lattice, the work list, graph-captured stepping, the autotuner and the
// export / diagnostics path.  Replaces the Poplar graph build + Engine of
// main.cpp:187-317 and the CUDA host loop of kernel.cu:18-280.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "batch.hpp"
#include "checkpoint.hpp"
#include "d3q19.hpp"
#include "errors.hpp"
#include "nflow.h"

namespace nflow {

// Extra elements between population planes (0: planes start every
// round_up(N, 128) elements).
#ifndef NFLOW_PLANE_PAD
#define NFLOW_PLANE_PAD 0
#endif

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
static constexpr unsigned long long NODE_LIMIT = (1ull << 31) - (1ull << 20);
static size_t elem_size(int storage) { return storage == NFLOW_STORAGE_F32 ? 4 : 2; }

int validate_input(const nflow_input& in, int storage) {
  if (in.nx < 1 || in.ny < 1 || in.nz < 1) return fail(NFLOW_EINVAL, "lattice extents must be >= 1");
  if (in.casetype != NFLOW_CAVITY && in.casetype != NFLOW_CHANNEL && in.casetype != NFLOW_PIPE)
    return fail(NFLOW_EINVAL, "casetype must be NFLOW_CAVITY, NFLOW_CHANNEL or NFLOW_PIPE");
  if (in.casetype != NFLOW_CAVITY && in.nx < 3) return fail(NFLOW_EINVAL, "channel and pipe need nx >= 3");
  if (in.casetype == NFLOW_PIPE && !(in.pipe_radius > 0.0))
    return fail(NFLOW_EINVAL, "pipe_radius must be positive (got %g)", in.pipe_radius);
  if (storage < NFLOW_STORAGE_F32 || storage > NFLOW_STORAGE_BF16) return fail(NFLOW_EINVAL, "bad storage");
  const unsigned long long n = (unsigned long long)in.nx * in.ny * in.nz;
  // node ids are 32-bit and divided exactly by FastDiv below 2^31; the
  // populations are addressed through windows with 32-bit offsets
  if (n >= NODE_LIMIT) return fail(NFLOW_EINVAL, "lattice of %llu nodes exceeds the %llu node ids the kernels address", n, NODE_LIMIT);
  // the step kernel addresses a window of two x-z planes plus one work item
  // with 32-bit byte offsets
  const unsigned long long plane = (unsigned long long)in.nx * in.nz + in.nx;
  if ((2 * plane + 1024) * 19ull * 4ull >= (1ull << 32))
    return fail(NFLOW_EINVAL, "x-z plane of %d x %d nodes exceeds the 32-bit window range", in.nx, in.nz);
  if (!(in.tau > 0.5)) return fail(NFLOW_EINVAL, "tau must exceed 0.5 (got %g)", in.tau);
  if (!(in.ref_den > 0.0)) return fail(NFLOW_EINVAL, "ref_den must be positive");
  if (in.enable_les && in.smagorinsky < 0.0) return fail(NFLOW_EINVAL, "smagorinsky must be >= 0");
  return 0;
}

struct Layout {
  size_t pstride, pop, side, ctr, mask, total;
};

static Layout layout_of(const nflow_input& in, int storage) {
  Layout l;
  const size_t n = (size_t)in.nx * in.ny * in.nz;
  l.pstride = round_up(n, 128) + NFLOW_PLANE_PAD;
  l.pop = round_up(19 * l.pstride * elem_size(storage), 256);
  l.side = round_up((size_t)SIDE_REC * in.ny * in.nz * sizeof(float), 256);
  l.ctr = 256;  // collision counters of the stochastic rounding, divergence flag
  l.mask = in.casetype == NFLOW_PIPE ? round_up(n, 256) : 0;  // one byte per node
  l.total = 2 * l.pop + 2 * l.side + l.ctr + l.mask;
  return l;
}

bool pipe_solid(const nflow_input& in, int y, int z) {
  const double dy = y - in.pipe_cy, dz = z - in.pipe_cz;
  return dy * dy + dz * dz > in.pipe_radius * in.pipe_radius;
}

size_t lattice_bytes(const nflow_input& in, int storage) { return layout_of(in, storage).total; }

void lattice_layout(HostLattice& h, const nflow_input& in, int storage, void* mem) {
  const Layout l = layout_of(in, storage);
  char* p = static_cast<char*>(mem);
  h.in = in;
  h.in.storage = storage;
  h.mem = mem;
  h.bytes = l.total;
  LatticeDev& d = h.dev;
  std::memset(&d, 0, sizeof(d));
  d.buf[0] = p;
  d.buf[1] = p + l.pop;
  d.side[0] = reinterpret_cast<float*>(p + 2 * l.pop);
  d.side[1] = reinterpret_cast<float*>(p + 2 * l.pop + l.side);
  d.ctr = reinterpret_cast<uint32_t*>(p + 2 * l.pop + 2 * l.side);
  d.mask = l.mask ? reinterpret_cast<const uint8_t*>(p + 2 * l.pop + 2 * l.side + l.ctr) : nullptr;
  d.pstride = (long long)l.pstride;
  d.nx = in.nx;
  d.ny = in.ny;
  d.nz = in.nz;
  d.nxz = in.nx * in.nz;
  d.n = (uint32_t)((size_t)in.nx * in.ny * in.nz);
  d.div_nx = make_fastdiv((uint32_t)in.nx);
  d.div_nxz = make_fastdiv((uint32_t)d.nxz);
  d.casetype = in.casetype == NFLOW_CAVITY ? CASE_CAVITY : in.casetype == NFLOW_PIPE ? CASE_PIPE : CASE_CHANNEL;
  d.model = in.enable_mrt ? MODEL_MRT : MODEL_SRT;
  // with Cs = 0 the Smagorinsky term vanishes identically (tau_eddy = 0)
  d.les = (in.enable_les && in.smagorinsky > 0.0) ? 1 : 0;
  d.parity0 = 0;
  d.init_lid = in.init_lid_velocity ? 1 : 0;
  d.sr = (in.stochastic_rounding && storage != NFLOW_STORAGE_F32) ? 1 : 0;
  d.tau = (float)in.tau;
  d.les_c = (float)(18.0 * 1.41421356237 * in.smagorinsky * in.smagorinsky);
  d.refden = (float)in.ref_den;
  d.inv_refden = (float)(1.0 / in.ref_den);
  d.weps = (float)in.w_eps;
  d.wepsj = (float)in.w_epsj;
  d.wxx = (float)in.w_xx;
  d.ulid[0] = (float)in.ref_vel[0];
  d.ulid[1] = (float)in.ref_vel[1];
  d.ulid[2] = (float)in.ref_vel[2];
}

// splitmix64 (Steele, Lea & Flood 2014), counter-based so the inlet
// perturbation of node (y, z) depends only on (seed, y, z).
static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int lattice_upload_inlet(HostLattice& h, hipStream_t st) {
  const nflow_input& in = h.in;
  if (in.casetype == NFLOW_CAVITY) return 0;
  // side records of both step parities: u_in in its slot, the rest zero until
  // the kernels write them
  std::vector<float> rec((size_t)in.ny * in.nz * SIDE_REC, 0.0f);
  for (int y = 0; y < in.ny; ++y)
    for (int z = 0; z < in.nz; ++z) {
      float& u = rec[((size_t)y * in.nz + z) * SIDE_REC + SIDE_UIN];
      if (in.casetype == NFLOW_PIPE) {  // keBoundaryInlet, kernel.cu:599-608: uniform RefVelx
        u = (float)in.ref_vel[0];
        continue;
      }
      const double eta = (y + 0.5) / in.ny, zeta = (z + 0.5) / in.nz;
      const uint64_t key = (uint64_t)in.seed ^ (((uint64_t)(uint32_t)y << 32) | (uint64_t)(uint32_t)z);
      const double xi = 2.0 * ((double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
      u = (float)(in.inlet_peak * 16.0 * eta * (1.0 - eta) * zeta * (1.0 - zeta) * (1.0 + in.perturb * xi));
    }
  for (int b = 0; b < 2; ++b)
    NFLOW_HIP(hipMemcpyAsync(h.dev.side[b], rec.data(), rec.size() * sizeof(float), hipMemcpyHostToDevice, st));
  std::vector<uint8_t> m;
  if (in.casetype == NFLOW_PIPE) {  // keSetNodeType, kernel.cu:1078-1094
    m.resize((size_t)h.dev.n);
    for (int y = 0; y < in.ny; ++y)
      for (int z = 0; z < in.nz; ++z) {
        const uint8_t s = pipe_solid(in, y, z) ? 1 : 0;
        std::memset(&m[(size_t)in.nx * ((size_t)z + (size_t)in.nz * y)], s, (size_t)in.nx);
      }
    NFLOW_HIP(hipMemcpyAsync(const_cast<uint8_t*>(h.dev.mask), m.data(), m.size(), hipMemcpyHostToDevice, st));
  }
  NFLOW_HIP(hipStreamSynchronize(st));
  return 0;
}

// ---------------------------------------------------------------------------
Batch::~Batch() {
  if (stream) (void)hipStreamSynchronize(stream);
  destroy_graphs();
  for (auto& h : lats) {
    if (h.own_mem && h.mem) (void)hipFree(h.mem);
    if (h.mask_mem) (void)hipFree(h.mask_mem);
  }
  (void)hipFree(d_lats);
  (void)hipFree(d_items);
  (void)hipFree(d_aos);
  (void)hipFree(d_macro);
  (void)hipFree(d_partials);
  (void)hipFree(d_out5);
  (void)hipFree(d_types);
  (void)hipFree(d_flags);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

int Batch::open() {
  int dev = 0;
  NFLOW_HIP(hipGetDevice(&dev));
  device = dev;
  NFLOW_HIP(hipDeviceGetAttribute(&cu_count, hipDeviceAttributeMultiprocessorCount, dev));
  NFLOW_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  NFLOW_HIP(hipEventCreate(&ev0));
  NFLOW_HIP(hipEventCreate(&ev1));
  return 0;
}

int Batch::add_lattice(const nflow_input& in, const std::vector<Tile>& tiles, void* arena_mem) {
  int rc = validate_input(in, storage);
  if (rc) return rc;
  rc = settle_pair();
  if (rc) return rc;
  HostLattice h;
  void* mem = arena_mem;
  const size_t bytes = lattice_bytes(in, storage);
  if (!mem) {
    if (hipMalloc(&mem, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(NFLOW_ENOMEM, "hipMalloc of %zu bytes for a %dx%dx%d lattice failed", bytes, in.nx, in.ny, in.nz);
    }
    h.own_mem = true;
  }
  lattice_layout(h, in, storage, mem);
  h.tiles = tiles;
  if (h.tiles.empty()) h.tiles.push_back(Tile{0, in.nx, 0, in.ny, 0, in.nz});
  rc = lattice_upload_inlet(h, stream);
  if (rc) {
    if (h.own_mem) (void)hipFree(mem);
    return rc;
  }
  lats.push_back(h);
  tables_dirty = true;
  return 0;
}

void Batch::remove_lattice(int idx) {
  if (idx < 0 || idx >= (int)lats.size()) return;
  if (lats[idx].own_mem && lats[idx].mem) (void)hipFree(lats[idx].mem);
  if (lats[idx].mask_mem) (void)hipFree(lats[idx].mask_mem);
  lats.erase(lats.begin() + idx);
  tables_dirty = true;
}

// An executable graph may still be running on the stream (its replays are
// only enqueued), so the stream drains before the graphs are destroyed.
void Batch::destroy_graphs() {
  bool any = false;
  for (auto& g : graph) any = any || g != nullptr;
  if (any && stream) (void)hipStreamSynchronize(stream);
  for (auto& g : graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  graph_steps_built = 0;
}

long long Batch::total_nodes() const {
  long long s = 0;
  for (const auto& h : lats) s += (long long)h.dev.n;
  return s;
}

size_t Batch::device_bytes() const {
  size_t s = 0;
  for (const auto& h : lats) s += h.bytes;
  return s;
}

int Batch::set_tuning(const nflow_tuning& t) {
  if (t.block != 64 && t.block != 128 && t.block != 256 && t.block != 512)
    return fail(NFLOW_EINVAL, "tuning.block must be 64, 128, 256 or 512");
  if (t.persistent < 0 || t.graph_steps < 0 || (t.graph_steps & 1))
    return fail(NFLOW_EINVAL, "tuning.persistent must be >= 0 and graph_steps even and >= 0");
  if (t.temporal < -1) return fail(NFLOW_EINVAL, "tuning.temporal must be >= -1");
  int rc = settle_pair();
  if (rc) return rc;
  tuning = t;
  tables_dirty = true;
  destroy_graphs();
  return 0;
}

// Work list: each tile's nodes as runs of consecutive node ids (whole slabs
// merge into one run), cut into workgroup-sized items.  With xcd_order the
// items are dealt so that the workgroups of XCD k (dispatch id = k mod 8)
// sweep one contiguous eighth of the list, keeping the y/z neighbour planes
// of a workgroup in the same XCD's L2.
int Batch::upload_tables() {
  if (!tables_dirty) return 0;
  NFLOW_HIP(hipStreamSynchronize(stream));
  destroy_graphs();
  const int bs = tuning.block;  // nodes per work item
  std::vector<WorkItem> items;
  std::vector<LatticeDev> devs(lats.size());
  masked = false;
  sr_any = false;
  for (size_t li = 0; li < lats.size(); ++li) {
    const HostLattice& h = lats[li];
    devs[li] = h.dev;
    if (h.dev.mask) masked = true;
    if (h.dev.sr) sr_any = true;
    const uint32_t nx = (uint32_t)h.in.nx, nxz = (uint32_t)h.in.nx * (uint32_t)h.in.nz;
    auto emit = [&](uint32_t a, uint32_t b) {
      for (uint32_t s = a; s < b; s += (uint32_t)bs)
        items.push_back(WorkItem{(int)li, s, std::min<uint32_t>((uint32_t)bs, b - s), 0u});
    };
    for (const Tile& t : h.tiles) {
      if (t.x0 == 0 && t.x1 == h.in.nx && t.z0 == 0 && t.z1 == h.in.nz) {
        emit((uint32_t)t.y0 * nxz, (uint32_t)t.y1 * nxz);
      } else if (t.x0 == 0 && t.x1 == h.in.nx) {
        for (int y = t.y0; y < t.y1; ++y) emit(y * nxz + t.z0 * nx, y * nxz + t.z1 * nx);
      } else {
        for (int y = t.y0; y < t.y1; ++y)
          for (int z = t.z0; z < t.z1; ++z) emit(y * nxz + z * nx + t.x0, y * nxz + z * nx + t.x1);
      }
    }
  }
  if (tuning.xcd_order && items.size() > 8) {
    const size_t m = items.size();
    std::vector<std::vector<WorkItem>> chunk(8);
    for (int k = 0; k < 8; ++k) chunk[k].assign(items.begin() + m * k / 8, items.begin() + m * (k + 1) / 8);
    size_t maxlen = 0;
    for (auto& c : chunk) maxlen = std::max(maxlen, c.size());
    std::vector<WorkItem> dealt(8 * maxlen, WorkItem{0, 0u, 0u, 0u});
    for (int k = 0; k < 8; ++k)
      for (size_t j = 0; j < chunk[k].size(); ++j) dealt[8 * j + k] = chunk[k][j];
    items.swap(dealt);
  }
  nitems = (int)items.size();
  if (tuning.persistent > 0) {
    grid = tuning.persistent * cu_count;
    grid = (grid + 7) / 8 * 8;
    grid = std::min(grid, std::max(nitems, 1));
  } else {
    grid = nitems;
  }
  if (devs.size() > d_lats_cap) {
    (void)hipFree(d_lats);
    d_lats = nullptr;
    NFLOW_HIP(hipMalloc(&d_lats, devs.size() * sizeof(LatticeDev)));
    d_lats_cap = devs.size();
  }
  if (items.size() > d_items_cap) {
    (void)hipFree(d_items);
    d_items = nullptr;
    NFLOW_HIP(hipMalloc(&d_items, items.size() * sizeof(WorkItem)));
    d_items_cap = items.size();
  }
  if (!devs.empty())
    NFLOW_HIP(hipMemcpy(d_lats, devs.data(), devs.size() * sizeof(LatticeDev), hipMemcpyHostToDevice));
  if (!items.empty())
    NFLOW_HIP(hipMemcpy(d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice));
  tables_dirty = false;
  return 0;
}

int Batch::init_lattice(int idx, const float* d_src, long long step0) {
  HostLattice& h = lats[idx];
  // counters and divergence flag start afresh (k_init sets the counter)
  NFLOW_HIP(hipMemsetAsync(h.dev.ctr, 0, CTR_WORDS * sizeof(uint32_t), stream));
  NFLOW_HIP(launch_init(storage, stream, h.dev, gparity, d_src, (uint32_t)step0));
  h.cur = gparity;
  h.steps_done = step0;
  h.prev_streamed = true;
  h.pair_last = false;
  h.initialized = true;
  return 0;
}

int Batch::init_all() {
  for (size_t i = 0; i < lats.size(); ++i) {
    int rc = init_lattice((int)i, nullptr);
    if (rc) return rc;
  }
  return 0;
}

int Batch::build_graph(int parity) {
  hipGraph_t g = nullptr;
  NFLOW_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int p = parity;
  hipError_t err = hipSuccess;
  for (int s = 0; s < tuning.graph_steps && err == hipSuccess; ++s) {
    err = launch_step(storage, masked, sr_any, tuning.block, grid, stream, d_lats, (int)lats.size(), d_items, nitems, p);
    p ^= 1;
  }
  hipError_t e2 = hipStreamEndCapture(stream, &g);
  if (err != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(NFLOW_EDEVICE, "graph capture failed: %s", hipGetErrorString(err != hipSuccess ? err : e2));
  }
  hipError_t e3 = hipGraphInstantiate(&graph[parity], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) return fail(NFLOW_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e3));
  graph_steps_built = tuning.graph_steps;
  return 0;
}

int Batch::launch_steps(int nsteps) {
  if (nsteps < 0) return fail(NFLOW_EINVAL, "nsteps must be >= 0");
  for (const auto& h : lats)
    if (!h.initialized) return fail(NFLOW_ESTATE, "run before initialize");
  int rc = upload_tables();
  if (rc) return rc;
  int left = nsteps;
  // two steps per launch: one FP32 cavity or channel lattice without solid
  // nodes.  The read-back streams the previous step's state, which a pair's
  // intermediate state is not (it never reaches memory): an odd count ends
  // with one single step, and an even count with a pair, which the read-back
  // re-runs as two single steps (settle_pair) -- unless the previous run's
  // state was read back, when an even count ends with two single steps
  // (a caller that reads back after every run pays for no pair twice)
  const int h2 = step2_height();
  const bool end_single = read_back;
  read_back = false;
  bool pair_last = false;
  if (h2 > 0 && left >= 2) {
    const int pairs = (left % 2 || end_single) ? (left - 1) / 2 : left / 2;
    for (int k = 0; k < pairs; ++k) {
      NFLOW_HIP(launch_step2(storage, stream, lats[0].dev, gparity, h2));
      gparity ^= 1;
      ++launches;
    }
    left -= 2 * pairs;
    pair_last = pairs > 0 && left == 0;
  }
  const int gs = tuning.graph_steps;
  if (gs > 0 && left >= gs && nitems > 0) {
    if (graph_steps_built != gs) destroy_graphs();
    if (!graph[gparity]) {
      rc = build_graph(gparity);
      if (rc) return rc;
    }
    const int reps = left / gs;
    for (int r = 0; r < reps; ++r) NFLOW_HIP(hipGraphLaunch(graph[gparity], stream));
    left -= reps * gs;
    launches += (long long)reps * gs;
  }
  for (int s = 0; s < left; ++s) {
    NFLOW_HIP(launch_step(storage, masked, sr_any, tuning.block, grid, stream, d_lats, (int)lats.size(), d_items, nitems, gparity));
    gparity ^= 1;
    ++launches;
  }
  for (auto& h : lats) {
    h.steps_done += nsteps;
    if (nsteps > 0) {
      h.prev_streamed = false;
      h.pair_last = pair_last;
    }
    h.cur = gparity;
  }
  return 0;
}

int Batch::settle_pair() {
  if (lats.empty() || !lats[0].pair_last) return 0;
  // a pair is launched for one lattice only, and every change of the lattice
  // set or the tuning settles first, so the tables are those of that launch
  if (tables_dirty || lats.size() != 1) return fail(NFLOW_ESTATE, "settle_pair: tables changed after a two-step launch");
  const int p = gparity ^ 1;  // the parity the pair was launched with
  NFLOW_HIP(launch_step(storage, masked, sr_any, tuning.block, grid, stream, d_lats, 1, d_items, nitems, p));
  NFLOW_HIP(launch_step(storage, masked, sr_any, tuning.block, grid, stream, d_lats, 1, d_items, nitems, p ^ 1));
  launches += 2;
  gparity = p;
  lats[0].cur = p;
  lats[0].pair_last = false;
  return 0;
}

int Batch::run(int nsteps) {
  NFLOW_HIP(hipEventRecord(ev0, stream));
  int rc = launch_steps(nsteps);
  if (rc) return rc;
  NFLOW_HIP(hipEventRecord(ev1, stream));
  last_steps = nsteps;
  return 0;
}

int Batch::synchronize() {
  NFLOW_HIP(hipStreamSynchronize(stream));
  return 0;
}

int Batch::timing(double* ms_per_step, int* steps) {
  NFLOW_HIP(hipEventSynchronize(ev1));
  float ms = 0.0f;
  NFLOW_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  if (ms_per_step) *ms_per_step = last_steps > 0 ? (double)ms / last_steps : 0.0;
  if (steps) *steps = last_steps;
  return 0;
}

int Batch::ensure_scratch(size_t aos, size_t macro, size_t partials, size_t types) {
  NFLOW_HIP(hipStreamSynchronize(stream));
  if (aos > d_aos_cap) {
    (void)hipFree(d_aos);
    d_aos = nullptr;
    if (hipMalloc(&d_aos, aos * sizeof(float)) != hipSuccess) return fail(NFLOW_ENOMEM, "export scratch");
    d_aos_cap = aos;
  }
  if (macro > d_macro_cap) {
    (void)hipFree(d_macro);
    d_macro = nullptr;
    if (hipMalloc(&d_macro, macro * sizeof(float)) != hipSuccess) return fail(NFLOW_ENOMEM, "export scratch");
    d_macro_cap = macro;
  }
  if (partials > d_partials_cap) {
    (void)hipFree(d_partials);
    d_partials = nullptr;
    if (hipMalloc(&d_partials, partials * sizeof(double)) != hipSuccess) return fail(NFLOW_ENOMEM, "export scratch");
    d_partials_cap = partials;
  }
  if (!d_out5) NFLOW_HIP(hipMalloc(&d_out5, 8 * sizeof(double)));
  if (types > d_types_cap) {
    (void)hipFree(d_types);
    d_types = nullptr;
    if (hipMalloc(&d_types, types * sizeof(int)) != hipSuccess) return fail(NFLOW_ENOMEM, "export scratch");
    d_types_cap = types;
  }
  return 0;
}

static size_t padded_count(const nflow_input& in) {
  return (size_t)(in.nx + 2) * (in.ny + 2) * (in.nz + 2) * 19;
}

int Batch::diverged_flags(std::vector<int>& out) {
  out.assign(lats.size(), 0);
  if (lats.empty()) return 0;
  int rc = upload_tables();
  if (rc) return rc;
  if (lats.size() > d_flags_cap) {
    (void)hipFree(d_flags);
    d_flags = nullptr;
    if (hipMalloc(&d_flags, lats.size() * sizeof(uint32_t)) != hipSuccess) {
      (void)hipGetLastError();
      d_flags_cap = 0;
      return fail(NFLOW_ENOMEM, "divergence flags");
    }
    d_flags_cap = lats.size();
  }
  h_flags.resize(lats.size());
  NFLOW_HIP(launch_probe(storage, stream, d_lats, d_items, nitems, gparity));
  NFLOW_HIP(launch_flags(stream, d_lats, (int)lats.size(), d_flags));
  NFLOW_HIP(hipMemcpyAsync(h_flags.data(), d_flags, lats.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  NFLOW_HIP(hipStreamSynchronize(stream));
  for (size_t i = 0; i < lats.size(); ++i) out[i] = h_flags[i] ? 1 : 0;
  return 0;
}

int Batch::export_one(int idx, float* aos, float* rho, float* ux, float* uy, float* uz, double* diag5,
                      int* diverged, bool consume) {
  if (idx < 0 || idx >= (int)lats.size()) return fail(NFLOW_EINVAL, "lattice index out of range");
  HostLattice& h = lats[idx];
  if (!h.initialized) return fail(NFLOW_ESTATE, "read before initialize");
  if (int rc = settle_pair()) return rc;
  read_back = true;
  const size_t n = h.dev.n;
  const size_t nb = (n + export_block_size() - 1) / export_block_size();
  float* macro = nullptr;
  double *partials = nullptr, *out5 = nullptr;
  // consume: the export reads buf[cur ^ 1] and the side records only, so the
  // look-ahead state in buf[cur] (>= 38 bytes per node) can hold rho/u and the
  // FP64 partials (16.2 bytes per node) when the lattice is not stepped again
  const size_t b_macro = round_up(4 * n * sizeof(float), 256), b_part = round_up(5 * nb * sizeof(double), 256);
  const size_t pop_bytes = (size_t)(static_cast<char*>(h.dev.buf[1]) - static_cast<char*>(h.dev.buf[0]));
  if (consume && !aos && b_macro + b_part + 256 <= pop_bytes) {
    char* base = static_cast<char*>(h.dev.buf[h.cur]);
    macro = reinterpret_cast<float*>(base);
    partials = reinterpret_cast<double*>(base + b_macro);
    out5 = reinterpret_cast<double*>(base + b_macro + b_part);
  } else {
    int rc = ensure_scratch(aos ? padded_count(h.in) : 0, rho ? 4 * n : 0, diag5 ? 5 * nb : 0, 0);
    if (rc) return rc;
    macro = d_macro;
    partials = d_partials;
    out5 = d_out5;
  }
  if (aos) NFLOW_HIP(hipMemsetAsync(d_aos, 0, padded_count(h.in) * sizeof(float), stream));
  const int prev = h.cur ^ 1;
  const int prev_streamed = h.prev_streamed ? 1 : 0;
  int nblocks = 0;
  float* m = rho ? macro : nullptr;
  NFLOW_HIP(launch_export(storage, stream, h.dev, prev, prev_streamed, aos ? d_aos : nullptr, m, m ? m + n : nullptr,
                          m ? m + 2 * n : nullptr, m ? m + 3 * n : nullptr, diag5 ? partials : nullptr, &nblocks));
  if (diag5) NFLOW_HIP(launch_reduce(stream, partials, nblocks, out5));
  if (aos) NFLOW_HIP(hipMemcpyAsync(aos, d_aos, padded_count(h.in) * sizeof(float), hipMemcpyDeviceToHost, stream));
  if (rho) {
    NFLOW_HIP(hipMemcpyAsync(rho, m, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    NFLOW_HIP(hipMemcpyAsync(ux, m + n, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    NFLOW_HIP(hipMemcpyAsync(uy, m + 2 * n, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    NFLOW_HIP(hipMemcpyAsync(uz, m + 3 * n, n * sizeof(float), hipMemcpyDeviceToHost, stream));
  }
  if (diag5) NFLOW_HIP(hipMemcpyAsync(diag5, out5, 5 * sizeof(double), hipMemcpyDeviceToHost, stream));
  uint32_t flag = 0;
  if (diverged)
    NFLOW_HIP(hipMemcpyAsync(&flag, h.dev.ctr + CTR_DIVERGED, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  NFLOW_HIP(hipStreamSynchronize(stream));
  if (diverged) *diverged = flag ? 1 : 0;
  return 0;
}

int Batch::import_one(int idx, const float* aos_host, long long step0) {
  if (idx < 0 || idx >= (int)lats.size()) return fail(NFLOW_EINVAL, "lattice index out of range");
  HostLattice& h = lats[idx];
  int rc = ensure_scratch(padded_count(h.in), 0, 0, 0);
  if (rc) return rc;
  NFLOW_HIP(hipMemcpyAsync(d_aos, aos_host, padded_count(h.in) * sizeof(float), hipMemcpyHostToDevice, stream));
  rc = init_lattice(idx, d_aos, step0);
  if (rc) return rc;
  NFLOW_HIP(hipStreamSynchronize(stream));
  return 0;
}

int Batch::node_types(int idx, int* type, int* dirstate) {
  if (idx < 0 || idx >= (int)lats.size()) return fail(NFLOW_EINVAL, "lattice index out of range");
  HostLattice& h = lats[idx];
  const size_t n = h.dev.n;
  int rc = ensure_scratch(0, 0, 0, 2 * n);
  if (rc) return rc;
  NFLOW_HIP(launch_node_type(stream, h.dev, d_types, d_types + n));
  NFLOW_HIP(hipMemcpyAsync(type, d_types, n * sizeof(int), hipMemcpyDeviceToHost, stream));
  NFLOW_HIP(hipMemcpyAsync(dirstate, d_types + n, n * sizeof(int), hipMemcpyDeviceToHost, stream));
  NFLOW_HIP(hipStreamSynchronize(stream));
  return 0;
}

int Batch::step2_height(const nflow_tuning* t) const {
  const nflow_tuning& tu = t ? *t : tuning;
  if (tu.temporal == 0 || storage != NFLOW_STORAGE_F32 || lats.size() != 1) return 0;
  const HostLattice& h = lats[0];
  if (h.in.casetype == NFLOW_PIPE || h.dev.mask) return 0;
  // a k_step2 workgroup reads planes ya - 2 .. ya + H + 1 through a window of
  // 32-bit byte offsets
  auto fits = [&](int hh) {
    return ((unsigned long long)(hh + 4) * (unsigned long long)h.dev.nxz + 64) * 19ull * 4ull < (1ull << 32);
  };
  if (tu.temporal > 0) return fits(tu.temporal) ? tu.temporal : 0;
  // automatic: the chunk height of the least estimated launch time, and two
  // steps per launch only where that estimate beats 40 G node updates per
  // second (one step per launch runs at 37-39.5 from 128^3 up).  Estimate:
  // one workgroup per CU (its LDS), each sweeping its chunk's planes plus
  // 0.86 of halo and pipeline overhead at 12.18 + 2.70 w us per plane, w the
  // share of the CUs busy in its round (memory contention); times
  // 0.84 + 0.16 fill for the idle lanes of partial x tiles (fill = nx over
  // the tiles' width) and 1.21 for a channel's boundary records.  Fitted to
  // the measured rates of 34 cavities and channels from 48^3 to 512^3 at
  // heights 8-64 (8% rms); it picks the fastest height in every one of them.
  const int ntx = (h.in.nx + STEP2_TX - 1) / STEP2_TX;
  const long cols = (long)ntx * ((h.in.nz + STEP2_TZ - 1) / STEP2_TZ);
  const double fill = (double)h.in.nx / ((double)ntx * STEP2_TX);
  const double scale = (0.841 + 0.159 * fill) * (h.in.casetype == NFLOW_CHANNEL ? 1.213 : 1.0);
  const double nodes = (double)h.in.nx * h.in.ny * h.in.nz;
  double best_us = 2.0 * nodes / 40.0e3;  // two steps at 40 nodes per ns
  int best = 0, prev = 0;
  for (int cand : {8, 12, 16, 24, 32, 48, 64}) {
    const int hh = std::min(cand, h.in.ny);
    if (hh == prev || !fits(hh)) continue;
    prev = hh;
    const long nch = (h.in.ny + hh - 1) / hh;
    const double planes = (double)h.in.ny / nch + 0.858;
    double us = 0.0;
    for (long left = cols * nch; left > 0; left -= cu_count)
      us += planes * (12.184 + 2.696 * (double)std::min<long>(left, cu_count) / cu_count);
    us *= scale;
    if (us < best_us) {
      best_us = us;
      best = hh;
    }
  }
  return best;
}

// Times every candidate launch shape on this batch's own lattices and keeps
// the fastest.  The clocks of an issue-bound kernel settle over the first
// few hundred steps, which would favour whichever shapes run first, so the
// device is warmed up with the first candidate and the candidates are timed
// twice, in opposite orders, each keeping its better time.  The trial steps
// advance the state, so every lattice is re-initialised afterwards (call
// before importing a state).
int Batch::autotune(int trial_steps, nflow_tuning* out) {
  if (trial_steps < 2) trial_steps = 2;
  const nflow_tuning base = tuning;
  std::vector<nflow_tuning> cand;
  for (int bs : {64, 128, 256, 512})
    for (int xo = 0; xo < 2; ++xo)
      for (int pe : {0, 2, 4, 8}) {
        nflow_tuning t = base;
        t.block = bs;
        t.xcd_order = xo;
        t.persistent = pe;
        t.graph_steps = 0;
        t.temporal = 0;
        t.ms_per_step = 1e30;
        cand.push_back(t);
      }
  // initialize before set_tuning: a fresh state has no two-step launch for
  // set_tuning to settle
  auto trial = [&](nflow_tuning& t, int steps) -> int {
    int rc = init_all();
    if (!rc) rc = set_tuning(t);
    if (!rc) rc = launch_steps(2);
    if (!rc) rc = run(steps);
    double ms = 0.0;
    if (!rc) rc = timing(&ms, nullptr);
    if (!rc) t.ms_per_step = std::min(t.ms_per_step, ms);
    return rc;
  };
  nflow_tuning warm = cand.back();
  int rc = trial(warm, 8 * trial_steps);
  if (rc) return rc;
  for (int pass = 0; pass < 2; ++pass)
    for (size_t k = 0; k < cand.size(); ++k) {
      rc = trial(cand[pass == 0 ? k : cand.size() - 1 - k], trial_steps);
      if (rc) return rc;
    }
  nflow_tuning best = cand[0];
  for (const auto& t : cand)
    if (t.ms_per_step < best.ms_per_step) best = t;
  // two steps per launch (where the batch allows it): the automatic chunk
  // height and a few fixed ones, each timed over pairs with the best shape
  nflow_tuning probe = best;
  probe.temporal = 16;
  if (step2_height(&probe) > 0) {
    probe.temporal = -1;
    std::vector<int> hs = {8, 16, 24, 32, 64};
    const int ah = step2_height(&probe);
    if (ah > 0 && std::find(hs.begin(), hs.end(), ah) == hs.end()) hs.push_back(ah);
    std::vector<nflow_tuning> tc;
    for (int hh : hs) {
      nflow_tuning t = best;
      t.temporal = hh;
      t.ms_per_step = 1e30;
      tc.push_back(t);
    }
    for (int pass = 0; pass < 2; ++pass)
      for (size_t k = 0; k < tc.size(); ++k) {
        rc = trial(tc[pass == 0 ? k : tc.size() - 1 - k], 4 * trial_steps + 1);
        if (rc) return rc;
      }
    for (const auto& t : tc)
      if (t.ms_per_step < best.ms_per_step) best = t;
  }
  best.graph_steps = base.graph_steps;
  rc = init_all();
  if (rc) return rc;
  rc = set_tuning(best);
  if (rc) return rc;
  tuning.ms_per_step = best.ms_per_step;
  NFLOW_HIP(hipStreamSynchronize(stream));
  if (out) *out = tuning;
  return 0;
}

}  // namespace nflow

// ===========================================================================
// C ABI
// ===========================================================================
using namespace nflow;

struct nflow_solver {
  Batch b;
};
struct nflow_batch {
  Batch b;
};

extern "C" const char* nflow_last_error(void) { return last_error_slot().c_str(); }

// Tiles must cover [0,nx) x [0,ny) x [0,nz) exactly once; empty tiles are
// skipped as slice_chunk's callers do (runSimulation.cpp:38).
static int tiles_from_slice(const nflow_input& in, int ntile, const int slice[][3][2], std::vector<Tile>& tiles) {
  tiles.clear();
  if (ntile <= 0 || !slice) {
    tiles.push_back(Tile{0, in.nx, 0, in.ny, 0, in.nz});
    return 0;
  }
  long long vol = 0;
  for (int t = 0; t < ntile; ++t) {
    Tile tl{slice[t][0][0], slice[t][0][1], slice[t][1][0], slice[t][1][1], slice[t][2][0], slice[t][2][1]};
    if (tl.x1 <= tl.x0 || tl.y1 <= tl.y0 || tl.z1 <= tl.z0) continue;
    if (tl.x0 < 0 || tl.y0 < 0 || tl.z0 < 0 || tl.x1 > in.nx || tl.y1 > in.ny || tl.z1 > in.nz)
      return fail(NFLOW_EINVAL, "tile %d lies outside the %dx%dx%d box", t, in.nx, in.ny, in.nz);
    vol += (long long)(tl.x1 - tl.x0) * (tl.y1 - tl.y0) * (tl.z1 - tl.z0);
    tiles.push_back(tl);
  }
  const long long n = (long long)in.nx * in.ny * in.nz;
  if (vol != n) return fail(NFLOW_EINVAL, "tiles cover %lld nodes, the box has %lld", vol, n);
  // equal volume + no overlap => exact partition
  std::vector<uint8_t> seen((size_t)n, 0);
  for (const Tile& tl : tiles)
    for (int y = tl.y0; y < tl.y1; ++y)
      for (int z = tl.z0; z < tl.z1; ++z)
        for (int x = tl.x0; x < tl.x1; ++x) {
          uint8_t& s = seen[(size_t)x + (size_t)in.nx * ((size_t)z + (size_t)in.nz * y)];
          if (s) return fail(NFLOW_EINVAL, "tiles overlap at node (%d, %d, %d)", x, y, z);
          s = 1;
        }
  return 0;
}

extern "C" int nflow_create(const nflow_input* in, int ntile, const int slice[][3][2], nflow_solver** out) {
  if (!in || !out) return fail(NFLOW_EINVAL, "nflow_create: null argument");
  *out = nullptr;
  int rc = validate_input(*in, in->storage);
  if (rc) return rc;
  std::vector<Tile> tiles;
  rc = tiles_from_slice(*in, ntile, slice, tiles);
  if (rc) return rc;
  nflow_solver* s = new nflow_solver();
  s->b.storage = in->storage;
  // one lattice: 512-node work items suit 16-bit storage (256^3 MRT-LES 2-5%
  // ahead of 256; FP32 within 1%); batches start at 256 and are autotuned,
  // the server fixes its own shape
  if (in->storage != NFLOW_STORAGE_F32) s->b.tuning.block = 512;
  rc = s->b.open();
  if (!rc) rc = s->b.add_lattice(*in, tiles, nullptr);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return 0;
}

extern "C" int nflow_initialize(nflow_solver* s) {
  if (!s) return fail(NFLOW_EINVAL, "null solver");
  int rc = s->b.init_all();
  if (rc) return rc;
  return s->b.synchronize();
}

extern "C" int nflow_run(nflow_solver* s, int nsteps) {
  if (!s) return fail(NFLOW_EINVAL, "null solver");
  return s->b.run(nsteps);
}

extern "C" int nflow_synchronize(nflow_solver* s) {
  if (!s) return fail(NFLOW_EINVAL, "null solver");
  return s->b.synchronize();
}

extern "C" long long nflow_steps_done(const nflow_solver* s) { return s ? s->b.lats[0].steps_done : -1; }

extern "C" int nflow_read_arF(nflow_solver* s, float* dst) {
  if (!s || !dst) return fail(NFLOW_EINVAL, "nflow_read_arF: null argument");
  return s->b.export_one(0, dst, nullptr, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int nflow_write_arF(nflow_solver* s, const float* src) {
  if (!s || !src) return fail(NFLOW_EINVAL, "nflow_write_arF: null argument");
  return s->b.import_one(0, src, 0);
}

extern "C" int nflow_read_macro(nflow_solver* s, float* rho, float* ux, float* uy, float* uz) {
  if (!s || !rho || !ux || !uy || !uz) return fail(NFLOW_EINVAL, "nflow_read_macro: null argument");
  return s->b.export_one(0, nullptr, rho, ux, uy, uz, nullptr);
}

extern "C" int nflow_read_node_type(nflow_solver* s, int* node_type, int* direction_state) {
  if (!s || !node_type || !direction_state) return fail(NFLOW_EINVAL, "nflow_read_node_type: null argument");
  return s->b.node_types(0, node_type, direction_state);
}

extern "C" int nflow_set_solid_mask(nflow_solver* s, const unsigned char* mask) {
  if (!s) return fail(NFLOW_EINVAL, "null solver");
  Batch& b = s->b;
  HostLattice& h = b.lats[0];
  if (h.initialized) return fail(NFLOW_ESTATE, "nflow_set_solid_mask: call before nflow_initialize");
  NFLOW_HIP(hipStreamSynchronize(b.stream));
  b.tables_dirty = true;
  if (!mask) {  // back to the case's own geometry
    if (h.mask_mem) {
      (void)hipFree(h.mask_mem);
      h.mask_mem = nullptr;
      h.dev.mask = nullptr;
    }
    return h.in.casetype == NFLOW_PIPE ? lattice_upload_inlet(h, b.stream) : 0;
  }
  if (!h.dev.mask) {
    if (hipMalloc(&h.mask_mem, (size_t)h.dev.n) != hipSuccess) {
      (void)hipGetLastError();
      h.mask_mem = nullptr;
      return fail(NFLOW_ENOMEM, "hipMalloc of a %u-byte solid mask failed", h.dev.n);
    }
    h.dev.mask = static_cast<const uint8_t*>(h.mask_mem);
  }
  NFLOW_HIP(hipMemcpy(const_cast<uint8_t*>(h.dev.mask), mask, (size_t)h.dev.n, hipMemcpyHostToDevice));
  return 0;
}

// Fills the FP64 sums of lattice idx; NFLOW_EDIVERGED when the lattice's
// divergence flag is set or a sum is not finite.
static int diagnostics_of(Batch& b, int idx, double* mass, double mom[3], double* drho_l1) {
  double d[5];
  int diverged = 0;
  int rc = b.export_one(idx, nullptr, nullptr, nullptr, nullptr, nullptr, d, &diverged);
  if (rc) return rc;
  *mass = d[0];
  mom[0] = d[1];
  mom[1] = d[2];
  mom[2] = d[3];
  *drho_l1 = d[4];
  bool finite = true;
  for (double v : d) finite = finite && std::isfinite(v);
  if (diverged || !finite)
    return fail(NFLOW_EDIVERGED, "lattice %d diverged: a node's density was not a positive finite number (mass %g)",
                idx, d[0]);
  return 0;
}

extern "C" int nflow_diagnostics(nflow_solver* s, double* mass, double mom[3], double* drho_l1) {
  if (!s || !mass || !mom || !drho_l1) return fail(NFLOW_EINVAL, "nflow_diagnostics: null argument");
  return diagnostics_of(s->b, 0, mass, mom, drho_l1);
}

// ---- diagnostics without allocation -----------------------------------------
// The FP64 block partials of the export and the five reduced sums live in a
// file-scope device array instead of library-allocated scratch.  Its name is a
// host-side stand-in: the array's address on a device comes from
// hipGetSymbolAddress, called once per device (with that device current) on
// the first call there, and only that address reaches kernels and copies.
// Every call on a device shares its array; each call rewrites the partials of
// the blocks it reduces and the five sums, so nothing needs clearing, but
// calls that overlap on one device race.
namespace {
constexpr size_t NOALLOC_MAX_NODES = size_t(1) << 25;  // 33,554,432 nodes
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace
static __device__ __align__(256) double scratch[(NOALLOC_MAX_NODES / 256) * 5 + 8];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

int scratch_on(int device, double** out) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES)
    return fail(NFLOW_EINVAL, "device %d out of the %d devices the static scratch supports", device,
                NOALLOC_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    NFLOW_HIP(hipGetDevice(&cur));
    if (cur != device) NFLOW_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return fail(NFLOW_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return 0;
}
}  // namespace

extern "C" int nflow_diagnostics_noalloc(nflow_solver* s, double* mass, double mom[3], double* drho_l1) {
  if (!s || !mass || !mom || !drho_l1) return fail(NFLOW_EINVAL, "nflow_diagnostics_noalloc: null argument");
  Batch& b = s->b;
  const HostLattice& h = b.lats[0];
  if (!h.initialized) return fail(NFLOW_ESTATE, "read before initialize");
  if ((size_t)h.dev.n > NOALLOC_MAX_NODES)
    return fail(NFLOW_EINVAL,
                "nflow_diagnostics_noalloc: the lattice has %u nodes, more than the %zu its static scratch holds "
                "(nflow_diagnostics has no limit)",
                h.dev.n, NOALLOC_MAX_NODES);
  double* ptr = nullptr;
  int rc = scratch_on(b.device, &ptr);
  if (rc) return rc;
  rc = b.settle_pair();
  if (rc) return rc;
  b.read_back = true;
  double* const partials = ptr;
  double* const sums = ptr + (NOALLOC_MAX_NODES / 256) * 5;
  // the kernels, grid and order of nflow_diagnostics: the same sums, bit for
  // bit; n <= 2^25 keeps the export's ceil(n / 256) block partials inside
  int nblocks = 0;
  NFLOW_HIP(launch_export(b.storage, b.stream, h.dev, h.cur ^ 1, h.prev_streamed ? 1 : 0, nullptr, nullptr, nullptr,
                          nullptr, nullptr, partials, &nblocks));
  NFLOW_HIP(launch_reduce(b.stream, partials, nblocks, sums));
  double d[5];
  uint32_t flag = 0;
  NFLOW_HIP(hipMemcpyAsync(d, sums, sizeof(d), hipMemcpyDeviceToHost, b.stream));
  NFLOW_HIP(hipMemcpyAsync(&flag, h.dev.ctr + CTR_DIVERGED, sizeof(flag), hipMemcpyDeviceToHost, b.stream));
  NFLOW_HIP(hipStreamSynchronize(b.stream));
  *mass = d[0];
  mom[0] = d[1];
  mom[1] = d[2];
  mom[2] = d[3];
  *drho_l1 = d[4];
  bool finite = true;
  for (double v : d) finite = finite && std::isfinite(v);
  if (flag || !finite)
    return fail(NFLOW_EDIVERGED, "lattice diverged: a node's density was not a positive finite number (mass %g)", d[0]);
  return 0;
}

extern "C" int nflow_timing(nflow_solver* s, double* ms_per_step, int* steps_timed) {
  if (!s) return fail(NFLOW_EINVAL, "null solver");
  return s->b.timing(ms_per_step, steps_timed);
}

extern "C" void nflow_destroy(nflow_solver* s) { delete s; }

// ---- checkpoint files (format and header parsing: checkpoint.hpp, io.cpp) ----
extern "C" int nflow_save_checkpoint(nflow_solver* s, const char* path) {
  if (!s || !path) return fail(NFLOW_EINVAL, "nflow_save_checkpoint: null argument");
  const HostLattice& L = s->b.lats[0];
  std::vector<float> a(padded_count(L.in));
  int rc = s->b.export_one(0, a.data(), nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  CkpHeader h{};
  std::memcpy(h.magic, CKP_MAGIC, 8);
  h.header_bytes = 64;
  h.version = CKP_VERSION;
  h.nx = L.in.nx;
  h.ny = L.in.ny;
  h.nz = L.in.nz;
  h.casetype = L.in.casetype;
  h.steps = L.steps_done;
  h.storage = L.in.storage;
  h.params = params_hash(L.in);
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(NFLOW_EINVAL, "cannot open %s for writing", path);
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(a.data(), sizeof(float), a.size(), f) == a.size();
  const bool closed = std::fclose(f) == 0;
  return ok && closed ? 0 : fail(NFLOW_EINVAL, "write to %s failed", path);
}

extern "C" int nflow_load_checkpoint(nflow_solver* s, const char* path) {
  if (!s || !path) return fail(NFLOW_EINVAL, "nflow_load_checkpoint: null argument");
  HostLattice& L = s->b.lats[0];
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(NFLOW_EINVAL, "cannot open %s", path);
  CkpHeader h;
  int rc = read_ckp_header(f, path, h);
  if (!rc && (h.nx != L.in.nx || h.ny != L.in.ny || h.nz != L.in.nz))
    rc = fail(NFLOW_EINVAL, "%s holds a %dx%dx%d lattice, the solver is %dx%dx%d", path, h.nx, h.ny, h.nz, L.in.nx,
              L.in.ny, L.in.nz);
  if (!rc && h.casetype != L.in.casetype)
    rc = fail(NFLOW_EINVAL, "%s holds casetype %d, the solver is casetype %d", path, h.casetype, L.in.casetype);
  if (!rc && h.version >= 2 && h.storage != L.in.storage)
    rc = fail(NFLOW_EINVAL, "%s was saved with storage %d, the solver uses storage %d", path, h.storage, L.in.storage);
  if (!rc && h.version >= 2 && h.params != params_hash(L.in))
    rc = fail(NFLOW_EINVAL,
              "%s was saved with other step parameters (tau, collision model, LES, lid/inlet, pipe or rounding)", path);
  if (!rc && (h.steps < 0 || h.steps > 0xFFFFFFFFll))
    rc = fail(NFLOW_EINVAL, "%s: step count %lld out of range", path, (long long)h.steps);
  std::vector<float> a;
  if (!rc) {
    a.resize(padded_count(L.in));
    if (std::fread(a.data(), sizeof(float), a.size(), f) != a.size())
      rc = fail(NFLOW_EINVAL, "%s: truncated population data", path);
  }
  std::fclose(f);
  if (rc) return rc;
  // the run continues from the saved step count (which also seeds the
  // stochastic rounding, so a restart continues it exactly)
  return s->b.import_one(0, a.data(), h.steps);
}

// ---- batch -----------------------------------------------------------------
extern "C" int nflow_batch_create(int count, const nflow_input* inputs, nflow_batch** out) {
  if (!inputs || !out || count < 1) return fail(NFLOW_EINVAL, "nflow_batch_create: bad argument");
  *out = nullptr;
  nflow_batch* b = new nflow_batch();
  b->b.storage = inputs[0].storage;
  int rc = b->b.open();
  for (int i = 0; i < count && !rc; ++i) rc = b->b.add_lattice(inputs[i], {}, nullptr);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return 0;
}

extern "C" int nflow_batch_initialize(nflow_batch* b) {
  if (!b) return fail(NFLOW_EINVAL, "null batch");
  int rc = b->b.init_all();
  return rc ? rc : b->b.synchronize();
}

extern "C" int nflow_batch_autotune(nflow_batch* b, int steps_per_trial, nflow_tuning* chosen) {
  if (!b) return fail(NFLOW_EINVAL, "null batch");
  return b->b.autotune(steps_per_trial, chosen);
}

extern "C" int nflow_batch_set_tuning(nflow_batch* b, const nflow_tuning* t) {
  if (!b || !t) return fail(NFLOW_EINVAL, "null argument");
  return b->b.set_tuning(*t);
}

extern "C" int nflow_batch_get_tuning(nflow_batch* b, nflow_tuning* t) {
  if (!b || !t) return fail(NFLOW_EINVAL, "null argument");
  *t = b->b.tuning;
  return 0;
}

extern "C" int nflow_set_tuning(nflow_solver* s, const nflow_tuning* t) {
  if (!s || !t) return fail(NFLOW_EINVAL, "null argument");
  return s->b.set_tuning(*t);
}

extern "C" int nflow_batch_step2_height(nflow_batch* b) {
  if (!b) return fail(NFLOW_EINVAL, "null batch");
  return b->b.step2_height();
}

extern "C" int nflow_batch_run(nflow_batch* b, int nsteps) {
  if (!b) return fail(NFLOW_EINVAL, "null batch");
  return b->b.run(nsteps);
}

extern "C" int nflow_batch_synchronize(nflow_batch* b) {
  if (!b) return fail(NFLOW_EINVAL, "null batch");
  return b->b.synchronize();
}

extern "C" int nflow_batch_read_arF(nflow_batch* b, int idx, float* dst) {
  if (!b || !dst) return fail(NFLOW_EINVAL, "null argument");
  return b->b.export_one(idx, dst, nullptr, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int nflow_batch_read_macro(nflow_batch* b, int idx, float* rho, float* ux, float* uy, float* uz) {
  if (!b || !rho || !ux || !uy || !uz) return fail(NFLOW_EINVAL, "null argument");
  return b->b.export_one(idx, nullptr, rho, ux, uy, uz, nullptr);
}

extern "C" int nflow_batch_diagnostics(nflow_batch* b, int idx, double* mass, double mom[3], double* drho_l1) {
  if (!b || !mass || !mom || !drho_l1) return fail(NFLOW_EINVAL, "null argument");
  return diagnostics_of(b->b, idx, mass, mom, drho_l1);
}

extern "C" int nflow_batch_diverged(nflow_batch* b, int* flags) {
  if (!b || !flags) return fail(NFLOW_EINVAL, "null argument");
  std::vector<int> f;
  int rc = b->b.diverged_flags(f);
  if (rc) return rc;
  std::copy(f.begin(), f.end(), flags);
  return 0;
}

extern "C" int nflow_batch_timing(nflow_batch* b, double* ms_per_step, int* steps_timed) {
  if (!b) return fail(NFLOW_EINVAL, "null batch");
  return b->b.timing(ms_per_step, steps_timed);
}

extern "C" long long nflow_batch_nodes(nflow_batch* b) { return b ? b->b.total_nodes() : -1; }
extern "C" size_t nflow_batch_device_bytes(nflow_batch* b) { return b ? b->b.device_bytes() : 0; }
extern "C" void nflow_batch_destroy(nflow_batch* b) { delete b; }

// ---- measurement helpers -----------------------------------------------------
extern "C" int nflow_copy_bandwidth(size_t bytes, int streams, int reps, double* gbps, double* ms_per_copy) {
  // streams = 0: the FP32 population layout (blocks of 19 runs of 64 floats)
  const bool runs = streams == 0;
  const size_t nblk = bytes / (19 * 64 * sizeof(float));
  if (streams < 0 || reps < 1 || (runs ? nblk == 0 : bytes < (size_t)streams * 16))
    return fail(NFLOW_EINVAL, "bad copy parameters");
  const int ns = runs ? 1 : streams;
  const size_t per = runs ? nblk * 19 * 64 : bytes / sizeof(float) / streams / 4 * 4;
  float *src = nullptr, *dst = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float ms = 0.0f;
  auto launch = [&]() { return runs ? launch_copy_runs(st, src, dst, nblk) : launch_copy_streams(st, src, dst, per, ns); };
  if (hipMalloc(&src, per * ns * sizeof(float)) != hipSuccess || hipMalloc(&dst, per * ns * sizeof(float)) != hipSuccess) {
    rc = fail(NFLOW_ENOMEM, "copy buffers");
    goto done;
  }
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "stream/event creation");
    goto done;
  }
  if (hipMemsetAsync(src, 0, per * ns * sizeof(float), st) != hipSuccess || launch() != hipSuccess ||
      hipEventRecord(e0, st) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "copy warm-up");
    goto done;
  }
  for (int r = 0; r < reps; ++r)
    if (launch() != hipSuccess) {
      rc = fail(NFLOW_EDEVICE, "copy launch");
      goto done;
    }
  if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "copy timing");
    goto done;
  }
  if (ms_per_copy) *ms_per_copy = ms / reps;
  if (gbps) *gbps = 2.0 * (double)per * ns * sizeof(float) / ((double)ms / reps * 1e-3) / 1e9;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}

extern "C" int nflow_valu_rate(int reps, double* g_wave_fma_per_s, double* ms_per_launch) {
  if (reps < 1) return fail(NFLOW_EINVAL, "reps must be >= 1");
  int dev = 0, cus = 0;
  NFLOW_HIP(hipGetDevice(&dev));
  NFLOW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = 16 * cus, iters = 2048;
  float* out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  float ms = 0.0f;
  if (hipMalloc(&out, (size_t)grid * 256 * sizeof(float)) != hipSuccess) {
    rc = fail(NFLOW_ENOMEM, "probe output");
    goto done;
  }
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "stream/event creation");
    goto done;
  }
  if (launch_fma_chains(st, out, grid, iters) != hipSuccess || hipEventRecord(e0, st) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "probe warm-up");
    goto done;
  }
  for (int r = 0; r < reps; ++r)
    if (launch_fma_chains(st, out, grid, iters) != hipSuccess) {
      rc = fail(NFLOW_EDEVICE, "probe launch");
      goto done;
    }
  if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
      hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "probe timing");
    goto done;
  }
  if (ms_per_launch) *ms_per_launch = ms / reps;
  if (g_wave_fma_per_s)  // grid * 4 waves, 64 FMAs per wave per iteration
    *g_wave_fma_per_s = (double)grid * 4 * 64 * iters / ((double)ms / reps * 1e-3) / 1e9;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(out);
  return rc;
}

extern "C" int nflow_round16(const float* src, const unsigned* seeds, size_t n, int storage, int stochastic,
                             unsigned short* dst) {
  if (!src || !dst || (stochastic && !seeds)) return fail(NFLOW_EINVAL, "nflow_round16: null argument");
  if (storage != NFLOW_STORAGE_F16 && storage != NFLOW_STORAGE_BF16)
    return fail(NFLOW_EINVAL, "nflow_round16: storage must be NFLOW_STORAGE_F16 or NFLOW_STORAGE_BF16");
  if (n == 0) return 0;
  if (n > (size_t)1 << 30) return fail(NFLOW_EINVAL, "nflow_round16: at most 2^30 values per call");
  float* d_src = nullptr;
  uint32_t* d_seeds = nullptr;
  uint16_t* d_dst = nullptr;
  hipStream_t st = nullptr;
  int rc = 0;
  if (hipMalloc(&d_src, n * sizeof(float)) != hipSuccess || hipMalloc(&d_dst, n * sizeof(uint16_t)) != hipSuccess ||
      (stochastic && hipMalloc(&d_seeds, n * sizeof(uint32_t)) != hipSuccess)) {
    (void)hipGetLastError();
    rc = fail(NFLOW_ENOMEM, "nflow_round16: device buffers");
  } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
             hipMemcpyAsync(d_src, src, n * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess ||
             (stochastic &&
              hipMemcpyAsync(d_seeds, seeds, n * sizeof(uint32_t), hipMemcpyHostToDevice, st) != hipSuccess) ||
             launch_round16(st, storage, d_src, d_seeds, d_dst, n, stochastic ? 1 : 0) != hipSuccess ||
             hipMemcpyAsync(dst, d_dst, n * sizeof(uint16_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess) {
    rc = fail(NFLOW_EDEVICE, "nflow_round16: %s", hipGetErrorString(hipGetLastError()));
  }
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(d_src);
  (void)hipFree(d_seeds);
  (void)hipFree(d_dst);
  return rc;
}

extern "C" int nflow_device_info(char* name, int name_len, int* cu_count, size_t* hbm_bytes) {
  int dev = 0;
  NFLOW_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  NFLOW_HIP(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    // some runtimes leave the marketing name empty: name the device by its
    // architecture, compute units and memory instead
    if (p.name[0])
      std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    else
      std::snprintf(name, (size_t)name_len, "AMD %s, %d CUs, %.0f GiB HBM", p.gcnArchName, p.multiProcessorCount,
                    (double)p.totalGlobalMem / (1024.0 * 1024.0 * 1024.0));
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
  return 0;
}
