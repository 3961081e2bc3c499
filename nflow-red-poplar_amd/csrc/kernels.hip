// kernels.hip -- CDNA4 (gfx950) kernels of the D3Q19 step.
//
// One fused kernel per time step: every thread owns one node, pulls the 19
// post-collision populations of the previous step from its neighbours
// (streaming, lbm_codelets.cpp:21-105), applies the wall / inlet / outlet
// rules in place for links whose donor lies outside the box
// (lbm_codelets.cpp:325-402, kernel.cu:566-657), forms the macroscopic
// quantities (lbm_codelets.cpp:248-318), the Smagorinsky relaxation time
// (lbm_codelets.cpp:558-668) and the SRT or MRT collision
// (kernel.cu:458-487, lbm_codelets.cpp:409-550), and writes its 19
// post-collision populations with unit-stride stores.  The reference runs
// the same maths as five (Poplar) or five-to-seven (CUDA) passes per step.
//
// Populations are stored in 256-byte runs of consecutive nodes (x fastest)
// per population (see NFLOW_AOSOA) in FP32, or as FP16 / BF16 of f - w_q
// with FP32 arithmetic.  Per node and step the kernel reads 19 and writes 19
// populations: 152 B at FP32, 76 B at FP16/BF16.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3q19.hpp"
#include "lattice.hpp"
#include "mrt.hpp"

// Compile-time memory-access policy of the population streams.  Every
// population is read once and written once per step and not touched again
// until the next step, after ~2.5 GB of other traffic at 256^3, so the
// streams are marked non-temporal.  Measured on MI355X at 256^3 FP32:
// 5.55 TB/s without the hint, 5.76 TB/s with non-temporal stores and
// 5.89 TB/s with non-temporal loads and stores (512-thread workgroups).
#ifndef NFLOW_NT_LOAD
#define NFLOW_NT_LOAD 1
#endif
#ifndef NFLOW_NT_STORE
#define NFLOW_NT_STORE 1
#endif
// Occupancy hint for the step kernel (0: compiler's choice).
#ifndef NFLOW_WAVES_PER_EU
#define NFLOW_WAVES_PER_EU 0
#endif

// Which object of the library this compilation builds: 1 = the FP32 kernels
// and the launch entry points, 2 = the FP16 / BF16 kernels (see Launchers).
#ifndef NFLOW_STORAGE_SET
#define NFLOW_STORAGE_SET 1
#endif

// Population layout.  NFLOW_AOSOA = 1 (default): blocks of B consecutive
// nodes, each block holding its 19 populations as 19 runs of B elements with
// B * sizeof(element) = 256 bytes (B = 64 at FP32, 128 at FP16/BF16).  A
// workgroup's stores then form one contiguous range and its pulls come from
// a handful of neighbour blocks, so the device sees ~10 concurrent DRAM
// streams instead of the 38 of plane-per-population SoA.  Measured at 256^3
// MRT-LES FP32 (512-thread workgroups): SoA 5.75 TB/s, B = 32 5.91,
// B = 64 6.05, B = 128 6.01, B = 256 5.87.  NFLOW_AOSOA = 0: SoA planes
// (element q * P + n).
#ifndef NFLOW_AOSOA
#define NFLOW_AOSOA 1
#endif
// Step-kernel population access through buffer instructions for the 16-bit
// storage formats (AoSoA only).  Measured at 256^3 MRT-LES: FP16 5.38 ->
// 5.80 TB/s; FP32 6.04 -> 5.87 TB/s, so FP32 keeps 64-bit addresses.
#ifndef NFLOW_BUFFER
#define NFLOW_BUFFER 1
#endif

namespace nflow {
using namespace d3q19;

// element index of population q of node n for element type T (ps = plane
// stride, SoA only).  (The same index as n + (Q - 1) B (n / B) + q B with one
// 24-bit multiply-add in place of the quarter-rate 32-bit multiply measured
// 5% faster in the two-step channel kernel but 5% slower in the 16-bit step
// and neutral elsewhere, so the form stays.)
template <class T>
__device__ __forceinline__ uint32_t pidx(int q, uint32_t n, uint32_t ps) {
#if NFLOW_AOSOA
  constexpr uint32_t B = 256 / sizeof(T);
  return ((n / B) * Q + (uint32_t)q) * B + (n & (B - 1));
#else
  return (uint32_t)q * ps + n;
#endif
}

// Element offset of node bn's block (bn a multiple of the block length B):
// pidx(q, n) = window(bn) + pidx(q, n - bn), so a kernel that views a
// buffer from that element on indexes it with 32-bit offsets whatever the
// lattice size (a lattice may hold more than 2^32 / 19 nodes).
template <class T>
__device__ __forceinline__ size_t window(uint32_t bn) {
#if NFLOW_AOSOA
  return (size_t)bn * Q;
#else
  return bn;
#endif
}

// The population buffers live in device global memory; saying so (address
// space 1) lets the compiler emit global_load/global_store instead of FLAT
// accesses, which it must otherwise use for pointers read out of a struct.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Population access.  GPops: a 64-bit pointer per lane (global_load/store).
// BPops: a buffer resource in scalar registers whose base is a per-work-item
// window of the population array, plus a 32-bit byte offset per lane
// (buffer_load/store ... offen); the window keeps offsets small for any
// lattice size, and loads past num_records return 0 rather than fault.
// ---------------------------------------------------------------------------
// NTL: loads carry the non-temporal hint (NFLOW_NT_LOAD); false for a kernel
// whose loads are read again soon, by a neighbouring workgroup, from L2.
template <class T, bool NTL = (NFLOW_NT_LOAD != 0)>
struct GPops {
  T* p;
  // element i as a 32-bit byte offset from the uniform base: the access takes
  // the scalar-base form, with no 64-bit address arithmetic per lane (every
  // view is a window of less than 4 GB).  Measured, same box: two-step 256^3
  // cavity 62.9 -> 63.9 GLUPS, channel 52.9 -> 53.1; one step within 0.5%.
  __device__ gptr<const T> at(uint32_t i) const {
    return (gptr<const T>)((gptr<const char>)p + i * (uint32_t)sizeof(T));
  }
  __device__ T get(uint32_t i) const {
    if constexpr (NTL)
      return __builtin_nontemporal_load(at(i));
    else
      return *at(i);
  }
  __device__ void put(uint32_t i, T v) const {
    const gptr<T> g = (gptr<T>)((gptr<char>)p + i * (uint32_t)sizeof(T));
#if NFLOW_NT_STORE
    __builtin_nontemporal_store(v, g);
#else
    *g = v;
#endif
  }
};

// Buffer resource word 3 for gfx9-family untyped access; cache-policy bit 1
// of the instruction is the non-temporal hint.
template <class T>
struct BPops {
  __amdgpu_buffer_rsrc_t r;
  __device__ BPops(T* base, uint64_t bytes) {
    r = __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)(uint32_t)(bytes < 0xFFFFFFFFull ? bytes : 0xFFFFFFFFull),
                                          0x00020000);
  }
  __device__ T get(uint32_t i) const {
    constexpr int aux = NFLOW_NT_LOAD ? 2 : 0;
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(i * 4u), 0, aux));
    else
      return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, (int)(i * 2u), 0, aux));
  }
  __device__ void put(uint32_t i, T v) const {
    constexpr int aux = NFLOW_NT_STORE ? 2 : 0;
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)(i * 4u), 0, aux);
    else
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, (int)(i * 2u), 0, aux);
  }
};

// ---------------------------------------------------------------------------
// Storage formats: element type and the float conversions
// ---------------------------------------------------------------------------
struct StF32 {
  using T = float;
  static constexpr bool shifted = false;
  static constexpr bool buffer = false;  // step-kernel access through buffer instructions
  __device__ static inline float dec(T v) { return v; }
  __device__ static inline T enc(float v) { return v; }
  __device__ static inline T enc_sr(float v, uint32_t) { return v; }
};

// Stochastic rounding to 16 bits with gfx950's v_cvt_sr_f16_f32 /
// v_cvt_sr_bf16_f32.  Measured behaviour of the instruction (every seed, for
// values at and beside binade edges, subnormals and both signs): the
// magnitude's bits below the 16-bit mantissa (13 for FP16, 16 for BF16) are
// added to the top 13 / 16 bits of the seed and the sum is truncated, so a
// value rounds away from zero with probability equal to its distance from the
// grid point below it, in units of the grid spacing -- exactly, with the carry
// into the next binade handled.  FP16 results in the subnormal range
// (|v| < 2^-14) use the same 13 seed bits for up to 23 discarded bits: the
// bias there is below 2^-13 of the 2^-24 spacing.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

struct StF16 {
  using T = _Float16;
  static constexpr bool shifted = true;
  static constexpr bool buffer = true;
  __device__ static inline float dec(T v) { return (float)v; }
  __device__ static inline T enc(float v) { return (_Float16)v; }
  __device__ static inline T enc_sr(float v, uint32_t seed) {
    return __builtin_amdgcn_cvt_sr_f16_f32(f16x2{0, 0}, v, seed, false)[0];
  }
};

// gfx950 converts float -> bf16 (round to nearest even) in one instruction
// (v_cvt_pk_bf16_f32); bf16 -> float is a 16-bit shift.
struct StBF16 {
  using T = __bf16;
  static constexpr bool shifted = true;
  static constexpr bool buffer = true;
  __device__ static inline float dec(T v) { return (float)v; }
  __device__ static inline T enc(float v) { return (__bf16)v; }
  __device__ static inline T enc_sr(float v, uint32_t seed) {
    return __builtin_amdgcn_cvt_sr_bf16_f32(bf16x2{0, 0}, v, seed, false)[0];
  }
};

template <class St>
__device__ constexpr float shift_of(int q) {
  return St::shifted ? W[q] : 0.0f;
}

// Stored element of population q -> its value: f = v + w_q for the shifted
// 16-bit formats; FP32 stores f itself (an added 0.0f would only turn -0 into
// +0, at one VALU instruction per population).
template <class St>
__device__ __forceinline__ float load_f(typename St::T v, int q) {
  if constexpr (St::shifted)
    return St::dec(v) + shift_of<St>(q);
  else
    return St::dec(v);
}

// Seeds of the stochastic rounding.  Each collision draws one 32-bit hash per
// node from (node id, collision count) and gives store q the seed
// hash + q * 0x9E3779B9 (a Weyl sequence, whose top bits -- the ones the
// conversion uses -- are uniform for every q).  The collision count makes the
// rounding of a value that repeats exactly differ from step to step, and
// since the seed depends on nothing else, a lattice rounds the same way alone
// and in any batch.  mix32 is the "lowbias32" integer finaliser
// (C. Wellons, 2018: xorshift-multiply rounds with full avalanche).
__host__ __device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
constexpr uint32_t SR_WEYL = 0x9E3779B9u;

// e_q . (a, b, c) with only the non-zero terms of the integer vector e_q.
__device__ inline float edot(int q, float a, float b, float c) {
  float r = 0.0f;
  bool any = false;
  if (EX[q] != 0) { r = EX[q] > 0 ? a : -a; any = true; }
  if (EY[q] != 0) { r = any ? (EY[q] > 0 ? r + b : r - b) : (EY[q] > 0 ? b : -b); any = true; }
  if (EZ[q] != 0) { r = any ? (EZ[q] > 0 ? r + c : r - c) : (EZ[q] > 0 ? c : -c); any = true; }
  return r;
}

__device__ inline int node_type(const LatticeDev& L, int x, int y, int z) {
  int t = NODE_FLUID;
  if (L.casetype == CASE_CAVITY) {  // lbm_codelets.cpp:183-201
    if (y == 0) t = NODE_FIXEDWALL;
    if (y == L.ny - 1) t = NODE_MOVINGWALL;
    if (x == 0 || x == L.nx - 1 || z == 0 || z == L.nz - 1) t = NODE_FIXEDWALL;
  } else {  // channel and pipe (a duct; the pipe adds its solid mask): inlet x = 0,
            // outlet x = nx-1, side faces are walls (kernel.cu:1081-1088 for the flags)
    if (x == 0) t = NODE_INLET;
    if (x == L.nx - 1) t = NODE_OUTLET;
    if (y == 0 || y == L.ny - 1 || z == 0 || z == L.nz - 1) t = NODE_FIXEDWALL;
  }
  return t;
}

// ---------------------------------------------------------------------------
// Streaming + boundary rules: f_q(x) for the step just completed.
// Donor inside: f_q(x) = f*_q(x - e_q).  Donor outside (the link x -> x + e_opp
// leaves the box): the rule of the node's type applied to f*_opp(q)(x).
// MASK: a solid donor bounces back, f_q(x) = f*_opp(q)(x) (keBoundaryStruct,
// kernel.cu:682-695, applied here to every non-solid node).
// ---------------------------------------------------------------------------
// `in` addresses the population array relative to a window starting at node
// n - nr (nr = n for a whole-array view); n itself indexes the solid mask.
// RUNS (AoSoA, nx a multiple of the block length B, no mask): a donor
// e_y planes and e_z rows away lies a whole number of blocks away from the
// node at x - e_x, so its index is one of three per-node bases (x - 1, x,
// x + 1) plus a wave-uniform term, one add per population in place of a full
// index computation.
template <class St, bool MASK, class Acc, bool RUNS = false>
__device__ inline void pull(const LatticeDev& L, const Acc& in, uint32_t n, uint32_t nr, int x, int y, int z,
                            float f[Q]) {
  using T = typename St::T;
  const bool xlo = x == 0, xhi = x == L.nx - 1;
  const bool ylo = y == 0, yhi = y == L.ny - 1;
  const bool zlo = z == 0, zhi = z == L.nz - 1;
  const uint32_t ps = (uint32_t)L.pstride;
  const gptr<const uint8_t> mk = (gptr<const uint8_t>)L.mask;
  static_assert(!RUNS || (!MASK && NFLOW_AOSOA), "run-aligned donors: AoSoA layout, no mask");
  constexpr uint32_t B = 256 / sizeof(T);
  uint32_t b0 = 0, bm = 0, bp = 0, ky = 0, kz = 0;
  if constexpr (RUNS) {
    b0 = pidx<T>(0, nr, ps);
    bm = pidx<T>(0, nr - 1u, ps);
    bp = pidx<T>(0, nr + 1u, ps);
    ky = ((uint32_t)L.nxz / B) * (Q * B);  // elements per plane
    kz = ((uint32_t)L.nx / B) * (Q * B);   // elements per row
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    bool out = (EX[q] > 0 && xlo) || (EX[q] < 0 && xhi) || (EY[q] > 0 && ylo) || (EY[q] < 0 && yhi) ||
               (EZ[q] > 0 && zlo) || (EZ[q] < 0 && zhi);
    uint32_t idx;
    if constexpr (RUNS) {
      const uint32_t bx = EX[q] > 0 ? bm : (EX[q] < 0 ? bp : b0);
      const uint32_t u = (uint32_t)q * B - (EY[q] > 0 ? ky : 0u) + (EY[q] < 0 ? ky : 0u) - (EZ[q] > 0 ? kz : 0u) +
                         (EZ[q] < 0 ? kz : 0u);
      idx = out ? b0 + (uint32_t)OPP[q] * B : bx + u;
    } else {
      // (a select of two full index computations: measured faster than
      // selecting the source node and the plane separately, which needs more
      // registers)
      const int off = EX[q] + EZ[q] * L.nx + EY[q] * L.nxz;
      if (MASK && q != 0) {
        if (mk && !out) out = mk[(long long)n - off] != 0;
      }
      idx = out ? pidx<T>(OPP[q], nr, ps) : pidx<T>(q, (uint32_t)((int)nr - off), ps);
    }
    f[q] = load_f<St>(in.get(idx), q);
  }
}

// Inputs of the inlet / outlet rules from the previous step's side planes:
// inlet, v[0] = rho(x+1) and v[1] = u_in; outlet, v[0..2] = u(nx-1) and
// v[3..5] = u(nx-2).  Loaded before the 19 pulls, so that a wave holding an
// inlet or outlet lane waits for both together rather than in sequence.
struct SideIn {
  float v[6];
};

__device__ inline SideIn side_inputs(const LatticeDev& L, const float* __restrict__ side, int x, int y, int z) {
  SideIn si = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
  if (L.casetype != CASE_CAVITY && (x == 0 || x == L.nx - 1) &&
      !(y == 0 || y == L.ny - 1 || z == 0 || z == L.nz - 1)) {  // side faces are walls (node_type)
    const gptr<const float> rec = (gptr<const float>)side + (size_t)(y * L.nz + z) * SIDE_REC;
    if (x == 0) {
      const f32x2 a = *(gptr<const f32x2>)(rec + SIDE_RHO);  // rho(x = 1), u_in
      si.v[0] = a.x;
      si.v[1] = a.y;
    } else {
      const f32x4 u0 = *(gptr<const f32x4>)(rec + SIDE_U0);
      const f32x4 u1 = *(gptr<const f32x4>)(rec + SIDE_U1);
      si.v[0] = u0.x;
      si.v[1] = u0.y;
      si.v[2] = u0.z;
      si.v[3] = u1.x;
      si.v[4] = u1.y;
      si.v[5] = u1.z;
    }
  }
  return si;
}

// Wall / inlet / outlet rules on the links pull() filled with f*_opp(q)(x).
// node_type gives the x, z and y = 0 faces priority as fixed walls, so a
// moving-wall node lies on y = ny-1 (and on y = 0 only when ny = 1), an inlet
// node on x = 0 alone and an outlet node on x = nx-1 alone (nx >= 3): each
// rule touches a fixed set of directions.
__device__ inline void boundary(const LatticeDev& L, const SideIn& si, int x, int y, int z, float f[Q]) {
  const bool xlo = x == 0, xhi = x == L.nx - 1;
  const bool ylo = y == 0, yhi = y == L.ny - 1;
  const bool zlo = z == 0, zhi = z == L.nz - 1;
  if (xlo | xhi | ylo | yhi | zlo | zhi) {
    const int t = node_type(L, x, y, z);
    if (t == NODE_MOVINGWALL) {  // f_q = f*_opp + 6 w rho0 (e_q . U), links leaving through y = ny-1
#pragma unroll
      for (int q = 1; q < Q; ++q)
        if (EY[q] < 0 || (EY[q] > 0 && ylo)) f[q] += 6.0f * W[q] * L.refden * edot(q, L.ulid[0], L.ulid[1], L.ulid[2]);
    } else if (t == NODE_INLET) {  // f_q = f*_opp + 6 w rho(x+1) (e_q . u_in), e_x = +1
      const float a = 6.0f * si.v[0] * si.v[1];  // rho(x = 1) and u_in of the previous step
#pragma unroll
      for (int q = 1; q < Q; ++q)
        if (EX[q] > 0) f[q] += W[q] * a;
    } else if (t == NODE_OUTLET) {  // anti-bounce-back at rho0, e_x = -1
      const float ubx = si.v[0] + 0.5f * (si.v[0] - si.v[3]);
      const float uby = si.v[1] + 0.5f * (si.v[1] - si.v[4]);
      const float ubz = si.v[2] + 0.5f * (si.v[2] - si.v[5]);
      const float usq = ubx * ubx + uby * uby + ubz * ubz;
#pragma unroll
      for (int q = 1; q < Q; ++q)
        if (EX[q] < 0) {
          const float eu = edot(q, ubx, uby, ubz);
          f[q] = -f[q] + 2.0f * W[q] * L.refden * (1.0f + 4.5f * eu * eu - 1.5f * usq);
        }
    }  // fixed walls: plain bounce-back, already in place
  }
}

template <class St, bool MASK, class Acc, bool RUNS = false>
__device__ inline void gather(const LatticeDev& L, const Acc& in, const float* __restrict__ side, uint32_t n,
                              uint32_t nr, int x, int y, int z, float f[Q]) {
  const SideIn si = side_inputs(L, side, x, y, z);
  pull<St, MASK, Acc, RUNS>(L, in, n, nr, x, y, z, f);
  boundary(L, si, x, y, z, f);
}

__device__ inline bool is_solid(const LatticeDev& L, uint32_t n) {
  return L.mask && ((gptr<const uint8_t>)L.mask)[n] != 0;
}

// ---------------------------------------------------------------------------
// Collision
// ---------------------------------------------------------------------------
// Hardware reciprocal and square root (v_rcp_f32, v_sqrt_f32, each within
// 1 ulp) for the per-node 1/rho, 1/(tau + tau_eddy) and the two LES roots:
// the correctly rounded forms expand to ~10 instructions each on the
// critical path of the collision.
__device__ inline float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ inline void moments(const float f[Q], float& rho, float& jx, float& jy, float& jz) {
  float r = f[0];
#pragma unroll
  for (int q = 1; q < Q; ++q) r += f[q];
  float a = 0.0f, b = 0.0f, c = 0.0f;
  bool sa = false, sb = false, sc = false;
#pragma unroll
  for (int q = 1; q < Q; ++q) {
    if (EX[q] != 0) { a = sa ? (EX[q] > 0 ? a + f[q] : a - f[q]) : (EX[q] > 0 ? f[q] : -f[q]); sa = true; }
    if (EY[q] != 0) { b = sb ? (EY[q] > 0 ? b + f[q] : b - f[q]) : (EY[q] > 0 ? f[q] : -f[q]); sb = true; }
    if (EZ[q] != 0) { c = sc ? (EZ[q] > 0 ? c + f[q] : c - f[q]) : (EZ[q] > 0 ? f[q] : -f[q]); sc = true; }
  }
  rho = r;
  jx = a;
  jy = b;
  jz = c;
}

// Smagorinsky: tau_eddy = -tau/2 + sqrt(tau^2 + 18 sqrt2 Cs^2 sqrt(Q)/rho)/2 with
// Q = Pi:Pi and Pi = sum_q e e (f - feq) = sum_q e e (f - w rho) - j j / rho.
__device__ inline float tau_eddy_les(const float f[Q], float rho, float ir, float jx, float jy, float jz, float tau,
                                     float les_c) {
  float pxx = 0.f, pyy = 0.f, pzz = 0.f, pxy = 0.f, pyz = 0.f, pxz = 0.f;
#pragma unroll
  for (int q = 1; q < Q; ++q) {
    const float g = f[q] - W[q] * rho;
    if (EX[q] != 0) pxx += g;
    if (EY[q] != 0) pyy += g;
    if (EZ[q] != 0) pzz += g;
    if (EX[q] * EY[q] != 0) pxy += (EX[q] * EY[q] > 0) ? g : -g;
    if (EY[q] * EZ[q] != 0) pyz += (EY[q] * EZ[q] > 0) ? g : -g;
    if (EX[q] * EZ[q] != 0) pxz += (EX[q] * EZ[q] > 0) ? g : -g;
  }
  pxx -= jx * jx * ir;
  pyy -= jy * jy * ir;
  pzz -= jz * jz * ir;
  pxy -= jx * jy * ir;
  pyz -= jy * jz * ir;
  pxz -= jx * jz * ir;
  const float qq = pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pyz * pyz + pxz * pxz);
  return 0.5f * (hw_sqrt(tau * tau + les_c * hw_sqrt(qq) * ir) - tau);
}

// The Smagorinsky relaxation time of tau_eddy_les from the MRT moments: the
// second moments P = sum_q e e f are m13 = P_xy, m14 = P_yz, m15 = P_xz,
// m9 = 3 P_xx - tr P, m11 = P_yy - P_zz and m1 = 19 tr P - 30 rho, and
// sum_q e e feq = rho/3 I + j j / rho, so Pi = P - rho/3 I - j j / rho.
__device__ inline float les_from_moments(const float m[Q], float rho, float ir, float jx, float jy, float jz,
                                         float tau, float les_c) {
  const float tr = (m[1] + 30.0f * rho) * (1.0f / 19.0f);
  const float pxx = (m[9] + tr) * (1.0f / 3.0f);
  const float ryz = tr - pxx;
  const float pyy = 0.5f * (ryz + m[11]);
  const float pzz = 0.5f * (ryz - m[11]);
  const float r3 = rho * (1.0f / 3.0f);
  const float qxx = pxx - r3 - jx * jx * ir;
  const float qyy = pyy - r3 - jy * jy * ir;
  const float qzz = pzz - r3 - jz * jz * ir;
  const float qxy = m[13] - jx * jy * ir;
  const float qyz = m[14] - jy * jz * ir;
  const float qxz = m[15] - jx * jz * ir;
  const float qq = qxx * qxx + qyy * qyy + qzz * qzz + 2.0f * (qxy * qxy + qyz * qyz + qxz * qxz);
  return 0.5f * (hw_sqrt(tau * tau + les_c * hw_sqrt(qq) * ir) - tau);
}

// BGK, kernel.cu:479-486.
__device__ inline void collide_srt(float f[Q], float rho, float ir, float jx, float jy, float jz, float omega) {
  const float ux = jx * ir, uy = jy * ir, uz = jz * ir;
  const float base = 1.0f - 1.5f * (ux * ux + uy * uy + uz * uz);
  {
    const float feq = W[0] * rho * base;
    f[0] -= omega * (f[0] - feq);
  }
#pragma unroll
  for (int i = 0; i < NPAIR; ++i) {
    const int a = PA[i], b = PB[i];
    const float eu = edot(a, ux, uy, uz);
    const float sym = W[a] * rho * (base + 4.5f * eu * eu);
    const float asym = W[a] * rho * 3.0f * eu;
    f[a] -= omega * (f[a] - (sym + asym));
    f[b] -= omega * (f[b] - (sym - asym));
  }
}

// MRT, lbm_codelets.cpp:449-537: f* = f - M^-1 S (M f - meq) with
// M^-1 = M^T diag(1/|row|^2) (the rows are orthogonal).  m: all 19 moments
// (mrt_moments).  The back transform is factored like mrt_moments: each row
// is even or odd under e -> -e, so f_a -= E + O and f_opp(a) -= E - O, with
// the even parts sharing their axis and edge-group terms.  With LES the eddy
// relaxation time comes from the same moments (les_from_moments).
__device__ inline float les_from_moments(const float m[Q], float rho, float ir, float jx, float jy, float jz,
                                         float tau, float les_c);
__device__ inline void collide_mrt(float f[Q], const float m[Q], float rho, float ir, float jx, float jy, float jz,
                                   const LatticeDev& L) {
  const float snu = hw_rcp(L.tau + (L.les ? les_from_moments(m, rho, ir, jx, jy, jz, L.tau, L.les_c) : 0.0f));
  const float ir0 = L.inv_refden;
  const float jxx = jx * jx, jyy = jy * jy, jzz = jz * jz, jj = jxx + jyy + jzz;
  float meq[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) meq[k] = 0.0f;
  meq[1] = -11.0f * rho + 19.0f * jj * ir0;
  meq[2] = L.weps * rho + L.wepsj * jj * ir0;
  meq[4] = (-2.0f / 3.0f) * jx;
  meq[6] = (-2.0f / 3.0f) * jy;
  meq[8] = (-2.0f / 3.0f) * jz;
  meq[9] = (2.0f * jxx - (jyy + jzz)) * ir0;
  meq[10] = L.wxx * meq[9];
  meq[11] = (jyy - jzz) * ir0;
  meq[12] = L.wxx * meq[11];
  meq[13] = jx * jy * ir0;
  meq[14] = jy * jz * ir0;
  meq[15] = jx * jz * ir0;
  // relaxation rates of lbm_codelets.cpp:488-506, folded with 1/|row|^2
  float dl[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) {
    float rate;
    switch (k) {
      case 1: rate = 1.19f; break;
      case 2: case 10: case 12: rate = 1.4f; break;
      case 4: case 6: case 8: rate = 1.2f; break;
      case 9: case 11: case 13: case 14: case 15: rate = -1.0f; break;
      case 16: case 17: case 18: rate = 1.98f; break;
      default: rate = 0.0f;
    }
    const float inv_n = 1.0f / (float)MNORM2[k];
    if (rate == 0.0f)
      dl[k] = 0.0f;
    else if (rate < 0.0f)
      dl[k] = (snu * inv_n) * (m[k] - meq[k]);
    else
      dl[k] = (rate * inv_n) * (m[k] - meq[k]);
  }
  mrt_back(f, dl);
}

// Macroscopic -> side planes -> LES -> collision, in place.
// SIDES = false: no side records written (cavity lattices, or a caller that
// keeps them itself).  mac: if given, receives rho and u = j / rho as the side
// records hold them.
template <bool SIDES = true>
__device__ __forceinline__ void relax(const LatticeDev& L, float f[Q], float* __restrict__ side_out, int x, int y,
                                      int z, float* mac = nullptr) {
  // macroscopic outputs: mac, and the inlet / outlet side records of channel
  // and pipe lattices
  auto outputs = [&](float rho, float ir, float jx, float jy, float jz) {
    if (mac) {
      mac[0] = rho;
      mac[1] = jx * ir;
      mac[2] = jy * ir;
      mac[3] = jz * ir;
    }
    if (SIDES && L.casetype != CASE_CAVITY) {
      const gptr<float> rec = (gptr<float>)side_out + (size_t)(y * L.nz + z) * SIDE_REC;
      if (x == 1) rec[SIDE_RHO] = rho;
      if (x == L.nx - 1) *(gptr<f32x4>)(rec + SIDE_U0) = f32x4{jx * ir, jy * ir, jz * ir, 0.0f};
      if (x == L.nx - 2) *(gptr<f32x4>)(rec + SIDE_U1) = f32x4{jx * ir, jy * ir, jz * ir, 0.0f};
    }
  };
  float rho, jx, jy, jz;
  if (L.model == MODEL_MRT) {
    float m[Q];
    mrt_moments(f, m);
    rho = m[0];
    jx = m[3];
    jy = m[5];
    jz = m[7];
    const float ir = hw_rcp(rho);
    outputs(rho, ir, jx, jy, jz);
    collide_mrt(f, m, rho, ir, jx, jy, jz, L);
  } else {
    moments(f, rho, jx, jy, jz);
    const float ir = hw_rcp(rho);
    outputs(rho, ir, jx, jy, jz);
    const float te = L.les ? tau_eddy_les(f, rho, ir, jx, jy, jz, L.tau, L.les_c) : 0.0f;
    collide_srt(f, rho, ir, jx, jy, jz, hw_rcp(L.tau + te));
  }
}

// relax, then the unit-stride store of the 19 post-collision populations.
// SR: the instantiation carries the stochastic-rounding store (taken for
// lattices with LatticeDev::sr set, seeded from sr_hash); without it the code
// is not generated.
template <class St, bool SR, bool SIDES = true, class Acc>
__device__ inline void relax_and_store(const LatticeDev& L, float f[Q], const Acc& out, float* __restrict__ side_out,
                                       uint32_t nr, int x, int y, int z, uint32_t sr_hash) {
  relax<SIDES>(L, f, side_out, x, y, z);
  const uint32_t ps = (uint32_t)L.pstride;
  if constexpr (St::shifted && SR) {
    if (L.sr) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        out.put(pidx<typename St::T>(q, nr, ps), St::enc_sr(f[q] - shift_of<St>(q), sr_hash + (uint32_t)q * SR_WEYL));
      return;
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) out.put(pidx<typename St::T>(q, nr, ps), St::enc(f[q] - shift_of<St>(q)));
}

__device__ inline void coords(const LatticeDev& L, uint32_t n, int& x, int& y, int& z) {
  const uint32_t yy = fastdiv(n, L.div_nxz);
  const uint32_t r = n - yy * (uint32_t)L.nxz;
  const uint32_t zz = fastdiv(r, L.div_nx);
  x = (int)(r - zz * (uint32_t)L.nx);
  y = (int)yy;
  z = (int)zz;
}

// ---------------------------------------------------------------------------
// The time step: one launch advances every lattice in the work list.
// ---------------------------------------------------------------------------
#if NFLOW_WAVES_PER_EU
#define NFLOW_STEP_ATTR __attribute__((amdgpu_waves_per_eu(NFLOW_WAVES_PER_EU, NFLOW_WAVES_PER_EU)))
#else
#define NFLOW_STEP_ATTR
#endif

// One node of a work item: pull + boundary rules, collide, store.  `in` and
// `out` address the arrays from node bn on (0: the whole array).
// sr_key: mix32 of this collision's count (stochastic rounding only).
template <class St, bool MASK, bool SR, class Acc>
__device__ __forceinline__ void step_node(const LatticeDev& L, const WorkItem& wi, int p, uint32_t bn, const Acc& in,
                                 const Acc& out, uint32_t sr_key) {
  if (threadIdx.x >= wi.count) return;
  const uint32_t n = wi.start + threadIdx.x;
  if (MASK && is_solid(L, n)) return;  // solid nodes are never updated
  int x, y, z;
  coords(L, n, x, y, z);
  float f[Q];
  gather<St, MASK>(L, in, L.side[p], n, n - bn, x, y, z, f);
  relax_and_store<St, SR>(L, f, out, L.side[p ^ 1], n - bn, x, y, z, SR ? mix32(n ^ sr_key) : 0u);
}

// MASK = true: at least one lattice of the launch has solid nodes (the plain
// instantiation carries no mask code at all).
//
// Population addressing: with buffer access each work item sets up (in
// scalar registers) a window of both population arrays starting at the
// B-node block that holds node start - (nxz + nx), the farthest donor any
// node of the item pulls from, and its lanes index the window with n - bn.
// pidx(q, n) - pidx(q, n - bn) = bn * Q for bn a multiple of B, so the
// element addressed is the same as with whole-array indexing.
template <class St, int BS, bool MASK, bool SR>
__global__ void __launch_bounds__(BS) NFLOW_STEP_ATTR k_step(const LatticeDev* __restrict__ lats, const WorkItem* __restrict__ items,
                                             int nitems, int parity) {
  using T = typename St::T;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem wi = items[it];
    const LatticeDev& L = lats[wi.lat];
    const int p = parity ^ L.parity0;
    uint32_t sr_key = 0;
    if constexpr (SR) {
      // this launch rounds with collision count ctr[p] (k_advance, launched
      // after it, hands ctr[p] + 1 to the next launch in the other slot).
      // Nothing in this kernel writes the counters, so the read goes through
      // the constant address space and becomes a scalar load: the step
      // kernel carries no store besides its populations and side records,
      // whose waits a store of its own would lengthen.
      if (L.sr) sr_key = mix32(((__attribute__((address_space(4))) const uint32_t*)L.ctr)[p]);
    }
    if constexpr (NFLOW_AOSOA && NFLOW_BUFFER && St::buffer) {
      constexpr uint32_t B = 256 / sizeof(T);
      const uint32_t reach = (uint32_t)L.nxz + (uint32_t)L.nx;
      const uint32_t bn = (wi.start > reach ? wi.start - reach : 0u) & ~(B - 1);
      const uint64_t skip = (uint64_t)bn * Q;  // elements before the window
      const uint64_t bytes = ((uint64_t)L.pstride * Q - skip) * sizeof(T);
      step_node<St, MASK, SR>(L, wi, p, bn, BPops<T>(static_cast<T*>(L.buf[p]) + skip, bytes),
                          BPops<T>(static_cast<T*>(L.buf[p ^ 1]) + skip, bytes), sr_key);
    } else {
      // the same window, from a per-item base pointer
      constexpr uint32_t B = 256 / sizeof(T);
      const uint32_t reach = (uint32_t)L.nxz + (uint32_t)L.nx;
      const uint32_t bn = (wi.start > reach ? wi.start - reach : 0u) & ~(B - 1);
      step_node<St, MASK, SR>(L, wi, p, bn, GPops<T>{static_cast<T*>(L.buf[p]) + window<T>(bn)},
                              GPops<T>{static_cast<T*>(L.buf[p ^ 1]) + window<T>(bn)}, sr_key);
    }
  }
}

// Collision counters of the stochastic rounding: after the launch that read
// buf[p] (p = parity ^ parity0), ctr[p ^ 1] = ctr[p] + 1.
#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_advance(const LatticeDev* __restrict__ lats, int nlat, int parity) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nlat) return;
  const LatticeDev& L = lats[i];
  if (!L.sr) return;
  const int p = parity ^ L.parity0;
  L.ctr[p ^ 1] = L.ctr[p] + 1u;
}
#endif

// ---------------------------------------------------------------------------
// Two time steps per launch (a cavity or channel lattice without solid nodes).
//
// A workgroup owns a TX x TZ tile of the x-z plane and a range [ya, yb) of y
// planes, and sweeps it in y.  TX spans whole x rows of the benchmark sizes:
// the pulls of a 128-byte line granularity make every x edge of a tile cost
// about 28 extra lines per row, so the tile has none where nx <= TX.  Thread
// t owns K nodes of column x0 + t % TX, stacked in z.  For each plane yi:
//   1. the first of the two steps (pull from buf[p] with the boundary rules,
//      collision) on the own nodes and on the halo around the tile (the rows
//      z0 - 1 and z0 + TZ, and the columns x0 - 1 and x0 + TX where they are
//      inside the box), spread over the threads; populations that cross to
//      another x-z column go to LDS (from the halo rows only the 5 that cross
//      into the tile in z); the own node's rest population, e = (0, +-1, 0)
//      and the bounce-back of its links that leave the box go straight to its
//      registers;
//   2. each thread collects the crossing populations its nodes pull from
//      plane yi: e_y = -1 ones for plane yi - 1, e_y = 0 ones for plane yi,
//      e_y = +1 ones for plane yi + 1;
//   3. the second step of plane yi - 1, whose 19 pulled values are now
//      complete, stored to buf[p ^ 1].
// The partial pulls of planes yi and yi + 1 stay in registers.
//
// Built for FP32 storage.  Per node and step the arithmetic and the stored
// values are those of k_step (both compiled with -ffp-contract=on, so an FMA
// never spans statements and the collision rounds the same in either kernel),
// so two launches of k_step and one of k_step2 give the same state bit for
// bit; the memory traffic of the pair is one read and one write of the
// populations (plus the halo) instead of two of each.  The intermediate state
// never reaches memory, so the launch before a read-back must be a k_step
// (the export streams the state the last launch read): a run that ends with
// a k_step2 is settled by re-running it as two k_step launches from the state
// it read, which it leaves untouched.
// ---------------------------------------------------------------------------
namespace step2 {
// Slot of crossing population q (e_x or e_z non-zero) in an LDS row: the 5
// with e_z = -1 first (3, 7, 8, 15, 17), then the 6 with e_z = 0, then the 5
// with e_z = +1 (4, 9, 10, 16, 18); -1: the own column keeps it (rest, e =
// (0, +-1, 0)).  The rows of the tile are stored one after another, the halo
// row z0 - 1 before them with only its e_z = +1 slots (the ones pulled into
// the tile) and the halo row z0 + TZ after them with only its e_z = -1 slots,
// so that every access is one per-thread base plus a constant.
constexpr int CIX[Q] = {-1, 5, 6, 0, 11, -1, -1, 1, 2, 12, 13, 7, 8, 9, 10, 3, 14, 4, 15};
constexpr int NC = 16;
constexpr int QUP = 5, QDN = 6;  // e = (0, +1, 0), (0, -1, 0)

template <int TX_, int TZ_, int NT_>
struct Geo {
  static constexpr int TX = TX_, TZ = TZ_, NT = NT_, WX = TX + 2;
  static constexpr int RP = NT / TX;  // x rows per pass of the threads
  static constexpr int K = TZ / RP;   // own nodes per thread
  static_assert(NT % TX == 0 && TZ % RP == 0, "whole rows per thread pass");
  static_assert(TX % 64 == 0, "a wave within one row group");
  // Tile row of own node j of row group r.  The halo rows of a tile pull the
  // same elements as the edge rows of the neighbouring tiles, which run on the
  // same XCD; the halo goes to the first row groups, so the edge rows (0 and
  // TZ - 1) go to row groups without halo work, as their first node, and are
  // computed in the same pass of the plane as the neighbours' halo rows: the
  // second load of those lines then finds them in L2 (measured at 256^3:
  // 302.4 -> 297.0 us per step).
  __device__ static constexpr int row(int r, int j) {
    if constexpr (TZ == 8 && RP == 4 && K == 2) {
      constexpr int R[4][2] = {{1, 2}, {3, 4}, {0, 5}, {7, 6}};
      return R[r][j];
    }
    return r * K + j;
  }
  // LDS element of slot c of row zs (-1 .. TZ) at column-frame x xs
  static constexpr int LN = (TZ * NC + 10) * WX;
  __device__ static constexpr int lidx(int zs, int c, int xs) { return ((zs + 1) * NC + c - 11) * WX + xs; }
};
}  // namespace step2

// CH: a channel lattice (velocity inlet at x = 0, pressure outlet at x =
// nx - 1).  The second step's inlet and outlet rules read the first step's
// rho at x = 1 and u at x = nx - 1 and nx - 2 on the same plane; the first
// step leaves them in LDS (two plane slots), where k_step would have written
// the side records, and the second step writes the side records the next
// launch reads.
// RUNS: nx is a multiple of the 64-node block, and the pulls index donors
// from three per-node bases (pull<..., RUNS>).
template <class St, class G, bool CH, bool RUNS>
__global__ void __launch_bounds__(G::NT) k_step2(LatticeDev L, int parity, int H, int ntx, int ntz) {
  static_assert(!St::shifted, "two steps per launch are built for FP32 storage");
  using namespace step2;
  using T = typename St::T;
  constexpr int TX = G::TX, TZ = G::TZ, NT = G::NT, K = G::K;
  // crossing populations of the own rows (with the x - 1 and x + TX columns)
  // and of the halo rows z0 - 1 (e_z = +1 ones) and z0 + TZ (e_z = -1 ones)
  __shared__ T cross[G::LN];
  // channel: the first step's rho(x = 1), u(nx - 1), u(nx - 2) of each tile
  // row, for planes of even and odd y
  __shared__ float sidein[CH ? 2 : 1][TZ][8];
  int bt = blockIdx.x;
  // workgroups are dispatched to the 8 XCDs round-robin: give each XCD a
  // contiguous range of tiles, so that neighbours share their halo in L2
  if ((gridDim.x & 7u) == 0) bt = (bt & 7) * (int)(gridDim.x >> 3) + (bt >> 3);
  const int tx = bt % ntx, tz = (bt / ntx) % ntz, ch = bt / (ntx * ntz);
  const int x0 = tx * TX, z0 = tz * TZ, ya = ch * H, yb = min(L.ny, ya + H);
  const int p = parity ^ L.parity0;
  // The buffers are viewed from the block of the lowest plane the workgroup
  // reads (y = ya - 2, pulled by the first step on plane ya - 1): 32-bit
  // element indices whatever the lattice size.  Loads without the
  // non-temporal hint: the halo rows, and the planes on either side of a y
  // chunk, are loaded again by the neighbouring tile or chunk, and then find
  // them in L2 (measured per launch, 256^3 cavity: 604 -> 564 us; 512 x 128
  // x 128 channel: 372 -> 356 us).  The stores keep it.
  constexpr uint32_t B = 256 / sizeof(T);
  const uint32_t bn = ((uint32_t)max(ya - 2, 0) * (uint32_t)L.nxz) & ~(B - 1);
  const GPops<T, false> in{static_cast<T*>(L.buf[p]) + window<T>(bn)};
  const GPops<T> out{static_cast<T*>(L.buf[p ^ 1]) + window<T>(bn)};
  const int t = threadIdx.x;
  // own nodes: (x0 + lx, z0 + zrow[j]), j < K; the row group rg is the same
  // for a whole wave
  const int lx = t % TX, rg = __builtin_amdgcn_readfirstlane(t / TX);
  int zrow[K];  // tile row of own node j
#pragma unroll
  for (int j = 0; j < K; ++j) zrow[j] = G::row(rg, j);
  const int ox = x0 + lx;
  const bool xin = ox < L.nx, xlo = ox == 0, xhi = ox == L.nx - 1;
  // halo nodes of the tile: row z0 - 1, row z0 + TZ (x from xa to xb), column
  // x0 - 1, column x0 + TX (z0 .. z0 + tzn - 1), each where inside the box
  const int xa = max(x0 - 1, 0), xb = min(x0 + TX, L.nx - 1), rowlen = xb - xa + 1;
  const int tzn = min(TZ, L.nz - z0);
  const int nlo = z0 > 0 ? rowlen : 0, nhi = z0 + TZ < L.nz ? rowlen : 0;
  const int nxl = x0 > 0 ? tzn : 0, nxr = x0 + TX < L.nx ? tzn : 0;
  const int nhalo = nlo + nhi + nxl + nxr;
  // element index base of node (x, y, z) in the window
  auto node = [&](int x, int y, int z) {
    return (uint32_t)x + (uint32_t)L.nx * ((uint32_t)z + (uint32_t)L.nz * (uint32_t)y) - bn;
  };
  // first step on node (x, yi, z): e = the stored values, g = them as the
  // next step reads them back
  auto step1 = [&](int x, int yi, int z, T e[Q], float g[Q]) {
    const uint32_t n = node(x, yi, z);
    if constexpr (CH) {
      gather<St, false, decltype(in), RUNS>(L, in, L.side[p], n, n, x, yi, z, g);
      float mac[4];
      relax<false>(L, g, nullptr, x, yi, z, mac);
      if (z >= z0 && z < z0 + TZ) {  // rows of the tile: what its second step reads
        float* r = sidein[yi & 1][z - z0];
        if (x == 1) r[0] = mac[0];
        if (x == L.nx - 1) { r[1] = mac[1]; r[2] = mac[2]; r[3] = mac[3]; }
        if (x == L.nx - 2) { r[4] = mac[1]; r[5] = mac[2]; r[6] = mac[3]; }
      }
    } else {
      pull<St, false, decltype(in), RUNS>(L, in, n, n, x, yi, z, g);
      boundary(L, SideIn{{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}}, x, yi, z, g);
      relax<false>(L, g, nullptr, x, yi, z);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      e[q] = St::enc(g[q] - shift_of<St>(q));
      g[q] = load_f<St>(e[q], q);
    }
  };

  // Pulled values of own node j: fp on plane yi - 1 (complete but for the
  // e_y = -1 crossing links from plane yi), fc on plane yi (so far its e_y =
  // +1 links), fn the e_y = +1 links of plane yi + 1.
  float fp[K][Q], fc[K][Q], fn[K][Q];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int q = 0; q < Q; ++q) fp[j][q] = fc[j][q] = fn[j][q] = 0.0f;
  // does the link of q at own node j on plane y leave the box?
  auto gone = [&](int q, int j, int y) {
    const int oz = z0 + zrow[j];
    return (EX[q] > 0 && xlo) || (EX[q] < 0 && xhi) || (EY[q] > 0 && y == 0) || (EY[q] < 0 && y == L.ny - 1) ||
           (EZ[q] > 0 && oz == 0) || (EZ[q] < 0 && oz == L.nz - 1);
  };
  for (int yi = ya - 1; yi <= yb; ++yi) {
    const bool plane = yi >= 0 && yi < L.ny;
    if (plane) {
      for (int h = t; h < nhalo; h += NT) {  // halo
        int x, z, kind;  // kind 0: row z0 - 1, 1: row z0 + TZ, 2: a column
        if (h < nlo) { x = xa + h; z = z0 - 1; kind = 0; }
        else if (h < nlo + nhi) { x = xa + (h - nlo); z = z0 + TZ; kind = 1; }
        else if (h < nlo + nhi + nxl) { x = x0 - 1; z = z0 + (h - nlo - nhi); kind = 2; }
        else { x = x0 + TX; z = z0 + (h - nlo - nhi - nxl); kind = 2; }
        T e[Q];
        float g[Q];
        step1(x, yi, z, e, g);
        if (kind == 2) {
#pragma unroll
          for (int q = 0; q < Q; ++q)
            if (CIX[q] >= 0) cross[G::lidx(z - z0, CIX[q], x - x0 + 1)] = e[q];
        } else {
#pragma unroll
          for (int q = 0; q < Q; ++q)
            if (EZ[q] != 0 && (EZ[q] > 0) == (kind == 0)) cross[G::lidx(z - z0, CIX[q], x - x0 + 1)] = e[q];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {  // own nodes
      const int oz = z0 + zrow[j];
      float g[Q];
      if (plane && xin && oz < L.nz) {
        T e[Q];
        step1(ox, yi, oz, e, g);
#pragma unroll
        for (int q = 0; q < Q; ++q)
          if (CIX[q] >= 0) cross[G::lidx(zrow[j], CIX[q], lx + 1)] = e[q];
      } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) g[q] = 0.0f;
      }
      fc[j][0] = g[0];
      fn[j][QUP] = g[QUP];
      fp[j][QDN] = yi - 1 == L.ny - 1 ? fp[j][QDN] : g[QDN];  // leaves only from plane ny - 1
      fc[j][QDN] = gone(QDN, j, yi) ? g[QUP] : 0.0f;
      fc[j][QUP] = gone(QUP, j, yi) ? g[QDN] : fc[j][QUP];
    }
    __syncthreads();
    // crossing population q pulled by own node j from plane yi
    auto pulled = [&](int q, int j) {
      return load_f<St>(cross[G::lidx(zrow[j] - EZ[q], CIX[q], lx + 1 - EX[q])], q);
    };
    // own node j's crossing population q after the first step on plane yi
    auto own = [&](int q, int j) { return load_f<St>(cross[G::lidx(zrow[j], CIX[q], lx + 1)], q); };
    const int yf = yi - 1;
#pragma unroll
    for (int j = 0; j < K; ++j) {  // second step on plane yf
      const int oz = z0 + zrow[j];
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (CIX[q] >= 0 && EY[q] < 0) fp[j][q] = gone(q, j, yf) ? fp[j][q] : pulled(q, j);
      if (xin && oz < L.nz && yf >= ya && yf < yb) {
        const uint32_t n = node(ox, yf, oz);
        SideIn si = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
        if constexpr (CH) {  // side_inputs(), from the first step's values
          if ((ox == 0 || ox == L.nx - 1) && !(yf == 0 || yf == L.ny - 1 || oz == 0 || oz == L.nz - 1)) {
            const float* r = sidein[yf & 1][zrow[j]];
            if (ox == 0) {
              si.v[0] = r[0];
              si.v[1] = ((gptr<const float>)L.side[p])[(size_t)(yf * L.nz + oz) * SIDE_REC + SIDE_UIN];
            } else {
              si.v[0] = r[1];
              si.v[1] = r[2];
              si.v[2] = r[3];
              si.v[3] = r[4];
              si.v[4] = r[5];
              si.v[5] = r[6];
            }
          }
        }
        boundary(L, si, ox, yf, oz, fp[j]);
        relax_and_store<St, false, CH>(L, fp[j], out, L.side[p ^ 1], n, ox, yf, oz, 0u);
      }
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {  // what planes yi and yi + 1 pull from plane yi
#pragma unroll
      for (int q = 1; q < Q; ++q) {
        if (CIX[q] < 0) continue;
        const bool out_q = gone(q, j, yi);
        const float b = out_q ? own(OPP[q], j) : 0.0f;
        if (EY[q] < 0) fc[j][q] = b;
        else if (EY[q] == 0) fc[j][q] = out_q ? b : pulled(q, j);
        else {
          fc[j][q] = out_q ? b : fc[j][q];
          fn[j][q] = pulled(q, j);
        }
      }
    }
    __syncthreads();  // the next plane's first step overwrites LDS
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        fp[j][q] = fc[j][q];
        if (EY[q] > 0) fc[j][q] = fn[j][q];
      }
  }
}

// Initialisation (Initialize vertex, lbm_codelets.cpp:140-241) or import of
// an arF array: the streamed state goes to buf[cur ^ 1] (what an export at
// step 0 returns), its collision to buf[cur] (what the first step streams).
// The state stands at step `step0`: its collision is rounded with that count
// and the first step launch reads step0 + 1 from ctr[cur].
template <class St>
__global__ void __launch_bounds__(256) k_init(LatticeDev L, int cur, const float* __restrict__ src, uint32_t step0) {
  using T = typename St::T;
  const uint32_t bn = blockIdx.x * 256u;  // the block's first node: buffers viewed from there
  const uint32_t n = bn + threadIdx.x, nr = threadIdx.x;
  if (n >= L.n) return;
  if (n == 0) L.ctr[cur] = step0 + 1u;
  int x, y, z;
  coords(L, n, x, y, z);
  if (is_solid(L, n)) {
    // solid: zero populations in both buffers (keInitialize skips them,
    // kernel.cu:294, leaving the memset zeros) and rho = u = 0 in both side
    // buffers, never touched again
    const uint32_t ps = (uint32_t)L.pstride;
    const GPops<T> b0{static_cast<T*>(L.buf[0]) + window<T>(bn)}, b1{static_cast<T*>(L.buf[1]) + window<T>(bn)};
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      b0.put(pidx<T>(q, nr, ps), St::enc(-shift_of<St>(q)));
      b1.put(pidx<T>(q, nr, ps), St::enc(-shift_of<St>(q)));
    }
    if (L.casetype != CASE_CAVITY && (x == 1 || x >= L.nx - 2)) {
      for (int b = 0; b < 2; ++b) {
        const gptr<float> rec = (gptr<float>)L.side[b] + (size_t)(y * L.nz + z) * SIDE_REC;
        if (x == 1) rec[SIDE_RHO] = 0.0f;
        if (x == L.nx - 1) *(gptr<f32x4>)(rec + SIDE_U0) = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        if (x == L.nx - 2) *(gptr<f32x4>)(rec + SIDE_U1) = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
    }
    return;
  }
  float f[Q];
  if (src) {
    const size_t b = (((size_t)(y + 1) * (L.nz + 2) + (z + 1)) * (L.nx + 2) + (x + 1)) * Q;
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = src[b + q];
  } else {
    float ux = 0.f, uy = 0.f, uz = 0.f;
    if (L.init_lid && node_type(L, x, y, z) == NODE_MOVINGWALL) {
      ux = L.ulid[0];
      uy = L.ulid[1];
      uz = L.ulid[2];
    }
    const float usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float eu = edot(q, ux, uy, uz);
      f[q] = W[q] * L.refden * (1.0f + 3.0f * eu + 4.5f * eu * eu - 1.5f * usq);
    }
  }
  const GPops<T> streamed{static_cast<T*>(L.buf[cur ^ 1]) + window<T>(bn)};
  const uint32_t ps = (uint32_t)L.pstride;
#pragma unroll
  for (int q = 0; q < Q; ++q) streamed.put(pidx<T>(q, nr, ps), St::enc(f[q] - shift_of<St>(q)));
  relax_and_store<St, true>(L, f, GPops<T>{static_cast<T*>(L.buf[cur]) + window<T>(bn)}, L.side[cur], nr, x, y, z,
                            mix32(n ^ mix32(step0)));
}

// ---------------------------------------------------------------------------
// Export: the streamed state of the last completed step (stream + boundary
// rules applied to buf[prev]), written as padded AoS arF and/or rho/u, and
// per-workgroup FP64 partial sums of mass, momentum and |rho - rho_prev|.
// ---------------------------------------------------------------------------
constexpr int EXPORT_BS = 256;

__device__ inline double block_sum(double v, double* sh) {
  // deterministic: fixed butterfly within the wave, fixed order across waves
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < EXPORT_BS / 64; ++w) r += sh[w];
  return r;
}

template <class St>
__global__ void __launch_bounds__(EXPORT_BS) k_export(LatticeDev L, int prev, int prev_streamed, float* __restrict__ aos,
                                                      float* __restrict__ rho_o, float* __restrict__ ux_o,
                                                      float* __restrict__ uy_o, float* __restrict__ uz_o,
                                                      double* __restrict__ partials) {
  __shared__ double sh[EXPORT_BS / 64];
  const uint32_t n = blockIdx.x * (uint32_t)EXPORT_BS + threadIdx.x;
  double dm = 0.0, dx = 0.0, dy = 0.0, dz = 0.0, dr = 0.0;
  if (n < L.n && is_solid(L, n)) {  // solid nodes read back as zeros, outside the sums
    int x, y, z;
    coords(L, n, x, y, z);
    if (aos) {
      const size_t b = (((size_t)(y + 1) * (L.nz + 2) + (z + 1)) * (L.nx + 2) + (x + 1)) * Q;
      for (int q = 0; q < Q; ++q) aos[b + q] = 0.0f;
    }
    if (rho_o) rho_o[n] = ux_o[n] = uy_o[n] = uz_o[n] = 0.0f;
  } else if (n < L.n) {
    int x, y, z;
    coords(L, n, x, y, z);
    // buf[prev] viewed from the block of the farthest donor of the
    // workgroup's first node (32-bit offsets from there)
    using T = typename St::T;
    constexpr uint32_t B = 256 / sizeof(T);
    const uint32_t first = blockIdx.x * (uint32_t)EXPORT_BS, reach = (uint32_t)L.nxz + (uint32_t)L.nx;
    const uint32_t bn = (first > reach ? first - reach : 0u) & ~(B - 1), nr = n - bn;
    const GPops<T> pb{static_cast<T*>(L.buf[prev]) + window<T>(bn)};
    const uint32_t ps = (uint32_t)L.pstride;
    float f[Q];
    if (prev_streamed) {
#pragma unroll
      for (int q = 0; q < Q; ++q) f[q] = load_f<St>(pb.get(pidx<T>(q, nr, ps)), q);
    } else {
      gather<St, true>(L, pb, L.side[prev], n, nr, x, y, z, f);
    }
    float rho, jx, jy, jz;
    moments(f, rho, jx, jy, jz);
    // instability: a density that is not a positive finite number marks the
    // lattice (NaN fails both comparisons)
    if (!(rho > 0.0f && rho < __builtin_inff())) L.ctr[CTR_DIVERGED] = 1u;
    const float ir = 1.0f / rho;
    if (aos) {
      const size_t b = (((size_t)(y + 1) * (L.nz + 2) + (z + 1)) * (L.nx + 2) + (x + 1)) * Q;
#pragma unroll
      for (int q = 0; q < Q; ++q) aos[b + q] = f[q];
    }
    if (rho_o) {
      rho_o[n] = rho;
      ux_o[n] = jx * ir;
      uy_o[n] = jy * ir;
      uz_o[n] = jz * ir;
    }
    if (partials) {
      double rn = 0.0, ro = 0.0;
#pragma unroll
      for (int q = 0; q < Q; ++q) rn += (double)f[q];
      if (!prev_streamed) {
        // the collision conserves mass, so sum_q f*_q(prev) is the previous step's density
#pragma unroll
        for (int q = 0; q < Q; ++q) ro += (double)load_f<St>(pb.get(pidx<T>(q, nr, ps)), q);
      }
      dm = rn;
      dx = (double)jx;
      dy = (double)jy;
      dz = (double)jz;
      dr = fabs(rn - ro);
    }
  }
  if (partials) {
    const double a = block_sum(dm, sh);
    const double b = block_sum(dx, sh);
    const double c = block_sum(dy, sh);
    const double d = block_sum(dz, sh);
    const double e = block_sum(dr, sh);
    if (threadIdx.x == 0) {
      double* o = partials + (size_t)blockIdx.x * 5;
      o[0] = a;
      o[1] = b;
      o[2] = c;
      o[3] = d;
      o[4] = e;
    }
  }
}

#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_reduce(const double* __restrict__ partials, int nblocks,
                                                double* __restrict__ out) {
  __shared__ double sh[256];
  for (int c = 0; c < 5; ++c) {
    double v = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += 256) v += partials[(size_t)i * 5 + c];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = sh[0];
    __syncthreads();
  }
}
#endif

// Instability probe over a batch's work list: the rest population of every
// node in the newest state (the buffer the next step reads).  A NaN or
// infinity in any population of a node reaches its rest population through
// the next collision, so a lattice that blows up is caught within a step, for
// one 19th of a buffer's traffic.  (The step kernel itself carries no check:
// one conditional store in it measured 19% slower at FP16 storage.)
template <class St>
__global__ void __launch_bounds__(512) k_probe(const LatticeDev* __restrict__ lats, const WorkItem* __restrict__ items,
                                               int nitems, int parity) {
  using T = typename St::T;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const WorkItem wi = items[it];
    if (threadIdx.x >= wi.count) continue;
    const LatticeDev& L = lats[wi.lat];
    constexpr uint32_t B = 256 / sizeof(T);
    const uint32_t bn = wi.start & ~(B - 1);
    const GPops<T> b{static_cast<T*>(L.buf[parity ^ L.parity0]) + window<T>(bn)};
    const float v = St::dec(b.get(pidx<T>(0, wi.start + threadIdx.x - bn, (uint32_t)L.pstride)));
    if (!__builtin_isfinite(v)) L.ctr[CTR_DIVERGED] = 1u;
  }
}

#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_flags(const LatticeDev* __restrict__ lats, int nlat, uint32_t* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < nlat) out[i] = lats[i].ctr[CTR_DIVERGED];
}
#endif

// keSetNodeType + keInitialNodeDirection (kernel.cu:1020-1097) for read-back.
#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_node_type(LatticeDev L, int* __restrict__ type, int* __restrict__ dirstate) {
  const uint32_t n = blockIdx.x * 256u + threadIdx.x;
  if (n >= L.n) return;
  int x, y, z;
  coords(L, n, x, y, z);
  int ds = 1;
#pragma unroll
  for (int q = 1; q < Q; ++q) {
    const int x1 = x + EX[q], y1 = y + EY[q], z1 = z + EZ[q];
    if (x1 >= 0 && x1 < L.nx && y1 >= 0 && y1 < L.ny && z1 >= 0 && z1 < L.nz) ds |= 1 << q;
  }
  type[n] = is_solid(L, n) ? NODE_STRUCT : node_type(L, x, y, z);
  dirstate[n] = ds;
}
#endif

// Streaming-copy ceiling: `streams` arrays of `per` floats (per % 4 == 0),
// each thread copying one 16-byte vector of every stream, as the step kernel
// touches one element of every population plane.
#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_copy_streams(const float4* __restrict__ src, float4* __restrict__ dst,
                                                      size_t per4, int streams) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= per4) return;
  for (int s = 0; s < streams; ++s) dst[(size_t)s * per4 + i] = src[(size_t)s * per4 + i];
}
#endif

// The same streaming ceiling in the FP32 population layout itself: every
// thread moves one 4-byte element of each of the 19 runs of its 64-node block,
// as k_step's own-node stores do, with the same non-temporal hints.
#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_copy_runs(const float* __restrict__ src, float* __restrict__ dst,
                                                   size_t nblk) {
  const size_t node = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t blk = node / 64, lane = node % 64;
  if (blk >= nblk) return;
  float v[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) v[q] = __builtin_nontemporal_load((gptr<const float>)src + (blk * Q + q) * 64 + lane);
#pragma unroll
  for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(v[q], (gptr<float>)dst + (blk * Q + q) * 64 + lane);
}
#endif

// The 16-bit store conversion of the step kernel applied to an array:
// to nearest (enc) or stochastically with the given seeds (enc_sr).
template <class St>
__global__ void __launch_bounds__(256) k_round16(const float* __restrict__ src, const uint32_t* __restrict__ seeds,
                                                 uint16_t* __restrict__ dst, size_t n, int stochastic) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const typename St::T v = stochastic ? St::enc_sr(src[i], seeds[i]) : St::enc(src[i]);
  dst[i] = __builtin_bit_cast(uint16_t, v);
}

// VALU issue ceiling: eight independent v_fma_f32 chains per lane, 64 FMAs
// per loop iteration, no memory traffic until the single store at the end.
#if NFLOW_STORAGE_SET != 2  // the FP32 object carries it
__global__ void __launch_bounds__(256) k_fma_chains(float* __restrict__ out, float a, float b, int iters) {
  float x[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) x[c] = (float)(threadIdx.x + c);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
#pragma unroll
      for (int c = 0; c < 8; ++c) x[c] = __builtin_fmaf(x[c], a, b);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) s += x[c];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}
#endif

// ---------------------------------------------------------------------------
// Launchers.  This file is compiled twice.  NFLOW_STORAGE_SET 1 (the default)
// builds the FP32 kernels, the kernels that do not depend on the storage
// format and the launch entry points; it is compiled with -ffp-contract=on,
// which the bitwise equality of one- and two-step launches needs (an FMA then
// never spans statements, so the collision rounds the same in every kernel it
// is inlined into).  NFLOW_STORAGE_SET 2 builds the FP16 / BF16 kernels,
// reached through the *_h functions, with the compiler's default contraction
// (fused across statements: 2.5% faster at FP16, and no two-step kernel there
// to match).  Each template instantiation lives in exactly one of the two.
// ---------------------------------------------------------------------------
template <class St, bool MASK, bool SR>
static hipError_t step_bs(int block, int grid, hipStream_t st, const LatticeDev* d_lats, const WorkItem* d_items,
                          int nitems, int parity) {
  switch (block) {
    case 64:
      hipLaunchKernelGGL((k_step<St, 64, MASK, SR>), dim3(grid), dim3(64), 0, st, d_lats, d_items, nitems, parity);
      break;
    case 128:
      hipLaunchKernelGGL((k_step<St, 128, MASK, SR>), dim3(grid), dim3(128), 0, st, d_lats, d_items, nitems, parity);
      break;
    case 256:
      hipLaunchKernelGGL((k_step<St, 256, MASK, SR>), dim3(grid), dim3(256), 0, st, d_lats, d_items, nitems, parity);
      break;
    case 512:
      hipLaunchKernelGGL((k_step<St, 512, MASK, SR>), dim3(grid), dim3(512), 0, st, d_lats, d_items, nitems, parity);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class St, bool SR>
static hipError_t step_mask(bool masked, int block, int grid, hipStream_t st, const LatticeDev* d_lats,
                            const WorkItem* d_items, int nitems, int parity) {
  return masked ? step_bs<St, true, SR>(block, grid, st, d_lats, d_items, nitems, parity)
                : step_bs<St, false, SR>(block, grid, st, d_lats, d_items, nitems, parity);
}

// 16-bit storage (NFLOW_STORAGE_SET 2)
hipError_t launch_step_h(int storage, bool masked, bool sr, int block, int grid, hipStream_t st,
                         const LatticeDev* d_lats, const WorkItem* d_items, int nitems, int parity);
hipError_t launch_init_h(int storage, hipStream_t st, const LatticeDev& L, int cur, const float* src, uint32_t step0);
hipError_t launch_export_h(int storage, hipStream_t st, const LatticeDev& L, int prev, int prev_streamed, float* aos,
                           float* rho, float* ux, float* uy, float* uz, double* partials, int grid);
hipError_t launch_probe_h(int storage, hipStream_t st, const LatticeDev* d_lats, const WorkItem* d_items, int nitems,
                          int grid, int parity);

#if NFLOW_STORAGE_SET == 2

hipError_t launch_step_h(int storage, bool masked, bool sr, int block, int grid, hipStream_t st,
                         const LatticeDev* d_lats, const WorkItem* d_items, int nitems, int parity) {
  switch (storage) {
    case ST_F16:
      return sr ? step_mask<StF16, true>(masked, block, grid, st, d_lats, d_items, nitems, parity)
                : step_mask<StF16, false>(masked, block, grid, st, d_lats, d_items, nitems, parity);
    case ST_BF16:
      return sr ? step_mask<StBF16, true>(masked, block, grid, st, d_lats, d_items, nitems, parity)
                : step_mask<StBF16, false>(masked, block, grid, st, d_lats, d_items, nitems, parity);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_init_h(int storage, hipStream_t st, const LatticeDev& L, int cur, const float* src, uint32_t step0) {
  const int grid = (int)((L.n + 255u) / 256u);
  switch (storage) {
    case ST_F16: hipLaunchKernelGGL(k_init<StF16>, dim3(grid), dim3(256), 0, st, L, cur, src, step0); break;
    case ST_BF16: hipLaunchKernelGGL(k_init<StBF16>, dim3(grid), dim3(256), 0, st, L, cur, src, step0); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_export_h(int storage, hipStream_t st, const LatticeDev& L, int prev, int prev_streamed, float* aos,
                           float* rho, float* ux, float* uy, float* uz, double* partials, int grid) {
  switch (storage) {
    case ST_F16:
      hipLaunchKernelGGL(k_export<StF16>, dim3(grid), dim3(EXPORT_BS), 0, st, L, prev, prev_streamed, aos, rho, ux, uy,
                         uz, partials);
      break;
    case ST_BF16:
      hipLaunchKernelGGL(k_export<StBF16>, dim3(grid), dim3(EXPORT_BS), 0, st, L, prev, prev_streamed, aos, rho, ux,
                         uy, uz, partials);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_probe_h(int storage, hipStream_t st, const LatticeDev* d_lats, const WorkItem* d_items, int nitems,
                          int grid, int parity) {
  switch (storage) {
    case ST_F16: hipLaunchKernelGGL(k_probe<StF16>, dim3(grid), dim3(512), 0, st, d_lats, d_items, nitems, parity); break;
    case ST_BF16:
      hipLaunchKernelGGL(k_probe<StBF16>, dim3(grid), dim3(512), 0, st, d_lats, d_items, nitems, parity);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_round16(hipStream_t st, int storage, const float* src, const uint32_t* seeds, uint16_t* dst, size_t n,
                          int stochastic) {
  const size_t blocks = (n + 255) / 256;
  if (storage == ST_F16)
    hipLaunchKernelGGL(k_round16<StF16>, dim3((unsigned)blocks), dim3(256), 0, st, src, seeds, dst, n, stochastic);
  else if (storage == ST_BF16)
    hipLaunchKernelGGL(k_round16<StBF16>, dim3((unsigned)blocks), dim3(256), 0, st, src, seeds, dst, n, stochastic);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

#else  // NFLOW_STORAGE_SET 1

hipError_t launch_step(int storage, bool masked, bool sr, int block, int grid, hipStream_t st,
                       const LatticeDev* d_lats, int nlat, const WorkItem* d_items, int nitems, int parity) {
  if (grid <= 0 || nitems <= 0) return hipSuccess;
  if (storage != ST_F32) {
    const hipError_t e = launch_step_h(storage, masked, sr, block, grid, st, d_lats, d_items, nitems, parity);
    if (e != hipSuccess || !sr) return e;
    hipLaunchKernelGGL(k_advance, dim3((unsigned)((nlat + 255) / 256)), dim3(256), 0, st, d_lats, nlat, parity);
    return hipGetLastError();
  }
  return step_mask<StF32, false>(masked, block, grid, st, d_lats, d_items, nitems, parity);
}

hipError_t launch_step2(int storage, hipStream_t st, const LatticeDev& L, int parity, int H) {
  using G = step2::Geo<STEP2_TX, STEP2_TZ, STEP2_NT>;
  // a workgroup's window (planes ya - 2 .. ya + H + 1) takes 32-bit byte offsets
  if (H < 1 || storage != ST_F32 || ((unsigned long long)(H + 4) * L.nxz + 64) * Q * 4 >= (1ull << 32))
    return hipErrorInvalidValue;
  const int ntx = (L.nx + G::TX - 1) / G::TX, ntz = (L.nz + G::TZ - 1) / G::TZ, nch = (L.ny + H - 1) / H;
  const dim3 grid((unsigned)(ntx * ntz * nch)), block(G::NT);
  const bool runs = L.nx % 64 == 0;  // whole 64-node blocks per row
  if (L.casetype == CASE_CHANNEL) {
    if (runs)
      hipLaunchKernelGGL((k_step2<StF32, G, true, true>), grid, block, 0, st, L, parity, H, ntx, ntz);
    else
      hipLaunchKernelGGL((k_step2<StF32, G, true, false>), grid, block, 0, st, L, parity, H, ntx, ntz);
  } else {
    if (runs)
      hipLaunchKernelGGL((k_step2<StF32, G, false, true>), grid, block, 0, st, L, parity, H, ntx, ntz);
    else
      hipLaunchKernelGGL((k_step2<StF32, G, false, false>), grid, block, 0, st, L, parity, H, ntx, ntz);
  }
  return hipGetLastError();
}

hipError_t launch_init(int storage, hipStream_t st, const LatticeDev& L, int cur, const float* src, uint32_t step0) {
  if (storage != ST_F32) return launch_init_h(storage, st, L, cur, src, step0);
  const int grid = (int)((L.n + 255u) / 256u);
  hipLaunchKernelGGL(k_init<StF32>, dim3(grid), dim3(256), 0, st, L, cur, src, step0);
  return hipGetLastError();
}

int export_block_size() { return EXPORT_BS; }

hipError_t launch_export(int storage, hipStream_t st, const LatticeDev& L, int prev, int prev_streamed, float* aos,
                         float* rho, float* ux, float* uy, float* uz, double* partials, int* nblocks_out) {
  const int grid = (int)((L.n + EXPORT_BS - 1) / EXPORT_BS);
  if (nblocks_out) *nblocks_out = grid;
  if (storage != ST_F32)
    return launch_export_h(storage, st, L, prev, prev_streamed, aos, rho, ux, uy, uz, partials, grid);
  hipLaunchKernelGGL(k_export<StF32>, dim3(grid), dim3(EXPORT_BS), 0, st, L, prev, prev_streamed, aos, rho, ux, uy, uz,
                     partials);
  return hipGetLastError();
}

hipError_t launch_reduce(hipStream_t st, const double* partials, int nblocks, double* out5) {
  hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, partials, nblocks, out5);
  return hipGetLastError();
}

hipError_t launch_probe(int storage, hipStream_t st, const LatticeDev* d_lats, const WorkItem* d_items, int nitems,
                        int parity) {
  if (nitems <= 0) return hipSuccess;
  const int grid = nitems < 65536 ? nitems : 65536;
  if (storage != ST_F32) return launch_probe_h(storage, st, d_lats, d_items, nitems, grid, parity);
  hipLaunchKernelGGL(k_probe<StF32>, dim3(grid), dim3(512), 0, st, d_lats, d_items, nitems, parity);
  return hipGetLastError();
}

hipError_t launch_flags(hipStream_t st, const LatticeDev* d_lats, int nlat, uint32_t* out) {
  if (nlat <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_flags, dim3((unsigned)((nlat + 255) / 256)), dim3(256), 0, st, d_lats, nlat, out);
  return hipGetLastError();
}

hipError_t launch_node_type(hipStream_t st, const LatticeDev& L, int* type, int* dirstate) {
  const int grid = (int)((L.n + 255u) / 256u);
  hipLaunchKernelGGL(k_node_type, dim3(grid), dim3(256), 0, st, L, type, dirstate);
  return hipGetLastError();
}

hipError_t launch_copy_runs(hipStream_t st, const float* src, float* dst, size_t nblk) {
  const size_t blocks = (nblk * 64 + 255) / 256;
  hipLaunchKernelGGL(k_copy_runs, dim3((unsigned)blocks), dim3(256), 0, st, src, dst, nblk);
  return hipGetLastError();
}

hipError_t launch_fma_chains(hipStream_t st, float* out, int grid, int iters) {
  hipLaunchKernelGGL(k_fma_chains, dim3((unsigned)grid), dim3(256), 0, st, out, 0.999f, 0.001f, iters);
  return hipGetLastError();
}

hipError_t launch_copy_streams(hipStream_t st, const float* src, float* dst, size_t per, int streams) {
  const size_t per4 = per / 4;
  const size_t blocks = (per4 + 255) / 256;
  hipLaunchKernelGGL(k_copy_streams, dim3((unsigned)blocks), dim3(256), 0, st, reinterpret_cast<const float4*>(src),
                     reinterpret_cast<float4*>(dst), per4, streams);
  return hipGetLastError();
}

#endif  // NFLOW_STORAGE_SET

}  // namespace nflow
