// checkpoint.hpp -- restart-file format (host only).
//
// Little-endian: 8-byte magic "NFLWCKP1", int32 header size (64), int32
// version, int32 nx, ny, nz, casetype, int64 steps, int32 storage, then
// (version 2) uint64 params: FNV-1a of every nflow_input field the step uses,
// zero padding to 64 bytes, then arF as float32 [ny+2][nz+2][nx+2][19].
// Version 1 files (no params word) still load, unchecked.
#pragma once
#include <cstdint>
#include <cstdio>

#include "nflow.h"

namespace nflow {

constexpr char CKP_MAGIC[8] = {'N', 'F', 'L', 'W', 'C', 'K', 'P', '1'};
constexpr int32_t CKP_VERSION = 2;
#pragma pack(push, 4)
struct CkpHeader {
  char magic[8];
  int32_t header_bytes, version, nx, ny, nz, casetype;
  int64_t steps;
  int32_t storage;
  uint64_t params;
  char pad[12];
};
#pragma pack(pop)
static_assert(sizeof(CkpHeader) == 64, "checkpoint header is 64 bytes");

// FNV-1a (64-bit) over the fields that change what a step computes; run
// lengths, output cadence and the initial-state flag are left out.
uint64_t params_hash(const nflow_input& in);
// Reads and checks the 64-byte header (magic, size, version, extents, steps).
int read_ckp_header(FILE* f, const char* path, CkpHeader& h);

}  // namespace nflow
