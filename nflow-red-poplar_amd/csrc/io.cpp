// io.cpp -- host-side output writers of the C ABI (no device code).
//
// nflow_write_vtu: the reference's binary VTK unstructured-grid output
// (saveFile, kernel.cu:699-970): points, hexahedral cells between
// neighbouring nodes, velocity, density and the six node-flag fields, as
// appended raw data with UInt64 block headers.  The reference declares its
// points and fields Float64 but writes 4-byte floats with an 8-byte stride
// (kernel.cu:716, 874-876, 906-914); here every array is written with the
// type its header declares (Float32).
// nflow_write_csv: result.csv of kernel.cu:972-997.
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>

#include "checkpoint.hpp"
#include "errors.hpp"
#include "nflow.h"

using namespace nflow;

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// (an empty block -- no cells in a box one node thick -- may come with a null
// pointer, which fwrite must not see even for zero bytes)
bool put(FILE* f, const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; }
bool put_str(FILE* f, const std::string& s) { return put(f, s.data(), s.size()); }

}  // namespace

extern "C" int nflow_write_vtu(const char* path, int nx, int ny, int nz, const float* rho, const float* ux,
                               const float* uy, const float* uz, const int* node_type) {
  if (!path || !rho || !ux || !uy || !uz || nx < 1 || ny < 1 || nz < 1)
    return fail(NFLOW_EINVAL, "nflow_write_vtu: bad argument");
  const uint64_t npts = (uint64_t)nx * ny * nz;
  // one hexahedron per node that has +x, +y and +z neighbours (kernel.cu:29-58)
  std::vector<int32_t> conn;
  conn.reserve((size_t)(nx > 1 && ny > 1 && nz > 1 ? (nx - 1) * (ny - 1) * (nz - 1) : 0) * 8);
  for (int y = 0; y + 1 < ny; ++y)
    for (int z = 0; z + 1 < nz; ++z)
      for (int x = 0; x + 1 < nx; ++x)
        for (int yy = y; yy < y + 2; ++yy) {
          const int c[4][2] = {{x, z}, {x + 1, z}, {x + 1, z + 1}, {x, z + 1}};
          for (auto& xz : c) conn.push_back((int32_t)(xz[0] + (int64_t)nx * (xz[1] + (int64_t)nz * yy)));
        }
  const uint64_t ncell = conn.size() / 8;
  const uint64_t b_pts = npts * 3 * sizeof(float), b_conn = conn.size() * sizeof(int32_t);
  const uint64_t b_off = ncell * sizeof(int32_t), b_types = ncell, b_vec = npts * 3 * sizeof(float);
  const uint64_t b_scal = npts * sizeof(float), b_flag = npts * sizeof(int32_t);
  static const char* flag_names[6] = {"Fluid", "Fixed Wall", "Moving Wall", "Inlet", "Outlet", "Struct"};
  static const int flag_bits[6] = {NFLOW_NODE_FLUID, NFLOW_NODE_FIXEDWALL, NFLOW_NODE_MOVINGWALL,
                                   NFLOW_NODE_INLET, NFLOW_NODE_OUTLET, NFLOW_NODE_STRUCT};
  const int nflags = node_type ? 6 : 0;

  std::string h;
  uint64_t off = 0;
  auto array = [&](const char* type, const char* name, int comps, uint64_t bytes) {
    h += std::string("<DataArray type=\"") + type + "\" Name=\"" + name + "\" NumberOfComponents=\"" +
         std::to_string(comps) + "\" format=\"appended\" offset=\"" + std::to_string(off) + "\"/>\n";
    off += bytes + sizeof(uint64_t);
  };
  h += "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n";
  h += "<UnstructuredGrid>\n<Piece NumberOfPoints=\"" + std::to_string(npts) + "\" NumberOfCells=\"" +
       std::to_string(ncell) + "\">\n<Points>\n";
  array("Float32", "Points", 3, b_pts);
  h += "</Points>\n<Cells>\n";
  array("Int32", "connectivity", 1, b_conn);
  array("Int32", "offsets", 1, b_off);
  array("UInt8", "types", 1, b_types);
  h += "</Cells>\n<PointData>\n";
  array("Float32", "Velocity", 3, b_vec);
  array("Float32", "Density", 1, b_scal);
  for (int k = 0; k < nflags; ++k) array("Int32", flag_names[k], 1, b_flag);
  h += "</PointData>\n</Piece>\n</UnstructuredGrid>\n<AppendedData encoding=\"raw\">\n_";

  File out(path, "wb");
  if (!out.f) return fail(NFLOW_EINVAL, "cannot open %s for writing", path);
  bool ok = put_str(out.f, h);
  auto block = [&](const void* p, uint64_t bytes) { ok = ok && put(out.f, &bytes, 8) && put(out.f, p, bytes); };
  {
    std::vector<float> pts((size_t)npts * 3);
    for (uint64_t i = 0; i < npts; ++i) {
      pts[3 * i + 0] = (float)(i % nx);
      pts[3 * i + 1] = (float)(i / ((uint64_t)nx * nz));
      pts[3 * i + 2] = (float)((i / nx) % nz);
    }
    block(pts.data(), b_pts);
  }
  block(conn.data(), b_conn);
  {
    std::vector<int32_t> offs((size_t)ncell);
    for (uint64_t i = 0; i < ncell; ++i) offs[i] = (int32_t)(8 * (i + 1));
    block(offs.data(), b_off);
    std::vector<uint8_t> types((size_t)ncell, 12);  // VTK_HEXAHEDRON
    block(types.data(), b_types);
  }
  {
    std::vector<float> vel((size_t)npts * 3);
    for (uint64_t i = 0; i < npts; ++i) {
      vel[3 * i] = ux[i];
      vel[3 * i + 1] = uy[i];
      vel[3 * i + 2] = uz[i];
    }
    block(vel.data(), b_vec);
  }
  block(rho, b_scal);
  for (int k = 0; k < nflags; ++k) {
    std::vector<int32_t> fl((size_t)npts);
    for (uint64_t i = 0; i < npts; ++i) fl[i] = (node_type[i] & flag_bits[k]) ? 1 : 0;
    block(fl.data(), b_flag);
  }
  ok = ok && put_str(out.f, "\n</AppendedData>\n</VTKFile>\n");
  if (!ok) return fail(NFLOW_EINVAL, "write to %s failed", path);
  return 0;
}

extern "C" int nflow_write_csv(const char* path, int nx, int ny, int nz, const float* rho, const float* ux,
                               const float* uy, const float* uz) {
  if (!path || !rho || !ux || !uy || !uz || nx < 1 || ny < 1 || nz < 1)
    return fail(NFLOW_EINVAL, "nflow_write_csv: bad argument");
  File out(path, "wt");
  if (!out.f) return fail(NFLOW_EINVAL, "cannot open %s for writing", path);
  bool ok = std::fprintf(out.f, "%s\n", "x,y,z,rho,u,v,w") > 0;
  const uint64_t n = (uint64_t)nx * ny * nz;
  for (uint64_t i = 0; i < n && ok; ++i) {
    const int x = (int)(i % nx), z = (int)((i / nx) % nz), y = (int)(i / ((uint64_t)nx * nz));
    ok = std::fprintf(out.f, "%d, %d, %d, %lf, %lf, %lf, %lf\n", x, y, z, rho[i], ux[i], uy[i], uz[i]) > 0;
  }
  if (!ok) return fail(NFLOW_EINVAL, "write to %s failed", path);
  return 0;
}

// ---- checkpoint headers (format: checkpoint.hpp) -------------------------------
namespace nflow {

uint64_t params_hash(const nflow_input& in) {
  uint64_t h = 0xCBF29CE484222325ull;
  auto add = [&h](const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
  };
  const int ints[] = {in.casetype, in.enable_mrt, in.enable_les, in.stochastic_rounding};
  const double dbl[] = {in.tau,    in.smagorinsky, in.ref_den,     in.ref_vel[0], in.ref_vel[1], in.ref_vel[2],
                        in.w_eps,  in.w_epsj,      in.w_xx,        in.inlet_peak, in.perturb,    in.pipe_radius,
                        in.pipe_cy, in.pipe_cz};
  add(ints, sizeof(ints));
  add(dbl, sizeof(dbl));
  add(&in.seed, sizeof(in.seed));
  return h;
}

int read_ckp_header(FILE* f, const char* path, CkpHeader& h) {
  if (std::fread(&h, sizeof(h), 1, f) != 1) return fail(NFLOW_EINVAL, "%s: short checkpoint header", path);
  if (std::memcmp(h.magic, CKP_MAGIC, 8) != 0 || h.header_bytes != 64 || (h.version != 1 && h.version != CKP_VERSION))
    return fail(NFLOW_EINVAL, "%s is not an nflow checkpoint", path);
  if (h.nx < 1 || h.ny < 1 || h.nz < 1 || h.steps < 0) return fail(NFLOW_EINVAL, "%s: corrupt header", path);
  return 0;
}

}  // namespace nflow

extern "C" int nflow_checkpoint_info(const char* path, int* nx, int* ny, int* nz, long long* steps) {
  using namespace nflow;
  if (!path) return fail(NFLOW_EINVAL, "null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(NFLOW_EINVAL, "cannot open %s", path);
  CkpHeader h;
  const int rc = read_ckp_header(f, path, h);
  std::fclose(f);
  if (rc) return rc;
  if (nx) *nx = h.nx;
  if (ny) *ny = h.ny;
  if (nz) *nz = h.nz;
  if (steps) *steps = h.steps;
  return 0;
}
