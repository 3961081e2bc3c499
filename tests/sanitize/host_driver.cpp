// The library's host-only code built with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitize_cpu.py): the decomposition
// API, the moment basis and its inverse, the VTU/CSV writers and the
// checkpoint-header parser, the latter on valid, truncated and random files.
// Prints "host ok" and exits 0 when every call returned what it should; the
// sanitizers abort on the first memory or undefined-behaviour error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "checkpoint.hpp"
#include "nflow.h"

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      std::exit(1);                                                      \
    }                                                                    \
  } while (0)

static void decompositions() {
  nflow_input in;
  nflow_input_defaults(&in);
  const int boxes[][3] = {{1, 1, 1}, {10, 10, 10}, {7, 5, 3}, {33, 17, 9}, {64, 3, 2}, {2, 2, 40}};
  std::vector<int> slice(64 * 3 * 2), move(64 * 19);
  for (const auto& b : boxes)
    for (int xd = 1; xd <= 4; ++xd)
      for (int yd = 1; yd <= 4; ++yd)
        for (int zd = 1; zd <= 4; ++zd) {
          in.nx = b[0];
          in.ny = b[1];
          in.nz = b[2];
          auto* sl = reinterpret_cast<int(*)[3][2]>(slice.data());
          const int nt = nflow_decomposition(&in, xd, yd, zd, sl, 64);
          CHECK(nt == xd * yd * zd);
          long long vol = 0;
          for (int t = 0; t < nt; ++t) {
            int v[10];
            const int empty = nflow_slice_chunk((unsigned)t, sl, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7],
                                                &v[8], &v[9]);
            if (!empty) vol += (long long)v[6] * v[7] * v[8];
          }
          CHECK(vol == (long long)b[0] * b[1] * b[2]);
          CHECK(nflow_set_move_decomp(nt, sl, reinterpret_cast<int(*)[19]>(move.data())) == 0);
          CHECK(nflow_decomposition(&in, xd, yd, zd, sl, nt - 1) < 0 || nt == 1);  // too few slots
        }
}

static void moment_basis() {
  float m[19][19], mi[19][19];
  nflow_mrt_matrix(m);
  CHECK(nflow_inverse_mrt(m, mi) == 0);
  for (int i = 0; i < 19; ++i)
    for (int j = 0; j < 19; ++j) {
      double s = 0.0;
      for (int k = 0; k < 19; ++k) s += (double)m[i][k] * mi[k][j];
      CHECK(std::fabs(s - (i == j ? 1.0 : 0.0)) < 1e-4);
    }
  float z[19][19] = {};
  CHECK(nflow_inverse_mrt(z, mi) == NFLOW_EINVAL);
}

static void writers(const std::string& dir) {
  std::mt19937 rng(1);
  std::uniform_real_distribution<float> u(-1.0f, 1.0f);
  const int dims[][3] = {{3, 2, 4}, {1, 1, 1}, {5, 1, 2}};
  for (const auto& d : dims) {
    const size_t n = (size_t)d[0] * d[1] * d[2];
    std::vector<float> r(n), a(n), b(n), c(n);
    std::vector<int> t(n);
    for (size_t i = 0; i < n; ++i) {
      r[i] = 1.0f + 0.01f * u(rng);
      a[i] = u(rng);
      b[i] = u(rng);
      c[i] = u(rng);
      t[i] = 1 << (1 + (int)(i % 6));
    }
    const std::string p = dir + "/f.vtu", q = dir + "/f.csv";
    CHECK(nflow_write_vtu(p.c_str(), d[0], d[1], d[2], r.data(), a.data(), b.data(), c.data(), t.data()) == 0);
    CHECK(nflow_write_vtu(p.c_str(), d[0], d[1], d[2], r.data(), a.data(), b.data(), c.data(), nullptr) == 0);
    CHECK(nflow_write_csv(q.c_str(), d[0], d[1], d[2], r.data(), a.data(), b.data(), c.data()) == 0);
  }
  float x = 0.0f;
  CHECK(nflow_write_csv((dir + "/no/such/dir/f.csv").c_str(), 1, 1, 1, &x, &x, &x, &x) < 0);
}

static void write_file(const std::string& path, const void* p, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  CHECK(f);
  if (n) CHECK(std::fwrite(p, 1, n, f) == n);
  std::fclose(f);
}

static void checkpoint_headers(const std::string& dir) {
  const std::string path = dir + "/c.ckp";
  nflow::CkpHeader h{};
  std::memcpy(h.magic, nflow::CKP_MAGIC, 8);
  h.header_bytes = 64;
  h.version = nflow::CKP_VERSION;
  h.nx = 4;
  h.ny = 3;
  h.nz = 2;
  h.casetype = NFLOW_CAVITY;
  h.steps = 17;
  nflow_input in;
  nflow_input_defaults(&in);
  h.params = nflow::params_hash(in);
  write_file(path, &h, sizeof(h));
  int nx = 0, ny = 0, nz = 0;
  long long steps = 0;
  CHECK(nflow_checkpoint_info(path.c_str(), &nx, &ny, &nz, &steps) == 0);
  CHECK(nx == 4 && ny == 3 && nz == 2 && steps == 17);
  nflow_input other = in;
  other.tau += 1e-9;
  CHECK(nflow::params_hash(other) != h.params);
  for (size_t len = 0; len < sizeof(h); ++len) {  // every truncation
    write_file(path, &h, len);
    CHECK(nflow_checkpoint_info(path.c_str(), &nx, &ny, &nz, &steps) == NFLOW_EINVAL);
  }
  std::mt19937_64 rng(7);
  int accepted = 0;
  for (int k = 0; k < 4000; ++k) {  // random headers, half with a valid magic
    unsigned char buf[96];
    for (auto& c : buf) c = (unsigned char)rng();
    if (k & 1) {
      std::memcpy(buf, nflow::CKP_MAGIC, 8);
      const int32_t hb = 64, ver = (k & 2) ? 1 : 2;
      std::memcpy(buf + 8, &hb, 4);
      std::memcpy(buf + 12, &ver, 4);
    }
    write_file(path, buf, 64 + rng() % 33);
    const int rc = nflow_checkpoint_info(path.c_str(), &nx, &ny, &nz, &steps);
    CHECK(rc == 0 || rc == NFLOW_EINVAL);
    if (rc == 0) {
      CHECK(nx >= 1 && ny >= 1 && nz >= 1 && steps >= 0);
      ++accepted;
    }
  }
  CHECK(accepted > 0);
  CHECK(nflow_checkpoint_info((dir + "/missing.ckp").c_str(), &nx, &ny, &nz, &steps) == NFLOW_EINVAL);
  CHECK(nflow_checkpoint_info(nullptr, &nx, &ny, &nz, &steps) == NFLOW_EINVAL);
}

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::string dir = argv[1];
  decompositions();
  moment_basis();
  writers(dir);
  checkpoint_headers(dir);
  std::printf("host ok\n");
  return 0;
}
