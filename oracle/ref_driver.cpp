// ref_driver.cpp -- C-ABI entry points over the reference's own
// decomposition.cpp (decomposition 98-161, slice_chunk 66-94,
// set_move_decomp 17-62) and inverseMRT.cpp (inverseMRT 8-92).
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this driver together
// with those two reference files, where they lie, into oracle/_ref/libref.so.
// Tests use it to pin the oracle's and the product library's decomposition
// and matrix inverse against the reference code itself.  The decomposition
// counts are the reference's compile-time XDECOMP/YDECOMP/ZDECOMP (common.h:17-19).
#include <cstring>
#include <vector>
#include <cmath>
#include "struct.h"

int decomposition(Inputval* pinput, int Slice[][DIMENSION][2]);
int set_move_decomp(const int NTILE, const int Slice[][DIMENSION][2], int Move[][DIRECTION_SIZE]);
int slice_chunk(const unsigned tile, const int Slice[][DIMENSION][2], int* xs, int* xe, int* ys, int* ye, int* zs,
                int* ze, int* chunk_nx, int* chunk_ny, int* chunk_nz, int* chunk_len);
void inverseMRT(std::vector<std::vector<float>>& matori, std::vector<std::vector<float>>& pInvMat);

extern "C" {

void ref_decomp_counts(int* xd, int* yd, int* zd) {
  *xd = XDECOMP;
  *yd = YDECOMP;
  *zd = ZDECOMP;
}

int ref_decomposition(int nx, int ny, int nz, int* slice_out, int max_tiles) {
  static int Slice[TILE_NUM][DIMENSION][2];
  std::memset(Slice, 0, sizeof(Slice));
  Inputval in{};
  in.nx = nx;
  in.ny = ny;
  in.nz = nz;
  const int nt = decomposition(&in, Slice);
  if (nt > max_tiles) return -1;
  std::memcpy(slice_out, Slice, sizeof(int) * DIMENSION * 2 * nt);
  return nt;
}

int ref_set_move_decomp(int ntile, const int* slice, int* move) {
  return set_move_decomp(ntile, reinterpret_cast<const int(*)[DIMENSION][2]>(slice),
                         reinterpret_cast<int(*)[DIRECTION_SIZE]>(move));
}

int ref_slice_chunk(int tile, const int* slice, int* out10) {
  return slice_chunk((unsigned)tile, reinterpret_cast<const int(*)[DIMENSION][2]>(slice), &out10[0], &out10[1],
                     &out10[2], &out10[3], &out10[4], &out10[5], &out10[6], &out10[7], &out10[8], &out10[9]);
}

void ref_inverse_mrt(const float* m, float* minv) {
  std::vector<std::vector<float>> a(DIRECTION_SIZE, std::vector<float>(DIRECTION_SIZE));
  std::vector<std::vector<float>> b(DIRECTION_SIZE, std::vector<float>(DIRECTION_SIZE, 0.0f));
  for (int i = 0; i < DIRECTION_SIZE; ++i)
    for (int j = 0; j < DIRECTION_SIZE; ++j) a[i][j] = m[i * DIRECTION_SIZE + j];
  inverseMRT(a, b);
  for (int i = 0; i < DIRECTION_SIZE; ++i)
    for (int j = 0; j < DIRECTION_SIZE; ++j) minv[i * DIRECTION_SIZE + j] = b[i][j];
}

}  // extern "C"
