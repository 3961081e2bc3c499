// mrt.hpp -- the two MRT moment transforms of the D3Q19 collision, in
// factored form, for the kernels and for a host test that checks them
// against the moment basis of d3q19.hpp (mrow) on every unit vector.
//
// Each row of the d'Humieres basis is even or odd under e -> -e, so with the
// opposite-pair sums s_i = f_a + f_opp(a) and differences d_i = f_a - f_opp(a)
// (a = PA[i]) the even rows read only s and the odd rows only d.  The rows
// also share terms: the axis-pair sums Sx, Sy, Sz, the edge-group sums Exz,
// Exy, Eyz and their totals; among the odd rows m4 = jx - 5 d0, m6 = jy - 5 d2
// and m8 = jz + 5 d1.  Forming each shared term once takes about half the
// adds of summing row by row, in both directions.  Pair order (d3q19.hpp):
// 0 the x axis (1, 2), 1 z (3, 4), 2 y (5, 6), 3-4 the xz edges (7, 10),
// (8, 9), 5-6 xy (11, 14), (12, 13), 7-8 yz (15, 18), (16, 17).
#pragma once
#include <hip/hip_runtime.h>

#include "d3q19.hpp"

namespace nflow {

// m = M f, all 19 rows (m0 = rho, m3, m5, m7 = j).
__host__ __device__ inline void mrt_moments(const float f[d3q19::Q], float m[d3q19::Q]) {
  using namespace d3q19;
  float s[NPAIR], d[NPAIR];
#pragma unroll
  for (int i = 0; i < NPAIR; ++i) {
    s[i] = f[PA[i]] + f[PB[i]];
    d[i] = f[PA[i]] - f[PB[i]];
  }
  const float sx = s[0], sz = s[1], sy = s[2];
  const float exz = s[3] + s[4], exy = s[5] + s[6], eyz = s[7] + s[8];
  const float a = sx + sy + sz, e = exz + exy + eyz;
  m[0] = f[0] + a + e;
  m[1] = -30.0f * f[0] - 11.0f * a + 8.0f * e;
  m[2] = 12.0f * f[0] - 4.0f * a + e;
  const float t = e - 3.0f * eyz;
  m[9] = 3.0f * sx - a + t;
  m[10] = -6.0f * sx + 2.0f * a + t;
  const float u = sy - sz, v = exy - exz;
  m[11] = u + v;
  m[12] = v - 2.0f * u;
  m[13] = s[5] - s[6];
  m[14] = s[8] - s[7];
  m[15] = s[4] - s[3];
  const float p1 = d[5] - d[6], p2 = d[4] - d[3], q1 = d[5] + d[6], q2 = d[7] + d[8], r1 = d[3] + d[4],
              r2 = d[7] - d[8];
  m[3] = d[0] - p2 + p1;
  m[5] = d[2] + q1 + q2;
  m[7] = -d[1] - r1 - r2;
  m[4] = m[3] - 5.0f * d[0];
  m[6] = m[5] - 5.0f * d[2];
  m[8] = m[7] + 5.0f * d[1];
  m[16] = p2 + p1;
  m[17] = q2 - q1;
  m[18] = r2 - r1;
}

// f -= M^T dl over the non-conserved rows (dl[0], dl[3], dl[5], dl[7] are
// not read): f_a -= E + O and f_opp(a) -= E - O per pair.
__host__ __device__ inline void mrt_back(float f[d3q19::Q], const float dl[d3q19::Q]) {
  using namespace d3q19;
  const float cax = -11.0f * dl[1] - 4.0f * dl[2], ced = 8.0f * dl[1] + dl[2];
  const float g = dl[9] - 2.0f * dl[10], h = dl[11] - 2.0f * dl[12], tt = cax - g;
  const float kk = dl[9] + dl[10], ll = dl[11] + dl[12], uu = ced + kk;
  const float xz = uu - ll, xy = uu + ll, yz = ced - 2.0f * kk;
  const float ev[NPAIR] = {cax + 2.0f * g, tt - h,      tt + h,      xz - dl[15], xz + dl[15],
                           xy + dl[13],    xy - dl[13], yz - dl[14], yz + dl[14]};
  const float oa = dl[4] - dl[16], ob = dl[8] + dl[18], oc = dl[4] + dl[16], oe = dl[6] - dl[17],
              og = dl[6] + dl[17], oh = dl[8] - dl[18];
  const float od[NPAIR] = {-4.0f * dl[4], 4.0f * dl[8], -4.0f * dl[6], oa - ob, -oa - ob,
                           oc + oe,       oe - oc,      og - oh,       og + oh};
  f[0] -= -30.0f * dl[1] + 12.0f * dl[2];
#pragma unroll
  for (int i = 0; i < NPAIR; ++i) {
    f[PA[i]] -= ev[i] + od[i];
    f[PB[i]] -= ev[i] - od[i];
  }
}

}  // namespace nflow
