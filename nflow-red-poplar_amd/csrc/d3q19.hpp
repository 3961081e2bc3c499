// d3q19.hpp -- D3Q19 velocity set, weights and the moment basis, as
// compile-time tables shared by host and device code.
//
// Velocity order, opposite map and weights follow main.cpp:87-146 (the
// reference's [XDIM/YDIM/ZDIM][q] tables); the moment basis is the
// d'Humieres Gram-Schmidt basis whose literal is main.cpp:152-170, written
// here as its polynomial rows.  The rows are mutually orthogonal, so
// M^-1 = M^T diag(1/|row_k|^2) and the back transform needs only M itself.
#pragma once

namespace d3q19 {

constexpr int Q = 19;
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int EY[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
constexpr int EZ[Q] = {0, 0, 0, -1, 1, 0, 0, -1, -1, 1, 1, 0, 0, 0, 0, -1, 1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
constexpr float W[Q] = {1.0f / 3.0f,  1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f,
                        1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                        1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                        1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};

// Opposite pairs (a, OPP[a]); every moment row is even or odd under e -> -e.
constexpr int NPAIR = 9;
constexpr int PA[NPAIR] = {1, 3, 5, 7, 8, 11, 12, 15, 16};
constexpr int PB[NPAIR] = {2, 4, 6, 10, 9, 14, 13, 18, 17};

constexpr int mrow(int k, int q) {
  const int x = EX[q], y = EY[q], z = EZ[q];
  const int e2 = x * x + y * y + z * z;
  switch (k) {
    case 0: return 1;
    case 1: return 19 * e2 - 30;
    case 2: return (21 * e2 * e2 - 53 * e2 + 24) / 2;
    case 3: return x;
    case 4: return (5 * e2 - 9) * x;
    case 5: return y;
    case 6: return (5 * e2 - 9) * y;
    case 7: return z;
    case 8: return (5 * e2 - 9) * z;
    case 9: return 3 * x * x - e2;
    case 10: return (3 * e2 - 5) * (3 * x * x - e2);
    case 11: return y * y - z * z;
    case 12: return (3 * e2 - 5) * (y * y - z * z);
    case 13: return x * y;
    case 14: return y * z;
    case 15: return x * z;
    case 16: return x * (y * y - z * z);
    case 17: return y * (z * z - x * x);
    default: return z * (x * x - y * y);
  }
}

// |row_k|^2 = (M M^T)_kk
constexpr int MNORM2[Q] = {19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8};
// rows that are odd under e -> -e
constexpr bool ODD[Q] = {false, false, false, true,  true,  true,  true,  true,  true, false,
                         false, false, false, false, false, false, true,  true,  true};

}  // namespace d3q19
