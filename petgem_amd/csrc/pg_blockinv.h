// Entity-block Jacobi: inverse of one diagonal block A_BB (the rows and columns
// of one mesh entity's DOFs) of a CSR matrix, shared by the Krylov solver and
// the multi-mesh batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pg_common.h"

namespace pg {

constexpr int kMaxInvBlock = 8;

// Gathers A_BB for DOFs [i0, i0 + d) (each row's columns i0 .. i0 + d - 1 must
// be consecutive CSR entries) and inverts it by Gauss-Jordan elimination with
// partial pivoting; writes the inverse row-major to out.  Returns false for a
// block it cannot gather or a singular block (out is then left unwritten).
__device__ inline bool invert_block(const int32_t* __restrict__ rowptr,
                                    const int32_t* __restrict__ cols,
                                    const double2* __restrict__ vals, int i0, int d,
                                    double2* __restrict__ out) {
  if (d <= 0 || d > kMaxInvBlock) return false;
  double2 a[kMaxInvBlock][kMaxInvBlock], inv[kMaxInvBlock][kMaxInvBlock];
  for (int t = 0; t < d; ++t) {
    const int row = i0 + t;
    int lo = rowptr[row], hi = rowptr[row + 1] - 1, pos = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int c = cols[mid];
      if (c == i0) {
        pos = mid;
        break;
      }
      if (c < i0) lo = mid + 1;
      else hi = mid - 1;
    }
    if (pos < 0 || pos + d > rowptr[row + 1] || cols[pos + d - 1] != i0 + d - 1) return false;
    for (int u = 0; u < d; ++u) {
      a[t][u] = vals[pos + u];
      inv[t][u] = make_double2(t == u ? 1.0 : 0.0, 0.0);
    }
  }
  for (int c = 0; c < d; ++c) {
    int piv = c;
    double best = hypot(a[c][c].x, a[c][c].y);
    for (int t = c + 1; t < d; ++t) {
      const double v = hypot(a[t][c].x, a[t][c].y);
      if (v > best) {
        best = v;
        piv = t;
      }
    }
    if (!(best > 0.0)) return false;
    if (piv != c)
      for (int u = 0; u < d; ++u) {
        const double2 t1 = a[c][u], t2 = inv[c][u];
        a[c][u] = a[piv][u];
        inv[c][u] = inv[piv][u];
        a[piv][u] = t1;
        inv[piv][u] = t2;
      }
    const double2 ip = zdiv(make_double2(1.0, 0.0), a[c][c]);
    for (int u = 0; u < d; ++u) {
      a[c][u] = zmul(a[c][u], ip);
      inv[c][u] = zmul(inv[c][u], ip);
    }
    for (int t = 0; t < d; ++t) {
      if (t == c) continue;
      const double2 f = a[t][c];
      for (int u = 0; u < d; ++u) {
        a[t][u] = zsub(a[t][u], zmul(f, a[c][u]));
        inv[t][u] = zsub(inv[t][u], zmul(f, inv[c][u]));
      }
    }
  }
  for (int t = 0; t < d; ++t)
    for (int u = 0; u < d; ++u) out[t * d + u] = inv[t][u];
  return true;
}

}  // namespace pg
