// Per resource element linear detection: ZF (pseudo-inverse for full column
// rank H) and MMSE, both as the Hermitian system (H^H H + s2 I) x = H^H y
// solved by a float32 Cholesky factorisation followed by one float32
// iterative-refinement step x += G^{-1} (H^H (y - H x) - s2 x).
#pragma once

#include "hpx_common.hpp"

namespace hpx {

// Lower Cholesky factor of an NT x NT Hermitian positive definite matrix held
// as its lower triangle g[i][j], i >= j (in place).  Diagonal stored real in .x;
// returns false if a pivot is not positive.
template <int NT>
__device__ __forceinline__ bool cholesky(float2 (&g)[NT][NT], float (&inv_diag)[NT]) {
  bool ok = true;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    float d = g[j][j].x;
#pragma unroll
    for (int k = 0; k < j; ++k) d -= cabs2(g[j][k]);
    ok = ok && (d > 0.0f);
    const float inv = rsqrtf(fmaxf(d, 1e-30f));
    inv_diag[j] = inv;
#pragma unroll
    for (int i = j + 1; i < NT; ++i) {
      float2 s = g[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s = csub(s, cmul(g[i][k], make_float2(g[j][k].x, -g[j][k].y)));
      g[i][j] = cscale(s, inv);
    }
  }
  return ok;
}

// Solve L L^H x = b with the factor of cholesky().
template <int NT>
__device__ __forceinline__ void cholesky_solve(const float2 (&l)[NT][NT], const float (&inv_diag)[NT],
                                               float2 (&b)[NT]) {
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    float2 s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s = csub(s, cmul(l[i][k], b[k]));
    b[i] = cscale(s, inv_diag[i]);
  }
#pragma unroll
  for (int i = NT - 1; i >= 0; --i) {
    float2 s = b[i];
#pragma unroll
    for (int k = i + 1; k < NT; ++k) s = csub(s, cmulc(l[k][i], b[k]));
    b[i] = cscale(s, inv_diag[i]);
  }
}

// Detection with H supplied row by row through hrow(r, t).  NT transmit
// streams, nr receive antennas (nr >= NT).  s2 = 0 gives ZF.
template <int NT, class HRow, class YAt>
__device__ __forceinline__ void detect_rows(int nr, HRow hrow, YAt yat, float s2, float2 (&x)[NT]) {
  float2 g[NT][NT];
  float2 b[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    b[i] = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < NT; ++j) g[i][j] = make_float2(0.f, 0.f);
  }
  for (int r = 0; r < nr; ++r) {
    float2 h[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) h[t] = hrow(r, t);
    const float2 y = yat(r);
#pragma unroll
    for (int i = 0; i < NT; ++i) {
      b[i] = cfmac(h[i], y, b[i]);
#pragma unroll
      for (int j = 0; j <= i; ++j) g[i][j] = cfmac(h[i], h[j], g[i][j]);
    }
  }
  float2 l[NT][NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
#pragma unroll
    for (int j = 0; j < NT; ++j) l[i][j] = g[i][j];
    l[i][i].x += s2;
    l[i][i].y = 0.f;
  }
  float inv_diag[NT];
  cholesky<NT>(l, inv_diag);
#pragma unroll
  for (int i = 0; i < NT; ++i) x[i] = b[i];
  cholesky_solve<NT>(l, inv_diag, x);

  // one refinement step: r = H^H y - (H^H H + s2 I) x = b - G x - s2 x
  float2 res[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    float2 s = csub(b[i], cscale(x[i], s2));
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const float2 gij = (j <= i) ? g[i][j] : make_float2(g[j][i].x, -g[j][i].y);
      s = csub(s, cmul(gij, x[j]));
    }
    res[i] = s;
  }
  cholesky_solve<NT>(l, inv_diag, res);
#pragma unroll
  for (int i = 0; i < NT; ++i) x[i] = cadd(x[i], res[i]);
}

}  // namespace hpx
