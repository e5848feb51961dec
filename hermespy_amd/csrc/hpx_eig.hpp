// Eigen-decomposition of a small complex Hermitian matrix (n <= 8) by the
// cyclic Jacobi method in float64 (Golub & Van Loan, "Matrix Computations",
// sec. 8.5, with a phase rotation that makes each off-diagonal pivot real).
//
// Used once per drop: with ideal CSI the reference's MIMO channel of every
// resource element is H_k = c_k S (hermespy/channel/fading/fading.py:360-364 and
// hermespy/simulation/modem/channel_estimation.py:242-247), so with
// S^H S = V diag(lambda) V^H the MMSE/ZF filter of every resource element is
//   W_k = V diag(conj(c_k) / (|c_k|^2 lambda_i + s2)) V^H S^H.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace hpx {

constexpr int kEigMax = 8;

// Parallel-order cyclic Jacobi: a sweep is n' - 1 rounds (n' = n rounded up to even) of n' / 2
// disjoint pivot pairs (the circle method of round-robin tournaments: index 0 fixed, the others
// rotating one place per round; a pair with the padding index n is skipped).  The rotations of a
// round touch disjoint rows and columns, so they commute and their parameters all come from the
// matrix before the round: applied one after another (host) or side by side (one wavefront) they
// give the same result.
__host__ __device__ inline void jacobi_round_pair(int n, int round, int i, int* p, int* q) {
  const int m = n + (n & 1);  // players 0 .. m-1; player n (odd n) sits out
  auto player = [&](int slot) { return slot == 0 ? 0 : 1 + (slot - 1 + round) % (m - 1); };
  int a = player(i), b = player(m - 1 - i);
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  *p = a;
  *q = b;
}

// The rotation that zeroes a_pq of a Hermitian A: with d = a_qq - a_pp, x = a_pq,
// h = sqrt(d^2 + 4 |x|^2), sigma = 2 sgn(d) / (|d| + h) (sgn(0) = 1), c = 1 / sqrt(1 + sigma^2 |x|^2)
// and z = c sigma x, the unitary J = [[c, z], [-conj(z), c]] on columns p, q makes (J^H A J)_pq = 0
// (t = z / (c x / |x|) = sigma |x| is the smaller root of |x| t^2 + d t - |x| = 0).  No division
// by |x| and two square roots in the dependent chain.
struct JacobiRotation {
  double c, zr, zi;
  bool skip;
};

__host__ __device__ inline JacobiRotation jacobi_rotation(double app, double aqq, double xr, double xi) {
  JacobiRotation j;
  const double r2 = xr * xr + xi * xi;
  j.skip = r2 == 0.0;
  const double d = aqq - app;
  const double h = sqrt(d * d + 4.0 * r2);
  const double sigma = (d >= 0.0 ? 2.0 : -2.0) / (fabs(d) + h);
  j.c = 1.0 / sqrt(1.0 + sigma * sigma * r2);
  const double cs = j.c * sigma;
  j.zr = cs * xr;
  j.zi = cs * xi;
  return j;
}

// column update of row k of a matrix (and of V = V J): m_kp' = c m_kp - conj(z) m_kq,
// m_kq' = z m_kp + c m_kq
__host__ __device__ inline void jacobi_cols(const JacobiRotation& j, double* mre, double* mim, int kp, int kq) {
  const double pr = mre[kp], pi = mim[kp], qr = mre[kq], qi = mim[kq];
  mre[kp] = j.c * pr - (j.zr * qr + j.zi * qi);
  mim[kp] = j.c * pi - (j.zr * qi - j.zi * qr);
  mre[kq] = (j.zr * pr - j.zi * pi) + j.c * qr;
  mim[kq] = (j.zr * pi + j.zi * pr) + j.c * qi;
}

// row update of column k (J^H A): a_pk' = c a_pk - z a_qk, a_qk' = conj(z) a_pk + c a_qk
__host__ __device__ inline void jacobi_rows(const JacobiRotation& j, double* are, double* aim, int pk, int qk) {
  const double pr = are[pk], pi = aim[pk], qr = are[qk], qi = aim[qk];
  are[pk] = j.c * pr - (j.zr * qr - j.zi * qi);
  aim[pk] = j.c * pi - (j.zr * qi + j.zi * qr);
  are[qk] = (j.zr * pr + j.zi * pi) + j.c * qr;
  aim[qk] = (j.zr * pi - j.zi * pr) + j.c * qi;
}

// a: row-major n x n Hermitian (re, im interleaved, only read/updated as a
// full matrix); on return the diagonal holds the eigenvalues.  v: n x n
// unitary, columns are eigenvectors.  Returns the number of sweeps used.
__host__ __device__ inline int hermitian_jacobi(int n, double* are, double* aim, double* vre, double* vim,
                                                int max_sweeps = 30) {
#pragma unroll 1
  for (int i = 0; i < n; ++i)
#pragma unroll 1
    for (int j = 0; j < n; ++j) {
      vre[i * n + j] = (i == j) ? 1.0 : 0.0;
      vim[i * n + j] = 0.0;
    }
  const int m = n + (n & 1);
  int sweep = 0;
#pragma unroll 1
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll 1
    for (int i = 0; i < n; ++i) {
      diag += are[i * n + i] * are[i * n + i];
#pragma unroll 1
      for (int j = i + 1; j < n; ++j) off += are[i * n + j] * are[i * n + j] + aim[i * n + j] * aim[i * n + j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
#pragma unroll 1
    for (int round = 0; round < m - 1; ++round) {
      JacobiRotation rot[kEigMax / 2];
      int ps[kEigMax / 2], qs[kEigMax / 2];
#pragma unroll 1
      for (int i = 0; i < m / 2; ++i) {
        jacobi_round_pair(n, round, i, &ps[i], &qs[i]);
        const int p = ps[i], q = qs[i];
        rot[i].skip = true;
        if (q < n) rot[i] = jacobi_rotation(are[p * n + p], are[q * n + q], are[p * n + q], aim[p * n + q]);
      }
#pragma unroll 1
      for (int i = 0; i < m / 2; ++i) {
        if (qs[i] >= n || rot[i].skip) continue;
        const int p = ps[i], q = qs[i];
#pragma unroll 1
        for (int k = 0; k < n; ++k) {
          jacobi_cols(rot[i], are, aim, k * n + p, k * n + q);
          jacobi_cols(rot[i], vre, vim, k * n + p, k * n + q);
        }
      }
#pragma unroll 1
      for (int i = 0; i < m / 2; ++i) {
        if (qs[i] >= n || rot[i].skip) continue;
        const int p = ps[i], q = qs[i];
#pragma unroll 1
        for (int k = 0; k < n; ++k) jacobi_rows(rot[i], are, aim, p * n + k, q * n + k);
        are[p * n + q] = aim[p * n + q] = are[q * n + p] = aim[q * n + p] = 0.0;
        aim[p * n + p] = aim[q * n + q] = 0.0;
      }
    }
  }
  return sweep;
}

// hermitian_jacobi executed by one wavefront: lane (i, k) = (lane / n, lane % n) for the i-th pair of
// a round applies that pair's column update to row k (of A and V), then, after a wavefront fence,
// its row update to column k; the same rotations and arithmetic as the serial version, so the
// results are identical.  Call from all 64 lanes of one wavefront (n <= 8).
__device__ inline int hermitian_jacobi_wave(int n, double* are, double* aim, double* vre, double* vim,
                                            int max_sweeps = 30) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  const int k = lane % n, pair = lane / n;
  auto sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  if (lane < n)
    for (int j = 0; j < n; ++j) {
      vre[lane * n + j] = (lane == j) ? 1.0 : 0.0;
      vim[lane * n + j] = 0.0;
    }
  sync();
  const int m = n + (n & 1);
  int sweep = 0;
#pragma unroll 1
  for (; sweep < max_sweeps; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll 1
    for (int i = 0; i < n; ++i) {
      diag += are[i * n + i] * are[i * n + i];
#pragma unroll 1
      for (int j = i + 1; j < n; ++j) off += are[i * n + j] * are[i * n + j] + aim[i * n + j] * aim[i * n + j];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
#pragma unroll 1
    for (int round = 0; round < m - 1; ++round) {
      int p = 0, q = n;
      if (pair < m / 2) jacobi_round_pair(n, round, pair, &p, &q);
      const bool active = pair < m / 2 && q < n;
      JacobiRotation rot;
      rot.skip = true;
      if (active) rot = jacobi_rotation(are[p * n + p], are[q * n + q], are[p * n + q], aim[p * n + q]);
      sync();  // every lane has read its pivot before the matrix changes
      if (active && !rot.skip) {
        jacobi_cols(rot, are, aim, k * n + p, k * n + q);
        jacobi_cols(rot, vre, vim, k * n + p, k * n + q);
      }
      sync();
      if (active && !rot.skip) jacobi_rows(rot, are, aim, p * n + k, q * n + k);
      sync();
      if (active && !rot.skip && k == 0) {
        are[p * n + q] = aim[p * n + q] = are[q * n + p] = aim[q * n + p] = 0.0;
        aim[p * n + p] = aim[q * n + q] = 0.0;
      }
      sync();
    }
  }
  return sweep;
}

// Doubles of workspace spatial_detector needs.
constexpr int kEigWork = 4 * kEigMax * kEigMax;

// Per-drop linear detector of the rank-one-in-frequency ideal-CSI channel
// H_k = c_k S (S: nr x nt, row-major, float64 interleaved complex):
// lambda (nt), V (nt x nt), A = V^H S^H (nt x nr).  work: kEigWork doubles
// (LDS in the fused kernel, so that no array lands in scratch memory).
__host__ __device__ inline void spatial_detector(int nr, int nt, const double* s_ri, double* lambda, double* v_ri,
                                                 double* a_ri, double* work) {
  double* gre = work;
  double* gim = work + kEigMax * kEigMax;
  double* vre = work + 2 * kEigMax * kEigMax;
  double* vim = work + 3 * kEigMax * kEigMax;
#pragma unroll 1
  for (int i = 0; i < nt; ++i)
#pragma unroll 1
    for (int j = 0; j < nt; ++j) {
      double sr = 0.0, si = 0.0;
#pragma unroll 1
      for (int r = 0; r < nr; ++r) {  // conj(S_ri) S_rj
        const double ar = s_ri[2 * (r * nt + i)], ai = s_ri[2 * (r * nt + i) + 1];
        const double br = s_ri[2 * (r * nt + j)], bi = s_ri[2 * (r * nt + j) + 1];
        sr += ar * br + ai * bi;
        si += ar * bi - ai * br;
      }
      gre[i * nt + j] = sr;
      gim[i * nt + j] = si;
    }
  hermitian_jacobi(nt, gre, gim, vre, vim);
#pragma unroll 1
  for (int i = 0; i < nt; ++i) lambda[i] = gre[i * nt + i];
#pragma unroll 1
  for (int i = 0; i < nt * nt; ++i) {
    v_ri[2 * i] = vre[i];
    v_ri[2 * i + 1] = vim[i];
  }
  // A[i][r] = sum_k conj(V[k][i]) conj(S[r][k])
#pragma unroll 1
  for (int i = 0; i < nt; ++i)
#pragma unroll 1
    for (int r = 0; r < nr; ++r) {
      double ar = 0.0, ai = 0.0;
#pragma unroll 1
      for (int k = 0; k < nt; ++k) {
        const double vr = vre[k * nt + i], vi = -vim[k * nt + i];
        const double sr = s_ri[2 * (r * nt + k)], si = -s_ri[2 * (r * nt + k) + 1];
        ar += vr * sr - vi * si;
        ai += vr * si + vi * sr;
      }
      a_ri[2 * (i * nr + r)] = ar;
      a_ri[2 * (i * nr + r) + 1] = ai;
    }
}


// spatial_detector executed by one wavefront (all 64 lanes call it; the same sums in the
// same order, so the results equal the serial version's): the Gram matrix and A = V^H S^H
// spread over the lanes, the Jacobi rotations by hermitian_jacobi_wave.
__device__ inline void spatial_detector_wave(int nr, int nt, const double* s_ri, double* lambda, double* v_ri,
                                             double* a_ri, double* work) {
  const int lane = static_cast<int>(threadIdx.x & 63);
  auto sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  double* gre = work;
  double* gim = work + kEigMax * kEigMax;
  double* vre = work + 2 * kEigMax * kEigMax;
  double* vim = work + 3 * kEigMax * kEigMax;
  for (int e = lane; e < nt * nt; e += 64) {
    const int i = e / nt, j = e - i * nt;
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < nr; ++r) {  // conj(S_ri) S_rj
      const double ar = s_ri[2 * (r * nt + i)], ai = s_ri[2 * (r * nt + i) + 1];
      const double br = s_ri[2 * (r * nt + j)], bi = s_ri[2 * (r * nt + j) + 1];
      sr += ar * br + ai * bi;
      si += ar * bi - ai * br;
    }
    gre[e] = sr;
    gim[e] = si;
  }
  sync();
  hermitian_jacobi_wave(nt, gre, gim, vre, vim);
  sync();
  for (int i = lane; i < nt; i += 64) lambda[i] = gre[i * nt + i];
  for (int i = lane; i < nt * nt; i += 64) {
    v_ri[2 * i] = vre[i];
    v_ri[2 * i + 1] = vim[i];
  }
  for (int e = lane; e < nt * nr; e += 64) {
    const int i = e / nr, r = e - i * nr;
    double ar = 0.0, ai = 0.0;
    for (int k = 0; k < nt; ++k) {  // A[i][r] = sum_k conj(V[k][i]) conj(S[r][k])
      const double vr = vre[k * nt + i], vi = -vim[k * nt + i];
      const double sr = s_ri[2 * (r * nt + k)], si = -s_ri[2 * (r * nt + k) + 1];
      ar += vr * sr - vi * si;
      ai += vr * si + vi * sr;
    }
    a_ri[2 * (i * nr + r)] = ar;
    a_ri[2 * (i * nr + r) + 1] = ai;
  }
  sync();
}

}  // namespace hpx
