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
    for (int p = 0; p < n - 1; ++p)
#pragma unroll 1
      for (int q = p + 1; q < n; ++q) {
        const double xr = are[p * n + q], xi = aim[p * n + q];
        const double r = sqrt(xr * xr + xi * xi);
        if (r == 0.0) continue;
        // e^{i phi} = a_pq / |a_pq|
        const double er = xr / r, ei = xi / r;
        const double tau = (are[q * n + q] - are[p * n + p]) / (2.0 * r);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        // G = diag(1, e^{-i phi}) [[c, s], [-s, c]] on columns p, q:
        //   col_p' = c col_p - s e^{-i phi} col_q,  col_q' = s col_p + c e^{-i phi} col_q
#pragma unroll 1
        for (int k = 0; k < n; ++k) {
          const double pr = are[k * n + p], pi = aim[k * n + p];
          const double qr = are[k * n + q], qi = aim[k * n + q];
          const double qer = qr * er + qi * ei, qei = qi * er - qr * ei;  // a_kq e^{-i phi}
          are[k * n + p] = c * pr - s * qer;
          aim[k * n + p] = c * pi - s * qei;
          are[k * n + q] = s * pr + c * qer;
          aim[k * n + q] = s * pi + c * qei;
          const double vpr = vre[k * n + p], vpi = vim[k * n + p];
          const double vqr = vre[k * n + q], vqi = vim[k * n + q];
          const double vqer = vqr * er + vqi * ei, vqei = vqi * er - vqr * ei;
          vre[k * n + p] = c * vpr - s * vqer;
          vim[k * n + p] = c * vpi - s * vqei;
          vre[k * n + q] = s * vpr + c * vqer;
          vim[k * n + q] = s * vpi + c * vqei;
        }
        //   row_p' = c row_p - s e^{i phi} row_q,  row_q' = s row_p + c e^{i phi} row_q
#pragma unroll 1
        for (int k = 0; k < n; ++k) {
          const double pr = are[p * n + k], pi = aim[p * n + k];
          const double qr = are[q * n + k], qi = aim[q * n + k];
          const double qer = qr * er - qi * ei, qei = qi * er + qr * ei;  // a_qk e^{i phi}
          are[p * n + k] = c * pr - s * qer;
          aim[p * n + k] = c * pi - s * qei;
          are[q * n + k] = s * pr + c * qer;
          aim[q * n + k] = s * pi + c * qei;
        }
        are[p * n + q] = aim[p * n + q] = are[q * n + p] = aim[q * n + p] = 0.0;
        aim[p * n + p] = aim[q * n + q] = 0.0;
      }
  }
  return sweep;
}

// hermitian_jacobi executed by one wavefront: the same rotations in the same order, lane k
// updating row k's entries of columns p, q (and of V), then, after a wavefront fence, row p's
// and q's entries of column k.  Every entry sees exactly the arithmetic of the serial version,
// so the results are identical; the serial dependency is one rotation, not one element.
// Call from all 64 lanes of one wavefront (n <= 8; other lanes idle in the updates).
__device__ inline int hermitian_jacobi_wave(int n, double* are, double* aim, double* vre, double* vim,
                                            int max_sweeps = 30) {
  const int k = static_cast<int>(threadIdx.x & 63);
  auto sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  if (k < n)
    for (int j = 0; j < n; ++j) {
      vre[k * n + j] = (k == j) ? 1.0 : 0.0;
      vim[k * n + j] = 0.0;
    }
  sync();
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
    for (int p = 0; p < n - 1; ++p)
#pragma unroll 1
      for (int q = p + 1; q < n; ++q) {
        const double xr = are[p * n + q], xi = aim[p * n + q];
        const double r = sqrt(xr * xr + xi * xi);
        if (r == 0.0) continue;
        const double er = xr / r, ei = xi / r;
        const double tau = (are[q * n + q] - are[p * n + p]) / (2.0 * r);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        sync();  // every lane has read the pivot before it changes
        if (k < n) {
          const double pr = are[k * n + p], pi = aim[k * n + p];
          const double qr = are[k * n + q], qi = aim[k * n + q];
          const double qer = qr * er + qi * ei, qei = qi * er - qr * ei;
          are[k * n + p] = c * pr - s * qer;
          aim[k * n + p] = c * pi - s * qei;
          are[k * n + q] = s * pr + c * qer;
          aim[k * n + q] = s * pi + c * qei;
          const double vpr = vre[k * n + p], vpi = vim[k * n + p];
          const double vqr = vre[k * n + q], vqi = vim[k * n + q];
          const double vqer = vqr * er + vqi * ei, vqei = vqi * er - vqr * ei;
          vre[k * n + p] = c * vpr - s * vqer;
          vim[k * n + p] = c * vpi - s * vqei;
          vre[k * n + q] = s * vpr + c * vqer;
          vim[k * n + q] = s * vpi + c * vqei;
        }
        sync();
        if (k < n) {
          const double pr = are[p * n + k], pi = aim[p * n + k];
          const double qr = are[q * n + k], qi = aim[q * n + k];
          const double qer = qr * er - qi * ei, qei = qi * er + qr * ei;
          are[p * n + k] = c * pr - s * qer;
          aim[p * n + k] = c * pi - s * qei;
          are[q * n + k] = s * pr + c * qer;
          aim[q * n + k] = s * pi + c * qei;
        }
        sync();
        if (k == 0) {
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
