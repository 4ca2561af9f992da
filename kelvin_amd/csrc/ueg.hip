// Uniform-electron-gas integrals built in HBM, and the finite-temperature
// mean-field quantities every system needs from its integrals.
//
// Spin-orbital <pq||rs> of the plane-wave basis (kelvin/ueg_utils.py:83-138):
// orbitals 0..nk-1 are alpha, nk..2nk-1 beta, both over the same k-points.
// With spatial P, Q, R, S and integer wave vectors X,
//   V[P,Q,R,S] = 4 pi / (L^3 |k_P - k_R|^2)  if X_P + X_Q = X_R + X_S, P != R
//   <pq||rs>   = [s_p = s_r, s_q = s_s] V[P,Q,R,S] - [s_p = s_s, s_q = s_r] V[P,Q,S,R].
// Every product and sum is rounded separately in the order of the
// reference's numpy expression, so the array is bitwise the one built on the
// host.  One thread per element; the n^4 array at n = 114 is 1.35 GB.
#include "ftcc_internal.h"

#include <algorithm>

namespace ftcc {
namespace {

constexpr int UB = 256;

__device__ inline double inv_k2(const double *k, int P, int R, double aaa) {
  const double d0 = __dsub_rn(k[3 * P], k[3 * R]);
  const double d1 = __dsub_rn(k[3 * P + 1], k[3 * R + 1]);
  const double d2 = __dsub_rn(k[3 * P + 2], k[3 * R + 2]);
  const double s = __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)),
                             __dmul_rn(d2, d2));
  return aaa / s;
}

__global__ __launch_bounds__(UB) void ueg_eri_kernel(int nk, const int *X,
                                                     const double *k,
                                                     double aaa, double *eri) {
  const int n = 2 * nk;
  const int64_t total = (int64_t)n * n * n * n;
  for (int64_t e = blockIdx.x * (int64_t)UB + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * UB) {
    int64_t t = e;
    const int s = (int)(t % n);
    t /= n;
    const int r = (int)(t % n);
    t /= n;
    const int q = (int)(t % n);
    const int p = (int)(t / n);
    const int sp = p >= nk, sq = q >= nk, sr = r >= nk, ss = s >= nk;
    const int P = p - sp * nk, Q = q - sq * nk, R = r - sr * nk,
              S = s - ss * nk;
    const bool conserve = X[3 * P] + X[3 * Q] == X[3 * R] + X[3 * S] &&
                          X[3 * P + 1] + X[3 * Q + 1] == X[3 * R + 1] + X[3 * S + 1] &&
                          X[3 * P + 2] + X[3 * Q + 2] == X[3 * R + 2] + X[3 * S + 2];
    double v = 0.0;
    if (conserve) {
      const double dir = (sp == sr && sq == ss && P != R) ? inv_k2(k, P, R, aaa) : 0.0;
      const double exc = (sp == ss && sq == sr && P != S) ? inv_k2(k, P, S, aaa) : 0.0;
      v = __dsub_rn(dir, exc);
    }
    eri[e] = v;
  }
}

// fock[p][q] = en[p] delta_pq + sum_r <pr||qr> fo[r]
__global__ __launch_bounds__(UB) void mean_field_kernel(int n,
                                                        const double *en,
                                                        const double *eri,
                                                        const double *fo,
                                                        double *fock) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)UB + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * UB) {
    const int p = (int)(e / n), q = (int)(e % n);
    const double *row = eri + ((int64_t)p * n * n * n + q * (int64_t)n);
    double acc = 0.0;
    for (int r = 0; r < n; ++r)
      acc += row[(int64_t)r * n * n + r] * fo[r];
    fock[e] = (p == q ? en[p] : 0.0) + acc;
  }
}

// per-block partial sums of <ij||ij> fo_i fo_j (first slot of the triples)
__global__ __launch_bounds__(UB) void mp1_kernel(int n, const double *eri,
                                                 const double *fo,
                                                 double *partials) {
  __shared__ double red[UB];
  const int64_t nn = (int64_t)n * n;
  double acc = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)UB + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * UB) {
    const int i = (int)(e / n), j = (int)(e % n);
    acc += eri[((int64_t)i * n + j) * nn + e] * fo[i] * fo[j];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = UB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x] = red[0];
    partials[3 * blockIdx.x + 1] = 0.0;
    partials[3 * blockIdx.x + 2] = 0.0;
  }
}

int blocks_for(int64_t n, int cap) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + UB - 1) / UB, cap));
}

}  // namespace

void ueg_eri(int nk, const int *X_dev, const double *k_dev, double aaa,
             double *eri, hipStream_t s) {
  const int64_t n = 2 * (int64_t)nk;
  hipLaunchKernelGGL(ueg_eri_kernel, dim3(blocks_for(n * n * n * n, 1 << 20)),
                     dim3(UB), 0, s, nk, X_dev, k_dev, aaa, eri);
  FTCC_CHECK_HIP(hipGetLastError());
}

void mean_field_fock(int n, const double *en, const double *eri,
                     const double *fo, double *fock, hipStream_t s) {
  hipLaunchKernelGGL(mean_field_kernel,
                     dim3(blocks_for((int64_t)n * n, 4096)), dim3(UB), 0, s,
                     n, en, eri, fo, fock);
  FTCC_CHECK_HIP(hipGetLastError());
}

int mp1_partials(int n, const double *eri, const double *fo, double *partials,
                 hipStream_t s) {
  const int blocks = blocks_for((int64_t)n * n, 1024);
  hipLaunchKernelGGL(mp1_kernel, dim3(blocks), dim3(UB), 0, s, n, eri, fo,
                     partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

}  // namespace ftcc
