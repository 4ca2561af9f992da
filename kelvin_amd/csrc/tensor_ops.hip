// HBM-bound helper kernels of the residual: permutations, Fermi weights,
// GEMVs and the antisymmetrising assembly of the doubles residual.
// All are grid-stride loops over the element range with 64-bit indexing.
#include "ftcc_internal.h"

#include <algorithm>
#include <math.h>

namespace ftcc {
namespace {

constexpr int TPB = 256;

inline int blocks_for(int64_t n, int cap = 1 << 20) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + TPB - 1) / TPB, cap));
}

#define GRID_STRIDE(i, n)                                                   \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

__global__ void permute4_kernel(const double *in, double *out, int d0, int d1,
                                int d2, int d3, int64_t o0, int64_t o1,
                                int64_t o2, int64_t o3, double alpha,
                                double beta, int nb, int64_t in_b,
                                int64_t out_b) {
  const int64_t per = (int64_t)d0 * d1 * d2 * d3;
  GRID_STRIDE(idx, per * nb) {
    const int64_t b = idx / per;
    int64_t r = idx % per;
    const int i3 = (int)(r % d3);
    r /= d3;
    const int i2 = (int)(r % d2);
    r /= d2;
    const int i1 = (int)(r % d1);
    const int i0 = (int)(r / d1);
    double *o = out + b * out_b + i0 * o0 + i1 * o1 + i2 * o2 + i3 * o3;
    const double v = alpha * in[b * in_b + (idx % per)];
    *o = (beta == 0.0) ? v : v + beta * *o;
  }
}

__global__ void gather_weighted4_kernel(const double *eri, int n,
                                        const int *p0, const int *p1,
                                        const int *p2, const int *p3,
                                        const double *w0, const double *w1,
                                        const double *w2, const double *w3,
                                        int d0, int d1, int d2, int d3,
                                        double *out) {
  const int64_t total = (int64_t)d0 * d1 * d2 * d3;
  GRID_STRIDE(idx, total) {
    int64_t r = idx;
    const int i3 = (int)(r % d3);
    r /= d3;
    const int i2 = (int)(r % d2);
    r /= d2;
    const int i1 = (int)(r % d1);
    const int i0 = (int)(r / d1);
    const int64_t src =
        (((int64_t)p0[i0] * n + p1[i1]) * n + p2[i2]) * n + p3[i3];
    out[idx] = eri[src] * w0[i0] * w1[i1] * w2[i2] * w3[i3];
  }
}

// one wave64 per row; deterministic shuffle reduction
__global__ void gemv_rows_kernel(const double *M, int64_t ldm, int64_t sMb,
                                 const double *x, int64_t sxb, double *y,
                                 int64_t syb, int rows, int K, int nb,
                                 double alpha, double beta) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t w = wave; w < (int64_t)rows * nb; w += nwaves) {
    const int64_t b = w / rows, r = w % rows;
    const double *row = M + b * sMb + r * ldm;
    const double *xb = x + b * sxb;
    double s = 0.0;
    for (int k = lane; k < K; k += 64) s += row[k] * xb[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
      double *o = y + b * syb + r;
      *o = (beta == 0.0) ? alpha * s : alpha * s + beta * *o;
    }
  }
}

// partial[chunk][b][c] = sum_{r in chunk} x[b][r] * M[b][r*ldm + c]
__global__ void gemv_cols_partial_kernel(const double *M, int64_t ldm,
                                         int64_t sMb, const double *x,
                                         int64_t sxb, int rows, int cols,
                                         int nb, int rchunk, double *partial) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int chunk = blockIdx.y;
  const int b = blockIdx.z;
  if (c >= cols) return;
  const int r0 = chunk * rchunk, r1 = min(rows, r0 + rchunk);
  const double *Mb = M + b * sMb;
  const double *xb = x + b * sxb;
  double s = 0.0;
  for (int r = r0; r < r1; ++r) s += xb[r] * Mb[(int64_t)r * ldm + c];
  partial[((int64_t)chunk * nb + b) * cols + c] = s;
}

__global__ void gemv_cols_reduce_kernel(const double *partial, int nchunk,
                                        int cols, int nb, double *y,
                                        int64_t syb, double alpha,
                                        double beta) {
  GRID_STRIDE(idx, (int64_t)cols * nb) {
    const int64_t b = idx / cols, c = idx % cols;
    double s = 0.0;
    for (int k = 0; k < nchunk; ++k) s += partial[((int64_t)k * nb + b) * cols + c];
    double *o = y + b * syb + c;
    *o = (beta == 0.0) ? alpha * s : alpha * s + beta * *o;
  }
}

__global__ void axpby_kernel(double *y, const double *x, int64_t n, double a,
                             double b) {
  GRID_STRIDE(i, n) y[i] = (b == 0.0) ? a * x[i] : a * x[i] + b * y[i];
}

__global__ void fill_kernel(double *y, int64_t n, double v) {
  GRID_STRIDE(i, n) y[i] = v;
}

__global__ void bcast_kernel(double *y, const double *x, int64_t n, int nb,
                             double a, double b) {
  GRID_STRIDE(i, n * nb) {
    const double v = a * x[i % n];
    y[i] = (b == 0.0) ? v : v + b * y[i];
  }
}

// fo = 1/(1+e^{x}), fv = 1/(1+e^{-x}), x = beta(e-mu), each evaluated with
// the exponential of a non-positive argument; GP0 = -(max(-x,0) +
// log1p(e^{-|x|}))/beta.
__global__ void fermi_kernel(const double *en, int n, double beta, double mu,
                             double *fo, double *fv, double *sfo, double *sfv,
                             double *gp0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double x = beta * (en[p] - mu);
  const double ex = exp(-fabs(x));
  double o, v;
  if (x > 0) {
    o = ex / (1.0 + ex);
    v = 1.0 / (1.0 + ex);
  } else {
    o = 1.0 / (1.0 + ex);
    v = ex / (1.0 + ex);
  }
  fo[p] = o;
  fv[p] = v;
  sfo[p] = sqrt(o);
  sfv[p] = sqrt(v);
  gp0[p] = -(fmax(-x, 0.0) + log1p(ex)) / beta;
}

__global__ void gather_vec_kernel(const double *src, const int *idx, int n,
                                  double *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void fock_block_kernel(const double *fock, const double *en, int n,
                                  const int *p, const int *q,
                                  const double *wi, const double *wj, int di,
                                  int dj, double *out) {
  GRID_STRIDE(idx, (int64_t)di * dj) {
    const int i = (int)(idx / dj), j = (int)(idx % dj);
    const int pi = p[i], qj = q[j];
    double f = fock[(int64_t)pi * n + qj];
    if (pi == qj) f -= en[pi];
    out[idx] = f * wi[i] * wj[j];
  }
}

__global__ void denominators_kernel(const double *ev, const double *eo, int nv,
                                    int no, double *D1, double *D2) {
  const int64_t total = (int64_t)nv * nv * no * no;
  GRID_STRIDE(idx, total) {
    int64_t r = idx;
    const int j = (int)(r % no);
    r /= no;
    const int i = (int)(r % no);
    r /= no;
    const int b = (int)(r % nv);
    const int a = (int)(r / nv);
    D2[idx] = ev[a] + ev[b] - eo[i] - eo[j];
    if (b == 0 && j == 0) D1[(int64_t)a * no + i] = ev[a] - eo[i];
  }
}

// One thread per t2 element (a,b,i,j) of each slice:
//   tau   [a,b,i,j] = t2 + t1[a,i]t1[b,j] - t1[b,i]t1[a,j]
//   t2r   [a,i,b,j] = t2[a,b,i,j]                    (ring layout)
//   aring [a,i,b,j] = -t2[a,b,i,j]/2 + t1[a,j]t1[b,i]
__global__ void prep_kernel(const double *t1s, const double *t2s, int nv,
                            int no, int nb, double *tau, double *t2r,
                            double *aring) {
  const int64_t per = (int64_t)nv * nv * no * no;
  const int64_t per1 = (int64_t)nv * no;
  GRID_STRIDE(idx, per * nb) {
    const int64_t y = idx / per;
    const int64_t e = idx % per;
    int64_t r = e;
    const int j = (int)(r % no);
    r /= no;
    const int i = (int)(r % no);
    r /= no;
    const int b = (int)(r % nv);
    const int a = (int)(r / nv);
    const double *t1 = t1s + y * per1;
    const double v = t2s[y * per + e];
    const double x = t1[a * no + i] * t1[b * no + j] -
                     t1[b * no + i] * t1[a * no + j];
    tau[y * per + e] = v + x;
    const int64_t ring = (((int64_t)a * no + i) * nv + b) * no + j;
    t2r[y * per + ring] = v;
    aring[y * per + ring] = -0.5 * v + t1[a * no + j] * t1[b * no + i];
  }
}

__global__ void antisym_add_kernel(double *W, const double *Q, int d0, int d1,
                                   int d2, int d3, int swap_last, double sgn,
                                   int nb) {
  const int64_t per = (int64_t)d0 * d1 * d2 * d3;
  GRID_STRIDE(idx, per * nb) {
    const int64_t y = idx / per;
    int64_t r = idx % per;
    const int x3 = (int)(r % d3);
    r /= d3;
    const int x2 = (int)(r % d2);
    r /= d2;
    const int x1 = (int)(r % d1);
    const int x0 = (int)(r / d1);
    int64_t other;
    if (swap_last)
      other = (((int64_t)x0 * d1 + x1) * d2 + x3) * d3 + x2;
    else
      other = (((int64_t)x1 * d1 + x0) * d2 + x2) * d3 + x3;
    W[idx] += sgn * (Q[idx] - Q[y * per + other]);
  }
}

// Wr[e,m,b,j] = ovvo_ring[e,m,b,j] - Q1[b,m,e,j] + Q2[m,b,j,e]
__global__ void ring_assemble_kernel(double *Wr, const double *ovvo_ring,
                                     const double *Q1, const double *Q2,
                                     int nv, int no, int nb) {
  const int64_t per = (int64_t)nv * no * nv * no;
  GRID_STRIDE(idx, per * nb) {
    const int64_t y = idx / per;
    const int64_t w = idx % per;
    int64_t r = w;
    const int j = (int)(r % no);
    r /= no;
    const int b = (int)(r % nv);
    r /= nv;
    const int m = (int)(r % no);
    const int e = (int)(r / no);
    const int64_t q1 = (((int64_t)b * no + m) * nv + e) * no + j;
    const int64_t q2 = (((int64_t)m * nv + b) * no + j) * nv + e;
    Wr[idx] = ovvo_ring[w] - Q1[y * per + q1] + Q2[y * per + q2];
  }
}

// t2bar[a,b,i,j] = -( acc1[a,b,i,j] + acc2[a,b,i,j]   (ladder products)
//                     [+ add[a,b,i,j]]         (one slice, all grid points)
//                     + S1[a,b,i,j] - S1[b,a,i,j]
//                     + S2[a,b,i,j] - S2[a,b,j,i]
//                     + Z[a,i,b,j] - Z[b,i,a,j] - Z[a,j,b,i] + Z[b,j,a,i] )
// One workgroup per (a <= b) pair and 32x32 (i,j) tile writes both (a,b)
// and (b,a): their S1 and Z terms are the same four and two reads with
// opposite signs.  The operands read transposed (S2[a,b,j,i], S2[b,a,j,i]
// and the difference Z[b,j,a,i] - Z[a,j,b,i]) are loaded along i into LDS
// so every global access is coalesced (three 32x33 tiles).  With spin
// labels (lv, lo), elements whose spins do not conserve
// (lv[a] + lv[b] != lo[i] + lo[j]) are written as zero and their operands
// are not read: every operand element an allowed output reads is itself
// allowed.
__global__ __launch_bounds__(256) void r2_final_pair_kernel(
    double *t2bar, const double *acc1, const double *acc2, const double *S1,
    const double *S2, const double *Z, const double *add, int nv, int no,
    const int *lv, const int *lo) {
  __shared__ double tA[32][33], tB[32][33], tZ[32][33];
  const int tiles = (no + 31) / 32;
  const int64_t npair = (int64_t)nv * (nv + 1) / 2;
  int64_t bid = blockIdx.x;
  const int tj = (int)(bid % tiles);
  bid /= tiles;
  const int ti = (int)(bid % tiles);
  bid /= tiles;
  const int64_t q = bid % npair;
  const int64_t y = bid / npair;
  // pair q -> (a, b), a <= b, q = b (b + 1) / 2 + a
  int b = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while ((int64_t)b * (b + 1) / 2 > q) --b;
  while ((int64_t)(b + 1) * (b + 2) / 2 <= q) ++b;
  const int a = (int)(q - (int64_t)b * (b + 1) / 2);
  const int64_t per = (int64_t)nv * nv * no * no;
  const double *s1 = S1 + y * per, *s2 = S2 + y * per, *z = Z + y * per;
  const double *c1 = acc1 + y * per, *c2 = acc2 + y * per;
  double *out = t2bar + y * per;
  const int i0 = ti * 32, j0 = tj * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int sab = lv ? lv[a] + lv[b] : 0;
  auto nat = [&](int p, int r, int k, int l) {
    return (((int64_t)p * nv + r) * no + k) * no + l;
  };
  auto ring = [&](int p, int k, int r, int l) {
    return (((int64_t)p * no + k) * nv + r) * no + l;
  };
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + tx, j = j0 + r;
    const bool ok = i < no && j < no && (!lv || lo[i] + lo[j] == sab);
    tA[r][tx] = ok ? s2[nat(a, b, j, i)] : 0.0;
    tB[r][tx] = ok ? s2[nat(b, a, j, i)] : 0.0;
    tZ[r][tx] = ok ? z[ring(b, j, a, i)] - z[ring(a, j, b, i)] : 0.0;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = i0 + r, j = j0 + tx;
    if (i >= no || j >= no) continue;
    const int64_t nab = nat(a, b, i, j), nba = nat(b, a, i, j);
    if (lv && lo[i] + lo[j] != sab) {
      out[nab] = 0.0;
      out[nba] = 0.0;
      continue;
    }
    const double s1ab = s1[nab], s1ba = s1[nba];
    const double zz = z[ring(a, i, b, j)] - z[ring(b, i, a, j)] + tZ[tx][r];
    double vab = c1[nab] + c2[nab] + s1ab - s1ba + s2[nab] - tA[tx][r] + zz;
    double vba = c1[nba] + c2[nba] + s1ba - s1ab + s2[nba] - tB[tx][r] - zz;
    if (add) {
      vab += add[nab];
      vba += add[nba];
    }
    out[nab] = -vab;
    if (a != b) out[nba] = -vba;
  }
}

// The same assembly in the packed spin layout, where every input holds
// one pair of each antisymmetric couple (x < y for equal spins, the alpha
// index first for mixed spins; the other follows by antisymmetry):
//   acc1, acc2  the stored (a,b) and the stored (i,j)
//   S1          every (a,b), the stored (i,j)   (S1 is antisymmetric in ij)
//   S2          the stored (a,b), every (i,j)   (S2 is antisymmetric in ab)
// and add (<ab||ij>) is antisymmetric, so that R2[b,a,i,j] = -R2[a,b,i,j]
// is written without further reads.  zh > 0 (spin-flip symmetric problem,
// alpha runs [0, zh) and [0, zo), beta runs after them): Z holds the rows of
// an alpha first index only, and a row of a beta one is read from its image
// Z[p - zh][fl(k)][fl(q)][fl(l)].  One workgroup per (a <= b) pair and
// R2P_ROWS rows i, all j: direct reads and writes run along whole rows of
// j; the four values read transposed (the stored element of a transposed
// (i,j), S2[.,.,j,i], the S1 and Z differences at (j,i)) are first staged
// in LDS as [j][i] strips (row stride R2P_LD doubles: two-way bank
// conflicts at most on the reads along j).  Four rows per workgroup
// measured fastest at n = 114 (2, 4, 8, 16 rows: 38.6, 28.5, 33.7, 71.0 ms
// per iteration).  The row tiles of one (a,b) pair share the cache lines of
// their staged reads (R2P_ROWS doubles of each line per tile), so the block
// index is regrouped to run consecutive tiles on one XCD (workgroups are
// dispatched to the eight XCDs round robin; each XCD has its own L2).
constexpr int R2P_ROWS = 4, R2P_LD = 5;

// Logical block of a hardware block id: the ids x, x + 8, x + 16, ... (XCD
// x) take a contiguous range of logical blocks; a bijection on [0, G).
__device__ inline int64_t xcd_grouped(int64_t bid, int64_t G) {
  const int64_t q = G / 8, r = G % 8, x = bid % 8, loc = bid / 8;
  return x < r ? x * (q + 1) + loc : r * (q + 1) + (x - r) * q + loc;
}
__global__ __launch_bounds__(256) void r2_packed_kernel(
    double *t2bar, const double *acc1, const double *acc2, const double *S1,
    const double *S2, const double *Z, const double *add, int nv, int no,
    const int *lv, const int *lo, int zh, int zo, int zero_forbidden,
    const int *jl, int jc0, int jc1) {
  extern __shared__ double sm[];   // 4 strips of no * R2P_LD
  double *tS2 = sm, *tS1 = sm + no * R2P_LD, *tZ = sm + 2 * no * R2P_LD,
         *tP = sm + 3 * no * R2P_LD;
  const int tiles = (no + R2P_ROWS - 1) / R2P_ROWS;
  const int64_t npair = (int64_t)nv * (nv + 1) / 2;
  int64_t bid = xcd_grouped(blockIdx.x, gridDim.x);
  const int ti = (int)(bid % tiles);
  bid /= tiles;
  const int64_t q = bid % npair;
  const int64_t y = bid / npair;
  int b = (int)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while ((int64_t)b * (b + 1) / 2 > q) --b;
  while ((int64_t)(b + 1) * (b + 2) / 2 <= q) ++b;
  const int a = (int)(q - (int64_t)b * (b + 1) / 2);
  const int64_t per = (int64_t)nv * nv * no * no;
  const double *s1 = S1 + y * per, *s2 = S2 + y * per, *z = Z + y * per;
  const double *c1 = acc1 + y * per, *c2 = acc2 + y * per;
  double *out = t2bar + y * per;
  const int i0 = ti * R2P_ROWS;
  const int sab = lv[a] + lv[b];
  // spin-flip symmetric step (zh > 0, forbidden doubles not written): the
  // pairs of two beta virtuals are the images of the alpha-alpha pairs,
  // whose workgroups write them; half of the mixed pairs likewise
  const bool fimage = zh > 0 && !zero_forbidden;
  if (fimage && a >= zh) return;   // a <= b: both beta
  const bool wimage = fimage && b < zh;   // both alpha
  // a mixed pair (a alpha, b beta) and (b - zh, a + zh) are images of each
  // other: the one with a < b - zh writes both.  A pair with a == b - zh is
  // its own image: R2[a,b,i,j] = R2[a,b,fl(j),fl(i)] = -R2[a,b,j,i]
  // = -R2[a,b,fl(i),fl(j)], so the four elements are written from one of
  // them, the one with i alpha and the smaller i (computed separately, they
  // would differ by rounding in the low-precision modes)
  bool mimage = false, simage = false;
  if (fimage && b >= zh) {
    if (a > b - zh) return;
    mimage = a < b - zh;
    simage = a == b - zh;
  }
  auto flo = [&](int k) { return k < zo ? k + zo : k - zo; };
  auto nat = [&](int p, int r, int k, int l) {
    return (((int64_t)p * nv + r) * no + k) * no + l;
  };
  auto ring = [&](int p, int k, int r, int l) {
    return (((int64_t)p * no + k) * nv + r) * no + l;
  };
  auto zr = [&](int p, int k, int r, int l) {
    if (zh > 0 && p >= zh) {
      p -= zh;
      k = k < zo ? k + zo : k - zo;
      r = r < zh ? r + zh : r - zh;
      l = l < zo ? l + zo : l - zo;
    }
    return z[ring(p, k, r, l)];
  };
  // the stored (a,b) pair and its sign; a == b is zero throughout
  int pa = a, pb = b;
  double sg = 1.0;
  const bool zab = a == b;
  if (lv[a] != lv[b] && lv[a] == 1) {
    pa = b;
    pb = a;
    sg = -1.0;
  }
  // (i,j): 1 = stored as (i,j), -1 = stored as (j,i), 0 = zero
  auto ij_sign = [&](int i, int j) {
    if (lo[i] == lo[j]) return i < j ? 1 : (i > j ? -1 : 0);
    return lo[i] == 0 ? 1 : -1;
  };
  // stage the values at (j, i) for the block's rows i, contiguous along i;
  // only the spin-allowed (i, j) are staged, the only ones read below: row
  // i allows the occupied j of label sab - lo[i] (jl lists them by label)
  const int cmax = jc0 > jc1 ? jc0 : jc1;
  auto allowed_j = [&](int il, int k) {   // k-th allowed j of row il, or -1
    const int i = i0 + il;
    if (i >= no) return -1;
    const int sj = sab - lo[i];
    if (sj < 0 || sj > 1 || k >= (sj ? jc1 : jc0)) return -1;
    return jl[sj ? jc0 + k : k];
  };
  if (!zab)
    for (int t = threadIdx.x; t < cmax * R2P_ROWS; t += blockDim.x) {
      const int il = t % R2P_ROWS, j = allowed_j(il, t / R2P_ROWS);
      if (j < 0) continue;
      const int i = i0 + il;
      const bool tr = ij_sign(i, j) < 0;
      const int64_t e = nat(pa, pb, j, i);
      tS2[j * R2P_LD + il] = s2[e];
      tS1[j * R2P_LD + il] =
          tr ? s1[nat(a, b, j, i)] - s1[nat(b, a, j, i)] : 0.0;
      tZ[j * R2P_LD + il] = zr(b, j, a, i) - zr(a, j, b, i);
      tP[j * R2P_LD + il] = tr ? c1[e] + c2[e] : 0.0;
    }
  __syncthreads();
  // one spin-allowed element (a != b): R2[a,b,i,j] and R2[b,a,i,j]
  auto element = [&](int il, int i, int j) {
    const int64_t nab = nat(a, b, i, j), nba = nat(b, a, i, j);
    const int sij = ij_sign(i, j);
    const int64_t e = nat(pa, pb, i, j);
    double acc = 0.0, s1d = 0.0;
    if (sij > 0) {
      acc = c1[e] + c2[e];
      s1d = s1[nab] - s1[nba];
    } else if (sij < 0) {
      acc = -tP[j * R2P_LD + il];
      s1d = -tS1[j * R2P_LD + il];
    }
    const double s2d = s2[e] - tS2[j * R2P_LD + il];
    const double zz = zr(a, i, b, j) - zr(b, i, a, j) + tZ[j * R2P_LD + il];
    double v = sg * (acc + s2d) + s1d + zz;
    if (add) v += add[nab];
    out[nab] = -v;
    out[nba] = v;
    if (wimage) {
      out[nat(a + zh, b + zh, flo(i), flo(j))] = -v;
      out[nat(b + zh, a + zh, flo(i), flo(j))] = v;
    }
    if (mimage) {
      out[nat(a + zh, b - zh, flo(i), flo(j))] = -v;
      out[nat(b - zh, a + zh, flo(i), flo(j))] = v;
    }
    if (simage) {
      out[nat(a, b, j, i)] = v;
      out[nat(b, a, j, i)] = -v;
      if (i != flo(j)) {
        out[nat(a, b, flo(j), flo(i))] = -v;
        out[nat(b, a, flo(j), flo(i))] = v;
        out[nat(a, b, flo(i), flo(j))] = v;
        out[nat(b, a, flo(i), flo(j))] = -v;
      }
    }
  };
  if (zero_forbidden) {
    const int lanes = blockDim.x / 2;   // two rows at a time
    for (int il = threadIdx.x / lanes; il < R2P_ROWS; il += 2) {
      const int i = i0 + il;
      if (i >= no) break;
      for (int j = threadIdx.x % lanes; j < no; j += lanes) {
        if (zab || lo[i] + lo[j] != sab) {
          out[nat(a, b, i, j)] = 0.0;
          out[nat(b, a, i, j)] = 0.0;
          continue;
        }
        element(il, i, j);
      }
    }
  } else {
    // the allowed (i, j) only, rows after one another along the threads (a
    // == b is allowed for some (i,j) and always zero: written)
    for (int t = threadIdx.x; t < R2P_ROWS * cmax; t += blockDim.x) {
      const int il = t / cmax, j = allowed_j(il, t % cmax);
      if (j < 0) continue;
      const int i = i0 + il;
      if (zab) {
        out[nat(a, b, i, j)] = 0.0;
        if (wimage) out[nat(a + zh, b + zh, flo(i), flo(j))] = 0.0;
        continue;
      }
      // a self-image pair: written by its element with i alpha and the
      // smaller i (i beta: from (j, i); fl(j) < i: from (fl(j), fl(i)))
      if (simage && (lo[i] != 0 || flo(j) < i)) continue;
      element(il, i, j);
    }
  }
}

// Ring tensor X[y][p][k][q][l] (virtual p, q; occupied k, l) of a
// spin-flip symmetric problem whose rows with p alpha are computed: the
// allowed elements with p beta are their flipped images
// X[y][p - hv][fl(k)][fl(q)][fl(l)] (fl: alpha <-> beta partner). Allowed
// means sk + sq == sp + sl for a ring intermediate W[e,m,b,j] (vv_pair = 0)
// and sp + sq == sk + sl for an amplitude-like Z[a,i,b,j] (vv_pair = 1).
// With p beta each (k, q) spin pair allows at most one l run, and three of
// the four pairs allow one:
//   vv_pair 0: (sk, sq) = (1,0), (0,1) -> sl = 0; (1,1) -> sl = 1
//   vv_pair 1: (sk, sq) = (1,0) -> sl = 0; (0,0), (1,1) -> sl = 1
// so the kernel enumerates the allowed elements only: per slice (blockIdx.y)
// hv x 3 x ho x hv x ho of them, l innermost (source and destination runs of
// ho contiguous elements).
__global__ void flip_copy_ring_kernel(double *X, int nv, int no, int hv,
                                      int ho, int64_t sb, int vv_pair) {
  const int per = 3 * hv * ho * hv * ho;
  double *x = X + blockIdx.y * sb;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < per;
       idx += gridDim.x * blockDim.x) {
    int r = idx;
    const int lh = r % ho;
    r /= ho;
    const int qh = r % hv;
    r /= hv;
    const int kh = r % ho;
    r /= ho;
    const int c = r % 3;
    const int ph = r / 3;
    const int sk = c == 1 ? 0 : 1;
    const int sq = c == 0 ? 0 : (c == 1 ? vv_pair ^ 1 : 1);
    const int sl = c == 0 ? 0 : (c == 1 ? vv_pair : 1);
    const int k = kh + sk * ho, q = qh + sq * hv, l = lh + sl * ho;
    const int fk = kh + (sk ^ 1) * ho, fq = qh + (sq ^ 1) * hv,
              fl = lh + (sl ^ 1) * ho;
    x[(((int64_t)(hv + ph) * no + k) * nv + q) * no + l] =
        x[(((int64_t)ph * no + fk) * nv + fq) * no + fl];
  }
}

// Spin-diagonal matrices X[y][p][q] (P x Q, slice stride sb) of a
// spin-flip symmetric problem whose alpha-alpha block is computed: the
// beta-beta block is its image, X[y][hp + p][hq + q] = X[y][p][q].
__global__ void flip_copy_mat_kernel(double *X, int Q, int hp, int hq,
                                     int64_t sb) {
  const int per = hp * hq;
  double *x = X + blockIdx.y * sb;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < per;
       idx += gridDim.x * blockDim.x) {
    const int p = idx / hq, q = idx % hq;
    x[(int64_t)(hp + p) * Q + hq + q] = x[(int64_t)p * Q + q];
  }
}

__global__ void transpose2_kernel(const double *in, double *out, int r, int c,
                                  int nb) {
  const int64_t per = (int64_t)r * c;
  GRID_STRIDE(idx, per * nb) {
    const int64_t b = idx / per;
    const int64_t e = idx % per;
    const int i = (int)(e / c), j = (int)(e % c);
    out[b * per + (int64_t)j * r + i] = in[idx];
  }
}

}  // namespace

#define LAUNCH(kernel, n, ...)                                            \
  do {                                                                    \
    hipLaunchKernelGGL(kernel, dim3(blocks_for(n)), dim3(TPB), 0, s,      \
                       __VA_ARGS__);                                      \
    FTCC_CHECK_HIP(hipGetLastError());                                    \
  } while (0)

void permute4(const double *in, double *out, int d0, int d1, int d2, int d3,
              int64_t o0, int64_t o1, int64_t o2, int64_t o3, double alpha,
              double beta, int nb, int64_t in_b, int64_t out_b,
              hipStream_t s) {
  const int64_t n = (int64_t)d0 * d1 * d2 * d3 * nb;
  if (n == 0) return;
  LAUNCH(permute4_kernel, n, in, out, d0, d1, d2, d3, o0, o1, o2, o3, alpha,
         beta, nb, in_b, out_b);
}

void gather_weighted4(const double *eri, int n, const int *p0, const int *p1,
                      const int *p2, const int *p3, const double *w0,
                      const double *w1, const double *w2, const double *w3,
                      int d0, int d1, int d2, int d3, double *out,
                      hipStream_t s) {
  const int64_t total = (int64_t)d0 * d1 * d2 * d3;
  if (total == 0) return;
  LAUNCH(gather_weighted4_kernel, total, eri, n, p0, p1, p2, p3, w0, w1, w2,
         w3, d0, d1, d2, d3, out);
}

void gemv_rows(const double *M, int64_t ldm, int64_t sMb, const double *x,
               int64_t sxb, double *y, int64_t syb, int rows, int K, int nb,
               double alpha, double beta, hipStream_t s) {
  const int64_t waves = (int64_t)rows * nb;
  if (waves == 0) return;
  const int blocks = (int)std::min<int64_t>((waves + 3) / 4, 1 << 16);
  hipLaunchKernelGGL(gemv_rows_kernel, dim3(blocks), dim3(TPB), 0, s, M, ldm,
                     sMb, x, sxb, y, syb, rows, K, nb, alpha, beta);
  FTCC_CHECK_HIP(hipGetLastError());
}

void gemv_cols(const double *M, int64_t ldm, int64_t sMb, const double *x,
               int64_t sxb, double *y, int64_t syb, int rows, int cols,
               int nb, double alpha, double beta, double *partial,
               size_t partial_elems, hipStream_t s) {
  if (rows == 0 || cols == 0 || nb == 0) return;
  const int cblocks = (cols + TPB - 1) / TPB;
  // aim for ~2048 workgroups in total, at least 64 rows per chunk
  int nchunk = (int)std::max<int64_t>(
      1, std::min<int64_t>(rows / 64, 2048 / std::max<int64_t>(1, (int64_t)cblocks * nb)));
  nchunk = std::max(1, std::min(nchunk, 1024));
  while ((size_t)nchunk * nb * cols > partial_elems && nchunk > 1) nchunk /= 2;
  require((size_t)nchunk * nb * cols <= partial_elems,
          "gemv scratch too small", FTCC_ENOMEM);
  const int rchunk = (rows + nchunk - 1) / nchunk;
  nchunk = (rows + rchunk - 1) / rchunk;
  hipLaunchKernelGGL(gemv_cols_partial_kernel, dim3(cblocks, nchunk, nb),
                     dim3(TPB), 0, s, M, ldm, sMb, x, sxb, rows, cols, nb,
                     rchunk, partial);
  FTCC_CHECK_HIP(hipGetLastError());
  LAUNCH(gemv_cols_reduce_kernel, (int64_t)cols * nb, partial, nchunk, cols,
         nb, y, syb, alpha, beta);
}

void axpby(double *y, const double *x, int64_t n, double a, double b,
           hipStream_t s) {
  if (n == 0) return;
  LAUNCH(axpby_kernel, n, y, x, n, a, b);
}

void fill(double *y, int64_t n, double v, hipStream_t s) {
  if (n == 0) return;
  LAUNCH(fill_kernel, n, y, n, v);
}

void bcast(double *y, const double *x, int64_t n, int nb, double a, double b,
           hipStream_t s) {
  if (n * nb == 0) return;
  LAUNCH(bcast_kernel, n * nb, y, x, n, nb, a, b);
}

void fermi_weights(const double *en, int n, double beta, double mu, double *fo,
                   double *fv, double *sfo, double *sfv, double *gp0,
                   hipStream_t s) {
  LAUNCH(fermi_kernel, n, en, n, beta, mu, fo, fv, sfo, sfv, gp0);
}

void gather_vec(const double *src, const int *idx, int n, double *dst,
                hipStream_t s) {
  if (n == 0) return;
  LAUNCH(gather_vec_kernel, n, src, idx, n, dst);
}

void fock_block(const double *fock, const double *en, int n, const int *p,
                const int *q, const double *wi, const double *wj, int di,
                int dj, double *out, hipStream_t s) {
  if ((int64_t)di * dj == 0) return;
  LAUNCH(fock_block_kernel, (int64_t)di * dj, fock, en, n, p, q, wi, wj, di,
         dj, out);
}

void denominators(const double *ev, const double *eo, int nv, int no,
                  double *D1, double *D2, hipStream_t s) {
  LAUNCH(denominators_kernel, (int64_t)nv * nv * no * no, ev, eo, nv, no, D1,
         D2);
}

void prep_amplitudes(const double *t1, const double *t2, int nv, int no,
                     int nb, const PrepOut &o, hipStream_t s) {
  LAUNCH(prep_kernel, (int64_t)nv * nv * no * no * nb, t1, t2, nv, no, nb,
         o.tau, o.t2r, o.aring);
}

void antisym_add(double *W, const double *Q, int d0, int d1, int d2, int d3,
                 bool swap_last, double sgn, int nb, hipStream_t s) {
  LAUNCH(antisym_add_kernel, (int64_t)d0 * d1 * d2 * d3 * nb, W, Q, d0, d1,
         d2, d3, swap_last ? 1 : 0, sgn, nb);
}

void ring_assemble(double *Wr, const double *ovvo_ring, const double *Q1,
                   const double *Q2, int nv, int no, int nb, hipStream_t s) {
  LAUNCH(ring_assemble_kernel, (int64_t)nv * no * nv * no * nb, Wr,
         ovvo_ring, Q1, Q2, nv, no, nb);
}

void r2_final(double *t2bar, const double *acc1, const double *acc2,
              const double *S1, const double *S2, const double *Z, int nv,
              int no, int nb, hipStream_t s, const int *lab_v,
              const int *lab_o, const double *add, bool packed, int zh,
              int zo, bool zero_forbidden, const int *jl, int jc0,
              int jc1) {
  require(!packed || (lab_v && lab_o), "packed pairs need spin labels");
  require(zh == 0 || (packed && 2 * zh == nv && 2 * zo == no),
          "r2 assembly: flipped Z needs the packed layout and equal halves");
  require(zero_forbidden || packed,
          "r2 assembly: skipping forbidden doubles needs the packed layout");
  require(!packed || (jl && jc0 >= 0 && jc1 >= 0 && jc0 + jc1 == no),
          "r2 assembly: packed pairs need the occupied lists by label");
  const int tiles = (no + 31) / 32;
  const int64_t npair = (int64_t)nv * (nv + 1) / 2;
  const int64_t blocks = (int64_t)nb * npair * tiles * tiles;
  require(blocks < (1LL << 31), "r2 assembly grid too large");
  if (packed) {
    const int64_t pblocks =
        (int64_t)nb * npair * ((no + R2P_ROWS - 1) / R2P_ROWS);
    require(pblocks < (1LL << 31), "r2 assembly grid too large");
    const size_t lds = sizeof(double) * 4 * (size_t)no * R2P_LD;
    require(lds <= 160 * 1024, "r2 assembly: too many occupied orbitals");
    static std::once_flag attr;
    set_max_lds_once(attr, (const void *)r2_packed_kernel, 160 * 1024);
    hipLaunchKernelGGL(r2_packed_kernel, dim3((unsigned)pblocks), dim3(256),
                       lds, s, t2bar, acc1, acc2, S1, S2, Z, add, nv, no,
                       lab_v, lab_o, zh, zo, zero_forbidden ? 1 : 0, jl, jc0,
                       jc1);
  } else
    hipLaunchKernelGGL(r2_final_pair_kernel, dim3((unsigned)blocks),
                       dim3(256), 0, s, t2bar, acc1, acc2, S1, S2, Z, add, nv,
                       no, lab_v, lab_o);
  FTCC_CHECK_HIP(hipGetLastError());
}

void flip_copy_mat(double *X, int P, int Q, int hp, int hq, int nb,
                   int64_t sb, hipStream_t s) {
  require(2 * hp == P && 2 * hq == Q, "flip copy needs equal spin halves");
  require(nb >= 1 && nb <= 65535 && (int64_t)P * Q <= sb,
          "flip copy: bad slice layout");
  const int per = hp * hq;
  const int blocks = std::max(1, std::min((per + TPB - 1) / TPB, 1 << 16));
  hipLaunchKernelGGL(flip_copy_mat_kernel, dim3(blocks, nb), dim3(TPB), 0, s,
                     X, Q, hp, hq, sb);
  FTCC_CHECK_HIP(hipGetLastError());
}

void flip_copy_ring(double *X, int nv, int no, int hv, int ho, int nb,
                    int64_t sb, bool vv_pair, hipStream_t s) {
  require(2 * hv == nv && 2 * ho == no, "flip copy needs equal spin halves");
  require(nb >= 1 && nb <= 65535, "flip copy: slice count out of range");
  const int64_t per = 3 * (int64_t)hv * ho * hv * ho;
  require(per < (int64_t)1 << 31, "flip copy: slice too large");
  const int blocks = (int)std::min<int64_t>((per + TPB - 1) / TPB, 1 << 16);
  hipLaunchKernelGGL(flip_copy_ring_kernel, dim3(blocks, nb), dim3(TPB), 0, s,
                     X, nv, no, hv, ho, sb, vv_pair ? 1 : 0);
  FTCC_CHECK_HIP(hipGetLastError());
}

void transpose2(const double *in, double *out, int r, int c, int nb,
                hipStream_t s) {
  LAUNCH(transpose2_kernel, (int64_t)r * c * nb, in, out, r, c, nb);
}

}  // namespace ftcc
