// exp(-tau D) integration, damping, norms and grand-potential partials.
//
// One thread owns one amplitude element across all ng grid points: it stages
// Tbar[0..ng) in its own LDS column, forms
//     T(tau_y) = sum_x G[y,x] * exp(dt_yx * D) * Tbar(tau_x),
//     dt_yx = tau_x - tau_y for x < y and 0 otherwise
// (the convention of kelvin/quadrature.py:292-317), and in the fused mode
// damps it into the stored iterate while accumulating |T - Told|^2,
// |Told_new|^2 and the energy contraction (kelvin/cc_utils.py:145-155,
// kelvin/ft_cc_energy.py:7-32).  Everything is FP64.  Per-block partial sums
// are reduced by a single fixed-order pass, so results are deterministic.
#include "ftcc_internal.h"

#include <algorithm>
#include <math.h>

namespace ftcc {
namespace {

constexpr int UT = 128;          // threads per block of the update kernels
constexpr int MAX_LDS_NG = 48;   // grids up to this size stage Tbar, Told in LDS
constexpr int MAX_BLOCKS = 8192;

__device__ inline void block_reduce3(double a, double b, double c,
                                     double *out) {
  __shared__ double red[3][UT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0, s1 = 0, s2 = 0;
    for (int i = 0; i < UT / 64; ++i) {
      s0 += red[0][i];
      s1 += red[1][i];
      s2 += red[2][i];
    }
    out[3 * blockIdx.x + 0] = s0;
    out[3 * blockIdx.x + 1] = s1;
    out[3 * blockIdx.x + 2] = s2;
  }
}

// Element decode: T2 element (a,b,i,j) or T1 element (a,i) (b=j=-1).
struct Elem {
  int a, b, i, j;
};

template <bool DOUBLES>
__device__ inline Elem decode(int64_t e, int nv, int no) {
  Elem r;
  if (DOUBLES) {
    r.j = (int)(e % no);
    e /= no;
    r.i = (int)(e % no);
    e /= no;
    r.b = (int)(e % nv);
    r.a = (int)(e / nv);
  } else {
    r.i = (int)(e % no);
    r.a = (int)(e / no);
    r.b = r.j = -1;
  }
  return r;
}

// Weight of Tbar(tau_x) in T(tau_y): exp((tau_x - tau_y) d) for x < y.  On
// an equally spaced grid (p.uniform, spacing p.delta) it is r^(y-x) with
// r = exp(-delta d), formed as a running product from x = y-1 downwards, so
// each element costs one exp instead of ng(ng-1)/2.
template <bool DOUBLES, bool STAGE_LDS>
__global__ __launch_bounds__(UT) void update_kernel(UpdateArgs p, int mode) {
  extern __shared__ double lds[];   // G[ng*ng], ti[ng], [ng][UT] Tbar, [ng][UT] Told
  const int ng = p.ng, nv = p.nv, no = p.no;
  const double *Gs = p.G, *tis = p.ti;
  double *col = lds + ng * ng + ng;
  double *ocol = col + ng * UT;
  if (STAGE_LDS) {
    for (int k = threadIdx.x; k < ng * ng; k += UT) lds[k] = p.G[k];
    for (int k = threadIdx.x; k < ng; k += UT) lds[ng * ng + k] = p.ti[k];
    __syncthreads();
    Gs = lds;
    tis = lds + ng * ng;
  }
  const int64_t per = DOUBLES ? (int64_t)nv * nv * no * no : (int64_t)nv * no;
  const int64_t per1 = (int64_t)nv * no;
  const double *src = p.src ? p.src : p.tbar;
  double dsum = 0.0, nsum = 0.0, esum = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)UT + threadIdx.x; e < per;
       e += (int64_t)gridDim.x * UT) {
    const Elem q = decode<DOUBLES>(e, nv, no);
    if (DOUBLES && p.lv && p.lv[q.a] + p.lv[q.b] != p.lo[q.i] + p.lo[q.j])
      continue;   // zero in Tbar and T: contributes nothing
    double d;
    if (p.D)
      d = p.D[e];
    else if (DOUBLES)
      d = p.ev[q.a] + p.ev[q.b] - p.eo[q.i] - p.eo[q.j];
    else
      d = p.ev[q.a] - p.eo[q.i];
    if (STAGE_LDS) {
      // issue every load of this element before any arithmetic
      for (int x = 0; x < ng; ++x)
        col[x * UT + threadIdx.x] =
            src[x * per + e] + (p.corr ? p.corr[x * per + e] : 0.0);
      if (mode == 1)
        for (int x = 0; x < ng; ++x)
          ocol[x * UT + threadIdx.x] = p.told[x * per + e];
    }
    const double r = p.uniform ? exp(-p.delta * d) : 0.0;
    double eacc = 0.0;
    for (int y = 0; y < ng; ++y) {
      double tn = 0.0;
      const double *Gy = Gs + y * ng;
      const double ty = tis[y];
      double w = 1.0;
      for (int x = y - 1; x >= 0; --x) {
        if (p.uniform) w *= r;
        const double gyx = Gy[x];
        if (gyx == 0.0) continue;
        const double tb =
            STAGE_LDS ? col[x * UT + threadIdx.x]
                      : src[x * per + e] + (p.corr ? p.corr[x * per + e] : 0.0);
        const double wx = p.uniform ? w : exp((tis[x] - ty) * d);
        tn += gyx * (wx * tb);
      }
      for (int x = y; x < ng; ++x) {
        const double gyx = Gy[x];
        if (gyx == 0.0) continue;
        const double tb =
            STAGE_LDS ? col[x * UT + threadIdx.x]
                      : src[x * per + e] + (p.corr ? p.corr[x * per + e] : 0.0);
        tn += gyx * tb;
      }
      if (mode == 0) {
        p.tbar[y * per + e] = tn;
      } else {
        const double to = STAGE_LDS ? ocol[y * UT + threadIdx.x] : p.told[y * per + e];
        const double df = tn - to;
        dsum += df * df;
        const double tnew = p.damp * to + (1.0 - p.damp) * tn;
        p.told[y * per + e] = tnew;
        nsum += tnew * tnew;
        if (DOUBLES) {
          const double *t1 = p.t1new + y * per1;
          const double qq = t1[q.a * no + q.i] * t1[q.b * no + q.j];
          eacc += p.g[y] * (0.25 * tnew + 0.5 * qq);
        } else {
          eacc += p.g[y] * tnew;
        }
      }
    }
    if (mode == 1) esum += eacc * p.eint[e];
  }
  if (mode == 1) block_reduce3(dsum, nsum, esum, p.partials);
}

// Recurrence form.  On an equally spaced grid whose weight matrix is zero
// above the diagonal and whose rows repeat the row s above them left of the
// last three columns (G[y][x] == G[y-s][x] for x < y-2; kelvin's Simpson
// rule builds G that way with s = 2, kelvin/quadrature.py:32-41, the
// midpoint and 'L' rules with s = 1), the causal sum obeys
//     T_y = r^s T_{y-s} + a_y r^2 Tbar_{y-2} + b_y r Tbar_{y-1} + c_y Tbar_y
// with r = exp(-delta d) and (a_y, b_y, c_y) = G[y][y-2..y] - G[y-s][y-2..y]
// (rows above the first taken as zero).  Each element then costs a few
// flops per grid point, and the kernel is bound by its HBM streams alone:
// Tbar in (and T out in mode 0), Told in and out in mode 1.  Grid points are
// taken in batches of RCH so that every load of a batch is in flight
// together (batches of 16 or 24 measured slower at n = 114, ng = 24).
// In-place integration is safe: slice y is read before it is written and
// earlier slices are carried in registers.
constexpr int RCH = 8;

// Allowed doubles of the spin layout when the labels are runs (alpha then
// beta in each list): the six (sa, sb, si, sj) boxes with sa + sb = si + sj,
// each [a range] x [b range] x [i range] x [j range], enumerated with j
// innermost.  Returns false past the last element.
// x / d and x % d for 0 <= x < 2^50 and 0 < d < 2^20 through one FP64
// multiply by 1/d and an exact integer correction (no 64-bit integer
// division, which the GPU expands into a long instruction sequence)
__device__ inline int64_t divmod(int64_t x, int d, int &rem) {
  int64_t q = (int64_t)((double)x * (1.0 / (double)d));
  int64_t r = x - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
  rem = (int)r;
  return q;
}
__device__ inline bool run_elem(const UpdateArgs &p, int64_t k, Elem &q) {
  const int nbox = p.flip_half ? 8 : 16;   // flip_half: alpha a only
  for (int box = 0; box < nbox; ++box) {
    const int sa = box >> 3, sb = (box >> 2) & 1, si = (box >> 1) & 1,
              sj = box & 1;
    if (sa + sb != si + sj) continue;
    const int na = p.vrun[2 * sa + 1], nb = p.vrun[2 * sb + 1],
              ni = p.orun[2 * si + 1], nj = p.orun[2 * sj + 1];
    const int64_t size = (int64_t)na * nb * ni * nj;
    if (k < size) {
      int r;
      k = divmod(k, nj, r);
      q.j = p.orun[2 * sj] + r;
      k = divmod(k, ni, r);
      q.i = p.orun[2 * si] + r;
      k = divmod(k, nb, r);
      q.b = p.vrun[2 * sb] + r;
      q.a = p.vrun[2 * sa] + (int)k;
      return true;
    }
    k -= size;
  }
  return false;
}

template <bool DOUBLES>
__global__ __launch_bounds__(UT) void update_rec_kernel(UpdateArgs p,
                                                        int mode) {
  extern __shared__ double co[];   // a[ng], b[ng], c[ng], g[ng]
  const int ng = p.ng, nv = p.nv, no = p.no;
  for (int k = threadIdx.x; k < 3 * ng; k += UT) co[k] = p.rc[k];
  if (mode == 1)
    for (int k = threadIdx.x; k < ng; k += UT) co[3 * ng + k] = p.g[k];
  __syncthreads();
  const double *ca = co, *cb = co + ng, *cc = co + 2 * ng, *cg = co + 3 * ng;
  const int64_t per = DOUBLES ? (int64_t)nv * nv * no * no : (int64_t)nv * no;
  const int64_t per1 = (int64_t)nv * no;
  const double *src = p.src ? p.src : p.tbar;
  double dsum = 0.0, nsum = 0.0, esum = 0.0;
  // with label runs only the allowed doubles are enumerated
  const bool runs = DOUBLES && p.nallowed > 0;
  const bool half = runs && p.flip_half && mode == 1;
  const int64_t count = runs ? (half ? p.nallowed / 2 : p.nallowed) : per;
  const int hv = p.vrun[1], ho = p.orun[1];
  for (int64_t k = blockIdx.x * (int64_t)UT + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * UT) {
    Elem q;
    int64_t e = k, eimg = 0;
    if (runs) {
      if (!run_elem(p, k, q)) break;
      e = (((int64_t)q.a * nv + q.b) * no + q.i) * no + q.j;
      if (half) {
        auto fv = [&](int x) { return x < hv ? x + hv : x - hv; };
        auto fo = [&](int x) { return x < ho ? x + ho : x - ho; };
        eimg = (((int64_t)fv(q.a) * nv + fv(q.b)) * no + fo(q.i)) * no +
               fo(q.j);
      }
    } else {
      q = decode<DOUBLES>(e, nv, no);
    }
    if (DOUBLES && p.lv && p.lv[q.a] + p.lv[q.b] != p.lo[q.i] + p.lo[q.j])
      continue;   // zero in Tbar and T: contributes nothing
    double d;
    if (p.D)
      d = p.D[e];
    else if (DOUBLES)
      d = p.ev[q.a] + p.ev[q.b] - p.eo[q.i] - p.eo[q.j];
    else
      d = p.ev[q.a] - p.eo[q.i];
    const double r = exp(-p.delta * d), r2 = r * r;
    const double rs = p.rec == 1 ? r : r2;
    double tbm1 = 0.0, tbm2 = 0.0, tm1 = 0.0, tm2 = 0.0, eacc = 0.0;
    for (int y0 = 0; y0 < ng; y0 += RCH) {
      double tb[RCH], to[RCH];
#pragma unroll
      for (int k = 0; k < RCH; ++k)
        if (y0 + k < ng) tb[k] = src[(y0 + k) * per + e];
      if (p.corr) {
#pragma unroll
        for (int k = 0; k < RCH; ++k)
          if (y0 + k < ng) tb[k] += p.corr[(y0 + k) * per + e];
      }
      if (mode == 1) {
#pragma unroll
        for (int k = 0; k < RCH; ++k)
          if (y0 + k < ng) to[k] = p.told[(y0 + k) * per + e];
      }
#pragma unroll
      for (int k = 0; k < RCH; ++k) {
        const int y = y0 + k;
        if (y < ng) {
          const double tn = rs * (p.rec == 1 ? tm1 : tm2) +
                            ca[y] * (r2 * tbm2) + cb[y] * (r * tbm1) +
                            cc[y] * tb[k];
          tm2 = tm1;
          tm1 = tn;
          tbm2 = tbm1;
          tbm1 = tb[k];
          if (mode == 0) {
            p.tbar[y * per + e] = tn;
          } else {
            const double df = tn - to[k];
            const double tnew = p.damp * to[k] + (1.0 - p.damp) * tn;
            p.told[y * per + e] = tnew;
            if (half) p.told[y * per + eimg] = tnew;
            const double w = half ? 2.0 : 1.0;
            dsum += w * (df * df);
            nsum += w * (tnew * tnew);
            if (DOUBLES) {
              const double *t1 = p.t1new + y * per1;
              const double qq = t1[q.a * no + q.i] * t1[q.b * no + q.j];
              eacc += w * (cg[y] * (0.25 * tnew + 0.5 * qq));
            } else {
              eacc += cg[y] * tnew;
            }
          }
        }
      }
    }
    if (mode == 1) esum += eacc * p.eint[e];
  }
  if (mode == 1) block_reduce3(dsum, nsum, esum, p.partials);
}

template <bool DOUBLES>
int launch_update(const UpdateArgs &a, int mode, hipStream_t s) {
  const int64_t per = DOUBLES ? (a.rec && a.nallowed > 0
                                     ? a.nallowed
                                     : (int64_t)a.nv * a.nv * a.no * a.no)
                              : (int64_t)a.nv * a.no;
  const int blocks =
      (int)std::max<int64_t>(1, std::min<int64_t>((per + UT - 1) / UT, MAX_BLOCKS));
  if (a.rec) {
    require(!a.flip_half ||
                (a.vrun[1] == a.vrun[3] && a.orun[1] == a.orun[3] &&
                 a.vrun[0] == 0 && a.vrun[2] == a.vrun[1] &&
                 a.orun[0] == 0 && a.orun[2] == a.orun[1]),
            "flip-half update needs equal alpha and beta runs");
    hipLaunchKernelGGL((update_rec_kernel<DOUBLES>), dim3(blocks), dim3(UT),
                       sizeof(double) * 4 * a.ng, s, a, mode);
    FTCC_CHECK_HIP(hipGetLastError());
    return blocks;
  }
  if (a.ng <= MAX_LDS_NG) {
    const size_t lds = sizeof(double) * (a.ng * a.ng + a.ng + 2 * a.ng * UT);
    static std::once_flag attr;
    set_max_lds_once(attr, (const void *)update_kernel<DOUBLES, true>,
                     (int)(sizeof(double) * (MAX_LDS_NG * MAX_LDS_NG +
                                             MAX_LDS_NG + 2 * MAX_LDS_NG * UT)));
    hipLaunchKernelGGL((update_kernel<DOUBLES, true>), dim3(blocks), dim3(UT),
                       lds, s, a, mode);
  } else {
    hipLaunchKernelGGL((update_kernel<DOUBLES, false>), dim3(blocks),
                       dim3(UT), 0, s, a, mode);
  }
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

// Single grid point y of the causal integral (the pointwise schedule):
//   T(tau_y) = sum_x G[y,x] exp((tau_x - tau_y) d) Tbar(tau_x)  (x < y;
//   weight 1 for x >= y), reading the ng slices of Tbar.
// mode 0 stores T into told (the slice of point y); mode 1 accumulates
// |T - told|^2 into the first partial and damps told = damp*told +
// (1-damp)*T (kelvin/cc_utils.py:226-233).
template <bool DOUBLES>
__global__ __launch_bounds__(UT) void point_kernel(UpdateArgs p, int y,
                                                   int mode) {
  extern __shared__ double grow[];   // G[y][0..ng), ti[0..ng)
  const int ng = p.ng, nv = p.nv, no = p.no;
  for (int k = threadIdx.x; k < ng; k += UT) {
    grow[k] = p.G[(int64_t)y * ng + k];
    grow[ng + k] = p.ti[k];
  }
  __syncthreads();
  const double *gy = grow, *tis = grow + ng;
  const int64_t per = DOUBLES ? (int64_t)nv * nv * no * no : (int64_t)nv * no;
  double dsum = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)UT + threadIdx.x; e < per;
       e += (int64_t)gridDim.x * UT) {
    const Elem q = decode<DOUBLES>(e, nv, no);
    if (DOUBLES && p.lv && p.lv[q.a] + p.lv[q.b] != p.lo[q.i] + p.lo[q.j])
      continue;   // zero in Tbar and T: contributes nothing
    double d;
    if (p.D)
      d = p.D[e];
    else if (DOUBLES)
      d = p.ev[q.a] + p.ev[q.b] - p.eo[q.i] - p.eo[q.j];
    else
      d = p.ev[q.a] - p.eo[q.i];
    const double r = p.uniform ? exp(-p.delta * d) : 0.0;
    double t = 0.0, w = 1.0;
    for (int x = y - 1; x >= 0; --x) {
      if (p.uniform) w *= r;
      const double gx = gy[x];
      if (gx == 0.0) continue;
      const double wx = p.uniform ? w : exp((tis[x] - tis[y]) * d);
      t += gx * (wx * p.tbar[x * per + e]);
    }
    for (int x = y; x < ng; ++x) {
      const double gx = gy[x];
      if (gx != 0.0) t += gx * p.tbar[x * per + e];
    }
    if (mode == 0) {
      p.told[e] = t;
    } else {
      const double to = p.told[e];
      const double df = t - to;
      dsum += df * df;
      p.told[e] = p.damp * to + (1.0 - p.damp) * t;
    }
  }
  if (mode == 1) block_reduce3(dsum, 0.0, 0.0, p.partials);
}

__global__ void reduce_partials_kernel(const double *partials, int nblocks,
                                       double *out3) {
  __shared__ double red[3][256];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int b = threadIdx.x; b < nblocks; b += 256) {
    s0 += partials[3 * b];
    s1 += partials[3 * b + 1];
    s2 += partials[3 * b + 2];
  }
  red[0][threadIdx.x] = s0;
  red[1][threadIdx.x] = s1;
  red[2][threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
      red[2][threadIdx.x] += red[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) out3[threadIdx.x] = red[threadIdx.x][0];
}

// partials: {0, 0, energy} so that the same reducer applies
__global__ __launch_bounds__(UT) void energy_t2_kernel(
    const double *T1, const double *T2, const double *g, const double *oovvT,
    int ng, int nv, int no, int qterm, double *partials) {
  const int64_t per = (int64_t)nv * nv * no * no;
  const int64_t per1 = (int64_t)nv * no;
  double esum = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)UT + threadIdx.x; e < per;
       e += (int64_t)gridDim.x * UT) {
    const Elem q = decode<true>(e, nv, no);
    double acc = 0.0;
    for (int y = 0; y < ng; ++y) {
      double v = 0.25 * T2[y * per + e];
      if (qterm) {
        const double *t1 = T1 + y * per1;
        v += 0.5 * t1[q.a * no + q.i] * t1[q.b * no + q.j];
      }
      acc += g[y] * v;
    }
    esum += acc * oovvT[e];
  }
  block_reduce3(0.0, 0.0, esum, partials);
}

__global__ __launch_bounds__(UT) void energy_t1_kernel(const double *T1,
                                                       const double *g,
                                                       const double *fovT,
                                                       int ng, int nv, int no,
                                                       double *partials) {
  const int64_t per = (int64_t)nv * no;
  double esum = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)UT + threadIdx.x; e < per;
       e += (int64_t)gridDim.x * UT) {
    double acc = 0.0;
    for (int y = 0; y < ng; ++y) acc += g[y] * T1[y * per + e];
    esum += acc * fovT[e];
  }
  block_reduce3(0.0, 0.0, esum, partials);
}

__global__ __launch_bounds__(UT) void sumsq_kernel(const double *x, int64_t n,
                                                   double *partials) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)UT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * UT)
    s += x[i] * x[i];
  block_reduce3(s, 0.0, 0.0, partials);
}

__global__ __launch_bounds__(UT) void vsum_kernel(const double *x, int64_t n,
                                                  double *partials) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)UT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * UT)
    s += x[i];
  block_reduce3(s, 0.0, 0.0, partials);
}

// Entries of nb dense (d0,d1,d2,d3) blocks that are nonzero although
// l0[i0] + l1[i1] != l2[i2] + l3[i3] (a null label array reads as 0): the
// spin-conservation test of integral blocks and amplitudes.
__global__ __launch_bounds__(UT) void spin_violations_kernel(
    const double *a, int d0, int d1, int d2, int d3, int nb, const int *l0,
    const int *l1, const int *l2, const int *l3, double *partials) {
  const int64_t per = (int64_t)d0 * d1 * d2 * d3;
  const int64_t n = per * nb;
  double cnt = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)UT + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * UT) {
    int64_t r = e % per;
    const int i3 = (int)(r % d3);
    r /= d3;
    const int i2 = (int)(r % d2);
    r /= d2;
    const int i1 = (int)(r % d1);
    const int i0 = (int)(r / d1);
    const int sx = (l0 ? l0[i0] : 0) + (l1 ? l1[i1] : 0);
    const int sy = (l2 ? l2[i2] : 0) + (l3 ? l3[i3] : 0);
    if (sx != sy && a[e] != 0.0) cnt += 1.0;
  }
  block_reduce3(0.0, 0.0, cnt, partials);
}

// Elements that differ from their image under the spin flip (index x of a
// dimension with half h maps to x + h below h, x - h above; h = 0: fixed).
__global__ __launch_bounds__(UT) void flip_violations_kernel(
    const double *a, int d0, int d1, int d2, int d3, int h0, int h1, int h2,
    int h3, int nb, double tol, double *partials) {
  const int64_t per = (int64_t)d0 * d1 * d2 * d3;
  double cnt = 0.0;
  auto fl = [](int x, int h) { return h ? (x < h ? x + h : x - h) : x; };
  for (int64_t g = blockIdx.x * (int64_t)UT + threadIdx.x; g < per * nb;
       g += (int64_t)gridDim.x * UT) {
    const int64_t base = (g / per) * per;
    int64_t r = g - base;
    const int i3 = (int)(r % d3);
    r /= d3;
    const int i2 = (int)(r % d2);
    r /= d2;
    const int i1 = (int)(r % d1);
    const int i0 = (int)(r / d1);
    const int64_t f =
        base +
        (((int64_t)fl(i0, h0) * d1 + fl(i1, h1)) * d2 + fl(i2, h2)) * d3 +
        fl(i3, h3);
    const double x = a[g], y = a[f];
    if (tol == 0.0 ? x != y
                   : !(fabs(x - y) <= tol * fmax(fabs(x), fabs(y))))
      cnt += 1.0;
  }
  block_reduce3(0.0, 0.0, cnt, partials);
}

int nblocks_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + UT - 1) / UT, MAX_BLOCKS));
}

}  // namespace

bool update_needs_copy(const UpdateArgs &a) {
  return !a.rec && a.ng > MAX_LDS_NG;
}

int update_t1(const UpdateArgs &a, int mode, hipStream_t s) {
  return launch_update<false>(a, mode, s);
}

int update_t2(const UpdateArgs &a, int mode, hipStream_t s) {
  return launch_update<true>(a, mode, s);
}

void reduce_partials(const double *partials, int nblocks, double *out3,
                     hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, s,
                     partials, nblocks, out3);
  FTCC_CHECK_HIP(hipGetLastError());
}

int energy_t2(const double *T1, const double *T2, const double *g,
              const double *oovvT, int ng, int nv, int no, int qterm,
              double *partials, hipStream_t s) {
  const int blocks = nblocks_for((int64_t)nv * nv * no * no);
  hipLaunchKernelGGL(energy_t2_kernel, dim3(blocks), dim3(UT), 0, s, T1, T2,
                     g, oovvT, ng, nv, no, qterm, partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int energy_t1(const double *T1, const double *g, const double *fovT, int ng,
              int nv, int no, double *partials, hipStream_t s) {
  const int blocks = nblocks_for((int64_t)nv * no);
  hipLaunchKernelGGL(energy_t1_kernel, dim3(blocks), dim3(UT), 0, s, T1, g,
                     fovT, ng, nv, no, partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int reduction_blocks(int64_t n) { return nblocks_for(n); }

int sumsq(const double *x, int64_t n, double *partials, hipStream_t s) {
  const int blocks = nblocks_for(n);
  hipLaunchKernelGGL(sumsq_kernel, dim3(blocks), dim3(UT), 0, s, x, n,
                     partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int vsum(const double *x, int64_t n, double *partials, hipStream_t s) {
  const int blocks = nblocks_for(n);
  hipLaunchKernelGGL(vsum_kernel, dim3(blocks), dim3(UT), 0, s, x, n,
                     partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

// One stage of an explicit Runge-Kutta step of dT/dtau = Tbar(T) - D T (the
// right-hand side of kelvin/td_ccsd.py:646-652), over p.ng independent
// slices: with k = Tbar - D*in, where in is T on the first stage and var
// after it, acc <- (first ? 0 : acc) + b*k; the last stage adds h*acc to T,
// the others write the next stage's input T + hnext*k into var.  zero: the
// right-hand side is zero (amplitudes held fixed).  With var == nullptr the
// kernel only forms k into acc (first stage, b = 1, not last).
template <bool DOUBLES>
__global__ __launch_bounds__(UT) void rk_stage_kernel(
    UpdateArgs p, const double *tb, double *T, double *var, double *acc,
    double b, double hnext, double h, int first, int last, int zero) {
  const int nv = p.nv, no = p.no;
  const int64_t per = DOUBLES ? (int64_t)nv * nv * no * no : (int64_t)nv * no;
  const int64_t total = per * p.ng;
  for (int64_t x = blockIdx.x * (int64_t)UT + threadIdx.x; x < total;
       x += (int64_t)gridDim.x * UT) {
    double k = 0.0;
    if (!zero) {
      const int64_t e = x % per;
      double d;
      if (p.D) {
        d = p.D[e];
      } else {
        const Elem q = decode<DOUBLES>(e, nv, no);
        d = DOUBLES ? p.ev[q.a] + p.ev[q.b] - p.eo[q.i] - p.eo[q.j]
                    : p.ev[q.a] - p.eo[q.i];
      }
      const double in = (first || !var) ? T[x] : var[x];
      k = tb[x] - d * in;
    }
    const double a = (first ? 0.0 : acc[x]) + b * k;
    if (last) {
      T[x] += h * a;
    } else {
      acc[x] = a;
      if (var) var[x] = T[x] + hnext * k;
    }
  }
}

// One pass of an implicit step of dT/dtau = Tbar(T) - D T (kelvin/
// propagation.py:55-128) over p.ng slices.  mode 0 (predictor): with
// k = Tbar - D T, k1 = dold = h k and var = T + h k.  mode 1 (iteration):
// k = Tbar - D var, d = h k (be, cn) or (h k + k1) / 2 (am2); the block's
// sums of (d - dold)^2 and d^2 go to p.partials; dold <- (1 - alpha) d +
// alpha dold and var = T + dold.  mode 2 (finish): T += dold, or
// (dold + k1) / 2 for cn.  zero: the right-hand side is zero.
template <bool DOUBLES>
__global__ __launch_bounds__(UT) void implicit_kernel(
    UpdateArgs p, const double *tb, double *T, double *var, double *dold,
    double *k1, double h, double alpha, int mode, int scheme, int zero) {
  const int nv = p.nv, no = p.no;
  const int64_t per = DOUBLES ? (int64_t)nv * nv * no * no : (int64_t)nv * no;
  const int64_t total = per * p.ng;
  double sdiff = 0.0, snorm = 0.0;
  for (int64_t x = blockIdx.x * (int64_t)UT + threadIdx.x; x < total;
       x += (int64_t)gridDim.x * UT) {
    if (mode == 2) {
      if (!zero) T[x] += scheme == 1 ? 0.5 * (dold[x] + k1[x]) : dold[x];
      continue;
    }
    double k = 0.0;
    if (!zero) {
      const int64_t e = x % per;
      double dd;
      if (p.D) {
        dd = p.D[e];
      } else {
        const Elem q = decode<DOUBLES>(e, nv, no);
        dd = DOUBLES ? p.ev[q.a] + p.ev[q.b] - p.eo[q.i] - p.eo[q.j]
                     : p.ev[q.a] - p.eo[q.i];
      }
      k = tb[x] - dd * (mode == 0 ? T[x] : var[x]);
    }
    if (mode == 0) {
      const double hk = h * k;
      k1[x] = hk;
      dold[x] = hk;
      var[x] = T[x] + hk;
    } else {
      const double d = scheme == 2 ? 0.5 * (h * k + k1[x]) : h * k;
      const double o = dold[x];
      sdiff += (d - o) * (d - o);
      snorm += d * d;
      const double n = (1.0 - alpha) * d + alpha * o;
      dold[x] = n;
      var[x] = T[x] + n;
    }
  }
  if (mode == 1) block_reduce3(sdiff, snorm, 0.0, p.partials);
}

int implicit_pass(const UpdateArgs &a, bool doubles, const double *tb,
                  double *T, double *var, double *dold, double *k1, double h,
                  double alpha, int mode, int scheme, bool zero,
                  hipStream_t s) {
  const int64_t per = doubles ? (int64_t)a.nv * a.nv * a.no * a.no
                              : (int64_t)a.nv * a.no;
  const int blocks = nblocks_for(per * a.ng);
  if (doubles)
    hipLaunchKernelGGL(implicit_kernel<true>, dim3(blocks), dim3(UT), 0, s, a,
                       tb, T, var, dold, k1, h, alpha, mode, scheme,
                       zero ? 1 : 0);
  else
    hipLaunchKernelGGL(implicit_kernel<false>, dim3(blocks), dim3(UT), 0, s,
                       a, tb, T, var, dold, k1, h, alpha, mode, scheme,
                       zero ? 1 : 0);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int rk_stage(const UpdateArgs &a, bool doubles, const double *tb, double *T,
             double *var, double *acc, double b, double hnext, double h,
             bool first, bool last, bool zero, hipStream_t s) {
  const int64_t per = doubles ? (int64_t)a.nv * a.nv * a.no * a.no
                              : (int64_t)a.nv * a.no;
  const int blocks = (int)std::max<int64_t>(
      1, std::min<int64_t>((per * a.ng + UT - 1) / UT, MAX_BLOCKS));
  if (doubles)
    hipLaunchKernelGGL(rk_stage_kernel<true>, dim3(blocks), dim3(UT), 0, s, a,
                       tb, T, var, acc, b, hnext, h, first ? 1 : 0,
                       last ? 1 : 0, zero ? 1 : 0);
  else
    hipLaunchKernelGGL(rk_stage_kernel<false>, dim3(blocks), dim3(UT), 0, s,
                       a, tb, T, var, acc, b, hnext, h, first ? 1 : 0,
                       last ? 1 : 0, zero ? 1 : 0);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int spin_violations(const double *a, int d0, int d1, int d2, int d3, int nb,
                    const int *l0, const int *l1, const int *l2, const int *l3,
                    double *partials, hipStream_t s) {
  const int blocks = nblocks_for((int64_t)d0 * d1 * d2 * d3 * nb);
  hipLaunchKernelGGL(spin_violations_kernel, dim3(blocks), dim3(UT), 0, s, a,
                     d0, d1, d2, d3, nb, l0, l1, l2, l3, partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int flip_violations(const double *a, int d0, int d1, int d2, int d3, int h0,
                    int h1, int h2, int h3, double *partials, hipStream_t s,
                    int nb, double tol) {
  const int blocks = nblocks_for((int64_t)d0 * d1 * d2 * d3 * nb);
  hipLaunchKernelGGL(flip_violations_kernel, dim3(blocks), dim3(UT), 0, s, a,
                     d0, d1, d2, d3, h0, h1, h2, h3, nb, tol, partials);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

int integrate_point(const UpdateArgs &a, bool doubles, int y, int mode,
                    hipStream_t s) {
  const int64_t per = doubles ? (int64_t)a.nv * a.nv * a.no * a.no
                              : (int64_t)a.nv * a.no;
  const int blocks =
      (int)std::max<int64_t>(1, std::min<int64_t>((per + UT - 1) / UT, MAX_BLOCKS));
  const size_t lds = sizeof(double) * 2 * a.ng;
  if (doubles)
    hipLaunchKernelGGL(point_kernel<true>, dim3(blocks), dim3(UT), lds, s, a,
                       y, mode);
  else
    hipLaunchKernelGGL(point_kernel<false>, dim3(blocks), dim3(UT), lds, s,
                       a, y, mode);
  FTCC_CHECK_HIP(hipGetLastError());
  return blocks;
}

}  // namespace ftcc
