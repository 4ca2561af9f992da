// Adjoint of the FT-CCSD fixed point: the transpose of the residual's
// Jacobian (the Lambda equations), the reverse-time integration, the damped
// lambda update, the derivative of the Fermi-weighted integrals along a
// change of beta or mu, and the derivative of the quadrature with beta.
//
// kelvin obtains E, S and N from lambda amplitudes and response densities
// (kelvin/ccsd.py:128-214, :882-961, :1075-1148, :1333-1492; the lambda
// residual is cqcpy cc_equations._Lambda_opt, called at
// kelvin/ft_cc_equations.py:385-409).  Here the same derivative of the
// grand potential is taken as the adjoint of the discretised fixed point
// T = Phi(T):
//     dOmega_cc/dtheta = dE/dtheta + sum_y <Lam_y, dPhi_y/dtheta>,
//     Lam = dE/dT + (dPhi/dT)^T Lam,
// with (dPhi/dT)^T Lam = -J^T (Int^T Lam) at every grid point, J the
// Jacobian of the Stanton-Gauss residual and Int^T the transpose of the
// exp(-tau D) integration.  The residual is linear in the integrals, so
// dPhi/dtheta through them is one residual evaluation with the derivative
// blocks (ftcc_set_system_response).
//
// J^T is the reverse-mode derivative of the residual (every term of the
// forward equations transposed; see the oracle's stanton_rhs_transpose),
// with the amplitude-only parts folded into intermediates that are fixed
// during the lambda iteration (computed once per grid point by
// ftcc_adjoint_begin and kept while they fit FTCC_ADJOINT_CACHE_GB,
// default 64; the other grid points recompute theirs at every application
// from the amplitudes, which are the checkpoint):
//   Wh_mnij = 1/2 (oooo + P(ij) t1.ooov) + 1/4 oovv.tau      (O^4)
//   Wh_efab = 1/2 (vvvv - P(ab) t1.vovv) + 1/4 tau.oovv      (V^4, [e,f,a,b])
//   Wmbej   = <mb||ej> - t1.vovv + t1.ooov - oovv.(t2/2 + t1 t1)
//   Zt[e,a,b,i] = sum_mn tau[a,b,m,n] ooov[m,n,i,e]
//   Zv[a,m,i,j] = sum_ef vovv[a,m,e,f] tau[e,f,i,j]
// and the Fock-type Fme, Fae, Fmi, Fbe, Fmj.  One application costs five
// O(n^6) products per grid point (two ladders, three rings) and O(n^5)
// contractions, all in FP64 on the GEMM engine (dense operands; every
// index is addressed through its own stride).
#include "ftcc_ctx.h"

#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

using namespace ftcc;

namespace ftcc_core {

namespace {

constexpr int AT = 256;            // threads per block of the kernels below
constexpr int MAX_RED_BLOCKS = 4096;
constexpr int MAX_STAGE_NG = 32;   // grids up to this size stage in LDS

#define ADJ_LOOP(i, n)                                                     \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
       i += (int64_t)gridDim.x * blockDim.x)

int grid_for(int64_t n, int cap = 65536) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + AT - 1) / AT, cap));
}

// per-block sums of four values into partials[4*block + k]
__device__ inline void block_sum4(double v0, double v1, double v2, double v3,
                                  double *out) {
  __shared__ double red[4][AT / 64];
  for (int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_xor(v0, off, 64);
    v1 += __shfl_xor(v1, off, 64);
    v2 += __shfl_xor(v2, off, 64);
    v3 += __shfl_xor(v3, off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    red[0][w] = v0;
    red[1][w] = v1;
    red[2][w] = v2;
    red[3][w] = v3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[4] = {0, 0, 0, 0};
    for (int k = 0; k < AT / 64; ++k)
      for (int q = 0; q < 4; ++q) s[q] += red[q][k];
    for (int q = 0; q < 4; ++q) out[4 * blockIdx.x + q] = s[q];
  }
}

// sums of the per-block quadruples in a fixed order (deterministic)
__global__ void sum4_kernel(const double *partials, int nblocks,
                            double *out) {
  __shared__ double red[4][AT];
  double s[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += AT)
    for (int q = 0; q < 4; ++q) s[q] += partials[4 * b + q];
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < 4; ++q) {
      double t = 0.0;
      for (int k = 0; k < AT; ++k) t += red[q][k];
      out[q] = t;
    }
  }
}

// tau = t2 + t1[a,i]t1[b,j] - t1[b,i]t1[a,j], taut = t2 + (...)/2,
// ar = t2/2 + t1[a,i]t1[b,j] (the ring amplitude of Wmbej, [f,b,j,n])
__global__ void tau_kernel(const double *t1, const double *t2, int V, int O,
                           double *tau, double *taut, double *ar) {
  const int64_t n = (int64_t)V * V * O * O;
  ADJ_LOOP(e, n) {
    int64_t r = e;
    const int j = (int)(r % O);
    r /= O;
    const int i = (int)(r % O);
    r /= O;
    const int b = (int)(r % V);
    const int a = (int)(r / V);
    const double p = t1[a * O + i] * t1[b * O + j];
    const double tt = p - t1[b * O + i] * t1[a * O + j];
    if (tau) tau[e] = t2[e] + tt;
    if (taut) taut[e] = t2[e] + 0.5 * tt;
    if (ar) ar[e] = 0.5 * t2[e] + p;
  }
}

// P(x)[a,b,i,j] = (x[abij] - x[baij] - x[abji] + x[baji]) / 4
__device__ inline double proj(const double *x, int a, int b, int i, int j,
                              int V, int O) {
  auto at = [&](int p, int q, int r, int s) {
    return x[(((int64_t)p * V + q) * O + r) * O + s];
  };
  return 0.25 * (at(a, b, i, j) - at(b, a, i, j) - at(a, b, j, i) +
                 at(b, a, j, i));
}

// S = P(gtau + gt/2) and out2 = P(acc + gt/2) + S: the doubles of J^T
// (the transposed residual is projected onto antisymmetric tensors) and
// the adjoint of tau/taut for the t1 chain rule
__global__ void assemble_kernel(const double *acc, const double *gtau,
                                const double *gt, int V, int O, double *S,
                                double *out2) {
  const int64_t n = (int64_t)V * V * O * O;
  ADJ_LOOP(e, n) {
    int64_t r = e;
    const int j = (int)(r % O);
    r /= O;
    const int i = (int)(r % O);
    r /= O;
    const int b = (int)(r % V);
    const int a = (int)(r / V);
    const double pt = proj(gt, a, b, i, j, V, O);
    const double s = proj(gtau, a, b, i, j, V, O) + 0.5 * pt;
    S[e] = s;
    out2[e] = proj(acc, a, b, i, j, V, O) + 0.5 * pt + s;
  }
}

// Transpose of the causal integration over the slices:
//   out[x] = sum_y G[y,x] P_yx lam[y],  P_yx = exp((ti_x - ti_y) d) for
// x < y and 1 for x >= y (the integrator's convention), d the denominator
// of the element (explicit D or from the energies).  On an equally spaced
// grid P_yx = r^(y-x), r = exp(-delta d).
template <bool DOUBLES>
__global__ __launch_bounds__(AT) void int_adjoint_kernel(
    const double *lam, double *out, int ng, int V, int O, const double *ti,
    const double *G, const double *ev, const double *eo, const double *D,
    int uniform, double delta) {
  extern __shared__ double col[];   // [ng][AT] when staged
  const int64_t per = DOUBLES ? (int64_t)V * V * O * O : (int64_t)V * O;
  const bool stage = ng <= MAX_STAGE_NG;
  ADJ_LOOP(e, per) {
    double d;
    if (D) {
      d = D[e];
    } else if (DOUBLES) {
      int64_t r = e;
      const int j = (int)(r % O);
      r /= O;
      const int i = (int)(r % O);
      r /= O;
      const int b = (int)(r % V);
      const int a = (int)(r / V);
      d = ev[a] + ev[b] - eo[i] - eo[j];
    } else {
      d = ev[e / O] - eo[e % O];
    }
    if (stage)
      for (int y = 0; y < ng; ++y) col[y * AT + threadIdx.x] = lam[y * per + e];
    const double r = uniform ? exp(-delta * d) : 0.0;
    for (int x = 0; x < ng; ++x) {
      double acc = 0.0, w = 1.0;
      for (int y = 0; y < ng; ++y) {
        if (y > x && uniform) w *= r;
        const double gyx = G[y * ng + x];
        if (gyx == 0.0) continue;
        const double l = stage ? col[y * AT + threadIdx.x] : lam[y * per + e];
        const double p =
            y > x ? (uniform ? w : exp((ti[x] - ti[y]) * d)) : 1.0;
        acc += gyx * (p * l);
      }
      out[x * per + e] = acc;
    }
  }
}

// Damped lambda update of ng slices: Lnew = dE - G (dE = src + scale *
// ebase per slice), L <- damp L + (1 - damp) Lnew; partial sums of
// w_y^2 |Lnew - L|^2 and w_y^2 |L_damped|^2 (w: the per-slice factor to
// kelvin's lambda scale)
__global__ void lambda_update_kernel(double *L, const double *Gj,
                                     const double *ebase, const double *escale,
                                     const double *w, int ng, int64_t per,
                                     int64_t ebase_stride, double damp,
                                     double *partials) {
  double d2 = 0.0, n2 = 0.0;
  ADJ_LOOP(g, per * ng) {
    const int y = (int)(g / per);
    const int64_t e = g - (int64_t)y * per;
    const double lnew = escale[y] * ebase[y * ebase_stride + e] - Gj[g];
    const double lo = L[g];
    const double df = lnew - lo;
    const double ld = damp * lo + (1.0 - damp) * lnew;
    L[g] = ld;
    const double ww = w[y] * w[y];
    d2 += ww * df * df;
    n2 += ww * ld * ld;
  }
  block_sum4(d2, n2, 0.0, 0.0, partials);
}

// sum over elements and (y, x) of lam[y] (Gd[y,x] - G[y,x] (ti_y - ti_x)
// d / beta) P_yx tb[x]
template <bool DOUBLES>
__global__ __launch_bounds__(AT) void grid_deriv_kernel(
    const double *lam, const double *tb, int ng, int V, int O,
    const double *ti, const double *G, const double *Gd, const double *ev,
    const double *eo, const double *D, double beta, double *partials) {
  const int64_t per = DOUBLES ? (int64_t)V * V * O * O : (int64_t)V * O;
  double acc = 0.0;
  ADJ_LOOP(e, per) {
    double d;
    if (D) {
      d = D[e];
    } else if (DOUBLES) {
      int64_t r = e;
      const int j = (int)(r % O);
      r /= O;
      const int i = (int)(r % O);
      r /= O;
      const int b = (int)(r % V);
      const int a = (int)(r / V);
      d = ev[a] + ev[b] - eo[i] - eo[j];
    } else {
      d = ev[e / O] - eo[e % O];
    }
    for (int y = 0; y < ng; ++y) {
      const double l = lam[y * per + e];
      if (l == 0.0) continue;
      double s = 0.0;
      for (int x = 0; x < ng; ++x) {
        const double gyx = G[y * ng + x], gd = Gd[y * ng + x];
        if (gyx == 0.0 && gd == 0.0) continue;
        const double p = x < y ? exp((ti[x] - ti[y]) * d) : 1.0;
        s += (gd - gyx * (ti[y] - ti[x]) * d / beta) * p * tb[x * per + e];
      }
      acc += l * s;
    }
  }
  block_sum4(acc, 0.0, 0.0, 0.0, partials);
}

// out[y][e] = s[y] * base[y * stride + e] over ng slices of per elements
__global__ void scale_slices_kernel(double *out, const double *base,
                                    const double *s, int ng, int64_t per,
                                    int64_t stride) {
  ADJ_LOOP(g, per * ng) {
    const int y = (int)(g / per);
    out[g] = s[y] * base[y * stride + (g - (int64_t)y * per)];
  }
}

__global__ void dot_kernel(const double *x, const double *y, int64_t n,
                           double *partials) {
  double s = 0.0;
  ADJ_LOOP(e, n) s += x[e] * y[e];
  block_sum4(s, 0.0, 0.0, 0.0, partials);
}

// d/dtheta of the Fermi-weighted blocks: eri[p0,p1,p2,p3] w0 w1 w2 w3
// (c0 + c1 + c2 + c3) with c = d ln w along the change
__global__ void response4_kernel(const double *eri, int n, const int *p0,
                                 const int *p1, const int *p2, const int *p3,
                                 const double *w0, const double *w1,
                                 const double *w2, const double *w3,
                                 const double *c0, const double *c1,
                                 const double *c2, const double *c3, int d0,
                                 int d1, int d2, int d3, double *out) {
  const int64_t total = (int64_t)d0 * d1 * d2 * d3;
  ADJ_LOOP(e, total) {
    int64_t r = e;
    const int i3 = (int)(r % d3);
    r /= d3;
    const int i2 = (int)(r % d2);
    r /= d2;
    const int i1 = (int)(r % d1);
    const int i0 = (int)(r / d1);
    const int64_t src =
        (((int64_t)p0[i0] * n + p1[i1]) * n + p2[i2]) * n + p3[i3];
    out[e] = eri[src] * w0[i0] * w1[i1] * w2[i2] * w3[i3] *
             (c0[i0] + c1[i1] + c2[i2] + c3[i3]);
  }
}

// dF[i,j] = (dfock[p,q] + (fock[p,q] - (p == q) en[p]) (ci + cj)) wi wj
__global__ void fock_response_kernel(const double *fock, const double *dfock,
                                     const double *en, int n, const int *p,
                                     const int *q, const double *wi,
                                     const double *wj, const double *ci,
                                     const double *cj, int di, int dj,
                                     double *out) {
  ADJ_LOOP(e, (int64_t)di * dj) {
    const int i = (int)(e / dj), j = (int)(e % dj);
    const int a = p[i], b = q[j];
    const double f = fock[(int64_t)a * n + b] - (a == b ? en[a] : 0.0);
    out[e] = (dfock[(int64_t)a * n + b] + f * (ci[i] + cj[j])) * wi[i] * wj[j];
  }
}

// co[k] = -fv[io[k]] dx[io[k]] / 2, cv[k] = fo[iv[k]] dx[iv[k]] / 2,
// the logarithmic derivatives of sqrt(fo) and sqrt(fv)
__global__ void log_weight_kernel(const double *fo, const double *fv,
                                  const double *dx, const int *io, int O,
                                  const int *iv, int V, double *co,
                                  double *cv) {
  ADJ_LOOP(k, (int64_t)O + V) {
    if (k < O) {
      const int p = io[k];
      co[k] = -0.5 * fv[p] * dx[p];
    } else {
      const int p = iv[k - O];
      cv[k - O] = 0.5 * fo[p] * dx[p];
    }
  }
}

// ---------------------------------------------------------------------------
// Contractions by index letters.  a..h are virtual-role indices (nv), i..n
// occupied-role ones (no); a tensor is dense in C order over its letters.
// con() maps C[...] = alpha sum_K A[...] B[...] + beta C[...] onto the GEMM
// engine: rows are C's letters in A, columns C's letters in B, K the
// letters A and B share; letters that are contiguous in every operand
// holding them merge into one level, and the level orders (and which
// operand is A) are chosen so that the engine's layout rules hold, writes
// of C are as contiguous as possible and K runs fastest.
// ---------------------------------------------------------------------------
struct Ten {
  const double *p = nullptr;
  std::string l;
  std::vector<int64_t> st;
  std::string rule;   // spin rule: letters left of '|' sum to those right
  int64_t stride(char x) const {
    const size_t k = l.find(x);
    return k == std::string::npos ? -1 : st[k];
  }
};

int letter_dim(const ftcc_ctx *c, char x) {
  require((x >= 'a' && x <= 'h') || (x >= 'i' && x <= 'n'),
          std::string("contraction: bad index letter ") + x);
  return (x >= 'i' && x <= 'n') ? c->no : c->nv;
}

// A dense tensor over `letters` (C order).  Its spin rule is `rule`, or for
// four letters "pq|rs" and for two "p|q" (true of the integrals, the
// amplitudes and every intermediate stored in the <pq||rs> order).
Ten ten(const ftcc_ctx *c, const double *p, const char *letters,
        const char *rule = nullptr) {
  Ten t;
  t.p = p;
  t.l = letters;
  t.st.assign(t.l.size(), 0);
  int64_t s = 1;
  for (int k = (int)t.l.size() - 1; k >= 0; --k) {
    t.st[k] = s;
    s *= letter_dim(c, t.l[k]);
  }
  for (size_t a = 0; a < t.l.size(); ++a)
    for (size_t b = a + 1; b < t.l.size(); ++b)
      require(t.l[a] != t.l[b], "contraction: repeated letter");
  if (rule) {
    t.rule = rule;
  } else if (t.l.size() == 4) {
    t.rule = t.l.substr(0, 2) + "|" + t.l.substr(2, 2);
  } else if (t.l.size() == 2) {
    t.rule = t.l.substr(0, 1) + "|" + t.l.substr(1, 1);
  }
  return t;
}

struct Level {
  int64_t dim = 1;
  std::vector<char> ls;   // letters, outer to inner
};

// the extent of every letter of one (sub-)contraction
struct Extents {
  int64_t d[128] = {};
  int64_t operator()(char x) const { return d[(unsigned char)x]; }
};

// merge consecutive letters of `order` that are contiguous in every operand
std::vector<Level> levels_of(const Extents &ext,
                             const std::vector<char> &order,
                             const std::vector<const Ten *> &ops) {
  std::vector<Level> out;
  for (char x : order) {
    bool merge = !out.empty();
    if (merge) {
      const char prev = out.back().ls.back();
      for (const Ten *t : ops)
        if (t->stride(prev) != t->stride(x) * ext(x)) merge = false;
    }
    if (merge) {
      out.back().ls.push_back(x);
      out.back().dim *= ext(x);
    } else {
      Level L;
      L.ls.push_back(x);
      L.dim = ext(x);
      out.push_back(L);
    }
  }
  return out;
}

int64_t lstride(const Ten &t, const Level &L) { return t.stride(L.ls.back()); }

struct Groups {
  std::vector<char> M, N, K;
};

Groups groups_of(const Ten &A0, const Ten &B0, const Ten &C) {
  Groups g;
  for (char x : C.l) {
    const bool inA = A0.stride(x) >= 0, inB = B0.stride(x) >= 0;
    require(inA != inB, "contraction: an output letter must be in A or B");
    (inA ? g.M : g.N).push_back(x);
  }
  for (char x : A0.l)
    if (C.stride(x) < 0) {
      require(B0.stride(x) >= 0, "contraction: a letter of A only");
      g.K.push_back(x);
    }
  for (char x : B0.l)
    if (C.stride(x) < 0)
      require(A0.stride(x) >= 0, "contraction: a letter of B only");
  return g;
}

struct Shape {
  int score = -1;
  bool swap = false;
  std::vector<Level> m, n, k;
};

// The level orders of C = A.B over the letter extents `ext` (score < 0:
// none satisfies the engine's layout rules); a search over the orders of
// each group and the choice of A, cached per thread by letters, strides
// and extents.
Shape shape_of(const Extents &ext, const Ten &A0, const Ten &B0,
               const Ten &C, const Groups &G) {
  static thread_local std::map<std::string, Shape> cache;
  std::string key = A0.l + "," + B0.l + ">" + C.l + ":";
  for (const Ten *t : {&A0, &B0, &C})
    for (int64_t x : t->st) key += std::to_string(x) + ",";
  for (const auto *grp : {&G.M, &G.N, &G.K})
    for (char x : *grp) key += std::to_string(ext(x)) + ";";
  const auto hit = cache.find(key);
  if (hit != cache.end()) return hit->second;
  Shape best;
  auto perms = [](std::vector<char> v) {
    std::vector<std::vector<char>> out;
    std::sort(v.begin(), v.end());
    do out.push_back(v);
    while (std::next_permutation(v.begin(), v.end()));
    return out;
  };
  const auto pM = perms(G.M), pN = perms(G.N), pK = perms(G.K);
  for (int sw = 0; sw < 2; ++sw) {
    const Ten &A = sw ? B0 : A0, &B = sw ? A0 : B0;
    const auto &pm = sw ? pN : pM, &pn = sw ? pM : pN;
    for (const auto &om : pm)
      for (const auto &on : pn)
        for (const auto &ok : pK) {
          const auto lm = levels_of(ext, om, {&A, &C});
          const auto ln = levels_of(ext, on, {&B, &C});
          const auto lk = levels_of(ext, ok, {&A, &B});
          if (lm.size() > 3 || ln.size() > 3 || lk.size() > 3) continue;
          const int64_t sAm = lm.empty() ? 0 : lstride(A, lm.back());
          const int64_t sBn = ln.empty() ? 0 : lstride(B, ln.back());
          const int64_t sAk = lk.empty() ? 0 : lstride(A, lk.back());
          const int64_t sBk = lk.empty() ? 0 : lstride(B, lk.back());
          const bool amf = sAm == 1 && sAk != 1, bnf = sBn == 1 && sBk != 1;
          if (lk.size() > 1 && (amf || bnf)) continue;
          int score = 0;
          const int64_t sCn = ln.empty() ? 0 : lstride(C, ln.back());
          const int64_t sCm = lm.empty() ? 0 : lstride(C, lm.back());
          if (sCn == 1) score += 4;          // contiguous writes along n
          else if (sCm == 1) score += 2;
          if (sAk == 1) score += 1;
          if (sBk == 1) score += 1;
          score = score * 16 + (12 - (int)(lm.size() + ln.size() + lk.size()));
          if (score > best.score) {
            best.score = score;
            best.swap = sw;
            best.m = lm;
            best.n = ln;
            best.k = lk;
          }
        }
  }
  cache[key] = best;
  return best;
}

// The GEMM of C = alpha A.B + beta C over the letter extents `ext`, or
// false when no level order satisfies the engine's layout rules.
bool plan(const Extents &ext, double alpha, const Ten &A0, const Ten &B0,
          double beta, const Ten &C, const Groups &G, Gemm &g) {
  const Shape best = shape_of(ext, A0, B0, C, G);
  if (best.score < 0) return false;
  const Ten &A = best.swap ? B0 : A0, &B = best.swap ? A0 : B0;
  g = Gemm();
  g.alpha = alpha;
  g.beta = beta;
  g.A = A.p;
  g.B = B.p;
  g.C = const_cast<double *>(C.p);
  g.nb1 = 1;
  auto prod = [](const std::vector<Level> &v) {
    int64_t s = 1;
    for (const Level &L : v) s *= L.dim;
    return s;
  };
  const int64_t Mt = prod(best.m), Nt = prod(best.n), Kt = prod(best.k);
  require(Mt < (1LL << 31) && Nt < (1LL << 31) && Kt < (1LL << 31),
          "contraction too large for one GEMM");
  g.M = (int)Mt;
  g.N = (int)Nt;
  g.K = (int)Kt;
  const auto &m = best.m;
  if (!m.empty()) {
    g.sAm = lstride(A, m.back());
    g.sCm = lstride(C, m.back());
    if (m.size() >= 2) {
      g.M2 = (int)m.back().dim;
      g.M1 = (int)m[m.size() - 2].dim;
      g.sAm1 = lstride(A, m[m.size() - 2]);
      g.sCm1 = lstride(C, m[m.size() - 2]);
    }
    if (m.size() == 3) {
      g.sAm0 = lstride(A, m[0]);
      g.sCm0 = lstride(C, m[0]);
    }
  }
  const auto &n = best.n;
  if (!n.empty()) {
    g.sBn = lstride(B, n.back());
    g.sCn = lstride(C, n.back());
    if (n.size() >= 2) {
      g.N2 = (int)n.back().dim;
      g.N1 = (int)n[n.size() - 2].dim;
      g.sBn1 = lstride(B, n[n.size() - 2]);
      g.sCn1 = lstride(C, n[n.size() - 2]);
    }
    if (n.size() == 3) {
      g.sBn0 = lstride(B, n[0]);
      g.sCn0 = lstride(C, n[0]);
    }
  }
  const auto &k = best.k;
  if (!k.empty()) {
    g.sAk = lstride(A, k.back());
    g.sBk = lstride(B, k.back());
    if (k.size() >= 2) {
      g.K2 = (int)k.back().dim;
      g.K1 = (int)k[k.size() - 2].dim;
      g.sAk1 = lstride(A, k[k.size() - 2]);
      g.sBk1 = lstride(B, k[k.size() - 2]);
    }
    if (k.size() == 3) {
      g.sAk0 = lstride(A, k[0]);
      g.sBk0 = lstride(B, k[0]);
    }
  }
  return true;
}

bool rule_ok(const std::string &rule, const int *sp) {
  if (rule.empty()) return true;
  int left = 0, right = 0;
  bool r = false;
  for (char q : rule) {
    if (q == '|') {
      r = true;
      continue;
    }
    (r ? right : left) += sp[(unsigned char)q];
  }
  return left == right;
}

// Element count of a tensor over the letter extents.
int64_t elems(const Ten &t, const Extents &ext) {
  int64_t s = 1;
  for (char x : t.l) s *= ext(x);
  return s;
}

// Operand element (r, k) of a GEMM whose rows (columns) and K run over up to
// three index levels: index x = (x0 * n1 + x1) * n2 + x2 sits at
// x0 * s0 + x1 * s1 + x2 * s2.
struct PackIdx {
  int n1, n2;
  int64_t s0, s1, s2;
};
__device__ inline int64_t pack_off(int x, const PackIdx &l) {
  const int x2 = x % l.n2, q = x / l.n2;
  return (int64_t)(q / l.n1) * l.s0 + (int64_t)(q % l.n1) * l.s1 +
         (int64_t)x2 * l.s2;
}
// dst[r][k] (row length K) = src(r, k): rows along the thread index when the
// source is contiguous along rows (through a 32x32 LDS tile), else along k.
__global__ __launch_bounds__(256) void pack_kernel(const double *src,
                                                   PackIdx rl, PackIdx kl,
                                                   int R, int K, double *dst,
                                                   int along_r) {
  __shared__ double tile[32][33];
  const int r0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (along_r) {
    for (int i = ty; i < 32; i += 8) {
      const int r = r0 + tx, k = k0 + i;
      tile[i][tx] =
          (r < R && k < K) ? src[pack_off(r, rl) + pack_off(k, kl)] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
      const int r = r0 + i, k = k0 + tx;
      if (r < R && k < K) dst[(int64_t)r * K + k] = tile[tx][i];
    }
  } else {
    for (int i = ty; i < 32; i += 8) {
      const int r = r0 + i, k = k0 + tx;
      if (r < R && k < K)
        dst[(int64_t)r * K + k] = src[pack_off(r, rl) + pack_off(k, kl)];
    }
  }
}

PackIdx pack_levels(int n1, int n2, int64_t s0, int64_t s1, int64_t s) {
  // Gemm convention: n2 == 0 means one level with stride s
  if (n2 == 0) return PackIdx{1, 1 << 30, 0, 0, s};
  return PackIdx{n1, n2, s0, s1, s};
}

// A large block with strided operands runs on contiguous copies ([M][K] and
// [N][K], K fastest): the engine's strided loads of such operands reach a
// third of its rate on the contiguous form.  Returns false when the block
// is left as it is (small, already contiguous, batched or level-mapped).
bool pack_block(ftcc_ctx *c, Gemm &g) {
  const double flops = 2.0 * g.M * (double)g.N * g.K;
  if (flops < 1e9 || g.nb1 != 1 || g.nb2 != 1 || g.mmap || g.nmap)
    return false;
  if (g.sAk == 1 && g.sBk == 1 && g.K2 == 0) return false;
  const int64_t need = ((int64_t)g.M + g.N) * g.K;
  if (need > c->pack_elems) {
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
    dfree(c, c->pack_buf);
    c->pack_buf = nullptr;
    c->pack_elems = 0;
    c->pack_buf = dalloc(c, need);
    c->pack_elems = need;
  }
  double *pa = c->pack_buf, *pb = c->pack_buf + (int64_t)g.M * g.K;
  const PackIdx kA = pack_levels(g.K1, g.K2, g.sAk0, g.sAk1, g.sAk);
  const PackIdx kB = pack_levels(g.K1, g.K2, g.sBk0, g.sBk1, g.sBk);
  const PackIdx rA = pack_levels(g.M1, g.M2, g.sAm0, g.sAm1, g.sAm);
  const PackIdx rB = pack_levels(g.N1, g.N2, g.sBn0, g.sBn1, g.sBn);
  require(g.M < (1 << 30) && g.N < (1 << 30) && g.K < (1 << 30),
          "pack: operand too large");
  const dim3 ga((g.K + 31) / 32, (g.M + 31) / 32), gb((g.K + 31) / 32,
                                                     (g.N + 31) / 32);
  require(ga.y < 65536 && gb.y < 65536, "pack: grid too large");
  hipLaunchKernelGGL(pack_kernel, ga, dim3(256), 0, c->stream, g.A, rA, kA,
                     g.M, g.K, pa, g.sAm == 1 ? 1 : 0);
  hipLaunchKernelGGL(pack_kernel, gb, dim3(256), 0, c->stream, g.B, rB, kB,
                     g.N, g.K, pb, g.sBn == 1 ? 1 : 0);
  FTCC_CHECK_HIP(hipGetLastError());
  // the packed rows keep the levels of C: row m at m * K
  const int64_t K = g.K;
  g.A = pa;
  g.sAm = K;
  g.sAm1 = g.M2 > 0 ? (int64_t)g.M2 * K : 0;
  g.sAm0 = g.M2 > 0 ? (int64_t)g.M1 * g.M2 * K : 0;
  g.B = pb;
  g.sBn = K;
  g.sBn1 = g.N2 > 0 ? (int64_t)g.N2 * K : 0;
  g.sBn0 = g.N2 > 0 ? (int64_t)g.N1 * g.N2 * K : 0;
  g.K1 = 1;
  g.K2 = 0;
  g.sAk = g.sBk = 1;
  g.sAk0 = g.sAk1 = g.sBk0 = g.sBk1 = 0;
  return true;
}

// dst_hi/dst_lo[r][k] (row length ld) = the split of src(r, k) for k < K,
// zero for K <= k < ld (the split-bf16 GEMM reads whole 32-deep slabs)
__global__ __launch_bounds__(256) void split_pack_kernel(
    const double *src, PackIdx rl, PackIdx kl, int R, int K, int ld,
    __bf16 *hi, __bf16 *lo, int along_r) {
  __shared__ double tile[32][33];
  const int r0 = blockIdx.y * 32, k0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  auto put = [&](int r, int k, double v) {
    const float f = (float)v;
    const __bf16 h = (__bf16)f;
    hi[(int64_t)r * ld + k] = h;
    lo[(int64_t)r * ld + k] = (__bf16)(f - (float)h);
  };
  if (along_r) {
    for (int i = ty; i < 32; i += 8) {
      const int r = r0 + tx, k = k0 + i;
      tile[i][tx] =
          (r < R && k < K) ? src[pack_off(r, rl) + pack_off(k, kl)] : 0.0;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
      const int r = r0 + i, k = k0 + tx;
      if (r < R && k < ld) put(r, k, tile[tx][i]);
    }
  } else {
    for (int i = ty; i < 32; i += 8) {
      const int r = r0 + i, k = k0 + tx;
      if (r < R && k < ld)
        put(r, k, k < K ? src[pack_off(r, rl) + pack_off(k, kl)] : 0.0);
    }
  }
}

// out[x] = add + offset of index x over the levels (int: the output spans
// fewer than 2^31 elements)
__global__ void level_offsets_kernel(PackIdx l, int n, int64_t add,
                                     int *out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < n) out[x] = (int)(add + pack_off(x, l));
}

PackIdx c_rows(const Gemm &g) {
  return pack_levels(g.M1, g.M2, g.sCm0, g.sCm1, g.sCm);
}
PackIdx c_cols(const Gemm &g) {
  return pack_levels(g.N1, g.N2, g.sCn0, g.sCn1, g.sCn);
}
// largest offset of an index of n values over the levels
int64_t level_extent(const PackIdx &l, int n) {
  if (l.n2 >= (1 << 30)) return (int64_t)(n - 1) * l.s2;
  const int n0 = n / (l.n1 * l.n2);
  return (int64_t)(n0 - 1) * l.s0 + (int64_t)(l.n1 - 1) * l.s1 +
         (int64_t)(l.n2 - 1) * l.s2;
}

// In the low-precision mode of the lambda iteration, a large block runs as a
// split-bf16 product (lo*hi + hi*lo + hi*hi, FP32 accumulation).  False when
// the block stays FP64 (the mode is off, the block is small, batched, mapped
// or has an addend).
bool lowp_ok(const ftcc_ctx *c, const Gemm &g) {
  if (c->adj_prec != FTCC_PREC_BF16X3) return false;
  const double flops = 2.0 * g.M * (double)g.N * g.K;
  return flops >= c->adj_lowp_min_flops && g.nb1 == 1 && g.nb2 == 1 &&
         !g.mmap && !g.nmap && !g.D && !g.rmap && !g.cmap && g.K >= 1;
}

// doubles of a block's split operands and of its offset tables in the
// low-precision buffer, each rounded to 256 bytes
int64_t lowp_split_doubles(const Gemm &g) {
  const int64_t bf = 2 * ((int64_t)g.M + g.N) * split_ld(g.K);
  return ((bf + 3) / 4 + 31) / 32 * 32;
}
int64_t lowp_table_doubles(const Gemm &g) {
  return (((int64_t)g.M + g.N + 1) / 2 + 31) / 32 * 32;
}

// Up to three such blocks of one contraction (equal alpha and beta, disjoint
// outputs) as one grouped launch, so that a single grid point's 3249^3
// blocks (169 tiles each) fill the device together: both operands of each
// are split into contiguous K-fast bf16 copies, and C is addressed from the
// lowest block's C through per-row and per-column offset tables.  False
// (nothing launched) when the offsets do not fit an int.
bool lowp_group(ftcc_ctx *c, const Gemm *const *gs, int n) {
  const Gemm &g0 = *gs[0];
  double *base = g0.C;
  for (int k = 1; k < n; ++k) base = std::min(base, gs[k]->C);
  int64_t need = 8;
  for (int k = 0; k < n; ++k) {
    const Gemm &g = *gs[k];
    if ((int64_t)(g.C - base) + level_extent(c_rows(g), g.M) +
            level_extent(c_cols(g), g.N) >=
        (1LL << 31))
      return false;
    need += lowp_split_doubles(g) + lowp_table_doubles(g);
  }
  if (need > c->lp_doubles) {
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
    dfree(c, c->lp_buf);
    c->lp_buf = nullptr;
    c->lp_doubles = 0;
    c->lp_buf = dalloc(c, need);
    c->lp_doubles = need;
  }
  double *p = c->lp_buf;
  SbProblem prob[3];
  double flops = 0.0;
  for (int k = 0; k < n; ++k) {
    const Gemm &g = *gs[k];
    const int64_t ld = split_ld(g.K);
    SplitMat A, B;
    A.ld = B.ld = ld;
    A.bstride = B.bstride = 0;
    A.hi = (__bf16 *)p;
    A.lo = A.hi + (int64_t)g.M * ld;
    B.hi = A.lo + (int64_t)g.M * ld;
    B.lo = B.hi + (int64_t)g.N * ld;
    p += lowp_split_doubles(g);
    int *rmap = (int *)p, *cmap = rmap + g.M;
    p += lowp_table_doubles(g);
    const PackIdx kA = pack_levels(g.K1, g.K2, g.sAk0, g.sAk1, g.sAk);
    const PackIdx kB = pack_levels(g.K1, g.K2, g.sBk0, g.sBk1, g.sBk);
    const PackIdx rA = pack_levels(g.M1, g.M2, g.sAm0, g.sAm1, g.sAm);
    const PackIdx rB = pack_levels(g.N1, g.N2, g.sBn0, g.sBn1, g.sBn);
    const dim3 ga((unsigned)((ld + 31) / 32), (g.M + 31) / 32),
        gb((unsigned)((ld + 31) / 32), (g.N + 31) / 32);
    require(ga.y < 65536 && gb.y < 65536, "split pack: grid too large");
    hipLaunchKernelGGL(split_pack_kernel, ga, dim3(256), 0, c->stream, g.A,
                       rA, kA, g.M, g.K, (int)ld, A.hi, A.lo,
                       g.sAm == 1 ? 1 : 0);
    hipLaunchKernelGGL(split_pack_kernel, gb, dim3(256), 0, c->stream, g.B,
                       rB, kB, g.N, g.K, (int)ld, B.hi, B.lo,
                       g.sBn == 1 ? 1 : 0);
    hipLaunchKernelGGL(level_offsets_kernel, dim3((g.M + 255) / 256),
                       dim3(256), 0, c->stream, c_rows(g), g.M,
                       (int64_t)(g.C - base), rmap);
    hipLaunchKernelGGL(level_offsets_kernel, dim3((g.N + 255) / 256),
                       dim3(256), 0, c->stream, c_cols(g), g.N, (int64_t)0,
                       cmap);
    FTCC_CHECK_HIP(hipGetLastError());
    prob[k].M = g.M;
    prob[k].N = g.N;
    prob[k].K = g.K;
    prob[k].A = A;
    prob[k].B = B;
    prob[k].rmap = rmap;
    prob[k].cmap = cmap;
    flops += 2.0 * g.M * (double)g.N * g.K;
  }
  c->cur_M = g0.M;
  c->cur_N = g0.N;
  c->cur_K = g0.K;
  c->cur_nb = n;
  const int tok = prof_begin(c, ftcc_ctx::PROF_PRODUCTS, flops, 0.0);
  sbgemm_grouped(n, prob, 1, g0.alpha, g0.beta, base, 1, 1, 0, c->stream);
  prof_end(c, tok);
  return true;
}

// C[...] = alpha sum_K A[...] B[...] + beta C[...] (beta 0 or 1).  When the
// orbital labels form one alpha run and one beta run in each list, only the
// spin blocks allowed by the three operands' rules are multiplied (a
// forbidden block of A or B is zero, one of C stays zero); otherwise, or
// when a block has no layout the engine takes, the whole product runs
// dense.
void con(ftcc_ctx *c, double alpha, const Ten &A0, const Ten &B0, double beta,
         const Ten &C, int line = __builtin_LINE()) {
  require(beta == 0.0 || beta == 1.0, "contraction: beta must be 0 or 1");
  // per-call-site timing: the GEMMs below belong to the caller's line
  struct Site {
    ftcc_ctx *c;
    int saved, policy;
    ~Site() {
      c->site_override = saved;
      c->tile_policy = policy;
    }
  } site{c, c->site_override, c->tile_policy};
  if (!c->site_override) c->site_override = line;
  // the adjoint's blocks run on narrow tiles: 41 against 24 TFLOP/s on its
  // 3249^3 blocks at n = 114 (three workgroups per CU instead of two)
  c->tile_policy = 2;
  const Groups G = groups_of(A0, B0, C);
  Extents full;
  std::vector<char> letters;
  for (const Ten *t : {&A0, &B0, &C})
    for (char x : t->l)
      if (!full(x)) {
        full.d[(unsigned char)x] = letter_dim(c, x);
        letters.push_back(x);
      }
  if (c->spin && c->spin5 && letters.size() <= 10) {
    struct Block {
      Gemm g;
      int ckey;   // spins of C's letters: blocks of equal key share outputs
    };
    std::vector<Block> blocks;
    bool ok = true;
    const int L = (int)letters.size();
    for (int mask = 0; mask < (1 << L) && ok; ++mask) {
      int sp[128] = {};
      for (int k = 0; k < L; ++k)
        sp[(unsigned char)letters[k]] = (mask >> k) & 1;
      if (!rule_ok(A0.rule, sp) || !rule_ok(B0.rule, sp) ||
          !rule_ok(C.rule, sp))
        continue;
      Extents ext;
      int64_t off[3] = {0, 0, 0};
      bool empty = false;
      for (char x : letters) {
        const bool occ = x >= 'i' && x <= 'n';
        const int s = sp[(unsigned char)x];
        const int64_t start = occ ? c->o_start[s] : c->v_start[s];
        const int64_t cnt = occ ? c->o_count[s] : c->v_count[s];
        ext.d[(unsigned char)x] = cnt;
        if (cnt == 0) empty = true;
        const Ten *ts[3] = {&A0, &B0, &C};
        for (int q = 0; q < 3; ++q)
          if (ts[q]->stride(x) >= 0) off[q] += start * ts[q]->stride(x);
      }
      if (empty) continue;
      Ten a = A0, b = B0, cc = C;
      a.p += off[0];
      b.p += off[1];
      cc.p += off[2];
      Block blk;
      if (!plan(ext, alpha, a, b, 1.0, cc, G, blk.g)) ok = false;
      blk.ckey = 0;
      for (int k = 0; k < L; ++k)
        if (C.stride(letters[k]) >= 0 && sp[(unsigned char)letters[k]])
          blk.ckey |= 1 << k;
      blocks.push_back(blk);
    }
    if (ok) {
      if (beta == 0.0) fill(const_cast<double *>(C.p), elems(C, full), 0.0,
                            c->stream);
      // split-bf16 blocks go out in groups of up to three with disjoint
      // outputs (a group's launch must not add twice into one element)
      std::vector<const Block *> group;
      auto flush = [&] {
        if (group.empty()) return;
        const Gemm *gs[3];
        for (size_t k = 0; k < group.size(); ++k) gs[k] = &group[k]->g;
        if (!lowp_group(c, gs, (int)group.size()))
          for (const Block *b : group) {   // offsets too large: FP64
            Gemm g = b->g;
            pack_block(c, g);
            run(c, g, FTCC_PREC_FP64, false);
          }
        group.clear();
      };
      for (const Block &blk : blocks) {
        if (lowp_ok(c, blk.g)) {
          bool clash = group.size() == 3;
          for (const Block *b : group) clash = clash || b->ckey == blk.ckey;
          if (clash) flush();
          group.push_back(&blk);
          continue;
        }
        Gemm g = blk.g;
        pack_block(c, g);
        run(c, g, FTCC_PREC_FP64, false);
      }
      flush();
      return;
    }
  }
  Gemm g;
  require(plan(full, alpha, A0, B0, beta, C, G, g),
          "contraction " + A0.l + "," + B0.l + "->" + C.l +
              ": no layout the GEMM engine takes");
  pack_block(c, g);
  run(c, g, FTCC_PREC_FP64, false);
}

}  // namespace

// ---------------------------------------------------------------------------
// Workspace of the adjoint: per-grid-point fixed intermediates (a cache of
// as many slices as the budget allows), and the buffers of one application.
// ---------------------------------------------------------------------------
struct AdjSlice {
  double *wmnij = nullptr, *wefab = nullptr, *wmbej = nullptr, *zt = nullptr,
         *zv = nullptr, *f = nullptr;   // f: Fme Fae Fmi Fbe Fmj
};

struct Adjoint {
  const double *T1 = nullptr, *T2 = nullptr;   // the checkpoint
  int ng = 0;
  std::vector<AdjSlice> cache;   // slices [0, cache.size()) are kept
  AdjSlice work;                 // intermediates of an uncached slice
  // application buffers
  double *tau = nullptr, *taut = nullptr, *ar = nullptr, *vT = nullptr,
         *acc = nullptr, *gtau = nullptr, *gt = nullptr, *gw = nullptr,
         *q = nullptr, *u = nullptr, *u2 = nullptr, *small = nullptr;
  double *partials = nullptr;
  // lambda update: per-slice scales and weights (4 * ng on the device, the
  // host copy alive until the stream passes the upload), singles energy
  // derivative base F.ov^T + oovv.t1 per slice
  std::vector<double> h_w;
  double *d_w = nullptr, *e1 = nullptr;
  int w_ng = 0;
  double *d_Gd = nullptr;        // grid derivative weights (ng * ng)
  std::vector<void *> mine;      // allocations of this workspace
};

namespace {

double *aalloc(ftcc_ctx *c, Adjoint *A, int64_t n) {
  double *p = dalloc(c, n);
  A->mine.push_back(p);
  return p;
}

int64_t small_elems(const ftcc_ctx *c) {
  const int64_t O = c->no, V = c->nv;
  return 3 * O * V + 2 * V * V + 2 * O * O;
}

void alloc_slice(ftcc_ctx *c, Adjoint *A, AdjSlice &s) {
  const int64_t O = c->no, V = c->nv;
  s.wmnij = aalloc(c, A, O * O * O * O);
  s.wefab = aalloc(c, A, V * V * V * V);
  s.wmbej = aalloc(c, A, O * V * V * O);
  s.zt = aalloc(c, A, V * V * V * O);
  s.zv = aalloc(c, A, V * O * O * O);
  s.f = aalloc(c, A, small_elems(c));
}

int64_t slice_bytes(const ftcc_ctx *c) {
  const int64_t O = c->no, V = c->nv;
  return 8 * (O * O * O * O + V * V * V * V + O * V * V * O + V * V * V * O +
              V * O * O * O + small_elems(c));
}

// the adjoint's reduction buffers (allocated on first use)
Adjoint *adjoint_ws(ftcc_ctx *c) {
  if (c->adj) return c->adj;
  Adjoint *A = new Adjoint();
  c->adj = A;
  A->partials = aalloc(c, A, 4 * MAX_RED_BLOCKS);
  return A;
}

// the buffers of one application of J^T (ftcc_adjoint_begin)
void ensure_apply_buffers(ftcc_ctx *c, Adjoint *A) {
  if (A->tau) return;
  const int64_t O = c->no, V = c->nv, S2 = V * V * O * O;
  A->tau = aalloc(c, A, S2);
  A->taut = aalloc(c, A, S2);
  A->ar = aalloc(c, A, S2);
  A->vT = aalloc(c, A, S2);
  A->acc = aalloc(c, A, S2);
  A->gtau = aalloc(c, A, S2);
  A->gt = aalloc(c, A, S2);
  A->gw = aalloc(c, A, O * V * O * V);
  A->q = aalloc(c, A, V * O * O * V);
  A->u = aalloc(c, A, O * V * O * O);
  A->u2 = aalloc(c, A, V * V * V * O);
  A->small = aalloc(c, A, 5 * small_elems(c));
  alloc_slice(c, A, A->work);
}

// Fock-type intermediates inside AdjSlice::f
struct Fs {
  double *me, *ae, *mi, *be, *mj;
};
Fs fs_of(const ftcc_ctx *c, double *f) {
  const int64_t O = c->no, V = c->nv;
  Fs r;
  r.me = f;
  r.ae = r.me + O * V;
  r.mi = r.ae + V * V;
  r.be = r.mi + O * O;
  r.mj = r.be + V * V;
  return r;
}

void copy(ftcc_ctx *c, double *dst, const double *src, int64_t n, double a) {
  axpby(dst, src, n, a, 0.0, c->stream);
}

// the fixed (amplitude-only) intermediates of one grid point
void prepare_slice(ftcc_ctx *c, Adjoint *A, const double *t1,
                   const double *t2, AdjSlice &s) {
  const int O = c->no, V = c->nv;
  const int64_t S2 = (int64_t)V * V * O * O;
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(tau_kernel, dim3(grid_for(S2)), dim3(AT), 0, st, t1, t2,
                     V, O, A->tau, A->taut, A->ar);
  FTCC_CHECK_HIP(hipGetLastError());
  double *const *I = c->I;   // vvvv vvvo vovv vvoo vovo oovv vooo ooov oooo
  const Fs F = fs_of(c, s.f);
  // Fme = F.ov + t1.oovv
  copy(c, F.me, c->F[1], (int64_t)O * V, 1.0);
  con(c, 1.0, ten(c, I[5], "mnef"), ten(c, t1, "fn"), 1.0, ten(c, F.me, "me"));
  // Fae = F.vv - 1/2 t1.F.ov - t1.vovv - 1/2 taut.oovv
  copy(c, F.ae, c->F[3], (int64_t)V * V, 1.0);
  con(c, -0.5, ten(c, t1, "am"), ten(c, c->F[1], "me"), 1.0,
      ten(c, F.ae, "ae"));
  con(c, -1.0, ten(c, I[2], "amfe"), ten(c, t1, "fm"), 1.0,
      ten(c, F.ae, "ae"));
  con(c, -0.5, ten(c, A->taut, "afmn"), ten(c, I[5], "mnef"), 1.0,
      ten(c, F.ae, "ae"));
  // Fmi = F.oo + 1/2 F.ov.t1 + t1.ooov + 1/2 oovv.taut
  copy(c, F.mi, c->F[0], (int64_t)O * O, 1.0);
  con(c, 0.5, ten(c, c->F[1], "me"), ten(c, t1, "ei"), 1.0,
      ten(c, F.mi, "mi"));
  con(c, 1.0, ten(c, I[7], "mnie"), ten(c, t1, "en"), 1.0,
      ten(c, F.mi, "mi"));
  con(c, 0.5, ten(c, I[5], "mnef"), ten(c, A->taut, "efin"), 1.0,
      ten(c, F.mi, "mi"));
  // Fbe = Fae - 1/2 t1.Fme, Fmj = Fmi + 1/2 Fme.t1
  copy(c, F.be, F.ae, (int64_t)V * V, 1.0);
  con(c, -0.5, ten(c, t1, "bm"), ten(c, F.me, "me"), 1.0, ten(c, F.be, "be"));
  copy(c, F.mj, F.mi, (int64_t)O * O, 1.0);
  con(c, 0.5, ten(c, F.me, "me"), ten(c, t1, "ej"), 1.0, ten(c, F.mj, "mj"));
  // Wh_mnij = 1/2 (oooo + P(ij) t1.ooov) + 1/4 oovv.tau
  copy(c, s.wmnij, I[8], (int64_t)O * O * O * O, 0.5);
  con(c, 0.5, ten(c, I[7], "mnie"), ten(c, t1, "ej"), 1.0,
      ten(c, s.wmnij, "mnij"));
  con(c, -0.5, ten(c, I[7], "mnje"), ten(c, t1, "ei"), 1.0,
      ten(c, s.wmnij, "mnij"));
  con(c, 0.25, ten(c, I[5], "mnef"), ten(c, A->tau, "efij"), 1.0,
      ten(c, s.wmnij, "mnij"));
  // Wh_efab = 1/2 (vvvv - P(ab) t1.vovv) + 1/4 tau.oovv, stored [e,f,a,b]
  // vvvv[a,b,e,f] / 2 into [e,f,a,b]
  permute4(I[0], s.wefab, V, V, V, V, (int64_t)V, 1, (int64_t)V * V * V,
           (int64_t)V * V, 0.5, 0.0, 1, 0, 0, st);
  con(c, -0.5, ten(c, I[2], "amef"), ten(c, t1, "bm"), 1.0,
      ten(c, s.wefab, "efab"));
  con(c, 0.5, ten(c, I[2], "bmef"), ten(c, t1, "am"), 1.0,
      ten(c, s.wefab, "efab"));
  con(c, 0.25, ten(c, A->tau, "abmn"), ten(c, I[5], "mnef"), 1.0,
      ten(c, s.wefab, "efab"));
  // Wmbej = <mb||ej> - t1.vovv + t1.ooov - oovv.(t2/2 + t1 t1)
  // <mb||ej> = -vovo[b,m,e,j] into [m,b,e,j]
  permute4(I[4], s.wmbej, V, O, V, O, (int64_t)V * O, (int64_t)V * V * O,
           (int64_t)O, 1, -1.0, 0.0, 1, 0, 0, st);
  con(c, -1.0, ten(c, I[2], "bmef"), ten(c, t1, "fj"), 1.0,
      ten(c, s.wmbej, "mbej"));
  con(c, 1.0, ten(c, I[7], "mnje"), ten(c, t1, "bn"), 1.0,
      ten(c, s.wmbej, "mbej"));
  con(c, -1.0, ten(c, I[5], "mnef"), ten(c, A->ar, "fbjn"), 1.0,
      ten(c, s.wmbej, "mbej"));
  // Zt[e,a,b,i] = tau.ooov, Zv[a,m,i,j] = vovv.tau
  con(c, 1.0, ten(c, A->tau, "abmn"), ten(c, I[7], "mnie"), 0.0,
      ten(c, s.zt, "eabi", "ab|ie"));
  con(c, 1.0, ten(c, I[2], "amef"), ten(c, A->tau, "efij"), 0.0,
      ten(c, s.zv, "amij"));
}

// (g1, g2) = J^T (v1, v2) at one grid point (v2 antisymmetric); g2 is
// projected onto antisymmetric tensors
void apply_slice(ftcc_ctx *c, Adjoint *A, const double *t1, const double *t2,
                 const AdjSlice &s, const double *v1, const double *v2,
                 double *g1, double *g2) {
  const int O = c->no, V = c->nv;
  const int64_t S2 = (int64_t)V * V * O * O;
  hipStream_t st = c->stream;
  double *const *I = c->I;
  const Fs F = fs_of(c, s.f);
  const int64_t nsm = small_elems(c);
  double *gFmj = A->small, *gFbe = gFmj + nsm, *gFae = gFbe + nsm,
         *gFmi = gFae + nsm, *gFme = gFmi + nsm;
  // v2 transposed [i,j,a,b]
  permute4(v2, A->vT, V, V, O, O, (int64_t)V, 1, (int64_t)O * V * V,
           (int64_t)V * V, 1.0, 0.0, 1, 0, 0, st);
  fill(g1, (int64_t)V * O, 0.0, st);
  const Ten V2 = ten(c, v2, "abij"), VT = ten(c, A->vT, "ijab");
  // doubles terms of the residual, transposed (antisymmetrisers give 2, 4)
  con(c, 2.0, V2, ten(c, I[6], "bmij"), 1.0, ten(c, g1, "am"));      // vooo
  con(c, 2.0, ten(c, I[1], "abej"), V2, 1.0, ten(c, g1, "ei"));      // vvvo
  con(c, 4.0, V2, ten(c, s.wmbej, "mbej"), 0.0, ten(c, A->acc, "aeim"));
  con(c, 4.0, ten(c, t2, "aeim"), V2, 0.0, ten(c, A->gw, "jbme", "mb|ej"));
  // -t1 t1 <mb||ej> term: <mb||ej> = -vovo[b,m,e,j]
  con(c, 1.0, ten(c, t1, "am"), V2, 0.0, ten(c, A->u, "mbij"));
  con(c, 4.0, ten(c, I[4], "bmej"), ten(c, A->u, "mbij"), 1.0,
      ten(c, g1, "ei"));
  con(c, 1.0, V2, ten(c, t1, "ei"), 0.0, ten(c, A->u2, "abej"));
  con(c, 4.0, ten(c, A->u2, "abej"), ten(c, I[4], "bmej"), 1.0,
      ten(c, g1, "am"));
  // ladders (tau adjoint), folded with the W_mnij / W_abef tau terms
  con(c, 1.0, ten(c, s.wefab, "efab"), VT, 0.0, ten(c, A->gtau, "efij"));
  con(c, 1.0, V2, ten(c, s.wmnij, "mnij"), 1.0, ten(c, A->gtau, "abmn"));
  // t1 parts of W_mnij and W_abef through Zt, Zv
  con(c, 1.0, ten(c, s.zt, "eabi", "ab|ie"), VT, 1.0, ten(c, g1, "ej"));
  con(c, -1.0, V2, ten(c, s.zv, "amij"), 1.0, ten(c, g1, "bm"));
  // Fock-type terms of the doubles
  con(c, -2.0, V2, ten(c, F.mj, "mj"), 1.0, ten(c, A->acc, "abim"));
  con(c, -2.0, ten(c, t2, "abim"), V2, 0.0, ten(c, gFmj, "mj"));
  con(c, 2.0, V2, ten(c, F.be, "be"), 1.0, ten(c, A->acc, "aeij"));
  con(c, 2.0, ten(c, t2, "aeij"), V2, 0.0, ten(c, gFbe, "be"));
  // Fmj = Fmi + 1/2 t1.Fme ; Fbe = Fae - 1/2 t1.Fme
  copy(c, gFmi, gFmj, (int64_t)O * O, 1.0);
  con(c, 0.5, ten(c, gFmj, "mj"), ten(c, F.me, "me"), 1.0, ten(c, g1, "ej"));
  con(c, 0.5, ten(c, gFmj, "mj"), ten(c, t1, "ej"), 0.0, ten(c, gFme, "me"));
  copy(c, gFae, gFbe, (int64_t)V * V, 1.0);
  con(c, -0.5, ten(c, gFbe, "be"), ten(c, F.me, "me"), 1.0,
      ten(c, g1, "bm"));
  con(c, -0.5, ten(c, gFbe, "be"), ten(c, t1, "bm"), 1.0,
      ten(c, gFme, "me"));
  // singles terms of the residual
  con(c, 1.0, ten(c, v1, "ai"), ten(c, t1, "ei"), 1.0, ten(c, gFae, "ae"));
  con(c, 1.0, ten(c, F.ae, "ae"), ten(c, v1, "ai"), 1.0, ten(c, g1, "ei"));
  con(c, -1.0, ten(c, t1, "am"), ten(c, v1, "ai"), 1.0, ten(c, gFmi, "mi"));
  con(c, -1.0, ten(c, v1, "ai"), ten(c, F.mi, "mi"), 1.0, ten(c, g1, "am"));
  con(c, 1.0, ten(c, v1, "ai"), ten(c, F.me, "me"), 1.0,
      ten(c, A->acc, "aeim"));
  con(c, 1.0, ten(c, v1, "ai"), ten(c, t2, "aeim"), 1.0,
      ten(c, gFme, "me"));
  con(c, -1.0, ten(c, v1, "ai"), ten(c, I[4], "anfi"), 1.0,
      ten(c, g1, "fn"));
  con(c, 0.5, ten(c, v1, "ai"), ten(c, I[2], "amef"), 1.0,
      ten(c, A->acc, "efim"));
  con(c, 0.5, ten(c, v1, "ai"), ten(c, I[7], "nmie"), 1.0,
      ten(c, A->acc, "aemn"));
  // Wmbej: t1.vovv, t1.ooov and the ring amplitude t2/2 + t1 t1
  con(c, -1.0, ten(c, A->gw, "jbme", "mb|ej"), ten(c, I[2], "bmef"), 1.0,
      ten(c, g1, "fj"));
  con(c, 1.0, ten(c, A->gw, "jbme", "mb|ej"), ten(c, I[7], "mnje"), 1.0,
      ten(c, g1, "bn"));
  con(c, 1.0, ten(c, A->gw, "jbme", "mb|ej"), ten(c, I[5], "mnef"), 0.0,
      ten(c, A->q, "bjnf", "bf|jn"));
  // acc[f,b,j,n] -= q[b,j,n,f] / 2
  permute4(A->q, A->acc, V, O, O, V, (int64_t)O * O, (int64_t)O, 1,
           (int64_t)V * O * O, -0.5, 1.0, 1, 0, 0, st);
  con(c, -1.0, ten(c, A->q, "bjnf", "bf|jn"), ten(c, t1, "bn"), 1.0,
      ten(c, g1, "fj"));
  con(c, -1.0, ten(c, A->q, "bjnf", "bf|jn"), ten(c, t1, "fj"), 1.0,
      ten(c, g1, "bn"));
  // Fmi, Fae, Fme
  con(c, 0.5, ten(c, c->F[1], "me"), ten(c, gFmi, "mi"), 1.0,
      ten(c, g1, "ei"));
  con(c, 1.0, ten(c, I[7], "mnie"), ten(c, gFmi, "mi"), 1.0,
      ten(c, g1, "en"));
  con(c, 0.5, ten(c, gFmi, "mi"), ten(c, I[5], "mnef"), 0.0,
      ten(c, A->gt, "efin"));
  con(c, -0.5, ten(c, gFae, "ae"), ten(c, c->F[1], "me"), 1.0,
      ten(c, g1, "am"));
  con(c, -1.0, ten(c, gFae, "ae"), ten(c, I[2], "amfe"), 1.0,
      ten(c, g1, "fm"));
  con(c, -0.5, ten(c, gFae, "ae"), ten(c, I[5], "mnef"), 1.0,
      ten(c, A->gt, "afmn"));
  con(c, 1.0, ten(c, gFme, "me"), ten(c, I[5], "mnef"), 1.0,
      ten(c, g1, "fn"));
  // projection and the tau / taut chain rule into t1
  hipLaunchKernelGGL(assemble_kernel, dim3(grid_for(S2)), dim3(AT), 0, st,
                     A->acc, A->gtau, A->gt, V, O, A->tau, g2);
  FTCC_CHECK_HIP(hipGetLastError());
  con(c, 4.0, ten(c, A->tau, "abij"), ten(c, t1, "bj"), 1.0,
      ten(c, g1, "ai"));
}

}  // namespace

// per-slice factors of the lambda update: [0,ng) g/beta, [ng,2ng)
// g/(4 beta), [2ng,3ng) beta/g, [3ng,4ng) 4 beta/g, and the singles base
// F.ov^T + oovv.t1 of every slice
void lambda_factors(ftcc_ctx *c, Adjoint *A, const double *T1,
                           const double *g_host, double beta) {
  const int O = c->no, V = c->nv, ng = c->ng;
  const int64_t VO = (int64_t)V * O;
  if (A->w_ng != ng) {
    for (double **p : {&A->d_w, &A->e1}) {
      if (*p) {
        dfree(c, *p);
        A->mine.erase(std::remove(A->mine.begin(), A->mine.end(),
                                  (void *)*p),
                      A->mine.end());
        *p = nullptr;
      }
    }
    A->d_w = aalloc(c, A, 4 * (int64_t)ng);
    A->e1 = aalloc(c, A, ng * VO);
    A->w_ng = ng;
  }
  A->h_w.assign(4 * (size_t)ng, 0.0);
  for (int y = 0; y < ng; ++y) {
    require(g_host[y] != 0.0, "zero quadrature weight");
    A->h_w[y] = g_host[y] / beta;
    A->h_w[ng + y] = g_host[y] / (4.0 * beta);
    A->h_w[2 * ng + y] = beta / g_host[y];
    A->h_w[3 * ng + y] = 4.0 * beta / g_host[y];
  }
  FTCC_CHECK_HIP(hipMemcpyAsync(A->d_w, A->h_w.data(),
                                sizeof(double) * 4 * ng,
                                hipMemcpyHostToDevice, c->stream));
  for (int y = 0; y < ng; ++y) {
    copy(c, A->e1 + y * VO, c->fovT, VO, 1.0);
    con(c, 1.0, ten(c, c->I[5], "ijab"), ten(c, T1 + y * VO, "bj"), 1.0,
        ten(c, A->e1 + y * VO, "ai"));
  }
}

void sum4(ftcc_ctx *c, Adjoint *A, int nblocks, double *dst) {
  hipLaunchKernelGGL(sum4_kernel, dim3(1), dim3(AT), 0, c->stream,
                     A->partials, nblocks, dst);
  FTCC_CHECK_HIP(hipGetLastError());
}

void adjoint_release(ftcc_ctx *c) {
  if (!c || !c->adj) return;
  Adjoint *A = c->adj;
  (void)hipStreamSynchronize(c->stream);
  for (void *p : A->mine) {
    try {
      dfree(c, p);
    } catch (...) {
    }
  }
  delete A;
  c->adj = nullptr;
}

}  // namespace ftcc_core

using namespace ftcc_core;

namespace {

// pia[i][a] = s_vo[a][i]; pba[b][a] = s_vv[a][b]; pji[j][i] = s_oo[i][j];
// pai[a][i] = s_ov[i][a] + t1sum[a][i]
__global__ void density1_out_kernel(const double *s_vo, const double *s_vv,
                                    const double *s_oo, const double *s_ov,
                                    const double *t1sum, int V, int O,
                                    double *pia, double *pba, double *pji,
                                    double *pai) {
  const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V,
                OO = (int64_t)O * O;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
       k < 2 * VO + VV + OO; k += (int64_t)gridDim.x * blockDim.x) {
    if (k < VO) {                       // pia[i][a]
      const int i = (int)(k / V), a = (int)(k % V);
      pia[k] = s_vo[(int64_t)a * O + i];
    } else if (k < VO + VV) {           // pba[b][a]
      const int64_t q = k - VO;
      const int b = (int)(q / V), a = (int)(q % V);
      pba[q] = s_vv[(int64_t)a * V + b];
    } else if (k < VO + VV + OO) {      // pji[j][i]
      const int64_t q = k - VO - VV;
      const int j = (int)(q / O), i = (int)(q % O);
      pji[q] = s_oo[(int64_t)i * O + j];
    } else {                            // pai[a][i]
      const int64_t q = k - VO - VV - OO;
      const int a = (int)(q / O), i = (int)(q % O);
      pai[q] = s_ov[(int64_t)i * V + a] + t1sum[q];
    }
  }
}

}  // namespace

extern "C" {

int ftcc_adjoint_begin(ftcc_ctx *c, const double *T1, const double *T2) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && T2, "null amplitude pointer");
    drop_graph(c);
    // the forward residual's operand buffers are rebuilt when next needed
    free_workspace(c);
    adjoint_release(c);
    Adjoint *A = adjoint_ws(c);
    ensure_apply_buffers(c, A);
    A->T1 = T1;
    A->T2 = T2;
    A->ng = c->ng;
    double budget_gb = 64.0;
    if (const char *env = getenv("FTCC_ADJOINT_CACHE_GB"))
      budget_gb = atof(env);
    const int64_t keep = std::min<int64_t>(
        c->ng, (int64_t)(budget_gb * (1LL << 30)) / slice_bytes(c));
    const int64_t VO = (int64_t)c->nv * c->no;
    A->cache.resize((size_t)std::max<int64_t>(0, keep));
    for (int y = 0; y < (int)A->cache.size(); ++y) {
      alloc_slice(c, A, A->cache[y]);
      prepare_slice(c, A, T1 + y * VO, T2 + y * VO * VO, A->cache[y]);
    }
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int ftcc_adjoint_cached(ftcc_ctx *c, int *nslices) {
  return guarded(c, [&] {
    require(nslices != nullptr, "null output pointer");
    *nslices = c->adj ? (int)c->adj->cache.size() : 0;
  });
}

int ftcc_adjoint_apply(ftcc_ctx *c, const double *Lt1, const double *Lt2,
                       double *G1, double *G2) {
  return guarded(c, [&] {
    Adjoint *A = c->adj;
    require(A && A->T1, "ftcc_adjoint_begin has not been called",
            FTCC_ESTATE);
    require(A->ng == c->ng, "the grid changed since ftcc_adjoint_begin",
            FTCC_ESTATE);
    require(Lt1 && Lt2 && G1 && G2, "null pointer");
    const int64_t VO = (int64_t)c->nv * c->no, S2 = VO * VO;
    for (int y = 0; y < c->ng; ++y) {
      const double *t1 = A->T1 + y * VO, *t2 = A->T2 + y * S2;
      const AdjSlice *s = &A->work;
      if (y < (int)A->cache.size())
        s = &A->cache[y];
      else
        prepare_slice(c, A, t1, t2, A->work);
      apply_slice(c, A, t1, t2, *s, Lt1 + y * VO, Lt2 + y * S2, G1 + y * VO,
                  G2 + y * S2);
    }
  });
}

int ftcc_adjoint_precision(ftcc_ctx *c, int precision) {
  return guarded(c, [&] {
    require(precision == FTCC_PREC_FP64 || precision == FTCC_PREC_BF16X3,
            "the adjoint's large contractions run in FP64 or bf16x3");
    c->adj_prec = precision;
    // blocks of at least this many flops take the split-bf16 product
    // (FTCC_ADJOINT_LOWP_MIN_GFLOP, default 10)
    c->adj_lowp_min_flops = 1e10;
    if (const char *env = getenv("FTCC_ADJOINT_LOWP_MIN_GFLOP"))
      c->adj_lowp_min_flops = atof(env) * 1e9;
  });
}

int ftcc_adjoint_end(ftcc_ctx *c) {
  return guarded(c, [&] { adjoint_release(c); });
}

int ftcc_integrate_adjoint(ftcc_ctx *c, const double *L1, const double *L2,
                           double *out1, double *out2) {
  return guarded(c, [&] {
    check_grid_D(c);
    require(L1 && L2 && out1 && out2 && L1 != out1 && L2 != out2,
            "integrate_adjoint needs distinct input and output arrays");
    const int O = c->no, V = c->nv, ng = c->ng;
    const Grid &g = c->grid;
    const size_t lds = ng <= MAX_STAGE_NG ? sizeof(double) * ng * AT : 0;
    const int64_t VO = (int64_t)V * O;
    hipLaunchKernelGGL(int_adjoint_kernel<false>, dim3(grid_for(VO)),
                       dim3(AT), lds, c->stream, L1, out1, ng, V, O, g.d_ti,
                       g.d_G, c->ev, c->eo, c->have_D ? c->D1 : nullptr,
                       g.uniform ? 1 : 0, g.delta);
    FTCC_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(int_adjoint_kernel<true>, dim3(grid_for(VO * VO)),
                       dim3(AT), lds, c->stream, L2, out2, ng, V, O, g.d_ti,
                       g.d_G, c->ev, c->eo, c->have_D ? c->D2 : nullptr,
                       g.uniform ? 1 : 0, g.delta);
    FTCC_CHECK_HIP(hipGetLastError());
  });
}

// The unrelaxed one-particle response densities of kelvin's
// ft_cc_equations.ccsd_1rdm (kelvin/ft_cc_equations.py:704-722): with
// l = int_L(L) (the reverse-integrated lambda amplitudes) at every grid
// point x and weights g_x,
//   pia = sum_x g_x l1,
//   pba[b,a] = sum_x g_x (sum_i l1[i,a] t1[b,i]
//                         + 1/2 sum_cij l2[i,j,c,a] t2[c,b,i,j]),
//   pji[j,i] = sum_x g_x (-sum_a l1[j,a] t1[a,i]
//                         - 1/2 sum_kab l2[k,j,a,b] t2[a,b,k,i]),
//   pai[a,i] = sum_x g_x (t1[a,i] - sum_jb l1[j,b] t1[b,i] t1[a,j]
//                + sum_jb l1[j,b] t2[b,a,j,i]
//                - 1/2 sum l2[j,k,b,c] t2[b,a,j,k] t1[c,i]
//                - 1/2 sum l2[j,k,b,c] t2[b,c,j,i] t1[a,k]),
// the derivatives of the Lagrangian E - (1/beta) sum_x g_x <L_x, T_new,x>
// with respect to the Fock blocks (the definition kelvin's tests check,
// kelvin/tests/test_ft_ccsd_rdm.py:12-64).  The engine's integrated adjoint
// amplitudes u = integrate_adjoint(Lam) carry the weights: g_x l1_x =
// -beta u1_x and g_x l2_x = -4 beta u2_x (transposed), so the terms are
// accumulated from u with those factors.  T and u in the engine layout
// ([ng][V][O], [ng][V][V][O][O]); outputs in kelvin's (pia [O][V], pba
// [V][V], pji [O][O], pai [V][O]), device memory.
int ftcc_density1(ftcc_ctx *c, const double *T1, const double *T2,
                  const double *Lt1, const double *Lt2, const double *g_host,
                  double beta, double *pia, double *pba, double *pji,
                  double *pai) {
  return guarded(c, [&] {
    require(c->have_ints, "integrals not set", FTCC_ESTATE);
    require(T1 && T2 && Lt1 && Lt2 && g_host && pia && pba && pji && pai,
            "null pointer");
    require(beta > 0.0, "beta must be positive");
    const int O = c->no, V = c->nv, ng = c->ng;
    const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V,
                  OO = (int64_t)O * O, S2 = VV * OO;
    hipStream_t s = c->stream;
    // accumulators and per-point intermediates
    const int64_t tot = VO + VV + OO + VO + VO + OO + VV + OO;
    double *buf = dalloc(c, tot);
    struct Free {
      ftcc_ctx *c;
      double *p;
      ~Free() {
        (void)hipStreamSynchronize(c->stream);
        try {
          dfree(c, p);
        } catch (...) {
        }
      }
    } fr{c, buf};
    double *s_vo = buf, *s_vv = s_vo + VO, *s_oo = s_vv + VV,
           *s_ov = s_oo + OO, *t1sum = s_ov + VO, *x1 = t1sum + VO,
           *y2 = x1 + OO, *z2 = y2 + VV;
    fill(buf, VO + VV + OO + VO + VO, 0.0, s);
    for (int x = 0; x < ng; ++x) {
      const double *t1 = T1 + x * VO, *t2 = T2 + x * S2;
      const double *u1 = Lt1 + x * VO, *u2 = Lt2 + x * S2;
      axpby(s_vo, u1, VO, -beta, 1.0, s);          // pia
      axpby(t1sum, t1, VO, g_host[x], 1.0, s);     // energy term of pai
      // y2[b,e] = sum u2[a,b,i,j] t2[a,e,i,j];  z2[m,j] = sum u2[a,b,i,j]
      // t2[a,b,i,m];  x1[i,m] = sum_a u1[a,i] t1[a,m]
      con(c, 1.0, ten(c, u2, "abij"), ten(c, t2, "aeij"), 0.0,
          ten(c, y2, "be"));
      con(c, 1.0, ten(c, u2, "abij"), ten(c, t2, "abim"), 0.0,
          ten(c, z2, "mj"));
      con(c, 1.0, ten(c, u1, "ai"), ten(c, t1, "am"), 0.0, ten(c, x1, "im"));
      // pba: s_vv[a,e] += -beta sum_i u1[a,i] t1[e,i] - 2 beta y2[a,e]
      con(c, -beta, ten(c, u1, "ai"), ten(c, t1, "ei"), 1.0,
          ten(c, s_vv, "ae"));
      axpby(s_vv, y2, VV, -2.0 * beta, 1.0, s);
      // pji: s_oo[m,j] += beta sum_a u1[a,j] t1[a,m] + 2 beta z2[m,j]
      con(c, beta, ten(c, u1, "aj"), ten(c, t1, "am"), 1.0,
          ten(c, s_oo, "mj"));
      axpby(s_oo, z2, OO, 2.0 * beta, 1.0, s);
      // pai: s_ov[m,e] += beta sum_i t1[e,i] x1[i,m]
      //   - beta sum u1[a,i] t2[a,e,i,m] + 2 beta sum_b y2[b,e] t1[b,m]
      //   + 2 beta sum_j z2[m,j] t1[e,j]
      con(c, beta, ten(c, t1, "ei"), ten(c, x1, "im"), 1.0,
          ten(c, s_ov, "me"));
      con(c, -beta, ten(c, u1, "ai"), ten(c, t2, "aeim"), 1.0,
          ten(c, s_ov, "me"));
      con(c, 2.0 * beta, ten(c, y2, "be"), ten(c, t1, "bm"), 1.0,
          ten(c, s_ov, "me"));
      con(c, 2.0 * beta, ten(c, z2, "mj"), ten(c, t1, "ej"), 1.0,
          ten(c, s_ov, "me"));
    }
    const int64_t n = 2 * VO + VV + OO;
    hipLaunchKernelGGL(density1_out_kernel,
                       dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)),
                       dim3(256), 0, s, s_vo, s_vv, s_oo, s_ov, t1sum, V, O,
                       pia, pba, pji, pai);
    FTCC_CHECK_HIP(hipGetLastError());
  });
}

int ftcc_lambda_update(ftcc_ctx *c, double *Lam1, double *Lam2,
                       const double *G1, const double *G2, const double *T1,
                       const double *g_host, double beta, double damp,
                       double out4[4]) {
  return guarded(c, [&] {
    check_ready(c);
    require(Lam1 && Lam2 && G1 && G2 && T1 && g_host && out4,
            "null pointer");
    require(beta > 0.0, "beta must be positive");
    Adjoint *A = adjoint_ws(c);
    const int ng = c->ng;
    const int64_t VO = (int64_t)c->nv * c->no;
    hipStream_t s = c->stream;
    // dE/dT: (g_y / beta) (F.ov^T + oovv.t1) for the singles and
    // (g_y / (4 beta)) oovv^T for the doubles; kelvin's lambda scale is
    // beta / g_y (singles) and 4 beta / g_y (doubles) times these
    lambda_factors(c, A, T1, g_host, beta);
    const int nb1 = grid_for(ng * VO, MAX_RED_BLOCKS);
    hipLaunchKernelGGL(lambda_update_kernel, dim3(nb1), dim3(AT), 0, s, Lam1,
                       G1, A->e1, A->d_w, A->d_w + 2 * ng, ng, VO, VO, damp,
                       A->partials);
    FTCC_CHECK_HIP(hipGetLastError());
    sum4(c, A, nb1, c->red);
    const int nb2 = grid_for(ng * VO * VO, MAX_RED_BLOCKS);
    hipLaunchKernelGGL(lambda_update_kernel, dim3(nb2), dim3(AT), 0, s, Lam2,
                       G2, c->oovvT, A->d_w + ng, A->d_w + 3 * ng, ng, VO * VO,
                       (int64_t)0, damp, A->partials);
    FTCC_CHECK_HIP(hipGetLastError());
    sum4(c, A, nb2, c->red + 4);
    read_scalars(c, 8);
    out4[0] = c->host_red[0];
    out4[1] = c->host_red[4];
    out4[2] = c->host_red[1];
    out4[3] = c->host_red[5];
  });
}

int ftcc_lambda_guess(ftcc_ctx *c, const double *T1, const double *g_host,
                      double beta, double *Lam1, double *Lam2) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && g_host && Lam1 && Lam2, "null pointer");
    require(beta > 0.0, "beta must be positive");
    Adjoint *A = adjoint_ws(c);
    lambda_factors(c, A, T1, g_host, beta);
    const int ng = c->ng;
    const int64_t VO = (int64_t)c->nv * c->no;
    // kelvin's guess (1/beta) dE_y/dT_y in its scale is -(1/beta) dE/dT here
    for (size_t k = 0; k < 2 * (size_t)ng; ++k) A->h_w[k] *= -1.0 / beta;
    FTCC_CHECK_HIP(hipMemcpyAsync(A->d_w, A->h_w.data(),
                                  sizeof(double) * 2 * ng,
                                  hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(scale_slices_kernel, dim3(grid_for(ng * VO)), dim3(AT),
                       0, c->stream, Lam1, A->e1, A->d_w, ng, VO, VO);
    FTCC_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(scale_slices_kernel, dim3(grid_for(ng * VO * VO)),
                       dim3(AT), 0, c->stream, Lam2, c->oovvT, A->d_w + ng,
                       ng, VO * VO, (int64_t)0);
    FTCC_CHECK_HIP(hipGetLastError());
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int ftcc_lambda_convert(ftcc_ctx *c, int to_kelvin, const double *g_host,
                        double beta, const double *src1, const double *src2,
                        double *dst1, double *dst2) {
  return guarded(c, [&] {
    require(g_host && src1 && src2 && dst1 && dst2 && src1 != dst1 &&
                src2 != dst2,
            "lambda conversion needs distinct arrays");
    require(beta > 0.0, "beta must be positive");
    const int O = c->no, V = c->nv, ng = c->ng;
    const int64_t VO = (int64_t)V * O, S2 = VO * VO;
    hipStream_t s = c->stream;
    for (int y = 0; y < ng; ++y) {
      require(g_host[y] != 0.0, "zero quadrature weight");
      const double f1 = to_kelvin ? -beta / g_host[y] : -g_host[y] / beta;
      const double f2 =
          to_kelvin ? -4.0 * beta / g_host[y] : -g_host[y] / (4.0 * beta);
      if (to_kelvin) {
        // L1[i,a] = f1 Lam1[a,i]; L2[i,j,a,b] = f2 Lam2[a,b,i,j]
        permute4(src1 + y * VO, dst1 + y * VO, V, O, 1, 1, 1, V, 0, 0, f1,
                 0.0, 1, 0, 0, s);
        permute4(src2 + y * S2, dst2 + y * S2, V, V, O, O, (int64_t)V, 1,
                 (int64_t)O * V * V, (int64_t)V * V, f2, 0.0, 1, 0, 0, s);
      } else {
        permute4(src1 + y * VO, dst1 + y * VO, O, V, 1, 1, 1, O, 0, 0, f1,
                 0.0, 1, 0, 0, s);
        permute4(src2 + y * S2, dst2 + y * S2, O, O, V, V, (int64_t)O, 1,
                 (int64_t)V * O * O, (int64_t)O * O, f2, 0.0, 1, 0, 0, s);
      }
    }
  });
}

int ftcc_grid_derivative(ftcc_ctx *c, const double *Lam1, const double *Lam2,
                         const double *T1b, const double *T2b,
                         const double *Gd_host, double beta, double *out) {
  return guarded(c, [&] {
    check_grid_D(c);
    require(Lam1 && Lam2 && T1b && T2b && Gd_host && out, "null pointer");
    require(beta > 0.0, "beta must be positive");
    Adjoint *A = adjoint_ws(c);
    const int O = c->no, V = c->nv, ng = c->ng;
    hipStream_t s = c->stream;
    if (A->d_Gd) {
      dfree(c, A->d_Gd);
      A->mine.erase(std::remove(A->mine.begin(), A->mine.end(),
                                (void *)A->d_Gd),
                    A->mine.end());
    }
    A->d_Gd = aalloc(c, A, (int64_t)ng * ng);
    FTCC_CHECK_HIP(hipMemcpyAsync(A->d_Gd, Gd_host, sizeof(double) * ng * ng,
                                  hipMemcpyHostToDevice, s));
    const Grid &g = c->grid;
    const int64_t VO = (int64_t)V * O;
    const int nb1 = grid_for(VO, MAX_RED_BLOCKS);
    hipLaunchKernelGGL(grid_deriv_kernel<false>, dim3(nb1), dim3(AT), 0, s,
                       Lam1, T1b, ng, V, O, g.d_ti, g.d_G, A->d_Gd, c->ev,
                       c->eo, c->have_D ? c->D1 : nullptr, beta, A->partials);
    FTCC_CHECK_HIP(hipGetLastError());
    sum4(c, A, nb1, c->red);
    const int nb2 = grid_for(VO * VO, MAX_RED_BLOCKS);
    hipLaunchKernelGGL(grid_deriv_kernel<true>, dim3(nb2), dim3(AT), 0, s,
                       Lam2, T2b, ng, V, O, g.d_ti, g.d_G, A->d_Gd, c->ev,
                       c->eo, c->have_D ? c->D2 : nullptr, beta, A->partials);
    FTCC_CHECK_HIP(hipGetLastError());
    sum4(c, A, nb2, c->red + 4);
    read_scalars(c, 8);
    *out = c->host_red[0] + c->host_red[4];
  });
}

int ftcc_dot(ftcc_ctx *c, const double *x, const double *y, int64_t n,
             double *out) {
  return guarded(c, [&] {
    require(x && y && out && n >= 0, "bad dot arguments");
    Adjoint *A = adjoint_ws(c);
    const int nb = grid_for(n, MAX_RED_BLOCKS);
    hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(AT), 0, c->stream, x, y, n,
                       A->partials);
    FTCC_CHECK_HIP(hipGetLastError());
    sum4(c, A, nb, c->red);
    read_scalars(c, 1);
    *out = c->host_red[0];
  });
}

int ftcc_set_system_response(ftcc_ctx *c, int n, const double *en,
                             const double *fock, const double *dfock,
                             const double *eri, const double *dx,
                             double beta, double mu, const int *iocc,
                             const int *ivir) {
  return guarded(c, [&] {
    drop_graph(c);
    require(n > 0 && en && fock && dfock && eri && dx,
            "bad system arguments");
    require(beta > 0.0 && std::isfinite(beta), "beta must be positive");
    const int O = c->no, V = c->nv;
    std::vector<int> io(O), iv(V);
    for (int k = 0; k < O; ++k) {
      io[k] = iocc ? iocc[k] : k;
      require(io[k] >= 0 && io[k] < n, "occupied index out of range");
    }
    for (int k = 0; k < V; ++k) {
      iv[k] = ivir ? ivir[k] : k;
      require(iv[k] >= 0 && iv[k] < n, "virtual index out of range");
    }
    hipStream_t s = c->stream;
    ensure_integrals(c);
    // scratch: fo fv sfo sfv gp0 [n each], wo cO [O], wv cV [V], index lists
    std::vector<void *> tmp;
    auto talloc = [&](size_t bytes) {
      void *p = nullptr;
      FTCC_CHECK_HIP(hipMalloc(&p, bytes));
      tmp.push_back(p);
      return p;
    };
    struct Free {
      std::vector<void *> &v;
      hipStream_t s;
      ~Free() {
        (void)hipStreamSynchronize(s);
        for (void *p : v) (void)hipFree(p);
      }
    } release{tmp, s};
    double *fo = (double *)talloc(sizeof(double) * 5 * n);
    double *wo = (double *)talloc(sizeof(double) * 2 * O);
    double *wv = (double *)talloc(sizeof(double) * 2 * V);
    double *co = wo + O, *cv = wv + V;
    int *d_io = (int *)talloc(sizeof(int) * O);
    int *d_iv = (int *)talloc(sizeof(int) * V);
    FTCC_CHECK_HIP(hipMemcpyAsync(d_io, io.data(), sizeof(int) * O,
                                  hipMemcpyHostToDevice, s));
    FTCC_CHECK_HIP(hipMemcpyAsync(d_iv, iv.data(), sizeof(int) * V,
                                  hipMemcpyHostToDevice, s));
    double *fv = fo + n, *sfo = fo + 2 * n, *sfv = fo + 3 * n,
           *gp0 = fo + 4 * n;
    fermi_weights(en, n, beta, mu, fo, fv, sfo, sfv, gp0, s);
    gather_vec(sfo, d_io, O, wo, s);
    gather_vec(sfv, d_iv, V, wv, s);
    gather_vec(en, d_io, O, c->eo, s);
    gather_vec(en, d_iv, V, c->ev, s);
    hipLaunchKernelGGL(log_weight_kernel, dim3(grid_for(O + V)), dim3(AT), 0,
                       s, fo, fv, dx, d_io, O, d_iv, V, co, cv);
    FTCC_CHECK_HIP(hipGetLastError());
    auto fblock = [&](bool po, bool qo, double *out) {
      const int di = po ? O : V, dj = qo ? O : V;
      hipLaunchKernelGGL(fock_response_kernel,
                         dim3(grid_for((int64_t)di * dj)), dim3(AT), 0, s,
                         fock, dfock, en, n, po ? d_io : d_iv,
                         qo ? d_io : d_iv, po ? wo : wv, qo ? wo : wv,
                         po ? co : cv, qo ? co : cv, di, dj, out);
      FTCC_CHECK_HIP(hipGetLastError());
    };
    fblock(true, true, c->F[0]);
    fblock(true, false, c->F[1]);
    fblock(false, true, c->F[2]);
    fblock(false, false, c->F[3]);
    const char *names[9] = {"vvvv", "vvvo", "vovv", "vvoo", "vovo",
                            "oovv", "vooo", "ooov", "oooo"};
    for (int k = 0; k < 9; ++k) {
      const int *p[4];
      const double *w[4], *cc[4];
      int d[4];
      for (int a = 0; a < 4; ++a) {
        const bool occ = names[k][a] == 'o';
        p[a] = occ ? d_io : d_iv;
        w[a] = occ ? wo : wv;
        cc[a] = occ ? co : cv;
        d[a] = occ ? O : V;
      }
      const int64_t total = (int64_t)d[0] * d[1] * d[2] * d[3];
      hipLaunchKernelGGL(response4_kernel, dim3(grid_for(total)), dim3(AT), 0,
                         s, eri, n, p[0], p[1], p[2], p[3], w[0], w[1], w[2],
                         w[3], cc[0], cc[1], cc[2], cc[3], d[0], d[1], d[2],
                         d[3], c->I[k]);
      FTCC_CHECK_HIP(hipGetLastError());
    }
    c->have_D = false;
    c->have_energies = true;
    build_derived(c);
    c->spin = 0;
    const char *env = getenv("FTCC_SPIN");
    if (n % 2 == 0 && !(env && atoi(env) == 0)) {
      std::vector<int> lo(O), lv(V);
      for (int k = 0; k < O; ++k) lo[k] = io[k] >= n / 2;
      for (int k = 0; k < V; ++k) lv[k] = iv[k] >= n / 2;
      try_spin(c, lo, lv);
    }
    FTCC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
