// C ABI of libftcc: context, Fermi-weighted integrals, the Stanton-Gauss
// residual at every grid point, integration, damping and energies.
//
// Residual schedule for a chunk of nb grid points (all indices over the
// occupied-role list o and virtual-role list v; O = no, V = nv).  Five
// O(n^6) products run on the matrix cores:
//   G1  Wmnij += 1/2 oovv[mn,ef] . tau[ef,ij]
//   G2  R2    += 1/2 tau[ab,mn] . Wmnij[mn,ij]
//   G3  R2    += 1/2 Wabef[ab,ef] . tau[ef,ij]
//   G4  Wr[em,bj] -= oovv[m,n,e,f](as [em,fn]) . A[fn,bj],
//                    A[f,n,b,j] = 1/2 t_jn^fb + t_j^f t_n^b
//   G5  Z[ai,bj]   = t2[a,e,i,m](as [ai,em]) . Wr[em,bj]
// ("ring" tensors are stored [p,k,q,l] for <pq..kl> so G4/G5 are plain
// products), and the O(n^5) terms are batched GEMMs on the same engine.
// The Stanton-Gauss intermediates carry 1/4 tau.oovv each in Wmnij and
// Wabef, so that R2 holds 1/2 . 1/4 tau.oovv.tau through each ladder;
// here Wmnij carries all of it (1/2) and Wabef none, which is the same
// 1/4 tau.oovv.tau in R2 for one product fewer.
// Equations: Stanton, Gauss, Watts, Bartlett, J. Chem. Phys. 94, 4334 (1991)
// in kelvin's block conventions (kelvin/ft_cc_equations.py:96-113).
#include "ftcc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <vector>

using namespace ftcc;

#include "ftcc_ctx.h"

namespace ftcc_core {

constexpr int MAX_PARTIAL_BLOCKS = 8192;

double *dalloc(ftcc_ctx *c, int64_t n) {
  if (n <= 0) n = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(double) * (size_t)n);
  if (e != hipSuccess)
    throw Error(FTCC_ENOMEM, "hipMalloc of " + std::to_string(8 * n) +
                                 " bytes failed: " + hipGetErrorString(e));
  c->allocs.push_back(p);
  c->alloc_doubles[p] = n;
  c->bytes += 8 * n;
  return (double *)p;
}

// release one allocation of dalloc
void dfree(ftcc_ctx *c, void *p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  auto sz = c->alloc_doubles.find(p);
  if (sz != c->alloc_doubles.end()) {
    c->bytes -= 8 * sz->second;
    c->alloc_doubles.erase(sz);
  }
  FTCC_CHECK_HIP(hipFree(p));
}


int64_t p4(int a, int b, int c, int d) {
  return (int64_t)a * b * c * d;
}

// Output strides for writing a dense (d0,d1,d2,d3) source into the axis
// order `perm` (perm[k] = source axis placed at output position k).
void derive(const double *src, double *dst, const int d[4], const int perm[4],
            double alpha, hipStream_t s) {
  int od[4];
  for (int k = 0; k < 4; ++k) od[k] = d[perm[k]];
  int64_t ostride[4];
  ostride[3] = 1;
  for (int k = 2; k >= 0; --k) ostride[k] = ostride[k + 1] * od[k + 1];
  int64_t sstr[4];
  for (int k = 0; k < 4; ++k) sstr[perm[k]] = ostride[k];
  permute4(src, dst, d[0], d[1], d[2], d[3], sstr[0], sstr[1], sstr[2],
           sstr[3], alpha, 0.0, 1, 0, 0, s);
}

void build_derived(ftcc_ctx *c) {
  const int O = c->no, V = c->nv;
  hipStream_t s = c->stream;
  const int doovv[4] = {O, O, V, V};
  const int pr[4] = {2, 0, 3, 1};
  derive(c->I[5], c->oovv_ring, doovv, pr, 1.0, s);
  const int pt[4] = {2, 3, 0, 1};
  derive(c->I[5], c->oovvT, doovv, pt, 1.0, s);
  const int pk[4] = {0, 2, 3, 1};
  derive(c->I[5], c->oovv_mefn, doovv, pk, 1.0, s);
  const int dvovo[4] = {V, O, V, O};
  const int pv[4] = {2, 1, 0, 3};
  derive(c->I[4], c->ovvo_ring, dvovo, pv, -1.0, s);
  const int dring[4] = {V, O, V, O};
  derive(c->ovvo_ring, c->ovvo_ringT, dring, pt, 1.0, s);
  const int dooov[4] = {O, O, O, V};
  const int p1[4] = {0, 2, 1, 3};
  derive(c->I[7], c->ooov_p1, dooov, p1, 1.0, s);
  const int pe[4] = {2, 3, 1, 0};
  derive(c->I[7], c->ooov_iemn, dooov, pe, 1.0, s);
  const int pj[4] = {3, 0, 1, 2};
  derive(c->I[7], c->ooov_emnj, dooov, pj, 1.0, s);
  const int dvovv[4] = {V, O, V, V};
  const int pb[4] = {2, 1, 0, 3};
  derive(c->I[2], c->vovv_embf, dvovv, pb, 1.0, s);
  derive(c->I[2], c->vovv_aefm, dvovv, pk, 1.0, s);
  transpose2(c->F[1], c->fovT, O, V, 1, s);
  c->have_ints = true;
  c->split_fixed_valid = false;
  c->b64_fixed_valid = false;
}

// GEMM helper: dense row-major operands with per-slot batch strides.
Gemm mk(int M, int N, int K, double alpha, double beta, const double *A,
        int64_t sAm, int64_t sAk, int64_t sA1, const double *B, int64_t sBk,
        int64_t sBn, int64_t sB1, double *C, int64_t sCm, int64_t sCn,
        int64_t sC1, int nb1) {
  Gemm g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.sAm = sAm;
  g.sAk = sAk;
  g.sA1 = sA1;
  g.B = B;
  g.sBk = sBk;
  g.sBn = sBn;
  g.sB1 = sB1;
  g.C = C;
  g.sCm = sCm;
  g.sCn = sCn;
  g.sC1 = sC1;
  g.nb1 = nb1;
  return g;
}

void ensure_partial(ftcc_ctx *c, const Gemm &g) {
  const size_t need = gemm_partial_elems(g);
  if (need > c->scr.partial_elems) {
    // grow: allocation happens outside the kernel stream order, so drain it
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (c->scr.partial) {
      auto it = std::find(c->allocs.begin(), c->allocs.end(),
                          (void *)c->scr.partial);
      if (it != c->allocs.end()) c->allocs.erase(it);
      c->bytes -= 8 * (int64_t)c->scr.partial_elems;
      FTCC_CHECK_HIP(hipFree(c->scr.partial));
    }
    c->scr.partial = dalloc(c, (int64_t)need);
    c->scr.partial_elems = need;
  }
}

hipEvent_t take_event(ftcc_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  FTCC_CHECK_HIP(hipEventCreate(&e));
  return e;
}

// Accumulate finished event pairs (call only after the stream drained).
void flush_profile(ftcc_ctx *c) {
  for (auto &p : c->prof_pending) {
    float ms = 0.f;
    FTCC_CHECK_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
    c->prof_ms[p.cls] += ms;
    c->prof_flops[p.cls] += p.flops;
    c->prof_bytes[p.cls] += p.bytes;
    c->prof_launches[p.cls] += 1;
    if (c->prof_sites) {
      ftcc_ctx::SiteStat &st = c->site_stats[std::array<int64_t, 6>{
          p.cls, p.line, p.M, p.N, p.K, p.nb}];
      st.launches += 1;
      st.ms += ms;
      st.flops += p.flops;
      st.bytes += p.bytes;
      st.M = p.M;
      st.N = p.N;
      st.K = p.K;
      st.nb = p.nb;
    }
    c->ev_pool.push_back(p.e0);
    c->ev_pool.push_back(p.e1);
  }
  c->prof_pending.clear();
}

// Run fn() bracketed by events of kernel class cls when profiling is on,
// with its algorithmic flops and bytes (cls < 0: never timed).
template <typename Fn>
void timed_class(ftcc_ctx *c, int cls, double flops, double bytes, Fn fn) {
  if (cls >= 0 && c->prof_on) {
    ftcc_ctx::Pending p;
    p.e0 = take_event(c);
    p.e1 = take_event(c);
    p.cls = cls;
    p.flops = flops;
    p.bytes = bytes;
    p.line = c->cur_line;
    p.M = c->cur_M;
    p.N = c->cur_N;
    p.K = c->cur_K;
    p.nb = c->cur_nb;
    FTCC_CHECK_HIP(hipEventRecord(p.e0, c->stream));
    fn();
    FTCC_CHECK_HIP(hipEventRecord(p.e1, c->stream));
    c->prof_pending.push_back(p);
  } else {
    fn();
  }
}

// the same for code outside this file: prof_begin returns a token for
// prof_end (-1: profiling is off)
int prof_begin(ftcc_ctx *c, int cls, double flops, double bytes) {
  if (cls < 0 || !c->prof_on) return -1;
  ftcc_ctx::Pending p;
  p.e0 = take_event(c);
  p.e1 = take_event(c);
  p.cls = cls;
  p.flops = flops;
  p.bytes = bytes;
  p.line = c->site_override ? c->site_override : c->cur_line;
  p.M = c->cur_M;
  p.N = c->cur_N;
  p.K = c->cur_K;
  p.nb = c->cur_nb;
  FTCC_CHECK_HIP(hipEventRecord(p.e0, c->stream));
  c->prof_pending.push_back(p);
  return (int)c->prof_pending.size() - 1;
}
void prof_end(ftcc_ctx *c, int token) {
  if (token < 0) return;
  FTCC_CHECK_HIP(hipEventRecord(c->prof_pending[token].e1, c->stream));
}

// the O(n^6) products (big) are class PROF_PRODUCTS
template <typename Fn>
void timed(ftcc_ctx *c, bool big, double flops, Fn fn) {
  timed_class(c, big ? (int)ftcc_ctx::PROF_PRODUCTS : -1, flops, 0.0, fn);
}

double gemm_flops(const Gemm &g) {
  return 2.0 * g.M * (double)g.N * g.K * g.nb1 * g.nb2;
}

// bytes of a GEMM that reads A and B and writes C once each (C also read
// when beta != 0); an operand shared by the slices is read once
double gemm_bytes(const Gemm &g) {
  const double sl = (double)g.nb1 * g.nb2;
  const double a = (double)g.M * g.K * ((g.sA1 || g.sA2) ? sl : 1.0);
  const double b = (double)g.K * g.N * ((g.sB1 || g.sB2) ? sl : 1.0);
  const double cm = (double)g.M * g.N * (g.reduce_b2 ? g.nb1 : sl);
  return 8.0 * (a + b + cm * (g.beta != 0.0 ? 2.0 : 1.0));
}

// A batched GEMM as one launch whose outer row (column) level runs over
// its slices: the operand shared by every slice (slice stride 0) is then
// read once for all of them.  g has one row (column) index.
Gemm fold_m(Gemm g) {
  g.M1 = g.nb1;
  g.M2 = g.M;
  g.M *= g.nb1;
  g.sAm1 = g.sA1;
  g.sCm1 = g.sC1;
  g.sDm1 = g.sD1;
  g.nb1 = 1;
  return g;
}
Gemm fold_n(Gemm g) {
  g.N1 = g.nb1;
  g.N2 = g.N;
  g.N *= g.nb1;
  g.sBn1 = g.sB1;
  g.sCn1 = g.sC1;
  g.sDn1 = g.sD1;
  g.nb1 = 1;
  return g;
}

void run(ftcc_ctx *c, const Gemm &g0, int prec, bool big, int line) {
  Gemm g = g0;
  g.tile_policy = c->tile_policy;
  g.split_target = c->split_target;
  ensure_partial(c, g);
  c->cur_line = c->site_override ? c->site_override : line;
  c->cur_M = g.M;
  c->cur_N = g.N;
  c->cur_K = g.K;
  c->cur_nb = g.nb1 * g.nb2;
  timed_class(c, big ? ftcc_ctx::PROF_PRODUCTS : ftcc_ctx::PROF_N5,
              gemm_flops(g), big ? 0.0 : gemm_bytes(g),
              [&] { gemm_run(g, prec, c->scr, c->stream); });
  c->cur_line = 0;
}


// Spin conservation rule of an operand, "mn|ie": the spins of the letters
// left of '|' sum to those right of it ("e|j": a spin-diagonal matrix).
bool rule_holds(const char *rule, const std::map<char, int> &sp) {
  int left = 0, right = 0;
  bool r = false;
  for (const char *q = rule; *q; ++q) {
    if (*q == '|') {
      r = true;
      continue;
    }
    (r ? right : left) += sp.at(*q);
  }
  return left == right;
}

// C[M,N] = alpha * A[M,K] B[K,N] + beta * C + gamma * D over the
// spin-allowed blocks only: M, N and K run over up to three orbital indices
// each (K over several only with both operands K-fastest), and every
// assignment of spins to the letters under which the rules of
// A, B and C all hold is one launch over the product of the letters' spin
// runs (the other blocks of A and B are zero, those of C are not written).
// `base` carries the pointers, alpha/beta, D, the slice count and the slice
// strides; the rest is filled here.  An M or N index of role 'y' runs over
// the base's slices (grid points) inside one launch: a fixed integral
// operand is then read once for all of them; it takes no spin, and D (an
// integral) does not move along it.  Role '1' is an index of extent one
// (a matrix-vector product).
void run_spin(ftcc_ctx *c, const Gemm &base, int prec,
              const std::vector<SIdx> &Mi, const std::vector<SIdx> &Ni,
              const std::vector<SIdx> &Ki, const char *ruleA,
              const char *ruleB, const char *ruleC, int line) {
  struct Site {
    ftcc_ctx *c;
    int saved;
    ~Site() { c->site_override = saved; }
  } site{c, c->site_override};
  if (!c->site_override) c->site_override = line;
  std::vector<char> letters;
  int batch = 0;
  int unit = 0;
  auto is_pair = [](const SIdx &x) { return x.role == 'P' || x.role == 'Q'; };
  auto add1 = [&](char l) {
    require(std::find(letters.begin(), letters.end(), l) == letters.end(),
            "spin-restricted GEMM: repeated index");
    letters.push_back(l);
  };
  auto add = [&](const SIdx &x) {
    if (x.role == 'y' || x.role == '1') {
      ++(x.role == 'y' ? batch : unit);
      return;
    }
    add1(x.l);
    if (is_pair(x)) add1(x.l2);
  };
  int pm = 0, pn = 0;
  for (const SIdx &x : Mi) {
    add(x);
    pm += is_pair(x);
  }
  for (const SIdx &x : Ni) {
    add(x);
    pn += is_pair(x);
  }
  for (const SIdx &x : Ki) {
    require(x.role != 'y' && x.role != '1' && !is_pair(x),
            "spin-restricted GEMM: slice or pair index in K");
    add(x);
  }
  require(Mi.size() >= 1 && Mi.size() <= 3 && Ni.size() >= 1 &&
              Ni.size() <= 3 && Ki.size() >= 1 && Ki.size() <= 3 &&
              batch <= 1 && pm <= 1 && pn <= 1 && (!(pm + pn) || c->pack),
          "spin-restricted GEMM: bad index lists");
  // a flip lead must index C (skipping a K spin would drop partial sums)
  const char lead = c->flip_lead;
  if (lead) {
    bool in_c = false;
    for (const auto *v : {&Mi, &Ni})
      for (const SIdx &x : *v)
        in_c |= (x.role == 'o' || x.role == 'v') && x.l == lead;
    require(in_c, "spin-restricted GEMM: flip lead is not an output index");
  }
  // output pairs of class beta-alpha (skip_ba) or beta-beta (skip_bb) that
  // no reader reads: both letters index C
  const char bx = c->skip_ba[0], by = c->skip_ba[1];
  const char qx = c->skip_bb[0], qy = c->skip_bb[1];
  auto in_c = [&](char l, bool pairs) {
    int found = 0;
    for (const auto *v : {&Mi, &Ni})
      for (const SIdx &x : *v)
        found += ((x.role == 'o' || x.role == 'v') && x.l == l) ||
                 (pairs && is_pair(x) && (x.l == l || x.l2 == l));
    return found == 1;
  };
  require(!bx || (in_c(bx, false) && in_c(by, false)),
          "spin-restricted GEMM: skipped pair is not in C");
  require(!qx || (in_c(qx, true) && in_c(qy, true)),
          "spin-restricted GEMM: skipped pair is not in C");
  auto spin_of = [](const std::map<char, int> &sp, char l) {
    const auto it = sp.find(l);
    return it == sp.end() ? -1 : it->second;
  };
  auto lead_beta = [&](const std::map<char, int> &sp) {
    if (lead && spin_of(sp, lead) == 1) return true;
    if (bx && spin_of(sp, bx) == 1 && spin_of(sp, by) == 0) return true;
    return qx && spin_of(sp, qx) == 1 && spin_of(sp, qy) == 1;
  };
  // a pair index runs over the packed pairs of its spin class (x < y for
  // equal spins, every mixed pair once, none for the beta-alpha class);
  // its level maps the packed index to the dense pair x * n + y
  auto pair_space = [&](const SIdx &x) -> const PairSpace & {
    return x.role == 'P' ? c->ps_oop : c->ps_vvp;
  };
  auto cls = [&](const SIdx &x, std::map<char, int> &sp) {
    return 2 * sp[x.l] + sp[x.l2];
  };
  auto start = [&](const SIdx &x, std::map<char, int> &sp) -> int64_t {
    if (x.role == 'y' || x.role == '1' || is_pair(x)) return 0;
    return x.role == 'o' ? c->o_start[sp[x.l]] : c->v_start[sp[x.l]];
  };
  auto count = [&](const SIdx &x, std::map<char, int> &sp) -> int64_t {
    if (x.role == 'y') return base.nb1;
    if (x.role == '1') return 1;
    if (is_pair(x)) return pair_space(x).count[cls(x, sp)];
    return x.role == 'o' ? c->o_count[sp[x.l]] : c->v_count[sp[x.l]];
  };
  // D's stride along an index: none along the slices
  auto sd = [](const SIdx &x) { return x.role == 'y' ? 0 : x.sc; };
  const bool writes_all = base.beta == 0.0 || base.D != nullptr;
  std::set<std::vector<int>> c_blocks;   // C spin blocks that were written
  const int L = (int)letters.size();
  const SIdx z{0, 'o', 0, 0, 0};
  for (int mask = 0; mask < (1 << L); ++mask) {
    std::map<char, int> sp;
    for (int k = 0; k < L; ++k) sp[letters[k]] = (mask >> k) & 1;
    if (!rule_holds(ruleA, sp) || !rule_holds(ruleB, sp) ||
        !rule_holds(ruleC, sp) || lead_beta(sp))
      continue;
    Gemm g = base;
    int64_t oa = 0, ob = 0, oc = 0;
    int64_t M = 1, N = 1, K = 1;
    std::vector<int> cm, cn, ck;
    for (const SIdx &x : Mi) {
      oa += start(x, sp) * x.sa;
      oc += start(x, sp) * x.sc;
      cm.push_back((int)count(x, sp));
      M *= count(x, sp);
    }
    for (const SIdx &x : Ni) {
      ob += start(x, sp) * x.sb;
      oc += start(x, sp) * x.sc;
      cn.push_back((int)count(x, sp));
      N *= count(x, sp);
    }
    for (const SIdx &x : Ki) {
      oa += start(x, sp) * x.sa;
      ob += start(x, sp) * x.sb;
      ck.push_back((int)count(x, sp));
      K *= count(x, sp);
    }
    std::vector<int> cb;
    for (const SIdx &x : Mi) {
      cb.push_back(sp[x.l]);
      if (is_pair(x)) cb.push_back(sp[x.l2]);
    }
    for (const SIdx &x : Ni) {
      cb.push_back(sp[x.l]);
      if (is_pair(x)) cb.push_back(sp[x.l2]);
    }
    // several K spin blocks onto one C block: the first applies beta and
    // D, the others accumulate
    const bool first = c_blocks.insert(cb).second;
    if (!first) {
      g.beta = 1.0;
      g.D = nullptr;
    }
    if (M == 0 || N == 0) continue;
    require(M < (1LL << 31) && N < (1LL << 31) && K < (1LL << 31),
            "spin block too large");
    g.M = (int)M;
    g.N = (int)N;
    g.K = (int)K;
    g.A = base.A + oa;
    g.B = base.B + ob;
    g.C = base.C + oc;
    g.D = (first && base.D) ? base.D + oc : nullptr;
    // levels: the listed indices fill the innermost slots
    const int nm = (int)Mi.size(), nn = (int)Ni.size(), nk = (int)Ki.size();
    for (int q = 0; q < nm; ++q)
      if (is_pair(Mi[q])) {
        g.mmap = pair_space(Mi[q]).d_map + pair_space(Mi[q]).start[cls(Mi[q], sp)];
        g.mlvl = 3 - nm + q;
      }
    for (int q = 0; q < nn; ++q)
      if (is_pair(Ni[q])) {
        g.nmap = pair_space(Ni[q]).d_map + pair_space(Ni[q]).start[cls(Ni[q], sp)];
        g.nlvl = 3 - nn + q;
      }
    const SIdx &m0 = nm == 3 ? Mi[0] : z, &m1 = nm >= 2 ? Mi[nm - 2] : z,
               &m2 = Mi[nm - 1];
    const SIdx &n0 = nn == 3 ? Ni[0] : z, &n1 = nn >= 2 ? Ni[nn - 2] : z,
               &n2 = Ni[nn - 1];
    const SIdx &k0 = nk == 3 ? Ki[0] : z, &k1 = nk >= 2 ? Ki[nk - 2] : z,
               &k2 = Ki[nk - 1];
    g.M1 = nm >= 2 ? cm[nm - 2] : 1;
    g.M2 = cm[nm - 1];
    g.N1 = nn >= 2 ? cn[nn - 2] : 1;
    g.N2 = cn[nn - 1];
    g.sAm0 = m0.sa;
    g.sAm1 = m1.sa;
    g.sAm = m2.sa;
    g.sCm0 = m0.sc;
    g.sCm1 = m1.sc;
    g.sCm = m2.sc;
    g.sDm0 = sd(m0);
    g.sDm1 = sd(m1);
    g.sDm = sd(m2);
    g.sBn0 = n0.sb;
    g.sBn1 = n1.sb;
    g.sBn = n2.sb;
    g.sCn0 = n0.sc;
    g.sCn1 = n1.sc;
    g.sCn = n2.sc;
    g.sDn0 = sd(n0);
    g.sDn1 = sd(n1);
    g.sDn = sd(n2);
    if (batch) g.nb1 = 1;   // the slices are an index of the launch
    g.sAk = k2.sa;
    g.sBk = k2.sb;
    if (nk >= 2) {
      g.K1 = ck[nk - 2];
      g.K2 = ck[nk - 1];
      g.sAk0 = k0.sa;
      g.sAk1 = k1.sa;
      g.sBk0 = k0.sb;
      g.sBk1 = k1.sb;
    }
    if (K == 0) {
      if (!writes_all || !first) continue;
      g.K = 0;   // C = beta * C + gamma * D only
      g.K1 = 1;
      g.K2 = 0;
    }
    run(c, g, prec);
  }
  if (writes_all) {
    // every spin-allowed, nonempty block of C must have been written
    const int nc = (int)(Mi.size() + Ni.size()) - batch - unit + pm + pn;
    for (int mask = 0; mask < (1 << nc); ++mask) {
      std::map<char, int> sp;
      std::vector<int> cb;
      int k = 0;
      int64_t cells = 1;
      auto take = [&](const SIdx &x) {
        sp[x.l] = (x.role == 'y' || x.role == '1') ? 0 : (mask >> k++) & 1;
        cb.push_back(sp[x.l]);
        if (is_pair(x)) {
          sp[x.l2] = (mask >> k++) & 1;
          cb.push_back(sp[x.l2]);
        }
        cells *= count(x, sp);
      };
      for (const SIdx &x : Mi) take(x);
      for (const SIdx &x : Ni) take(x);
      if (rule_holds(ruleC, sp) && cells > 0 && !lead_beta(sp))
        require(c_blocks.count(cb) == 1,
                "spin-restricted GEMM leaves an allowed output block unset");
    }
  }
}

void run_spin(ftcc_ctx *c, const Gemm &base, int prec,
              const std::vector<SIdx> &Mi, const std::vector<SIdx> &Ni,
              const SIdx &Ki, const char *ruleA, const char *ruleB,
              const char *ruleC, int line = __builtin_LINE()) {
  run_spin(c, base, prec, Mi, Ni, std::vector<SIdx>{Ki}, ruleA, ruleB, ruleC,
           line);
}

// One of the five O(n^6) products: split-bf16 operands in bf16x3 mode
// (A as [M][K], B as [N][K]), the generic engine otherwise.
void big(ftcc_ctx *c, const Gemm &g, int prec, const SplitMat &A,
         const SplitMat &B) {
  if (prec != FTCC_PREC_BF16X3) {
    run(c, g, prec, true);
    return;
  }
  timed(c, true, gemm_flops(g), [&] {
    sbgemm(g.M, g.N, g.K, g.nb1, g.alpha, A, B, g.beta, g.C, g.sCm, g.sCn,
           g.sC1, c->stream);
  });
}

struct Range {
  int start = 0, len = 0, count = 0;
};
struct Blk {
  Range r, k, c;
};

// classes c1 (and c2, adjacent after c1) of a pair space
Range group(const PairSpace &p, int c1, int c2 = -1) {
  Range g;
  g.start = p.start[c1];
  g.len = p.len[c1];
  g.count = p.count[c1];
  if (c2 >= 0) {
    g.len += p.len[c2];
    g.count += p.count[c2];
  }
  return g;
}

// ---------------------------------------------------------------------------
// Operands of the five O(n^6) products.  Each is a matrix over two pair
// spaces (rows, K); A operands are stored [M][K], B operands transposed as
// [N][K].  Pair spaces: 0 = (a,b) virtual pairs, 1 = (i,j) occupied pairs,
// 2 = ring (a,i) pairs.  Ids 7 and 8 depend on the integrals only.
enum Op {
  OP_TAU_A, OP_TAU_B, OP_T2R_A, OP_ARING_B, OP_WMNIJ_B, OP_WABEF_A, OP_WR_B,
  OP_OOVV_A, OP_RING_A, NOPS
};
constexpr int OP_ROWS[NOPS] = {0, 1, 2, 2, 1, 0, 2, 1, 2};
constexpr int OP_K[NOPS] = {1, 0, 2, 2, 1, 0, 2, 0, 2};
constexpr bool OP_FIXED[NOPS] = {false, false, false, false, false,
                                 false, false, true,  true};

const PairSpace &space(const ftcc_ctx *c, int sp) {
  return sp == 0   ? c->ps_vv
         : sp == 1 ? c->ps_oo
         : sp == 2 ? c->ps_vo
         : sp == 3 ? c->ps_vvp
                   : c->ps_oop;
}

// pair space of operand `id`'s K index: the packed ladder spaces (3, 4)
// when packing is on
int op_k(const ftcc_ctx *c, int id) {
  return (c->spin && c->pack && OP_K[id] < 2) ? OP_K[id] + 3 : OP_K[id];
}
// and of its rows: every ladder operand and output is antisymmetric in both
// of its pairs, so packing covers its rows too (the outputs then hold one
// pair of each couple)
int op_rows(const ftcc_ctx *c, int id) {
  return (c->spin && c->pack && OP_ROWS[id] < 2) ? OP_ROWS[id] + 3
                                                 : OP_ROWS[id];
}

// layout of the product operands: 0 dense, 1 spin blocks, 2 spin blocks
// with packed ladder K
int op_layout(const ftcc_ctx *c) { return c->spin ? (c->pack ? 2 : 1) : 0; }

// length of a pair space in the current layout
int space_len(const ftcc_ctx *c, int sp) {
  if (c->spin) return space(c, sp).total;
  const int O = c->no, V = c->nv;
  return sp == 0 ? V * V : sp == 1 ? O * O : V * O;
}

SplitMat &split_of(ftcc_ctx *c, int id) {
  SplitMat *t[NOPS] = {&c->sp_tauA,   &c->sp_tauB,   &c->sp_t2rA,
                       &c->sp_aringB, &c->sp_wmnijB, &c->sp_wabefA,
                       &c->sp_wrB,    &c->sp_oovvA,  &c->sp_ringA};
  return *t[id];
}

// elements of one array of a split operand (hi or lo), all slots
int64_t split_elems(const ftcc_ctx *c, int id) {
  const int64_t rows = space_len(c, op_rows(c, id));
  return rows * split_ld(space_len(c, op_k(c, id))) *
         (OP_FIXED[id] ? 1 : c->nb);
}

// doubles of one carved buffer, rounded to 256-byte boundaries
int64_t carve_len(int64_t doubles) { return (doubles + 31) / 32 * 32; }

// arena doubles of the operand set of precision `prec` at chunk nb (0: the
// dense FP64/FP32 layout reads its operands in place)
int64_t operand_doubles(const ftcc_ctx *c, int prec, int nb) {
  int64_t tot = 0;
  for (int id = 0; id < NOPS; ++id) {
    const int64_t rows = space_len(c, op_rows(c, id));
    const int64_t slots = OP_FIXED[id] ? 1 : nb;
    if (prec == FTCC_PREC_BF16X3) {
      const int64_t elems = rows * split_ld(space_len(c, op_k(c, id))) * slots;
      tot += 2 * carve_len((elems + 3) / 4 + 1);   // four bf16 per double
    } else if (c->spin) {
      tot += carve_len(rows * space_len(c, op_k(c, id)) * slots);
    }
  }
  return tot;
}

// The operand arena, at least `need` doubles.  Its first allocation also
// holds the other precision's set at half the chunk, so that the switches
// of a mixed-precision solve find room without allocating.
double *operand_arena(ftcc_ctx *c, int64_t need) {
  if (need <= c->arena_doubles) return c->arena;
  if (!c->arena) {
    const int other = c->prec == FTCC_PREC_BF16X3 ? FTCC_PREC_FP64
                                                  : FTCC_PREC_BF16X3;
    need = std::max(need,
                    operand_doubles(c, other, std::max(1, c->nb / 2)));
  }
  FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  dfree(c, c->arena);
  c->arena = nullptr;
  c->arena_doubles = 0;
  c->arena_owner = 0;
  c->arena = dalloc(c, need);
  c->arena_doubles = need;
  return c->arena;
}

// Split-bf16 operand buffers for the current layout and chunk, carved from
// the arena (the FP64 copies there are given up).
void ensure_split_buffers(ftcc_ctx *c) {
  if (c->arena_owner == 1 && c->split_layout == op_layout(c) &&
      c->split_nb == c->nb)
    return;
  double *p = operand_arena(c, operand_doubles(c, FTCC_PREC_BF16X3, c->nb));
  for (int id = 0; id < NOPS; ++id) c->b64[id] = nullptr;
  c->b64_layout = -1;
  for (int id = 0; id < NOPS; ++id) {
    SplitMat &m = split_of(c, id);
    m = SplitMat();
    m.ld = split_ld(space_len(c, op_k(c, id)));
    const int64_t rows = space_len(c, op_rows(c, id));
    m.bstride = OP_FIXED[id] ? 0 : rows * m.ld;
    const int64_t len = carve_len((split_elems(c, id) + 3) / 4 + 1);
    m.hi = (__bf16 *)p;
    p += len;
    m.lo = (__bf16 *)p;
    p += len;
  }
  c->arena_owner = 1;
  c->split_layout = op_layout(c);
  c->split_nb = c->nb;
  c->split_fixed_valid = false;
}

// FP64 spin-blocked operand buffers (FP64 and FP32 modes), carved from the
// arena (the split copies there are given up).
void ensure_b64_buffers(ftcc_ctx *c) {
  if (c->arena_owner == 2 && c->b64_layout == op_layout(c) &&
      c->b64_nb == c->nb)
    return;
  double *p = operand_arena(c, operand_doubles(c, FTCC_PREC_FP64, c->nb));
  for (int id = 0; id < NOPS; ++id) split_of(c, id) = SplitMat();
  c->split_layout = -1;
  for (int id = 0; id < NOPS; ++id) {
    c->b64[id] = p;
    p += carve_len((int64_t)space_len(c, op_rows(c, id)) *
                   space_len(c, op_k(c, id)) * (OP_FIXED[id] ? 1 : c->nb));
  }
  c->arena_owner = 2;
  c->b64_layout = op_layout(c);
  c->b64_nb = c->nb;
  c->b64_fixed_valid = false;
}

// The (row classes, K classes) blocks of operand `id` that its products read
// (block_list below): ladder operands the three blocks of equal total spin;
// ring operands the spin-keeping block and the two spin-flip blocks, crossed
// (flip classes 1 x 2 and 2 x 1) except for G5's B operand Wr, whose K and
// column classes agree.  Only these blocks are written in the spin layout.
void operand_blocks(const ftcc_ctx *c, int id, Range rows[3], Range ks[3]) {
  const PairSpace &R = space(c, op_rows(c, id)), &K = space(c, op_k(c, id));
  if (OP_ROWS[id] != 2) {
    rows[0] = group(R, 0), ks[0] = group(K, 0);
    rows[1] = group(R, 1, 2), ks[1] = group(K, 1, 2);
    rows[2] = group(R, 3), ks[2] = group(K, 3);
    // (the fixed integral operand is gathered once for every later call)
    if (c->flip_step_now && !OP_FIXED[id]) rows[2] = Range(), ks[2] = Range();
    return;
  }
  rows[0] = group(R, 0, 3), ks[0] = group(K, 0, 3);
  if (id == OP_T2R_A && c->flip_now) {
    // the spin-flip symmetric G5 reads the rows with an alpha virtual only
    rows[0] = group(R, 0);
    rows[1] = group(R, 1), ks[1] = group(K, 2);
    rows[2] = Range(), ks[2] = Range();
    return;
  }
  if (id == OP_WR_B) {
    rows[1] = group(R, 1), ks[1] = group(K, 1);
    rows[2] = group(R, 2), ks[2] = group(K, 2);
    // G5 over the alpha-virtual rows never reads the class-1 columns
    if (c->flip_now) rows[1] = Range(), ks[1] = Range();
  } else {
    rows[1] = group(R, 1), ks[1] = group(K, 2);
    rows[2] = group(R, 2), ks[2] = group(K, 1);
    // G4 over the alpha-virtual rows never reads the class-2 columns of B
    if (c->flip_now && id == OP_ARING_B) rows[2] = Range(), ks[2] = Range();
  }
}

// The spin blocks of operand `id` (operand_blocks) through the pair maps:
// split(rmap, rows, kmap, K, dst) in bf16x3 mode, f64(rmap, rows, kmap, K,
// ld, dst, slice stride) otherwise.
template <typename SplitFn, typename F64Fn>
void gather_blocks(ftcc_ctx *c, int id, SplitFn split, F64Fn f64,
                   int nb = 1) {
  const PairSpace &R = space(c, op_rows(c, id)), &K = space(c, op_k(c, id));
  Range rows[3], ks[3];
  operand_blocks(c, id, rows, ks);
  // algorithmic bytes: one FP64 read and one write (two bf16 halves or an
  // FP64 copy) per gathered element
  double elems = 0.0;
  for (int q = 0; q < 3; ++q)
    elems += (double)rows[q].count * ks[q].count * nb;
  const double per = c->prec == FTCC_PREC_BF16X3 ? 12.0 : 16.0;
  timed_class(c, ftcc_ctx::PROF_GATHER, 0.0, per * elems, [&] {
  for (int q = 0; q < 3; ++q) {
    if (rows[q].len == 0 || ks[q].len == 0) continue;
    const int *rmap = R.d_map + rows[q].start, *kmap = K.d_map + ks[q].start;
    if (c->prec == FTCC_PREC_BF16X3) {
      SplitMat m = split_of(c, id);
      const int64_t off = (int64_t)rows[q].start * m.ld + ks[q].start;
      m.hi += off;
      m.lo += off;
      split(rmap, rows[q].len, kmap, ks[q].len, m);
    } else {
      const int64_t ld = K.total;
      f64(rmap, rows[q].len, kmap, ks[q].len, ld,
          c->b64[id] + (int64_t)rows[q].start * ld + ks[q].start,
          (int64_t)R.total * ld);
    }
  }
  });
}

// Write operand `id` of nb slots from its FP64 source: src[b*sb + row*sr +
// k*sc] with row, k flat pair indices of the dense layout.
void put_operand(ftcc_ctx *c, int id, const double *src, int64_t sr,
                 int64_t sc, int64_t sb, int nb) {
  hipStream_t s = c->stream;
  if (!c->spin) {
    if (c->prec == FTCC_PREC_BF16X3)
      split_convert(src, sr, sc, sb, space_len(c, OP_ROWS[id]),
                    space_len(c, op_k(c, id)), split_of(c, id), nb, s);
    return;
  }
  gather_blocks(
      c, id,
      [&](const int *rmap, int R, const int *kmap, int W, const SplitMat &m) {
        split_gather(src, sr, sc, sb, rmap, R, kmap, W, m, nb, s);
      },
      [&](const int *rmap, int R, const int *kmap, int W, int64_t ld,
          double *dst, int64_t db) {
        gather64(src, sr, sc, sb, rmap, R, kmap, W, ld, dst, db, nb, s);
      },
      nb);
}

// G5's B operand Wr of a spin-flip symmetric residual, whose rows of an
// alpha e only are computed: the beta ones are read from their images
void put_flip_wr(ftcc_ctx *c, const double *wr, int64_t sb, int nb) {
  const int V = c->nv, O = c->no, hv = c->v_count[0], ho = c->o_count[0];
  hipStream_t s = c->stream;
  gather_blocks(
      c, OP_WR_B,
      [&](const int *rmap, int R, const int *kmap, int W, const SplitMat &m) {
        flip_ring_gather(wr, V, O, hv, ho, sb, rmap, R, kmap, W, &m, 0,
                         nullptr, 0, nb, s);
      },
      [&](const int *rmap, int R, const int *kmap, int W, int64_t ld,
          double *dst, int64_t db) {
        flip_ring_gather(wr, V, O, hv, ho, sb, rmap, R, kmap, W, nullptr, ld,
                         dst, db, nb, s);
      },
      nb);
}

// Spin layout: operand `id` gathered straight from the amplitudes t1, t2 of
// nb slices (tau, t2 in ring order, the ring A of G4: see AmpKind)
void put_amp_operand(ftcc_ctx *c, int id, int kind, const double *t1,
                     const double *t2, int nb) {
  require(c->spin, "amplitude gathers need the spin layout");
  hipStream_t s = c->stream;
  const int V = c->nv, O = c->no;
  gather_blocks(
      c, id,
      [&](const int *rmap, int R, const int *kmap, int W, const SplitMat &m) {
        amp_split_gather(kind, t1, t2, V, O, rmap, R, kmap, W, m, nb, s);
      },
      [&](const int *rmap, int R, const int *kmap, int W, int64_t ld,
          double *dst, int64_t db) {
        amp_gather64(kind, t1, t2, V, O, rmap, R, kmap, W, ld, dst, db, nb,
                     s);
      },
      nb);
}

// operands that depend on the integrals only
void ensure_fixed_operands(ftcc_ctx *c) {
  const int O = c->no, V = c->nv;
  const int64_t VV = (int64_t)V * V, VO = (int64_t)V * O;
  if (c->prec == FTCC_PREC_BF16X3) {
    ensure_split_buffers(c);
    if (c->split_fixed_valid) return;
    c->split_fixed_valid = true;
  } else if (c->spin) {
    ensure_b64_buffers(c);
    if (c->b64_fixed_valid) return;
    c->b64_fixed_valid = true;
  } else {
    return;
  }
  put_operand(c, OP_OOVV_A, c->I[5], VV, 1, 0, 1);   // [mn][ef]
  put_operand(c, OP_RING_A, c->oovv_ring, VO, 1, 0, 1);
}

// Pairs of (lx, ly) grouped by class 2*lx + ly in `order`, each class padded
// to a multiple of 256 (the product tile).
// packed (lx and ly the same list): only x < y within a class of equal
// spins, class 01 whole and class 10 empty -- one pair of each
// antisymmetric couple.
PairSpace build_pairs(ftcc_ctx *c, const std::vector<int> &lx,
                      const std::vector<int> &ly, const int order[4],
                      bool packed = false) {
  PairSpace ps;
  std::vector<int> map;
  const int ny = (int)ly.size();
  for (int q = 0; q < 4; ++q) {
    const int cls = order[q], sx = cls >> 1, sy = cls & 1;
    ps.start[cls] = (int)map.size();
    int cnt = 0;
    for (int x = 0; x < (int)lx.size(); ++x) {
      if (lx[x] != sx || (packed && sx > sy)) continue;
      for (int y = packed && sx == sy ? x + 1 : 0; y < ny; ++y)
        if (ly[y] == sy) {
          map.push_back(x * ny + y);
          ++cnt;
        }
    }
    while (map.size() % 256) map.push_back(-1);
    ps.count[cls] = cnt;
    ps.len[cls] = (int)map.size() - ps.start[cls];
  }
  ps.total = (int)map.size();
  ps.d_map = (int *)dalloc(c, ((int64_t)map.size() + 1) / 2 + 1);
  if (!map.empty())
    FTCC_CHECK_HIP(hipMemcpyAsync(ps.d_map, map.data(),
                                  sizeof(int) * map.size(),
                                  hipMemcpyHostToDevice, c->stream));
  FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  return ps;
}

// Class orders: pairs of equal total spin are adjacent in the ladder spaces
// ((a,b), (i,j): 00 | 01 10 | 11); pairs of equal spin flip in the ring
// space ((a,i): 00 11 | 01 | 10).
constexpr int LADDER_ORDER[4] = {0, 1, 2, 3};
constexpr int RING_ORDER[4] = {0, 3, 1, 2};

// The three nonzero blocks of a product C[R,Cs] = A[R,K] B[K,Cs].
//   ladder products (G1-G3): every operand conserves total spin of its pairs;
//   G4: Wr[em,bj] -= oovv_ring[em,fn] aring[fn,bj]: ring pairs of A and B
//       flip spin with opposite signs, so C keeps it;
//   G5: Z[ai,bj] = t2r[ai,em] Wr[em,bj]: A flips sign, B keeps it.
enum BlockKind { LADDER, RING_G4, RING_G5 };
void block_list(int kind, const PairSpace &R, const PairSpace &K,
                const PairSpace &C, Blk out[3]) {
  if (kind == LADDER) {
    out[0] = {group(R, 0), group(K, 0), group(C, 0)};
    out[1] = {group(R, 1, 2), group(K, 1, 2), group(C, 1, 2)};
    out[2] = {group(R, 3), group(K, 3), group(C, 3)};
    return;
  }
  // ring classes: 0 and 3 keep spin, 1 lowers it (a alpha, i beta), 2 raises
  out[0] = {group(R, 0, 3), group(K, 0, 3), group(C, 0, 3)};
  if (kind == RING_G4) {
    out[1] = {group(R, 1), group(K, 2), group(C, 1)};
    out[2] = {group(R, 2), group(K, 1), group(C, 2)};
  } else {
    out[1] = {group(R, 1), group(K, 2), group(C, 2)};
    out[2] = {group(R, 2), group(K, 1), group(C, 1)};
  }
}

// The rows of a ring tensor (layout [e][m][b][j], nb slices of V*O*V*O)
// whose virtual index e is beta, from their spin-flipped images: Wr[e,m,b,j]
// conserves se + sj = sm + sb, Z[a,i,b,j] sa + sb = si + sj (vv_pair).
void flip_rows(ftcc_ctx *c, double *X, int nb, bool vv_pair) {
  const int O = c->no, V = c->nv;
  flip_copy_ring(X, V, O, c->v_count[0], c->o_count[0], nb,
                 (int64_t)V * O * V * O, vv_pair, c->stream);
}

// One O(n^6) product C = alpha A.B + beta C over g.nb1 slots, g describing
// it in the dense layout and ia/ib naming its operands: dense, or as the
// three nonzero spin blocks written straight into C through the pair maps.
void product(ftcc_ctx *c, const Gemm &g, int ia, int ib, int kind) {
  const int P = c->prec;
  if (!c->spin) {
    big(c, g, P, split_of(c, ia), split_of(c, ib));
    return;
  }
  const PairSpace &R = space(c, op_rows(c, ia)), &K = space(c, op_k(c, ia));
  const PairSpace &Cs = space(c, op_rows(c, ib));
  Blk blk[3];
  block_list(kind, R, K, Cs, blk);
  // spin-flip symmetry: the ring products compute the rows whose virtual
  // index is alpha (classes 00 and 01 of the ring pairs); the others are
  // read from their flipped images once the output is complete (G5's B
  // gather for Wr, the assembly for Z)
  const bool flip = c->flip_now && kind != LADDER;
  if (flip) {
    blk[0].r = group(R, 0);
    blk[2].r = Range();
  }
  // the beta-beta | beta-beta block of a ladder product that nothing reads
  if (kind == LADDER && c->ladder_skip_bb) blk[2].r = Range();
  // a packed ladder K sums one pair of each antisymmetric couple: twice
  const double alpha = (kind == LADDER && op_k(c, ia) >= 3) ? 2.0 * g.alpha
                                                            : g.alpha;
  if (P == FTCC_PREC_BF16X3) {
    // the three blocks write disjoint parts of C: one grouped launch, the
    // largest block first (its tiles are the longest)
    SbProblem pb[3];
    int np = 0;
    double flops = 0.0;
    for (const Blk &b : blk) {
      // an empty K range still writes beta * C (zero for beta = 0); a
      // block without rows or columns has nothing to write
      if (!b.r.count || !b.c.count || (!b.k.count && g.beta != 0.0))
        continue;
      SbProblem &q = pb[np++];
      q.M = b.r.len;
      q.N = b.c.len;
      q.K = b.k.len;
      q.A = split_of(c, ia);
      q.B = split_of(c, ib);
      const int64_t ao = (int64_t)b.r.start * q.A.ld + b.k.start;
      const int64_t bo = (int64_t)b.c.start * q.B.ld + b.k.start;
      q.A.hi += ao;
      q.A.lo += ao;
      q.B.hi += bo;
      q.B.lo += bo;
      q.rmap = R.d_map + b.r.start;
      q.cmap = Cs.d_map + b.c.start;
      flops += 2.0 * b.r.count * (double)b.c.count * b.k.count * g.nb1;
    }
    std::sort(pb, pb + np, [](const SbProblem &x, const SbProblem &y) {
      return (double)x.M * x.N * x.K > (double)y.M * y.N * y.K;
    });
    if (np > 0)
      timed(c, true, flops, [&] {
        sbgemm_grouped(np, pb, g.nb1, alpha, g.beta, g.C, g.sCm, g.sCn,
                       g.sC1, c->stream);
      });
    return;
  }
  for (const Blk &b : blk) {
    if (!b.r.count || !b.c.count || (!b.k.count && g.beta != 0.0)) continue;
    const double flops = 2.0 * b.r.count * (double)b.c.count * b.k.count * g.nb1;
    const int64_t ld = K.total;
    Gemm gb = mk(b.r.len, b.c.len, b.k.len, alpha, g.beta,
                 c->b64[ia] + (int64_t)b.r.start * ld + b.k.start, ld, 1,
                 OP_FIXED[ia] ? 0 : (int64_t)R.total * ld,
                 c->b64[ib] + (int64_t)b.c.start * ld + b.k.start, 1, ld,
                 OP_FIXED[ib] ? 0 : (int64_t)Cs.total * ld, g.C, g.sCm,
                 g.sCn, g.sC1, g.nb1);
    gb.rmap = R.d_map + b.r.start;
    gb.cmap = Cs.d_map + b.c.start;
    ensure_partial(c, gb);
    timed(c, true, flops, [&] { gemm_run(gb, P, c->scr, c->stream); });
  }
}

void read_scalars(ftcc_ctx *c, int n);

// Spin labels (0/1) of the occupied and virtual index lists: blocking is
// switched on when every Fock and integral block conserves spin under them.
bool try_spin(ftcc_ctx *c, const std::vector<int> &lo,
              const std::vector<int> &lv) {
  const int O = c->no, V = c->nv;
  hipStream_t s = c->stream;
  c->spin = 0;
  if (!c->d_lab_o) {
    c->d_lab_o = (int *)dalloc(c, (O + 1) / 2 + 1);
    c->d_lab_v = (int *)dalloc(c, (V + 1) / 2 + 1);
    c->d_jl = (int *)dalloc(c, (O + 1) / 2 + 1);
  }
  FTCC_CHECK_HIP(hipMemcpyAsync(c->d_lab_o, lo.data(), sizeof(int) * O,
                                hipMemcpyHostToDevice, s));
  FTCC_CHECK_HIP(hipMemcpyAsync(c->d_lab_v, lv.data(), sizeof(int) * V,
                                hipMemcpyHostToDevice, s));
  std::vector<int> jl;
  for (int sp = 0; sp < 2; ++sp) {
    c->jl_cnt[sp] = 0;
    for (int k = 0; k < O; ++k)
      if (lo[k] == sp) {
        jl.push_back(k);
        ++c->jl_cnt[sp];
      }
  }
  if (!jl.empty())
    FTCC_CHECK_HIP(hipMemcpyAsync(c->d_jl, jl.data(), sizeof(int) * jl.size(),
                                  hipMemcpyHostToDevice, s));
  FTCC_CHECK_HIP(hipStreamSynchronize(s));   // the host lists are local
  auto lab = [&](char x) { return x == 'o' ? c->d_lab_o : c->d_lab_v; };
  auto dim = [&](char x) { return x == 'o' ? O : V; };
  double bad = 0.0;
  const char *fn[4] = {"oo", "ov", "vo", "vv"};
  for (int k = 0; k < 4; ++k) {
    // f[p][q] as a (P,1,Q,1) block: label(p) must equal label(q)
    const int nbk = spin_violations(c->F[k], dim(fn[k][0]), 1, dim(fn[k][1]),
                                    1, 1, lab(fn[k][0]), nullptr,
                                    lab(fn[k][1]), nullptr, c->partials, s);
    reduce_partials(c->partials, nbk, c->red, s);
    read_scalars(c, 3);
    bad += c->host_red[2];
  }
  const char *in[9] = {"vvvv", "vvvo", "vovv", "vvoo", "vovo",
                       "oovv", "vooo", "ooov", "oooo"};
  for (int k = 0; k < 9 && bad == 0.0; ++k) {
    const char *b = in[k];
    const int nbk = spin_violations(c->I[k], dim(b[0]), dim(b[1]), dim(b[2]),
                                    dim(b[3]), 1, lab(b[0]), lab(b[1]),
                                    lab(b[2]), lab(b[3]), c->partials, s);
    reduce_partials(c->partials, nbk, c->red, s);
    read_scalars(c, 3);
    bad += c->host_red[2];
  }
  if (bad != 0.0) return false;
  for (PairSpace *ps :
       {&c->ps_vv, &c->ps_oo, &c->ps_vo, &c->ps_vvp, &c->ps_oop}) {
    dfree(c, ps->d_map);
    *ps = PairSpace();
  }
  c->ps_vv = build_pairs(c, lv, lv, LADDER_ORDER);
  c->ps_oo = build_pairs(c, lo, lo, LADDER_ORDER);
  c->ps_vo = build_pairs(c, lv, lo, RING_ORDER);
  c->ps_vvp = build_pairs(c, lv, lv, LADDER_ORDER, true);
  c->ps_oop = build_pairs(c, lo, lo, LADDER_ORDER, true);
  const char *penv = getenv("FTCC_PACK");
  c->pack = (penv && atoi(penv) == 0) ? 0 : 1;
  c->spin = 1;
  // one run of alpha then one run of beta in each list
  auto runs = [](const std::vector<int> &l, int start[2], int count[2]) {
    int n0 = 0;
    while (n0 < (int)l.size() && l[n0] == 0) ++n0;
    for (size_t k = n0; k < l.size(); ++k)
      if (l[k] != 1) return false;
    start[0] = 0;
    count[0] = n0;
    start[1] = n0;
    count[1] = (int)l.size() - n0;
    return true;
  };
  const char *env = getenv("FTCC_SPIN5");
  const bool want5 = !(env && atoi(env) == 0);
  c->spin5 = (want5 && runs(lo, c->o_start, c->o_count) &&
              runs(lv, c->v_start, c->v_count))
                 ? 1
                 : 0;
  c->flip = 0;
  const char *fenv = getenv("FTCC_FLIP");
  if (c->spin5 && !(fenv && atoi(fenv) == 0) && c->have_energies &&
      !c->have_D && c->o_count[0] == c->o_count[1] &&
      c->v_count[0] == c->v_count[1]) {
    const int ho = c->o_count[0], hv = c->v_count[0];
    auto half = [&](char x) { return x == 'o' ? ho : hv; };
    double fbad = 0.0;
    auto check = [&](const double *a, int d0, int d1, int d2, int d3, int h0,
                     int h1, int h2, int h3) {
      const int nbk =
          flip_violations(a, d0, d1, d2, d3, h0, h1, h2, h3, c->partials, s);
      reduce_partials(c->partials, nbk, c->red, s);
      read_scalars(c, 3);
      fbad += c->host_red[2];
    };
    check(c->eo, O, 1, 1, 1, ho, 0, 0, 0);
    check(c->ev, V, 1, 1, 1, hv, 0, 0, 0);
    for (int k = 0; k < 4 && fbad == 0.0; ++k)
      check(c->F[k], dim(fn[k][0]), dim(fn[k][1]), 1, 1, half(fn[k][0]),
            half(fn[k][1]), 0, 0);
    for (int k = 0; k < 9 && fbad == 0.0; ++k) {
      const char *b = in[k];
      check(c->I[k], dim(b[0]), dim(b[1]), dim(b[2]), dim(b[3]), half(b[0]),
            half(b[1]), half(b[2]), half(b[3]));
    }
    c->flip = fbad == 0.0 ? 1 : 0;
  }
  return true;
}

void ensure_integrals(ftcc_ctx *c) {
  if (c->F[0]) return;
  const int O = c->no, V = c->nv;
  const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V, OO = (int64_t)O * O;
  const int64_t S2 = VV * OO;
  c->F[0] = dalloc(c, OO);
  c->F[1] = dalloc(c, VO);
  c->F[2] = dalloc(c, VO);
  c->F[3] = dalloc(c, VV);
  const int64_t isz[9] = {VV * VV, VV * VO, VO * VV, S2, S2,
                          S2,      V * OO * O, OO * VO, OO * OO};
  for (int k = 0; k < 9; ++k) c->I[k] = dalloc(c, isz[k]);
  c->oovv_ring = dalloc(c, S2);
  c->oovvT = dalloc(c, S2);
  c->oovv_mefn = dalloc(c, S2);
  c->vovv_aefm = dalloc(c, VO * VV);
  c->ovvo_ring = dalloc(c, S2);
  c->ovvo_ringT = dalloc(c, S2);
  c->ooov_p1 = dalloc(c, OO * VO);
  c->ooov_iemn = dalloc(c, OO * VO);
  c->ooov_emnj = dalloc(c, OO * VO);
  c->vovv_embf = dalloc(c, VO * VV);
  c->fovT = dalloc(c, VO);
  c->eo = dalloc(c, O);
  c->ev = dalloc(c, V);
}

// Residual workspace: as many grid points per chunk as fit the budget
// (FTCC_WORKSPACE_GB, default 120 GiB: eight grid points at n = 114 in
// bf16x3, six in FP64, the whole grid at n <= 66) -- larger chunks
// measured faster at every size tried (fewer, fuller launches; integrals
// read once per chunk).
// grid points per chunk for the current precision within the budget
int chunk_for_budget(const ftcc_ctx *c) {
  const int O = c->no, V = c->nv, ng = c->ng;
  const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V, OO = (int64_t)O * O;
  const int64_t S2 = VV * OO;
  const int64_t sm = std::max<int64_t>(VV, std::max<int64_t>(VO, OO));
  // five n^4 intermediates (seven in the dense layout, which also forms tau
  // and t2r), Wmnij/Wabef/U, and the seven per-slot product operands of the
  // current precision: split into bf16 pairs (bf16x3) or as FP64
  // spin-blocked copies (FP64/FP32 spin layout), one set at a time
  const int64_t operands =
      c->prec == FTCC_PREC_BF16X3
          ? 4 * 7 *
                (S2 + 64 * std::max<int64_t>(VV, std::max<int64_t>(VO, OO)))
          : (c->spin ? 8 * 7 * S2 : 0);
  const int64_t per_slot =
      8 * ((c->spin ? 5 : 7) * S2 + OO * OO + VV * VV + OO * VO + 8 * sm) +
      operands;
  double budget_gb = 120.0;
  if (const char *env = getenv("FTCC_WORKSPACE_GB")) budget_gb = atof(env);
  const int64_t budget = (int64_t)(budget_gb * (1LL << 30));
  const int64_t cap = std::max<int64_t>(1, budget / per_slot);
  int nb = (int)std::min<int64_t>(ng, cap);
  if (c->nb_fixed > 0)
    nb = (int)std::min<int64_t>(std::min<int64_t>(ng, c->nb_fixed), cap);
  return nb;
}

void ensure_workspace(ftcc_ctx *c) {
  if (c->aring) {
    if (c->ws_prec == c->prec) return;
    // a precision switch keeps every buffer: the chunk is what the
    // allocated slots and the operand arena hold (a mixed-precision solve
    // runs its FP64 steps at about half the low-precision chunk)
    int nb = std::min(chunk_for_budget(c), c->nb_alloc);
    if (c->arena)
      while (nb > 1 && operand_doubles(c, c->prec, nb) > c->arena_doubles)
        --nb;
    c->nb = nb;
    c->ws_prec = c->prec;
    return;
  }
  const int O = c->no, V = c->nv;
  const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V, OO = (int64_t)O * O;
  const int64_t S2 = VV * OO;
  const int64_t sm = std::max<int64_t>(VV, std::max<int64_t>(VO, OO));
  const int nb = chunk_for_budget(c);
  c->nb = nb;
  c->nb_alloc = nb;
  c->ws_prec = c->prec;
  c->aring = dalloc(c, S2 * nb);
  c->wr = dalloc(c, S2 * nb);
  c->z = dalloc(c, S2 * nb);
  c->s1 = dalloc(c, S2 * nb);
  c->s2 = dalloc(c, S2 * nb);
  c->wmnij = dalloc(c, OO * OO * nb);
  c->wabef = dalloc(c, VV * VV * nb);
  c->u = dalloc(c, OO * VO * nb);
  c->t1T = dalloc(c, VO * nb);
  c->Fme = dalloc(c, VO * nb);
  c->FmeT = dalloc(c, VO * nb);
  c->Fae = dalloc(c, VV * nb);
  c->Fmi = dalloc(c, OO * nb);
  c->Fbe = dalloc(c, VV * nb);
  c->Fmj = dalloc(c, OO * nb);
  c->ytmp = dalloc(c, sm * nb);
}

// Release the residual workspace and the operand copies sized by it (they
// are rebuilt on the next residual with the current chunk size).
void free_workspace(ftcc_ctx *c) {
  FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (double **p : {&c->tau, &c->t2r, &c->aring, &c->wr, &c->z,
                     &c->s1, &c->s2, &c->wmnij,
                     &c->wabef, &c->u, &c->t1T, &c->Fme, &c->FmeT, &c->Fae,
                     &c->Fmi, &c->Fbe, &c->Fmj, &c->ytmp}) {
    dfree(c, *p);
    *p = nullptr;
  }
  for (int id = 0; id < NOPS; ++id) {
    split_of(c, id) = SplitMat();
    c->b64[id] = nullptr;
  }
  dfree(c, c->arena);
  c->arena = nullptr;
  c->arena_doubles = 0;
  c->arena_owner = 0;
  c->split_layout = -1;
  c->b64_layout = -1;
  c->nb_alloc = 0;
  c->ws_prec = -1;
}

// Forget the captured step (its buffers or kernels may change).
void drop_graph(ftcc_ctx *c) {
  if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
  if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
  c->step_exec = nullptr;
  c->step_graph = nullptr;
  c->gwarm = false;
}

// Residual for `nb` consecutive slices starting at t1/t2 (slice strides
// V*O and V*V*O*O); writes t1bar/t2bar.
void residual_chunk(ftcc_ctx *c, const double *t1, const double *t2,
                    double *t1bar, double *t2bar, int nb) {
  const int O = c->no, V = c->nv;
  const int P64 = FTCC_PREC_FP64;  // small terms
  // O(n^5) contractions run in FP64 whatever the precision of the O(n^6)
  // products: they stream n^4 operands once, and on these skinny shapes
  // (N, K ~ n) the FP32 matrix-core variant of the GEMM measured 1.4-2x
  // slower than the FP64 one at n = 114
  const int P5 = FTCC_PREC_FP64;
  // O(n^5) contractions over spin-allowed blocks only (forbidden blocks of
  // their outputs are left unwritten; no consumer reads them)
  const bool s5 = c->spin && c->spin5;
  // Under the spin-flip symmetry (flip_now, spin5 implied) the spin-
  // restricted terms of FmeT, Fae, Fmi, Wr, Z and t1bar compute the blocks
  // of an alpha lead index only (flip_lead); once each tensor is complete
  // its beta blocks are copied from their images (the Fock-type matrices
  // and t1bar) or read from them in place (Wr, Z)
  const bool flip = c->flip_now;
  require(!flip || s5, "flip symmetry needs the spin-restricted terms");
  const int hv = c->v_count[0], ho = c->o_count[0];
  auto lead = [&](char l) { c->flip_lead = flip ? l : 0; };
  // ftcc_step's assembly under the flip symmetry reads nothing of a pair of
  // two beta virtuals (their R2 is written as the image of the alpha-alpha
  // pair): the beta-beta | beta-beta blocks of the ladder products, of
  // W_mnij and W_abef, and the beta-beta (a,b) of S1 and S2 are skipped
  const bool flip_step = flip && c->pack && c->tbar_scratch;
  c->flip_step_now = flip_step;
  auto sgem = [&](double alpha, double beta, const double *A, int64_t sA1,
                  const double *B, int64_t sB1, double *C, int64_t sC1,
                  const double *D) {
    Gemm g = mk(0, 0, 0, alpha, beta, A, 0, 0, sA1, B, 0, 0, sB1, C, 0, 0,
                sC1, nb);
    g.D = D;
    return g;
  };
  hipStream_t s = c->stream;
  const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V, OO = (int64_t)O * O;
  const int64_t S2 = VV * OO;           // doubles slice
  const int64_t O4 = OO * OO, V4 = VV * VV;
  const double *Foo = c->F[0], *Fov = c->F[1], *Fvo = c->F[2], *Fvv = c->F[3];
  const double *vvvv = c->I[0], *vvvo = c->I[1], *vovv = c->I[2],
               *vvoo = c->I[3], *oovv = c->I[5], *vooo = c->I[6],
               *ooov = c->I[7], *oooo = c->I[8];

  transpose2(t1, c->t1T, V, O, nb, s);   // t1T[m,f] = t1[f,m]
  // product operands (split into bf16 pairs, and/or reordered by spin
  // class): A as [M][K], B as [N][K] (K contiguous).  The spin layout
  // gathers the amplitude-derived ones straight from t1 and t2; the dense
  // layout reads them whole from their FP64 forms.
  ensure_fixed_operands(c);
  if (c->spin) {
    put_amp_operand(c, OP_TAU_A, AMP_TAU_AB, t1, t2, nb);
    put_amp_operand(c, OP_TAU_B, AMP_TAU_IJ, t1, t2, nb);
    put_amp_operand(c, OP_T2R_A, AMP_T2R, t1, t2, nb);
    put_amp_operand(c, OP_ARING_B, AMP_ARING, t1, t2, nb);
  } else {
    // dense layout (also a spin-blocked system's rare dense fallback):
    // tau and t2r are formed whole
    if (!c->tau) {
      c->tau = dalloc(c, S2 * c->nb_alloc);
      c->t2r = dalloc(c, S2 * c->nb_alloc);
    }
    PrepOut po{c->tau, c->t2r, c->aring};
    prep_amplitudes(t1, t2, V, O, nb, po, s);
    put_operand(c, OP_TAU_A, c->tau, OO, 1, S2, nb);
    put_operand(c, OP_TAU_B, c->tau, 1, OO, S2, nb);
    put_operand(c, OP_T2R_A, c->t2r, VO, 1, S2, nb);
    put_operand(c, OP_ARING_B, c->aring, 1, VO, S2, nb);
  }

  // Fme[m,e] = Fov[m,e] + sum_{fn} <mn||ef> t_n^f: FmeT[y][e,m] over the
  // grid points y as columns (oovv_ring[e,m,f,n] read once)
  if (s5) {
    fill(c->FmeT, VO * nb, 0.0, s);   // spin-forbidden entries stay zero
    lead('e');
    run_spin(c, sgem(1.0, 0.0, c->oovv_ring, 0, t1, VO, c->FmeT, VO,
                     nullptr),
             P5, {{'e', 'v', O * VO, 0, O}, {'m', 'o', VO, 0, 1}},
             {{'y', 'y', 0, VO, VO}},
             {{'f', 'v', O, O, 0}, {'n', 'o', 1, 1, 0}}, "mn|ef", "f|n",
             "e|m");
    lead(0);
    if (flip) flip_copy_mat(c->FmeT, V, O, hv, ho, nb, VO, s);
  } else {
    run(c, mk((int)VO, nb, (int)VO, 1.0, 0.0, c->oovv_ring, VO, 1, 0, t1, 1,
              VO, 0, c->FmeT, 1, VO, 0, 1),
        P5);
  }
  transpose2(c->FmeT, c->Fme, V, O, nb, s);
  bcast(c->Fme, Fov, VO, nb, 1.0, 1.0, s);
  transpose2(c->Fme, c->FmeT, O, V, nb, s);

  // Fae[a,e] = Fvv - 1/2 t1.Fov + sum t_m^f <ma||fe> - 1/2 taut.oovv with
  // taut = t2 + 1/2 (t1 t1 - t1 t1): its t1 t1 part contracted with oovv is
  // 1/2 t1.(Fme - Fov), so Fae = Fvv - 1/2 t1.Fme + sum t1.vovv
  // - 1/2 t2.oovv (and likewise for Fmi below): no taut is formed
  bcast(c->Fae, Fvv, VV, nb, 1.0, 0.0, s);
  run(c, mk(V, V, O, -0.5, 1.0, t1, O, 1, VO, c->Fme, V, 1, VO, c->Fae, V, 1,
            VV, nb), P64);
  // + sum_{mf} t_m^f <ma||fe> = - sum_{fm} t1[f,m] vovv[a,m,f,e], vovv[a,m,f,e]
  // = vovv_aefm[a,f,e,m]: rows y, columns (a,e), K = (f,m)
  if (s5) {
    lead('a');
    run_spin(c, sgem(-1.0, 1.0, t1, VO, c->vovv_aefm, 0, c->Fae, VV,
                     nullptr),
             P5, {{'y', 'y', VO, 0, VV}},
             {{'a', 'v', 0, V * VO, V}, {'e', 'v', 0, O, 1}},
             {{'f', 'v', O, VO, 0}, {'m', 'o', 1, 1, 0}}, "f|m", "am|fe",
             "a|e");
  } else {
    Gemm g = mk(nb, (int)VV, (int)VO, -1.0, 1.0, t1, VO, 1, 0, c->vovv_aefm,
                1, O, 0, c->Fae, VV, 1, 0, 1);
    g.N1 = V;        // columns (a, e)
    g.N2 = V;
    g.sBn1 = V * VO;
    g.sCn1 = V;
    g.K1 = V;        // k = f * O + m
    g.K2 = O;
    g.sAk1 = O;
    g.sBk1 = VO;
    run(c, g, P5);
  }
  // - 1/2 sum_{fmn} t2[a,f,m,n] <mn||ef> (oovvT[e,f,m,n]: K fastest)
  // (the chunk's grid points are rows of one launch: oovvT is read once)
  if (s5) {
    run_spin(c, sgem(-0.5, 1.0, t2, S2, c->oovvT, 0, c->Fae, VV,
                     nullptr),
             P5, {{'y', 'y', S2, 0, VV}, {'a', 'v', V * OO, 0, V}},
             {{'e', 'v', 0, V * OO, 1}},
             {{'f', 'v', OO, OO, 0}, {'m', 'o', O, O, 0}, {'n', 'o', 1, 1, 0}},
             "af|mn", "mn|ef", "a|e");
    lead(0);
    if (flip) flip_copy_mat(c->Fae, V, V, hv, hv, nb, VV, s);
  } else
    run(c, fold_m(mk(V, V, (int)(V * OO), -0.5, 1.0, t2, V * OO, 1, S2,
                     c->oovvT, 1, V * OO, 0, c->Fae, V, 1, VV, nb)),
        P5);

  // Fmi[m,i] = Foo + 1/2 Fov.t1 + sum t_n^e <mn||ie> + 1/2 oovv.taut
  //         = Foo + 1/2 Fme.t1 + sum t_n^e <mn||ie> + 1/2 oovv.t2
  bcast(c->Fmi, Foo, OO, nb, 1.0, 0.0, s);
  run(c, mk(O, O, V, 0.5, 1.0, c->Fme, V, 1, VO, t1, O, 1, VO, c->Fmi, O, 1,
            OO, nb), P64);
  // + sum_{ne} <mn||ie> t_n^e: ooov_p1[m,i,n,e] rows (m,i), columns y
  if (s5) {
    lead('m');
    run_spin(c, sgem(1.0, 1.0, c->ooov_p1, 0, c->t1T, VO, c->Fmi, OO,
                     nullptr),
             P5, {{'m', 'o', O * VO, 0, O}, {'i', 'o', VO, 0, 1}},
             {{'y', 'y', 0, VO, OO}},
             {{'n', 'o', V, V, 0}, {'e', 'v', 1, 1, 0}}, "mn|ie", "e|n",
             "m|i");
  } else
    run(c, mk((int)OO, nb, (int)VO, 1.0, 1.0, c->ooov_p1, VO, 1, 0, c->t1T, 1,
              VO, 0, c->Fmi, 1, OO, 0, 1),
        P5);
  // + 1/2 sum_{nef} <mn||ef> t2[e,f,i,n] over K = (e, f, n), n fastest in
  // both (oovv_mefn[m,e,f,n]); the grid points are columns of one launch
  if (s5) {
    run_spin(c, sgem(0.5, 1.0, c->oovv_mefn, 0, t2, S2, c->Fmi, OO,
                     nullptr),
             P5, {{'m', 'o', V * VO, 0, O}},
             {{'y', 'y', 0, S2, OO}, {'i', 'o', 0, O, 1}},
             {{'e', 'v', VO, V * OO, 0}, {'f', 'v', O, OO, 0},
              {'n', 'o', 1, 1, 0}},
             "mn|ef", "ef|in", "m|i");
    lead(0);
    if (flip) flip_copy_mat(c->Fmi, O, O, ho, ho, nb, OO, s);
  } else {
    Gemm g = fold_n(mk(O, O, (int)(O * VV), 0.5, 1.0, c->oovv_mefn, V * VO, 1,
                       0, t2, 1, O, S2, c->Fmi, O, 1, OO, nb));
    g.K1 = (int)VV;   // k = (e,f) * O + n
    g.K2 = O;
    g.sAk1 = O;
    g.sBk1 = OO;
    run(c, g, P5);
  }

  // Fbe = Fae - 1/2 t1.Fme ; Fmj = Fmi + 1/2 Fme.t1
  axpby(c->Fbe, c->Fae, VV * nb, 1.0, 0.0, s);
  run(c, mk(V, V, O, -0.5, 1.0, t1, O, 1, VO, c->Fme, V, 1, VO, c->Fbe, V, 1,
            VV, nb), P64);
  axpby(c->Fmj, c->Fmi, OO * nb, 1.0, 0.0, s);
  run(c, mk(O, O, V, 0.5, 1.0, c->Fme, V, 1, VO, t1, O, 1, VO, c->Fmj, O, 1,
            OO, nb), P64);

  // Each W intermediate starts with its O(n^6) product written with
  // beta = 0 (the split-bf16 GEMM then never reads C, whose latency it
  // cannot hide); the O(n^5) terms and the integral (an addend D) accumulate
  // onto it in the FP64 GEMM.  In the spin layout the product writes the
  // allowed blocks only, and nothing reads the others.
  // Wmnij = oooo + P(ij) t_j^e <mn||ie> + 1/2 oovv.tau          (G1)
  c->ladder_skip_bb = flip_step;
  product(c, mk((int)OO, (int)OO, (int)VV, 0.5, 0.0, oovv, VV, 1, 0, c->tau,
                OO, 1, S2, c->wmnij, OO, 1, O4, nb), OP_OOVV_A, OP_TAU_B,
          LADDER);
  c->ladder_skip_bb = false;
  // + oooo + sum_e <mn||ie> t_j^e into [(m,n,i), j]; with packed pairs
  // only the stored (m,n) of each couple (all the G2 operand reads)
  if (s5) {
    const std::vector<SIdx> mn =
        c->pack ? std::vector<SIdx>{{'m', 'P', O * V, 0, OO, 'n'}}
                : std::vector<SIdx>{{'m', 'o', OO * V, 0, O * OO},
                                    {'n', 'o', O * V, 0, OO}};
    std::vector<SIdx> Mi = mn;
    Mi.push_back({'i', 'o', V, 0, O});
    // G2 reads the stored (i,j) of each couple only: no (i beta, j alpha)
    if (c->pack) c->skip_ba[0] = 'i', c->skip_ba[1] = 'j';
    c->flip_lead = flip_step ? 'i' : 0;   // i beta: (m,n) and (i,j) beta-beta
    run_spin(c, sgem(1.0, 1.0, ooov, 0, t1, VO, c->wmnij, O4, oooo), P5, Mi,
             {{'y', 'y', 0, VO, O4}, {'j', 'o', 0, 1, 1}}, {'e', 'v', 1, O, 0},
             "mn|ie", "e|j", "mn|ij");
    // - sum_e t_i^e <mn||je>: C[i][(m,n,j)] -= t1^T[i,e] ooov[m,n,j,e]
    std::vector<SIdx> Ni;
    for (SIdx x : mn) {   // the same (m,n) as columns: strides of B and C
      x.sb = x.sa;
      x.sa = 0;
      Ni.push_back(x);
    }
    Ni.push_back({'j', 'o', 0, V, 1});
    run_spin(c, sgem(-1.0, 1.0, t1, VO, ooov, 0, c->wmnij, O4, nullptr), P5,
             {{'y', 'y', VO, 0, O4}, {'i', 'o', 1, 0, O}}, Ni,
             {'e', 'v', O, 1, 0}, "e|i", "mn|je", "mn|ij");
    c->skip_ba[0] = c->skip_ba[1] = 0;
    c->flip_lead = 0;
  } else {
    Gemm g = mk((int)(OO * O), O, V, 1.0, 1.0, ooov, V, 1, 0, t1, O, 1, VO,
                c->wmnij, O, 1, O4, nb);
    g.D = oooo;
    g.sDm = O;
    g.sDn = 1;
    run(c, g, P5);
    // - sum_e t_i^e <mn||je>: per (m,n), C[i][j] -= t1^T[i,e] ooov[m,n,j,e]
    g = mk(O, O, V, -1.0, 1.0, t1, 1, O, VO, ooov, 1, V, 0, c->wmnij, O, 1,
           O4, nb);
    g.nb2 = (int)OO;
    g.sA2 = 0;
    g.sB2 = O * V;
    g.sC2 = OO;
    run(c, g, P5);
  }

  // Wabef = vvvv - P(ab) t_m^b <am||ef> (its tau.oovv share is in Wmnij);
  // with packed pairs only the stored (e,f) of each couple
  if (s5) {
    const std::vector<SIdx> ef =
        c->pack ? std::vector<SIdx>{{'e', 'Q', 0, 1, 1, 'f'}}
                : std::vector<SIdx>{{'e', 'v', 0, V, V}, {'f', 'v', 0, 1, 1}};
    // vvvv - sum_m t_m^b <am||ef>: C[b][(a,e,f)] = vvvv - t1[b,m] vovv[a,m,ef]
    std::vector<SIdx> Ni{{'a', 'v', 0, O * VV, V * VV}};
    Ni.insert(Ni.end(), ef.begin(), ef.end());
    // G3 reads the stored (a,b) of each couple only: no (a beta, b alpha)
    if (c->pack) c->skip_ba[0] = 'a', c->skip_ba[1] = 'b';
    c->flip_lead = flip_step ? 'a' : 0;   // a beta: (a,b) and (e,f) beta-beta
    run_spin(c, sgem(-1.0, 0.0, t1, VO, vovv, 0, c->wabef, V4, vvvv), P5,
             {{'y', 'y', VO, 0, V4}, {'b', 'v', O, 0, VV}}, Ni,
             {'m', 'o', 1, VV, 0}, "b|m", "am|ef", "ab|ef");
    // + sum_m t_m^a <bm||ef>: C[a][(b,e,f)]
    Ni = {{'b', 'v', 0, O * VV, VV}};
    Ni.insert(Ni.end(), ef.begin(), ef.end());
    run_spin(c, sgem(1.0, 1.0, t1, VO, vovv, 0, c->wabef, V4, nullptr), P5,
             {{'y', 'y', VO, 0, V4}, {'a', 'v', O, 0, V * VV}}, Ni,
             {'m', 'o', 1, VV, 0}, "a|m", "bm|ef", "ab|ef");
    c->skip_ba[0] = c->skip_ba[1] = 0;
    c->flip_lead = 0;
  } else {
    // vvvv - sum_m t_m^b <am||ef>: per a, C[b][ef] = vvvv - t1[b,m] vovv[a,m,ef]
    Gemm g = mk(V, (int)VV, O, -1.0, 0.0, t1, O, 1, VO, vovv, VV, 1, 0,
                c->wabef, VV, 1, V4, nb);
    g.nb2 = V;
    g.sA2 = 0;
    g.sB2 = O * VV;
    g.sC2 = V * VV;
    g.D = vvvv;
    g.sDm = VV;
    g.sDn = 1;
    g.sD2 = V * VV;
    run(c, g, P5);
    g.D = nullptr;
    // + sum_m t_m^a <bm||ef>: per b, rows a of W[a,b,ef]
    g.alpha = 1.0;
    g.beta = 1.0;
    g.sCm = V * VV;
    g.sC2 = VV;
    run(c, g, P5);
  }

  // Wr[e,m,b,j] = <mb||ej> + t_j^f <mb||ef> - t_n^b <mn||ej> - A.<mn||ef>
  product(c, mk((int)VO, (int)VO, (int)VO, -1.0, 0.0, c->oovv_ring, VO, 1, 0,
                c->aring, VO, 1, S2, c->wr, VO, 1, S2, nb), OP_RING_A,
          OP_ARING_B, RING_G4);                                          // G4
  // <mb||ej> + t_j^f <mb||ef> = ovvo_ring - sum_f vovv[b,m,e,f] t1[f,j]:
  // rows (e,m,b) of Wr (like G4 and G5, over the rows of an alpha e, a
  // under the flip symmetry)
  if (s5) {
    lead('e');
    run_spin(c, sgem(-1.0, 1.0, c->vovv_embf, 0, t1, VO, c->wr, S2,
                     c->ovvo_ring),
             P5,
             {{'e', 'v', O * VV, 0, O * VO}, {'m', 'o', VV, 0, VO},
              {'b', 'v', V, 0, O}},
             {{'y', 'y', 0, VO, S2}, {'j', 'o', 0, 1, 1}}, {'f', 'v', 1, O, 0},
             "bm|ef", "f|j", "mb|ej");
    // -t_n^b <mn||ej> = sum_n t1[b,n] ooov[m,n,j,e]: C[b][(e,m,j)]
    run_spin(c, sgem(1.0, 1.0, t1, VO, c->ooov_emnj, 0, c->wr, S2, nullptr),
             P5, {{'y', 'y', VO, 0, S2}, {'b', 'v', O, 0, O}},
             {{'e', 'v', 0, O * OO, O * VO}, {'m', 'o', 0, OO, VO},
              {'j', 'o', 0, 1, 1}},
             {'n', 'o', 1, O, 0}, "b|n", "mn|je", "mb|ej");
    c->flip_lead = 0;
  } else {
    Gemm g = mk((int)(V * VO), O, V, -1.0, 1.0, c->vovv_embf, V, 1, 0, t1, O,
                1, VO, c->wr, O, 1, S2, nb);
    g.D = c->ovvo_ring;
    g.sDm = O;
    g.sDn = 1;
    run(c, g, P5);
  }
  if (!s5) {
    // -t_n^b <mn||ej> = sum_n t1[b,n] ooov[m,n,j,e]: per (e,m), C[b][j]
    Gemm g = mk(V, O, O, 1.0, 1.0, t1, O, 1, VO, c->ooov_emnj, O, 1, 0, c->wr,
                O, 1, S2, nb);
    g.nb2 = (int)VO;
    g.sA2 = 0;
    g.sB2 = OO;
    g.sC2 = VO;
    run(c, g, P5);
  }

  // Z[ai,bj] = t2r.Wr - t_i^e t_m^a <mb||ej>                    (G5)
  // (under the flip symmetry the gather reads Wr's beta rows from their
  // images: nothing else reads them)
  if (flip)
    put_flip_wr(c, c->wr, S2, nb);
  else
    put_operand(c, OP_WR_B, c->wr, 1, VO, S2, nb);
  // spin blocks leave the spin-forbidden part of Z unwritten; the
  // spin-aware assembly below never reads it
  product(c, mk((int)VO, (int)VO, (int)VO, 1.0, 0.0, c->t2r, VO, 1, S2, c->wr,
                VO, 1, S2, c->z, VO, 1, S2, nb), OP_T2R_A, OP_WR_B, RING_G5);
  if (s5) {
    // u[i,(m,b,j)] = t1[e,i] ovvo_ring[e,m,b,j]; Z[a,(i,b,j)] -= t1[a,m] u
    // (Z rows of an alpha a read the u of an alpha m only)
    lead('m');
    run_spin(c, sgem(1.0, 0.0, t1, VO, c->ovvo_ring, 0, c->u, OO * VO,
                     nullptr),
             P5, {{'y', 'y', VO, 0, OO * VO}, {'i', 'o', 1, 0, O * VO}},
             {{'m', 'o', 0, VO, VO}, {'b', 'v', 0, O, O}, {'j', 'o', 0, 1, 1}},
             {'e', 'v', O, O * VO, 0}, "e|i", "mb|ej", "mb|ij");
    lead('a');
    run_spin(c, sgem(-1.0, 1.0, t1, VO, c->u, OO * VO, c->z, S2, nullptr), P5,
             {{'a', 'v', O, 0, O * VO}},
             {{'i', 'o', 0, O * VO, VO}, {'b', 'v', 0, O, O},
              {'j', 'o', 0, 1, 1}},
             {'m', 'o', 1, VO, 0}, "a|m", "mb|ij", "ab|ij");
    c->flip_lead = 0;
    // the packed assembly reads the beta rows of Z from their images
    if (flip && !c->pack) flip_rows(c, c->z, nb, true);
  } else {
    run(c, mk(O, (int)(O * VO), V, 1.0, 0.0, t1, 1, O, VO, c->ovvo_ring,
              O * VO, 1, 0, c->u, O * VO, 1, OO * VO, nb),
        P5);
    Gemm g = mk(V, (int)VO, O, -1.0, 1.0, t1, O, 1, VO, c->u, VO, 1, OO * VO,
                c->z, O * VO, 1, S2, nb);
    g.nb2 = O;
    g.sA2 = 0;
    g.sB2 = O * VO;
    g.sC2 = VO;
    run(c, g, P5);
  }

  // doubles: G2 = 1/2 tau.Wmnij into wr (free after the G5 operand was
  // gathered) and G3 = 1/2 Wabef.tau into aring (free after G4), both with
  // beta = 0 (with packed pairs: one pair of each antisymmetric couple);
  // the assembly adds them, vvoo and the rest into t2bar
  put_operand(c, OP_WMNIJ_B, c->wmnij, 1, OO, O4, nb);
  put_operand(c, OP_WABEF_A, c->wabef, VV, 1, V4, nb);
  c->ladder_skip_bb = flip_step;
  product(c, mk((int)VV, (int)OO, (int)OO, 0.5, 0.0, c->tau, OO, 1, S2,
                c->wmnij, OO, 1, O4, c->wr, OO, 1, S2, nb), OP_TAU_A,
          OP_WMNIJ_B, LADDER);                                           // G2
  product(c, mk((int)VV, (int)OO, (int)VV, 0.5, 0.0, c->wabef, VV, 1, V4,
                c->tau, OO, 1, S2, c->aring, OO, 1, S2, nb), OP_WABEF_A,
          OP_TAU_B, LADDER);                                             // G3
  c->ladder_skip_bb = false;

  // S1 (antisymmetrised in ab): Fbe[b,e] t2[a,e,i,j] + t1[a,m] vooo[b,m,i,j]
  // S2 (antisymmetrised in ij): -t2[a,b,i,m] Fmj[m,j] + t1[e,i] vvvo[a,b,e,j]
  // S1 is antisymmetric in (i,j) and S2 in (a,b): with packed pairs each is
  // formed for the stored pair of that couple only
  if (s5) {
    const std::vector<SIdx> ij =
        c->pack ? std::vector<SIdx>{{'i', 'P', 0, 1, 1, 'j'}}
                : std::vector<SIdx>{{'i', 'o', 0, O, O}, {'j', 'o', 0, 1, 1}};
    std::vector<SIdx> Ni{{'a', 'v', 0, V * OO, V * OO}};
    Ni.insert(Ni.end(), ij.begin(), ij.end());
    if (flip_step) c->skip_bb[0] = 'a', c->skip_bb[1] = 'b';
    run_spin(c, sgem(1.0, 0.0, c->Fbe, VV, t2, S2, c->s1, S2, nullptr), P5,
             {{'b', 'v', V, 0, OO}}, Ni, {'e', 'v', 1, OO, 0}, "b|e",
             "ae|ij", "ab|ij");
    Ni = {{'b', 'v', 0, O * OO, OO}};
    Ni.insert(Ni.end(), ij.begin(), ij.end());
    run_spin(c, sgem(1.0, 1.0, t1, VO, vooo, 0, c->s1, S2, nullptr), P5,
             {{'y', 'y', VO, 0, S2}, {'a', 'v', O, 0, V * OO}}, Ni,
             {'m', 'o', 1, OO, 0}, "a|m", "bm|ij", "ab|ij");
    std::vector<SIdx> Mi =
        c->pack ? std::vector<SIdx>{{'a', 'Q', OO, 0, OO, 'b'}}
                : std::vector<SIdx>{{'a', 'v', V * OO, 0, V * OO},
                                    {'b', 'v', OO, 0, OO}};
    Mi.push_back({'i', 'o', O, 0, O});
    run_spin(c, sgem(-1.0, 0.0, t2, S2, c->Fmj, OO, c->s2, S2, nullptr), P5,
             Mi, {{'j', 'o', 0, 1, 1}}, {'m', 'o', 1, O, 0}, "ab|im", "m|j",
             "ab|ij");
    Ni = c->pack ? std::vector<SIdx>{{'a', 'Q', 0, VO, OO, 'b'}}
                 : std::vector<SIdx>{{'a', 'v', 0, V * VO, V * OO},
                                     {'b', 'v', 0, VO, OO}};
    Ni.push_back({'j', 'o', 0, 1, 1});
    run_spin(c, sgem(1.0, 1.0, t1, VO, vvvo, 0, c->s2, S2, nullptr), P5,
             {{'y', 'y', VO, 0, S2}, {'i', 'o', 1, 0, O}}, Ni,
             {'e', 'v', O, O, 0}, "e|i", "ab|ej", "ab|ij");
    c->skip_bb[0] = c->skip_bb[1] = 0;
  } else {
    Gemm g = mk(V, (int)OO, V, 1.0, 0.0, c->Fbe, V, 1, VV, t2, OO, 1, S2,
                c->s1, OO, 1, S2, nb);
    g.nb2 = V;
    g.sA2 = 0;
    g.sB2 = V * OO;
    g.sC2 = V * OO;
    run(c, g, P5);
    g = mk(V, (int)OO, O, 1.0, 1.0, t1, O, 1, VO, vooo, OO, 1, 0, c->s1,
           V * OO, 1, S2, nb);
    g.nb2 = V;
    g.sA2 = 0;
    g.sB2 = O * OO;
    g.sC2 = OO;
    run(c, g, P5);
    run(c, mk((int)(VV * O), O, O, -1.0, 0.0, t2, O, 1, S2, c->Fmj, O, 1, OO,
              c->s2, O, 1, S2, nb),
        P5);
    g = mk(O, O, V, 1.0, 1.0, t1, 1, O, VO, vvvo, O, 1, 0, c->s2, O, 1, S2,
           nb);
    g.nb2 = (int)VV;
    g.sA2 = 0;
    g.sB2 = VO;
    g.sC2 = OO;
    run(c, g, P5);
  }
  // forbidden blocks (spin layout): written as zero without reading
  const bool zflip = flip && c->pack;
  // algorithmic bytes per spin-allowed double: one read of each of the six
  // inputs and one write (56); with packed pairs the four packed inputs
  // hold one of each antisymmetric couple (40); the flip-symmetric step
  // reads them for the pairs it computes, half of them, and writes every
  // double (24, the measured traffic of that form)
  double allowed = (double)S2;
  if (c->spin5) {
    allowed = 0.0;
    for (int box = 0; box < 16; ++box) {
      const int sa = box >> 3, sb = (box >> 2) & 1, si = (box >> 1) & 1,
                sj = box & 1;
      if (sa + sb == si + sj)
        allowed += (double)c->v_count[sa] * c->v_count[sb] *
                   c->o_count[si] * c->o_count[sj];
    }
  }
  const double per_double =
      flip_step ? 24.0 : (c->spin && c->pack ? 40.0 : 56.0);
  timed_class(c, ftcc_ctx::PROF_ASSEMBLY, 0.0, per_double * allowed * nb,
              [&] {
    r2_final(t2bar, c->wr, c->aring, c->s1, c->s2, c->z, V, O, nb, s,
             c->spin ? c->d_lab_v : nullptr, c->spin ? c->d_lab_o : nullptr,
             vvoo, c->spin && c->pack, zflip ? hv : 0, zflip ? ho : 0,
             !(c->tbar_scratch && c->spin && c->pack), c->d_jl, c->jl_cnt[0],
             c->jl_cnt[1]);
  });

  // singles: t1bar = -(Fvo + Fae.t1 - t1.Fmi + t2r.FmeT + ovvo_ring^T.t1
  //                    + 1/2 vovv.t2 + 1/2 t2.ooov)
  bcast(t1bar, Fvo, VO, nb, 1.0, 0.0, s);
  run(c, mk(V, O, V, 1.0, 1.0, c->Fae, V, 1, VV, t1, O, 1, VO, t1bar, O, 1, VO,
            nb), P64);
  run(c, mk(V, O, O, -1.0, 1.0, t1, O, 1, VO, c->Fmi, O, 1, OO, t1bar, O, 1,
            VO, nb), P64);
  // + sum_{em} t2[a,e,i,m] FmeT[e,m] (per grid point) and
  // + sum_{em} t1[e,m] ovvo_ring[e,m,a,i] (ovvo_ringT[a,i,e,m]; rows y)
  if (s5) {
    lead('a');
    run_spin(c, sgem(1.0, 1.0, t2, S2, c->FmeT, VO, t1bar, VO, nullptr), P5,
             {{'a', 'v', V * OO, 0, O}, {'i', 'o', O, 0, 1}},
             {{'_', '1', 0, 0, 0}}, {{'e', 'v', OO, O, 0}, {'m', 'o', 1, 1, 0}},
             "ae|im", "e|m", "a|i");
    run_spin(c, sgem(1.0, 1.0, t1, VO, c->ovvo_ringT, 0, t1bar, VO, nullptr),
             P5, {{'y', 'y', VO, 0, VO}},
             {{'a', 'v', 0, O * VO, O}, {'i', 'o', 0, VO, 1}},
             {{'e', 'v', O, O, 0}, {'m', 'o', 1, 1, 0}}, "e|m", "ma|ei",
             "a|i");
  } else {
    // rows (a,i) and K = (e,m) over two index levels each
    Gemm g = mk((int)VO, 1, (int)VO, 1.0, 1.0, t2, O, 1, S2, c->FmeT, 1, 0, VO,
                t1bar, 1, 0, VO, nb);
    g.M1 = V;
    g.M2 = O;
    g.sAm1 = V * OO;
    g.sCm1 = O;
    g.K1 = V;
    g.K2 = O;
    g.sAk1 = OO;
    g.sBk1 = O;
    run(c, g, P5);
    run(c, mk(nb, (int)VO, (int)VO, 1.0, 1.0, t1, VO, 1, 0, c->ovvo_ringT, 1,
              VO, 0, t1bar, VO, 1, 0, 1),
        P5);
  }
  // + 1/2 sum_{mef} <am||ef> t2[e,f,i,m] over K = (e, f, m), m fastest in
  // both (vovv_aefm[a,e,f,m])
  // + 1/2 sum_{emn} t2[a,e,m,n] <nm||ie> (ooov_iemn[i,e,m,n])
  // (the grid points are columns, resp. rows, of one launch)
  if (s5) {
    run_spin(c, sgem(0.5, 1.0, c->vovv_aefm, 0, t2, S2, t1bar, VO, nullptr),
             P5, {{'a', 'v', V * VO, 0, O}},
             {{'y', 'y', 0, S2, VO}, {'i', 'o', 0, O, 1}},
             {{'e', 'v', VO, V * OO, 0}, {'f', 'v', O, OO, 0},
              {'m', 'o', 1, 1, 0}},
             "am|ef", "ef|im", "a|i");
    run_spin(c, sgem(0.5, 1.0, t2, S2, c->ooov_iemn, 0, t1bar, VO, nullptr),
             P5, {{'y', 'y', S2, 0, VO}, {'a', 'v', V * OO, 0, O}},
             {{'i', 'o', 0, V * OO, 1}},
             {{'e', 'v', OO, OO, 0}, {'m', 'o', O, O, 0}, {'n', 'o', 1, 1, 0}},
             "ae|mn", "nm|ie", "a|i");
    lead(0);
    if (flip) flip_copy_mat(t1bar, V, O, hv, ho, nb, VO, s);
  } else {
    Gemm g = fold_n(mk(V, O, (int)(O * VV), 0.5, 1.0, c->vovv_aefm, V * VO, 1,
                       0, t2, 1, O, S2, t1bar, O, 1, VO, nb));
    g.K1 = (int)VV;   // k = (e,f) * O + m
    g.K2 = O;
    g.sAk1 = O;
    g.sBk1 = OO;
    run(c, g, P5);
    run(c, fold_m(mk(V, O, (int)(V * OO), 0.5, 1.0, t2, V * OO, 1, S2,
                     c->ooov_iemn, 1, V * OO, 0, t1bar, O, 1, VO, nb)),
        P5);
  }
  axpby(t1bar, t1bar, VO * nb, -1.0, 0.0, s);
}

void check_grid_D(ftcc_ctx *c) {
  require(c->have_energies || c->have_D,
          "denominators not set (ftcc_set_system/ftcc_set_denominators)",
          FTCC_ESTATE);
  require(c->grid.ng == c->ng, "grid not set (ftcc_set_grid)", FTCC_ESTATE);
}

void check_ready(ftcc_ctx *c) {
  require(c->have_ints, "integrals not set (ftcc_set_system/ftcc_set_blocks)",
          FTCC_ESTATE);
  require(c->have_energies || c->have_D,
          "denominators not set (ftcc_set_system/ftcc_set_denominators)",
          FTCC_ESTATE);
  require(c->grid.ng == c->ng, "grid not set (ftcc_set_grid)", FTCC_ESTATE);
}

void upload_g(ftcc_ctx *c, const double *g) {
  FTCC_CHECK_HIP(hipMemcpyAsync(c->grid.d_g, g, 8 * c->ng,
                                hipMemcpyHostToDevice, c->stream));
}

// true when all ng slices of T1 [ng][V][O] and T2 [ng][V][V][O][O] conserve
// spin under the labels (the blocked products rely on it)
bool amplitudes_conserve_spin(ftcc_ctx *c, const double *T1,
                              const double *T2, int nslices) {
  const int O = c->no, V = c->nv;
  hipStream_t s = c->stream;
  double bad = 0.0;
  int nbk = spin_violations(T1, V, 1, O, 1, nslices, c->d_lab_v, nullptr,
                            c->d_lab_o, nullptr, c->partials, s);
  reduce_partials(c->partials, nbk, c->red, s);
  read_scalars(c, 3);
  bad += c->host_red[2];
  nbk = spin_violations(T2, V, V, O, O, nslices, c->d_lab_v, c->d_lab_v,
                        c->d_lab_o, c->d_lab_o, c->partials, s);
  reduce_partials(c->partials, nbk, c->red, s);
  read_scalars(c, 3);
  bad += c->host_red[2];
  return bad == 0.0;
}

// The trusted promise, checked once on the first amplitudes after it is
// given: every slice conserves spin under the labels and, for a spin-flip
// symmetric system, equals its flipped image (to 1e-10 relative: iterates of
// an untrusted run are symmetric up to rounding).  A broken promise is
// withdrawn and reported.
void verify_trust(ftcc_ctx *c, const double *T1, const double *T2) {
  c->trust_pending = false;
  if (!c->spin) return;   // dense layout: nothing relies on the promise
  hipStream_t s = c->stream;
  if (s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    FTCC_CHECK_HIP(hipStreamIsCapturing(s, &st));
    require(st == hipStreamCaptureStatusNone,
            "ftcc_set_trusted: the first residual after the promise checks "
            "the amplitudes and cannot be captured");
  }
  if (!amplitudes_conserve_spin(c, T1, T2, c->ng)) {
    c->trusted = false;
    throw Error(FTCC_EINVAL,
                "ftcc_set_trusted: the amplitudes do not conserve spin under "
                "the orbital labels (promise withdrawn)");
  }
  if (!c->flip) return;
  const int O = c->no, V = c->nv;
  const int ho = c->o_count[0], hv = c->v_count[0];
  double bad = 0.0;
  int nbk = flip_violations(T1, V, O, 1, 1, hv, ho, 0, 0, c->partials, s,
                            c->ng, 1e-10);
  reduce_partials(c->partials, nbk, c->red, s);
  read_scalars(c, 3);
  bad += c->host_red[2];
  nbk = flip_violations(T2, V, V, O, O, hv, hv, ho, ho, c->partials, s, c->ng,
                        1e-10);
  reduce_partials(c->partials, nbk, c->red, s);
  read_scalars(c, 3);
  bad += c->host_red[2];
  if (bad != 0.0) {
    c->trusted = false;
    throw Error(FTCC_EINVAL,
                "ftcc_set_trusted: the amplitudes are not symmetric under the "
                "spin flip of this flip-symmetric system (" +
                    std::to_string((long long)bad) +
                    " entries differ; promise withdrawn)");
  }
}

// Returns whether the spin layout was used (the amplitudes conserve spin).
bool residual_all(ftcc_ctx *c, const double *T1, const double *T2,
                  double *T1b, double *T2b, uint64_t skip) {
  ensure_workspace(c);
  if (c->trusted && c->trust_pending) verify_trust(c, T1, T2);
  // spin blocks are exact only for spin-conserving amplitudes (every CC
  // iterate is); anything else runs dense
  const int spin_saved = c->spin;
  if (c->spin && !c->trusted && !amplitudes_conserve_spin(c, T1, T2, c->ng))
    c->spin = 0;
  // the flip symmetry of the amplitudes is part of the trusted promise
  c->flip_now = c->flip && c->spin && c->trusted;
  c->last_flip_step = c->flip_now && c->spin5 && c->pack && c->tbar_scratch;
  struct Restore {
    ftcc_ctx *c;
    int v;
    ~Restore() {
      c->spin = v;
      c->flip_now = false;
      c->flip_lead = 0;
      c->skip_ba[0] = c->skip_ba[1] = 0;
      c->skip_bb[0] = c->skip_bb[1] = 0;
      c->ladder_skip_bb = false;
      c->flip_step_now = false;
    }
  } restore{c, spin_saved};
  const int O = c->no, V = c->nv;
  const int64_t VO = (int64_t)V * O, S2 = (int64_t)V * V * O * O;
  int y = 0;
  while (y < c->ng) {
    if (y < 64 && (skip >> y) & 1) {
      bcast(T1b + y * VO, c->F[2], VO, 1, -1.0, 0.0, c->stream);
      bcast(T2b + y * S2, c->I[3], S2, 1, -1.0, 0.0, c->stream);
      ++y;
      continue;
    }
    int nb = 0;
    while (y + nb < c->ng && nb < c->nb &&
           !((y + nb) < 64 && (skip >> (y + nb)) & 1))
      ++nb;
    residual_chunk(c, T1 + y * VO, T2 + y * S2, T1b + y * VO, T2b + y * S2,
                   nb);
    y += nb;
  }
  return c->spin != 0;
}

UpdateArgs upd(ftcc_ctx *c) {
  UpdateArgs a;
  a.ng = c->ng;
  a.nv = c->nv;
  a.no = c->no;
  a.ti = c->grid.d_ti;
  a.G = c->grid.d_G;
  a.g = c->grid.d_g;
  a.eo = c->eo;
  a.ev = c->ev;
  a.D = nullptr;
  a.tbar = nullptr;
  a.told = nullptr;
  a.damp = 0.0;
  a.t1new = nullptr;
  a.eint = nullptr;
  a.partials = c->partials;
  a.src = nullptr;
  a.uniform = c->grid.uniform ? 1 : 0;
  a.delta = c->grid.delta;
  a.rec = c->grid.rec;
  a.rc = c->grid.d_rc;
  return a;
}

void read_scalars(ftcc_ctx *c, int n) {
  FTCC_CHECK_HIP(hipMemcpyAsync(c->host_red, c->red, sizeof(double) * n,
                                hipMemcpyDeviceToHost, c->stream));
  FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  flush_profile(c);
}

}  // namespace ftcc_core

using namespace ftcc_core;

extern "C" {

int ftcc_create(ftcc_ctx **out, int no, int nv, int ng, int precision,
                void *stream) {
  if (!out) return FTCC_EINVAL;
  *out = nullptr;
  ftcc_ctx *c = new ftcc_ctx();
  int rc = guarded(c, [&] {
    require(no > 0 && nv > 0 && ng > 0, "no, nv and ng must be positive");
    require(precision >= 0 && precision <= 2, "unknown precision");
    c->no = no;
    c->nv = nv;
    c->ng = ng;
    c->prec = precision;
    c->stream = (hipStream_t)stream;
    const int64_t VO = (int64_t)nv * no;
    // integrals
    // split-K / GEMV partial sums: start with a modest buffer, grown on use
    c->scr.partial_elems = (size_t)std::max<int64_t>(1 << 20, 64 * VO);
    c->scr.partial = dalloc(c, (int64_t)c->scr.partial_elems);
    c->partials = dalloc(c, 3 * MAX_PARTIAL_BLOCKS);
    c->red = dalloc(c, 16);
    FTCC_CHECK_HIP(hipHostMalloc((void **)&c->host_red, 16 * sizeof(double),
                                 hipHostMallocDefault));
    // the allocation-free entry point's scratch on this device
    int device = 0;
    FTCC_CHECK_HIP(hipGetDevice(&device));
    noalloc_scratch(device);
  });
  if (rc != FTCC_OK) {
    fprintf(stderr, "ftcc_create: %s\n", c->err.c_str());
    ftcc_destroy(c);
    return rc;
  }
  *out = c;
  return FTCC_OK;
}

int ftcc_destroy(ftcc_ctx *c) {
  if (!c) return FTCC_OK;
  (void)hipStreamSynchronize(c->stream);
  adjoint_release(c);
  for (auto &p : c->prof_pending) {
    (void)hipEventDestroy(p.e0);
    (void)hipEventDestroy(p.e1);
  }
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
  if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->grid.d_ti) (void)hipFree(c->grid.d_ti);
  if (c->grid.d_g) (void)hipFree(c->grid.d_g);
  if (c->grid.d_G) (void)hipFree(c->grid.d_G);
  if (c->grid.d_rc) (void)hipFree(c->grid.d_rc);
  if (c->host_red) (void)hipHostFree(c->host_red);
  delete c;
  return FTCC_OK;
}

const char *ftcc_last_error(const ftcc_ctx *c) {
  return c ? c->err.c_str() : "null context";
}

int ftcc_set_precision(ftcc_ctx *c, int precision) {
  return guarded(c, [&] {
    drop_graph(c);   // buffers or kernels of the captured step change
    require(precision >= 0 && precision <= 2, "unknown precision");
    if (precision == c->prec) return;
    // the workspace stays: the next residual carves this precision's
    // operands from the same arena and picks the chunk that fits
    c->prec = precision;
  });
}

int ftcc_set_system(ftcc_ctx *c, int n, const double *en, const double *fock,
                    const double *eri, double beta, double mu,
                    const int *iocc, const int *ivir, double *omega0) {
  return guarded(c, [&] {
    drop_graph(c);   // buffers or kernels of the captured step change
    require(n > 0 && en && fock && eri, "bad system arguments");
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
    int *d_io = nullptr, *d_iv = nullptr;
    double *fo = dalloc(c, 5 * (int64_t)n);   // fo fv sfo sfv gp0
    double *wo = dalloc(c, O), *wv = dalloc(c, V);
    FTCC_CHECK_HIP(hipMalloc(&d_io, sizeof(int) * O));
    FTCC_CHECK_HIP(hipMalloc(&d_iv, sizeof(int) * V));
    c->allocs.push_back(d_io);
    c->allocs.push_back(d_iv);
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
    // Fock blocks oo ov vo vv
    fock_block(fock, en, n, d_io, d_io, wo, wo, O, O, c->F[0], s);
    fock_block(fock, en, n, d_io, d_iv, wo, wv, O, V, c->F[1], s);
    fock_block(fock, en, n, d_iv, d_io, wv, wo, V, O, c->F[2], s);
    fock_block(fock, en, n, d_iv, d_iv, wv, wv, V, V, c->F[3], s);
    const char *names[9] = {"vvvv", "vvvo", "vovv", "vvoo", "vovo",
                            "oovv", "vooo", "ooov", "oooo"};
    for (int k = 0; k < 9; ++k) {
      const int *p[4];
      const double *w[4];
      int d[4];
      for (int a = 0; a < 4; ++a) {
        const bool occ = names[k][a] == 'o';
        p[a] = occ ? d_io : d_iv;
        w[a] = occ ? wo : wv;
        d[a] = occ ? O : V;
      }
      gather_weighted4(eri, n, p[0], p[1], p[2], p[3], w[0], w[1], w[2], w[3],
                       d[0], d[1], d[2], d[3], c->I[k], s);
    }
    c->have_D = false;
    c->have_energies = true;
    build_derived(c);
    // spin-orbital systems list the alpha orbitals first, then the beta
    // ones: try those labels (FTCC_SPIN=0 keeps every product dense)
    c->spin = 0;
    const char *env = getenv("FTCC_SPIN");
    if (n % 2 == 0 && !(env && atoi(env) == 0)) {
      std::vector<int> lo(O), lv(V);
      for (int k = 0; k < O; ++k) lo[k] = io[k] >= n / 2;
      for (int k = 0; k < V; ++k) lv[k] = iv[k] >= n / 2;
      try_spin(c, lo, lv);
    }
    if (omega0) {
      const int nbk = vsum(gp0, n, c->partials, s);
      reduce_partials(c->partials, nbk, c->red, s);
      read_scalars(c, 1);
      *omega0 = c->host_red[0];
    } else {
      FTCC_CHECK_HIP(hipStreamSynchronize(s));
    }
  });
}

int ftcc_set_blocks(ftcc_ctx *c, const double *const F[4],
                    const double *const I[9]) {
  return guarded(c, [&] {
    drop_graph(c);   // buffers or kernels of the captured step change
    require(F && I, "null block pointer");
    const int O = c->no, V = c->nv;
    const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V, OO = (int64_t)O * O;
    const int64_t S2 = VV * OO;
    const int64_t fsz[4] = {OO, VO, VO, VV};
    const int64_t isz[9] = {VV * VV, VV * VO, VO * VV, S2, S2,
                            S2,      V * OO * O, OO * VO, OO * OO};
    hipStream_t s = c->stream;
    ensure_integrals(c);
    for (int k = 0; k < 4; ++k) {
      require(F[k] != nullptr, "null F block");
      FTCC_CHECK_HIP(hipMemcpyAsync(c->F[k], F[k], 8 * fsz[k],
                                    hipMemcpyDeviceToDevice, s));
    }
    for (int k = 0; k < 9; ++k) {
      require(I[k] != nullptr, "null I block");
      FTCC_CHECK_HIP(hipMemcpyAsync(c->I[k], I[k], 8 * isz[k],
                                    hipMemcpyDeviceToDevice, s));
    }
    build_derived(c);
    c->spin = 0;   // labels unknown: ftcc_set_spin may switch blocking on
    // the orbital energies of an earlier ftcc_set_system belong to other
    // integrals: denominators must come from ftcc_set_denominators, and the
    // flip symmetry (checked on those energies) no longer holds
    c->have_energies = false;
    c->flip = 0;
    FTCC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

int ftcc_set_spin(ftcc_ctx *c, const int *lab_o, const int *lab_v) {
  return guarded(c, [&] {
    drop_graph(c);   // buffers or kernels of the captured step change
    require(c->have_ints, "integrals not set (ftcc_set_system/ftcc_set_blocks)",
            FTCC_ESTATE);
    if (!lab_o && !lab_v) {
      c->spin = 0;
      return;
    }
    require(lab_o && lab_v, "give both label arrays or neither");
    std::vector<int> lo(lab_o, lab_o + c->no), lv(lab_v, lab_v + c->nv);
    for (int x : lo) require(x == 0 || x == 1, "spin labels must be 0 or 1");
    for (int x : lv) require(x == 0 || x == 1, "spin labels must be 0 or 1");
    require(try_spin(c, lo, lv),
            "the integral blocks do not conserve spin under these labels");
  });
}

int ftcc_get_spin(ftcc_ctx *c, int *on) {
  return guarded(c, [&] {
    require(on != nullptr, "null output pointer");
    *on = c->spin ? (1 | (c->spin5 ? 2 : 0) | (c->pack ? 4 : 0) |
                     (c->flip ? 8 : 0))
                  : 0;
  });
}

int ftcc_set_denominators(ftcc_ctx *c, const double *D1, const double *D2) {
  return guarded(c, [&] {
    drop_graph(c);   // buffers or kernels of the captured step change
    require(D1 && D2, "null denominator pointer");
    const int O = c->no, V = c->nv;
    const int64_t VO = (int64_t)V * O, S2 = VO * VO;
    if (!c->D1) c->D1 = dalloc(c, VO);
    if (!c->D2) c->D2 = dalloc(c, S2);
    FTCC_CHECK_HIP(hipMemcpyAsync(c->D1, D1, 8 * VO, hipMemcpyDeviceToDevice,
                                  c->stream));
    FTCC_CHECK_HIP(hipMemcpyAsync(c->D2, D2, 8 * S2, hipMemcpyDeviceToDevice,
                                  c->stream));
    c->have_D = true;
    c->flip = 0;   // the symmetry check covered the orbital energies only
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int ftcc_get_block(ftcc_ctx *c, int which, double *dst) {
  return guarded(c, [&] {
    require(c->have_ints, "integrals not set", FTCC_ESTATE);
    require(which >= 0 && which < 13 && dst, "bad block request");
    const int O = c->no, V = c->nv;
    const int64_t VO = (int64_t)V * O, VV = (int64_t)V * V, OO = (int64_t)O * O;
    const int64_t S2 = VV * OO;
    const int64_t sz[13] = {OO, VO, VO, VV, VV * VV, VV * VO, VO * VV, S2, S2,
                            S2, V * OO * O, OO * VO, OO * OO};
    const double *src = which < 4 ? c->F[which] : c->I[which - 4];
    FTCC_CHECK_HIP(hipMemcpyAsync(dst, src, 8 * sz[which],
                                  hipMemcpyDeviceToDevice, c->stream));
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int ftcc_driving_terms(ftcc_ctx *c, double *T1b, double *T2b) {
  return guarded(c, [&] {
    require(c->have_ints, "integrals not set", FTCC_ESTATE);
    require(T1b && T2b, "null pointer");
    const int64_t VO = (int64_t)c->nv * c->no;
    bcast(T1b, c->F[2], VO, c->ng, -1.0, 0.0, c->stream);
    bcast(T2b, c->I[3], VO * VO, c->ng, -1.0, 0.0, c->stream);
  });
}

int ftcc_set_grid(ftcc_ctx *c, int ng, const double *ti, const double *G) {
  return guarded(c, [&] {
    drop_graph(c);   // buffers or kernels of the captured step change
    require(ti && G, "null grid pointer");
    require(ng > 0, "ng must be positive");
    Grid &gr = c->grid;
    if (gr.d_ti && ng != gr.ng) {
      FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
      FTCC_CHECK_HIP(hipFree(gr.d_ti));
      FTCC_CHECK_HIP(hipFree(gr.d_g));
      FTCC_CHECK_HIP(hipFree(gr.d_G));
      FTCC_CHECK_HIP(hipFree(gr.d_rc));
      gr.d_ti = gr.d_g = gr.d_G = gr.d_rc = nullptr;
    }
    c->ng = ng;
    gr.ng = ng;
    gr.ti.assign(ti, ti + ng);
    gr.G.assign(G, G + (size_t)ng * ng);
    // equally spaced abscissae ('lin', 'mid', 'L' rules): the propagator
    // between points x < y is exp(-delta d)^(y-x)
    gr.uniform = ng > 1;
    gr.delta = ng > 1 ? ti[1] - ti[0] : 0.0;
    double tmax = 0.0;
    for (int k = 0; k < ng; ++k) tmax = std::max(tmax, std::fabs(ti[k]));
    for (int k = 0; k < ng && gr.uniform; ++k)
      if (std::fabs(ti[k] - ti[0] - k * gr.delta) > 8e-16 * (tmax + 1e-300) * (k + 1))
        gr.uniform = false;
    if (!(gr.delta > 0.0)) gr.uniform = false;
    // Simpson ('lin'), midpoint and 'L' weight matrices: zero above the
    // diagonal, each row equal to the row s above it left of the last three
    // columns; integration then runs as a recurrence of step s.  The
    // comparisons are exact: the rules build each row by copying another.
    auto above = [&](int y, int s, int x) {
      return y - s >= 0 ? G[(size_t)(y - s) * ng + x] : 0.0;
    };
    auto fits = [&](int s) {
      for (int y = 0; y < ng; ++y)
        for (int x = 0; x < ng; ++x) {
          const double v = G[(size_t)y * ng + x];
          if (x > y && v != 0.0) return false;
          if (x < y - 2 && v != above(y, s, x)) return false;
        }
      return true;
    };
    gr.rec = 0;
    if (gr.uniform && ng <= MAX_REC_NG) gr.rec = fits(1) ? 1 : fits(2) ? 2 : 0;
    std::vector<double> rc(3 * (size_t)ng, 0.0);
    for (int y = 0; y < ng && gr.rec; ++y)
      for (int k = 0; k < 3; ++k) {
        const int x = y - 2 + k;
        if (x >= 0)
          rc[(size_t)k * ng + y] = G[(size_t)y * ng + x] - above(y, gr.rec, x);
      }
    if (!gr.d_ti) {
      FTCC_CHECK_HIP(hipMalloc(&gr.d_ti, 8 * ng));
      FTCC_CHECK_HIP(hipMalloc(&gr.d_g, 8 * ng));
      FTCC_CHECK_HIP(hipMalloc(&gr.d_G, 8 * (size_t)ng * ng));
      FTCC_CHECK_HIP(hipMalloc(&gr.d_rc, 8 * 3 * (size_t)ng));
    }
    FTCC_CHECK_HIP(hipMemcpyAsync(gr.d_rc, rc.data(), 8 * 3 * (size_t)ng,
                                  hipMemcpyHostToDevice, c->stream));
    FTCC_CHECK_HIP(hipMemcpyAsync(gr.d_ti, ti, 8 * ng, hipMemcpyHostToDevice,
                                  c->stream));
    FTCC_CHECK_HIP(hipMemcpyAsync(gr.d_G, G, 8 * (size_t)ng * ng,
                                  hipMemcpyHostToDevice, c->stream));
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int ftcc_residual(ftcc_ctx *c, const double *T1, const double *T2,
                  double *T1b, double *T2b, uint64_t skip) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && T2 && T1b && T2b, "null amplitude pointer");
    residual_all(c, T1, T2, T1b, T2b, skip);
  });
}

int ftcc_defect(ftcc_ctx *c, const double *T1, const double *T2,
                int low_precision, double *dT1b, double *dT2b, double *S1,
                double *S2, uint64_t skip) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && T2 && dT1b && dT2b && S1 && S2, "null amplitude pointer");
    require(low_precision >= 0 && low_precision <= 2, "unknown precision");
    require(c->corr1 != dT1b && c->corr2 != dT2b,
            "the defect cannot overwrite the correction in use");
    drop_graph(c);
    const int keep = c->prec;
    struct Back {
      ftcc_ctx *c;
      int p;
      ~Back() { c->prec = p; }
    } back{c, keep};
    // the residual of ftcc_step (spin-forbidden doubles left unwritten,
    // the step's flip images), so that the correction cancels the step's
    // own low-precision residual at T exactly
    auto with = [&](int prec, double *o1, double *o2) {
      c->prec = prec;
      c->tbar_scratch = true;
      try {
        residual_all(c, T1, T2, o1, o2, skip);
      } catch (...) {
        c->tbar_scratch = false;
        throw;
      }
      c->tbar_scratch = false;
    };
    const int64_t VO = (int64_t)c->nv * c->no;
    // zeros where the step's residual leaves forbidden doubles unwritten
    fill(dT2b, c->ng * VO * VO, 0.0, c->stream);
    fill(S2, c->ng * VO * VO, 0.0, c->stream);
    with(FTCC_PREC_FP64, dT1b, dT2b);
    with(low_precision, S1, S2);
    axpby(dT1b, S1, c->ng * VO, -1.0, 1.0, c->stream);
    axpby(dT2b, S2, c->ng * VO * VO, -1.0, 1.0, c->stream);
  });
}

int ftcc_integrate(ftcc_ctx *c, double *T1, double *T2) {
  return guarded(c, [&] {
    check_grid_D(c);
    UpdateArgs a = upd(c);
    const int64_t VO = (int64_t)c->nv * c->no;
    // large grids: integrate from a transient copy of Tbar
    double *copy = nullptr;
    if (update_needs_copy(a)) {
      const int64_t elems = (int64_t)c->ng * (T2 ? VO * VO : VO);
      FTCC_CHECK_HIP(hipMalloc(&copy, 8 * elems));
    }
    auto integrate_one = [&](double *T, int64_t per, bool doubles) {
      a.tbar = T;
      a.src = nullptr;
      if (copy) {
        FTCC_CHECK_HIP(hipMemcpyAsync(copy, T, 8 * per * c->ng,
                                      hipMemcpyDeviceToDevice, c->stream));
        a.src = copy;
      }
      if (doubles) {
        a.D = c->have_D ? c->D2 : nullptr;
        update_t2(a, 0, c->stream);
      } else {
        a.D = c->have_D ? c->D1 : nullptr;
        update_t1(a, 0, c->stream);
      }
    };
    try {
      if (T1) integrate_one(T1, VO, false);
      if (T2) integrate_one(T2, VO * VO, true);
    } catch (...) {
      if (copy) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(copy);
      }
      throw;
    }
    if (copy) {
      FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
      FTCC_CHECK_HIP(hipFree(copy));
    }
  });
}

int ftcc_point_begin(ftcc_ctx *c, int ig, double *T1, double *T2,
                     double *T1b, double *T2b) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && T2 && T1b && T2b, "null amplitude pointer");
    require(ig >= 0 && ig < c->ng, "grid point out of range");
    const int64_t VO = (int64_t)c->nv * c->no, S2 = VO * VO;
    hipStream_t s = c->stream;
    if (ig <= 1) {
      // Tbar[ig] = driving terms; T[1] = their single-point integral
      bcast(T1b + ig * VO, c->F[2], VO, 1, -1.0, 0.0, s);
      bcast(T2b + ig * S2, c->I[3], S2, 1, -1.0, 0.0, s);
      if (ig == 1) {
        UpdateArgs a = upd(c);
        a.tbar = T1b;
        a.told = T1 + VO;
        a.D = c->have_D ? c->D1 : nullptr;
        integrate_point(a, false, 1, 0, s);
        a.tbar = T2b;
        a.told = T2 + S2;
        a.D = c->have_D ? c->D2 : nullptr;
        integrate_point(a, true, 1, 0, s);
      }
    } else {
      // linear extrapolation from the two previous points
      const double *ti = c->grid.ti.data();
      const double f = (ti[ig] - ti[ig - 1]) / (ti[ig - 2] - ti[ig - 1]);
      axpby(T1 + ig * VO, T1 + (ig - 1) * VO, VO, 1.0 - f, 0.0, s);
      axpby(T1 + ig * VO, T1 + (ig - 2) * VO, VO, f, 1.0, s);
      axpby(T2 + ig * S2, T2 + (ig - 1) * S2, S2, 1.0 - f, 0.0, s);
      axpby(T2 + ig * S2, T2 + (ig - 2) * S2, S2, f, 1.0, s);
    }
    FTCC_CHECK_HIP(hipStreamSynchronize(s));
  });
}

int ftcc_td_rhs(ftcc_ctx *c, const double *T1, const double *T2, double *K1,
                double *K2, int singles) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && T2 && K1 && K2, "null amplitude pointer");
    residual_all(c, T1, T2, K1, K2, 0);
    UpdateArgs a = upd(c);
    a.D = c->have_D ? c->D1 : nullptr;
    rk_stage(a, false, K1, const_cast<double *>(T1), nullptr, K1, 1.0, 0.0,
             0.0, true, false, singles == 0, c->stream);
    a.D = c->have_D ? c->D2 : nullptr;
    rk_stage(a, true, K2, const_cast<double *>(T2), nullptr, K2, 1.0, 0.0,
             0.0, true, false, false, c->stream);
  });
}

int ftcc_td_step(ftcc_ctx *c, int order, double h, int singles, double *T1,
                 double *T2) {
  return guarded(c, [&] {
    check_ready(c);
    require(order == 1 || order == 2 || order == 4,
            "Runge-Kutta order must be 1, 2 or 4");
    require(T1 && T2, "null amplitude pointer");
    const int64_t VO = (int64_t)c->nv * c->no, S2 = VO * VO;
    const int64_t len[2] = {VO * c->ng, S2 * c->ng};
    for (int k = 0; k < 2; ++k)
      if (!c->td_var[k]) {
        c->td_var[k] = dalloc(c, len[k]);
        c->td_acc[k] = dalloc(c, len[k]);
        c->td_k[k] = dalloc(c, len[k]);
      }
    // kelvin/propagation.py:5-52: rk1 (Euler), rk2 (Heun), rk4 (classical)
    static const double b1[1] = {1.0}, b2[2] = {0.5, 0.5},
                        b4[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};
    static const double a2[1] = {1.0}, a4[3] = {0.5, 0.5, 1.0};
    const double *bs = order == 1 ? b1 : order == 2 ? b2 : b4;
    const double *an = order == 2 ? a2 : a4;
    const double *in1 = T1, *in2 = T2;
    for (int st = 0; st < order; ++st) {
      residual_all(c, in1, in2, c->td_k[0], c->td_k[1], 0);
      const bool first = st == 0, last = st == order - 1;
      const double hn = last ? 0.0 : h * an[st];
      UpdateArgs a = upd(c);
      a.D = c->have_D ? c->D1 : nullptr;
      rk_stage(a, false, c->td_k[0], T1, c->td_var[0], c->td_acc[0], bs[st],
               hn, h, first, last, singles == 0, c->stream);
      a.D = c->have_D ? c->D2 : nullptr;
      rk_stage(a, true, c->td_k[1], T2, c->td_var[1], c->td_acc[1], bs[st], hn,
               h, first, last, false, c->stream);
      in1 = c->td_var[0];
      in2 = c->td_var[1];
    }
  });
}

int ftcc_td_implicit_step(ftcc_ctx *c, int scheme, double h, int max_iter,
                          double damp, double thresh, int singles, double *T1,
                          double *T2, int *iters) {
  return guarded(c, [&] {
    check_ready(c);
    require(scheme >= 0 && scheme <= 2, "implicit scheme must be 0, 1 or 2");
    require(max_iter >= 1, "max_iter must be positive");
    require(T1 && T2, "null amplitude pointer");
    hipStream_t s = c->stream;
    const int64_t VO = (int64_t)c->nv * c->no, S2 = VO * VO;
    const int64_t len[2] = {VO * c->ng, S2 * c->ng};
    for (int k = 0; k < 2; ++k) {
      if (!c->td_var[k]) {
        c->td_var[k] = dalloc(c, len[k]);
        c->td_acc[k] = dalloc(c, len[k]);
        c->td_k[k] = dalloc(c, len[k]);
      }
      if (!c->td_k1[k]) c->td_k1[k] = dalloc(c, len[k]);
    }
    double *T[2] = {T1, T2};
    auto pass = [&](int k, int mode) {
      UpdateArgs a = upd(c);
      a.D = c->have_D ? (k ? c->D2 : c->D1) : nullptr;
      a.partials = c->partials;
      return implicit_pass(a, k == 1, c->td_k[k], T[k], c->td_var[k],
                           c->td_acc[k], c->td_k1[k], h, damp, mode, scheme,
                           k == 0 && singles == 0, s);
    };
    // predictor: k1 = h f(T)   (kelvin/propagation.py:83-128)
    residual_all(c, T1, T2, c->td_k[0], c->td_k[1], 0);
    pass(0, 0);
    pass(1, 0);
    bool converged = false;
    int it = 0;
    while (it < max_iter && !converged) {
      residual_all(c, c->td_var[0], c->td_var[1], c->td_k[0], c->td_k[1], 0);
      int nb = pass(0, 1);
      reduce_partials(c->partials, nb, c->red, s);
      nb = pass(1, 1);
      reduce_partials(c->partials, nb, c->red + 3, s);
      read_scalars(c, 6);
      const double err =
          sqrt(c->host_red[0]) / (sqrt(c->host_red[1]) + 0.01) +
          sqrt(c->host_red[3]) / (sqrt(c->host_red[4]) + 0.01);
      ++it;
      converged = err < thresh;
    }
    if (iters) *iters = it;
    require(converged, scheme == 2 ? "AM2: Failed to compute implicit step"
                                   : "BE: Failed to compute implicit step");
    pass(0, 2);
    pass(1, 2);
  });
}

int ftcc_point_step(ftcc_ctx *c, int ig, double *T1, double *T2, double *T1b,
                    double *T2b, double damp, double out2[2]) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1 && T2 && T1b && T2b && out2, "null pointer");
    require(ig >= 1 && ig < c->ng, "grid point out of range");
    ensure_workspace(c);
    const int64_t VO = (int64_t)c->nv * c->no, S2 = VO * VO;
    hipStream_t s = c->stream;
    // residual at point ig alone (spin blocks if its amplitudes allow)
    const int spin_saved = c->spin;
    if (c->spin &&
        !amplitudes_conserve_spin(c, T1 + ig * VO, T2 + ig * S2, 1))
      c->spin = 0;
    struct Restore {
      ftcc_ctx *c;
      int v;
      ~Restore() { c->spin = v; }
    } restore{c, spin_saved};
    residual_chunk(c, T1 + ig * VO, T2 + ig * S2, T1b + ig * VO,
                   T2b + ig * S2, 1);
    UpdateArgs a = upd(c);
    a.damp = damp;
    a.tbar = T1b;
    a.told = T1 + ig * VO;
    a.D = c->have_D ? c->D1 : nullptr;
    int nbk = integrate_point(a, false, ig, 1, s);
    reduce_partials(c->partials, nbk, c->red, s);
    a.tbar = T2b;
    a.told = T2 + ig * S2;
    a.D = c->have_D ? c->D2 : nullptr;
    nbk = integrate_point(a, true, ig, 1, s);
    reduce_partials(c->partials, nbk, c->red + 3, s);
    read_scalars(c, 6);
    out2[0] = c->host_red[0];
    out2[1] = c->host_red[3];
  });
}

int ftcc_ueg_eri(ftcc_ctx *c, int nk, const int *ivec, const double *kvec,
                 double L, double *eri) {
  return guarded(c, [&] {
    require(nk > 0 && ivec && kvec && eri && L > 0.0, "bad UEG arguments");
    int *dX = nullptr;
    double *dk = nullptr;
    FTCC_CHECK_HIP(hipMalloc(&dX, sizeof(int) * 3 * nk));
    FTCC_CHECK_HIP(hipMalloc(&dk, sizeof(double) * 3 * nk));
    hipStream_t s = c->stream;
    FTCC_CHECK_HIP(hipMemcpyAsync(dX, ivec, sizeof(int) * 3 * nk,
                                  hipMemcpyHostToDevice, s));
    FTCC_CHECK_HIP(hipMemcpyAsync(dk, kvec, sizeof(double) * 3 * nk,
                                  hipMemcpyHostToDevice, s));
    const double aaa = 4.0 * M_PI / (L * L * L);
    ueg_eri(nk, dX, dk, aaa, eri, s);
    const hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(dX);
    (void)hipFree(dk);
    FTCC_CHECK_HIP(e);
  });
}

int ftcc_mean_field(ftcc_ctx *c, int n, const double *en, const double *eri,
                    const double *fo, double *fock, double *e1) {
  return guarded(c, [&] {
    require(n > 0 && en && eri && fo, "bad mean-field arguments");
    hipStream_t s = c->stream;
    if (fock) mean_field_fock(n, en, eri, fo, fock, s);
    if (e1) {
      const int nbk = mp1_partials(n, eri, fo, c->partials, s);
      reduce_partials(c->partials, nbk, c->red, s);
      read_scalars(c, 1);
      *e1 = 0.5 * c->host_red[0];
    } else {
      FTCC_CHECK_HIP(hipStreamSynchronize(s));
    }
  });
}

int ftcc_energy(ftcc_ctx *c, const double *T1, const double *T2,
                const double *g, double beta, int qterm, double *energy) {
  return guarded(c, [&] {
    require(c->have_ints, "integrals not set", FTCC_ESTATE);
    require(c->grid.d_g != nullptr, "grid not set (ftcc_set_grid)",
            FTCC_ESTATE);
    require(T1 && T2 && g && energy, "null pointer");
    require(beta > 0.0, "beta must be positive");
    hipStream_t s = c->stream;
    upload_g(c, g);
    int nb = energy_t1(T1, c->grid.d_g, c->fovT, c->ng, c->nv, c->no,
                       c->partials, s);
    reduce_partials(c->partials, nb, c->red, s);
    nb = energy_t2(T1, T2, c->grid.d_g, c->oovvT, c->ng, c->nv, c->no, qterm,
                   c->partials, s);
    reduce_partials(c->partials, nb, c->red + 3, s);
    read_scalars(c, 6);
    *energy = (c->host_red[2] + c->host_red[5]) / beta;
  });
}

int ftcc_sumsq(ftcc_ctx *c, const double *x, int64_t n, double *out) {
  return guarded(c, [&] {
    require(x && out && n >= 0, "bad sumsq arguments");
    const int nb = sumsq(x, n, c->partials, c->stream);
    reduce_partials(c->partials, nb, c->red, c->stream);
    read_scalars(c, 1);
    *out = c->host_red[0];
  });
}

// The device work of one step (residual, integration, damping, norms and
// energy partial sums reduced into c->red), with no host synchronisation
// when the amplitudes are trusted.
void step_device(ftcc_ctx *c, double *T1old, double *T2old, double *T1b,
                 double *T2b, double damp, uint64_t skip) {
  hipStream_t s = c->stream;
  // Tbar is this step's scratch: its spin-forbidden doubles are read by
  // nobody when the residual ran in the spin layout (the update below skips
  // them), so the assembly need not write their zeros
  struct Scratch {
    ftcc_ctx *c;
    Scratch(ftcc_ctx *c_) : c(c_) { c->tbar_scratch = true; }
    ~Scratch() { c->tbar_scratch = false; }
  };
  bool spin_used;
  {
    Scratch sc(c);
    spin_used = residual_all(c, T1old, T2old, T1b, T2b, skip);
  }
  UpdateArgs a = upd(c);
  a.damp = damp;
  a.tbar = T1b;
  a.told = T1old;
  a.corr = c->corr1;
  a.D = c->have_D ? c->D1 : nullptr;
  a.eint = c->fovT;
  // algorithmic bytes: Tbar read, Told read and written at every grid point
  // (and the correction read), the energy integral read once, per element
  const double per_elem = (24.0 + (c->corr1 ? 8.0 : 0.0)) * c->ng + 8.0;
  int nb = 0;
  timed_class(c, ftcc_ctx::PROF_UPDATE, 0.0,
              per_elem * c->nv * (double)c->no,
              [&] { nb = update_t1(a, 1, s); });
  reduce_partials(c->partials, nb, c->red, s);
  a.tbar = T2b;
  a.told = T2old;
  a.corr = c->corr2;
  a.D = c->have_D ? c->D2 : nullptr;
  a.eint = c->oovvT;
  a.t1new = T1old;
  if (spin_used && c->spin5) {
    // label runs: only the spin-allowed doubles are enumerated (the others
    // are zero in Tbar and T)
    int64_t n = 0;
    for (int sp = 0; sp < 2; ++sp) {
      a.vrun[2 * sp] = c->v_start[sp];
      a.vrun[2 * sp + 1] = c->v_count[sp];
      a.orun[2 * sp] = c->o_start[sp];
      a.orun[2 * sp + 1] = c->o_count[sp];
    }
    for (int box = 0; box < 16; ++box) {
      const int sa = box >> 3, sb = (box >> 2) & 1, si = (box >> 1) & 1,
                sj = box & 1;
      if (sa + sb == si + sj)
        n += (int64_t)c->v_count[sa] * c->v_count[sb] * c->o_count[si] *
             c->o_count[sj];
    }
    a.nallowed = n;
    if (!a.rec) {   // the per-grid-point update skips by label instead
      a.lv = c->d_lab_v;
      a.lo = c->d_lab_o;
    }
    // flip-symmetric step: the alpha-lead half, results written twice
    a.flip_half = (a.rec && c->last_flip_step && c->v_count[0] == c->v_count[1] &&
                   c->o_count[0] == c->o_count[1] && c->v_start[0] == 0 &&
                   c->o_start[0] == 0) ? 1 : 0;
  } else if (spin_used) {   // spin-forbidden doubles are zero: skipped
    a.lv = c->d_lab_v;
    a.lo = c->d_lab_o;
  }
  const double elems2 = a.nallowed > 0 && a.rec
                            ? (double)a.nallowed
                            : (double)c->nv * c->nv * c->no * c->no;
  // flip half: Tbar, Told (and the correction) read for half the elements,
  // Told written for all
  const double bytes2 =
      a.flip_half ? (0.5 * (per_elem - 8.0 * c->ng) + 8.0 * c->ng) * elems2
                  : per_elem * elems2;
  timed_class(c, ftcc_ctx::PROF_UPDATE, 0.0, bytes2,
              [&] { nb = update_t2(a, 1, s); });
  reduce_partials(c->partials, nb, c->red + 3, s);
}

void step_outputs(ftcc_ctx *c, double beta, double out5[5]) {
  read_scalars(c, 6);
  out5[0] = c->host_red[0];
  out5[1] = c->host_red[3];
  out5[2] = c->host_red[1];
  out5[3] = c->host_red[4];
  out5[4] = (c->host_red[2] + c->host_red[5]) / beta;
}

int ftcc_set_chunk(ftcc_ctx *c, int nb) {
  return guarded(c, [&] {
    require(nb >= 0, "chunk size must be >= 0");
    if (nb == c->nb_fixed) return;
    drop_graph(c);
    c->nb_fixed = nb;
    free_workspace(c);
  });
}

int ftcc_get_chunk(ftcc_ctx *c, int *nb) {
  return guarded(c, [&] {
    require(nb != nullptr, "null output pointer");
    ensure_workspace(c);
    *nb = c->nb;
  });
}

int ftcc_set_trusted(ftcc_ctx *c, int on) {
  return guarded(c, [&] {
    drop_graph(c);
    c->trusted = on != 0;
    c->trust_pending = c->trusted;
  });
}

int ftcc_set_correction(ftcc_ctx *c, const double *dT1bar,
                        const double *dT2bar) {
  return guarded(c, [&] {
    require((dT1bar == nullptr) == (dT2bar == nullptr),
            "give both correction arrays or neither");
    drop_graph(c);   // the captured update reads the pointers
    c->corr1 = dT1bar;
    c->corr2 = dT2bar;
  });
}

int ftcc_set_weights(ftcc_ctx *c, const double *g) {
  return guarded(c, [&] {
    check_ready(c);
    require(g != nullptr, "null weights");
    upload_g(c, g);
  });
}

int ftcc_step_graph(ftcc_ctx *c, double *T1old, double *T2old, double *T1b,
                    double *T2b, double beta, double damp, uint64_t skip,
                    double out5[5]) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1old && T2old && T1b && T2b && out5, "null pointer");
    require(beta > 0.0, "beta must be positive");
    require(!c->spin || c->trusted,
            "graph steps need trusted amplitudes (ftcc_set_trusted) in the "
            "spin layout");
    require(!c->prof_on, "product profiling is off in graph steps");
    require(c->stream != nullptr,
            "graph steps need a context created on a non-default stream");
    const void *key[4] = {T1old, T2old, T1b, T2b};
    const bool same = c->gwarm && std::equal(key, key + 4, c->gkey) &&
                      c->gbeta == beta && c->gdamp == damp &&
                      c->gskip == skip && c->gprec == c->prec;
    if (same && c->step_exec) {
      FTCC_CHECK_HIP(hipGraphLaunch(c->step_exec, c->stream));
    } else if (same) {
      // second call with these arguments: capture, then launch
      FTCC_CHECK_HIP(hipStreamBeginCapture(c->stream,
                                           hipStreamCaptureModeThreadLocal));
      try {
        step_device(c, T1old, T2old, T1b, T2b, damp, skip);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      FTCC_CHECK_HIP(hipStreamEndCapture(c->stream, &c->step_graph));
      FTCC_CHECK_HIP(
          hipGraphInstantiate(&c->step_exec, c->step_graph, nullptr, nullptr, 0));
      FTCC_CHECK_HIP(hipGraphLaunch(c->step_exec, c->stream));
    } else {
      // new arguments: run eagerly (allocations happen here)
      drop_graph(c);
      step_device(c, T1old, T2old, T1b, T2b, damp, skip);
      std::copy(key, key + 4, c->gkey);
      c->gbeta = beta;
      c->gdamp = damp;
      c->gskip = skip;
      c->gprec = c->prec;
      c->gwarm = true;
    }
    step_outputs(c, beta, out5);
  });
}

int ftcc_step(ftcc_ctx *c, double *T1old, double *T2old, double *T1b,
              double *T2b, const double *g, double beta, double damp,
              uint64_t skip, double out5[5]) {
  return guarded(c, [&] {
    check_ready(c);
    require(T1old && T2old && T1b && T2b && g && out5, "null pointer");
    require(beta > 0.0, "beta must be positive");
    upload_g(c, g);
    step_device(c, T1old, T2old, T1b, T2b, damp, skip);
    step_outputs(c, beta, out5);
  });
}

int ftcc_workspace_bytes(const ftcc_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return FTCC_EINVAL;
  *bytes = c->bytes;
  return FTCC_OK;
}

int ftcc_set_gemm_plan(ftcc_ctx *c, int tile_policy, int split_target) {
  return guarded(c, [&] {
    require(tile_policy >= 0 && tile_policy <= 2, "tile policy is 0, 1 or 2");
    require(split_target >= 1 && split_target <= (1 << 20),
            "split target out of range");
    if (tile_policy == c->tile_policy && split_target == c->split_target)
      return;
    drop_graph(c);   // a captured step holds the old launches
    c->tile_policy = tile_policy;
    c->split_target = split_target;
  });
}

int ftcc_profile_enable(ftcc_ctx *c, int on) {
  return guarded(c, [&] {
    if (on && !c->prof_on) {
      const char *e = getenv("FTCC_PROFILE_SITES");
      c->prof_sites = e && atoi(e) != 0;
    }
    if (!on && c->prof_on && c->prof_sites) {
      FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
      flush_profile(c);
      std::vector<std::pair<std::array<int64_t, 6>, ftcc_ctx::SiteStat>> v(
          c->site_stats.begin(), c->site_stats.end());
      std::sort(v.begin(), v.end(), [](const auto &a, const auto &b) {
        return a.second.ms > b.second.ms;
      });
      fprintf(stderr, "ftcc sites: class line launches ms GB GFLOP "
                      "M N K slices\n");
      for (const auto &x : v)
        fprintf(stderr, "ftcc site %d %d %lld %.3f %.3f %.3f %d %d %d %d\n",
                (int)x.first[0], (int)x.first[1],
                (long long)x.second.launches, x.second.ms,
                x.second.bytes / 1e9, x.second.flops / 1e9, x.second.M,
                x.second.N, x.second.K, x.second.nb);
      c->site_stats.clear();
    }
    c->prof_on = (on != 0);
  });
}

int ftcc_profile_read(ftcc_ctx *c, int64_t *launches, double *ms,
                      double *flops) {
  return guarded(c, [&] {
    require(launches && ms && flops, "null pointer");
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
    flush_profile(c);
    *launches = c->prof_launches[0];
    *ms = c->prof_ms[0];
    *flops = c->prof_flops[0];
    c->prof_launches[0] = 0;
    c->prof_ms[0] = 0.0;
    c->prof_flops[0] = 0.0;
    c->prof_bytes[0] = 0.0;
  });
}

int ftcc_profile_class(ftcc_ctx *c, int cls, int64_t *launches, double *ms,
                       double *flops, double *bytes) {
  return guarded(c, [&] {
    require(launches && ms && flops && bytes, "null pointer");
    require(cls >= 0 && cls < ftcc_ctx::NPROF, "unknown kernel class");
    FTCC_CHECK_HIP(hipStreamSynchronize(c->stream));
    flush_profile(c);
    *launches = c->prof_launches[cls];
    *ms = c->prof_ms[cls];
    *flops = c->prof_flops[cls];
    *bytes = c->prof_bytes[cls];
    c->prof_launches[cls] = 0;
    c->prof_ms[cls] = 0.0;
    c->prof_flops[cls] = 0.0;
    c->prof_bytes[cls] = 0.0;
  });
}

int ftcc_gemm(ftcc_ctx *c, int M, int N, int K, int batch, double alpha,
              const double *A, int64_t sAm, int64_t sAk, int64_t sAb,
              const double *B, int64_t sBk, int64_t sBn, int64_t sBb,
              double beta, double *C, int64_t sCm, int64_t sCn, int64_t sCb,
              int precision) {
  return guarded(c, [&] {
    require(M >= 0 && N >= 0 && K >= 0 && batch >= 1, "bad GEMM sizes");
    require(A && B && C, "null GEMM operand");
    Gemm g = mk(M, N, K, alpha, beta, A, sAm, sAk, sAb, B, sBk, sBn, sBb, C,
                sCm, sCn, sCb, batch);
    if (precision != FTCC_PREC_BF16X3 || (int64_t)M * N * K == 0) {
      run(c, g, precision);
      return;
    }
    // bf16x3: split both operands into transient K-contiguous buffers and
    // run the split-bf16 kernel of the O(n^6) products
    SplitMat As, Bs;
    As.ld = Bs.ld = split_ld(K);
    As.bstride = (int64_t)M * As.ld;
    Bs.bstride = (int64_t)N * Bs.ld;
    const size_t abytes = 2 * (size_t)As.bstride * batch;
    const size_t bbytes = 2 * (size_t)Bs.bstride * batch;
    char *buf = nullptr;
    FTCC_CHECK_HIP(hipMalloc(&buf, 2 * (abytes + bbytes)));
    As.hi = (__bf16 *)buf;
    As.lo = (__bf16 *)(buf + abytes);
    Bs.hi = (__bf16 *)(buf + 2 * abytes);
    Bs.lo = (__bf16 *)(buf + 2 * abytes + bbytes);
    split_convert(A, sAm, sAk, sAb, M, K, As, batch, c->stream);
    split_convert(B, sBn, sBk, sBb, N, K, Bs, batch, c->stream);
    timed(c, true, gemm_flops(g), [&] {
      sbgemm(M, N, K, batch, alpha, As, Bs, beta, C, sCm, sCn, sCb,
             c->stream);
    });
    const hipError_t e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    FTCC_CHECK_HIP(e);
  });
}

}  // extern "C"
