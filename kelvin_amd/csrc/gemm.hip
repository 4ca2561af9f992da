// Batched strided GEMM on the CDNA4 matrix cores.
//
// One workgroup = 4 wave64s computes a 128x128 tile of C; each wave owns a
// 64x64 quadrant as 4x4 MFMA 16x16 blocks.  K advances in slabs of 16 staged
// through LDS (double buffered, one barrier per slab) with the next slab
// prefetched into registers while the current one feeds the MFMAs.
//
// LDS holds both operands "K-inner": As[m][k], Bs[n][k], rows padded so the
// 16 lanes that share a k-group hit distinct banks.  Inside a slab the
// physical k index used by lane group g (= lane/16) in MFMA step s is 4g+s, so
// each lane fetches its four k values for the four MFMA steps with one
// 128-bit (FP32) or two 128-bit (FP64) LDS reads; A and B use the same map, so
// the contraction is unchanged.
//
// Global loads are scalar and coalesced along whichever operand dimension is
// contiguous in memory (template flags), so every stride pattern is legal and
// the common layouts (A k-contiguous or m-contiguous, B n-contiguous or
// k-contiguous) stream whole cache lines.
//
// Precisions: FP64 (v_mfma_f64_16x16x4_f64) and FP32 (v_mfma_f32_16x16x4_f32,
// FP64 operands rounded to FP32 on load).  Split-BF16 products run in
// sbgemm.hip on operands split ahead of time.
//
// The output may be scattered through row and column maps (spin-blocked
// products write their blocks straight into the caller's layout).
#include "ftcc_internal.h"

#include <algorithm>

namespace ftcc {
namespace {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));

constexpr int BK = 16, NT = 256;

// Output tile: 128x128, or 64 on the side of a dimension of at most 64
// (products over one spin run of an orbital index)
struct Tile {
  int m, n;
};
// Short K (the O(n^5) contractions over one index): the narrow,
// single-buffered tiles (three workgroups per CU) stream the output faster
// than 128x128 (3.5% on those launches at n = 114).
inline Tile pick_tile(int M, int N, int K, int policy) {
  if (M <= 64 && N > 64) return {64, 128};
  if (N <= 64 && M > 64) return {128, 64};
  if (policy == 2 || (policy == 0 && K <= 64))
    return M >= N ? Tile{128, 64} : Tile{64, 128};
  return {128, 128};
}

struct GemmDev {
  int M, N, K, nb1, nb2, splits, kchunk, tilesM, tilesN;
  int m_inner;               // tile rows vary fastest (fewer of them)
  int reduce_b2;
  double alpha, beta;
  const double *A;
  int64_t sAm, sAk, sA1, sA2;
  const double *B;
  int64_t sBk, sBn, sB1, sB2;
  double *C;
  int64_t sCm, sCn, sC1, sC2;
  double *partial;   // [R][nbC][M][N] when splits*R > 1
  int nbC;           // number of distinct C batches (nb1 or nb1*nb2)
  const int *rmap, *cmap;   // output scatter (nullptr: identity)
  const double *D;          // addend gamma*D laid out like C (or nullptr)
  double gamma;
  int64_t sDm, sDn, sD1, sD2;
  // rows m = (m0*M1 + m1)*M2 + m2 and columns likewise (one index: M1 = 1,
  // M2 = 2^30); the strides above belong to the innermost index
  int M1, M2, N1, N2;
  int64_t sAm0, sAm1, sCm0, sCm1, sDm0, sDm1;
  int64_t sBn0, sBn1, sCn0, sCn1, sDn0, sDn1;
  int K1, K2;                // K levels (kernels instantiated with KL)
  int64_t sAk0, sAk1, sBk0, sBk1;
  // optional map of one row level and one column level (level 0, 1 or 2;
  // -1: none): that level's index x becomes map[x] before the strides
  const int *mmap, *nmap;
  int mlvl, nlvl;
};

struct Lvl {
  int x0, x1, x2;
};
__device__ inline Lvl map_level(Lvl l, const int *map, int lvl) {
  if (lvl == 0) l.x0 = map[l.x0];
  else if (lvl == 1) l.x1 = map[l.x1];
  else if (lvl == 2) l.x2 = map[l.x2];
  return l;
}
__device__ inline Lvl levels(int x, int n1, int n2) {
  Lvl l;
  l.x2 = x % n2;
  const int t = x / n2;
  l.x1 = t % n1;
  l.x0 = t / n1;
  return l;
}
__device__ inline int64_t at(const Lvl &l, int64_t s0, int64_t s1,
                             int64_t s2) {
  return l.x0 * s0 + l.x1 * s1 + l.x2 * s2;
}

template <typename T> struct Mma;
template <> struct Mma<double> {
  using acc_t = d4_t;
  static constexpr int LDK = BK + 2;   // 36-dword rows: conflict-free reads
  // output element v of lane group lg sits in row lg + 4v of the 16x16 block
  __device__ static int crow(int lg, int v) { return lg + 4 * v; }
  __device__ static acc_t op(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mma<float> {
  using acc_t = f4_t;
  static constexpr int LDK = BK + 4;   // 20-dword rows: conflict-free reads
  // output element v of lane group lg sits in row 4*lg + v of the block
  __device__ static int crow(int lg, int v) { return 4 * lg + v; }
  __device__ static acc_t op(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Decode the flat block index into (tile n, tile m, split, b2, b1), the
// tiles along the dimension with fewer of them innermost: those tiles share
// one panel of the operand with more panels, which is then read from HBM
// once instead of once per tile row (an O(n^5) term of 456 x 185193 x 57
// read 1.72 GB with columns innermost against 0.84 GB of distinct data).
// The index is first regrouped per XCD: workgroups are dispatched to the
// eight XCDs round robin and each XCD has its own L2; the hardware ids x,
// x + 8, x + 16, ... of XCD x take a contiguous range of tiles (a bijection
// on [0, gridDim.x)).
struct BlockPos {
  int tn, tm, ks, b1, b2;
};
__device__ inline BlockPos decode(const GemmDev &p) {
  const int64_t G = gridDim.x, hw = blockIdx.x;
  const int64_t q8 = G / 8, r8 = G % 8, x = hw % 8, loc = hw / 8;
  int64_t bid = x < r8 ? x * (q8 + 1) + loc : r8 * (q8 + 1) + (x - r8) * q8 + loc;
  BlockPos q;
  if (p.m_inner) {
    q.tm = (int)(bid % p.tilesM);
    bid /= p.tilesM;
    q.tn = (int)(bid % p.tilesN);
    bid /= p.tilesN;
  } else {
    q.tn = (int)(bid % p.tilesN);
    bid /= p.tilesN;
    q.tm = (int)(bid % p.tilesM);
    bid /= p.tilesM;
  }
  q.ks = (int)(bid % p.splits);
  bid /= p.splits;
  q.b2 = (int)(bid % p.nb2);
  q.b1 = (int)(bid / p.nb2);
  return q;
}

// Element offsets of the tile's rows of A and columns of B, one table entry
// per row/column of the tile (-1 past the end), kept in LDS so they hold no
// registers through the K loop.
template <int TM, int TN>
__device__ inline void tile_offsets(const GemmDev &p, int m0, int n0,
                                    int64_t *aoff, int64_t *boff) {
  const int t = threadIdx.x;
  if (t < TM) {
    const int m = m0 + t;
    aoff[t] = m < p.M ? at(map_level(levels(m, p.M1, p.M2), p.mmap, p.mlvl),
                           p.sAm0, p.sAm1, p.sAm)
                      : -1;
  } else if (t < TM + TN) {
    const int n = n0 + t - TM;
    boff[t - TM] =
        n < p.N ? at(map_level(levels(n, p.N1, p.N2), p.nmap, p.nlvl),
                     p.sBn0, p.sBn1, p.sBn)
                : -1;
  }
}

// One K slab (16 deep) of the TM-row A tile and the TN-column B tile: each
// of the 256 threads loads TM/16 values of A and TN/16 of B.  Along the
// contiguous index: m = t % TM, k = t / TM + (256 / TM) r; otherwise
// m = t / 16 + 16 r, k = t % 16.
// KL (K over several indices, both operands K-fastest): every load of a
// thread shares k = k0 + t % 16, decoded once per slab.
template <typename T, bool A_MFAST, bool B_NFAST, int TM, int TN, bool KL>
__device__ inline void load_slab(const GemmDev &p, const double *A,
                                 const double *B, const int64_t *aoff,
                                 const int64_t *boff, int k0, int kend,
                                 double ra[TM / 16], double rb[TN / 16]) {
  static_assert(!KL || (!A_MFAST && !B_NFAST), "K levels need K-fast loads");
  const int t = threadIdx.x;
  int64_t ka = 0, kb = 0;
  bool kin = true;
  if (KL) {
    const int k = k0 + (t & 15);
    kin = k < kend;
    const Lvl l = levels(kin ? k : 0, p.K1, p.K2);
    ka = at(l, p.sAk0, p.sAk1, p.sAk);
    kb = at(l, p.sBk0, p.sBk1, p.sBk);
  }
#pragma unroll
  for (int r = 0; r < TM / 16; ++r) {
    const int64_t off = aoff[A_MFAST ? (t % TM) : (t >> 4) + 16 * r];
    double v = 0.0;
    if (KL) {
      if (off >= 0 && kin) v = A[off + ka];
    } else {
      const int k = A_MFAST ? k0 + t / TM + (NT / TM) * r : k0 + (t & 15);
      if (off >= 0 && k < kend) v = A[off + (int64_t)k * p.sAk];
    }
    ra[r] = v;
  }
#pragma unroll
  for (int r = 0; r < TN / 16; ++r) {
    const int64_t off = boff[B_NFAST ? (t % TN) : (t >> 4) + 16 * r];
    double v = 0.0;
    if (KL) {
      if (off >= 0 && kin) v = B[kb + off];
    } else {
      const int k = B_NFAST ? k0 + t / TN + (NT / TN) * r : k0 + (t & 15);
      if (off >= 0 && k < kend) v = B[(int64_t)k * p.sBk + off];
    }
    rb[r] = v;
  }
}

template <typename T, bool A_MFAST, bool B_NFAST, int TM, int TN>
// (the prefetched values stay FP64 until here, so that the FP32 form's
// rounding does not wait for the loads right where they are issued)
__device__ inline void store_slab(T *As, T *Bs, const double ra[TM / 16],
                                  const double rb[TN / 16]) {
  constexpr int LDK = Mma<T>::LDK;
  const int t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < TM / 16; ++r) {
    if (A_MFAST)
      As[(t % TM) * LDK + t / TM + (NT / TM) * r] = (T)ra[r];
    else
      As[((t >> 4) + 16 * r) * LDK + (t & 15)] = (T)ra[r];
  }
#pragma unroll
  for (int r = 0; r < TN / 16; ++r) {
    if (B_NFAST)
      Bs[(t % TN) * LDK + t / TN + (NT / TN) * r] = (T)rb[r];
    else
      Bs[((t >> 4) + 16 * r) * LDK + (t & 15)] = (T)rb[r];
  }
}

template <typename T>
__device__ inline void load4(const T *src, T out[4]) {
#pragma unroll
  for (int s = 0; s < 4; ++s) out[s] = src[s];
}

// TM x TN output tile per workgroup (128x128, 64x128 or 128x64): the four
// waves form a 2x2 grid, each owning (TM/2) x (TN/2) as RT x CT MFMA blocks.
template <typename T, bool A_MFAST, bool B_NFAST, int TM, int TN, bool KL>
__global__ __launch_bounds__(NT, 2) void gemm_mfma_kernel(GemmDev p) {
  using acc_t = typename Mma<T>::acc_t;
  constexpr int LDK = Mma<T>::LDK;
  constexpr int RT = TM / 32, CT = TN / 32;
  // the 128x128 tile is register-bound at two workgroups per CU and double
  // buffers its slabs; the narrow tiles single-buffer so that three
  // workgroups fit the LDS (two barriers per slab)
  constexpr bool DB = (TM == 128 && TN == 128);
  constexpr int NBUF = DB ? 2 : 1;
  static_assert(TM + TN <= NT, "one offset-table entry per thread");
  __shared__ T As[NBUF][TM * LDK];
  __shared__ T Bs[NBUF][TN * LDK];
  __shared__ int64_t aoff[TM], boff[TN], drow[TM], dcol[TN];

  const BlockPos q = decode(p);
  const double *A = p.A + q.b1 * p.sA1 + q.b2 * p.sA2;
  const double *B = p.B + q.b1 * p.sB1 + q.b2 * p.sB2;
  const int m0 = q.tm * TM, n0 = q.tn * TN;
  const int kbeg = q.ks * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);
  const int nslab = (kend - kbeg + BK - 1) / BK;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int li = lane & 15, lg = lane >> 4;

  acc_t acc[RT][CT];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < CT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  double ra[TM / 16], rb[TN / 16];
  tile_offsets<TM, TN>(p, m0, n0, aoff, boff);
  __syncthreads();
  if (nslab > 0) {
    load_slab<T, A_MFAST, B_NFAST, TM, TN, KL>(p, A, B, aoff, boff, kbeg,
                                               kend, ra, rb);
    store_slab<T, A_MFAST, B_NFAST, TM, TN>(As[0], Bs[0], ra, rb);
  }
  __syncthreads();
  for (int s = 0; s < nslab; ++s) {
    const int buf = DB ? (s & 1) : 0;
    if (s + 1 < nslab)
      load_slab<T, A_MFAST, B_NFAST, TM, TN, KL>(p, A, B, aoff, boff,
                                                 kbeg + (s + 1) * BK, kend,
                                                 ra, rb);
    T a[RT][4], b[CT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
      load4<T>(&As[buf][(wm * (TM / 2) + rt * 16 + li) * LDK + 4 * lg],
               a[rt]);
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
      load4<T>(&Bs[buf][(wn * (TN / 2) + ct * 16 + li) * LDK + 4 * lg],
               b[ct]);
    // single buffer: every wave has its fragments before the slab is
    // overwritten
    if (!DB) __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          acc[rt][ct] = Mma<T>::op(a[rt][kk], b[ct][kk], acc[rt][ct]);
    if (s + 1 < nslab)
      store_slab<T, A_MFAST, B_NFAST, TM, TN>(As[DB ? buf ^ 1 : 0],
                                              Bs[DB ? buf ^ 1 : 0], ra, rb);
    __syncthreads();
  }

  // epilogue: lane holds column li of each 16x16 block, rows Mma<T>::crow.
  // Output offsets of the tile's rows and columns (after the scatter maps;
  // -1 where dropped or past the end) go to LDS first, so the unrolled
  // stores below only look them up.  The K loop ended on a barrier, so the
  // offset tables of A and B are free.
  const bool direct = (p.partial == nullptr);
  double *C = p.C + q.b1 * p.sC1 + q.b2 * p.sC2;
  const double *D = p.D ? p.D + q.b1 * p.sD1 + q.b2 * p.sD2 : nullptr;
  int64_t *crow = aoff, *ccol = boff;
  if (direct) {
    const int t = threadIdx.x;
    if (t < TM) {
      const int row = m0 + t;
      const int rr = row < p.M ? (p.rmap ? p.rmap[row] : row) : -1;
      const Lvl l =
          map_level(levels(rr < 0 ? 0 : rr, p.M1, p.M2), p.mmap, p.mlvl);
      crow[t] = rr < 0 ? -1 : at(l, p.sCm0, p.sCm1, p.sCm);
      drow[t] = rr < 0 ? -1 : at(l, p.sDm0, p.sDm1, p.sDm);
    } else if (t < TM + TN) {
      const int col = n0 + t - TM;
      const int cc = col < p.N ? (p.cmap ? p.cmap[col] : col) : -1;
      const Lvl l =
          map_level(levels(cc < 0 ? 0 : cc, p.N1, p.N2), p.nmap, p.nlvl);
      ccol[t - TM] = cc < 0 ? -1 : at(l, p.sCn0, p.sCn1, p.sCn);
      dcol[t - TM] = cc < 0 ? -1 : at(l, p.sDn0, p.sDn1, p.sDn);
    }
    __syncthreads();
  }
  const int cb = p.reduce_b2 ? q.b1 : q.b1 * p.nb2 + q.b2;
  const int rsel = p.reduce_b2 ? q.b2 * p.splits + q.ks : q.ks;
  double *P = direct ? nullptr
                     : p.partial + ((int64_t)rsel * p.nbC + cb) *
                                       (int64_t)p.M * p.N;
  // Direct stores go one 16-row block at a time: all its reads of D and C
  // are issued before any store (a store may alias them as far as the
  // compiler knows, which would otherwise serialise every element on a
  // round trip to memory).
  // The reads of D and C are unconditional (a dropped element reads
  // element 0 and is not stored) and every value is formed before the first
  // store: with reads or stores under divergent branches the compiler can
  // no longer count which memory operations are outstanding and waits for
  // all of them before each store, serialising the stores one round trip
  // at a time.
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    if (direct) {
      int64_t oc[CT][4];
      double res[CT][4];
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int lc = wn * (TN / 2) + ct * 16 + li;
          const int lr = wm * (TM / 2) + rt * 16 + Mma<T>::crow(lg, v);
          const bool ok = crow[lr] >= 0 && ccol[lc] >= 0;
          oc[ct][v] = ok ? crow[lr] + ccol[lc] : -1;
          res[ct][v] = p.alpha * (double)acc[rt][ct][v];
        }
      if (D) {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int lc = wn * (TN / 2) + ct * 16 + li;
            const int lr = wm * (TM / 2) + rt * 16 + Mma<T>::crow(lg, v);
            const double dv =
                D[oc[ct][v] >= 0 ? drow[lr] + dcol[lc] : (int64_t)0];
            res[ct][v] += p.gamma * dv;
          }
      }
      if (p.beta != 0.0) {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double cv = C[oc[ct][v] >= 0 ? oc[ct][v] : (int64_t)0];
            res[ct][v] += p.beta * cv;
          }
      }
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (oc[ct][v] >= 0) C[oc[ct][v]] = res[ct][v];
    } else {
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int lc = wn * (TN / 2) + ct * 16 + li;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int lr = wm * (TM / 2) + rt * 16 + Mma<T>::crow(lg, v);
          if (m0 + lr < p.M && n0 + lc < p.N)
            P[(int64_t)(m0 + lr) * p.N + n0 + lc] = (double)acc[rt][ct][v];
        }
      }
    }
  }
}

// C[b][m,n] = alpha * sum_r partial[r][b][m][n] + beta*C  (fixed order)
// Sum of the R split-K partials of each output element, then the epilogue.
// A block of RG x 64 threads takes 64 consecutive elements (coalesced rows
// of the partial buffer); its RG groups sum interleaved partials (several
// loads in flight per thread) and combine in LDS, so that even a small
// output (a few thousand elements) spreads its R-deep sums over the chip.
constexpr int RG = 16;
__global__ __launch_bounds__(64 * RG) void reduce_partial_kernel(GemmDev p,
                                                                 int R) {
  __shared__ double part[RG][64];
  const int64_t MN = (int64_t)p.M * p.N;
  const int64_t total = MN * p.nbC;
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  for (int64_t base = (int64_t)blockIdx.x * 64; base < total;
       base += (int64_t)gridDim.x * 64) {
    const int64_t idx = base + lane;
    const int64_t cb = idx / MN;
    const int64_t mn = idx % MN;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (idx < total) {
      const double *src = p.partial + cb * MN + mn;
      const int64_t step = (int64_t)p.nbC * MN;
      int r = g;
      for (; r + 3 * RG < R; r += 4 * RG) {
        s0 += src[(int64_t)r * step];
        s1 += src[(int64_t)(r + RG) * step];
        s2 += src[(int64_t)(r + 2 * RG) * step];
        s3 += src[(int64_t)(r + 3 * RG) * step];
      }
      for (; r < R; r += RG) s0 += src[(int64_t)r * step];
    }
    part[g][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (g == 0 && idx < total) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < RG; ++q) s += part[q][lane];
      const int m0 = (int)(mn / p.N), n0 = (int)(mn % p.N);
      const int m = p.rmap ? p.rmap[m0] : m0, n = p.cmap ? p.cmap[n0] : n0;
      if (m >= 0 && n >= 0) {
        int64_t b1, b2;
        if (p.reduce_b2) {
          b1 = cb;
          b2 = 0;
        } else {
          b1 = cb / p.nb2;
          b2 = cb % p.nb2;
        }
        const Lvl lr = map_level(levels(m, p.M1, p.M2), p.mmap, p.mlvl),
                  lc = map_level(levels(n, p.N1, p.N2), p.nmap, p.nlvl);
        double *c = p.C + b1 * p.sC1 + b2 * p.sC2 +
                    at(lr, p.sCm0, p.sCm1, p.sCm) +
                    at(lc, p.sCn0, p.sCn1, p.sCn);
        double y = p.alpha * s;
        if (p.D)
          y += p.gamma * p.D[b1 * p.sD1 + b2 * p.sD2 +
                             at(lr, p.sDm0, p.sDm1, p.sDm) +
                             at(lc, p.sDn0, p.sDn1, p.sDn)];
        if (p.beta != 0.0) y += p.beta * *c;
        *c = y;
      }
    }
    __syncthreads();
  }
}

int choose_splits(const Gemm &g, int tiles, int bk) {
  // Few output tiles over a long K (the O(n^5) contractions onto n x n
  // matrices) stream their operands from HBM once: split K until ~4
  // workgroups per CU keep enough loads in flight; >= 8 slabs per split.
  const int64_t blocks = (int64_t)tiles * g.nb1 * g.nb2;
  const int64_t target = std::max(1, g.split_target);
  if (blocks >= target) return 1;
  const int maxs = std::max(1, g.K / (8 * bk));
  int want = (int)((target + blocks - 1) / blocks);
  return std::max(1, std::min(want, std::min(maxs, 2048)));
}

template <typename K3>
void launch(K3 kernel, const GemmDev &d, hipStream_t s) {
  const int64_t nblocks =
      (int64_t)d.tilesM * d.tilesN * d.splits * d.nb1 * d.nb2;
  hipLaunchKernelGGL(kernel, dim3((unsigned)nblocks), dim3(NT), 0, s, d);
  FTCC_CHECK_HIP(hipGetLastError());
}

template <typename T, int TM, int TN>
void launch_tile(bool amf, bool bnf, bool kl, const GemmDev &d,
                 hipStream_t s) {
  if (kl) launch(gemm_mfma_kernel<T, false, false, TM, TN, true>, d, s);
  else if (amf && bnf)
    launch(gemm_mfma_kernel<T, true, true, TM, TN, false>, d, s);
  else if (amf) launch(gemm_mfma_kernel<T, true, false, TM, TN, false>, d, s);
  else if (bnf) launch(gemm_mfma_kernel<T, false, true, TM, TN, false>, d, s);
  else launch(gemm_mfma_kernel<T, false, false, TM, TN, false>, d, s);
}

template <typename T>
void launch_layout(Tile tl, bool amf, bool bnf, bool kl, const GemmDev &d,
                   hipStream_t s) {
  if (tl.m == 64) launch_tile<T, 64, 128>(amf, bnf, kl, d, s);
  else if (tl.n == 64) launch_tile<T, 128, 64>(amf, bnf, kl, d, s);
  else launch_tile<T, 128, 128>(amf, bnf, kl, d, s);
}

}  // namespace

size_t gemm_partial_elems(const Gemm &g) {
  const Tile tl = pick_tile(g.M, g.N, g.K, g.tile_policy);
  const int tiles = ((g.M + tl.m - 1) / tl.m) * ((g.N + tl.n - 1) / tl.n);
  const int bk = BK;  // conservative: smallest slab gives the most splits
  const int splits = choose_splits(g, tiles, bk);
  const int R = splits * (g.reduce_b2 ? g.nb2 : 1);
  if (R <= 1) return 0;
  const int nbC = g.reduce_b2 ? g.nb1 : g.nb1 * g.nb2;
  return (size_t)R * nbC * g.M * g.N;
}

void gemm_run(const Gemm &g, int precision, Scratch &scr, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.nb1 <= 0 || g.nb2 <= 0) return;
  require(precision == FTCC_PREC_FP64 || precision == FTCC_PREC_FP32,
          "the generic GEMM runs FP64 or FP32 (bf16x3 uses sbgemm)");
  const int bk = BK;
  GemmDev d;
  d.M = g.M;
  d.N = g.N;
  d.K = g.K;
  d.nb1 = g.nb1;
  d.nb2 = g.nb2;
  d.reduce_b2 = g.reduce_b2 ? 1 : 0;
  d.alpha = g.alpha;
  d.beta = g.beta;
  d.A = g.A;
  d.sAm = g.sAm;
  d.sAk = g.sAk;
  d.sA1 = g.sA1;
  d.sA2 = g.sA2;
  d.B = g.B;
  d.sBk = g.sBk;
  d.sBn = g.sBn;
  d.sB1 = g.sB1;
  d.sB2 = g.sB2;
  d.C = g.C;
  d.sCm = g.sCm;
  d.sCn = g.sCn;
  d.sC1 = g.sC1;
  d.sC2 = g.sC2;
  const Tile tl = pick_tile(g.M, g.N, g.K, g.tile_policy);
  d.tilesM = (g.M + tl.m - 1) / tl.m;
  d.tilesN = (g.N + tl.n - 1) / tl.n;
  d.m_inner = d.tilesM < d.tilesN ? 1 : 0;
  if (g.K <= 0) {
    // C = beta*C only
    d.splits = 1;
    d.kchunk = 0;
  } else {
    d.splits = choose_splits(g, d.tilesM * d.tilesN, BK);
    // round the chunk to whole slabs of the active precision
    int kc = (g.K + d.splits - 1) / d.splits;
    kc = ((kc + bk - 1) / bk) * bk;
    d.kchunk = kc;
    d.splits = (g.K + kc - 1) / kc;
  }
  const int R = d.splits * (g.reduce_b2 ? g.nb2 : 1);
  d.nbC = g.reduce_b2 ? g.nb1 : g.nb1 * g.nb2;
  d.partial = nullptr;
  d.rmap = g.rmap;
  d.cmap = g.cmap;
  d.D = g.D;
  d.gamma = g.gamma;
  d.sDm = g.sDm;
  d.sDn = g.sDn;
  d.sD1 = g.sD1;
  d.sD2 = g.sD2;
  // index levels of rows and columns (one index: everything in the last)
  require(g.M2 >= 0 && g.N2 >= 0 && g.M1 >= 1 && g.N1 >= 1,
          "GEMM index levels must be positive");
  if (g.M2 > 0)
    require(g.M % ((int64_t)g.M1 * g.M2) == 0,
            "GEMM rows are not a product of their index levels");
  if (g.N2 > 0)
    require(g.N % ((int64_t)g.N1 * g.N2) == 0,
            "GEMM columns are not a product of their index levels");
  d.M1 = g.M2 > 0 ? g.M1 : 1;
  d.M2 = g.M2 > 0 ? g.M2 : (1 << 30);
  d.N1 = g.N2 > 0 ? g.N1 : 1;
  d.N2 = g.N2 > 0 ? g.N2 : (1 << 30);
  d.sAm0 = g.sAm0;
  d.sAm1 = g.sAm1;
  d.sCm0 = g.sCm0;
  d.sCm1 = g.sCm1;
  d.sDm0 = g.sDm0;
  d.sDm1 = g.sDm1;
  d.sBn0 = g.sBn0;
  d.sBn1 = g.sBn1;
  d.sCn0 = g.sCn0;
  d.sCn1 = g.sCn1;
  d.sDn0 = g.sDn0;
  d.sDn1 = g.sDn1;
  require(g.K2 >= 0 && g.K1 >= 1, "GEMM K levels must be positive");
  if (g.K2 > 0)
    require(g.K % ((int64_t)g.K1 * g.K2) == 0,
            "GEMM K is not a product of its index levels");
  d.K1 = g.K2 > 0 ? g.K1 : 1;
  d.K2 = g.K2 > 0 ? g.K2 : (1 << 30);
  d.sAk0 = g.sAk0;
  d.sAk1 = g.sAk1;
  require((g.mlvl < 0 || (g.mmap && g.mlvl <= 2)) &&
              (g.nlvl < 0 || (g.nmap && g.nlvl <= 2)),
          "GEMM level map without a map or level");
  d.mmap = g.mmap;
  d.nmap = g.nmap;
  d.mlvl = g.mmap ? g.mlvl : -1;
  d.nlvl = g.nmap ? g.nlvl : -1;
  d.sBk0 = g.sBk0;
  d.sBk1 = g.sBk1;
  if (R > 1) {
    const size_t need = (size_t)R * d.nbC * g.M * g.N;
    require(need <= scr.partial_elems, "GEMM split-K scratch too small",
            FTCC_ENOMEM);
    d.partial = scr.partial;
  }
  const bool amf = (g.sAm == 1 && g.sAk != 1);
  const bool bnf = (g.sBn == 1 && g.sBk != 1);
  const bool kl = g.K2 > 0;
  require(!kl || (!amf && !bnf),
          "K over several indices needs both operands K-fastest");
  if (precision == FTCC_PREC_FP64)
    launch_layout<double>(tl, amf, bnf, kl, d, s);
  else
    launch_layout<float>(tl, amf, bnf, kl, d, s);
  if (R > 1) {
    const int64_t total = (int64_t)d.nbC * g.M * g.N;
    const int blocks = (int)std::min<int64_t>((total + 63) / 64, 65536);
    hipLaunchKernelGGL(reduce_partial_kernel, dim3(blocks), dim3(64 * RG), 0,
                       s, d, R);
    FTCC_CHECK_HIP(hipGetLastError());
  }
}

}  // namespace ftcc
