// Split-BF16 GEMM ("bf16x3") for the O(n^6) products, and the kernels that
// produce its operands.
//
// Operands are stored in HBM already split: x = hi + lo with hi = bf16(x),
// lo = bf16(x - hi), as two bf16 arrays with K contiguous (A as [M][ld],
// B transposed as [N][ld]); K is zero-padded to a multiple of 32 so the
// inner loop needs no bounds checks along K.  C = alpha * A.B + beta * C is
// accumulated in FP32 by v_mfma_f32_32x32x16_bf16 from lo*hi + hi*lo + hi*hi
// (in that order for every accumulator) and written as FP64.
//
// Tiling: a 512-thread workgroup (8 wave64s, two per SIMD) owns a 256x256
// tile of C, each wave a 128x64 quadrant (4x2 MFMA 32x32 blocks, 128 FP32
// accumulators per lane).  K advances in slabs of 32 through two LDS stages
// of 64 KB; each LDS row holds the row's 32 hi values and 32 lo values as
// eight 16-byte chunks, chunk c stored at slot c ^ ((row >> 1) & 7) so that
// the MFMA fragment reads (32 rows at one chunk per half-wave) hit distinct
// banks.  Slabs go from HBM straight into LDS (global_load_lds_dwordx4: each
// lane fetches the global chunk that belongs at its LDS slot, no register
// staging), two slabs ahead of their use, into the stage the previous slab
// has left.  Per slab a wave reads the fragments of the next 16-deep K step
// while the MFMAs of the current one run (fragments double-buffered in
// registers), so one barrier per slab is the only synchronisation; the loop
// body has no branches (the loads past the last slab re-read the last one
// into a stage that is never read).  Rows past M (N) read the last row:
// their products land in accumulators that are never stored.  Workgroup
// ids are remapped so that each XCD works on a compact group of tiles
// (shared A/B panels stay in that XCD's L2).  Against staging the slabs
// through registers with padded LDS rows (the earlier form of this
// kernel), 8-13% shorter launches at n = 114 shapes with bit-identical
// results; one wave per SIMD with 128x128 wave tiles was slower.
#include "ftcc_internal.h"

#include <algorithm>

namespace ftcc {
namespace {

typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 256, BN = 256, BK = 32;
constexpr int GROUP_M = 8;
constexpr int ROWB = 128;              // bytes per LDS row: 32 hi, 32 lo
constexpr int OPB = BM * ROWB;         // one operand of one stage: 32 KB
constexpr int STAGEB = 2 * OPB;        // A then B
constexpr size_t LDS_BYTES = 2 * (size_t)STAGEB;
typedef __attribute__((address_space(3))) void lds_void;

// One product of a grouped launch (the spin blocks of one O(n^6) product
// share C, alpha, beta and the batch; up to MAXP problems per launch).
constexpr int MAXP = 3;
struct SbProb {
  int M, N, Kp, tilesM, tilesN;
  int tile0;                 // first linear tile of this problem
  const __bf16 *Ahi, *Alo;
  int64_t lda, sAb;
  const __bf16 *Bhi, *Blo;
  int64_t ldb, sBb;
  const int *rmap, *cmap;    // output scatter (nullptr: identity; -1 drops)
};
struct SbDev {
  int nprob, nb;
  int group_m;               // tile rows per band of the launch order
  SbProb pr[MAXP];
  double *C;
  int64_t sCm, sCn, sCb;
  double alpha, beta;
};

// Workgroup -> (problem, batch, tile m, tile n).  Consecutive workgroup ids
// go to different XCDs, so ids are first regrouped per XCD (each XCD then
// takes a contiguous run of linear tiles); inside a problem, tiles run in
// GROUP_M-row bands so that an XCD's concurrent tiles share A/B panels.
__device__ inline SbProb tile_of(const SbDev &p, int &bt, int &tm, int &tn) {
  const int nblk = gridDim.x, b = blockIdx.x;
  const int q = nblk / 8, r = nblk % 8, x = b % 8, j = b / 8;
  const int lid = (x < r) ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
  SbProb pr = p.pr[0];
  if (p.nprob > 1 && lid >= p.pr[1].tile0) pr = p.pr[1];
  if (p.nprob > 2 && lid >= p.pr[2].tile0) pr = p.pr[2];
  const int local = lid - pr.tile0;
  const int per = pr.tilesM * pr.tilesN;
  bt = local / per;
  const int t = local % per;
  const int span = p.group_m * pr.tilesN;
  const int first_m = (t / span) * p.group_m;
  const int gsz = min(pr.tilesM - first_m, p.group_m);
  tm = first_m + (t % span) % gsz;
  tn = (t % span) / gsz;
  return pr;
}

// LDS slot of chunk c (0-3 hi, 4-7 lo, eight bf16 each) of row r
__device__ inline int lds_slot(int row, int c) { return c ^ ((row >> 1) & 7); }

// One wave's share of a slab of one operand: rows [32w, 32w + 32) of the
// tile, four loads of eight rows x eight chunks; the lane that lands at
// slot (lane & 7) of row r fetches the chunk stored there.
__device__ inline void load_slab(const __bf16 *hi, const __bf16 *lo,
                                 int64_t ld, int row0, int rows, int k0,
                                 char *lds_op, int wave, int lane) {
  const int r8 = lane >> 3, slot = lane & 7;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int rowl = wave * 32 + j * 8 + r8;
    const int c = lds_slot(rowl, slot);   // the slot map is an involution
    int g = row0 + rowl;
    g = g < rows ? g : rows - 1;
    const __bf16 *src = (c < 4 ? hi : lo) + (int64_t)g * ld + k0 + 8 * (c & 3);
    __builtin_amdgcn_global_load_lds(
        (const void *)src, (lds_void *)(lds_op + (wave * 32 + j * 8) * ROWB),
        16, 0, 0);
  }
}

__device__ inline bf8_t lds_frag(const char *lds_op, int row, int c) {
  return *(const bf8_t *)(lds_op + row * ROWB + (lds_slot(row, c) << 4));
}

// the A (4 row blocks) and B (2 column blocks) fragments, hi and lo, of one
// 16-deep K step (chunk c: 0 or 2, plus the lane's half)
struct Frag {
  bf8_t ah[4], al[4], bh[2], bl[2];
};
__device__ inline void read_frag(Frag &f, const char *As, const char *Bs,
                                 int wm, int wn, int li, int c) {
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const int row = wm * 128 + rt * 32 + li;
    f.ah[rt] = lds_frag(As, row, c);
    f.al[rt] = lds_frag(As, row, c + 4);
  }
#pragma unroll
  for (int ct = 0; ct < 2; ++ct) {
    const int row = wn * 64 + ct * 32 + li;
    f.bh[ct] = lds_frag(Bs, row, c);
    f.bl[ct] = lds_frag(Bs, row, c + 4);
  }
}
// lo*hi + hi*lo + hi*hi, in that order for every accumulator
__device__ inline void mfma_step(f16_t (&acc)[4][2], const Frag &f) {
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
          f.al[rt], f.bh[ct], acc[rt][ct], 0, 0, 0);
      acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
          f.ah[rt], f.bl[ct], acc[rt][ct], 0, 0, 0);
      acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
          f.ah[rt], f.bh[ct], acc[rt][ct], 0, 0, 0);
    }
}

__global__ __launch_bounds__(512, 1) void sbgemm_kernel(SbDev p) {
  extern __shared__ char lds[];   // [2 stages][A | B][256 rows][128 B]
  // the tile's output rows and columns after the scatter maps (-1: past
  // the end or dropped), read once here, alongside the first slab's loads
  __shared__ int emap[BM + BN];
  int bt, tm, tn;
  const SbProb pr = tile_of(p, bt, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  {
    const int t = threadIdx.x;   // 512 threads: 256 rows, then 256 columns
    if (t < BM) {
      const int row = m0 + t;
      emap[t] = row < pr.M ? (pr.rmap ? pr.rmap[row] : row) : -1;
    } else {
      const int col = n0 + t - BM;
      emap[t] = col < pr.N ? (pr.cmap ? pr.cmap[col] : col) : -1;
    }
  }
  const __bf16 *Ahi = pr.Ahi + bt * pr.sAb, *Alo = pr.Alo + bt * pr.sAb;
  const __bf16 *Bhi = pr.Bhi + bt * pr.sBb, *Blo = pr.Blo + bt * pr.sBb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  const int li = lane & 31, lg = lane >> 5;
  const int nslab = pr.Kp / BK;

  f16_t acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.f;

  // an empty K (a spin class without pairs) only writes beta * C
  Frag f0, f1;
  if (nslab > 0) {
    load_slab(Ahi, Alo, pr.lda, m0, pr.M, 0, lds, wave, lane);
    load_slab(Bhi, Blo, pr.ldb, n0, pr.N, 0, lds + OPB, wave, lane);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const int k1 = (nslab > 1 ? 1 : 0) * BK;
    load_slab(Ahi, Alo, pr.lda, m0, pr.M, k1, lds + STAGEB, wave, lane);
    load_slab(Bhi, Blo, pr.ldb, n0, pr.N, k1, lds + STAGEB + OPB, wave,
              lane);
    read_frag(f0, lds, lds + OPB, wm, wn, li, lg);
  }
  for (int s = 0; s < nslab; ++s) {
    char *cur = lds + (s & 1) * STAGEB, *nxt = lds + ((s + 1) & 1) * STAGEB;
    // second K step of slab s from LDS while the first one's MFMAs run
    read_frag(f1, cur, cur + OPB, wm, wn, li, 2 + lg);
    mfma_step(acc, f0);
    // slab s + 1 has landed (every wave's loads) and every wave is done
    // with stage s: slab s + 2 goes there, slab s + 1's first K step is read
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const int k2 = min(s + 2, nslab - 1) * BK;
    load_slab(Ahi, Alo, pr.lda, m0, pr.M, k2, cur, wave, lane);
    load_slab(Bhi, Blo, pr.ldb, n0, pr.N, k2, cur + OPB, wave, lane);
    read_frag(f0, nxt, nxt + OPB, wm, wn, li, lg);
    mfma_step(acc, f1);
  }
  // no load may still be writing LDS when the workgroup ends
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  // an empty K passed no barrier since the scatter maps went to LDS
  if (nslab == 0) __syncthreads();

  // lane holds column li, rows 8*(v/4) + 4*lg + v%4 of each 32x32 block.
  // Per block, the reads of C (beta != 0) are unconditional (a dropped
  // element reads element 0) and every value is formed before the first
  // store: reads or stores under divergent branches make the compiler wait
  // for all outstanding memory operations before each store.
  double *C = p.C + bt * p.sCb;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int cc = emap[BM + wn * 64 + ct * 32 + li];
      int64_t off[16];
      double res[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rr =
            emap[wm * 128 + rt * 32 + 8 * (v >> 2) + 4 * lg + (v & 3)];
        off[v] = (rr >= 0 && cc >= 0)
                     ? (int64_t)rr * p.sCm + (int64_t)cc * p.sCn
                     : (int64_t)-1;
        res[v] = p.alpha * (double)acc[rt][ct][v];
      }
      if (p.beta != 0.0) {
#pragma unroll
        for (int v = 0; v < 16; ++v)
          res[v] += p.beta * C[off[v] >= 0 ? off[v] : (int64_t)0];
      }
#pragma unroll
      for (int v = 0; v < 16; ++v)
        if (off[v] >= 0) C[off[v]] = res[v];
    }
}

__device__ inline void split1(double v, __bf16 &hi, __bf16 &lo) {
  const float f = (float)v;
  hi = (__bf16)f;
  lo = (__bf16)(f - (float)hi);
}

// Destination writers of the gather kernels: split into (hi, lo) bf16, or
// plain FP64 (the spin-blocked operands of the FP64 and FP32 modes).
// store4 writes four consecutive elements (o a multiple of 4).
struct SplitOut {
  __bf16 *hi, *lo;
  __device__ void operator()(int64_t o, double v) const {
    __bf16 h, l;
    split1(v, h, l);
    hi[o] = h;
    lo[o] = l;
  }
  __device__ void store4(int64_t o, const double v[4]) const {
    typedef __bf16 b4_t __attribute__((ext_vector_type(4)));
    b4_t h, l;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      __bf16 hk, lk;
      split1(v[k], hk, lk);
      h[k] = hk;
      l[k] = lk;
    }
    *(b4_t *)(hi + o) = h;
    *(b4_t *)(lo + o) = l;
  }
};
struct F64Out {
  double *p;
  __device__ void operator()(int64_t o, double v) const { p[o] = v; }
  __device__ void store4(int64_t o, const double v[4]) const {
    typedef double d2_t __attribute__((ext_vector_type(2)));
    *(d2_t *)(p + o) = d2_t{v[0], v[1]};
    *(d2_t *)(p + o + 2) = d2_t{v[2], v[3]};
  }
};

// source row / column of destination index i: map[i] (-1: zero) or i itself
// below the bound n
__device__ inline int src_index(const int *map, int i, int n) {
  return map ? map[i] : (i < n ? i : -1);
}

// The source columns of columns c0, c0 + step, ... (NQ of them): -1 past W
// or where the map drops them.  The map test is hoisted (it is uniform) and
// the map reads are unconditional, so they are in flight together.
template <int NQ>
__device__ inline void map_cols(const int *map, int n, int W, int c0,
                                int step, int out[NQ]) {
  if (map) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int c = c0 + q * step;
      const int m = map[c < W ? c : 0];
      out[q] = c < W ? m : -1;
    }
  } else {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int c = c0 + q * step;
      out[q] = (c < W && c < n) ? c : -1;
    }
  }
}

// Sources of the gather kernels: value of the dense matrix at (slice b,
// row pair rr, column pair cc).  Plain: a strided FP64 tensor.
struct PlainSrc {
  const double *src;
  int64_t sr, sc, sb;
  __device__ double operator()(int64_t b, int rr, int cc) const {
    return src[b * sb + (int64_t)rr * sr + (int64_t)cc * sc];
  }
};
// Ring tensor W[e][m][b][j] (slice stride sb) of a spin-flip symmetric
// problem whose rows of an alpha e are stored (alpha runs [0, hv), [0, ho),
// beta runs after them), read by rows (b,j) and columns (e,m) (pair indices
// x * no + y): an element of a beta e is read from its image
// W[e - hv][fl m][fl b][fl j].
struct FlipRingSrc {
  const double *src;
  int64_t sb;
  int nv, no, hv, ho;
  __device__ double operator()(int64_t b, int rr, int cc) const {
    int e = cc / no, m = cc - e * no;
    int q = rr / no, l = rr - q * no;
    if (e >= hv) {
      e -= hv;
      m = m < ho ? m + ho : m - ho;
      q = q < hv ? q + hv : q - hv;
      l = l < ho ? l + ho : l - ho;
    }
    return src[b * sb + ((int64_t)e * no + m) * ((int64_t)nv * no) +
               (int64_t)q * no + l];
  }
};
// Amplitude-derived operands computed from t1 [V][O] and t2 [V][V][O][O]
// of each slice as they are gathered (nothing is materialised):
//   AMP_TAU_AB: rows (a,b), columns (i,j): tau = t2 + t1 t1 - t1 t1
//   AMP_TAU_IJ: rows (i,j), columns (a,b): the same tau
//   AMP_T2R:    rows (a,i), columns (e,m): t2[a,e,i,m]
//   AMP_ARING:  rows (b,j), columns (f,n): -1/2 t2[f,b,n,j] + t1[f,j] t1[b,n]
// (pair indices x * ny + y of the pair maps), one kind per instantiation (no
// branch on the kind inside the gathers' loads).
template <int KIND>
struct AmpSrc {
  const double *t1, *t2;
  int nv, no;
  __device__ double operator()(int64_t b, int rr, int cc) const {
    const int64_t VO = (int64_t)nv * no;
    const double *x1 = t1 + b * VO, *x2 = t2 + b * VO * VO;
    auto T2 = [&](int a, int e, int i, int m) {
      return x2[(((int64_t)a * nv + e) * no + i) * no + m];
    };
    auto T1 = [&](int a, int i) { return x1[(int64_t)a * no + i]; };
    if (KIND == AMP_TAU_AB || KIND == AMP_TAU_IJ) {
      const int ab = KIND == AMP_TAU_AB ? rr : cc;
      const int ij = KIND == AMP_TAU_AB ? cc : rr;
      const int a = ab / nv, bb = ab % nv, i = ij / no, j = ij % no;
      return T2(a, bb, i, j) + T1(a, i) * T1(bb, j) - T1(bb, i) * T1(a, j);
    } else if (KIND == AMP_T2R) {
      const int a = rr / no, i = rr % no, e = cc / no, m = cc % no;
      return T2(a, e, i, m);
    } else {   // AMP_ARING
      const int bb = rr / no, j = rr % no, f = cc / no, n = cc % no;
      return -0.5 * T2(f, bb, n, j) + T1(f, j) * T1(bb, n);
    }
  }
};

// dst[b][r][c] (row stride os) = src(b, R(r), C(c)) for r < R, c < W, zero
// where a map gives -1 (or, without maps, for c >= C); threads run along c
// (use when the source is contiguous along columns).
// Each thread takes four consecutive columns (W, os and db multiples of 4,
// destinations aligned to four elements).
template <typename Src, typename Out>
__global__ void gather_rows_kernel(Src src, const int *rmap, int R,
                                   const int *cmap, int C, int W, int64_t os,
                                   int64_t db, Out out, int nb) {
  const int W4 = W / 4;
  const int64_t per = (int64_t)R * W4;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
       idx < per * nb; idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = idx / per, e = idx % per;
    const int r = (int)(e / W4);
    const int c = 4 * (int)(e % W4);
    const int rr = src_index(rmap, r, R);
    int cc[4];
    map_cols<4>(cmap, C, W, c, 1, cc);
    double v[4];
    // branch-free loads (out of range reads element (0, 0), then zero)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double x = src(b, rr < 0 ? 0 : rr, cc[k] < 0 ? 0 : cc[k]);
      v[k] = (rr >= 0 && cc[k] >= 0) ? x : 0.0;
    }
    out.store4(b * db + (int64_t)r * os + c, v);
  }
}

// Same map through a GT x GT LDS tile: reads run along r (the source is
// contiguous along rows), writes along c, four columns per thread.  32x32
// tiles measured faster than 64x64 at n = 114 (24.2 against 30.0 ms of
// gathers per iteration: more workgroups in flight).
constexpr int GT = 32;
template <typename Src, typename Out>
__global__ void gather_transpose_kernel(Src src, const int *rmap, int R,
                                        const int *cmap, int C, int W,
                                        int64_t os, int64_t db, Out out) {
  __shared__ double tile[GT][GT + 1];
  const int b = blockIdx.z;
  const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
  const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;   // 256 threads
  // a fixed trip count with branch-free loads (an out-of-range element
  // reads element (0, 0) and is replaced by zero): every row's map and
  // source loads are in flight together instead of one round trip each
  constexpr int NQ = GT / (256 / GT);
  const int r = r0 + tx;
  const int rr = r < R ? src_index(rmap, r, R) : -1;
  int cc[NQ];
  map_cols<NQ>(cmap, C, W, c0 + ty, 256 / GT, cc);
  double v[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const double x = src(b, rr < 0 ? 0 : rr, cc[q] < 0 ? 0 : cc[q]);
    v[q] = (rr >= 0 && cc[q] >= 0) ? x : 0.0;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) tile[ty + q * (256 / GT)][tx] = v[q];
  __syncthreads();
  const int cq = threadIdx.x % (GT / 4), rq = threadIdx.x / (GT / 4);
  for (int i = rq; i < GT; i += 1024 / GT) {
    const int r = r0 + i, c = c0 + 4 * cq;
    if (r < R && c < W) {
      double v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) v[k] = tile[4 * cq + k][i];
      out.store4(b * db + (int64_t)r * os + c, v);
    }
  }
}

template <typename Src, typename Out>
void gather_launch(Src src, bool along_rows, const int *rmap, int R,
                   const int *cmap, int C, int W, int64_t os, int64_t db,
                   Out out, int nb, hipStream_t s) {
  if ((int64_t)R * W == 0 || nb <= 0) return;
  require(W % 4 == 0 && os % 4 == 0 && db % 4 == 0,
          "gather: widths and strides must be multiples of 4");
  if (along_rows) {
    const int gx = (W + GT - 1) / GT, gy = (R + GT - 1) / GT;
    require(gy < 65536 && nb < 65536, "gather grid too large");
    hipLaunchKernelGGL((gather_transpose_kernel<Src, Out>), dim3(gx, gy, nb),
                       dim3(256), 0, s, src, rmap, R, cmap, C, W, os, db,
                       out);
  } else {
    const int64_t total = (int64_t)R * (W / 4) * nb;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL((gather_rows_kernel<Src, Out>), dim3(blocks),
                       dim3(256), 0, s, src, rmap, R, cmap, C, W, os, db, out,
                       nb);
  }
  FTCC_CHECK_HIP(hipGetLastError());
}

// a plain source is read along rows when its row stride is the unit one
PlainSrc plain(const double *src, int64_t sr, int64_t sc, int64_t sb,
               bool &along_rows) {
  along_rows = (sr == 1 && sc != 1);
  return PlainSrc{src, sr, sc, sb};
}
// gather_launch over the amplitude source of `kind` (tau by (i,j) rows and
// aring by (b,j) rows are contiguous along rows)
template <typename Out>
void amp_launch(int kind, const double *t1, const double *t2, int nv, int no,
                const int *rmap, int R, const int *cmap, int W, int64_t os,
                int64_t db, Out out, int nb, hipStream_t s) {
  const bool ar = (kind == AMP_TAU_IJ || kind == AMP_ARING);
  switch (kind) {
    case AMP_TAU_AB:
      gather_launch(AmpSrc<AMP_TAU_AB>{t1, t2, nv, no}, ar, rmap, R, cmap, W,
                    W, os, db, out, nb, s);
      break;
    case AMP_TAU_IJ:
      gather_launch(AmpSrc<AMP_TAU_IJ>{t1, t2, nv, no}, ar, rmap, R, cmap, W,
                    W, os, db, out, nb, s);
      break;
    case AMP_T2R:
      gather_launch(AmpSrc<AMP_T2R>{t1, t2, nv, no}, ar, rmap, R, cmap, W, W,
                    os, db, out, nb, s);
      break;
    default:
      gather_launch(AmpSrc<AMP_ARING>{t1, t2, nv, no}, ar, rmap, R, cmap, W,
                    W, os, db, out, nb, s);
  }
}

}  // namespace

int64_t split_ld(int K) { return ((int64_t)K + 63) / 64 * 64; }

void split_convert(const double *src, int64_t sr, int64_t sc, int64_t sb,
                   int R, int C, const SplitMat &d, int nb, hipStream_t s) {
  require(d.ld >= C && d.ld % 32 == 0, "split operand leading dimension");
  if ((int64_t)R * C == 0) return;
  bool ar;
  const PlainSrc src_ = plain(src, sr, sc, sb, ar);
  gather_launch(src_, ar, nullptr, R, nullptr, C, (int)d.ld, d.ld, d.bstride,
                SplitOut{d.hi, d.lo}, nb, s);
}

void split_gather(const double *src, int64_t sr, int64_t sc, int64_t sb,
                  const int *rmap, int R, const int *cmap, int W,
                  const SplitMat &d, int nb, hipStream_t s) {
  require(d.ld % 32 == 0 && W <= d.ld, "split operand leading dimension");
  bool ar;
  const PlainSrc src_ = plain(src, sr, sc, sb, ar);
  gather_launch(src_, ar, rmap, R, cmap, W, W, d.ld, d.bstride,
                SplitOut{d.hi, d.lo}, nb, s);
}

void amp_split_gather(int kind, const double *t1, const double *t2, int nv,
                      int no, const int *rmap, int R, const int *cmap, int W,
                      const SplitMat &d, int nb, hipStream_t s) {
  require(d.ld % 32 == 0 && W <= d.ld, "split operand leading dimension");
  require(rmap && cmap, "amplitude operands are gathered through pair maps");
  amp_launch(kind, t1, t2, nv, no, rmap, R, cmap, W, d.ld, d.bstride,
             SplitOut{d.hi, d.lo}, nb, s);
}

void amp_gather64(int kind, const double *t1, const double *t2, int nv,
                  int no, const int *rmap, int R, const int *cmap, int W,
                  int64_t os, double *dst, int64_t db, int nb,
                  hipStream_t s) {
  require(rmap && cmap, "amplitude operands are gathered through pair maps");
  amp_launch(kind, t1, t2, nv, no, rmap, R, cmap, W, os, db, F64Out{dst}, nb,
             s);
}

void flip_ring_gather(const double *src, int nv, int no, int hv, int ho,
                      int64_t sb, const int *rmap, int R, const int *cmap,
                      int W, const SplitMat *d, int64_t os, double *dst,
                      int64_t db, int nb, hipStream_t s) {
  require(2 * hv == nv && 2 * ho == no, "flip gather needs equal spin halves");
  require(rmap && cmap, "flip gather reads through pair maps");
  const FlipRingSrc src_{src, sb, nv, no, hv, ho};
  if (d) {
    require(d->ld % 32 == 0 && W <= d->ld, "split operand leading dimension");
    gather_launch(src_, true, rmap, R, cmap, W, W, d->ld, d->bstride,
                  SplitOut{d->hi, d->lo}, nb, s);
  } else {
    gather_launch(src_, true, rmap, R, cmap, W, W, os, db, F64Out{dst}, nb, s);
  }
}

void gather64(const double *src, int64_t sr, int64_t sc, int64_t sb,
              const int *rmap, int R, const int *cmap, int W, int64_t os,
              double *dst, int64_t db, int nb, hipStream_t s) {
  bool ar;
  const PlainSrc src_ = plain(src, sr, sc, sb, ar);
  gather_launch(src_, ar, rmap, R, cmap, W, W, os, db, F64Out{dst}, nb, s);
}

void sbgemm_grouped(int nprob, const SbProblem *prob, int nb, double alpha,
                    double beta, double *C, int64_t sCm, int64_t sCn,
                    int64_t sCb, hipStream_t s) {
  require(nprob >= 1 && nprob <= MAXP, "1 to 3 problems per grouped launch");
  static std::once_flag attr;
  set_max_lds_once(attr, (const void *)sbgemm_kernel, (int)LDS_BYTES);
  // band height of the launch order (FTCC_SB_GROUP_M tunes it)
  static const int group_m = [] {
    const char *e = getenv("FTCC_SB_GROUP_M");
    const int v = e ? atoi(e) : GROUP_M;
    return v >= 1 && v <= 64 ? v : GROUP_M;
  }();
  SbDev p;
  p.nb = nb;
  p.group_m = group_m;
  p.C = C;
  p.sCm = sCm;
  p.sCn = sCn;
  p.sCb = sCb;
  p.alpha = alpha;
  p.beta = beta;
  p.nprob = 0;
  int64_t tiles = 0;
  for (int k = 0; k < nprob; ++k) {
    const SbProblem &q = prob[k];
    if (q.M <= 0 || q.N <= 0) continue;
    require(q.A.ld % 32 == 0 && q.B.ld % 32 == 0 && q.A.ld >= q.K &&
                q.B.ld >= q.K,
            "split operands need K padded to 32");
    SbProb &d = p.pr[p.nprob++];
    d.M = q.M;
    d.N = q.N;
    d.Kp = (int)((q.K + BK - 1) / BK * BK);
    d.tilesM = (q.M + BM - 1) / BM;
    d.tilesN = (q.N + BN - 1) / BN;
    d.tile0 = (int)tiles;
    d.Ahi = q.A.hi;
    d.Alo = q.A.lo;
    d.lda = q.A.ld;
    d.sAb = q.A.bstride;
    d.Bhi = q.B.hi;
    d.Blo = q.B.lo;
    d.ldb = q.B.ld;
    d.sBb = q.B.bstride;
    d.rmap = q.rmap;
    d.cmap = q.cmap;
    tiles += (int64_t)d.tilesM * d.tilesN * nb;
  }
  if (p.nprob == 0 || nb <= 0) return;
  require(tiles < (1LL << 31), "too many GEMM tiles");
  hipLaunchKernelGGL(sbgemm_kernel, dim3((unsigned)tiles), dim3(512),
                     LDS_BYTES, s, p);
  FTCC_CHECK_HIP(hipGetLastError());
}

void sbgemm(int M, int N, int K, int nb, double alpha, const SplitMat &A,
            const SplitMat &B, double beta, double *C, int64_t sCm,
            int64_t sCn, int64_t sCb, hipStream_t s, const int *rmap,
            const int *cmap) {
  if (M <= 0 || N <= 0 || nb <= 0) return;
  SbProblem q{M, N, K, A, B, rmap, cmap};
  sbgemm_grouped(1, &q, nb, alpha, beta, C, sCm, sCn, sCb, s);
}

}  // namespace ftcc
