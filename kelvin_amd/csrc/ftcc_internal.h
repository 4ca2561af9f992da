// Internal declarations shared by the libftcc translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ftcc.h"

#define FTCC_CHECK_HIP(expr)                                                  \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      throw ftcc::Error(FTCC_EHIP, std::string(#expr) + ": " +                \
                                      hipGetErrorString(e_));                 \
    }                                                                         \
  } while (0)

namespace ftcc {

struct Error {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

inline void require(bool ok, const std::string &msg, int code = FTCC_EINVAL) {
  if (!ok) throw Error(code, msg);
}

// Raise a kernel's dynamic LDS limit to `bytes` once per process; safe when
// several host threads launch (a failed attempt is retried by the next).
inline void set_max_lds_once(std::once_flag &flag, const void *kernel,
                             int bytes) {
  std::call_once(flag, [&] {
    FTCC_CHECK_HIP(hipFuncSetAttribute(
        kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  });
}

// ---------------------------------------------------------------------------
// Batched strided GEMM:  C = alpha * A.B + beta * C  (all float64 in memory)
//   A(m,k) at A + b1*sA1 + b2*sA2 + m*sAm + k*sAk   (b1 < nb1, b2 < nb2)
//   B(k,n) at B + b1*sB1 + b2*sB2 + k*sBk + n*sBn
//   C(m,n) at C + b1*sC1 + b2*sC2 + m*sCm + n*sCn
// If reduce_b2 is set, the nb2 products are summed into one C per b1
// (sC2 is ignored): used for contractions whose K index is split over a
// stride break.
// ---------------------------------------------------------------------------
struct Gemm {
  int M = 0, N = 0, K = 0;
  int nb1 = 1, nb2 = 1;
  bool reduce_b2 = false;
  double alpha = 1.0, beta = 0.0;
  const double *A = nullptr;
  int64_t sAm = 0, sAk = 0, sA1 = 0, sA2 = 0;
  const double *B = nullptr;
  int64_t sBk = 0, sBn = 0, sB1 = 0, sB2 = 0;
  double *C = nullptr;
  int64_t sCm = 0, sCn = 0, sC1 = 0, sC2 = 0;
  // optional scatter of the output: row m of the product is written to row
  // rmap[m] of C (column n to cmap[n]); -1 drops it
  const int *rmap = nullptr, *cmap = nullptr;
  // optional addend: C = alpha*A.B + beta*C + gamma*D with D laid out like
  // C (its own strides; sD1 = 0 broadcasts one matrix over the slices); C is
  // read only when beta != 0
  const double *D = nullptr;
  double gamma = 1.0;
  int64_t sDm = 0, sDn = 0, sD1 = 0, sD2 = 0;
  // optional rows/columns over up to three tensor indices: row
  // m = (m0*M1 + m1)*M2 + m2 sits at m0*sXm0 + m1*sXm1 + m2*sXm for operand
  // X (A, C, D); columns likewise with N1, N2 and sXn0, sXn1, sXn for B, C,
  // D.  M2 = 0 (N2 = 0) means one index with stride sXm (sXn).
  int M1 = 1, M2 = 0, N1 = 1, N2 = 0;
  int64_t sAm0 = 0, sAm1 = 0, sCm0 = 0, sCm1 = 0, sDm0 = 0, sDm1 = 0;
  int64_t sBn0 = 0, sBn1 = 0, sCn0 = 0, sCn1 = 0, sDn0 = 0, sDn1 = 0;
  // K over up to three indices likewise (K1, K2 and sAk0, sAk1 / sBk0,
  // sBk1 with sAk, sBk innermost; K2 = 0: one index); needs A stored with
  // K fastest and B with K fastest ([N][K])
  int K1 = 1, K2 = 0;
  // optional map of one row (column) level: index x of level mlvl (nlvl,
  // 0..2 as above) is replaced by mmap[x] (nmap[x]) before the strides
  const int *mmap = nullptr, *nmap = nullptr;
  int mlvl = -1, nlvl = -1;
  int64_t sAk0 = 0, sAk1 = 0, sBk0 = 0, sBk1 = 0;
  // launch plan: tile shape policy (0 by the shape: narrow tiles when one
  // dimension is at most 64 or K is; 1 wide: 128x128 unless one dimension is
  // at most 64; 2 narrow: 128x64 or 64x128 by the aspect) and the workgroup
  // count below which K is split
  int tile_policy = 0;
  int split_target = 1024;
};

// Scratch the GEMM engine may use for split-K partial sums and low-precision
// operand copies.  Grown on demand by the context (hipMalloc outside any
// capture).
struct Scratch {
  double *partial = nullptr;
  size_t partial_elems = 0;
  void *lowp = nullptr;
  size_t lowp_bytes = 0;
};

void gemm_run(const Gemm &g, int precision, Scratch &scr, hipStream_t s);
size_t gemm_partial_elems(const Gemm &g);

// Split-BF16 operand: x = hi + lo, two bf16 arrays [rows][ld] (K contiguous,
// zero-padded to ld), batch stride in elements.
struct SplitMat {
  __bf16 *hi = nullptr, *lo = nullptr;
  int64_t ld = 0;
  int64_t bstride = 0;
};
// Spin-class ordering of a pair space (x, y), x from a set with spin labels
// lx, y from a set with labels ly: pairs are grouped by class
// c = 2*lx[x] + ly[y] in a chosen class order, lexicographic inside a class,
// and each class is padded to a multiple of 256.  map[r] is the flat pair
// index x*ny + y of blocked position r (-1 for padding).
struct PairSpace {
  int total = 0;
  int start[4] = {0, 0, 0, 0};   // per class id: blocked offset,
  int len[4] = {0, 0, 0, 0};     //   padded length,
  int count[4] = {0, 0, 0, 0};   //   number of pairs
  int *d_map = nullptr;          // device, length total
};

// padded leading dimension for a K-long operand row
int64_t split_ld(int K);
// d[b][r][c] = src[b*sb + r*sr + c*sc] split into (hi, lo), r < R, c < C
void split_convert(const double *src, int64_t sr, int64_t sc, int64_t sb,
                   int R, int C, const SplitMat &d, int nb, hipStream_t s);
// d[b][r][c] = src[b*sb + rmap[r]*sr + cmap[c]*sc] split into (hi, lo) for
// r < R, c < W (row stride d.ld); entries whose map is -1 are zero
void split_gather(const double *src, int64_t sr, int64_t sc, int64_t sb,
                  const int *rmap, int R, const int *cmap, int W,
                  const SplitMat &d, int nb, hipStream_t s);
// the same gather into FP64: dst[b*db + r*os + c], c < W
void gather64(const double *src, int64_t sr, int64_t sc, int64_t sb,
              const int *rmap, int R, const int *cmap, int W, int64_t os,
              double *dst, int64_t db, int nb, hipStream_t s);
// The gather of a ring tensor W[nb][nv][no][nv][no] (slice stride sb) by
// rows (b,j) and columns (e,m) whose rows of an alpha e only are stored
// (spin-flip symmetric problem; alpha runs [0, hv) and [0, ho), beta runs
// after them): a beta e is read from its image.  Into the split pair *d
// when d is given, else into FP64 dst as gather64.
void flip_ring_gather(const double *src, int nv, int no, int hv, int ho,
                      int64_t sb, const int *rmap, int R, const int *cmap,
                      int W, const SplitMat *d, int64_t os, double *dst,
                      int64_t db, int nb, hipStream_t s);
// Amplitude-derived operands gathered straight from t1 [nb][V][O] and
// t2 [nb][V][V][O][O] (rows / columns are pair indices x * ny + y):
//   AMP_TAU_AB  rows (a,b), columns (i,j): t2 + t1[a,i]t1[b,j] - t1[b,i]t1[a,j]
//   AMP_TAU_IJ  rows (i,j), columns (a,b): the same
//   AMP_T2R     rows (a,i), columns (e,m): t2[a,e,i,m]
//   AMP_ARING   rows (b,j), columns (f,n): -1/2 t2[f,b,n,j] + t1[f,j]t1[b,n]
enum AmpKind { AMP_TAU_AB, AMP_TAU_IJ, AMP_T2R, AMP_ARING };
void amp_split_gather(int kind, const double *t1, const double *t2, int nv,
                      int no, const int *rmap, int R, const int *cmap, int W,
                      const SplitMat &d, int nb, hipStream_t s);
void amp_gather64(int kind, const double *t1, const double *t2, int nv,
                  int no, const int *rmap, int R, const int *cmap, int W,
                  int64_t os, double *dst, int64_t db, int nb, hipStream_t s);
// C[b] = alpha * A[b] . B[b]^T + beta * C[b] with A as [M][K], B as [N][K];
// output row m goes to row rmap[m] of C and column n to cmap[n] when the
// maps are given (-1 drops the element)
void sbgemm(int M, int N, int K, int nb, double alpha, const SplitMat &A,
            const SplitMat &B, double beta, double *C, int64_t sCm,
            int64_t sCn, int64_t sCb, hipStream_t s,
            const int *rmap = nullptr, const int *cmap = nullptr);
// Up to three products sharing C, alpha, beta and the batch in one launch
// (their outputs must be disjoint); tiles are issued in the order given.
struct SbProblem {
  int M, N, K;
  SplitMat A, B;
  const int *rmap, *cmap;
};
void sbgemm_grouped(int nprob, const SbProblem *prob, int nb, double alpha,
                    double beta, double *C, int64_t sCm, int64_t sCn,
                    int64_t sCb, hipStream_t s);

// ---------------------------------------------------------------------------
// System integrals (ueg.hip)
// ---------------------------------------------------------------------------
// antisymmetrised spin-orbital UEG integrals, n = 2*nk (alpha, then beta);
// X_dev: integer wave vectors [nk][3], k_dev: wave vectors [nk][3]
void ueg_eri(int nk, const int *X_dev, const double *k_dev, double aaa,
             double *eri, hipStream_t s);
// fock = diag(en) + sum_r <pr||qr> fo[r]
void mean_field_fock(int n, const double *en, const double *eri,
                     const double *fo, double *fock, hipStream_t s);
// partial triples whose first slots sum to sum_ij <ij||ij> fo_i fo_j
int mp1_partials(int n, const double *eri, const double *fo, double *partials,
                 hipStream_t s);

// ---------------------------------------------------------------------------
// Elementwise / permutation / reduction kernels (tensor_ops.hip)
// ---------------------------------------------------------------------------
// out[i0*o0 + i1*o1 + i2*o2 + i3*o3] = alpha*in[i0,i1,i2,i3] + beta*out[...]
// with `in` dense of shape (d0,d1,d2,d3); batched over nb with strides
// in_b / out_b.
void permute4(const double *in, double *out, int d0, int d1, int d2, int d3,
              int64_t o0, int64_t o1, int64_t o2, int64_t o3, double alpha,
              double beta, int nb, int64_t in_b, int64_t out_b, hipStream_t s);

// out[i0,i1,i2,i3] (dense) = eri[p0[i0],p1[i1],p2[i2],p3[i3]] *
//                            w0[i0]*w1[i1]*w2[i2]*w3[i3]
void gather_weighted4(const double *eri, int n, const int *p0, const int *p1,
                      const int *p2, const int *p3, const double *w0,
                      const double *w1, const double *w2, const double *w3,
                      int d0, int d1, int d2, int d3, double *out,
                      hipStream_t s);

// y[b][r] = alpha * sum_k M[b][r*ldm + k] * x[b][k] + beta*y[b][r]
void gemv_rows(const double *M, int64_t ldm, int64_t sMb, const double *x,
               int64_t sxb, double *y, int64_t syb, int rows, int K, int nb,
               double alpha, double beta, hipStream_t s);
// y[b][c] = alpha * sum_r x[b][r] * M[b][r*ldm + c] + beta*y[b][c]
void gemv_cols(const double *M, int64_t ldm, int64_t sMb, const double *x,
               int64_t sxb, double *y, int64_t syb, int rows, int cols,
               int nb, double alpha, double beta, double *partial,
               size_t partial_elems, hipStream_t s);

// y = a*x + b*y over n elements
void axpby(double *y, const double *x, int64_t n, double a, double b,
           hipStream_t s);
void fill(double *y, int64_t n, double v, hipStream_t s);
// y[k][i] = a*x[i] + b*y[k][i] for k < nb (x shared by all batches)
void bcast(double *y, const double *x, int64_t n, int nb, double a, double b,
           hipStream_t s);

// Fermi factors, sqrt weights and GP0 on the device (FP64)
void fermi_weights(const double *en, int n, double beta, double mu, double *fo,
                   double *fv, double *sfo, double *sfv, double *gp0,
                   hipStream_t s);
void gather_vec(const double *src, const int *idx, int n, double *dst,
                hipStream_t s);
// F block: out[i,j] = (fock[p[i],q[j]] - (p[i]==q[j])*en[p[i]]) * wi[i]*wj[j]
void fock_block(const double *fock, const double *en, int n, const int *p,
                const int *q, const double *wi, const double *wj, int di,
                int dj, double *out, hipStream_t s);
void denominators(const double *ev, const double *eo, int nv, int no,
                  double *D1, double *D2, hipStream_t s);

// Per-slice preparation of the amplitude-derived operands (see ftcc_api.hip)
struct PrepOut {
  double *tau, *t2r, *aring;
};
void prep_amplitudes(const double *t1, const double *t2, int nv, int no,
                     int nb, const PrepOut &o, hipStream_t s);
// W[x0,x1,x2,x3] += sgn*(Q[x0,x1,x2,x3] - Q[x0,x1,x3,x2])   (swap_last)
// W[x0,x1,x2,x3] += sgn*(Q[x0,x1,x2,x3] - Q[x1,x0,x2,x3])   (!swap_last)
void antisym_add(double *W, const double *Q, int d0, int d1, int d2, int d3,
                 bool swap_last, double sgn, int nb, hipStream_t s);
// Wr[e,m,b,j] = ovvo_ring[e,m,b,j] - Q1[b,m,e,j] + Q2[m,b,j,e]
void ring_assemble(double *Wr, const double *ovvo_ring, const double *Q1,
                   const double *Q2, int nv, int no, int nb, hipStream_t s);
// T2bar[a,b,i,j] = -( acc1 + acc2 [+ add] + S1 - S1^(ab) + S2 - S2^(ij)
//                    + PP(Z ring) ) with acc1/acc2 one per grid point and
// add one slice for every grid point (or nullptr); with spin labels
// (device arrays, or nullptr), spin-forbidden elements are written as zero
// without reading the operands.  packed (needs the labels): every input
// holds one pair of each antisymmetric couple (x < y for equal spins, the
// alpha index first for mixed spins): acc1/acc2 in (a,b) and (i,j), S1 in
// (i,j), S2 in (a,b); add must be antisymmetric in (a,b).  zh > 0 (packed
// only; alpha runs [0, zh) and [0, zo) with beta runs of the same lengths
// after them): Z holds the rows of an alpha first index only, the others
// are read from their spin-flipped images.  zero_forbidden = false (packed
// only): the spin-forbidden doubles of t2bar are left unwritten.  packed
// needs jl: the occupied indices of label 0 (jc0 of them), then those of
// label 1 (jc1).
void r2_final(double *t2bar, const double *acc1, const double *acc2,
              const double *S1, const double *S2, const double *Z, int nv,
              int no, int nb, hipStream_t s, const int *lab_v,
              const int *lab_o, const double *add, bool packed, int zh = 0,
              int zo = 0, bool zero_forbidden = true,
              const int *jl = nullptr, int jc0 = 0, int jc1 = 0);
// Spin-flip images of a ring tensor X[nb][nv][no][nv][no] (slice stride
// sb) whose rows with a virtual alpha index are computed: every allowed
// element with that index beta is copied from its flipped partner (alpha
// runs [0, hv) and [0, ho), beta runs of the same lengths after them).
// vv_pair: X[a,i,b,j] conserves sa + sb = si + sj (else se + sj = sm + sb).
void flip_copy_ring(double *X, int nv, int no, int hv, int ho, int nb,
                    int64_t sb, bool vv_pair, hipStream_t s);
// Spin-flip image of a spin-diagonal matrix X[nb][P][Q] (slice stride sb;
// alpha runs [0, hp) and [0, hq), beta runs after them): the beta-beta
// block is copied from the alpha-alpha one.
void flip_copy_mat(double *X, int P, int Q, int hp, int hq, int nb,
                   int64_t sb, hipStream_t s);
// x[n][c] (nb batches of n*c) -> out[c][n] transposed copy for small mats
void transpose2(const double *in, double *out, int r, int c, int nb,
                hipStream_t s);

// ---------------------------------------------------------------------------
// Integration, damping, norms and energy (update.hip)
// ---------------------------------------------------------------------------
struct Grid {
  int ng = 0;
  std::vector<double> ti, G;
  bool uniform = false;         // tau_x - tau_y == (x - y) * delta to ~1 ulp
  double delta = 0.0;
  // step s of the integration recurrence (0: none): uniform, G zero above
  // the diagonal and G[y][x] == G[y-s][x] for x < y-2; coefficients d_rc
  // (a[ng], b[ng], c[ng])
  int rec = 0;
  double *d_ti = nullptr, *d_g = nullptr, *d_G = nullptr, *d_rc = nullptr;
};
// Largest grid integrated by the recurrence (its coefficients sit in LDS).
constexpr int MAX_REC_NG = 1024;

// mode 0: integrate in place (Tbar -> T).  mode 1: integrate Tbar, then damp
// into Told with norms and energy partials.
struct UpdateArgs {
  int ng, nv, no;
  const double *ti, *G, *g;     // device
  const double *eo, *ev;        // device energies for D (if D == nullptr)
  const double *D;              // explicit D1 or D2 (device) or nullptr
  double *tbar;                 // residual in; integrated out in mode 0
  const double *src;            // if set, Tbar is read from here (grids too
                                // large for LDS staging integrate in place
                                // only from a copy)
  double *told;                 // damped iterate (mode 1)
  // a fixed correction added to every element of Tbar before integration
  // (same layout as tbar; nullptr: none)
  const double *corr = nullptr;
  double damp;
  int uniform;                  // grid equally spaced: tau_x - tau_y =
  double delta;                 //   (x - y) * delta
  int rec;                      // step (1, 2) of the integration recurrence
  const double *rc;             //   (0: none), coefficients a, b, c [ng]
  const double *t1new;          // mode 1 T2 pass: updated T1 (ng,nv,no)
  const double *eint;           // energy integral in amplitude layout
                                // (F.ov^T for T1, oovv^T for T2)
  double *partials;             // per-block {d, n, e} triples
  // spin labels of the virtual / occupied lists (device) when the doubles
  // are known to conserve spin: spin-forbidden elements (zero in Tbar and
  // T) are then skipped; nullptr: every element
  const int *lv = nullptr, *lo = nullptr;
  // the same when each label list is one run of 0s then one of 1s: {start,
  // count} of the alpha and beta runs of the virtual and occupied lists and
  // the number of spin-allowed doubles (> 0: only those are enumerated)
  int vrun[4] = {0, 0, 0, 0}, orun[4] = {0, 0, 0, 0};
  int64_t nallowed = 0;
  // spin-flip symmetric Tbar and Told with label runs (mode 1 of the
  // recurrence update): only the doubles of an alpha first virtual are
  // enumerated (half of nallowed); each result is also written at its
  // image (alpha and beta runs of equal length hv, ho swapped) and the
  // partial sums count it twice
  int flip_half = 0;
};
// One grid point y of the integral from the ng slices of a.tbar into the
// single slice a.told: mode 0 stores it, mode 1 damps into it and leaves
// |T - told|^2 in the first partial of each block (returns #blocks)
int integrate_point(const UpdateArgs &a, bool doubles, int y, int mode,
                    hipStream_t s);
// One explicit Runge-Kutta stage of dT/dtau = Tbar - D T over a.ng slices
// (a.D or a.eo/a.ev give D): k = tb - D*(first ? T : var);
// acc = (first ? 0 : acc) + b*k; last: T += h*acc, else var = T + hnext*k
// (var == nullptr: only acc = k is formed).  zero: k = 0.
int rk_stage(const UpdateArgs &a, bool doubles, const double *tb, double *T,
             double *var, double *acc, double b, double hnext, double h,
             bool first, bool last, bool zero, hipStream_t s);
// One pass of an implicit step (be = 0, cn = 1, am2 = 2) of dT/dtau =
// Tbar - D T over a.ng slices: mode 0 predictor (k1 = dold = h k, var =
// T + h k), mode 1 iteration (d from Tbar(var); sums of (d - dold)^2 and
// d^2 per block in a.partials; dold, var updated), mode 2 finish (T +=
// dold, or the cn average).  Returns #blocks.
int implicit_pass(const UpdateArgs &a, bool doubles, const double *tb,
                  double *T, double *var, double *dold, double *k1, double h,
                  double alpha, int mode, int scheme, bool zero,
                  hipStream_t s);
int update_t1(const UpdateArgs &a, int mode, hipStream_t s);   // returns #blocks
int update_t2(const UpdateArgs &a, int mode, hipStream_t s);
// true when in-place integration needs a source copy in UpdateArgs::src
// (the general kernel on a grid too large to stage Tbar in LDS)
bool update_needs_copy(const UpdateArgs &a);
// out[k] = sum over blocks of partials[3*b+k], k=0..2 (fixed order)
void reduce_partials(const double *partials, int nblocks, double *out3,
                     hipStream_t s);
// energy of given amplitudes: per-block partials of
// sum_y g_y [T1.F.ov^T] and sum_y g_y [(T2/4 + q T1T1/2).oovv^T]
int energy_t2(const double *T1, const double *T2, const double *g,
              const double *oovvT, int ng, int nv, int no, int qterm,
              double *partials, hipStream_t s);
int energy_t1(const double *T1, const double *g, const double *fovT, int ng,
              int nv, int no, double *partials, hipStream_t s);
int sumsq(const double *x, int64_t n, double *partials, hipStream_t s);
// blocks (= partial triples) of the reductions above over n elements
int reduction_blocks(int64_t n);
int vsum(const double *x, int64_t n, double *partials, hipStream_t s);
// count (third slot of the partial triples) of the nonzero entries of nb
// consecutive (d0,d1,d2,d3) blocks with l0[i0] + l1[i1] != l2[i2] + l3[i3];
// null labels read as 0; returns the number of partial triples
int spin_violations(const double *a, int d0, int d1, int d2, int d3, int nb,
                    const int *l0, const int *l1, const int *l2, const int *l3,
                    double *partials, hipStream_t s);
// count (third slot) of the entries of nb consecutive (d0,d1,d2,d3) blocks
// that differ from their image under the spin flip: index x of dimension k
// maps to x + h_k below h_k and x - h_k from h_k on (h_k = 0: not flipped).
// tol = 0 counts any difference; tol > 0 only |x - y| > tol * max(|x|, |y|).
int flip_violations(const double *a, int d0, int d1, int d2, int d3, int h0,
                    int h1, int h2, int h3, double *partials, hipStream_t s,
                    int nb = 1, double tol = 0.0);

}  // namespace ftcc
