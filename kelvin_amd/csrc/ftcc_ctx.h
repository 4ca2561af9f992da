// The context behind the C ABI, shared by the translation units that
// implement its entry points (ftcc_api.hip, adjoint.hip), and the helpers
// of ftcc_api.hip they call.
#pragma once
#include "ftcc_internal.h"

#include <array>
#include <map>
#include <string>
#include <vector>

using namespace ftcc;

namespace ftcc_core {
struct Adjoint;
}

struct ftcc_ctx {
  int no = 0, nv = 0, ng = 0, prec = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void *> allocs;
  std::map<void *, int64_t> alloc_doubles;   // sizes of dalloc's buffers
  int64_t bytes = 0;

  // integrals in the reference layouts
  double *F[4] = {nullptr, nullptr, nullptr, nullptr};   // oo ov vo vv
  double *I[9] = {};  // vvvv vvvo vovv vvoo vovo oovv vooo ooov oooo
  // derived layouts
  double *oovv_ring = nullptr;  // [e,m,f,n] = oovv[m,n,e,f]
  double *ovvo_ring = nullptr;  // [e,m,b,j] = <mb||ej> = -vovo[b,m,e,j]
  double *ovvo_ringT = nullptr; // [b,j,e,m] = ovvo_ring[e,m,b,j]
  double *ooov_p1 = nullptr;    // [m,i,n,e] = ooov[m,n,i,e]
  double *ooov_iemn = nullptr;  // [i,e,m,n] = ooov[n,m,i,e]
  double *oovvT = nullptr;      // [a,b,i,j] = oovv[i,j,a,b]
  double *oovv_mefn = nullptr;  // [m,e,f,n] = oovv[m,n,e,f]
  double *vovv_aefm = nullptr;  // [a,e,f,m] = vovv[a,m,e,f]
  double *vovv_embf = nullptr;  // [e,m,b,f] = vovv[b,m,e,f]
  double *ooov_emnj = nullptr;  // [e,m,n,j] = ooov[m,n,j,e]
  double *fovT = nullptr;       // [a,i] = F.ov[i,a]
  double *eo = nullptr, *ev = nullptr;
  double *D1 = nullptr, *D2 = nullptr;   // explicit denominators (optional)
  bool have_D = false;
  bool have_ints = false;
  bool have_energies = false;
  Grid grid;

  // per-slot workspace (nb slots)
  int nb = 1;
  int nb_fixed = 0;          // grid points per chunk set by the caller (0: auto)
  bool trusted = false;      // caller promises spin-conserving amplitudes
  // the promise was given and the next residual call checks it once on the
  // amplitudes it receives
  bool trust_pending = false;
  // fixed corrections added to Tbar before the integration of ftcc_step
  // (caller-owned, amplitude layout; nullptr: none)
  const double *corr1 = nullptr, *corr2 = nullptr;
  // one FT-CCSD step captured as a HIP graph, and the arguments it was
  // captured for (the first call with new arguments runs eagerly, so that
  // every lazily allocated buffer exists before capture)
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  const void *gkey[4] = {nullptr, nullptr, nullptr, nullptr};
  double gbeta = 0.0, gdamp = 0.0;
  uint64_t gskip = 0;
  int gprec = -1;
  bool gwarm = false;
  // imaginary-time propagation: stage input, accumulator and right-hand
  // side (singles / doubles), allocated on first use
  double *td_var[2] = {}, *td_acc[2] = {}, *td_k[2] = {}, *td_k1[2] = {};
  double *tau = nullptr, *t2r, *aring, *wmnij, *wabef,
      *wr, *z, *u, *s1, *s2;
  double *t1T, *Fme, *FmeT, *Fae, *Fmi, *Fbe, *Fmj, *ytmp;
  Scratch scr;
  double *partials = nullptr;   // 3*MAX blocks
  double *red = nullptr;        // device scalars
  double *host_red = nullptr;   // pinned

  // event timing of the O(n^6) products
  // kernel classes timed with events while profiling is on
  enum { PROF_PRODUCTS = 0, PROF_N5 = 1, PROF_GATHER = 2, PROF_ASSEMBLY = 3,
         PROF_UPDATE = 4, NPROF = 5 };
  struct Pending {
    hipEvent_t e0, e1;
    int cls;
    double flops, bytes;
    int line, M, N, K, nb;   // call site and shape (per-site table)
  };
  // per call site (source line) totals when FTCC_PROFILE_SITES is set,
  // printed to stderr when profiling is switched off
  struct SiteStat {
    int64_t launches = 0;
    double ms = 0.0, flops = 0.0, bytes = 0.0;
    int M = 0, N = 0, K = 0, nb = 0;
  };
  bool prof_sites = false;
  // keyed by (class, line, M, N, K, slices): one entry per shape of a site
  std::map<std::array<int64_t, 6>, SiteStat> site_stats;
  // shape and line of the GEMM being launched (0: none)
  int cur_line = 0, cur_M = 0, cur_N = 0, cur_K = 0, cur_nb = 0;
  int site_override = 0;   // the line of an enclosing run_spin call
  // split-bf16 operands of the five O(n^6) products (bf16x3 mode)
  SplitMat sp_tauA, sp_tauB, sp_t2rA, sp_aringB, sp_wmnijB, sp_wabefA, sp_wrB;
  SplitMat sp_oovvA, sp_ringA;
  bool split_fixed_valid = false;
  int split_layout = -1;        // spin state the split buffers are sized for
  int split_nb = 0;             // and the chunk
  // the product operands of either precision (split-bf16 pairs or FP64
  // blocked copies, one set at a time) are carved from one allocation, so a
  // precision switch allocates nothing: owner 0 none, 1 split, 2 FP64
  double *arena = nullptr;
  int64_t arena_doubles = 0;
  int arena_owner = 0;
  // slots of the per-chunk workspace as allocated, and the precision the
  // current chunk was chosen for (-1: none)
  int nb_alloc = 0;
  int ws_prec = -1;

  // Spin blocking of the O(n^6) products: when every integral block
  // conserves spin under the orbital labels, the pair indices of the ten
  // product operands are reordered by spin class and each product runs as
  // three dense block products (10/64 of the dense work).
  int spin = 0;
  int *d_lab_o = nullptr, *d_lab_v = nullptr;
  // occupied indices grouped by spin label: d_jl[0, jl_cnt[0]) the label-0
  // ones, then the label-1 ones (in index order)
  int *d_jl = nullptr;
  int jl_cnt[2] = {0, 0};
  // O(n^5) contractions restricted to spin-allowed blocks: needs each label
  // list to be one run of 0s then one run of 1s (spin5 = 1); starts/counts
  // of the two runs in the occupied and virtual index lists
  int spin5 = 0;
  int o_start[2] = {0, 0}, o_count[2] = {0, 0};
  int v_start[2] = {0, 0}, v_count[2] = {0, 0};
  // Spin-flip symmetry (flip = 1): equal alpha and beta runs and every
  // integral, Fock block and orbital energy unchanged when each index moves
  // to its partner of the other spin.  Trusted amplitudes (the solver's own
  // iterates) then share the symmetry, and the ring products and the
  // alpha-lead O(n^5) terms compute half their blocks, the rest being their
  // flipped images (flip_now: in force for the residual being evaluated).
  int flip = 0;
  bool flip_now = false;
  // under flip_now: an output index letter of the spin-restricted GEMMs
  // whose beta assignments are skipped (their blocks are flipped images,
  // copied or read in place once the output is complete); 0 = none
  char flip_lead = 0;
  // output index letters (x, y) of the spin-restricted GEMMs whose
  // assignments with x beta and y alpha are skipped: the packed ladder
  // operands never read that class of a pair; 0 = none
  char skip_ba[2] = {0, 0};
  // output index letters (x, y) whose assignments with both beta are
  // skipped (letters of a pair index allowed); 0 = none
  char skip_bb[2] = {0, 0};
  // the ladder products skip their beta-beta | beta-beta block
  bool ladder_skip_bb = false;
  // ftcc_step under the flip symmetry (for the chunk being evaluated): the
  // ladder operands' beta-beta blocks are read by no product
  bool flip_step_now = false;
  // the last residual_all ran the flip-symmetric step form (its Tbar, like
  // the amplitudes, is exactly symmetric under the spin flip)
  bool last_flip_step = false;
  // set by ftcc_step around its residual: its update reads the spin-allowed
  // doubles of Tbar only, so the assembly leaves the forbidden ones unwritten
  bool tbar_scratch = false;
  PairSpace ps_vv, ps_oo, ps_vo;   // (a,b), (i,j) and ring (a,i) pairs
  // Ladder products contract over an antisymmetric pair, so they sum over
  // the packed pairs only (e < f within a spin, alpha-beta once) and double
  // the result: packed (e,f) and (m,n) spaces, used for their K index
  int pack = 0;
  PairSpace ps_vvp, ps_oop;
  double *b64[10] = {};            // FP64 blocked operands (fp64/fp32 modes)
  int b64_layout = -1;
  int b64_nb = 0;
  bool b64_fixed_valid = false;

  // contiguous copies of large strided GEMM operands (adjoint contractions)
  double *pack_buf = nullptr;
  int64_t pack_elems = 0;
  // precision of the adjoint's large contractions (ftcc_adjoint_precision)
  // and their split-bf16 operand copies and C offset tables
  int adj_prec = 0;   // FTCC_PREC_FP64
  double adj_lowp_min_flops = 1e10;
  double *lp_buf = nullptr;
  int64_t lp_doubles = 0;
  // launch plan of the generic GEMMs (ftcc_set_gemm_plan)
  int tile_policy = 0, split_target = 1024;

  bool prof_on = false;
  std::vector<Pending> prof_pending;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[NPROF] = {}, prof_flops[NPROF] = {}, prof_bytes[NPROF] = {};
  int64_t prof_launches[NPROF] = {};

  // adjoint (lambda) workspace and the T-dependent intermediates it keeps
  // between ftcc_adjoint_begin and ftcc_adjoint_end (adjoint.hip)
  ftcc_core::Adjoint *adj = nullptr;
};

namespace ftcc_core {
using namespace ftcc;

// device allocation owned by the context (freed by ftcc_destroy)
double *dalloc(ftcc_ctx *c, int64_t n);
void dfree(ftcc_ctx *c, void *p);
// run fn, turning exceptions into the status code and message of c
template <typename Fn>
int guarded(ftcc_ctx *c, Fn fn) {
  if (!c) return FTCC_EINVAL;
  try {
    fn();
    c->err.clear();
    return FTCC_OK;
  } catch (const Error &e) {
    c->err = e.msg;
    return e.code;
  } catch (const std::exception &e) {
    c->err = e.what();
    return FTCC_EINVAL;
  }
}
// one GEMM of the engine (split-K scratch grown as needed)
void run(ftcc_ctx *c, const Gemm &g, int prec, bool big = false,
         int line = __builtin_LINE());
// One index of a spin-restricted contraction: its letter, its orbital role
// ('o' or 'v'; 'y' the slices of the base, '1' extent one) and its stride
// in A, B and C (C's strides also address D).
struct SIdx {
  char l, role;
  int64_t sa, sb, sc;
  char l2 = 0;   // pair roles: the second letter
};
// C[M,N] = alpha A[M,K] B[K,N] + beta C (+ gamma D) over the spin-allowed
// blocks of the index lists (see ftcc_api.hip)
void run_spin(ftcc_ctx *c, const Gemm &base, int prec,
              const std::vector<SIdx> &Mi, const std::vector<SIdx> &Ni,
              const std::vector<SIdx> &Ki, const char *ruleA,
              const char *ruleB, const char *ruleC,
              int line = __builtin_LINE());
void check_ready(ftcc_ctx *c);
void check_grid_D(ftcc_ctx *c);
void drop_graph(ftcc_ctx *c);
void free_workspace(ftcc_ctx *c);
void read_scalars(ftcc_ctx *c, int n);
// the residual at every grid point (returns whether the spin layout ran)
bool residual_all(ftcc_ctx *c, const double *T1, const double *T2,
                  double *T1b, double *T2b, uint64_t skip);
void ensure_integrals(ftcc_ctx *c);
void build_derived(ftcc_ctx *c);
bool try_spin(ftcc_ctx *c, const std::vector<int> &lo,
              const std::vector<int> &lv);
UpdateArgs upd(ftcc_ctx *c);
// HIP-event bracket of one launch of kernel class cls while profiling is on
// (token -1 when off)
int prof_begin(ftcc_ctx *c, int cls, double flops, double bytes);
void prof_end(ftcc_ctx *c, int token);
// release the adjoint workspace (adjoint.hip)
void adjoint_release(ftcc_ctx *c);
// device address of the allocation-free entry point's scratch array on
// `device` (noalloc.hip; looked up once per device)
double *noalloc_scratch(int device);

}  // namespace ftcc_core
