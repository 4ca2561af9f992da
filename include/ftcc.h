/*
 * ftcc.h -- C ABI of libftcc.so, the MI355X (gfx950) finite-temperature CCSD
 * amplitude engine.
 *
 * Every pointer argument named *_dev is a device pointer (HBM, e.g. owned by a
 * torch tensor); *_host pointers are host memory.  Amplitudes use the
 * reference's layouts, C order, float64:
 *     T1[y][a][i]          shape (ng, nv, no)
 *     T2[y][a][b][i][j]    shape (ng, nv, nv, no, no)
 * All calls are ordered on the stream given to ftcc_create; calls that return
 * host scalars synchronise that stream.  A context is not thread safe: use
 * one context per stream.  Every function returns 0 on success and a nonzero
 * FTCC_E* code on failure, with a message available from ftcc_last_error.
 *
 * Citations name the reference function each entry point replaces (paths are
 * relative to the kelvin checkout).
 */
#ifndef FTCC_H
#define FTCC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ftcc_ctx ftcc_ctx;

/* contraction precision of the O(n^6)/O(n^5) products; everything else
 * (Fermi factors, propagators, quadrature, damping, norms, energies) is FP64 */
enum {
  FTCC_PREC_FP64 = 0,   /* FP64 operands, FP64 MFMA                          */
  FTCC_PREC_FP32 = 1,   /* operands rounded to FP32, FP32 MFMA and accum.     */
  FTCC_PREC_BF16X3 = 2  /* x = hi+lo in bf16; hi*hi+hi*lo+lo*hi, FP32 accum.  */
};

enum {
  FTCC_OK = 0,
  FTCC_EINVAL = 1,      /* bad argument (size, pointer, grid)                 */
  FTCC_ESTATE = 2,      /* call order: integrals / grid not set               */
  FTCC_EHIP = 3,        /* HIP runtime error (launch, copy)                   */
  FTCC_ENOMEM = 4       /* device allocation failed                           */
};

/* Create a context for no occupied-role and nv virtual-role orbitals (at
 * finite temperature with no active-space cut, no = nv = n) and ng grid
 * points.  stream is a hipStream_t (NULL = default stream).  Replaces the
 * set-up half of kelvin/ccsd.py:_ft_ccsd (ccsd.py:595-675). */
int ftcc_create(ftcc_ctx **ctx, int no, int nv, int ng, int precision,
                void *stream);
int ftcc_destroy(ftcc_ctx *ctx);
const char *ftcc_last_error(const ftcc_ctx *ctx);
/* Contraction precision of the O(n^6) products (FTCC_PREC_*).  A change
 * keeps the residual workspace: the next residual carves this precision's
 * product operands from the same arena and runs the largest chunk that the
 * allocated slots and arena hold (ftcc_set_chunk releases the workspace, so
 * that the next residual plans it for the current precision). */
int ftcc_set_precision(ftcc_ctx *ctx, int precision);

/* Fermi-weighted integrals (kelvin/cc_utils.py:569-602 ft_integrals and
 * :825-858 ft_active_integrals).  en_dev[n], fock_dev[n*n] (= g_fock_tot,
 * diagonal included), eri_dev[n^4] (= g_aint_tot, antisymmetrised).
 * iocc_host[no] / ivir_host[nv] list the orbitals kept in each role (NULL =
 * all n).  Fermi factors and sqrt weights are formed on the device in FP64;
 * the denominators e_a-e_i, e_a+e_b-e_i-e_j (kelvin/ccsd.py:633-634) are
 * formed on the fly from the role-ordered energies.  If omega0 is not NULL
 * it receives sum_p GP0(e_p) over all n orbitals (kelvin/ft_mp.py:12-14 over
 * cqcpy ft_utils.GP0, kelvin/ccsd.py:637-638). */
int ftcc_set_system(ftcc_ctx *ctx, int n, const double *en_dev,
                    const double *fock_dev, const double *eri_dev,
                    double beta, double mu, const int *iocc_host,
                    const int *ivir_host, double *omega0);

/* Arbitrary blocks in the reference's layouts (cqcpy one_e_blocks /
 * two_e_blocks as passed to kelvin/ft_cc_equations.py:96 ccsd_stanton):
 *   F: oo[no,no] ov[no,nv] vo[nv,no] vv[nv,nv]
 *   I: vvvv vvvo vovv vvoo vovo oovv vooo ooov oooo
 * (pointer order exactly as listed; device pointers, copied).  The orbital
 * energies of an earlier ftcc_set_system are dropped with its integrals:
 * give the denominators with ftcc_set_denominators before integrating. */
int ftcc_set_blocks(ftcc_ctx *ctx, const double *const F_dev[4],
                    const double *const I_dev[9]);

/* Spin blocking of the O(n^6) products (no counterpart in the reference,
 * whose einsums are dense; the results are the same).  Labels (0 or 1,
 * host arrays) of the occupied-role and virtual-role orbitals: if every
 * Fock and integral block is exactly zero wherever spin is not conserved,
 * each product runs as its three nonzero spin blocks (10/64 of the dense
 * work); otherwise FTCC_EINVAL and the products stay dense.  Both
 * NULL switches blocking off.  ftcc_set_system tries the labels "first
 * half alpha, second half beta" of the n orbitals by itself (environment
 * FTCC_SPIN=0 prevents it); ftcc_set_blocks switches blocking off.
 * The blocks are exact for amplitudes that conserve spin too (every
 * iterate started from the driving terms does): each residual call counts
 * the amplitude entries that do not and runs dense if there are any. */
int ftcc_set_spin(ftcc_ctx *ctx, const int *lab_o_host,
                  const int *lab_v_host);
/* *on = 0 when everything runs dense; otherwise bit flags: 1 the O(n^6)
 * products run spin-blocked; 2 the O(n^5) contractions run over the
 * spin-allowed blocks only (labels form one run of 0s then one of 1s in the
 * occupied and in the virtual list); 4 the ladder products run over packed
 * antisymmetric pairs in rows, columns and K (e < f within a spin,
 * alpha-beta once: the amplitudes must be antisymmetric, as every CC
 * iterate is); 8 the system is symmetric under the spin flip (equal alpha
 * and beta runs; every integral, Fock block and orbital energy unchanged
 * when each index moves to its partner of the other spin): for trusted
 * amplitudes (ftcc_set_trusted) the two ring products and the
 * spin-restricted terms of the Fock-type, ring and singles intermediates
 * compute the blocks of an alpha lead index and take the others from their
 * flipped images, and ftcc_step also skips the beta-beta blocks its update
 * never reads.  FTCC_SPIN5=0, FTCC_PACK=0 and FTCC_FLIP=0 in the environment at
 * ftcc_set_system/ftcc_set_spin time switch off flags 2, 4 and 8. */
int ftcc_get_spin(ftcc_ctx *ctx, int *on);

/* Uniform-electron-gas integrals built on the device (the provider of
 * kelvin/ueg_utils.py:83-138 and kelvin/ueg_system.py:400-419): nk plane
 * waves with integer vectors ivec_host[nk][3] and wave vectors
 * kvec_host[nk][3] = 2 pi/L * ivec in a box of side L.  Writes the
 * antisymmetrised spin-orbital <pq||rs> (n = 2 nk; alpha orbitals, then
 * beta) into eri_dev[n^4], bitwise equal to the host construction. */
int ftcc_ueg_eri(ftcc_ctx *ctx, int nk, const int *ivec_host,
                 const double *kvec_host, double L, double *eri_dev);

/* Mean-field quantities of a system whose integrals are on the device:
 * fock_dev[n,n] = diag(en) + sum_r <pr||qr> fo[r] (kelvin/ueg_system.py:
 * 400-419) and *e1 = 1/2 sum_ij <ij||ij> fo_i fo_j (kelvin/ueg_system.py:
 * 126-131).  Either output may be NULL. */
int ftcc_mean_field(ftcc_ctx *ctx, int n, const double *en_dev,
                    const double *eri_dev, const double *fo_dev,
                    double *fock_dev, double *e1);

/* Explicit denominators D1_dev[nv,no], D2_dev[nv,nv,no,no] (the arguments
 * D1, D2 of kelvin/ft_cc_equations.py:96), copied. */
int ftcc_set_denominators(ftcc_ctx *ctx, const double *D1_dev,
                          const double *D2_dev);

/* Copy one stored block into dst_dev: which = 0..3 for F oo ov vo vv,
 * 4..12 for I vvvv vvvo vovv vvoo vovo oovv vooo ooov oooo. */
int ftcc_get_block(ftcc_ctx *ctx, int which, double *dst_dev);

/* Quadrature abscissae and Volterra weights (kelvin/quadrature.py ft_quad
 * output): ti_host[ng], G_host[ng*ng] row-major; ng may differ from the
 * value given to ftcc_create (amplitude arguments then have ng slices).
 * Any G is accepted (compute_ESN passes Gd and G[i,j]*(t_j-t_i)/beta,
 * kelvin/ccsd.py:1118-1137). */
int ftcc_set_grid(ftcc_ctx *ctx, int ng, const double *ti_host,
                  const double *G_host);

/* Tbar = -F.vo, -I.vvoo at every grid point (the source of the initial
 * guess, kelvin/ccsd.py:680-686). */
int ftcc_driving_terms(ftcc_ctx *ctx, double *T1bar_dev, double *T2bar_dev);

/* Tbar = -F.vo - R1(T), -I.vvoo - R2(T) at every grid point: the loop of
 * kelvin/ft_cc_equations.py:101-107 (cqcpy cc_equations._Stanton, fac=-1).
 * Bit y of skip_mask (y < 64) promises T(tau_y) == 0 exactly; such slices
 * get the driving term alone. */
int ftcc_residual(ftcc_ctx *ctx, const double *T1_dev, const double *T2_dev,
                  double *T1bar_dev, double *T2bar_dev, uint64_t skip_mask);

/* T(tau_y) = sum_x G[y,x] exp(-(tau_y - tau_x) D) Tbar(tau_x), in place
 * (kelvin/quadrature.py:292-317 int_tbar1 / int_tbar2).  Either pointer may
 * be NULL to integrate only the other. */
int ftcc_integrate(ftcc_ctx *ctx, double *T1_dev, double *T2_dev);

/* (1/beta) sum_y g_y [ T1.F.ov + (T2/4 + q*T1 T1/2).I.oovv ]
 * (kelvin/ft_cc_energy.py:7-32; g_host[ng]; qterm = Qterm). */
int ftcc_energy(ftcc_ctx *ctx, const double *T1_dev, const double *T2_dev,
                const double *g_host, double beta, int qterm,
                double *energy);

/* ftcc_energy without allocating, copying from the host or synchronising:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The weights
 * g_dev[ng] and the result energy_dev[1] are device memory; the partial
 * sums live in a fixed device array per GPU (196,656 bytes: 8192 partial
 * triples, the block cap of the grid-stride reductions, so every problem
 * size fits; a reduction needing more blocks would get FTCC_EINVAL before
 * any launch).  Calls that overlap on one GPU share that array and race. */
int ftcc_energy_noalloc(ftcc_ctx *ctx, const double *T1_dev,
                        const double *T2_dev, const double *g_dev,
                        double beta, int qterm, double *energy_dev);

/* sum of squares of n doubles (numpy.linalg.norm(x)**2 of
 * kelvin/cc_utils.py:139-140,145-151). */
int ftcc_sumsq(ftcc_ctx *ctx, const double *x_dev, int64_t n, double *out);

/* One body of the damped fixed point (kelvin/cc_utils.py:141-155):
 *   (T1,T2) = integrate(residual(T1old, T2old))
 *   d1 = |T1 - T1old|^2, d2 = |T2 - T2old|^2
 *   Told <- damp*Told + (1-damp)*T      (in place in T1old_dev / T2old_dev)
 *   n1 = |T1old|^2, n2 = |T2old|^2, energy = ft_cc_energy(T1old, T2old, g)
 * T1bar_dev / T2bar_dev are caller-provided scratch of amplitude size:
 * with spin blocking on, the spin-forbidden doubles of T2bar are left as
 * they were (the update skips them), and the whole T2bar of the last call
 * is unspecified.  out5 receives {d1, d2, n1, n2, energy}. */
int ftcc_step(ftcc_ctx *ctx, double *T1old_dev, double *T2old_dev,
              double *T1bar_dev, double *T2bar_dev, const double *g_host,
              double beta, double damp, uint64_t skip_mask, double out5[5]);

/* Pointwise schedule, rt_iter='point' (kelvin/cc_utils.py:176-242,
 * quadrature.py:348-369).  Amplitudes T1/T2 and residuals T1bar/T2bar
 * have ng slices.  ftcc_point_begin prepares grid point ig: ig = 0 stores
 * the driving terms -F.vo, -I.vvoo as Tbar[0]; ig = 1 stores them as
 * Tbar[1] and sets T[1] to their single-point integral; ig >= 2 sets T[ig]
 * by linear extrapolation from T[ig-1] and T[ig-2]. */
int ftcc_point_begin(ftcc_ctx *ctx, int ig, double *T1_dev, double *T2_dev,
                     double *T1bar_dev, double *T2bar_dev);
/* One iteration at point ig >= 1: Tbar[ig] = residual(T[ig]);
 * T = sum_x G[ig,x] exp((t_x - t_ig) D) Tbar[x]; out2 = {|T1 - T1[ig]|^2,
 * |T2 - T2[ig]|^2}; T[ig] <- damp*T[ig] + (1-damp)*T. */
int ftcc_point_step(ftcc_ctx *ctx, int ig, double *T1_dev, double *T2_dev,
                    double *T1bar_dev, double *T2bar_dev, double damp,
                    double out2[2]);

/* Imaginary-time propagation (TD-CCSD, kelvin/td_ccsd.py:583-681 with the
 * propagators of kelvin/propagation.py:5-52).  Each of the ng slices of
 * the context is an independent state.  ftcc_td_rhs writes the right-hand
 * side of td_ccsd.py:646-652, K = Tbar(T) - D T (K1 = 0 when singles == 0).
 * ftcc_td_step advances T in place by one explicit Runge-Kutta step of
 * size h: order 1 (rk1, Euler), 2 (rk2, Heun) or 4 (rk4, classical); with
 * singles == 0, T1 is held fixed as the reference's singles=False does. */
int ftcc_td_rhs(ftcc_ctx *ctx, const double *T1_dev, const double *T2_dev,
                double *K1_dev, double *K2_dev, int singles);
int ftcc_td_step(ftcc_ctx *ctx, int order, double h, int singles,
                 double *T1_dev, double *T2_dev);
/* One implicit step of size h (kelvin/propagation.py:55-128): scheme 0 be
 * (backward Euler), 1 cn (Crank-Nicolson: the backward-Euler iterate
 * averaged with the explicit predictor), 2 am2 (trapezoid).  From the
 * predictor d = h f(T), d <- (1 - damp) d_new + damp d with d_new = h f(T + d)
 * (am2: (h f(T + d) + h f(T)) / 2) until sum over T1, T2 of
 * ||d_new - d|| / (||d_new|| + 0.01) < thresh (norms over all slices), at
 * most max_iter times; *iters (may be NULL) receives the count.  Not
 * converging returns FTCC_EINVAL ("BE: / AM2: Failed to compute implicit
 * step", the reference's message) and leaves T unchanged. */
int ftcc_td_implicit_step(ftcc_ctx *ctx, int scheme, double h, int max_iter,
                          double damp, double thresh, int singles,
                          double *T1_dev, double *T2_dev, int *iters);

/* Grid points per residual chunk (the slots whose O(n^6) products go out as
 * one launch): nb = 0 lets the context choose from its workspace budget
 * (FTCC_WORKSPACE_GB, default 120 GiB, read when the workspace is
 * allocated; it counts the product operands of the current precision) --
 * the default; nb > 0 fixes it, at most ng and what that budget allows.
 * Changing it releases the residual workspace.  ftcc_get_chunk reports the
 * value in use (allocating the workspace).  No reference counterpart: the
 * reference evaluates the grid points one at a time. */
int ftcc_set_chunk(ftcc_ctx *ctx, int nb);
int ftcc_get_chunk(ftcc_ctx *ctx, int *nb);

/* on != 0: the caller promises that the amplitudes it passes conserve spin
 * and, for a spin-flip symmetric system (ftcc_get_spin flag 8), are
 * unchanged under the spin flip (both true of every iterate of the solver
 * started from the driving terms), so residual calls skip the counting
 * pass over T, need no host synchronisation and use the flip symmetry.
 * The first residual-evaluating call after the promise (ftcc_residual,
 * ftcc_step, the eager first call of ftcc_step_graph, ...) checks it once on
 * the amplitudes it receives (one host synchronisation; entries must match
 * their flipped images to 1e-10 relative): if they break it, the promise is
 * withdrawn and that call returns FTCC_EINVAL without computing.  Later
 * calls are not checked.  No counterpart in the reference (whose residual
 * is dense). */
int ftcc_set_trusted(ftcc_ctx *ctx, int on);

/* Fixed corrections dT1bar_dev[ng][nv][no], dT2bar_dev[ng][nv][nv][no][no]
 * (device, caller-owned, read by every later step until cleared with NULL,
 * NULL) added to the residual of ftcc_step / ftcc_step_graph before the
 * integration: the step then iterates Tbar(T) + dTbar.  With dTbar =
 * Tbar_fp64(T0) - Tbar_low(T0) a low-precision step reproduces the FP64 step
 * at T0 (mixed-precision defect correction; no counterpart in the
 * reference, which iterates in FP64 only). */
int ftcc_set_correction(ftcc_ctx *ctx, const double *dT1bar_dev,
                        const double *dT2bar_dev);
/* The defect of a low-precision residual at T:
 *   dTbar = Tbar_fp64(T) - Tbar_low(T)      (every grid point, skip_mask as
 * ftcc_residual), with S1_dev/S2_dev amplitude-size scratch.  Both residuals
 * are evaluated as ftcc_step evaluates its own, so that a low-precision step
 * from T with this correction is the FP64 step from T to rounding (the
 * spin-forbidden doubles of dT2bar, which no step reads, are zero).  The
 * context's precision is unchanged on return (its workspace is rebuilt for
 * the other precision in between).  The source of ftcc_set_correction's
 * dTbar. */
int ftcc_defect(ftcc_ctx *ctx, const double *T1_dev, const double *T2_dev,
                int low_precision, double *dT1bar_dev, double *dT2bar_dev,
                double *S1_dev, double *S2_dev, uint64_t skip_mask);

/* ---- Adjoint of the converged fixed point (E, S, N) ----------------------
 * kelvin's compute_ESN (kelvin/ccsd.py:128-214) solves lambda equations
 * (ft_lambda_iter, kelvin/cc_utils.py:414-480, over cqcpy _Lambda_opt via
 * kelvin/ft_cc_equations.py:385-409) and contracts response densities with
 * the derivative integrals (kelvin/ccsd.py:1075-1148, 1333-1492).  These
 * entry points give the same derivative as the adjoint of T = Phi(T):
 *   Lam = dE/dT - J^T(Int^T Lam),  dOmega_cc = dE + sum_y <Lam_y, dPhi_y>.
 * Adjoint amplitudes Lam1[y][a][i], Lam2[y][a][b][i][j] (the amplitude
 * layout) are kelvin's L up to a factor per grid point:
 *   Lam1_y = -(g_y/beta) L1_y^T,   Lam2_y = -(g_y/(4 beta)) L2_y^T. */

/* Fix the amplitudes T1_dev/T2_dev (read again by ftcc_adjoint_apply: keep
 * them unchanged until ftcc_adjoint_end) and build the amplitude-only
 * intermediates of J^T for as many grid points as FTCC_ADJOINT_CACHE_GB
 * (default 64) holds; the others rebuild theirs at each application.
 * Releases the residual workspace of the forward iteration. */
int ftcc_adjoint_begin(ftcc_ctx *ctx, const double *T1_dev,
                       const double *T2_dev);
/* Grid points whose intermediates are kept. */
int ftcc_adjoint_cached(ftcc_ctx *ctx, int *nslices);
/* G = J^T Lt at every grid point: the transpose of the Jacobian of the
 * residual (cqcpy _Lambda_opt's operator), Lt2 antisymmetric in (a,b) and
 * in (i,j); G2 is projected onto such tensors.  FP64, or with its large
 * O(n^6) blocks as split-bf16 products after ftcc_adjoint_precision. */
int ftcc_adjoint_apply(ftcc_ctx *ctx, const double *Lt1_dev,
                       const double *Lt2_dev, double *G1_dev, double *G2_dev);
int ftcc_adjoint_end(ftcc_ctx *ctx);
/* The unrelaxed one-particle response densities of the converged lambda
 * amplitudes (kelvin/ft_cc_equations.py:704-722 ccsd_1rdm, as used by
 * kelvin/ccsd.py:1333-1387 _g_ft_1rdm): from T1/T2 and the integrated
 * adjoint amplitudes Lt1/Lt2 = ftcc_integrate_adjoint(Lam) (engine layout),
 * the host weights g and beta, writes pia [no][nv], pba [nv][nv],
 * pji [no][no] and pai [nv][no] (device), each the derivative of kelvin's
 * Lagrangian with respect to one Fock block times beta. */
int ftcc_density1(ftcc_ctx *ctx, const double *T1_dev, const double *T2_dev,
                  const double *Lt1_dev, const double *Lt2_dev,
                  const double *g_host, double beta, double *pia_dev,
                  double *pba_dev, double *pji_dev, double *pai_dev);
/* Contraction precision of the large blocks of ftcc_adjoint_apply:
 * FTCC_PREC_FP64 (default) or FTCC_PREC_BF16X3 (the low-precision phase of
 * a refined lambda iteration).  Unbatched blocks of at least
 * FTCC_ADJOINT_LOWP_MIN_GFLOP GFLOP (environment, read by this call; default
 * 10) run split-bf16.  No counterpart in the reference (its lambda equations
 * are FP64 numpy). */
int ftcc_adjoint_precision(ftcc_ctx *ctx, int precision);
/* out[x] = sum_y G[y,x] exp((t_x - t_y) D) L[y] (x < y; weight 1 for
 * x >= y): the transpose of ftcc_integrate over the grid points
 * (kelvin/quadrature.py:320-345 int_L1/int_L2 without their g weights). */
int ftcc_integrate_adjoint(ftcc_ctx *ctx, const double *L1_dev,
                           const double *L2_dev, double *out1_dev,
                           double *out2_dev);
/* One damped lambda step after ftcc_adjoint_apply:
 *   Lam <- damp Lam + (1 - damp) (dE/dT - G)
 * with dE/dT1_y = (g_y/beta)(F.ov^T + oovv.t1_y), dE/dT2_y = (g_y/(4 beta))
 * oovv^T; out4 = {|dL1|^2, |dL2|^2, |L1|^2, |L2|^2} in kelvin's lambda
 * scale (the norms of kelvin/cc_utils.py:440-470). */
int ftcc_lambda_update(ftcc_ctx *ctx, double *Lam1_dev, double *Lam2_dev,
                       const double *G1_dev, const double *G2_dev,
                       const double *T1_dev, const double *g_host,
                       double beta, double damp, double out4[4]);
/* kelvin's first lambda amplitudes (kelvin/ft_cc_equations.py:502-512) in
 * the adjoint layout. */
int ftcc_lambda_guess(ftcc_ctx *ctx, const double *T1_dev,
                      const double *g_host, double beta, double *Lam1_dev,
                      double *Lam2_dev);
/* Adjoint amplitudes <-> kelvin's L1[y][i][a], L2[y][i][j][a][b]
 * (to_kelvin != 0: src is Lam, dst is L). */
int ftcc_lambda_convert(ftcc_ctx *ctx, int to_kelvin, const double *g_host,
                        double beta, const double *src1_dev,
                        const double *src2_dev, double *dst1_dev,
                        double *dst2_dev);
/* The change of the integrated amplitudes with beta through the grid,
 * contracted with Lam: sum_yx Lam_y (Gd[y,x] - G[y,x] (t_y - t_x) D / beta)
 * P_yx Tbar_x, with Tbar the residual at the solution and Gd_host = dG/dbeta
 * (kelvin/quadrature.py:238-256 d_ft_quad); the grid terms of
 * kelvin/ccsd.py:1113-1146. */
int ftcc_grid_derivative(ftcc_ctx *ctx, const double *Lam1_dev,
                         const double *Lam2_dev, const double *T1bar_dev,
                         const double *T2bar_dev, const double *Gd_host,
                         double beta, double *out);
/* sum_k x[k] y[k] over n doubles. */
int ftcc_dot(ftcc_ctx *ctx, const double *x_dev, const double *y_dev,
             int64_t n, double *out);
/* Replace the context's integrals by their derivative along a change of
 * x_p = beta (e_p - mu) by dx_dev[p], the bare Fock matrix moving by
 * dfock_dev[n*n]: every Fermi-weighted block of ftcc_set_system
 * differentiated through sqrt(fo), sqrt(fv) and the Fock matrix (what
 * kelvin contracts with its response densities in g_d_on_oo, g_d_on_vv and
 * g_Fd_on, kelvin/cc_utils.py:1628-1700).  The residual is linear in the
 * integrals, so ftcc_residual then gives dTbar/dtheta and ftcc_energy the
 * explicit derivative of the grand potential.  ftcc_set_system restores
 * the integrals. */
int ftcc_set_system_response(ftcc_ctx *ctx, int n, const double *en_dev,
                             const double *fock_dev, const double *dfock_dev,
                             const double *eri_dev, const double *dx_dev,
                             double beta, double mu, const int *iocc_host,
                             const int *ivir_host);

/* Quadrature weights g_host[ng] for ftcc_step_graph. */
int ftcc_set_weights(ftcc_ctx *ctx, const double *g_host);

/* ftcc_step with the weights of ftcc_set_weights, replayed as a HIP graph:
 * the first call with a set of arguments (pointers, beta, damp, skip mask,
 * precision) runs eagerly, the second captures the step's device work into
 * a graph and launches it, later calls relaunch it; out5 as ftcc_step (one
 * host synchronisation per call).  The context's stream must not be the
 * null stream; in the spin layout it needs ftcc_set_trusted; product
 * profiling is off.  Any ftcc_set_* call drops
 * the graph.  No reference counterpart (host-side launch overhead). */
int ftcc_step_graph(ftcc_ctx *ctx, double *T1old_dev, double *T2old_dev,
                    double *T1bar_dev, double *T2bar_dev, double beta,
                    double damp, uint64_t skip_mask, double out5[5]);

/* Device bytes held by the context (integrals + workspace). */
int ftcc_workspace_bytes(const ftcc_ctx *ctx, int64_t *bytes);

/* Kernel timing with HIP events on the context's stream (no counterpart in
 * the reference; used by the benchmark).  While enabled, every launch of an
 * O(n^6) product of the residual is bracketed by events.  ftcc_profile_read
 * synchronises the stream and returns, since the last read, the number of
 * such launches, their summed device time in ms and their summed
 * algorithmic flop count (2*M*N*K per product and grid point). */
int ftcc_profile_enable(ftcc_ctx *ctx, int on);
int ftcc_profile_read(ftcc_ctx *ctx, int64_t *launches, double *ms,
                      double *flops);
/* Launch plan of the generic GEMMs (the O(n^5) terms, and everything in the
 * FP64/FP32 modes): tile_policy 0 picks the tile by the shape (128x64 or
 * 64x128 when a dimension or K is at most 64, else 128x128), 1 keeps
 * 128x128 unless a dimension is at most 64, 2 always narrow; K is split
 * while a launch has fewer than split_target workgroups (default 1024).
 * Results are the same up to the summation order of split K.  No
 * counterpart in the reference (the tile autotuner of the batch sets it). */
int ftcc_set_gemm_plan(ftcc_ctx *ctx, int tile_policy, int split_target);

/* The same for every kernel class while profiling is on: 0 the O(n^6)
 * products (as ftcc_profile_read), 1 the other GEMMs (O(n^5) terms), 2 the
 * product operand gathers, 3 the doubles assembly, 4 the integration and
 * damping update of ftcc_step; flops and bytes are algorithmic (2MNK; each
 * operand read and each output written once). */
int ftcc_profile_class(ftcc_ctx *ctx, int cls, int64_t *launches, double *ms,
                       double *flops, double *bytes);

/* Batched strided GEMM of the residual engine, exposed for testing:
 * C[b] = alpha * A[b] . B[b] + beta * C[b] with element strides
 * A(m,k) = A + b*sAb + m*sAm + k*sAk, likewise B(k,n), C(m,n);
 * float64 operands in memory, contraction precision FTCC_PREC_*. */
int ftcc_gemm(ftcc_ctx *ctx, int M, int N, int K, int batch, double alpha,
              const double *A, int64_t sAm, int64_t sAk, int64_t sAb,
              const double *B, int64_t sBk, int64_t sBn, int64_t sBb,
              double beta, double *C, int64_t sCm, int64_t sCn, int64_t sCb,
              int precision);

#ifdef __cplusplus
}
#endif
#endif /* FTCC_H */
