/*
 * specmc_hip.h -- C ABI of the MI355X (gfx950) kernels behind SpecMC's
 * time-step hot path.
 *
 * The reference (edwardaud/SpecMC) is pure numpy/scipy and has no FFI; the
 * drop-in boundary is its Python model protocol (step / cfl_dt / stokes /
 * diagnostics, see specmc/api.py:147-190 and core/rbc_galerkin.py:258-497).
 * The host side of that protocol lives in the Python package specmc_amd; every
 * numerical operation it performs per step goes through the entry points
 * below.  Each declaration names the reference code whose arithmetic it
 * replaces.
 *
 * Conventions
 *  - All entry points return 0 on success and a nonzero status otherwise;
 *    smx_last_error() returns a message for the calling thread.  No C++
 *    exception crosses the ABI.
 *  - Pointers are device pointers (HBM), owned by the caller; sizes are
 *    element counts.  `stream` is a hipStream_t passed as void*.
 *  - `prec` selects the arithmetic type: 0 = double (FP64), 1 = float (FP32).
 *  - Fields are real: with the mirror (box) symmetry every theta-type field
 *    has real x-mode coefficients and every stream-function-type field purely
 *    imaginary ones; the latter are stored by their imaginary part.
 *  - `active` (int[B], may be NULL) masks runs: entries != 0 are processed.
 */
#ifndef SPECMC_HIP_H
#define SPECMC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* library identification and error text */
int smx_version(void);
const char *smx_last_error(void);

/* Launch clock (measurement only; no counterpart in the reference): after
 * smx_clock_start(tags, capacity), with tags a bit mask of (1 << tag), HIP
 * events are recorded on the launch stream immediately before and after each
 * of the next `capacity` launches of the clocked entry points (tag 1:
 * smx_fg_post, 2: smx_inv_apply, 3: smx_dense_lu_solve), and those launches
 * count the algorithmic units they process on the device (fg_post: vectors
 * of length n read or written, j + 5 per Arnoldi step j, 2 per restart;
 * inv_apply: systems solved; dense_lu_solve: matrices solved).
 * smx_clock_stop(tag, ...) waits for that tag's launches, returns their
 * summed duration (ms), the launch count and the unit count, and releases
 * its events.  Process-wide. */
int smx_clock_start(int tags, int capacity);
int smx_clock_stop(int tag, double *total_ms, int *count, long long *units);

/* General strided batched GEMM
 *   C[z](i,j) = alpha * rs(i) * sum_k A[z](i,k) B[z](k,j) + beta * C[z](i,j)
 * with A(i,k) = A[i*ars + k*acs], B(k,j) = B[k*brs + j*bcs],
 * C(i,j) = C[i*crs + j*ccs]; batch z = z1*nz2 + z2 offsets
 * (sa1*z1 + sa2*z2) etc.; rs(i) = rscale[i % rperiod] (rscale may be NULL).
 * Masking: mask_mode 0 none, 1 run = i / rows_per_run, 2 run = z1, 3 run = z2.
 * Replaces the dense products of the reference: the z-transforms V@, Vi@,
 * D@, D2@ (rbc_galerkin.py:322-339, implicit_theta.py:296-304), the x-mode
 * DFTs of np.fft.rfft/irfft on the mirror-symmetric half grid
 * (rbc_galerkin.py:322-339, coupled_model.py:440-449), the Green applies
 * (stokes_coupled.py:318-325, lagged_stokes.py:195-200) and the operator
 * builds (vv_stokes.py:266-307, bfbt.py:157-180). */
int smx_gemm(int prec, int M, int N, int K, double alpha, double beta,
             const void *A, long ars, long acs, long sa1, long sa2,
             const void *B, long brs, long bcs, long sb1, long sb2,
             void *C, long crs, long ccs, long sc1, long sc2,
             int nz1, int nz2, const void *rscale, int rperiod,
             const int *active, int mask_mode, int rows_per_run,
             void *stream);

/* smx_gemm with a role tag (no counterpart in the reference): the role only
 * names the kernel symbol, so kernel traces separate the work: 0 z-transform
 * or z-derivative (what smx_gemm launches), 1 x-transform (the half-grid
 * DFTs), 2 per-mode operators (Green applies, BFBT / reduced-operator
 * builds), 3 dense-matrix work of the rescue route (Schur assembly, LU
 * trailing update). */
int smx_gemm_ex(int role, int prec, int M, int N, int K, double alpha, double beta,
                const void *A, long ars, long acs, long sa1, long sa2,
                const void *B, long brs, long bcs, long sb1, long sb2,
                void *C, long crs, long ccs, long sc1, long sc2,
                int nz1, int nz2, const void *rscale, int rperiod,
                const int *active, int mask_mode, int rows_per_run,
                void *stream);

/* Kernel family of the GEMMs of smx_gemm without a tuned choice,
 * process-wide: 0 vector FMA tiles, 1 (default) matrix cores
 * (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32) with 16-wide K chunks, 2
 * matrix cores with the whole K extent staged at once when K <= 160.  All
 * give the same products to roundoff. */
int smx_set_gemm_mfma(int mode);

/* Kernel variants, process-wide (no counterpart in the reference; both forms
 * give the same results to roundoff): smx_set_xprod_mfma(1) (default) runs
 * the FP64 smx_xprod on the matrix cores, 0 on the vector-FMA kernel;
 * smx_set_gj_rows(1) inverts with the row-owner Gauss-Jordan kernel, 0
 * (default) with the first one; smx_set_inv_pairs(1) (default) applies FP64
 * inverses of even order with the row-pair kernel that issues all of a
 * block's loads first, 2 with the row-pair kernel that loads as it
 * accumulates, 0 with the first one. */
int smx_set_xprod_mfma(int on);
int smx_set_gj_rows(int on);
/* smx_set_gj_reg(1) (default) inverts FP64 blocks of order 16..128 with the
 * matrix held in registers (4N threads), 0 with the kernels above. */
int smx_set_gj_reg(int on);
/* smx_set_gj_mfma(1) (default) inverts FP64 blocks of order 16..128 by
 * Gauss-Jordan blocked by 16 columns: each 16-column panel is factored in one
 * wavefront and applied to the other columns as a K = 16 product on the
 * matrix cores (v_mfma_f64_16x16x4f64); 0 uses the kernels above.  The same
 * pivot choice; the inverses agree to roundoff. */
int smx_set_gj_mfma(int on);
int smx_set_inv_pairs(int on);

/* Autotuner table (no counterpart in the reference; the choices are measured
 * by specmc_amd.autotune): the kernel used for GEMMs of precision `prec`,
 * shape M x N x K, nz batch entries and operand layout (a_kfast: acs == 1,
 * b_jfast: bcs == 1).  code: 0..6 vector FMA tiles 128x96, 128x64, 64x96,
 * 64x64, 64x32, 32x64, 32x32; 10 matrix cores 64x64 (K chunks of 16);
 * 11 matrix cores 32x32 (K chunks of 16); 12 matrix cores 32x32 with the
 * whole K (<= 160) staged at once; a negative code removes the entry.
 * smx_gemm_tune_clear empties the table.  Process-wide. */
int smx_gemm_tune(int prec, int M, int N, int K, int nz, int a_kfast, int b_jfast,
                  int code);
int smx_gemm_tune_clear(void);

/* Batched LU with partial pivoting, in place, column-major N x N matrices
 * (LAPACK getrf semantics).  Replaces scipy.linalg.lu_factor in
 * ChebTau._factors_for (rbc_galerkin.py:119-144), VVStokesReduced
 * (vv_stokes.py:245-278) and SpectralBFBTPrecond._build (bfbt.py:157-181).
 * Matrix m of batch belongs to run m / per_run. */
int smx_lu_factor(int prec, int nmat, int N, void *LU, int *piv,
                  int per_run, const int *active, void *stream);

/* Batched LU solve (getrs, one right-hand side per matrix):
 *   x = LU^{-1} ( scale * b ) with b[N-2] = b[N-1] = 0 when tau_zero,
 * matrix m, rhs at rhs + m*N; run = m / per_run; the scale is
 * rhs_scale * kscale[m % per_run] * bscale[run] (either may be NULL).  With
 * shared != 0 the factors of matrix m are those of m % per_run (one set of
 * factors serves every run, e.g. the isoviscous lam = 0 operators).
 * Replaces sla.lu_solve in ChebTau.solve (rbc_galerkin.py:171-185),
 * VVStokesReduced.solve (vv_stokes.py:309-321) and
 * SpectralBFBTPrecond.apply (bfbt.py:185-191).  In-place allowed. */
int smx_lu_solve(int prec, int nmat, int N, const void *LU, const int *piv,
                 const void *rhs, void *x, int tau_zero, double rhs_scale,
                 const void *kscale, const void *bscale, int per_run,
                 int shared, const int *active, void *stream);

/* Batched explicit inverses (in-place Gauss-Jordan with partial row
 * pivoting, column-major N x N) and their application
 *   x = inv(A_m) (scale * b) with the tau entries of b zeroed when tau_zero,
 * scale / run / shared conventions as smx_lu_solve.  The FGMRES
 * preconditioners of the BFBT route (SpectralBFBTPrecond.apply,
 * bfbt.py:185-191) and of the implicit theta solve (ImplicitThetaSolver.precond,
 * implicit_theta.py:452-474) apply the same per-mode operators as a
 * bandwidth-bound matrix-vector product. */
int smx_invert(int prec, int nmat, int N, void *A, int per_run,
               const int *active, void *stream);
int smx_inv_apply(int prec, int nmat, int N, const void *Inv, const void *rhs,
                  void *x, int tau_zero, double rhs_scale, const void *kscale,
                  const void *bscale, int per_run, int shared,
                  const int *active, void *stream);

/* smx_set_lud_fuse(1) (default): in the FP32 factorisation of smx_dense_lu
 * the sub-panel kernel also applies the sub-panel's row swaps to the rest of
 * its 128-column block and solves that block's U rows, and the rank-16
 * update below is a matrix-core GEMM; 0 runs the separate block-update
 * kernel.  No counterpart in the reference (a kernel variant; the factors
 * agree to roundoff). */
int smx_set_lud_fuse(int on);

/* Batched dense LU with partial pivoting, FP64 (prec 0, n <= 8192) or FP32
 * (prec 1, n <= 32768), row-major n x n matrices (matrix q at
 * A + q*mstride).  piv int[nmat][n]:
 * at step i rows i and piv[i] were exchanged (0-based, LAPACK order);
 * info int[nmat]: 0, or 1 + the first column with a zero pivot.  Replaces
 * sla.lu_factor of the assembled Schur matrix in LaggedStokes.refresh
 * (lagged_stokes.py:274-294), which the BFBT rescue calls at the current eta
 * (coupled_model.py:613-654), and the direct dense solve
 * (coupled_model.py:649-650). */
int smx_dense_lu(int prec, int nmat, int n, void *A, long mstride, int *piv,
                 int *info, void *stream);
/* x_q = A_q^{-1} b_q from smx_dense_lu factors; b and x are [nmat][n]
 * (sla.lu_solve of the defect correction, coupled_model.py:544-554).
 * prec 0: FP64 factor and vectors; 1: FP32 factor and vectors; 2: FP32
 * factor, FP64 b and x (the preconditioner of the mixed-precision rescue).
 * active: optional int[nmat]; matrices whose flag is 0 leave x untouched. */
int smx_dense_lu_solve(int prec, int nmat, int n, const void *LU, long mstride,
                       const int *piv, const void *b, void *x, const int *active,
                       void *stream);
/* The row order of smx_dense_lu's pivots: perm[q][i] is the row of b that
 * row i of P_q b takes (the swaps of piv applied in order).  Computed once
 * per factorisation for smx_dense_lu_solve_mw; the reference applies the
 * pivots inside sla.lu_solve (lagged_stokes.py:302-343). */
int smx_dense_lu_perm(int nmat, int n, const int *piv, int *perm, const int *active,
                      void *stream);
/* Bytes of device workspace smx_dense_lu_solve_mw needs (0 for bad sizes). */
long smx_dense_lu_solve_mw_bytes(int prec, int nmat, int n);
/* x_q = A_q^{-1} b_q as smx_dense_lu_solve (same prec codes, same factors),
 * with each matrix's two triangular sweeps split over workgroups of 64 rows,
 * so that a few active matrices still occupy the whole device (the
 * sla.lu_solve of the defect correction and of the frozen-factor FGMRES
 * preconditioner, coupled_model.py:544-554, lagged_stokes.py:345-360).
 * perm from smx_dense_lu_perm; b and x must not overlap; ws holds at least
 * smx_dense_lu_solve_mw_bytes(prec, nmat, n) bytes of device memory, which
 * the call clears and uses on its stream: concurrent calls need their own
 * workspaces.  Its 32-bit word 2 is nonzero after a solve whose wait between
 * workgroups ran past its bound (the solution is then not valid). */
int smx_dense_lu_solve_mw(int prec, int nmat, int n, const void *LU, long mstride,
                          const int *perm, const void *b, void *x, const int *active,
                          void *ws, long ws_bytes, void *stream);

/* Chebyshev-tau operators, column-major, per (run, mode):
 *   kind 0: D2 - (k^2 + lam_b) I, rows N-2, N-1 = (-1)^n, 1
 *           (ChebTau._assemble, rbc_galerkin.py:164-169);
 *   kind 1: L_k M_b + 2 k^2 C_b,k with rows N-2, N-1 = V[0], V[-1] where
 *           LM = D2@M_b is given per run (lm) and C = Comm_b @ G_k per
 *           (run, mode) (cg): VVStokesReduced._assemble (vv_stokes.py:302-307)
 *           and the BFBT block (bfbt.py:157-180). */
int smx_tau_assemble(int prec, int kind, int B, int Nk, int N,
                     const void *D2, const void *kx, const void *lam,
                     const void *Mb, const void *lm, const void *cg,
                     const void *V0, const void *V1, void *out,
                     const int *active, void *stream);

/* Exact truncated Chebyshev multiplication matrices M[b] (row-major N x N)
 * from coefficient vectors a[b] (cheb_mult_matrix, vv_stokes.py:81-109). */
int smx_cheb_mult(int prec, int B, int N, const void *a, void *M,
                  const int *active, void *stream);

/* Dense Schur matrix of one run in real form (CoupledStokes.assemble_schur,
 * stokes_coupled.py:242-289), used by the BFBT rescue route
 * (coupled_model.py:613-654).  From the per-shift Chebyshev multiplication
 * matrices M_d and P = D2 M, Qm = M D2, R = D M D ([Nk][N][N] row-major):
 * schur_blocks writes A_ij = (P_d - k_i^2 M_d) + s(P_d' - k_i^2 M_d') into the
 * dense row-major A (ld = Nk*N) and the bracket U_ij that multiplies G_j into
 * U [Nk][Nk][N][N] (d = |i - j|, d' = i + j, s = partner sign); the caller
 * adds U_ij G_j with smx_gemm; schur_tau writes the tau rows. */
int smx_schur_blocks(int prec, int Nk, int N, const void *M, const void *P,
                     const void *Qm, const void *R, const void *kx, double sgn,
                     void *A, void *U, void *stream);
int smx_schur_tau(int prec, int Nk, int N, const void *V0, const void *V1,
                  void *A, void *stream);

/* The viscosity law on the padded half grid (coupled_model.py:333-399):
 * T = 1 - z + theta, counters of T outside [0,1] / [0,2], T clipped to
 * [0,2], eta = exp(-ln(deta) T) (law 0), exp(-ln(deta)(1+Ts) T/(T+Ts))
 * (law 1) or the pressure-dependent exp(-E T/(Ts (T+Ts)) + V (1-z)/(T+Ts))
 * (law 2, arrhenius.py:283-325), then the window [lo, hi] with the 1e-12
 * dead-band (eta_bounds.py:360-379) or the C^2 smooth clip
 * (eta_bounds.py:411-480).  One block per run; grid rows carry mirror
 * weights (row 0 and row nrow-1 count once, the others twice).  law params
 * per run (double[8]): [ln_deta, Ts, lo, hi, smooth_w, law, E, V]; counters
 * per run (double[4]): n_T_out, n_clip_T, n_clip_eta, max_T_excursion
 * (accumulated). */
int smx_eta_law(int prec, int B, int nrow, int N, const void *theta,
                const void *z, const double *params, void *eta,
                double *counters, const int *active, void *stream);

/* Positivity guard and BFBT profile on the unpadded half grid
 * (coupled_model.py:451-496, bfbt.py:99-108): if min(eta) <= 0 the window is
 * re-applied (counted in counters[4]); then prof(z) = 1/mean_x(1/eta)
 * (profile 0, 'harm'), mean_x(eta) (1, 'diag') or exp(mean_x(log eta))
 * (2, 'geom') over the full Nx grid. */
int smx_eta_guard_profile(int prec, int B, int Nx, int N, void *eta,
                          const double *params /* double[8] per run, as above */, void *prof, int profile,
                          double *counters, const int *active, void *stream);

/* Pseudo-spectral eta product of the coupled Schur operator
 * (FreeSchurOperator.mult_eta, lagged_stokes.py:162-193, both halves of the
 * one-sided convolution, partner sign -1, the cross term flipped): for each
 * run, field f (0..nf-1) and padded z node q: synthesise the x-modes of
 * y[b][f][:, q] on the mirror half of the 2Nx grid (odd parity for f < nodd,
 * even otherwise), multiply by eta[b][:, q] and analyse back to Nk modes.
 * y/out layout [B][Nk][nf][Q]; fscale [nf][Nk] (may be NULL) scales the
 * input modes (the 2 k^2 and k factors inside the convolution); eta of run
 * b starts at eta + b*eta_bs (eta_bs = 0: one eta field for every column of
 * a dense assembly). */
int smx_xprod(int prec, int B, int nf, int nodd, int Nk, int nrow, int Q,
              const void *y, const void *fscale, const void *eta, long eta_bs,
              void *out,
              const void *So, const void *Ce, const void *Ao, const void *Ae,
              const int *active, void *stream);

/* The coupled Schur operator apply (FreeSchurOperator.matvec,
 * lagged_stokes.py:202-220, with mult_eta of lagged_stokes.py:162-193
 * evaluated pseudo-spectrally) for B runs in three launches:
 *   yz[b][k] = W_k v[b][k]     (Wz [Nk][Nz][3Q]: the transposed stacks
 *                               [Vp + 2k^2 Vp G_k; 2 Vp D2 G_k - Vp;
 *                                k Vp D G_k])
 *   tz = smx_xprod(yz, eta)    (3 fields, the first 2 odd in x)
 *   out[b][k] = Mc^T (u0, k^2 u1, -4k u2) on rows < Nz-2
 *               (Mc [3Q][Nz] = [(D2 Tp)^T; Tp^T; (D Tp)^T]), and the tau rows
 *               out[Nz-2] = V0 . v[b][k], out[Nz-1] = V1 . v[b][k].
 * v and out are [B][Nk][Nz]; eta of run b at eta + b*eta_bs, [nrow][Q];
 * So/Ce [nrow][Nk], Ao/Ae [Nk][nrow] as smx_xprod; kx [Nk].  With a mask,
 * the z-synthesis runs over the active runs only, listed on the device.  The
 * workspace holds yz, tz and that list: at least smx_schur_apply_work(B, Nk,
 * Q) elements of the precision's type, (6 Nk Q + 2) per run. */
long long smx_schur_apply_work(int B, int Nk, int Q);
/* smx_set_schur_compact(1) (default): with a mask, the z-synthesis of
 * smx_schur_apply runs over the active runs listed on the device (a 64-run
 * tile is no longer computed whole for one active run in it); 0 runs it over
 * every run and skips only tiles without an active run.  No counterpart in
 * the reference (a kernel variant; the same outputs bit for bit). */
int smx_set_schur_compact(int on);
int smx_schur_apply(int prec, int B, int Nk, int N, int nrow, int Q, const void *v,
                    const void *eta, long eta_bs, void *out, const void *Wz,
                    const void *Mc, const void *kx, const void *So, const void *Ce,
                    const void *Ao, const void *Ae, const void *V0, const void *V1,
                    void *work, long long work_elems, const int *active, void *stream);

/* smx_schur_apply without a workspace argument; no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * intermediates live in the library's own device array of
 * smx_schur_apply_static_bytes() bytes (64 MiB, one copy per device, held
 * until the process ends); the call allocates nothing.  Sizes whose
 * workspace (smx_schur_apply_work elements) exceeds it are rejected with an
 * error.  The first call on a device looks up the array's address there (make
 * that call before a capture).  All calls on one device share the array, so
 * calls that overlap on one device race; the caller serialises them (one
 * stream, or events). */
long long smx_schur_apply_static_bytes(void);
int smx_schur_apply_static(int prec, int B, int Nk, int N, int nrow, int Q, const void *v,
                           const void *eta, long eta_bs, void *out, const void *Wz,
                           const void *Mc, const void *kx, const void *So, const void *Ce,
                           const void *Ao, const void *Ae, const void *V0, const void *V1,
                           const int *active, void *stream);

/* Advection product of the implicit theta operator
 * (ImplicitThetaSolver.advect, implicit_theta.py:300-304):
 *   a = Ae * ( -(ux * So (ks0 y0) + uz * Ce y1) [+ e1] [+ h_b] [+ e2] )
 * per run and z node; y [B][Nk][2][N] holds z-values of the odd (k v) and
 * even (D v) inputs; e1/e2 extra grid terms and h a per-run constant (the
 * explicit right-hand side of RBC.explicit_rhs, rbc_galerkin.py:396-423). */
int smx_advect_prod(int prec, int B, int Nk, int nrow, int N,
                    const void *y, const void *ks0, const void *ux,
                    const void *uz,
                    const void *e1, const void *h, const void *e2,
                    void *out, const void *So, const void *Ce,
                    const void *Ae, const int *active, void *stream);

/* out[b] = max over the run's n_per_run grid values of
 * |eta - eta_lag| / max(|eta_lag|, 1e-300): the pointwise relative drift of
 * the viscosity since the frozen factorisation (RBCCoupled._eta_drift,
 * coupled_model.py:509-522). */
int smx_eta_drift(int prec, int B, long n_per_run, const void *eta, const void *eta_lag,
                  double *out, const int *active, void *stream);
/* out[b] = ||x[b, 0:n]||_2 for rows of stride ld (the np.linalg.norm of the
 * Picard loop in RBCCoupled._defect_solve, coupled_model.py:543-566). */
int smx_rownorm(int prec, int B, long n, long ld, const void *x, double *out,
                const int *active, void *stream);
/* out[b*ldo] = sum_i w[i] x[b*ldx + i] y[b*ldy + i] (y may be NULL), FP64
 * accumulation, for i < len: the quadratures of the energy budget -- the
 * volume mean <theta> (heating.py:275-276), <Phi> of the tau-zeroed
 * dissipation field (dissipation.py:384-392) and Ra <u_z theta> on the
 * padded grid (dissipation.py:423-444). */
int smx_wdot(int prec, int B, long len, const void *x, long ldx, const void *y, long ldy,
             const void *w, double *out, long ldo, const int *active, void *stream);
/* Elementwise / small kernels of the step.
 * axpby: out = (a*ab[b]) x + (b*bb[b]) y per run (ab, bb, y may be NULL).
 * rowscale: out[b][k][j] = (alpha*ab[b]) ks[k] x[b][k][j], tau entries zeroed
 *   when tau_zero -- the Stokes right-hand sides i k Ra theta
 *   (stokes_coupled.py:292-302, vv_stokes.py:311-312, rbc_galerkin.py:349).
 * tau_rows: the Dirichlet rows v(0), v(1) of an operator output.
 * stokes_combine: Lm(eta v) + D2(eta 2k^2 G v) + 2k^2 eta D2 G v
 *   - 4k D(eta k D G v) from the truncated products t[B][Nk][4][N] and
 *   dt = (D2 t0, D2 t1, D t3), tau rows from v (lagged_stokes.py:202-220).
 * theta_combine: a0 v - (D2 v - k^2 v) - adv, tau rows from v
 *   (implicit_theta.py:354-383).
 * theta_rhs: (b0 th - b1 thp) + k psi, tau entries zeroed, + H at (0, 0),
 *   + src on the equation rows (implicit_theta.py:385-449).
 * phi_grid: 2 eta ((gxx^2 + gzz^2 + 2 gxz^2) - tr^2/3), gzz = -gxx, times
 *   scale[b] (dissipation.py:313-362).
 * diag: Nu_top, Nu_bot, Nu_energy, urms, umax, amp (rbc_galerkin.py:439-453);
 *   dth0 = D theta_0 of run b at dth0 + b*ldd.
 * maxabs: per-run max |x|. */
int smx_axpby(int prec, long n_per_run, int B, double a, const void *ab,
              const void *x, double b, const void *bb, const void *y,
              void *out, const int *active, void *stream);
int smx_rowscale(int prec, int B, int Nk, int N, const void *x,
                 const void *ks, double alpha, const void *ab, int tau_zero,
                 void *out, const int *active, void *stream);
int smx_tau_rows(int prec, int B, int Nk, int N, const void *v,
                 const void *V0, const void *V1, void *out,
                 const int *active, void *stream);
int smx_stokes_combine(int prec, int B, int Nk, int N, const void *t,
                       const void *dt, const void *kx, const void *v,
                       const void *V0, const void *V1, void *out,
                       const int *active, void *stream);
int smx_theta_combine(int prec, int B, int Nk, int N, const void *v,
                      const void *d2v, const void *adv, const void *kx,
                      const void *a0, const void *V0, const void *V1,
                      void *out, const int *active, void *stream);
int smx_theta_rhs(int prec, int B, int Nk, int N, const void *th,
                  const void *thp, const void *psi, const void *kx,
                  const void *b0, const void *b1, const void *H,
                  const void *src, void *out, const int *active,
                  void *stream);
int smx_phi_grid(int prec, int B, long np, const void *gxx, const void *gxz,
                 const void *eta, const void *scale, void *out,
                 const int *active, void *stream);
int smx_diag(int prec, int B, int Nx, int N, const void *ux, const void *uz,
             const void *th, const void *dth0, long ldd, const void *V0,
             const void *V1, const void *wq, double *out,
             const int *active, void *stream);
int smx_maxabs(int prec, int B, long n_per_run, const void *x, double *out,
               const int *active, void *stream);

/* States with an x-antisymmetric part (isoviscous model only).  RBC.step,
 * RBC.cfl_dt and RBC.diagnostics (rbc_galerkin.py:396-471) take any complex
 * coefficient array; the device carries it as two real channels, a = Re
 * that (x-even fields) and b = Im that (x-odd fields).
 * advect_prod_general: the explicit right-hand side of RBC.explicit_rhs
 *   (rbc_galerkin.py:396-423) and its truncation (_trunc_pad,
 *   rbc_galerkin.py:338-339) for both channels: ya / yb [B][Nk][2][N] hold
 *   z-values of (theta, D theta) of channel a / b; ux_e..uz_o the parts of
 *   the velocity on the padded half grid [B][nrow][N]; out_e = Ae N_e,
 *   out_o = Ao N_o (h, e2_e, e2_o may be NULL).
 * phi_grid_general: even and odd parts of phi_grid (eta = 1) from the parts
 *   of gxx and gxz (dissipation.py:336-362).
 * diag_general: diag with the odd-parity fields of the general state added
 *   (ux_o, uz_e, th_e are the mirror-case inputs of smx_diag).
 * maxabs_pair: per-run max |x| + |y|, the full-grid max |f| of f = x +- y. */
int smx_advect_prod_general(int prec, int B, int Nk, int nrow, int N, const void *ya,
                            const void *yb, const void *ks, const void *ux_e,
                            const void *ux_o, const void *uz_e, const void *uz_o,
                            const void *h, const void *e2_e, const void *e2_o,
                            void *out_e, void *out_o, const void *So, const void *Ce,
                            const void *Ae, const void *Ao, const int *active, void *stream);
int smx_phi_grid_general(int prec, int B, long np, const void *gxx_e, const void *gxz_o,
                         const void *gxx_o, const void *gxz_e, const void *scale,
                         void *out_e, void *out_o, const int *active, void *stream);
int smx_diag_general(int prec, int B, int Nx, int N, const void *ux_o, const void *uz_e,
                     const void *th_e, const void *ux_e, const void *uz_o,
                     const void *th_o, const void *dth0, long ldd, const void *V0,
                     const void *V1, const void *wq, double *out, const int *active,
                     void *stream);
int smx_maxabs_pair(int prec, int B, long n_per_run, const void *x, const void *y,
                    double *out, const int *active, void *stream);

/* FGMRES state machine (krylov_fgmres.py:60-179), one solver instance per
 * run.  The per-run control block ctl (int[B][8]) and scalars sc
 * (double[B][8]) live in device memory; see the implementation for the
 * layout.  fg_init: x = 0, r = b, beta = ||b||, tol = max(atol, rtol*beta).
 * fg_cycle_end: y = H^{-1} g, x += Z y, then the inputs of this round's
 * preconditioner (V_j) and operator (x) apply.  fg_post: Z_j store, the
 * orthogonalisation (ortho 0: modified Gram-Schmidt in the reference's order;
 * 1: block Gram-Schmidt with reorthogonalisation, two passes over blocks of
 * basis vectors, for n <= 8192) + Givens + the stop tests, or the restart
 * from the true residual.  The input
 * pointer of the next operator apply (Z_j or x).  fg_count: number of runs
 * not yet finished (written to *out).  ctl is int[10][B] and sc double[8][B]
 * (struct of arrays); ctl + 5*B / ctl + 6*B are the preconditioner and
 * operator masks of the current round. */
int smx_fg_init(int prec, int B, long n, int m, const void *b, void *x,
                void *Vb, void *g, int *ctl, double *sc, double rtol,
                double atol, int maxiter, const int *active, void *stream);
int smx_fg_cycle_end(int prec, int B, long n, int m, void *x, const void *Zb,
                     const void *Vb, void *H, void *g, void *y, void *vin,
                     void *zout, int *ctl, void *stream);
int smx_fg_post(int prec, int B, long n, int m, const void *b, void *w,
                const void *zin, void *Zb, void *Vb, void *H, void *cs,
                void *sn, void *g, int *ctl, double *sc, int ortho,
                void *stream);
int smx_fg_count(int B, const int *ctl, int *out, void *stream);

/* CFL dt (rbc_galerkin.py:465-471, cfl_policy.py:64-74): per run
 * dt = min(dtmax, clip(safety * h / max(umax, 1e-12), dtmin, dtmax)). */
int smx_cfl_dt(int B, const double *umax, const double *params, double *dt,
               const int *active, void *stream);

#ifdef __cplusplus
}
#endif
#endif
