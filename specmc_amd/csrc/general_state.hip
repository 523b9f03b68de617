// Kernels for isoviscous states with an x-antisymmetric part.
//
// RBC.step, RBC.cfl_dt and RBC.diagnostics (rbc_galerkin.py:396-471) take
// any complex coefficient array that[k][:].  The device keeps such a state
// as two real channels, a = Re that (x-even grid fields, cosine synthesis)
// and b = Im that (x-odd, sine synthesis).  Every per-mode operator of the
// model is real, so the linear solves run on each channel with the same
// factors; what couples the channels is the quadratic part, evaluated here
// on the half grid j = 0..nrow-1 from the even part e and the odd part o of
// each factor: a product f g has the even part f_e g_e + f_o g_o and the odd
// part f_e g_o + f_o g_e, and a full-grid maximum of |f| is max |f_e|+|f_o|
// over the half grid (x_j and x_{-j} carry f_e + f_o and f_e - f_o).
#include <hip/hip_runtime.h>

#include "smx_common.h"
#include "specmc_hip.h"

// padded-grid columns (z nodes) per workgroup; the LDS holds
// (4 Nk + 2 nrow) * XG values
#define XG 8

// explicit right-hand side of RBC.explicit_rhs (rbc_galerkin.py:396-423) for
// a two-channel state, followed by its truncated analysis (_trunc_pad):
//   ya / yb [B][Nk][2][N]: z-values of (theta, D theta) of channels a / b;
//   ux_e, ux_o, uz_e, uz_o [B][nrow][N]: velocity parts on the padded half grid;
//   theta_x = Ce(-k b) + So(k a), theta_z = Ce(D a) + So(D b);
//   N = -(ux theta_x + uz theta_z) + uz [+ h_b] [+ e2];
//   out_e = Ae N_e (real coefficients), out_o = Ao N_o (imaginary ones).
template <typename T>
__global__ void __launch_bounds__(256)
    k_advect_gen(int Nk, int nrow, int N, const T *ya, const T *yb, const T *ks, const T *uxe,
                 const T *uxo, const T *uze, const T *uzo, const T *h, const T *e2e,
                 const T *e2o, T *oute, T *outo, const T *So, const T *Ce, const T *Ae,
                 const T *Ao, const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *KA = (T *)smem_raw;   // k (theta a)        [Nk][XG]
  T *DA = KA + Nk * XG;    // D (theta a)
  T *KB = DA + Nk * XG;    // -k (theta b)
  T *DB = KB + Nk * XG;    // D (theta b)
  T *PE = DB + Nk * XG;    // N_e               [nrow][XG]
  T *PO = PE + nrow * XG;  // N_o
  const int b = blockIdx.y, q0 = blockIdx.x * XG;
  if (active && !active[b]) return;
  const long yo = (long)b * Nk * 2 * N;
  for (int e = threadIdx.x; e < Nk * XG; e += blockDim.x) {
    const int m = e / XG, c = e % XG, q = q0 + c;
    const bool in = q < N;
    const long r = yo + (long)m * 2 * N + q;
    const T k = ks[m];
    KA[e] = in ? k * ya[r] : T(0);
    DA[e] = in ? ya[r + N] : T(0);
    KB[e] = in ? -(k * yb[r]) : T(0);
    DB[e] = in ? yb[r + N] : T(0);
  }
  __syncthreads();
  const long gb = (long)b * nrow * N;
  for (int e = threadIdx.x; e < nrow * XG; e += blockDim.x) {
    const int j = e / XG, c = e % XG, q = q0 + c;
    T xo = 0, xe = 0, ze = 0, zo = 0;
    const T *Sr = So + (long)j * Nk;
    const T *Cr = Ce + (long)j * Nk;
    for (int m = 0; m < Nk; ++m) {
      const T s = Sr[m], cc = Cr[m];
      xo = fma(s, KA[m * XG + c], xo);
      xe = fma(cc, KB[m * XG + c], xe);
      ze = fma(cc, DA[m * XG + c], ze);
      zo = fma(s, DB[m * XG + c], zo);
    }
    T pe = 0, po = 0;
    if (q < N) {
      const long o = gb + (long)j * N + q;
      const T ue = uxe[o], uo = uxo[o], we = uze[o], wo = uzo[o];
      pe = -(uo * xo + we * ze) - (ue * xe + wo * zo) + we;
      po = -(ue * xo + uo * xe) - (we * zo + wo * ze) + wo;
      if (h) pe = pe + h[b];
      if (e2e) pe = pe + e2e[o];
      if (e2o) po = po + e2o[o];
    }
    PE[e] = pe;
    PO[e] = po;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < Nk * XG; e += blockDim.x) {
    const int m = e / XG, c = e % XG, q = q0 + c;
    if (q >= N) continue;
    const T *Er = Ae + (long)m * nrow;
    const T *Or = Ao + (long)m * nrow;
    T se = 0, so = 0;
    for (int j = 0; j < nrow; ++j) {
      se = fma(Er[j], PE[j * XG + c], se);
      so = fma(Or[j], PO[j * XG + c], so);
    }
    const long o = (long)b * Nk * N + (long)m * N + q;
    oute[o] = se;
    outo[o] = so;
  }
}

extern "C" int smx_advect_prod_general(int prec, int B, int Nk, int nrow, int N, const void *ya,
                                       const void *yb, const void *ks, const void *ux_e,
                                       const void *ux_o, const void *uz_e, const void *uz_o,
                                       const void *h, const void *e2_e, const void *e2_o,
                                       void *out_e, void *out_o, const void *So,
                                       const void *Ce, const void *Ae, const void *Ao,
                                       const int *active, void *stream) {
  if (B <= 0) return 0;
  if (!ya || !yb || !ks || !ux_e || !ux_o || !uz_e || !uz_o || !out_e || !out_o)
    return fail_msg("smx_advect_prod_general: null operand");
  dim3 grid(cdiv(N, XG), B);
  size_t sh = (size_t)(4 * Nk + 2 * nrow) * XG * (prec == 0 ? 8 : 4);
  if (!(prec == 0 ? lds_optin(k_advect_gen<double>, sh) : lds_optin(k_advect_gen<float>, sh)))
    return fail_msg("smx_advect_prod_general: grid too large for LDS");
#define ADV_ARGS(T)                                                                         \
  Nk, nrow, N, (const T *)ya, (const T *)yb, (const T *)ks, (const T *)ux_e, (const T *)ux_o, \
      (const T *)uz_e, (const T *)uz_o, (const T *)h, (const T *)e2_e, (const T *)e2_o,      \
      (T *)out_e, (T *)out_o, (const T *)So, (const T *)Ce, (const T *)Ae, (const T *)Ao,    \
      active
  if (prec == 0)
    hipLaunchKernelGGL(k_advect_gen<double>, grid, dim3(256), sh, STREAM(stream),
                       ADV_ARGS(double));
  else
    hipLaunchKernelGGL(k_advect_gen<float>, grid, dim3(256), sh, STREAM(stream),
                       ADV_ARGS(float));
#undef ADV_ARGS
  CHECK_LAUNCH("smx_advect_prod_general");
  return 0;
}

// viscous dissipation (dissipation.py:336-362, eta = 1) of a two-channel
// stream function: with gzz = -gxx and a zero trace, 2 (gxx^2 + gzz^2 + 2 gxz^2)
// has the even part 2 (2 (xx_e^2 + xx_o^2) + 2 (xz_e^2 + xz_o^2)) and the odd
// part 2 (4 xx_e xx_o + 4 xz_e xz_o); both times scale[b].
template <typename T>
__global__ void k_phi_gen(long np, const T *xxe, const T *xzo, const T *xxo, const T *xze,
                          const T *scale, T *oute, T *outo, const int *active) {
  const int b = blockIdx.y;
  if (active && !active[b]) return;
  const long o = (long)b * np;
  const T sc = scale ? scale[b] : T(1);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < np;
       e += (long)gridDim.x * blockDim.x) {
    const T a = xxe[o + e], c = xxo[o + e], s = xzo[o + e], d = xze[o + e];
    const T xx2 = a * a + c * c, xz2 = s * s + d * d;
    const T de = xx2 + xx2 + T(2) * xz2;
    const T dd = T(4) * (a * c) + T(4) * (s * d);
    oute[o + e] = sc * (T(2) * de);
    outo[o + e] = sc * (T(2) * dd);
  }
}
extern "C" int smx_phi_grid_general(int prec, int B, long np, const void *gxx_e,
                                    const void *gxz_o, const void *gxx_o, const void *gxz_e,
                                    const void *scale, void *out_e, void *out_o,
                                    const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(min(cdiv(np, 256), 1024), B);
  if (prec == 0)
    hipLaunchKernelGGL(k_phi_gen<double>, grid, dim3(256), 0, STREAM(stream), np,
                       (const double *)gxx_e, (const double *)gxz_o, (const double *)gxx_o,
                       (const double *)gxz_e, (const double *)scale, (double *)out_e,
                       (double *)out_o, active);
  else
    hipLaunchKernelGGL(k_phi_gen<float>, grid, dim3(256), 0, STREAM(stream), np,
                       (const float *)gxx_e, (const float *)gxz_o, (const float *)gxx_o,
                       (const float *)gxz_e, (const float *)scale, (float *)out_e,
                       (float *)out_o, active);
  CHECK_LAUNCH("smx_phi_grid_general");
  return 0;
}

// per-run max of |x| + |y| (the full-grid max |f| of f = x + y on one half
// and x - y on the other), written as double
template <typename T>
__global__ void __launch_bounds__(1024)
    k_maxabs_pair(long n, const T *x, const T *y, double *out, const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  double m = 0.0;
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    const long o = (long)b * n + e;
    m = fmax(m, fabs((double)x[o]) + fabs((double)y[o]));
  }
  m = block_max(m, red);
  if (threadIdx.x == 0) out[b] = m;
}
extern "C" int smx_maxabs_pair(int prec, int B, long n_per_run, const void *x, const void *y,
                               double *out, const int *active, void *stream) {
  if (B <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_maxabs_pair<double>, dim3(B), dim3(1024), 0, STREAM(stream),
                       n_per_run, (const double *)x, (const double *)y, out, active);
  else
    hipLaunchKernelGGL(k_maxabs_pair<float>, dim3(B), dim3(1024), 0, STREAM(stream),
                       n_per_run, (const float *)x, (const float *)y, out, active);
  CHECK_LAUNCH("smx_maxabs_pair");
  return 0;
}
