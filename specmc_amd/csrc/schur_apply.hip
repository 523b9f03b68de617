// schur_apply.hip -- the coupled Schur operator apply of the BFBT route
// (FreeSchurOperator.matvec, lagged_stokes.py:202-220) in three launches:
//
//   1. z-synthesis   yz[b][k] = W_k v[b][k]           per-mode GEMM, K = Nz, over
//                                                    the active runs (listed)
//      W_k = [Vp + 2k^2 Vp G_k; 2 Vp D2 G_k - Vp; k Vp D G_k] (3Q x Nz): the
//      Green apply w = G_k v and the synthesis on the Q padded Chebyshev
//      nodes of the three combinations of v, 2k^2 w, D2 w and k D w that the
//      operator uses (the first three share one x-parity, and the x-product
//      is linear, so they are summed before it)
//   2. x-products    tz = Ae/Ao (eta . (Ce/So yz))     smx_xprod (mirror half
//      of the padded x grid, all three fields)
//   3. combine       out[b][k] = D2Tp u0 + Tp k^2 u1 - 4k DTp u2
//                                                    matrix-core GEMM, K = 3Q
//      with the tau rows out[Nz-2] = V0 . v, out[Nz-1] = V1 . v
//      (lagged_stokes.py:212-220; Tp truncates a degree < 2Nz-1 product on the
//      Q nodes to its first Nz Chebyshev coefficients)
//
// The intermediates yz and tz ([B][Nk][4][Q] each) live in a workspace: the
// caller's (smx_schur_apply) or the library's own device array
// (smx_schur_apply_static).
#include <mutex>

#include "smx_common.h"
#include "specmc_hip.h"

// combine: rows i = b*Nk + k (run-major), columns j < Nz, K = 3Q; A(i, kk) is
// formed while loading from the three product fields of row i:
//   kk <  Q : u0               (against (D2 Tp)^T)
//   kk < 2Q : k^2 u1           (against Tp^T)
//   kk < 3Q : -4k u2           (against (D Tp)^T)
// 32 x 32 tiles, four wavefronts each owning one 16 x 16 matrix-core block,
// K chunks of 16 double-buffered through registers.
#define SC_TM 32
#define SC_TN 32
#define SC_KC 16
template <typename T>
__global__ void __launch_bounds__(256) k_schur_combine(int Nk, int N, int Q, int M,
                                                       const T *tz, const T *Mc, const T *kx,
                                                       const T *v, const T *V0, const T *V1,
                                                       T *out, const int *active) {
  typedef typename MfmaAcc<T>::V VT;
  __shared__ T As[SC_KC][SC_TM + 1];
  __shared__ T Bs[SC_KC][SC_TN + 1];
  __shared__ T tau[SC_TM][2];
  int ti, tj;
  xcd_tile(ti, tj);
  const int i0 = ti * SC_TM, j0 = tj * SC_TN;
  if (active) {
    const int r0 = i0 / Nk, r1 = min(i0 + SC_TM - 1, M - 1) / Nk;
    bool any = false;
    for (int r = r0; r <= r1; ++r) any |= (active[r] != 0);
    if (!any) return;
  }
  const int K = 3 * Q, F = 3 * Q;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * 16, wc = (w & 1) * 16;
  const int lr = lane & 15, lk = lane >> 4;
  // tau rows of this tile's rows, when the tile holds columns Nz-2 or Nz-1
  if (j0 + SC_TN > N - 2 && tid < 2 * SC_TM) {
    const int ii = tid >> 1, which = tid & 1, i = i0 + ii;
    T s = T(0);
    if (i < M) {
      const T *vr = v + (long)i * N;
      const T *Vt = which ? V1 : V0;
      for (int n = 0; n < N; ++n) s = fma(vr[n], Vt[n], s);
    }
    tau[ii][which] = s;
  }
  constexpr int NA = SC_KC * SC_TM / 256, NB = SC_KC * SC_TN / 256;
  T ra[NA], rb[NB];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = u * 256 + tid, ii = e / SC_KC, kk = e % SC_KC;
      const int i = i0 + ii, c = k0 + kk;
      T a = T(0);
      if (i < M && c < K) {
        const T *tr = tz + (long)i * F;
        const int seg = c / Q;
        const T kv = kx[i % Nk];
        a = tr[c];
        if (seg == 1) a = kv * kv * a;
        else if (seg == 2) a = T(-4) * kv * a;
      }
      ra[u] = a;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int e = u * 256 + tid, kb = e / SC_TN, jj = e % SC_TN;
      const int j = j0 + jj, c = k0 + kb;
      rb[u] = (j < N && c < K) ? Mc[(long)c * N + j] : T(0);
    }
  };
  VT acc = VT{T(0), T(0), T(0), T(0)};
  load_chunk(0);
  for (int k0 = 0; k0 < K; k0 += SC_KC) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = u * 256 + tid;
      As[e % SC_KC][e / SC_KC] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int e = u * 256 + tid;
      Bs[e / SC_TN][e % SC_TN] = rb[u];
    }
    __syncthreads();
    if (k0 + SC_KC < K) load_chunk(k0 + SC_KC);
#pragma unroll
    for (int ks = 0; ks < SC_KC; ks += 4)
      acc = MfmaAcc<T>::step(As[ks + lk][wr + lr], Bs[ks + lk][wc + lr], acc);
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ii = wr + MfmaAcc<T>::row(lk, r), i = i0 + ii;
    const int j = j0 + wc + lr;
    if (i >= M || j >= N) continue;
    if (active && !active[i / Nk]) continue;
    out[(long)i * N + j] = j < N - 2 ? acc[r] : tau[ii][j - (N - 2)];
  }
}

// the active runs as a list (ascending) and its length, one wavefront
__global__ void __launch_bounds__(64) k_active_list(int B, const int *active, int *list,
                                                    int *cnt) {
  const int lane = threadIdx.x;
  int base = 0;
  for (int b0 = 0; b0 < B; b0 += 64) {
    const int b = b0 + lane;
    const bool on = b < B && active[b] != 0;
    const unsigned long long m = __ballot(on);
    if (on) list[base + __popcll(m & ((1ull << lane) - 1ull))] = b;
    base += __popcll(m);
  }
  if (lane == 0) *cnt = base;
}

int gemm_rows(int role, int prec, int M, int N, int K, double alpha, double beta,
              const void *A, long ars, long acs, long sa1, const void *B, long brs, long bcs,
              long sb1, void *C, long crs, long ccs, long sc1, int nz1, const int *rowlist,
              const int *rowcnt, hipStream_t st);
// z-synthesis over the listed active runs: 1 (default), or over all runs: 0
static int g_schur_compact = 1;
extern "C" int smx_set_schur_compact(int on) {
  g_schur_compact = on ? 1 : 0;
  return 0;
}

// yz and tz, then the active-run list and its length (B + 1 ints), in
// elements of at least 4 bytes; proportional to B, so that the capacity of a
// fixed workspace in runs is its size over the work of one run
extern "C" long long smx_schur_apply_work(int B, int Nk, int Q) {
  return (2LL * Nk * 3 * Q + 2) * B;
}

static int schur_apply_run(int prec, int B, int Nk, int N, int nrow, int Q, const void *v,
                           const void *eta, long eta_bs, void *out, const void *Wz,
                           const void *Mc, const void *kx, const void *So, const void *Ce,
                           const void *Ao, const void *Ae, const void *V0, const void *V1,
                           void *work, const int *active, void *stream) {
  const long F = 3L * Q;
  const size_t es = prec == 0 ? 8 : 4;
  char *yz = (char *)work;
  char *tz = yz + (size_t)B * Nk * F * es;
  // 1. yz[b][k][:] = W_k v[b][k]: M = B runs, N = 4Q, K = Nz, batch over k;
  //    A(b, n) = v[b][k][n], B(n, c) = Wz[k][n][c], C(b, c) = yz[b][k][c].
  //    With a mask, the rows are the active runs, compacted: a tile of 64 runs
  //    would otherwise be computed whole for one active run in it
  int rc;
  if (active && g_schur_compact) {
    int *list = (int *)(tz + (size_t)B * Nk * F * es), *cnt = list + B;
    hipLaunchKernelGGL(k_active_list, dim3(1), dim3(64), 0, STREAM(stream), B, active, list,
                       cnt);
    rc = gemm_rows(2, prec, B, (int)F, N, 1.0, 0.0, v, (long)Nk * N, 1, N, Wz, F, 1,
                   (long)N * F, yz, (long)Nk * F, 1, F, Nk, list, cnt, STREAM(stream));
  } else {
    rc = smx_gemm_ex(2, prec, B, (int)F, N, 1.0, 0.0, v, (long)Nk * N, 1, N, 0, Wz, F, 1,
                     (long)N * F, 0, yz, (long)Nk * F, 1, F, 0, Nk, 1, nullptr, 1, active,
                     active ? 1 : 0, 1, stream);
  }
  if (rc) return rc;
  // 2. the three x-products with eta on the padded grid (factors folded in W)
  rc = smx_xprod(prec, B, 3, 2, Nk, nrow, Q, yz, nullptr, eta, eta_bs, tz, So, Ce, Ao, Ae,
                 active, stream);
  if (rc) return rc;
  // 3. truncation, derivatives, combination and tau rows
  const int M = B * Nk;
  const dim3 grid(cdiv(N, SC_TN), cdiv(M, SC_TM));
  if (prec == 0)
    hipLaunchKernelGGL(k_schur_combine<double>, grid, dim3(256), 0, STREAM(stream), Nk, N, Q,
                       M, (const double *)tz, (const double *)Mc, (const double *)kx,
                       (const double *)v, (const double *)V0, (const double *)V1,
                       (double *)out, active);
  else
    hipLaunchKernelGGL(k_schur_combine<float>, grid, dim3(256), 0, STREAM(stream), Nk, N, Q,
                       M, (const float *)tz, (const float *)Mc, (const float *)kx,
                       (const float *)v, (const float *)V0, (const float *)V1, (float *)out,
                       active);
  CHECK_LAUNCH("smx_schur_apply");
  return 0;
}

static int schur_apply_check(int prec, int B, int Nk, int N, int nrow, int Q) {
  if (prec != 0 && prec != 1) return fail_msg("smx_schur_apply: prec must be 0 or 1");
  if (Nk <= 0 || N < 3 || nrow <= 0 || Q < N)
    return fail_msg("smx_schur_apply: bad sizes (Nk >= 1, Nz >= 3, Q >= Nz, nrow >= 1)");
  if ((long)cdiv((long)B * Nk, SC_TM) > 65535)
    return fail_msg("smx_schur_apply: B * Nk above 65535 * 32 rows");
  return 0;
}

extern "C" int smx_schur_apply(int prec, int B, int Nk, int N, int nrow, int Q, const void *v,
                               const void *eta, long eta_bs, void *out, const void *Wz,
                               const void *Mc, const void *kx, const void *So, const void *Ce,
                               const void *Ao, const void *Ae, const void *V0, const void *V1,
                               void *work, long long work_elems, const int *active,
                               void *stream) {
  if (B <= 0) return 0;
  if (int rc = schur_apply_check(prec, B, Nk, N, nrow, Q)) return rc;
  if (!work || work_elems < smx_schur_apply_work(B, Nk, Q))
    return fail_msg("smx_schur_apply: workspace smaller than smx_schur_apply_work(B, Nk, Q)");
  return schur_apply_run(prec, B, Nk, N, nrow, Q, v, eta, eta_bs, out, Wz, Mc, kx, So, Ce, Ao,
                         Ae, V0, V1, work, active, stream);
}

// ---------------------------------------------------------------------------
// smx_schur_apply_static: the same launches on the library's own workspace, a
// file-scope device array (one copy per device, held until the process
// ends), for callers that cannot allocate.  In host code the array's name is
// only a stand-in for its symbol; the device address comes from
// hipGetSymbolAddress, taken once per device on the first call there (with
// that device current) and kept.  Every call on one device shares the array:
// calls that overlap on one device race.  Nothing in it has to start at zero:
// every entry a launch reads for an active run was written earlier in the
// same call.
// ---------------------------------------------------------------------------
#define SCHUR_STATIC_BYTES (64ull << 20)
#define SCHUR_STATIC_MAXDEV 64
static __device__ __align__(256) unsigned char scratch[SCHUR_STATIC_BYTES];
static void *g_scratch_ptr[SCHUR_STATIC_MAXDEV];
static std::mutex g_scratch_mu;

static int scratch_ptr(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail("smx_schur_apply_static: hipGetDevice", e);
  if (dev < 0 || dev >= SCHUR_STATIC_MAXDEV)
    return fail_msg("smx_schur_apply_static: device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return fail("smx_schur_apply_static: hipGetSymbolAddress", e);
    if (!p) return fail_msg("smx_schur_apply_static: no device address for the workspace");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" long long smx_schur_apply_static_bytes(void) { return (long long)SCHUR_STATIC_BYTES; }

extern "C" int smx_schur_apply_static(int prec, int B, int Nk, int N, int nrow, int Q,
                                      const void *v, const void *eta, long eta_bs, void *out,
                                      const void *Wz, const void *Mc, const void *kx,
                                      const void *So, const void *Ce, const void *Ao,
                                      const void *Ae, const void *V0, const void *V1,
                                      const int *active, void *stream) {
  if (B <= 0) return 0;
  if (int rc = schur_apply_check(prec, B, Nk, N, nrow, Q)) return rc;
  const unsigned long long need =
      (unsigned long long)smx_schur_apply_work(B, Nk, Q) * (prec == 0 ? 8u : 4u);
  if (need > SCHUR_STATIC_BYTES) {
    snprintf(g_err, sizeof(g_err),
             "smx_schur_apply_static: B=%d runs of Nk=%d, Q=%d need %llu bytes of workspace, "
             "above the library's %llu; split the batch or pass a workspace to "
             "smx_schur_apply",
             B, Nk, Q, need, (unsigned long long)SCHUR_STATIC_BYTES);
    return -1;
  }
  void *work = nullptr;
  if (int rc = scratch_ptr(&work)) return rc;
  return schur_apply_run(prec, B, Nk, N, nrow, Q, v, eta, eta_bs, out, Wz, Mc, kx, So, Ce, Ao,
                         Ae, V0, V1, work, active, stream);
}
