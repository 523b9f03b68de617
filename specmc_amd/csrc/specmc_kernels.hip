// specmc_kernels.hip -- gfx950 kernels of SpecMC's time-step hot path.
//
// Every field is real (mirror-symmetric box): theta-type fields carry real
// x-mode coefficients, stream-function-type fields purely imaginary ones
// stored by their imaginary part.  All arrays are run-major; one launch
// serves a bucket of B runs of one (Nx, Nz) shape.  The C ABI is declared in
// include/specmc_hip.h.
#include "smx_common.h"

#define SMX_VERSION 1

thread_local char g_err[512] = "";

// ---------------------------------------------------------------------------
// general strided batched GEMM (64x64 tile, K step 16, 256 threads, 4x4/thread)
// ---------------------------------------------------------------------------
struct GemmArgs {
  int M, N, K;
  double alpha, beta;
  const void *A; long ars, acs, sa1, sa2;
  const void *B; long brs, bcs, sb1, sb2;
  void *C; long crs, ccs, sc1, sc2;
  int nz2;
  const void *rscale; int rperiod;
  const int *active; int mask_mode, rows_per_run;
  // compacted rows: when rowlist is set, row i < *rowcnt of the product is
  // row rowlist[i] of A and C (the active runs, listed on the device), and
  // the mask is not consulted
  const int *rowlist, *rowcnt;
};

// R: role of the product, part of the kernel symbol only (so kernel traces
// separate them): 0 z-transform / z-derivative, 1 x-transform, 2 per-mode
// operators (Green applies, block builds), 3 dense-matrix work of the rescue
// (assembly, LU update)
template <typename T, int TM, int TN, int R>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  constexpr int PM = TM / 16, PN = TN / 16;
  __shared__ T As[16][TM + 1];
  __shared__ T Bs[16][TN + 1];
  const int z = blockIdx.z, z1 = z / g.nz2, z2 = z % g.nz2;
  int ti, tj;
  xcd_tile(ti, tj);
  const int i0 = ti * TM, j0 = tj * TN;
  const int Mv = g.rowlist ? min(g.M, *g.rowcnt) : g.M;
  if (i0 >= Mv) return;
  if (g.active && !g.rowlist) {
    if (g.mask_mode == 2 && !g.active[z1]) return;
    if (g.mask_mode == 3 && !g.active[z2]) return;
    if (g.mask_mode == 1) {
      int ilast = min(i0 + TM - 1, g.M - 1);
      int r0 = i0 / g.rows_per_run, r1 = ilast / g.rows_per_run;
      bool any = false;
      for (int r = r0; r <= r1; ++r) any |= (g.active[r] != 0);
      if (!any) return;
    }
  }
  const T *A = (const T *)g.A + z1 * g.sa1 + z2 * g.sa2;
  const T *Bm = (const T *)g.B + z1 * g.sb1 + z2 * g.sb2;
  T *C = (T *)g.C + z1 * g.sc1 + z2 * g.sc2;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  T acc[PM][PN];
#pragma unroll
  for (int a = 0; a < PM; ++a)
#pragma unroll
    for (int b = 0; b < PN; ++b) acc[a][b] = 0;
  const bool a_kfast = (g.acs == 1), b_jfast = (g.bcs == 1);
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    // all loads of the chunk are issued before any LDS store (clamped
    // addresses, zero-selected afterwards), so they are in flight together
    {
      constexpr int NA = 16 * TM / 256, NB = 16 * TN / 256;
      T ra[NA], rb[NB];
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        int e = u * 256 + tid, ii, kk;
        if (a_kfast) { ii = e >> 4; kk = e & 15; } else { kk = e / TM; ii = e % TM; }
        const int i = i0 + ii, k = k0 + kk;
        const bool ok = (i < Mv && k < g.K);
        const long ir = ok && g.rowlist ? g.rowlist[i] : i;
        ra[u] = A[ok ? ir * g.ars + (long)k * g.acs : 0];
        ra[u] = ok ? ra[u] : T(0);
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        int e = u * 256 + tid, jj, kb;
        if (b_jfast) { kb = e / TN; jj = e % TN; } else { jj = e >> 4; kb = e & 15; }
        const int j = j0 + jj, kq = k0 + kb;
        const bool ok = (j < g.N && kq < g.K);
        rb[u] = Bm[ok ? (long)kq * g.brs + (long)j * g.bcs : 0];
        rb[u] = ok ? rb[u] : T(0);
      }
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        int e = u * 256 + tid, ii, kk;
        if (a_kfast) { ii = e >> 4; kk = e & 15; } else { kk = e / TM; ii = e % TM; }
        As[kk][ii] = ra[u];
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        int e = u * 256 + tid, jj, kb;
        if (b_jfast) { kb = e / TN; jj = e % TN; } else { jj = e >> 4; kb = e & 15; }
        Bs[kb][jj] = rb[u];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      T a[PM], b[PN];
#pragma unroll
      for (int q = 0; q < PM; ++q) a[q] = As[kk][ty + 16 * q];
#pragma unroll
      for (int q = 0; q < PN; ++q) b[q] = Bs[kk][tx + 16 * q];
#pragma unroll
      for (int p = 0; p < PM; ++p)
#pragma unroll
        for (int q = 0; q < PN; ++q) acc[p][q] = fma(a[p], b[q], acc[p][q]);
    }
    __syncthreads();
  }
  const T alpha = (T)g.alpha, beta = (T)g.beta;
  const T *rs = (const T *)g.rscale;
#pragma unroll
  for (int p = 0; p < PM; ++p) {
    int i = i0 + ty + 16 * p;
    if (i >= Mv) continue;
    if (g.rowlist) i = g.rowlist[i];
    else if (g.active && g.mask_mode == 1 && !g.active[i / g.rows_per_run]) continue;
    T al = rs ? alpha * rs[i % g.rperiod] : alpha;
#pragma unroll
    for (int q = 0; q < PN; ++q) {
      int j = j0 + tx + 16 * q;
      if (j >= g.N) continue;
      long o = (long)i * g.crs + (long)j * g.ccs;
      T v = al * acc[p][q];
      if (beta != T(0)) v += beta * C[o];
      C[o] = v;
    }
  }
}

// Matrix-core GEMM, same arguments and masking as k_gemm.  Four wavefronts in
// a 2 x 2 arrangement; each owns a (TM/2) x (TN/2) sub-tile made of 16 x 16
// blocks computed by v_mfma_f64_16x16x4f64 (T = double) or
// v_mfma_f32_16x16x4f32 (T = float).  Operand layout of both: lane l supplies
// A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]; accumulator element r of
// lane l is C[l / 16 + 4 r][l % 16] for the FP64 form and
// C[4 (l / 16) + r][l % 16] for the FP32 form (measured on gfx950).
template <typename T, int TM, int TN, int KC, int R>
__global__ void __launch_bounds__(256) k_gemm_mfma(GemmArgs g) {
  constexpr int WM = TM / 2, WN = TN / 2, BM = WM / 16, BN = WN / 16;
  typedef typename MfmaAcc<T>::V VT;
  __shared__ T As[KC][TM + 1];
  __shared__ T Bs[KC][TN + 1];
  const int z = blockIdx.z, z1 = z / g.nz2, z2 = z % g.nz2;
  int ti, tj;
  xcd_tile(ti, tj);
  const int i0 = ti * TM, j0 = tj * TN;
  const int Mv = g.rowlist ? min(g.M, *g.rowcnt) : g.M;
  if (i0 >= Mv) return;
  if (g.active && !g.rowlist) {
    if (g.mask_mode == 2 && !g.active[z1]) return;
    if (g.mask_mode == 3 && !g.active[z2]) return;
    if (g.mask_mode == 1) {
      int ilast = min(i0 + TM - 1, g.M - 1);
      int r0 = i0 / g.rows_per_run, r1 = ilast / g.rows_per_run;
      bool any = false;
      for (int r = r0; r <= r1; ++r) any |= (g.active[r] != 0);
      if (!any) return;
    }
  }
  const T *A = (const T *)g.A + z1 * g.sa1 + z2 * g.sa2;
  const T *Bm = (const T *)g.B + z1 * g.sb1 + z2 * g.sb2;
  T *C = (T *)g.C + z1 * g.sc1 + z2 * g.sc2;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wr = (w >> 1) * WM, wc = (w & 1) * WN;
  const int lr = lane & 15, lk = lane >> 4;
  VT acc[BM][BN];
#pragma unroll
  for (int a = 0; a < BM; ++a)
#pragma unroll
    for (int b = 0; b < BN; ++b) acc[a][b] = VT{T(0), T(0), T(0), T(0)};
  const bool a_kfast = (g.acs == 1), b_jfast = (g.bcs == 1);
  // register double buffering: the global loads of chunk k0 + KC are in
  // flight while the matrix cores work on chunk k0 from LDS
  constexpr int NA = KC * TM / 256, NB = KC * TN / 256;
  T ra[NA], rb[NB];
  // the A row each thread loads is the same in every K chunk: its offset
  // (through the row list, when there is one) is found once; -1 past the rows
  long arow[NA];
#pragma unroll
  for (int u = 0; u < NA; ++u) {
    const int e = u * 256 + tid, i = i0 + (a_kfast ? e / KC : e % TM);
    arow[u] = i < Mv ? (long)(g.rowlist ? g.rowlist[i] : i) * g.ars : -1;
  }
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int e = u * 256 + tid, k = k0 + (a_kfast ? e % KC : e / TM);
      const bool ok = (arow[u] >= 0 && k < g.K);
      ra[u] = A[ok ? arow[u] + (long)k * g.acs : 0];
      ra[u] = ok ? ra[u] : T(0);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      int e = u * 256 + tid, jj, kb;
      if (b_jfast) { kb = e / TN; jj = e % TN; } else { jj = e / KC; kb = e % KC; }
      const int j = j0 + jj, kq = k0 + kb;
      const bool ok = (j < g.N && kq < g.K);
      rb[u] = Bm[ok ? (long)kq * g.brs + (long)j * g.bcs : 0];
      rb[u] = ok ? rb[u] : T(0);
    }
  };
  load_chunk(0);
  for (int k0 = 0; k0 < g.K; k0 += KC) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      int e = u * 256 + tid, ii, kk;
      if (a_kfast) { ii = e / KC; kk = e % KC; } else { kk = e / TM; ii = e % TM; }
      As[kk][ii] = ra[u];
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      int e = u * 256 + tid, jj, kb;
      if (b_jfast) { kb = e / TN; jj = e % TN; } else { jj = e / KC; kb = e % KC; }
      Bs[kb][jj] = rb[u];
    }
    __syncthreads();
    if (k0 + KC < g.K) load_chunk(k0 + KC);
#pragma unroll
    for (int ks = 0; ks < KC; ks += 4) {
      T a[BM], b[BN];
#pragma unroll
      for (int q = 0; q < BM; ++q) a[q] = As[ks + lk][wr + 16 * q + lr];
#pragma unroll
      for (int q = 0; q < BN; ++q) b[q] = Bs[ks + lk][wc + 16 * q + lr];
#pragma unroll
      for (int p = 0; p < BM; ++p)
#pragma unroll
        for (int q = 0; q < BN; ++q)
          acc[p][q] = MfmaAcc<T>::step(a[p], b[q], acc[p][q]);
    }
    __syncthreads();
  }
  const T alpha = (T)g.alpha, beta = (T)g.beta;
  const T *rs = (const T *)g.rscale;
#pragma unroll
  for (int p = 0; p < BM; ++p) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int i = i0 + wr + 16 * p + MfmaAcc<T>::row(lk, r);
      if (i >= Mv) continue;
      if (g.rowlist) i = g.rowlist[i];
      else if (g.active && g.mask_mode == 1 && !g.active[i / g.rows_per_run]) continue;
      const T al = rs ? alpha * rs[i % g.rperiod] : alpha;
#pragma unroll
      for (int q = 0; q < BN; ++q) {
        const int j = j0 + wc + 16 * q + lr;
        if (j >= g.N) continue;
        const long o = (long)i * g.crs + (long)j * g.ccs;
        T v = al * acc[p][q][r];
        if (beta != T(0)) v += beta * C[o];
        C[o] = v;
      }
    }
  }
}

// FP64 inversion blocked by 16 columns on the matrix cores (default for
// 16 <= N <= 128; gj_mfma.hip)
int gj_invert_mfma(int nmat, int N, double *A, int per_run, const int *active, hipStream_t st);
static int g_gj_mfma = 1;
extern "C" int smx_set_gj_mfma(int on) {
  g_gj_mfma = on ? 1 : 0;
  return 0;
}

// FP64 inversion in registers (for 16 <= N <= 128 when the matrix-core form is off)
static int g_gj_reg = 1;
extern "C" int smx_set_gj_reg(int on) {
  g_gj_reg = on ? 1 : 0;
  return 0;
}

// Gauss-Jordan variant: 1 row-owner kernel, 0 the first kernel (default: at
// 96 x 96 the row-owner kernel measured 40 ms for 8448 inverses against 11.7)
static int g_gj_rows = 0;
extern "C" int smx_set_gj_rows(int on) {
  g_gj_rows = on ? 1 : 0;
  return 0;
}

// inverse-apply variant for FP64 and even N: 1 the row-pair kernel with the
// whole block's loads issued before the right-hand side is staged (default),
// 2 the row-pair kernel loading as it accumulates, 0 the first kernel
static int g_inv_pairs = 1;
extern "C" int smx_set_inv_pairs(int on) {
  g_inv_pairs = (on == 1 || on == 2) ? on : 0;
  return 0;
}

// k_xprod variant for FP64: 1 matrix cores (default), 0 the vector kernel
static int g_xprod_mfma = 1;
extern "C" int smx_set_xprod_mfma(int on) {
  g_xprod_mfma = on ? 1 : 0;
  return 0;
}

// GEMM kernel family: 0 vector FMA, 1 matrix cores with 16-wide K chunks
// (default), 2 matrix cores with single-chunk staging of small K
static int g_gemm_mfma = 1;

// Tile choice: the largest tile that still gives every CU several
// workgroups (the batched shapes of one step are small: K <= 3Nz/2 and
// M*N per problem of a few 10^5), else the smallest tile.
struct GemmTile { int tm, tn; };
static const GemmTile kGemmTiles[] = {{128, 96}, {128, 64}, {64, 96}, {64, 64},
                                      {64, 32}, {32, 64}, {32, 32}};
static const int kGemmWantWG = 1024;

// Explicit kernel choices per problem shape, set by the autotuner
// (smx_gemm_tune).  Codes: 0..6 vector FMA with tile kGemmTiles[code];
// 10 matrix cores 64 x 64, K chunks of 16; 11 matrix cores 32 x 32, K chunks
// of 16; 12 matrix cores 32 x 32, the whole K (<= 160) staged in one chunk.
struct GemmTune { int prec, M, N, K, nz, akf, bjf, code; };
#define GEMM_TUNE_MAX 256
static GemmTune g_tune[GEMM_TUNE_MAX];
static int g_ntune = 0;

static int tune_lookup(int prec, const GemmArgs &g, int nz) {
  const int akf = g.acs == 1, bjf = g.bcs == 1;
  for (int i = 0; i < g_ntune; ++i) {
    const GemmTune &t = g_tune[i];
    if (t.prec == prec && t.M == g.M && t.N == g.N && t.K == g.K && t.nz == nz &&
        t.akf == akf && t.bjf == bjf)
      return t.code;
  }
  return -1;
}

static bool gemm_code_ok(int code, int K) {
  return (code >= 0 && code <= 6) || code == 10 || code == 11 || (code == 12 && K <= 160);
}

template <typename T, int R>
static void launch_gemm_code(const GemmArgs &g, int nz, hipStream_t st, int code) {
  const int M = g.M, N = g.N;
  if (code == 10) {
    hipLaunchKernelGGL((k_gemm_mfma<T, 64, 64, 16, R>), dim3(cdiv(N, 64), cdiv(M, 64), nz),
                       dim3(256), 0, st, g);
    return;
  }
  if (code == 11 || code == 12) {
    const dim3 grid(cdiv(N, 32), cdiv(M, 32), nz);
#define GK(KC_) hipLaunchKernelGGL((k_gemm_mfma<T, 32, 32, KC_, R>), grid, dim3(256), 0, st, g)
    const int K = code == 11 ? 1 << 30 : g.K;
    if (K <= 32) { GK(32); return; }
    if (K <= 48) { GK(48); return; }
    if (K <= 64) { GK(64); return; }
    if (K <= 96) { GK(96); return; }
    if (K <= 128) { GK(128); return; }
    if (K <= 160) { GK(160); return; }
#undef GK
    hipLaunchKernelGGL((k_gemm_mfma<T, 32, 32, 16, R>), grid, dim3(256), 0, st, g);
    return;
  }
#define GL(TM_, TN_)                                                                      \
  hipLaunchKernelGGL((k_gemm<T, TM_, TN_, R>), dim3(cdiv(N, TN_), cdiv(M, TM_), nz),       \
                     dim3(256), 0, st, g)
  switch (code) {
    case 0: GL(128, 96); break;
    case 1: GL(128, 64); break;
    case 2: GL(64, 96); break;
    case 3: GL(64, 64); break;
    case 4: GL(64, 32); break;
    case 5: GL(32, 64); break;
    default: GL(32, 32); break;
  }
#undef GL
}

// the default choice when the autotuner has not set one for this shape
static int gemm_default_code(const GemmArgs &g, int nz) {
  const int M = g.M, N = g.N;
  if (g_gemm_mfma > 0) {
    // 64 x 64 tiles when they give every CU work and pad little, else 32 x 32
    const long wg64 = (long)cdiv(M, 64) * cdiv(N, 64) * nz;
    const long pad64 = (long)cdiv(M, 64) * 64 * cdiv(N, 64) * 64;
    if (wg64 >= 512 && pad64 * 4 <= (long)M * N * 5) return 10;
    // small problems: 32 x 32 tiles, with the whole K extent staged in one
    // chunk (family 2: one memory latency per launch instead of one per
    // 16-wide chunk)
    return (g_gemm_mfma == 2 && g.K <= 160) ? 12 : 11;
  }
  const int nt = (int)(sizeof(kGemmTiles) / sizeof(kGemmTiles[0]));
  // the largest tile with enough workgroups and at most 25% padding; else
  // the tile with the most workgroups
  for (int t = 0; t < nt; ++t) {
    const int tm = kGemmTiles[t].tm, tn = kGemmTiles[t].tn;
    const long wg = (long)cdiv(M, tm) * cdiv(N, tn) * nz;
    const long padded = (long)cdiv(M, tm) * tm * cdiv(N, tn) * tn;
    if (wg >= kGemmWantWG && padded * 4 <= (long)M * N * 5) return t;
  }
  return nt - 1;
}

template <typename T, int R>
static void launch_gemm(const GemmArgs &g, int nz, hipStream_t st) {
  int code = tune_lookup(sizeof(T) == 8 ? 0 : 1, g, nz);
  if (code < 0) code = gemm_default_code(g, nz);
  launch_gemm_code<T, R>(g, nz, st, code);
}

// ---------------------------------------------------------------------------
// launch clock: HIP events recorded on the launch stream immediately around
// every launch of the chosen entry points (tag 1: smx_fg_post, 2:
// smx_inv_apply and smx_lu_solve, 3: smx_dense_lu_solve), for measuring a
// kernel's average duration inside a timed region without host
// synchronisation.  Several tags can run at once, each with its own events
// and its own device counter of the algorithmic units its launches process
// (fg_post: vectors of length n moved; inv_apply: systems solved;
// dense_lu_solve: matrices solved).
// ---------------------------------------------------------------------------
#define CLK_TAGS 4
struct ClkTag {
  int cap = 0, n = 0;
  hipEvent_t *ev = nullptr;
};
static ClkTag g_clk[CLK_TAGS];
static unsigned long long *g_clk_units = nullptr;   // [CLK_TAGS], device

// the first failure of an event record, reported by smx_clock_stop
static hipError_t g_clk_err = hipSuccess;

bool clk_on(int tag) {
  return tag > 0 && tag < CLK_TAGS && g_clk[tag].n < g_clk[tag].cap;
}
static inline void clk_rec(hipEvent_t ev, hipStream_t st) {
  const hipError_t e = hipEventRecord(ev, st);
  if (e != hipSuccess && g_clk_err == hipSuccess) g_clk_err = e;
}
void clk_begin(int tag, hipStream_t st) { clk_rec(g_clk[tag].ev[2 * g_clk[tag].n], st); }
void clk_end(int tag, hipStream_t st) {
  clk_rec(g_clk[tag].ev[2 * g_clk[tag].n + 1], st);
  ++g_clk[tag].n;
}

unsigned long long *clk_units(int tag) {
  return clk_on(tag) ? g_clk_units + tag : nullptr;
}

static void clk_free_tag(int tag) {
  // teardown of the clock's own events: nothing to report
  ClkTag &c = g_clk[tag];
  for (int i = 0; i < 2 * c.cap; ++i) (void)hipEventDestroy(c.ev[i]);
  free(c.ev);
  c = ClkTag();
}

static void clk_free_all() {
  for (int t = 1; t < CLK_TAGS; ++t) clk_free_tag(t);
  if (g_clk_units) (void)hipFree(g_clk_units);
  g_clk_units = nullptr;
}

extern "C" int smx_clock_start(int tags, int capacity) {
  clk_free_all();
  g_clk_err = hipSuccess;
  if (tags <= 0 || capacity <= 0) return 0;
  hipError_t e = hipMalloc((void **)&g_clk_units, CLK_TAGS * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(g_clk_units, 0, CLK_TAGS * sizeof(unsigned long long));
  if (e != hipSuccess) {
    clk_free_all();
    return fail("smx_clock_start", e);
  }
  for (int t = 1; t < CLK_TAGS; ++t) {
    if (!(tags & (1 << t))) continue;
    ClkTag &c = g_clk[t];
    c.ev = (hipEvent_t *)calloc((size_t)2 * capacity, sizeof(hipEvent_t));
    if (!c.ev) {
      clk_free_all();
      return fail_msg("smx_clock_start: out of host memory");
    }
    for (int i = 0; i < 2 * capacity; ++i) {
      e = hipEventCreate(&c.ev[i]);
      if (e != hipSuccess) {
        c.cap = i / 2;
        clk_free_all();
        return fail("smx_clock_start", e);
      }
    }
    c.cap = capacity;
  }
  return 0;
}

extern "C" int smx_clock_stop(int tag, double *total_ms, int *count, long long *units) {
  if (g_clk_err != hipSuccess) {
    const hipError_t e = g_clk_err;
    g_clk_err = hipSuccess;
    clk_free_all();
    return fail("smx_clock_stop (an event record failed)", e);
  }
  if (tag <= 0 || tag >= CLK_TAGS) return fail_msg("smx_clock_stop: tag must be 1, 2 or 3");
  ClkTag &c = g_clk[tag];
  double tot = 0.0;
  if (units) {
    *units = 0;
    if (g_clk_units) {
      unsigned long long u = 0;
      hipError_t e = hipMemcpy(&u, g_clk_units + tag, sizeof(u), hipMemcpyDeviceToHost);
      if (e != hipSuccess) return fail("smx_clock_stop", e);
      *units = (long long)u;
    }
  }
  for (int i = 0; i < c.n; ++i) {
    hipError_t e = hipEventSynchronize(c.ev[2 * i + 1]);
    if (e != hipSuccess) return fail("smx_clock_stop", e);
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, c.ev[2 * i], c.ev[2 * i + 1]);
    if (e != hipSuccess) return fail("smx_clock_stop", e);
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (count) *count = c.n;
  clk_free_tag(tag);
  bool any = false;
  for (int t = 1; t < CLK_TAGS; ++t) any |= g_clk[t].cap > 0;
  if (!any) clk_free_all();
  return 0;
}

extern "C" int smx_version(void) { return SMX_VERSION; }
extern "C" int smx_set_gemm_mfma(int mode) {
  if (mode < 0 || mode > 2) return fail_msg("smx_set_gemm_mfma: mode must be 0, 1 or 2");
  g_gemm_mfma = mode;
  return 0;
}
extern "C" const char *smx_last_error(void) { return g_err; }

template <typename T>
static void launch_gemm_role(int role, const GemmArgs &g, int nz, hipStream_t st) {
  switch (role) {
    case 1: launch_gemm<T, 1>(g, nz, st); break;
    case 2: launch_gemm<T, 2>(g, nz, st); break;
    case 3: launch_gemm<T, 3>(g, nz, st); break;
    default: launch_gemm<T, 0>(g, nz, st); break;
  }
}

extern "C" int smx_gemm_ex(int role, int prec, int M, int N, int K, double alpha, double beta,
                           const void *A, long ars, long acs, long sa1, long sa2,
                           const void *B, long brs, long bcs, long sb1, long sb2,
                           void *C, long crs, long ccs, long sc1, long sc2,
                           int nz1, int nz2, const void *rscale, int rperiod,
                           const int *active, int mask_mode, int rows_per_run,
                           void *stream) {
  if (M <= 0 || N <= 0 || nz1 <= 0 || nz2 <= 0) return 0;
  if (role < 0 || role > 3) return fail_msg("smx_gemm_ex: role must be 0..3");
  if (K < 0) return fail_msg("smx_gemm: negative K");
  if (rscale && rperiod <= 0) return fail_msg("smx_gemm: rperiod must be > 0");
  if (mask_mode == 1 && rows_per_run <= 0) return fail_msg("smx_gemm: rows_per_run");
  if ((long)nz1 * nz2 > 65535) return fail_msg("smx_gemm: batch > 65535");
  GemmArgs g{M, N, K, alpha, beta, A, ars, acs, sa1, sa2, B, brs, bcs, sb1, sb2,
             C, crs, ccs, sc1, sc2, nz2, rscale, rperiod, active, mask_mode,
             rows_per_run};
  if (prec == 0) launch_gemm_role<double>(role, g, nz1 * nz2, STREAM(stream));
  else launch_gemm_role<float>(role, g, nz1 * nz2, STREAM(stream));
  CHECK_LAUNCH("smx_gemm");
  return 0;
}

// smx_gemm_ex over a compacted list of rows (rowlist / rowcnt on the device,
// at most M rows); used by the Schur apply for the active runs of a batch
int gemm_rows(int role, int prec, int M, int N, int K, double alpha, double beta,
              const void *A, long ars, long acs, long sa1, const void *B, long brs, long bcs,
              long sb1, void *C, long crs, long ccs, long sc1, int nz1, const int *rowlist,
              const int *rowcnt, hipStream_t st) {
  if (M <= 0 || N <= 0 || nz1 <= 0) return 0;
  if (nz1 > 65535) return fail_msg("smx_gemm: batch > 65535");
  GemmArgs g{M, N, K, alpha, beta, A, ars, acs, sa1, 0, B, brs, bcs, sb1, 0,
             C, crs, ccs, sc1, 0, 1, nullptr, 1, nullptr, 0, 1, rowlist, rowcnt};
  if (prec == 0) launch_gemm_role<double>(role, g, nz1, st);
  else launch_gemm_role<float>(role, g, nz1, st);
  CHECK_LAUNCH("smx_gemm");
  return 0;
}

extern "C" int smx_gemm(int prec, int M, int N, int K, double alpha, double beta,
                        const void *A, long ars, long acs, long sa1, long sa2,
                        const void *B, long brs, long bcs, long sb1, long sb2,
                        void *C, long crs, long ccs, long sc1, long sc2,
                        int nz1, int nz2, const void *rscale, int rperiod,
                        const int *active, int mask_mode, int rows_per_run,
                        void *stream) {
  return smx_gemm_ex(0, prec, M, N, K, alpha, beta, A, ars, acs, sa1, sa2, B, brs, bcs, sb1,
                     sb2, C, crs, ccs, sc1, sc2, nz1, nz2, rscale, rperiod, active, mask_mode,
                     rows_per_run, stream);
}

extern "C" int smx_gemm_tune(int prec, int M, int N, int K, int nz, int a_kfast, int b_jfast,
                             int code) {
  if (prec != 0 && prec != 1) return fail_msg("smx_gemm_tune: prec must be 0 or 1");
  if (code >= 0 && !gemm_code_ok(code, K)) return fail_msg("smx_gemm_tune: bad kernel code");
  const int akf = a_kfast != 0, bjf = b_jfast != 0;
  for (int i = 0; i < g_ntune; ++i) {
    GemmTune &t = g_tune[i];
    if (t.prec == prec && t.M == M && t.N == N && t.K == K && t.nz == nz && t.akf == akf &&
        t.bjf == bjf) {
      if (code >= 0) {
        t.code = code;
      } else {
        t = g_tune[--g_ntune];
      }
      return 0;
    }
  }
  if (code < 0) return 0;
  if (g_ntune >= GEMM_TUNE_MAX) return fail_msg("smx_gemm_tune: table full");
  g_tune[g_ntune++] = GemmTune{prec, M, N, K, nz, akf, bjf, code};
  return 0;
}

extern "C" int smx_gemm_tune_clear(void) {
  g_ntune = 0;
  return 0;
}

// ---------------------------------------------------------------------------
// batched LU (column-major, partial pivoting, LAPACK getrf semantics)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_lu_factor(T *LU, int *piv, int N, int per_run,
                                                   const int *active) {
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  T *A = LU + (long)m * N * N;
  int *p = piv + (long)m * N;
  __shared__ T rv[8];
  __shared__ int ri[8];
  __shared__ int s_piv;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int c = 0; c < N; ++c) {
    // pivot: first index of max |a| over rows c..N-1 (idamax)
    T best = T(-1);
    int bi = N;
    for (int r = c + tid; r < N; r += nt) {
      T v = fabs(A[(long)c * N + r]);
      if (v > best) { best = v; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      T ov = __shfl_xor(best, o, 64);
      int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      T b = rv[0];
      int i = ri[0];
      for (int w = 1; w < (nt >> 6); ++w)
        if (rv[w] > b || (rv[w] == b && ri[w] < i)) { b = rv[w]; i = ri[w]; }
      if (i >= N) i = c;
      s_piv = i;
      p[c] = i;
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != c)
      for (int col = tid; col < N; col += nt) {
        T t = A[(long)col * N + c];
        A[(long)col * N + c] = A[(long)col * N + pr];
        A[(long)col * N + pr] = t;
      }
    __syncthreads();
    const T d = A[(long)c * N + c];
    if (d != T(0)) {
      const T inv = T(1) / d;
      for (int r = c + 1 + tid; r < N; r += nt) A[(long)c * N + r] *= inv;
    }
    __syncthreads();
    const int w = N - c - 1;
    for (int e = tid; e < w * w; e += nt) {
      int q = c + 1 + e / w, r = c + 1 + e % w;
      A[(long)q * N + r] -= A[(long)c * N + r] * A[(long)q * N + c];
    }
    __syncthreads();
  }
}

// LDS-resident variant: the whole matrix is staged in LDS, factorised with the
// same partial-pivoting sweep as k_lu_factor, and written back.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_factor_lds(T *LU, int *piv, int N, int per_run,
                                                       const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *A = (T *)smem_raw;
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  T *G = LU + (long)m * N * N;
  int *p = piv + (long)m * N;
  __shared__ T rv[8];
  __shared__ int ri[8];
  __shared__ int s_piv;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int NN = N * N;
  for (int e = tid; e < NN; e += nt) A[e] = G[e];
  __syncthreads();
  for (int c = 0; c < N; ++c) {
    T best = T(-1);
    int bi = N;
    for (int r = c + tid; r < N; r += nt) {
      T v = fabs(A[c * N + r]);
      if (v > best) { best = v; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      T ov = __shfl_xor(best, o, 64);
      int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      T b = rv[0];
      int i = ri[0];
      for (int w = 1; w < (nt >> 6); ++w)
        if (rv[w] > b || (rv[w] == b && ri[w] < i)) { b = rv[w]; i = ri[w]; }
      if (i >= N) i = c;
      s_piv = i;
      p[c] = i;
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != c)
      for (int col = tid; col < N; col += nt) {
        T t = A[col * N + c];
        A[col * N + c] = A[col * N + pr];
        A[col * N + pr] = t;
      }
    __syncthreads();
    const T d = A[c * N + c];
    if (d != T(0)) {
      const T inv = T(1) / d;
      for (int r = c + 1 + tid; r < N; r += nt) A[c * N + r] *= inv;
    }
    __syncthreads();
    const int w = N - c - 1;
    for (int e = tid; e < w * w; e += nt) {
      int q = c + 1 + e / w, r = c + 1 + e % w;
      A[q * N + r] -= A[c * N + r] * A[q * N + c];
    }
    __syncthreads();
  }
  for (int e = tid; e < NN; e += nt) G[e] = A[e];
}

// LDS-resident solve: one block per system; same row interchanges and
// column-oriented sweeps as k_lu_solve.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_solve_lds(const T *LU, const int *piv, const T *rhs,
                                                      T *x, int N, int tau_zero, T scale,
                                                      const T *kscale, const T *bscale,
                                                      int per_run, int shared,
                                                      const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *A = (T *)smem_raw;
  T *y = A + N * N;
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  const long mi = shared ? (m % per_run) : m;
  const T *G = LU + mi * N * N;
  const int tid = threadIdx.x, nt = blockDim.x, NN = N * N;
  for (int e = tid; e < NN; e += nt) A[e] = G[e];
  T s = kscale ? scale * kscale[m % per_run] : scale;
  if (bscale) s = s * bscale[m / per_run];
  for (int r = tid; r < N; r += nt) {
    T v = rhs[(long)m * N + r] * s;
    if (tau_zero && r >= N - 2) v = T(0);
    y[r] = v;
  }
  __syncthreads();
  if (tid == 0) {
    const int *p = piv + mi * N;
    for (int c = 0; c < N; ++c) {
      int pc = p[c];
      if (pc != c) { T t = y[c]; y[c] = y[pc]; y[pc] = t; }
    }
  }
  __syncthreads();
  for (int c = 0; c < N - 1; ++c) {
    const T yc = y[c];
    for (int r = c + 1 + tid; r < N; r += nt) y[r] -= A[c * N + r] * yc;
    __syncthreads();
  }
  for (int c = N - 1; c >= 0; --c) {
    if (tid == 0) y[c] = y[c] / A[c * N + c];
    __syncthreads();
    const T xc = y[c];
    for (int r = tid; r < c; r += nt) y[r] -= A[c * N + r] * xc;
    __syncthreads();
  }
  for (int r = tid; r < N; r += nt) x[(long)m * N + r] = y[r];
}

// largest dynamic LDS a block may use (gfx950: 160 KiB per CU)
#define LUS_MAXN 512

// In-place Gauss-Jordan inversion with partial (row) pivoting, LDS-resident,
// column-major N x N matrices (one block per matrix).  The preconditioners of
// the FGMRES loops are applied as inv(A) r (smx_inv_apply) instead of two
// sequential triangular sweeps.
template <typename T>
__global__ void __launch_bounds__(256) k_gj_invert(T *Am, int N, int per_run, const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *A = (T *)smem_raw;                 // [col][row]
  int *sw = (int *)(A + N * N);         // row swap of each step
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  T *G = Am + (long)m * N * N;
  __shared__ T rv[8];
  __shared__ int ri[8];
  __shared__ int s_piv;
  __shared__ T s_inv;
  const int tid = threadIdx.x, nt = blockDim.x, NN = N * N;
  for (int e = tid; e < NN; e += nt) A[e] = G[e];
  __syncthreads();
  for (int c = 0; c < N; ++c) {
    T best = T(-1);
    int bi = N;
    for (int r = c + tid; r < N; r += nt) {
      T v = fabs(A[c * N + r]);
      if (v > best) { best = v; bi = r; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      T ov = __shfl_xor(best, o, 64);
      int oi = __shfl_xor(bi, o, 64);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { rv[tid >> 6] = best; ri[tid >> 6] = bi; }
    __syncthreads();
    if (tid == 0) {
      T b = rv[0];
      int i = ri[0];
      for (int w = 1; w < (nt >> 6); ++w)
        if (rv[w] > b || (rv[w] == b && ri[w] < i)) { b = rv[w]; i = ri[w]; }
      if (i >= N) i = c;
      s_piv = i;
      sw[c] = i;
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != c)
      for (int col = tid; col < N; col += nt) {
        T t = A[col * N + c];
        A[col * N + c] = A[col * N + pr];
        A[col * N + pr] = t;
      }
    __syncthreads();
    if (tid == 0) s_inv = T(1) / A[c * N + c];
    __syncthreads();
    const T inv = s_inv;
    // pivot row: a_cq /= a_cc, the pivot entry becomes 1/a_cc
    for (int q = tid; q < N; q += nt) A[q * N + c] = (q == c) ? inv : A[q * N + c] * inv;
    __syncthreads();
    // eliminate column c from the other rows; column c takes -a_rc / a_cc.
    // Entry e = tid + i nt is (column q, row r); both advance by the fixed
    // (dq, dr) = (nt / N, nt % N) per i, so no division per entry.
    {
      int q = tid / N, r = tid % N;
      const int dq = nt / N, dr = nt % N;
      for (int e = tid; e < NN; e += nt) {
        if (r != c && q != c) A[q * N + r] -= A[c * N + r] * A[q * N + c];
        q += dq;
        r += dr;
        if (r >= N) { r -= N; ++q; }
      }
    }
    __syncthreads();
    for (int r = tid; r < N; r += nt)
      if (r != c) A[c * N + r] = -A[c * N + r] * inv;
    __syncthreads();
  }
  // undo the row swaps as column swaps, last step first
  for (int c = N - 1; c >= 0; --c) {
    const int pr = sw[c];
    if (pr != c)
      for (int r = tid; r < N; r += nt) {
        T t = A[c * N + r];
        A[c * N + r] = A[pr * N + r];
        A[pr * N + r] = t;
      }
    __syncthreads();
  }
  for (int e = tid; e < NN; e += nt) G[e] = A[e];
}

// FP64 Gauss-Jordan inversion with the matrix in registers: 4N threads;
// thread (r = tid % N, g = tid / N) holds row r of the columns q = g + 4 i
// (i < CPT).  Per column step only the pivot column, row c and the scaled
// pivot row pass through LDS.  Same pivot choice and the same arithmetic per entry as
// k_gj_invert: the pivot row is scaled by 1/a_cc first, every other row
// subtracts (its column-c entry) x (the scaled pivot row), column c becomes
// -a_rc / a_cc with 1/a_cc on the diagonal.  The row swaps are undone at the
// end as a permutation of the columns on the way out.
// Two barriers per column: every wavefront finds the pivot itself from the
// pivot column (the same search, so the same row), and the LDS rows are
// double-buffered by the parity of the column, so step c + 2 rewrites a
// buffer only after every thread has passed step c + 1's first barrier,
// that is after it finished reading step c.  The block is rounded up to whole
// wavefronts (the search uses every lane); threads past 4N own no entries.
// FULL: N == 4 CPT, so every column slot of a thread is a column of the
// matrix and the per-entry bounds tests vanish at compile time.
#define GJR_MAXN 128
template <int CPT, bool FULL>
__global__ void __launch_bounds__(512) k_gj_invert_reg(double *Am, int N, int per_run,
                                                       const int *active) {
  __shared__ double colv[2][GJR_MAXN], rowa[2][GJR_MAXN], rows_[2][GJR_MAXN];
  __shared__ int sw[GJR_MAXN], dest[GJR_MAXN];
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  double *Gm = Am + (long)m * N * N;
  const int tid = threadIdx.x, r = tid % N, lane = tid & 63;
  const bool own = tid < 4 * N;
  const int g = own ? tid / N : 0;
  double a[CPT];
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int q = g + 4 * i;
    a[i] = (own && (FULL || q < N)) ? Gm[(long)q * N + r] : 0.0;
  }
  for (int c = 0; c < N; ++c) {
    const int b = c & 1, gc = c & 3, ic = c >> 2;
    double *cv = colv[b], *ra = rowa[b], *rs = rows_[b];
    if (own && g == gc) {
      double v = a[0];
#pragma unroll
      for (int i = 1; i < CPT; ++i) v = (i == ic) ? a[i] : v;
      cv[r] = v;
    }
    __syncthreads();
    // pivot: the first largest |a| at or below the diagonal (every wavefront)
    double best = -1.0;
    int bi = N;
    for (int rr = c + lane; rr < N; rr += 64) {
      const double v = fabs(cv[rr]);
      if (v > best) { best = v; bi = rr; }
    }
    wave_argmax_first(best, bi);
    const int p = bi >= N ? c : bi;
    if (tid == 0) sw[c] = p;
    // a_pc (the pivot after the swap) is cv[p]: every thread forms 1 / a_cc
    const double inv = 1.0 / cv[p];
    if (own && r == c) {
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int q = g + 4 * i;
        if (FULL || q < N) ra[q] = a[i];
      }
    }
    if (own && r == p) {    // also when p == c: rs is the scaled pivot row
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int q = g + 4 * i;
        if (FULL || q < N) rs[q] = a[i] * inv;
      }
    }
    __syncthreads();
    // after the swap, row p holds the old row c (ra); row c becomes the old
    // row p scaled (rs), written in full below
    if (p != c && r == p) {
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int q = g + 4 * i;
        if (FULL || q < N) a[i] = ra[q];
      }
    }
    const double fr = cv[(r == c) ? p : (r == p ? c : r)];
    if (r == c) {
      // the pivot row: scaled, with 1/a_cc in column c
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int q = g + 4 * i;
        if (!FULL && q >= N) continue;
        a[i] = (q == c) ? inv : rs[q];
      }
    } else {
      // every entry takes the elimination; column c (slot ic of the threads
      // with g == gc) is then replaced by -a_rc / a_cc
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const int q = g + 4 * i;
        if (!FULL && q >= N) continue;
        a[i] = a[i] - fr * rs[q];
      }
      if (g == gc) {
        const double dg = -fr * inv;
#pragma unroll
        for (int i = 0; i < CPT; ++i) a[i] = (i == ic) ? dg : a[i];
      }
    }
  }
  __syncthreads();
  // undo the row swaps as column swaps, last step first: position j ends up
  // holding column colat[j]; column q goes to dest[q]
  if (tid == 0) {
    int *colat = dest;        // reused: built in place, then inverted below
    for (int j = 0; j < N; ++j) colat[j] = j;
    for (int c = N - 1; c >= 0; --c) {
      const int s2 = sw[c];
      const int t = colat[c];
      colat[c] = colat[s2];
      colat[s2] = t;
    }
    for (int j = 0; j < N; ++j) sw[colat[j]] = j;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int q = g + 4 * i;
    if (own && (FULL || q < N)) Gm[(long)sw[q] * N + r] = a[i];
  }
}

// Gauss-Jordan inversion, row-owner form: thread (r = tid % N, g = tid / N)
// owns row r of the columns q = g, g + G, ... (G = blockDim / N), so the
// elimination needs no index arithmetic per entry.  Four barriers per
// column: pivot search (wavefront 0), row swap, pivot column / scaled pivot
// row copies, update.  Same pivot choice (first largest |a| at or below the
// diagonal) and the same arithmetic per entry as k_gj_invert.
template <typename T>
__global__ void __launch_bounds__(512) k_gj_invert_rows(T *Am, int N, int per_run,
                                                        const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *A = (T *)smem_raw;                 // [col][row]
  T *fcol = A + N * N;                  // pivot column after the swap
  T *prow = fcol + N;                   // scaled pivot row
  int *sw = (int *)(prow + N);          // row swap of each step
  __shared__ int s_piv;
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  T *Gm = Am + (long)m * N * N;
  const int tid = threadIdx.x, nt = blockDim.x, G = nt / N;
  const int r = tid % N, g = tid / N;   // every thread of the block has g < G
  for (int e = tid; e < N * N; e += nt) A[e] = Gm[e];
  __syncthreads();
  for (int c = 0; c < N; ++c) {
    if (tid < 64) {
      T best = T(-1);
      int bi = N;
      for (int rr = c + tid; rr < N; rr += 64) {
        const T v = fabs(A[c * N + rr]);
        if (v > best) { best = v; bi = rr; }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const T ov = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
      }
      if (tid == 0) {
        if (bi >= N) bi = c;
        s_piv = bi;
        sw[c] = bi;
      }
    }
    __syncthreads();
    const int p = s_piv;
    if (p != c)
      for (int q = tid; q < N; q += nt) {
        const T t = A[q * N + c];
        A[q * N + c] = A[q * N + p];
        A[q * N + p] = t;
      }
    __syncthreads();
    const T inv = T(1) / A[c * N + c];
    for (int e = tid; e < N; e += nt) {
      fcol[e] = A[c * N + e];
      prow[e] = A[e * N + c] * inv;
    }
    __syncthreads();
    const T fr = fcol[r];
#pragma unroll 4
    for (int q = g; q < N; q += G) {
      T v;
      if (q == c) v = (r == c) ? inv : -fr * inv;
      else v = (r == c) ? prow[q] : A[q * N + r] - fr * prow[q];
      A[q * N + r] = v;
    }
    __syncthreads();
  }
  // undo the row swaps as column swaps, last step first
  for (int c = N - 1; c >= 0; --c) {
    const int pr = sw[c];
    if (pr != c)
      for (int rr = tid; rr < N; rr += nt) {
        const T t = A[c * N + rr];
        A[c * N + rr] = A[pr * N + rr];
        A[pr * N + rr] = t;
      }
    __syncthreads();
  }
  for (int e = tid; e < N * N; e += nt) Gm[e] = A[e];
}

// x[m] = inv[m] (scale * b[m]) (b[N-2] = b[N-1] = 0 when tau_zero); column-major inv.
// One block per system: the 4 wavefronts split the columns, lanes the rows.
template <typename T>
__global__ void __launch_bounds__(256) k_inv_apply(const T *Inv, const T *rhs, T *x, int N,
                                                   int tau_zero, T scale, const T *kscale,
                                                   const T *bscale, int per_run, int shared,
                                                   const int *active,
                                                   unsigned long long *units) {
  __shared__ T bs[LUS_MAXN];
  __shared__ T part[4][LUS_MAXN];
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  if (units && threadIdx.x == 0) atomicAdd(units, 1ull);
  const long mi = shared ? (m % per_run) : m;
  const T *A = Inv + mi * (long)N * N;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  T s = kscale ? scale * kscale[m % per_run] : scale;
  if (bscale) s = s * bscale[m / per_run];
  for (int r = tid; r < N; r += 256) {
    T v = rhs[(long)m * N + r] * s;
    if (tau_zero && r >= N - 2) v = T(0);
    bs[r] = v;
  }
  __syncthreads();
  const int cpw = (N + 3) / 4, c0 = wv * cpw, c1 = min(N, c0 + cpw);
  for (int r0 = 0; r0 < N; r0 += 64) {
    const int r = r0 + lane;
    T acc = 0;
    if (r < N)
      for (int c = c0; c < c1; ++c) acc = fma(A[(long)c * N + r], bs[c], acc);
    if (r < N) part[wv][r] = acc;
  }
  __syncthreads();
  for (int r = tid; r < N; r += 256)
    x[(long)m * N + r] = (part[0][r] + part[1][r]) + (part[2][r] + part[3][r]);
}

// FP64 form of k_inv_apply for even N: thread t owns the row pair
// (2 (t % RP), 2 (t % RP) + 1), RP = N / 2, for the columns c = t / RP,
// t / RP + CG, ... (CG = 256 / RP column groups), reading each column as
// 16-byte pieces, so a wavefront's loads cover whole column segments and
// nearly every lane works; partial sums meet in LDS.
__global__ void __launch_bounds__(256) k_inv_apply_pairs(const double *Inv, const double *rhs,
                                                         double *x, int N, int tau_zero,
                                                         double scale, const double *kscale,
                                                         const double *bscale, int per_run,
                                                         int shared, const int *active,
                                                         unsigned long long *units) {
  __shared__ double bs[LUS_MAXN];
  __shared__ double part[2 * 256];
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  if (units && threadIdx.x == 0) atomicAdd(units, 1ull);
  const long mi = shared ? (m % per_run) : m;
  const double *A = Inv + mi * (long)N * N;
  const int tid = threadIdx.x;
  double s = kscale ? scale * kscale[m % per_run] : scale;
  if (bscale) s = s * bscale[m / per_run];
  for (int r = tid; r < N; r += 256) {
    double v = rhs[(long)m * N + r] * s;
    if (tau_zero && r >= N - 2) v = 0.0;
    bs[r] = v;
  }
  __syncthreads();
  const int RP = N >> 1, CG = 256 / RP;
  const int rp = tid % RP, cg = tid / RP;
  if (cg < CG) {
    double a0 = 0.0, a1 = 0.0;
    const double *col = A + 2 * rp;
#pragma unroll 4
    for (int c = cg; c < N; c += CG) {
      const double2 v = *(const double2 *)(col + (long)c * N);
      const double b = bs[c];
      a0 = fma(v.x, b, a0);
      a1 = fma(v.y, b, a1);
    }
    part[cg * N + 2 * rp] = a0;
    part[cg * N + 2 * rp + 1] = a1;
  }
  __syncthreads();
  for (int r = tid; r < N; r += 256) {
    double acc = 0.0;
    for (int g = 0; g < CG; ++g) acc += part[g * N + r];
    x[(long)m * N + r] = acc;
  }
}

// The row-pair apply with every column load of the thread in flight at once:
// the block's MAXIT double2 loads per thread are issued first, the right-hand
// side is staged in LDS meanwhile, and the accumulation then runs from
// registers in the same order as k_inv_apply_pairs (bit-identical results).
// Needs ceil(N / CG) <= MAXIT with CG = 256 / (N / 2) column groups.
template <int MAXIT>
__global__ void __launch_bounds__(256) k_inv_apply_pairs_pf(
    const double *Inv, const double *rhs, double *x, int N, int tau_zero, double scale,
    const double *kscale, const double *bscale, int per_run, int shared, const int *active,
    unsigned long long *units) {
  __shared__ double bs[LUS_MAXN];
  __shared__ double part[2 * 256];
  const int m = blockIdx.x;
  if (active && !active[m / per_run]) return;
  if (units && threadIdx.x == 0) atomicAdd(units, 1ull);
  const long mi = shared ? (m % per_run) : m;
  const double *A = Inv + mi * (long)N * N;
  const int tid = threadIdx.x;
  const int RP = N >> 1, CG = 256 / RP;
  const int rp = tid % RP, cg = tid / RP;
  const bool on = cg < CG;
  double2 v[MAXIT];
#pragma unroll
  for (int i = 0; i < MAXIT; ++i) {
    const int c = cg + i * CG;
    v[i] = (on && c < N) ? *(const double2 *)(A + 2 * rp + (long)c * N)
                         : make_double2(0.0, 0.0);
  }
  double s = kscale ? scale * kscale[m % per_run] : scale;
  if (bscale) s = s * bscale[m / per_run];
  for (int r = tid; r < N; r += 256) {
    double u = rhs[(long)m * N + r] * s;
    if (tau_zero && r >= N - 2) u = 0.0;
    bs[r] = u;
  }
  __syncthreads();
  if (on) {
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int i = 0; i < MAXIT; ++i) {
      const int c = cg + i * CG;
      if (c < N) {
        const double b = bs[c];
        a0 = fma(v[i].x, b, a0);
        a1 = fma(v[i].y, b, a1);
      }
    }
    part[cg * N + 2 * rp] = a0;
    part[cg * N + 2 * rp + 1] = a1;
  }
  __syncthreads();
  for (int r = tid; r < N; r += 256) {
    double acc = 0.0;
    for (int g = 0; g < CG; ++g) acc += part[g * N + r];
    x[(long)m * N + r] = acc;
  }
}

extern "C" int smx_lu_factor(int prec, int nmat, int N, void *LU, int *piv,
                             int per_run, const int *active, void *stream) {
  if (nmat <= 0) return 0;
  if (N <= 0 || per_run <= 0) return fail_msg("smx_lu_factor: bad sizes");
  const size_t sh = (size_t)N * N * (prec == 0 ? 8 : 4);
  if (prec == 0) {
    if (lds_optin(k_lu_factor_lds<double>, sh))
      hipLaunchKernelGGL(k_lu_factor_lds<double>, dim3(nmat), dim3(256), sh, STREAM(stream),
                         (double *)LU, piv, N, per_run, active);
    else
      hipLaunchKernelGGL(k_lu_factor<double>, dim3(nmat), dim3(256), 0, STREAM(stream),
                         (double *)LU, piv, N, per_run, active);
  } else {
    if (lds_optin(k_lu_factor_lds<float>, sh))
      hipLaunchKernelGGL(k_lu_factor_lds<float>, dim3(nmat), dim3(256), sh, STREAM(stream),
                         (float *)LU, piv, N, per_run, active);
    else
      hipLaunchKernelGGL(k_lu_factor<float>, dim3(nmat), dim3(256), 0, STREAM(stream),
                         (float *)LU, piv, N, per_run, active);
  }
  CHECK_LAUNCH("smx_lu_factor");
  return 0;
}

// one wavefront per system; LU column-major
#define LUS_WAVES 4
template <typename T>
__global__ void __launch_bounds__(64 * LUS_WAVES)
k_lu_solve(const T *LU, const int *piv, const T *rhs, T *x, int nmat, int N, int tau_zero,
           T scale, const T *kscale, const T *bscale, int per_run, int shared,
           const int *active) {
  __shared__ T buf[LUS_WAVES][LUS_MAXN];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m = blockIdx.x * LUS_WAVES + wv;
  const bool live = (m < nmat) && !(active && !active[m / per_run]);
  T *y = buf[wv];
  const long mi = live ? (shared ? (m % per_run) : m) : 0;
  const T *A = LU + mi * N * N;
  if (live) {
    T s = kscale ? scale * kscale[m % per_run] : scale;
    if (bscale) s = s * bscale[m / per_run];
    for (int r = lane; r < N; r += 64) {
      T v = rhs[(long)m * N + r] * s;
      if (tau_zero && r >= N - 2) v = T(0);
      y[r] = v;
    }
  }
  __syncthreads();
  if (live && lane == 0) {
    const int *p = piv + mi * N;
    for (int c = 0; c < N; ++c) {
      int pc = p[c];
      if (pc != c) { T t = y[c]; y[c] = y[pc]; y[pc] = t; }
    }
  }
  __syncthreads();
  for (int c = 0; c < N - 1; ++c) {
    if (live) {
      T yc = y[c];
      for (int r = c + 1 + lane; r < N; r += 64) y[r] -= A[(long)c * N + r] * yc;
    }
    __syncthreads();
  }
  for (int c = N - 1; c >= 0; --c) {
    if (live && lane == 0) y[c] = y[c] / A[(long)c * N + c];
    __syncthreads();
    if (live) {
      T xc = y[c];
      for (int r = lane; r < c; r += 64) y[r] -= A[(long)c * N + r] * xc;
    }
    __syncthreads();
  }
  if (live)
    for (int r = lane; r < N; r += 64) x[(long)m * N + r] = y[r];
}

extern "C" int smx_lu_solve(int prec, int nmat, int N, const void *LU, const int *piv,
                            const void *rhs, void *x, int tau_zero, double rhs_scale,
                            const void *kscale, const void *bscale, int per_run, int shared,
                            const int *active, void *stream) {
  if (nmat <= 0) return 0;
  if (N <= 0 || N > LUS_MAXN || per_run <= 0) return fail_msg("smx_lu_solve: bad sizes");
  const size_t sh = (size_t)(N * N + N) * (prec == 0 ? 8 : 4);
  if (prec == 0 && lds_optin(k_lu_solve_lds<double>, sh)) {
    hipLaunchKernelGGL(k_lu_solve_lds<double>, dim3(nmat), dim3(256), sh, STREAM(stream),
                       (const double *)LU, piv, (const double *)rhs, (double *)x, N, tau_zero,
                       rhs_scale, (const double *)kscale, (const double *)bscale, per_run,
                       shared, active);
    CHECK_LAUNCH("smx_lu_solve");
    return 0;
  }
  if (prec == 1 && lds_optin(k_lu_solve_lds<float>, sh)) {
    hipLaunchKernelGGL(k_lu_solve_lds<float>, dim3(nmat), dim3(256), sh, STREAM(stream),
                       (const float *)LU, piv, (const float *)rhs, (float *)x, N, tau_zero,
                       (float)rhs_scale, (const float *)kscale, (const float *)bscale, per_run,
                       shared, active);
    CHECK_LAUNCH("smx_lu_solve");
    return 0;
  }
  dim3 grid(cdiv(nmat, LUS_WAVES));
  if (prec == 0)
    hipLaunchKernelGGL(k_lu_solve<double>, grid, dim3(64 * LUS_WAVES), 0, STREAM(stream),
                       (const double *)LU, piv, (const double *)rhs, (double *)x, nmat, N,
                       tau_zero, rhs_scale, (const double *)kscale,
                       (const double *)bscale, per_run, shared, active);
  else
    hipLaunchKernelGGL(k_lu_solve<float>, grid, dim3(64 * LUS_WAVES), 0, STREAM(stream),
                       (const float *)LU, piv, (const float *)rhs, (float *)x, nmat, N,
                       tau_zero, (float)rhs_scale, (const float *)kscale,
                       (const float *)bscale, per_run, shared, active);
  CHECK_LAUNCH("smx_lu_solve");
  return 0;
}


extern "C" int smx_invert(int prec, int nmat, int N, void *A, int per_run, const int *active,
                          void *stream) {
  if (nmat <= 0) return 0;
  if (N <= 0 || per_run <= 0) return fail_msg("smx_invert: bad sizes");
  if (prec == 0 && N >= 16 && N <= 128 && g_gj_mfma) {
    int rc = gj_invert_mfma(nmat, N, (double *)A, per_run, active, STREAM(stream));
    if (rc) return rc;
    CHECK_LAUNCH("smx_invert");
    return 0;
  }
  if (prec == 0 && N >= 16 && N <= GJR_MAXN && g_gj_reg) {
    // register-resident kernel: 4N threads, ceil(N / 4) columns per thread
    const int cpt = (N + 3) / 4;
    const dim3 grid(nmat), blk((4 * N + 63) / 64 * 64);
    hipStream_t st = STREAM(stream);
#define GJR(C_)                                                                        \
  do {                                                                                 \
    if (N == 4 * (C_))                                                                 \
      hipLaunchKernelGGL((k_gj_invert_reg<C_, true>), grid, blk, 0, st, (double *)A, N, \
                         per_run, active);                                             \
    else                                                                               \
      hipLaunchKernelGGL((k_gj_invert_reg<C_, false>), grid, blk, 0, st, (double *)A,   \
                         N, per_run, active);                                          \
  } while (0)
    if (cpt <= 6) GJR(6);
    else if (cpt <= 12) GJR(12);
    else if (cpt <= 24) GJR(24);
    else GJR(32);
#undef GJR
    CHECK_LAUNCH("smx_invert");
    return 0;
  }
  if (N >= 8 && N <= 512 && g_gj_rows) {
    // row-owner kernel: N * G threads, G = min(N, 512 / N)
    const int G = max(1, min(N, 512 / N));
    const size_t es = prec == 0 ? 8 : 4;
    const size_t sh = (size_t)N * N * es + 2 * (size_t)N * es + (size_t)N * 4;
    if (prec == 0) {
      if (!lds_optin(k_gj_invert_rows<double>, sh))
        return fail_msg("smx_invert: matrix exceeds LDS");
      hipLaunchKernelGGL(k_gj_invert_rows<double>, dim3(nmat), dim3(N * G), sh, STREAM(stream),
                         (double *)A, N, per_run, active);
    } else {
      if (!lds_optin(k_gj_invert_rows<float>, sh))
        return fail_msg("smx_invert: matrix exceeds LDS");
      hipLaunchKernelGGL(k_gj_invert_rows<float>, dim3(nmat), dim3(N * G), sh, STREAM(stream),
                         (float *)A, N, per_run, active);
    }
    CHECK_LAUNCH("smx_invert");
    return 0;
  }
  const size_t sh = (size_t)N * N * (prec == 0 ? 8 : 4) + (size_t)N * 4;
  if (prec == 0) {
    if (!lds_optin(k_gj_invert<double>, sh)) return fail_msg("smx_invert: matrix exceeds LDS");
    hipLaunchKernelGGL(k_gj_invert<double>, dim3(nmat), dim3(256), sh, STREAM(stream),
                       (double *)A, N, per_run, active);
  } else {
    if (!lds_optin(k_gj_invert<float>, sh)) return fail_msg("smx_invert: matrix exceeds LDS");
    hipLaunchKernelGGL(k_gj_invert<float>, dim3(nmat), dim3(256), sh, STREAM(stream),
                       (float *)A, N, per_run, active);
  }
  CHECK_LAUNCH("smx_invert");
  return 0;
}

extern "C" int smx_inv_apply(int prec, int nmat, int N, const void *Inv, const void *rhs,
                             void *x, int tau_zero, double rhs_scale, const void *kscale,
                             const void *bscale, int per_run, int shared, const int *active,
                             void *stream) {
  if (nmat <= 0) return 0;
  if (N <= 0 || N > LUS_MAXN || per_run <= 0) return fail_msg("smx_inv_apply: bad sizes");
  const bool clk = clk_on(2);
  if (clk) clk_begin(2, STREAM(stream));
  const bool pairs = prec == 0 && (N % 2) == 0 && N >= 4 && N <= 512 && g_inv_pairs;
  const int cgs = pairs ? 256 / (N / 2) : 1;
  const int iters = pairs && cgs > 0 ? (N + cgs - 1) / cgs : 1 << 30;
  if (pairs && g_inv_pairs == 1 && cgs > 0 && iters <= 24)
    hipLaunchKernelGGL(k_inv_apply_pairs_pf<24>, dim3(nmat), dim3(256), 0, STREAM(stream),
                       (const double *)Inv, (const double *)rhs, (double *)x, N, tau_zero,
                       rhs_scale, (const double *)kscale, (const double *)bscale, per_run,
                       shared, active, clk_units(2));
  else if (pairs && cgs > 0)
    hipLaunchKernelGGL(k_inv_apply_pairs, dim3(nmat), dim3(256), 0, STREAM(stream),
                       (const double *)Inv, (const double *)rhs, (double *)x, N, tau_zero,
                       rhs_scale, (const double *)kscale, (const double *)bscale, per_run,
                       shared, active, clk_units(2));
  else if (prec == 0)
    hipLaunchKernelGGL(k_inv_apply<double>, dim3(nmat), dim3(256), 0, STREAM(stream),
                       (const double *)Inv, (const double *)rhs, (double *)x, N, tau_zero,
                       rhs_scale, (const double *)kscale, (const double *)bscale, per_run,
                       shared, active, clk_units(2));
  else
    hipLaunchKernelGGL(k_inv_apply<float>, dim3(nmat), dim3(256), 0, STREAM(stream),
                       (const float *)Inv, (const float *)rhs, (float *)x, N, tau_zero,
                       (float)rhs_scale, (const float *)kscale, (const float *)bscale,
                       per_run, shared, active, clk_units(2));
  if (clk) clk_end(2, STREAM(stream));
  CHECK_LAUNCH("smx_inv_apply");
  return 0;
}

// ---------------------------------------------------------------------------
// operator assembly (column-major out[b][k][col*N + row])
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_tau_assemble(int kind, int Nk, int N, const T *D2, const T *kx, const T *lam,
                               const T *Mb, const T *lm, const T *cg, const T *V0, const T *V1,
                               T *out, const int *active) {
  const int bk = blockIdx.y, b = bk / Nk, k = bk % Nk;
  if (active && !active[b]) return;
  const long NN = (long)N * N;
  T *o = out + (long)bk * NN;
  const T kk = kx[k] * kx[k];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < NN; e += (long)gridDim.x * blockDim.x) {
    int row = (int)(e % N), col = (int)(e / N);
    T v;
    if (kind == 0) {
      if (row == N - 2) v = (col & 1) ? T(-1) : T(1);
      else if (row == N - 1) v = T(1);
      else v = D2[(long)row * N + col] - (row == col ? (kk + lam[b]) : T(0));
    } else {
      if (row == N - 2) v = V0[col];
      else if (row == N - 1) v = V1[col];
      else {
        long rc = (long)row * N + col;
        v = (lm[(long)b * NN + rc] - kk * Mb[(long)b * NN + rc]) + T(2) * kk * cg[(long)bk * NN + rc];
      }
    }
    o[e] = v;
  }
}

extern "C" int smx_tau_assemble(int prec, int kind, int B, int Nk, int N, const void *D2,
                                const void *kx, const void *lam, const void *Mb,
                                const void *lm, const void *cg, const void *V0,
                                const void *V1, void *out, const int *active, void *stream) {
  if (B <= 0) return 0;
  if ((long)B * Nk > 65535) return fail_msg("smx_tau_assemble: B*Nk > 65535");
  dim3 grid(cdiv((long)N * N, 256 * 4), B * Nk);
  if (prec == 0)
    hipLaunchKernelGGL(k_tau_assemble<double>, grid, dim3(256), 0, STREAM(stream), kind, Nk, N,
                       (const double *)D2, (const double *)kx, (const double *)lam,
                       (const double *)Mb, (const double *)lm, (const double *)cg,
                       (const double *)V0, (const double *)V1, (double *)out, active);
  else
    hipLaunchKernelGGL(k_tau_assemble<float>, grid, dim3(256), 0, STREAM(stream), kind, Nk, N,
                       (const float *)D2, (const float *)kx, (const float *)lam,
                       (const float *)Mb, (const float *)lm, (const float *)cg,
                       (const float *)V0, (const float *)V1, (float *)out, active);
  CHECK_LAUNCH("smx_tau_assemble");
  return 0;
}

template <typename T>
__global__ void k_cheb_mult(int N, const T *a, T *M, const int *active) {
  const int b = blockIdx.y;
  if (active && !active[b]) return;
  const T *ab = a + (long)b * N;
  T *Mo = M + (long)b * N * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)N * N;
       e += (long)gridDim.x * blockDim.x) {
    int p = (int)(e / N), n = (int)(e % N);
    T s = T(0);
    if (p >= n) s += T(0.5) * ab[p - n];
    if (p + n < N) s += T(0.5) * ab[p + n];
    if (p > 0 && n >= p) s += T(0.5) * ab[n - p];
    Mo[e] = s;
  }
}

extern "C" int smx_cheb_mult(int prec, int B, int N, const void *a, void *M,
                             const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(cdiv((long)N * N, 256 * 4), B);
  if (prec == 0)
    hipLaunchKernelGGL(k_cheb_mult<double>, grid, dim3(256), 0, STREAM(stream), N,
                       (const double *)a, (double *)M, active);
  else
    hipLaunchKernelGGL(k_cheb_mult<float>, grid, dim3(256), 0, STREAM(stream), N,
                       (const float *)a, (float *)M, active);
  CHECK_LAUNCH("smx_cheb_mult");
  return 0;
}

// Dense Schur blocks (stokes_coupled.py:242-289) for one run, real form.
// M, P = D2 M, Qm = M D2, R = D M D are [Nk][N][N] (row-major) per x-shift d.
// For block (i, j): d = i - j (direct half, M_{-d} = M_d for a real eta) and,
// when j >= 1 and i + j < Nk, d' = i + j (partner half, sign s):
//   A_ij  = (P_d - k_i^2 M_d) + s (P_d' - k_i^2 M_d')
//   U_ij  = (2k_j^2 P_d + 2k_i^2 Q_d - 4k_i k_j R_d)
//           + s (2k_j^2 P_d' + 2k_i^2 Q_d' + 4k_i k_j R_d')
// A is written into the dense row-major matrix (ld = Nk*N); U into [Nk][Nk][N][N].
// The tau rows (N-2, N-1 of every block row) are set separately.
template <typename T>
__global__ void k_schur_blocks(int Nk, int N, const T *M, const T *P, const T *Qm, const T *R,
                               const T *kx, T sgn, T *A, T *U) {
  const int ij = blockIdx.y, i = ij / Nk, j = ij % Nk;
  const long NN = (long)N * N, n = (long)Nk * N;
  const int d = i >= j ? i - j : j - i;
  const int dp = i + j;
  const bool part = (j >= 1) && (dp < Nk) && (sgn != T(0));
  const T ki = kx[i], kj = kx[j];
  const T ki2 = ki * ki, kj2 = kj * kj;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < NN; e += (long)gridDim.x * blockDim.x) {
    const int a = (int)(e / N), b = (int)(e % N);
    const long od = (long)d * NN + e;
    T av = P[od] - ki2 * M[od];
    T uv = (T(2) * kj2 * P[od] + T(2) * ki2 * Qm[od]) - T(4) * ki * kj * R[od];
    if (part) {
      const long op = (long)dp * NN + e;
      av = av + sgn * (P[op] - ki2 * M[op]);
      uv = uv + sgn * ((T(2) * kj2 * P[op] + T(2) * ki2 * Qm[op]) + T(4) * ki * kj * R[op]);
    }
    A[((long)i * N + a) * n + (long)j * N + b] = av;
    U[(long)ij * NN + e] = uv;
  }
}

// tau rows of the dense Schur matrix: rows i*N + N-2 / N-1 are V0 / V1 on the
// diagonal block and zero elsewhere (stokes_coupled.py:282-289)
template <typename T>
__global__ void k_schur_tau(int Nk, int N, const T *V0, const T *V1, T *A) {
  const int i = blockIdx.y, which = blockIdx.z;
  const long n = (long)Nk * N;
  T *row = A + ((long)i * N + N - 2 + which) * n;
  const T *v = which ? V1 : V0;
  for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < n; c += (long)gridDim.x * blockDim.x) {
    long jb = c / N;
    row[c] = (jb == i) ? v[c % N] : T(0);
  }
}

extern "C" int smx_schur_blocks(int prec, int Nk, int N, const void *M, const void *P,
                                const void *Qm, const void *R, const void *kx, double sgn,
                                void *A, void *U, void *stream) {
  if ((long)Nk * Nk > 65535) return fail_msg("smx_schur_blocks: Nk too large");
  dim3 grid(cdiv((long)N * N, 256 * 4), Nk * Nk);
  if (prec == 0)
    hipLaunchKernelGGL(k_schur_blocks<double>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const double *)M, (const double *)P, (const double *)Qm,
                       (const double *)R, (const double *)kx, sgn, (double *)A, (double *)U);
  else
    hipLaunchKernelGGL(k_schur_blocks<float>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const float *)M, (const float *)P, (const float *)Qm,
                       (const float *)R, (const float *)kx, (float)sgn, (float *)A, (float *)U);
  CHECK_LAUNCH("smx_schur_blocks");
  return 0;
}

extern "C" int smx_schur_tau(int prec, int Nk, int N, const void *V0, const void *V1, void *A,
                             void *stream) {
  dim3 grid(cdiv((long)Nk * N, 256), Nk, 2);
  if (prec == 0)
    hipLaunchKernelGGL(k_schur_tau<double>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const double *)V0, (const double *)V1, (double *)A);
  else
    hipLaunchKernelGGL(k_schur_tau<float>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const float *)V0, (const float *)V1, (float *)A);
  CHECK_LAUNCH("smx_schur_tau");
  return 0;
}

// ---------------------------------------------------------------------------
// viscosity law on the padded half grid (one block per run)
// params per run (double[6]): ln_deta, Ts, lo, hi, smooth_w, law
// ---------------------------------------------------------------------------
__device__ __forceinline__ double smooth5(double t) {
  t = fmin(fmax(t, 0.0), 1.0);
  return t * t * t * (t * (6.0 * t - 15.0) + 10.0);
}

// doubles per run in the viscosity-law parameter rows: ln deta_T, Ts, eta_min,
// eta_max, eta_smooth, law (0 exp, 1 Arrhenius, 2 pV Arrhenius), E, V
#define LAW_ROW 8

template <typename T>
__global__ void __launch_bounds__(1024) k_eta_law(int nrow, int N, const T *theta, const T *z,
                                                  const double *params, T *eta, double *counters,
                                                  const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  const double *pp = params + LAW_ROW * b;
  const double L = pp[0], Ts = pp[1], lo = pp[2], hi = pp[3], w = pp[4];
  const int law = (int)pp[5];
  const double elo = lo * (1.0 - 1e-12), ehi = hi * (1.0 + 1e-12);
  const long np = (long)nrow * N;
  const T *th = theta + (long)b * np;
  T *et = eta + (long)b * np;
  double n_out = 0, n_clipT = 0, n_clipE = 0, exc = 0;
  for (long e = threadIdx.x; e < np; e += blockDim.x) {
    int j = (int)(e / N), q = (int)(e % N);
    double wgt = (j == 0 || j == nrow - 1) ? 1.0 : 2.0;
    double T_ = (1.0 - (double)z[q]) + (double)th[e];
    exc = fmax(exc, fmax(0.0 - T_, T_ - 1.0));
    if (T_ < 0.0 || T_ > 1.0) n_out += wgt;
    if (T_ < 0.0 || T_ > 2.0) n_clipT += wgt;
    double Tc = fmin(fmax(T_, 0.0), 2.0);
    double v;
    if (law == 2)        // pV: -E T / (Ts (T + Ts)) + V (1 - z) / (T + Ts)
      v = exp(-pp[6] * Tc / (Ts * (Tc + Ts)) + pp[7] * (1.0 - (double)z[q]) / (Tc + Ts));
    else if (law == 1)
      v = exp(-L * (1.0 + Ts) * Tc / (Tc + Ts));
    else
      v = exp(-L * Tc);
    if (v < elo || v > ehi) n_clipE += wgt;
    et[e] = (T)v;
  }
  n_out = block_sum(n_out, red);
  n_clipT = block_sum(n_clipT, red);
  n_clipE = block_sum(n_clipE, red);
  exc = block_max(exc, red);
  double n_band = 0;
  if (w > 0.0) {
    // C^2 smooth clamp in ln(eta); n_clip_eta keeps the hard-clip count and
    // n_band counts the points strictly inside a transition band
    const double ylo = log(lo), yhi = log(hi);
    for (long e = threadIdx.x; e < np; e += blockDim.x) {
      double y = log((double)et[e]);
      if ((y > ylo && y < ylo + w) || (y > yhi - w && y < yhi))
        n_band += ((e / N) == 0 || (e / N) == nrow - 1) ? 1.0 : 2.0;
      double Y = ylo + (y - ylo) * smooth5((y - ylo) / w);
      double yo = yhi - (yhi - Y) * smooth5((yhi - Y) / w);
      et[e] = (T)exp(yo);
    }
    n_band = block_sum(n_band, red);
  } else if (n_clipE > 0.0) {
    for (long e = threadIdx.x; e < np; e += blockDim.x)
      et[e] = (T)fmin(fmax((double)et[e], lo), hi);
  }
  if (threadIdx.x == 0) {
    double *c = counters + 8 * b;
    c[0] += n_out;
    c[1] += n_clipT;
    c[2] += n_clipE;
    c[3] = fmax(c[3], fmax(exc, 0.0));
    c[5] += n_band;
  }
}

extern "C" int smx_eta_law(int prec, int B, int nrow, int N, const void *theta,
                           const void *z, const double *params, void *eta,
                           double *counters, const int *active, void *stream) {
  if (B <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_eta_law<double>, dim3(B), dim3(1024), 0, STREAM(stream), nrow, N,
                       (const double *)theta, (const double *)z, params, (double *)eta,
                       counters, active);
  else
    hipLaunchKernelGGL(k_eta_law<float>, dim3(B), dim3(1024), 0, STREAM(stream), nrow, N,
                       (const float *)theta, (const float *)z, params, (float *)eta, counters,
                       active);
  CHECK_LAUNCH("smx_eta_law");
  return 0;
}

// guard + profile on the unpadded half grid (nrow = Nx/2 + 1)
template <typename T>
__global__ void __launch_bounds__(1024) k_eta_guard(int Nx, int N, T *eta, const double *params,
                                                    T *prof, int profile, double *counters,
                                                    const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  const int nrow = Nx / 2 + 1;
  const long np = (long)nrow * N;
  T *et = eta + (long)b * np;
  const double *pp = params + LAW_ROW * b;
  const double lo = pp[2], hi = pp[3], w = pp[4];
  double mn = 1e300;
  for (long e = threadIdx.x; e < np; e += blockDim.x) mn = fmin(mn, (double)et[e]);
  mn = block_min(mn, red);
  if (!(mn > 0.0)) {
    const double elo = lo * (1.0 - 1e-12), ehi = hi * (1.0 + 1e-12);
    double nc = 0;
    for (long e = threadIdx.x; e < np; e += blockDim.x) {
      int j = (int)(e / N);
      double wgt = (j == 0 || j == nrow - 1) ? 1.0 : 2.0;
      double v = (double)et[e];
      if (v < elo || v > ehi) nc += wgt;
    }
    nc = block_sum(nc, red);
    __syncthreads();
    double nb = 0;
    if (w > 0.0) {
      const double ylo = log(lo), yhi = log(hi);
      for (long e = threadIdx.x; e < np; e += blockDim.x) {
        double y = log((double)et[e]);
        if ((y > ylo && y < ylo + w) || (y > yhi - w && y < yhi))
          nb += ((e / N) == 0 || (e / N) == nrow - 1) ? 1.0 : 2.0;
        double Y = ylo + (y - ylo) * smooth5((y - ylo) / w);
        double yo = yhi - (yhi - Y) * smooth5((yhi - Y) / w);
        et[e] = (T)exp(yo);
      }
      nb = block_sum(nb, red);
    } else if (nc > 0.0) {
      for (long e = threadIdx.x; e < np; e += blockDim.x)
        et[e] = (T)fmin(fmax((double)et[e], lo), hi);
    }
    if (threadIdx.x == 0) {
      counters[8 * b + 4] += nc;
      counters[8 * b + 6] += nb;
    }
    __syncthreads();
  }
  // profile per z node: reduce over the full Nx grid (mirror weights)
  if (prof) {
    for (int q = threadIdx.x; q < N; q += blockDim.x) {
      double s = 0.0;
      for (int j = 0; j < nrow; ++j) {
        double v = (double)et[(long)j * N + q];
        double f = (profile == 0) ? 1.0 / v : (profile == 1 ? v : log(v));
        s += ((j == 0 || j == nrow - 1) ? 1.0 : 2.0) * f;
      }
      double mean = s / (double)Nx;
      double r = (profile == 0) ? 1.0 / mean : (profile == 1 ? mean : exp(mean));
      prof[(long)b * N + q] = (T)r;
    }
  }
}

extern "C" int smx_eta_guard_profile(int prec, int B, int Nx, int N, void *eta,
                                     const double *params, void *prof, int profile,
                                     double *counters, const int *active, void *stream) {
  if (B <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_eta_guard<double>, dim3(B), dim3(1024), 0, STREAM(stream), Nx, N,
                       (double *)eta, params, (double *)prof, profile, counters, active);
  else
    hipLaunchKernelGGL(k_eta_guard<float>, dim3(B), dim3(1024), 0, STREAM(stream), Nx, N,
                       (float *)eta, params, (float *)prof, profile, counters, active);
  CHECK_LAUNCH("smx_eta_guard_profile");
  return 0;
}

// ---------------------------------------------------------------------------
// pseudo-spectral x products (synthesis on the mirror half of the padded
// grid, pointwise product, analysis), one block per (run, field, 16 columns)
// y/out layout [B][Nk][nf][Q]; eta [B][nrow][Q]; So/Ce [nrow][Nk]; Ao/Ae [Nk][nrow]
// ---------------------------------------------------------------------------
#define XC 16
template <typename T>
__global__ void __launch_bounds__(256) k_xprod(int nf, int nodd, int Nk, int nrow, int Q,
                                               const T *y, const T *fscale, const T *eta,
                                               long eta_bs, T *out,
                                               const T *So, const T *Ce, const T *Ao,
                                               const T *Ae, const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *Ys = (T *)smem_raw;        // [Nk][XC]
  T *Gs = Ys + Nk * XC;         // [nrow][XC]
  const int b = blockIdx.z, f = blockIdx.y, q0 = blockIdx.x * XC;
  if (active && !active[b]) return;
  const bool odd = f < nodd;
  const T *S = odd ? So : Ce;
  const T *An = odd ? Ao : Ae;
  const long rs = (long)nf * Q;  // stride between modes
  const T *yb = y + (long)b * Nk * rs + (long)f * Q;
  for (int e = threadIdx.x; e < Nk * XC; e += blockDim.x) {
    int m = e / XC, c = e % XC, q = q0 + c;
    T v = (q < Q) ? yb[(long)m * rs + q] : T(0);
    Ys[e] = fscale ? fscale[(long)f * Nk + m] * v : v;
  }
  __syncthreads();
  const T *eb = eta + (long)b * eta_bs;
  for (int e = threadIdx.x; e < nrow * XC; e += blockDim.x) {
    int j = e / XC, c = e % XC, q = q0 + c;
    const T *Sr = S + (long)j * Nk;
    T s = 0;
    for (int m = 0; m < Nk; ++m) s = fma(Sr[m], Ys[m * XC + c], s);
    Gs[e] = (q < Q) ? s * eb[(long)j * Q + q] : T(0);
  }
  __syncthreads();
  T *ob = out + (long)b * Nk * rs + (long)f * Q;
  for (int e = threadIdx.x; e < Nk * XC; e += blockDim.x) {
    int m = e / XC, c = e % XC, q = q0 + c;
    if (q >= Q) continue;
    const T *Ar = An + (long)m * nrow;
    T s = 0;
    for (int j = 0; j < nrow; ++j) s = fma(Ar[j], Gs[j * XC + c], s);
    ob[(long)m * rs + q] = s;
  }
}

// FP64 matrix-core form of k_xprod: one block per (run, field, 32-column
// tile).  The synthesis G = S Y (rows nrow, K = Nk) and the analysis
// out = A (eta .* G) (rows Nk, K = nrow) run on v_mfma_f64_16x16x4f64: each
// wavefront owns 16 x 16 output tiles; A fragments come from L1/L2 (the
// transform matrices are a few tens of KB per shape), B fragments from LDS.
// The product with eta happens on the synthesis accumulators before they
// are staged in LDS as the analysis operand.
#define XT 32
// one 16 x 16 output tile: acc += Arow[0..K) (lane row) x Bs[k][col], K in
// steps of 4; with KTC > 0 the K loop is unrolled and every A fragment is
// loaded before the first MFMA (the fragments come from L1/L2)
template <int KTC>
__device__ __forceinline__ v4f64 xp_tile(const double *Ar, bool row_ok, int Klim, int KT,
                                         const double *Bs, int ld, int col, int lk) {
  v4f64 acc = v4f64{0.0, 0.0, 0.0, 0.0};
  if (KTC > 0) {
    double af[KTC > 0 ? KTC : 1];
#pragma unroll
    for (int ks = 0; ks < KTC; ++ks) {
      const int k = ks * 4 + lk;
      af[ks] = (row_ok && k < Klim) ? Ar[k] : 0.0;
    }
#pragma unroll
    for (int ks = 0; ks < KTC; ++ks)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(af[ks], Bs[(ks * 4 + lk) * ld + col], acc, 0,
                                                 0, 0);
  } else {
    for (int ks = 0; ks < KT; ++ks) {
      const int k = ks * 4 + lk;
      const double a = (row_ok && k < Klim) ? Ar[k] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bs[k * ld + col], acc, 0, 0, 0);
    }
  }
  return acc;
}

template <int KT1C, int KT2C>
__global__ void __launch_bounds__(256) k_xprod_mfma(int nf, int nodd, int Nk, int nrow, int Q,
                                                    const double *y, const double *fscale,
                                                    const double *eta, long eta_bs, double *out,
                                                    const double *So, const double *Ce,
                                                    const double *Ao, const double *Ae,
                                                    const int *active) {
  extern __shared__ unsigned char smem_raw[];
  const int KP1 = (Nk + 3) & ~3;          // synthesis K, padded to the MFMA K step
  const int MP1 = (nrow + 15) & ~15;      // grid rows, padded to the MFMA tile
  double *Ys = (double *)smem_raw;        // [KP1][XT + 1]
  double *Hs = Ys + KP1 * (XT + 1);       // [MP1][XT + 1]
  const int b = blockIdx.z, f = blockIdx.y, q0 = blockIdx.x * XT;
  if (active && !active[b]) return;
  const bool odd = f < nodd;
  const double *S = odd ? So : Ce;        // [nrow][Nk]
  const double *An = odd ? Ao : Ae;       // [Nk][nrow]
  const long rs = (long)nf * Q;
  const double *yb = y + (long)b * Nk * rs + (long)f * Q;
  const double *eb = eta + (long)b * eta_bs;
  double *ob = out + (long)b * Nk * rs + (long)f * Q;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  for (int e = tid; e < KP1 * XT; e += 256) {
    const int m = e / XT, c = e % XT, q = q0 + c;
    double v = 0.0;
    if (m < Nk && q < Q) {
      v = yb[(long)m * rs + q];
      if (fscale) v *= fscale[(long)f * Nk + m];
    }
    Ys[m * (XT + 1) + c] = v;
  }
  __syncthreads();
  const int MT1 = MP1 / 16;
  for (int t = w; t < MT1 * 2; t += 4) {
    const int ti = t >> 1, tj = t & 1;
    const int row = ti * 16 + lr;
    const v4f64 acc = xp_tile<KT1C>(S + (long)(row < nrow ? row : 0) * Nk, row < nrow, Nk,
                                    KP1 / 4, Ys, XT + 1, tj * 16 + lr, lk);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = ti * 16 + lk + 4 * r, c = tj * 16 + lr, q = q0 + c;
      const double ev = (j < nrow && q < Q) ? eb[(long)j * Q + q] : 0.0;
      Hs[j * (XT + 1) + c] = acc[r] * ev;
    }
  }
  __syncthreads();
  const int MT2 = (Nk + 15) / 16;
  for (int t = w; t < MT2 * 2; t += 4) {
    const int ti = t >> 1, tj = t & 1;
    const int row = ti * 16 + lr;
    const v4f64 acc = xp_tile<KT2C>(An + (long)(row < Nk ? row : 0) * nrow, row < Nk, nrow,
                                    MP1 / 4, Hs, XT + 1, tj * 16 + lr, lk);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = ti * 16 + lk + 4 * r, q = q0 + tj * 16 + lr;
      if (m < Nk && q < Q) ob[(long)m * rs + q] = acc[r];
    }
  }
}

extern "C" int smx_xprod(int prec, int B, int nf, int nodd, int Nk, int nrow, int Q,
                         const void *y, const void *fscale, const void *eta, long eta_bs,
                         void *out,
                         const void *So,
                         const void *Ce, const void *Ao, const void *Ae, const int *active,
                         void *stream) {
  if (B <= 0) return 0;
  if (prec == 0 && g_xprod_mfma) {
    const int KP1 = (Nk + 3) & ~3, MP1 = (nrow + 15) & ~15;
    const size_t sh = (size_t)(KP1 + MP1) * (XT + 1) * 8;
    const dim3 grid(cdiv(Q, XT), nf, B);
    hipStream_t st = STREAM(stream);
#define XPM(A_, B_)                                                                           \
  do {                                                                                        \
    if (!lds_optin(k_xprod_mfma<A_, B_>, sh))                                                 \
      return fail_msg("smx_xprod: grid too large for LDS");                                  \
    hipLaunchKernelGGL((k_xprod_mfma<A_, B_>), grid, dim3(256), sh, st, nf, nodd, Nk, nrow,   \
                       Q, (const double *)y, (const double *)fscale, (const double *)eta,     \
                       eta_bs, (double *)out, (const double *)So, (const double *)Ce,         \
                       (const double *)Ao, (const double *)Ae, active);                       \
  } while (0)
    // unrolled K loops for the grid widths Nx = 16, 32, 64, 128
    const int KT1 = KP1 / 4, KT2 = MP1 / 4;
    if (KT1 == 3 && KT2 == 8) XPM(3, 8);
    else if (KT1 == 5 && KT2 == 12) XPM(5, 12);
    else if (KT1 == 9 && KT2 == 20) XPM(9, 20);
    else if (KT1 == 17 && KT2 == 36) XPM(17, 36);
    else XPM(0, 0);
#undef XPM
    CHECK_LAUNCH("smx_xprod");
    return 0;
  }
  dim3 grid(cdiv(Q, XC), nf, B);
  size_t sh = (size_t)(Nk + nrow) * XC * (prec == 0 ? 8 : 4);
  if (!(prec == 0 ? lds_optin(k_xprod<double>, sh) : lds_optin(k_xprod<float>, sh)))
    return fail_msg("smx_xprod: grid too large for LDS");
  if (prec == 0)
    hipLaunchKernelGGL(k_xprod<double>, grid, dim3(256), sh, STREAM(stream), nf, nodd, Nk,
                       nrow, Q, (const double *)y, (const double *)fscale,
                       (const double *)eta, eta_bs, (double *)out,
                       (const double *)So, (const double *)Ce, (const double *)Ao,
                       (const double *)Ae, active);
  else
    hipLaunchKernelGGL(k_xprod<float>, grid, dim3(256), sh, STREAM(stream), nf, nodd, Nk, nrow,
                       Q, (const float *)y, (const float *)fscale, (const float *)eta,
                       eta_bs, (float *)out,
                       (const float *)So, (const float *)Ce, (const float *)Ao,
                       (const float *)Ae, active);
  CHECK_LAUNCH("smx_xprod");
  return 0;
}

// advect product: y [B][Nk][2][N] (f=0 odd: k v, f=1 even: D v, values in z),
// ux/uz [B][nrow][N]; out [B][Nk][N] = Ae( -(ux So y0 + uz Ce y1) [+ e1] [+ h] [+ e2] )
template <typename T>
__global__ void __launch_bounds__(256) k_advect(int Nk, int nrow, int N, const T *y,
                                                const T *ks0, const T *ux,
                                                const T *uz, const T *e1, const T *h,
                                                const T *e2, T *out, const T *So, const T *Ce,
                                                const T *Ae, const int *active) {
  extern __shared__ unsigned char smem_raw[];
  T *Y0 = (T *)smem_raw;     // [Nk][XC]
  T *Y1 = Y0 + Nk * XC;      // [Nk][XC]
  T *P = Y1 + Nk * XC;       // [nrow][XC]
  const int b = blockIdx.y, q0 = blockIdx.x * XC;
  if (active && !active[b]) return;
  const T *yb = y + (long)b * Nk * 2 * N;
  for (int e = threadIdx.x; e < Nk * XC; e += blockDim.x) {
    int m = e / XC, c = e % XC, q = q0 + c;
    T v0 = (q < N) ? yb[(long)m * 2 * N + q] : T(0);
    Y0[e] = ks0 ? ks0[m] * v0 : v0;
    Y1[e] = (q < N) ? yb[(long)m * 2 * N + N + q] : T(0);
  }
  __syncthreads();
  const long gb = (long)b * nrow * N;
  for (int e = threadIdx.x; e < nrow * XC; e += blockDim.x) {
    int j = e / XC, c = e % XC, q = q0 + c;
    T vx = 0, vz = 0;
    const T *Sr = So + (long)j * Nk;
    const T *Cr = Ce + (long)j * Nk;
    for (int m = 0; m < Nk; ++m) {
      vx = fma(Sr[m], Y0[m * XC + c], vx);
      vz = fma(Cr[m], Y1[m * XC + c], vz);
    }
    T p = 0;
    if (q < N) {
      long o = gb + (long)j * N + q;
      p = -(ux[o] * vx + uz[o] * vz);
      if (e1) p = p + e1[o];
      if (h) p = p + h[b];
      if (e2) p = p + e2[o];
    }
    P[e] = p;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < Nk * XC; e += blockDim.x) {
    int m = e / XC, c = e % XC, q = q0 + c;
    if (q >= N) continue;
    const T *Ar = Ae + (long)m * nrow;
    T s = 0;
    for (int j = 0; j < nrow; ++j) s = fma(Ar[j], P[j * XC + c], s);
    out[(long)b * Nk * N + (long)m * N + q] = s;
  }
}

extern "C" int smx_advect_prod(int prec, int B, int Nk, int nrow, int N, const void *y,
                               const void *ks0, const void *ux, const void *uz, const void *e1, const void *h,
                               const void *e2, void *out, const void *So, const void *Ce,
                               const void *Ae, const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(cdiv(N, XC), B);
  size_t sh = (size_t)(2 * Nk + nrow) * XC * (prec == 0 ? 8 : 4);
  if (!(prec == 0 ? lds_optin(k_advect<double>, sh) : lds_optin(k_advect<float>, sh)))
    return fail_msg("smx_advect_prod: grid too large for LDS");
  if (prec == 0)
    hipLaunchKernelGGL(k_advect<double>, grid, dim3(256), sh, STREAM(stream), Nk, nrow, N,
                       (const double *)y, (const double *)ks0, (const double *)ux,
                       (const double *)uz,
                       (const double *)e1, (const double *)h, (const double *)e2,
                       (double *)out, (const double *)So, (const double *)Ce,
                       (const double *)Ae, active);
  else
    hipLaunchKernelGGL(k_advect<float>, grid, dim3(256), sh, STREAM(stream), Nk, nrow, N,
                       (const float *)y, (const float *)ks0, (const float *)ux,
                       (const float *)uz,
                       (const float *)e1, (const float *)h, (const float *)e2, (float *)out,
                       (const float *)So, (const float *)Ce, (const float *)Ae, active);
  CHECK_LAUNCH("smx_advect_prod");
  return 0;
}

// ---------------------------------------------------------------------------
// elementwise
// ---------------------------------------------------------------------------
// out = (a * a_b) x + (b * b_b) y per run (a_b, b_b, y may be NULL)
template <typename T>
__global__ void k_axpby(long n, T a, const T *ab, const T *x, T b, const T *bb, const T *y,
                        T *out, const int *active) {
  const int r = blockIdx.y;
  if (active && !active[r]) return;
  const long off = (long)r * n;
  const T ca = ab ? a * ab[r] : a;
  const T cb = bb ? b * bb[r] : b;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    T v = ca * x[off + e];
    if (y) v = v + cb * y[off + e];
    out[off + e] = v;
  }
}
extern "C" int smx_axpby(int prec, long n_per_run, int B, double a, const void *ab,
                         const void *x, double b, const void *bb, const void *y, void *out,
                         const int *active, void *stream) {
  if (B <= 0 || n_per_run <= 0) return 0;
  dim3 grid(min(cdiv(n_per_run, 256), 1024), B);
  if (prec == 0)
    hipLaunchKernelGGL(k_axpby<double>, grid, dim3(256), 0, STREAM(stream), n_per_run, a,
                       (const double *)ab, (const double *)x, b, (const double *)bb,
                       (const double *)y, (double *)out, active);
  else
    hipLaunchKernelGGL(k_axpby<float>, grid, dim3(256), 0, STREAM(stream), n_per_run, (float)a,
                       (const float *)ab, (const float *)x, (float)b, (const float *)bb,
                       (const float *)y, (float *)out, active);
  CHECK_LAUNCH("smx_axpby");
  return 0;
}

// out[b][k][j] = (alpha * ab[b]) * ks[k] * x[b][k][j]; tau entries zeroed when tau_zero
template <typename T>
__global__ void k_rowscale(int Nk, int N, const T *x, const T *ks, T alpha, const T *ab,
                           int tau_zero, T *out, const int *active) {
  const int b = blockIdx.y;
  if (active && !active[b]) return;
  const long n = (long)Nk * N;
  const T al = ab ? alpha * ab[b] : alpha;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    int k = (int)(e / N), j = (int)(e % N);
    T v = (tau_zero && j >= N - 2) ? T(0) : (ks ? ks[k] * al : al) * x[(long)b * n + e];
    out[(long)b * n + e] = v;
  }
}
extern "C" int smx_rowscale(int prec, int B, int Nk, int N, const void *x, const void *ks,
                            double alpha, const void *ab, int tau_zero, void *out,
                            const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(cdiv((long)Nk * N, 256), B);
  if (prec == 0)
    hipLaunchKernelGGL(k_rowscale<double>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const double *)x, (const double *)ks, alpha, (const double *)ab,
                       tau_zero, (double *)out, active);
  else
    hipLaunchKernelGGL(k_rowscale<float>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const float *)x, (const float *)ks, (float)alpha, (const float *)ab,
                       tau_zero, (float *)out, active);
  CHECK_LAUNCH("smx_rowscale");
  return 0;
}

// theta right-hand side (implicit_theta.py:385-449, heating.py:219-230,
// dissipation.py:474-487): b = (b0 th - b1 thp) + k psi; tau entries zeroed;
// b[0][0] += H; b[:, :N-2] += src
template <typename T>
__global__ void k_theta_rhs(int Nk, int N, const T *th, const T *thp, const T *psi,
                            const T *kx, const T *b0, const T *b1, const T *H, const T *src,
                            T *out, const int *active) {
  const int b = blockIdx.y;
  if (active && !active[b]) return;
  const long n = (long)Nk * N, o = (long)b * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    int k = (int)(e / N), j = (int)(e % N);
    T v = b0[b] * th[o + e];
    if (b1[b] != T(0)) v = v - b1[b] * thp[o + e];
    v = v + kx[k] * psi[o + e];
    if (j >= N - 2) v = T(0);
    if (e == 0 && H && H[b] != T(0)) v = v + H[b];
    if (src && j < N - 2) v = v + src[o + e];
    out[o + e] = v;
  }
}
extern "C" int smx_theta_rhs(int prec, int B, int Nk, int N, const void *th, const void *thp,
                             const void *psi, const void *kx, const void *b0, const void *b1,
                             const void *H, const void *src, void *out, const int *active,
                             void *stream) {
  if (B <= 0) return 0;
  dim3 grid(cdiv((long)Nk * N, 256), B);
  if (prec == 0)
    hipLaunchKernelGGL(k_theta_rhs<double>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const double *)th, (const double *)thp, (const double *)psi,
                       (const double *)kx, (const double *)b0, (const double *)b1,
                       (const double *)H, (const double *)src, (double *)out, active);
  else
    hipLaunchKernelGGL(k_theta_rhs<float>, grid, dim3(256), 0, STREAM(stream), Nk, N,
                       (const float *)th, (const float *)thp, (const float *)psi,
                       (const float *)kx, (const float *)b0, (const float *)b1,
                       (const float *)H, (const float *)src, (float *)out, active);
  CHECK_LAUNCH("smx_theta_rhs");
  return 0;
}

// viscous dissipation on the padded half grid (dissipation.py:336-362):
// phi = scale_b * 2 * eta * ((gxx^2 + gzz^2 + 2 gxz^2) - (1/3) tr^2), gzz = -gxx, tr = 0
template <typename T>
__global__ void k_phi_grid(long np, const T *gxx, const T *gxz, const T *eta, const T *scale,
                           T *out, const int *active) {
  const int b = blockIdx.y;
  if (active && !active[b]) return;
  const long o = (long)b * np;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < np; e += (long)gridDim.x * blockDim.x) {
    T xx = gxx[o + e], zz = -xx, xz = gxz[o + e], tr = xx + zz;
    T dev = (xx * xx + zz * zz + T(2) * xz * xz) - (T(1) / T(3)) * tr * tr;
    T ph = eta ? T(2) * eta[o + e] * dev : T(2) * dev;
    out[o + e] = scale ? scale[b] * ph : ph;
  }
}
extern "C" int smx_phi_grid(int prec, int B, long np, const void *gxx, const void *gxz,
                            const void *eta, const void *scale, void *out, const int *active,
                            void *stream) {
  if (B <= 0) return 0;
  dim3 grid(min(cdiv(np, 256), 1024), B);
  if (prec == 0)
    hipLaunchKernelGGL(k_phi_grid<double>, grid, dim3(256), 0, STREAM(stream), np,
                       (const double *)gxx, (const double *)gxz, (const double *)eta,
                       (const double *)scale, (double *)out, active);
  else
    hipLaunchKernelGGL(k_phi_grid<float>, grid, dim3(256), 0, STREAM(stream), np,
                       (const float *)gxx, (const float *)gxz, (const float *)eta,
                       (const float *)scale, (float *)out, active);
  CHECK_LAUNCH("smx_phi_grid");
  return 0;
}

// diagnostics (rbc_galerkin.py:439-453) from the unpadded half grids
// ux, uz, th [B][Nx/2+1][N] and dth0 = D theta_0 [B][N]:
// out[b] = Nu_top, Nu_bot, Nu_energy, urms, umax, amp
template <typename T>
__global__ void __launch_bounds__(256) k_diag(int Nx, int N, const T *ux, const T *uz, const T *th,
                                              const T *uxo, const T *uzo, const T *tho,
                                              const T *dth0, long ldd, const T *V0, const T *V1,
                                              const T *wq, double *out, const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  const int nrow = Nx / 2 + 1;
  const long o = (long)b * nrow * N;
  double ek = 0, en = 0, um = 0, am = 0;
  for (int q = threadIdx.x; q < N; q += blockDim.x) {
    double sk = 0, sn = 0;
    for (int j = 0; j < nrow; ++j) {
      double w = (j == 0 || j == nrow - 1) ? 1.0 : 2.0;
      double a = (double)ux[o + (long)j * N + q], c = (double)uz[o + (long)j * N + q];
      double t = (double)th[o + (long)j * N + q];
      if (uxo) {
        // x-odd parts: squares and products of matching parity survive the
        // x-mean; x_j and x_{-j} carry e + o and e - o
        const double ao = (double)uxo[o + (long)j * N + q], co = (double)uzo[o + (long)j * N + q];
        const double to = (double)tho[o + (long)j * N + q];
        sk += w * ((a * a + c * c) + (ao * ao + co * co));
        sn += w * (c * t + co * to);
        um = fmax(um, fmax(fabs(a) + fabs(ao), fabs(c) + fabs(co)));
        am = fmax(am, fabs(t) + fabs(to));
        continue;
      }
      sk += w * (a * a + c * c);
      sn += w * (c * t);
      um = fmax(um, fmax(fabs(a), fabs(c)));
      am = fmax(am, fabs(t));
    }
    ek += (double)wq[q] * (sk / Nx);
    en += (double)wq[q] * (sn / Nx);
  }
  ek = block_sum(ek, red);
  en = block_sum(en, red);
  um = block_max(um, red);
  am = block_max(am, red);
  double t0 = 0, t1 = 0;
  for (int q = threadIdx.x; q < N; q += blockDim.x) {
    double d = (double)dth0[(long)b * ldd + q];
    t1 += (double)V1[q] * d;
    t0 += (double)V0[q] * d;
  }
  t0 = block_sum(t0, red);
  t1 = block_sum(t1, red);
  if (threadIdx.x == 0) {
    double *r = out + 6 * b;
    r[0] = 1.0 - t1;
    r[1] = 1.0 - t0;
    r[2] = 1.0 + en;
    r[3] = sqrt(fmax(ek, 0.0));
    r[4] = um;
    r[5] = am;
  }
}
static int diag_launch(int prec, int B, int Nx, int N, const void *ux, const void *uz,
                       const void *th, const void *uxo, const void *uzo, const void *tho,
                       const void *dth0, long ldd, const void *V0, const void *V1,
                       const void *wq, double *out, const int *active, void *stream) {
  if (B <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_diag<double>, dim3(B), dim3(256), 0, STREAM(stream), Nx, N,
                       (const double *)ux, (const double *)uz, (const double *)th,
                       (const double *)uxo, (const double *)uzo, (const double *)tho,
                       (const double *)dth0, ldd, (const double *)V0, (const double *)V1,
                       (const double *)wq, out, active);
  else
    hipLaunchKernelGGL(k_diag<float>, dim3(B), dim3(256), 0, STREAM(stream), Nx, N,
                       (const float *)ux, (const float *)uz, (const float *)th,
                       (const float *)uxo, (const float *)uzo, (const float *)tho,
                       (const float *)dth0, ldd, (const float *)V0, (const float *)V1,
                       (const float *)wq, out, active);
  CHECK_LAUNCH("smx_diag");
  return 0;
}
extern "C" int smx_diag(int prec, int B, int Nx, int N, const void *ux, const void *uz,
                        const void *th, const void *dth0, long ldd, const void *V0,
                        const void *V1,
                        const void *wq, double *out, const int *active, void *stream) {
  return diag_launch(prec, B, Nx, N, ux, uz, th, nullptr, nullptr, nullptr, dth0, ldd, V0, V1,
                     wq, out, active, stream);
}
extern "C" int smx_diag_general(int prec, int B, int Nx, int N, const void *ux_o,
                                const void *uz_e, const void *th_e, const void *ux_e,
                                const void *uz_o, const void *th_o, const void *dth0, long ldd,
                                const void *V0, const void *V1, const void *wq, double *out,
                                const int *active, void *stream) {
  if (!ux_e || !uz_o || !th_o) return fail_msg("smx_diag_general: null odd-part operand");
  return diag_launch(prec, B, Nx, N, ux_o, uz_e, th_e, ux_e, uz_o, th_o, dth0, ldd, V0, V1, wq,
                     out, active, stream);
}

// out[b][k][N-2] = v[b][k] . V0, out[b][k][N-1] = v[b][k] . V1
template <typename T>
__global__ void k_tau_rows(int Nk, int N, const T *v, const T *V0, const T *V1, T *out,
                           const int *active) {
  const int b = blockIdx.y;
  if (active && !active[b]) return;
  const int k = blockIdx.x;
  const T *vr = v + ((long)b * Nk + k) * N;
  T s0 = 0, s1 = 0;
  for (int j = threadIdx.x; j < N; j += 64) { s0 += vr[j] * V0[j]; s1 += vr[j] * V1[j]; }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (threadIdx.x == 0) {
    out[((long)b * Nk + k) * N + N - 2] = s0;
    out[((long)b * Nk + k) * N + N - 1] = s1;
  }
}
extern "C" int smx_tau_rows(int prec, int B, int Nk, int N, const void *v, const void *V0,
                            const void *V1, void *out, const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(Nk, B);
  if (prec == 0)
    hipLaunchKernelGGL(k_tau_rows<double>, grid, dim3(64), 0, STREAM(stream), Nk, N,
                       (const double *)v, (const double *)V0, (const double *)V1,
                       (double *)out, active);
  else
    hipLaunchKernelGGL(k_tau_rows<float>, grid, dim3(64), 0, STREAM(stream), Nk, N,
                       (const float *)v, (const float *)V0, (const float *)V1, (float *)out,
                       active);
  CHECK_LAUNCH("smx_tau_rows");
  return 0;
}

// Schur matvec combine (lagged_stokes.py:212-218) from truncated products
// t [B][Nk][4][N] and the derivative products dt [B][Nk][3][N]:
// dt0 = D2 t0, dt1 = D2 t1, dt3 = D t3.
// out = (dt0 - k^2 t0) + dt1 + 2k^2 t2 - 4k dt3 ; tau rows from v.
template <typename T>
__global__ void k_stokes_combine(int Nk, int N, const T *t, const T *dt, const T *kx, const T *v,
                                 const T *V0, const T *V1, T *out, const int *active) {
  const int b = blockIdx.y, k = blockIdx.x;
  if (active && !active[b]) return;
  const T kk = kx[k] * kx[k], kv = kx[k];
  const T *tr = t + ((long)b * Nk + k) * 4 * N;
  const T *dr = dt + ((long)b * Nk + k) * 3 * N;
  T *o = out + ((long)b * Nk + k) * N;
  for (int j = threadIdx.x; j < N - 2; j += blockDim.x) {
    T r = dr[j] - kk * tr[j];
    r = r + dr[N + j];
    r = r + T(2) * kk * tr[2 * N + j];
    r = r - T(4) * kv * dr[2 * N + j];
    o[j] = r;
  }
  const T *vr = v + ((long)b * Nk + k) * N;
  T s0 = 0, s1 = 0;
  for (int j = threadIdx.x; j < N; j += blockDim.x) { s0 += vr[j] * V0[j]; s1 += vr[j] * V1[j]; }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (threadIdx.x == 0) { o[N - 2] = s0; o[N - 1] = s1; }
}
extern "C" int smx_stokes_combine(int prec, int B, int Nk, int N, const void *t,
                                  const void *dt, const void *kx, const void *v,
                                  const void *V0, const void *V1, void *out,
                                  const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(Nk, B);
  if (prec == 0)
    hipLaunchKernelGGL(k_stokes_combine<double>, grid, dim3(64), 0, STREAM(stream), Nk, N,
                       (const double *)t, (const double *)dt, (const double *)kx,
                       (const double *)v, (const double *)V0, (const double *)V1,
                       (double *)out, active);
  else
    hipLaunchKernelGGL(k_stokes_combine<float>, grid, dim3(64), 0, STREAM(stream), Nk, N,
                       (const float *)t, (const float *)dt, (const float *)kx,
                       (const float *)v, (const float *)V0, (const float *)V1, (float *)out,
                       active);
  CHECK_LAUNCH("smx_stokes_combine");
  return 0;
}

// theta matvec combine (implicit_theta.py:354-383):
// out = (a0 v - (d2v - k^2 v)) - adv ; tau rows from v.
template <typename T>
__global__ void k_theta_combine(int Nk, int N, const T *v, const T *d2v, const T *adv,
                                const T *kx, const T *a0, const T *V0, const T *V1, T *out,
                                const int *active) {
  const int b = blockIdx.y, k = blockIdx.x;
  if (active && !active[b]) return;
  const T kk = kx[k] * kx[k], a = a0[b];
  const long off = ((long)b * Nk + k) * N;
  for (int j = threadIdx.x; j < N - 2; j += blockDim.x) {
    T lap = d2v[off + j] - kk * v[off + j];
    out[off + j] = (a * v[off + j] - lap) - adv[off + j];
  }
  T s0 = 0, s1 = 0;
  for (int j = threadIdx.x; j < N; j += blockDim.x) { s0 += v[off + j] * V0[j]; s1 += v[off + j] * V1[j]; }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  if (threadIdx.x == 0) { out[off + N - 2] = s0; out[off + N - 1] = s1; }
}
extern "C" int smx_theta_combine(int prec, int B, int Nk, int N, const void *v,
                                 const void *d2v, const void *adv, const void *kx,
                                 const void *a0, const void *V0, const void *V1, void *out,
                                 const int *active, void *stream) {
  if (B <= 0) return 0;
  dim3 grid(Nk, B);
  if (prec == 0)
    hipLaunchKernelGGL(k_theta_combine<double>, grid, dim3(64), 0, STREAM(stream), Nk, N,
                       (const double *)v, (const double *)d2v, (const double *)adv,
                       (const double *)kx, (const double *)a0, (const double *)V0,
                       (const double *)V1, (double *)out, active);
  else
    hipLaunchKernelGGL(k_theta_combine<float>, grid, dim3(64), 0, STREAM(stream), Nk, N,
                       (const float *)v, (const float *)d2v, (const float *)adv,
                       (const float *)kx, (const float *)a0, (const float *)V0,
                       (const float *)V1, (float *)out, active);
  CHECK_LAUNCH("smx_theta_combine");
  return 0;
}

// per-run max |x| (written as double)
template <typename T>
__global__ void __launch_bounds__(1024) k_maxabs(long n, const T *x, double *out, const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  double m = 0.0;
  for (long e = threadIdx.x; e < n; e += blockDim.x) m = fmax(m, fabs((double)x[(long)b * n + e]));
  m = block_max(m, red);
  if (threadIdx.x == 0) out[b] = m;
}
extern "C" int smx_maxabs(int prec, int B, long n_per_run, const void *x, double *out,
                          const int *active, void *stream) {
  if (B <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_maxabs<double>, dim3(B), dim3(1024), 0, STREAM(stream), n_per_run,
                       (const double *)x, out, active);
  else
    hipLaunchKernelGGL(k_maxabs<float>, dim3(B), dim3(1024), 0, STREAM(stream), n_per_run,
                       (const float *)x, out, active);
  CHECK_LAUNCH("smx_maxabs");
  return 0;
}

// per-run max over the field of |eta - eta_lag| / max(|eta_lag|, 1e-300): the
// pointwise relative drift of the viscosity since a frozen factorisation
// (coupled_model.py:509-522)
template <typename T>
__global__ void __launch_bounds__(1024) k_eta_drift(long n, const T *eta, const T *lag,
                                                    double *out, const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  double m = 0.0;
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    const double l = (double)lag[(long)b * n + e];
    const double d = fabs((double)eta[(long)b * n + e] - l) / fmax(fabs(l), 1e-300);
    m = fmax(m, d);
  }
  m = block_max(m, red);
  if (threadIdx.x == 0) out[b] = m;
}
extern "C" int smx_eta_drift(int prec, int B, long n_per_run, const void *eta,
                             const void *eta_lag, double *out, const int *active,
                             void *stream) {
  if (B <= 0) return 0;
  if (n_per_run <= 0) return fail_msg("smx_eta_drift: n_per_run must be positive");
  if (prec == 0)
    hipLaunchKernelGGL(k_eta_drift<double>, dim3(B), dim3(1024), 0, STREAM(stream), n_per_run,
                       (const double *)eta, (const double *)eta_lag, out, active);
  else
    hipLaunchKernelGGL(k_eta_drift<float>, dim3(B), dim3(1024), 0, STREAM(stream), n_per_run,
                       (const float *)eta, (const float *)eta_lag, out, active);
  CHECK_LAUNCH("smx_eta_drift");
  return 0;
}

// per-run Euclidean norm of row b of x (row stride ld), accumulated in FP64
template <typename T>
__global__ void __launch_bounds__(1024) k_rownorm(long n, long ld, const T *x, double *out,
                                                  const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  double s = 0.0;
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    const double v = (double)x[(long)b * ld + e];
    s = fma(v, v, s);
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[b] = sqrt(s);
}
extern "C" int smx_rownorm(int prec, int B, long n, long ld, const void *x, double *out,
                           const int *active, void *stream) {
  if (B <= 0) return 0;
  if (n <= 0 || ld < n) return fail_msg("smx_rownorm: need 0 < n <= ld");
  if (prec == 0)
    hipLaunchKernelGGL(k_rownorm<double>, dim3(B), dim3(1024), 0, STREAM(stream), n, ld,
                       (const double *)x, out, active);
  else
    hipLaunchKernelGGL(k_rownorm<float>, dim3(B), dim3(1024), 0, STREAM(stream), n, ld,
                       (const float *)x, out, active);
  CHECK_LAUNCH("smx_rownorm");
  return 0;
}

// per-run weighted sums out[b*ldo] = sum_i w[i] x[b*ldx + i] (y[b*ldy + i]),
// accumulated in FP64: quadratures of the energy budget (volume means of
// theta, of the dissipation field and of u_z theta)
template <typename T>
__global__ void __launch_bounds__(1024) k_wdot(long len, const T *x, long ldx, const T *y,
                                               long ldy, const T *w, double *out, long ldo,
                                               const int *active) {
  const int b = blockIdx.x;
  if (active && !active[b]) return;
  __shared__ double red[16];
  double s = 0.0;
  const T *xb = x + (long)b * ldx;
  const T *yb = y ? y + (long)b * ldy : nullptr;
  for (long e = threadIdx.x; e < len; e += blockDim.x) {
    double v = (double)w[e] * (double)xb[e];
    if (yb) v *= (double)yb[e];
    s += v;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[(long)b * ldo] = s;
}
extern "C" int smx_wdot(int prec, int B, long len, const void *x, long ldx, const void *y,
                        long ldy, const void *w, double *out, long ldo, const int *active,
                        void *stream) {
  if (B <= 0) return 0;
  if (len <= 0 || ldx < len || (y && ldy < len) || ldo < 1)
    return fail_msg("smx_wdot: need 0 < len <= ldx (and <= ldy), ldo >= 1");
  if (prec == 0)
    hipLaunchKernelGGL(k_wdot<double>, dim3(B), dim3(1024), 0, STREAM(stream), len,
                       (const double *)x, ldx, (const double *)y, ldy, (const double *)w, out,
                       ldo, active);
  else
    hipLaunchKernelGGL(k_wdot<float>, dim3(B), dim3(1024), 0, STREAM(stream), len,
                       (const float *)x, ldx, (const float *)y, ldy, (const float *)w, out,
                       ldo, active);
  CHECK_LAUNCH("smx_wdot");
  return 0;
}

// dt = min(dtmax, clip(safety * h / max(umax, 1e-12), dtmin, dtmax));
// params per run: safety, h, dtmin, dtmax
__global__ void k_cfl(int B, const double *umax, const double *params, double *dt,
                      const int *active) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B || (active && !active[b])) return;
  const double *p = params + 4 * b;
  double v = fmax(umax[b], 1e-12);
  double d = p[0] * p[1] / v;
  d = fmin(fmax(d, p[2]), p[3]);
  dt[b] = fmin(p[3], d);
}
extern "C" int smx_cfl_dt(int B, const double *umax, const double *params, double *dt,
                          const int *active, void *stream) {
  if (B <= 0) return 0;
  hipLaunchKernelGGL(k_cfl, dim3(cdiv(B, 256)), dim3(256), 0, STREAM(stream), B, umax, params,
                     dt, active);
  CHECK_LAUNCH("smx_cfl_dt");
  return 0;
}

// ---------------------------------------------------------------------------
// FGMRES state machine (krylov_fgmres.py:60-179), one solver per run.
// ctl: int[10][B] (struct of arrays)
//   0 mode (0 arnoldi, 1 cycle end, 2 true residual, 3 final residual, 4 done)
//   1 j, 2 kend, 3 total, 4 info, 5 act_pre, 6 act_mv, 7 maxiter, 8 cycles
// sc: double[8][B]: 0 bnorm, 1 tol, 2 beta, 3 res (true relative residual),
//   4 last quasi-residual
// H: [B][(m+1)*m] row-major (i, j) -> i*m + j
// ---------------------------------------------------------------------------
enum { FG_ARN = 0, FG_CYC = 1, FG_RES = 2, FG_FIN = 3, FG_DONE = 4 };
#define FG_MAXM 128   // largest restart length (Krylov dimension) supported

template <typename T>
__global__ void __launch_bounds__(1024) k_fg_init(int B, long n, int m, const T *b, T *x, T *Vb,
                                                  T *g, int *ctl, double *sc, double rtol,
                                                  double atol, int maxiter, const int *active) {
  const int r = blockIdx.x;
  __shared__ double red[16];
  if (active && !active[r]) {
    if (threadIdx.x == 0) { ctl[0 * B + r] = FG_DONE; ctl[5 * B + r] = 0; ctl[6 * B + r] = 0; }
    return;
  }
  const T *br = b + (long)r * n;
  double s = 0;
  for (long e = threadIdx.x; e < n; e += blockDim.x) { double v = (double)br[e]; s += v * v; }
  s = block_sum(s, red);
  const double beta = sqrt(s);
  const double tol = fmax(atol, rtol * beta);
  T *xr = x + (long)r * n;
  T *v0 = Vb + (long)r * (m + 1) * n;
  const bool done = (beta <= tol);
  for (long e = threadIdx.x; e < n; e += blockDim.x) {
    xr[e] = T(0);
    if (!done) v0[e] = (T)((double)br[e] / beta);
  }
  T *gr = g + (long)r * (m + 1);
  for (int i = threadIdx.x; i <= m; i += blockDim.x) gr[i] = T(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    gr[0] = (T)beta;
    ctl[0 * B + r] = done ? FG_DONE : FG_ARN;
    ctl[1 * B + r] = 0;
    ctl[2 * B + r] = 0;
    ctl[3 * B + r] = 0;
    ctl[4 * B + r] = done ? 0 : 1;
    ctl[5 * B + r] = 0;
    ctl[6 * B + r] = 0;
    ctl[7 * B + r] = maxiter;
    ctl[8 * B + r] = 0;
    sc[0 * B + r] = beta;
    sc[1 * B + r] = tol;
    sc[2 * B + r] = beta;
    sc[3 * B + r] = done ? 0.0 : 1.0;
    sc[4 * B + r] = beta;
  }
}

extern "C" int smx_fg_init(int prec, int B, long n, int m, const void *b, void *x, void *Vb,
                           void *g, int *ctl, double *sc, double rtol, double atol,
                           int maxiter, const int *active, void *stream) {
  if (B <= 0) return 0;
  if (prec == 0)
    hipLaunchKernelGGL(k_fg_init<double>, dim3(B), dim3(1024), 0, STREAM(stream), B, n, m,
                       (const double *)b, (double *)x, (double *)Vb, (double *)g, ctl, sc,
                       rtol, atol, maxiter, active);
  else
    hipLaunchKernelGGL(k_fg_init<float>, dim3(B), dim3(1024), 0, STREAM(stream), B, n, m,
                       (const float *)b, (float *)x, (float *)Vb, (float *)g, ctl, sc, rtol,
                       atol, maxiter, active);
  CHECK_LAUNCH("smx_fg_init");
  return 0;
}

// cycle end: y = H[:k,:k]^{-1} g[:k] (upper triangular), x += Z^T y; then the
// mode of every run decides the input of this round's operator apply:
// arnoldi -> vin = V_j (preconditioner input); residual -> zout = x.
template <typename T>
__global__ void __launch_bounds__(1024) k_fg_cycle(int B, long n, int m, T *x, const T *Zb,
                                                   const T *Vb, T *H, T *g, T *y, T *vin,
                                                   T *zout, int *ctl) {
  const int r = blockIdx.x;
  __shared__ T ys[1024];
  int mode = ctl[0 * B + r];
  if (mode == FG_CYC) {
    const int k = ctl[2 * B + r];
    T *Hr = H + (long)r * (m + 1) * m;
    const T *gr = g + (long)r * (m + 1);
    if (threadIdx.x == 0) {
      for (int i = 0; i < k; ++i) ys[i] = gr[i];
      for (int c = k - 1; c >= 0; --c) {
        ys[c] = ys[c] / Hr[c * m + c];
        T yc = ys[c];
        for (int i = 0; i < c; ++i) ys[i] -= yc * Hr[i * m + c];
      }
      for (int i = 0; i < k; ++i) y[(long)r * m + i] = ys[i];
    }
    __syncthreads();
    T *xr = x + (long)r * n;
    const T *Zr = Zb + (long)r * m * n;
    for (long e = threadIdx.x; e < n; e += blockDim.x) {
      T s = 0;
      for (int i = 0; i < k; ++i) s = fma(Zr[(long)i * n + e], ys[i], s);
      xr[e] = xr[e] + s;
    }
    const int info = ctl[4 * B + r];
    mode = (info == 0 || info == 2) ? FG_FIN : FG_RES;
    __syncthreads();
    if (threadIdx.x == 0) ctl[0 * B + r] = mode;
  }
  const int act_pre = (mode == FG_ARN);
  const int act_mv = (mode == FG_ARN || mode == FG_RES || mode == FG_FIN);
  if (threadIdx.x == 0) { ctl[5 * B + r] = act_pre; ctl[6 * B + r] = act_mv; }
  if (mode == FG_ARN) {
    const int j = ctl[1 * B + r];
    const T *vj = Vb + ((long)r * (m + 1) + j) * n;
    for (long e = threadIdx.x; e < n; e += blockDim.x) vin[(long)r * n + e] = vj[e];
  } else if (mode == FG_RES || mode == FG_FIN) {
    const T *xr = x + (long)r * n;
    for (long e = threadIdx.x; e < n; e += blockDim.x) zout[(long)r * n + e] = xr[e];
  }
}

extern "C" int smx_fg_cycle_end(int prec, int B, long n, int m, void *x, const void *Zb,
                                const void *Vb, void *H, void *g, void *y, void *vin,
                                void *zout, int *ctl, void *stream) {
  if (B <= 0) return 0;
  if (m > 1024) return fail_msg("smx_fg_cycle_end: restart > 1024");
  if (prec == 0)
    hipLaunchKernelGGL(k_fg_cycle<double>, dim3(B), dim3(1024), 0, STREAM(stream), B, n, m,
                       (double *)x, (const double *)Zb, (const double *)Vb, (double *)H,
                       (double *)g, (double *)y, (double *)vin, (double *)zout, ctl);
  else
    hipLaunchKernelGGL(k_fg_cycle<float>, dim3(B), dim3(1024), 0, STREAM(stream), B, n, m,
                       (float *)x, (const float *)Zb, (const float *)Vb, (float *)H,
                       (float *)g, (float *)y, (float *)vin, (float *)zout, ctl);
  CHECK_LAUNCH("smx_fg_cycle_end");
  return 0;
}

// post: arnoldi -> store Z_j, MGS, Givens, stop tests; residual -> restart
// or finish; final residual -> record ||b - A x|| / ||b||.
// E > 0: the operator output w lives in registers (E elements per thread);
// E == 0: w is updated in place in global memory (long vectors).
template <typename T, int E>
struct WVec {
  T v[E > 0 ? E : 1];
  T *g;
  long n;
  int per, tid, nt;
  __device__ void load(T *w, long n_, int tid_, int nt_) {
    g = w; n = n_; tid = tid_; nt = nt_;
    per = (int)((n + nt - 1) / nt);
    if constexpr (E > 0) {
#pragma unroll
      for (int i = 0; i < E; ++i) {
        long e = tid + (long)i * nt;
        v[i] = (i < per && e < n) ? w[e] : T(0);
      }
    }
  }
  __device__ __forceinline__ T get(int i, long e) const {
    if constexpr (E > 0) return v[i]; else return g[e];
  }
  __device__ __forceinline__ void set(int i, long e, T x) {
    if constexpr (E > 0) v[i] = x; else g[e] = x;
  }
  __device__ __forceinline__ int count() const { return E > 0 ? E : per; }
};

// ORTHO 0: modified Gram-Schmidt in the reference's order (1024 threads);
// 1: block Gram-Schmidt with reorthogonalisation (512 threads; blocks of 64/E
// basis vectors held in registers: one reduction of the block's dot products
// per block, MGS between blocks, everything twice, H = sum of both passes)
template <typename T, int E, int ORTHO>
__global__ void __launch_bounds__(ORTHO ? 512 : 1024) k_fg_post(int B, long n, int m,
                                                                const T *bvec, T *w,
                                                  const T *zin, T *Zb, T *Vb, T *H, T *cs,
                                                  T *sn, T *g, int *ctl, double *sc,
                                                  unsigned long long *units) {
  const int r = blockIdx.x;
  const int mode = ctl[0 * B + r];
  if (mode != FG_ARN && mode != FG_RES && mode != FG_FIN) return;
  // vectors moved: Arnoldi step j reads z, w, V_0..V_j, writes Z_j, V_{j+1};
  // a restart or finish reads the residual and writes V_0
  if (units && threadIdx.x == 0)
    atomicAdd(units, (unsigned long long)(mode == FG_ARN ? ctl[1 * B + r] + 5 : 2));
  __shared__ double red[16];
  __shared__ int s_next;
  const int tid = threadIdx.x, nt = blockDim.x;
  WVec<T, E> W;
  W.load(w + (long)r * n, n, tid, nt);
  const int per = W.per;
  const int cnt = W.count();
  if (mode == FG_ARN) {
    const int j = ctl[1 * B + r];
    T *Vr = Vb + (long)r * (m + 1) * n;
    T *Hr = H + (long)r * (m + 1) * m;
    // column j of H and the previous rotations live in LDS for this step
    __shared__ T hcol[FG_MAXM + 1], scs[FG_MAXM], ssn[FG_MAXM];
    {
      const T *csr = cs + (long)r * m, *snr = sn + (long)r * m;
      for (int i = tid; i < j; i += nt) {
        scs[i] = csr[i];
        ssn[i] = snr[i];
      }
    }
    {
      const T *zr = zin + (long)r * n;
      T *Zj = Zb + ((long)r * m + j) * n;
      for (long e = tid; e < n; e += nt) Zj[e] = zr[e];
    }
    if constexpr (E > 0 && ORTHO == 1) {
      constexpr int CH = 64 / E;
      __shared__ double redc[2][16][CH];
      const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
      int par = 0;
      for (int pass = 0; pass < 2; ++pass) {
        for (int i0 = 0; i0 <= j; i0 += CH) {
          T vv[CH][E];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            const bool okc = (i0 + c <= j);
            const T *vi = Vr + (long)(okc ? i0 + c : 0) * n;
#pragma unroll
            for (int q = 0; q < E; ++q) {
              long e = tid + (long)q * nt;
              const bool ok = okc && q < per && e < n;
              vv[c][q] = vi[ok ? e : 0];
              vv[c][q] = ok ? vv[c][q] : T(0);
            }
          }
          double part[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            double a = 0;
#pragma unroll
            for (int q = 0; q < E; ++q) a += (double)vv[c][q] * (double)W.v[q];
            part[c] = wave_sum(a);
          }
          if (lane == 0) {
#pragma unroll
            for (int c = 0; c < CH; ++c) redc[par][wv][c] = part[c];
          }
          __syncthreads();
          T h[CH];
#pragma unroll
          for (int c = 0; c < CH; ++c) {
            double a = 0;
            for (int u = 0; u < nw; ++u) a += redc[par][u][c];
            h[c] = (T)a;
          }
          par ^= 1;
#pragma unroll
          for (int q = 0; q < E; ++q) {
            T x = W.v[q];
#pragma unroll
            for (int c = 0; c < CH; ++c) x -= h[c] * vv[c][q];
            W.v[q] = x;
          }
          if (tid == 0) {
            for (int c = 0; c < CH && i0 + c <= j; ++c)
              hcol[i0 + c] = pass == 0 ? h[c] : hcol[i0 + c] + h[c];
          }
        }
      }
    } else if constexpr (E > 0) {
      // modified Gram-Schmidt in the reference's order, with the next PF-1
      // basis vectors already loading while h_i is reduced (a register ring
      // of PF vectors: one step's reduction is short next to an HBM load)
      // (E = 8 with 1024 threads has 128 registers: a ring of 3 spills)
      constexpr int PF = E <= 4 ? 6 : (E <= 8 ? 2 : 3);
      __shared__ double red2[32];
      __syncthreads();
      T vb[PF][E];
      auto ld = [&](T *v, int i) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          long e = tid + (long)q * nt;
          v[q] = (i <= j && q < per && e < n) ? Vr[(long)i * n + e] : T(0);
        }
      };
      // entries past n are zero in both W and the loaded v, so the inner
      // product and the update need no per-element guards (the valid terms
      // come first, in the same order)
      auto mgs = [&](const T *v, int i) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < E; ++q) s += (double)v[q] * (double)W.v[q];
        s = block_sum_db(s, red2, i);
        const T h = (T)s;
#pragma unroll
        for (int q = 0; q < E; ++q) W.v[q] = W.v[q] - h * v[q];
        if (tid == 0) hcol[i] = h;
      };
#pragma unroll
      for (int p = 0; p < PF; ++p) ld(vb[p], p);
      for (int i0 = 0; i0 <= j; i0 += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
          if (i0 + p > j) break;
          mgs(vb[p], i0 + p);
          ld(vb[p], i0 + p + PF);
        }
      }
    } else {
      for (int i = 0; i <= j; ++i) {
        const T *vi = Vr + (long)i * n;
        double s = 0;
        for (int q = 0; q < cnt; ++q) {
          long e = tid + (long)q * nt;
          if (q < per && e < n) s += (double)vi[e] * (double)W.get(q, e);
        }
        s = block_sum(s, red);
        const T h = (T)s;
        for (int q = 0; q < cnt; ++q) {
          long e = tid + (long)q * nt;
          if (q < per && e < n) W.set(q, e, W.get(q, e) - h * vi[e]);
        }
        if (tid == 0) hcol[i] = h;
      }
    }
    double s2 = 0;
#pragma unroll
    for (int q = 0; q < cnt; ++q) {
      long e = tid + (long)q * nt;
      if (q < per && e < n) { double x = (double)W.get(q, e); s2 += x * x; }
    }
    s2 = block_sum(s2, red);
    const double hn = sqrt(s2);
    if (tid == 0) {
      T *csr = cs + (long)r * m, *snr = sn + (long)r * m, *gr = g + (long)r * (m + 1);
      hcol[j + 1] = (T)hn;
      // the previous rotations applied to column j, in LDS
      T hi = hcol[0];
      for (int i = 0; i < j; ++i) {
        const T a = hi, bb = hcol[i + 1];
        hcol[i] = scs[i] * a + ssn[i] * bb;
        hi = -ssn[i] * a + scs[i] * bb;
      }
      hcol[j] = hi;
      // Givens (krylov_fgmres.py:42-57) for real entries
      T a = hcol[j], bb = hcol[j + 1];
      T aa = fabs(a), ab = fabs(bb), c, s, rr;
      if (ab == T(0)) { c = T(1); s = T(0); rr = a; }
      else if (aa == T(0)) { c = T(0); s = bb / ab; rr = ab; }
      else {
        T den = sqrt(aa * aa + ab * ab);
        c = aa / den;
        s = c * bb / a;
        rr = c * (aa * aa + ab * ab) / a;
      }
      csr[j] = c;
      snr[j] = s;
      hcol[j] = rr;
      hcol[j + 1] = T(0);
      T t = c * gr[j] + s * gr[j + 1];
      gr[j + 1] = -s * gr[j] + c * gr[j + 1];
      gr[j] = t;
      int total = ctl[3 * B + r] + 1;
      ctl[3 * B + r] = total;
      ctl[2 * B + r] = j + 1;
      double rn = fabs((double)gr[j + 1]);
      sc[4 * B + r] = rn;
      int next;
      if (rn <= sc[1 * B + r]) { ctl[4 * B + r] = 0; next = FG_CYC; }
      else if (hn == 0.0) { ctl[4 * B + r] = 2; next = FG_CYC; }
      else {
        next = FG_ARN;
        ctl[1 * B + r] = j + 1;
        if (j + 1 == m) next = FG_CYC;
        else if (total >= ctl[7 * B + r]) { ctl[4 * B + r] = 1; next = FG_CYC; }
      }
      ctl[0 * B + r] = next;
      s_next = (rn <= sc[1 * B + r] || hn == 0.0) ? 0 : 1;
    }
    __syncthreads();
    for (int i = tid; i <= j + 1; i += nt) Hr[(long)i * m + j] = hcol[i];
    if (s_next) {
      T *vn = Vr + (long)(j + 1) * n;
      const T hv = (T)hn;
#pragma unroll
      for (int q = 0; q < cnt; ++q) {
        long e = tid + (long)q * nt;
        if (q < per && e < n) vn[e] = W.get(q, e) / hv;
      }
    }
    return;
  }
  // residual modes: r = b - w
  const T *br = bvec + (long)r * n;
  double s2 = 0;
#pragma unroll
  for (int q = 0; q < cnt; ++q) {
    long e = tid + (long)q * nt;
    if (q < per && e < n) {
      T x = br[e] - W.get(q, e);
      W.set(q, e, x);
      s2 += (double)x * (double)x;
    }
  }
  s2 = block_sum(s2, red);
  const double beta = sqrt(s2);
  const double bn = fmax(sc[0 * B + r], 1e-300);
  if (mode == FG_FIN) {
    if (tid == 0) { sc[3 * B + r] = beta / bn; ctl[0 * B + r] = FG_DONE; }
    return;
  }
  if (tid == 0) {
    sc[2 * B + r] = beta;
    sc[3 * B + r] = beta / bn;
    int next;
    if (beta <= sc[1 * B + r]) { ctl[4 * B + r] = 0; next = FG_DONE; }
    else if (ctl[3 * B + r] >= ctl[7 * B + r]) next = FG_DONE;
    else next = FG_ARN;
    ctl[0 * B + r] = next;
    s_next = (next == FG_ARN);
  }
  __syncthreads();
  if (s_next) {
    T *v0 = Vb + (long)r * (m + 1) * n;
#pragma unroll
    for (int q = 0; q < cnt; ++q) {
      long e = tid + (long)q * nt;
      if (q < per && e < n) v0[e] = (T)((double)W.get(q, e) / beta);
    }
    T *gr = g + (long)r * (m + 1);
    for (int i = tid; i <= m; i += nt) gr[i] = T(0);
    __syncthreads();
    if (tid == 0) {
      gr[0] = (T)beta;
      ctl[1 * B + r] = 0;
      ctl[8 * B + r] += 1;
    }
  }
}

template <typename T>
static void launch_fg_post(int B, long n, int m, const void *b, void *w, const void *zin,
                           void *Zb, void *Vb, void *H, void *cs, void *sn, void *g, int *ctl,
                           double *sc, int ortho, unsigned long long *units, hipStream_t st) {
#define FGP(E, O)                                                                             \
  hipLaunchKernelGGL((k_fg_post<T, E, O>), dim3(B), dim3(O ? 512 : 1024), 0, st, B, n, m,     \
                     (const T *)b, (T *)w, (const T *)zin, (T *)Zb, (T *)Vb, (T *)H, (T *)cs, \
                     (T *)sn, (T *)g, ctl, sc, units)
  if (n <= 4L * 1024) { if (ortho) FGP(8, 1); else FGP(4, 0); }
  else if (n <= 8L * 1024) { if (ortho) FGP(16, 1); else FGP(8, 0); }
  else FGP(0, 0);
#undef FGP
}

extern "C" int smx_fg_post(int prec, int B, long n, int m, const void *b, void *w,
                           const void *zin, void *Zb, void *Vb, void *H, void *cs, void *sn,
                           void *g, int *ctl, double *sc, int ortho, void *stream) {
  if (B <= 0) return 0;
  if (ortho != 0 && ortho != 1) return fail_msg("smx_fg_post: ortho must be 0 or 1");
  if (m <= 0 || m > FG_MAXM) return fail_msg("smx_fg_post: restart length must be 1..128");
  const bool clk = clk_on(1);
  if (clk) clk_begin(1, STREAM(stream));
  unsigned long long *units = clk_units(1);
  if (prec == 0)
    launch_fg_post<double>(B, n, m, b, w, zin, Zb, Vb, H, cs, sn, g, ctl, sc, ortho, units,
                           STREAM(stream));
  else
    launch_fg_post<float>(B, n, m, b, w, zin, Zb, Vb, H, cs, sn, g, ctl, sc, ortho, units,
                          STREAM(stream));
  if (clk) clk_end(1, STREAM(stream));
  CHECK_LAUNCH("smx_fg_post");
  return 0;
}

__global__ void k_fg_count(int B, const int *ctl, int *out) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int b = threadIdx.x; b < B; b += blockDim.x)
    if (ctl[b] != FG_DONE) atomicAdd(&cnt, 1);
  __syncthreads();
  if (threadIdx.x == 0) out[0] = cnt;
}
extern "C" int smx_fg_count(int B, const int *ctl, int *out, void *stream) {
  hipLaunchKernelGGL(k_fg_count, dim3(1), dim3(256), 0, STREAM(stream), B, ctl, out);
  CHECK_LAUNCH("smx_fg_count");
  return 0;
}
