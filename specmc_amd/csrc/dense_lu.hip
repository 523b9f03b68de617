// dense_lu.hip -- batched dense LU with partial pivoting and its solves, for
// the dense Schur systems of the BFBT rescue route (n = Nk*Nz: 3168 at
// 64x96).  FP64 or FP32 (T), row-major n x n matrices, matrix q at
// A + q*mstride.  The FP32 factor serves the mixed-precision rescue: its
// solves take and return FP64 vectors and the caller's FP64 defect
// correction against the matrix-free operator refines the result.
//
// Blocked right-looking factorisation, block width NB = 64 (FP64) or 128
// (FP32):
//   for each block [j0, j0+jb):
//     for each sub-panel of PB columns [s0, s0+PB):
//       k_lud_panel        factor the m x PB sub-panel (rows s0..n-1) with
//                          partial pivoting; the sub-panel lives in registers
//                          of one 1024-thread workgroup per matrix
//       k_lud_block_update apply its swaps to the block's other columns, the
//                          U rows (unit-lower PB x PB solve) and the rank-PB
//                          update of the block columns to its right
//     k_lud_laswp          the block's swaps on every column outside it
//     k_lud_trsm           U12 = L11^{-1} A12 (in 64-row halves for NB = 128)
//     smx_gemm             A22 -= L21 U12 (K = NB)
// Row swaps are contiguous in row-major storage, so every kernel reads and
// writes whole row segments.  Pivots are global row indices: at step i rows i
// and piv[i] were exchanged (LAPACK getrf order, 0-based).
#include "smx_common.h"
#include "specmc_hip.h"

#define LUD_MAXN 8192         // FP64 (sub-panel rows held in registers)
#define LUD_MAXN_F32 32768    // FP32 (narrower sub-panels; solve vector in LDS)

// ---------------------------------------------------------------------------
// sub-panel factorisation: rows s0 + tid + 1024 k (k < R), columns s0..s0+PB-1
// ---------------------------------------------------------------------------
// FUSE: the block [j0, j0+jb) is finished here too (the FP32 factor): the
// sub-panel's row swaps on the block's other columns, and the U rows of the
// block's columns right of the sub-panel (unit-lower PB x PB solve); the
// rank-PB update below them is then a matrix-core GEMM.
template <typename T, int R, int PB, bool FUSE>
__global__ void __launch_bounds__(1024) k_lud_panel(T *A, long ms, int n, int s0, int *piv,
                                                    int *info, int j0, int jb) {
  const int q = blockIdx.x;
  T *M = A + (long)q * ms;
  int *P = piv + (long)q * n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  __shared__ T rowp[PB], rowc[PB];
  __shared__ T rv[16];
  __shared__ int ri[16];
  __shared__ int s_p;
  __shared__ int spv[PB];
  T v[R][PB];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int r = s0 + tid + k * 1024;
#pragma unroll
    for (int c = 0; c < PB; ++c)
      v[k][c] = (r < n && s0 + c < n) ? M[(long)r * n + s0 + c] : T(0);
  }
  // fully unrolled (no early exit), so every v[k][c] is a register: a loop
  // left rolled indexes v dynamically and puts it in scratch memory
#pragma unroll
  for (int c = 0; c < PB; ++c) {
    const int col = s0 + c;
    if (col >= n) continue;   // uniform: the barriers below are skipped by all
    // pivot: first row of largest |a| at or below the diagonal
    T best = T(-1);
    int bi = n;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const int r = s0 + tid + k * 1024;
      const T a = fabs(v[k][c]);
      if (r >= col && r < n && a > best) {
        best = a;
        bi = r;
      }
    }
    wave_argmax_first(best, bi);
    if (lane == 0) {
      rv[w] = best;
      ri[w] = bi;
    }
    __syncthreads();
    if (w == 0) {
      // the 16 wavefronts' choices (the same total order)
      T b = lane < 16 ? rv[lane] : T(-1);
      int i = lane < 16 ? ri[lane] : 0x7fffffff;
      wave_argmax_first(b, i);
      if (i < col || i >= n) i = col;
      if (lane == 0) {
        s_p = i;
        P[col] = i;
        spv[c] = i;
      }
    }
    __syncthreads();
    const int p = s_p;
    const int oc = col - s0, op = p - s0;
    if (tid == (op & 1023)) {
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (k == (op >> 10)) {
#pragma unroll
          for (int cc = 0; cc < PB; ++cc) rowp[cc] = v[k][cc];
        }
    }
    if (tid == (oc & 1023)) {
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (k == (oc >> 10)) {
#pragma unroll
          for (int cc = 0; cc < PB; ++cc) rowc[cc] = v[k][cc];
        }
    }
    __syncthreads();
    if (p != col) {
      if (tid == (oc & 1023)) {
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (k == (oc >> 10)) {
#pragma unroll
            for (int cc = 0; cc < PB; ++cc) v[k][cc] = rowp[cc];
          }
      }
      if (tid == (op & 1023)) {
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (k == (op >> 10)) {
#pragma unroll
            for (int cc = 0; cc < PB; ++cc) v[k][cc] = rowc[cc];
          }
      }
    }
    const T u = rowp[c];
    if (u != T(0)) {
      const T inv = T(1) / u;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const int r = s0 + tid + k * 1024;
        if (r > col && r < n) {
          const T l = v[k][c] * inv;
          v[k][c] = l;
#pragma unroll
          for (int cc = c + 1; cc < PB; ++cc) v[k][cc] = fma(-l, rowp[cc], v[k][cc]);
        }
      }
    } else if (tid == 0 && info[q] == 0) {
      info[q] = col + 1;
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int r = s0 + tid + k * 1024;
    if (r < n) {
#pragma unroll
      for (int c = 0; c < PB; ++c)
        if (s0 + c < n) M[(long)r * n + s0 + c] = v[k][c];
    }
  }
  if constexpr (FUSE) {
    __shared__ T Ls[PB][PB + 1];
    __shared__ int cand[2 * PB], csrc[2 * PB];
    const int npiv = min(PB, n - s0), subw = min(PB, j0 + jb - s0);
    const int nleft = s0 - j0, nc = jb - subw, ncr = j0 + jb - s0 - subw;
    // the unit-lower L11 of the sub-panel (rows s0.., held by threads 0..npiv-1)
    if (tid < npiv) {
#pragma unroll
      for (int c = 0; c < PB; ++c) Ls[tid][c] = v[0][c];
    }
    // the swaps as one permutation of the rows they touch: position x ends
    // up holding row src(x) = S_0(S_1(...S_{npiv-1}(x)))
    if (tid < 2 * npiv) {
      const int x = tid < npiv ? s0 + tid : spv[tid - npiv];
      int src = x;
      for (int i = npiv - 1; i >= 0; --i) {
        const int a = s0 + i, b = spv[i];
        src = src == a ? b : (src == b ? a : src);
      }
      cand[tid] = x;
      csrc[tid] = src;
    }
    __syncthreads();
    // gather every moved entry of the block's other columns, then scatter
    constexpr int NI = (2 * PB * 128 + 1023) / 1024;
    T val[NI];
    const int items = 2 * npiv * nc;
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int e = tid + u * 1024;
      if (e < items) {
        const int t = e / nc, ci = e - t * nc;
        const int col = ci < nleft ? j0 + ci : s0 + subw + (ci - nleft);
        if (csrc[t] != cand[t]) val[u] = M[(long)csrc[t] * n + col];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      const int e = tid + u * 1024;
      if (e < items) {
        const int t = e / nc, ci = e - t * nc;
        const int col = ci < nleft ? j0 + ci : s0 + subw + (ci - nleft);
        if (csrc[t] != cand[t]) M[(long)cand[t] * n + col] = val[u];
      }
    }
    __syncthreads();
    // U rows of the columns right of the sub-panel: L11^{-1} (their rows s0..)
    if (tid < ncr) {
      const int c = s0 + subw + tid;
      T x[PB];
#pragma unroll
      for (int i = 0; i < PB; ++i) x[i] = (i < npiv) ? M[(long)(s0 + i) * n + c] : T(0);
#pragma unroll
      for (int i = 1; i < PB; ++i) {
        if (i < npiv) {
          T a = x[i];
#pragma unroll
          for (int j = 0; j < i; ++j) a = fma(-Ls[i][j], x[j], a);
          x[i] = a;
        }
      }
#pragma unroll
      for (int i = 0; i < PB; ++i)
        if (i < npiv) M[(long)(s0 + i) * n + c] = x[i];
    }
  }
}

// ---------------------------------------------------------------------------
// after sub-panel s0 of block [j0, j0+jb): one workgroup per PB-column group
// of the block.  Left groups take the swaps; right groups take the swaps, the
// U rows and the rank-PB update of rows s0+PB..n-1.
// ---------------------------------------------------------------------------
template <typename T, int PB>
__global__ void __launch_bounds__(256) k_lud_block_update(T *A, long ms, int n, int j0,
                                                          int jb, int s0, const int *piv) {
  const int q = blockIdx.y;
  T *M = A + (long)q * ms;
  const int *P = piv + (long)q * n;
  const int g0 = j0 + blockIdx.x * PB;
  if (g0 == s0 || g0 >= j0 + jb) return;
  const int ncol = min(PB, j0 + jb - g0);
  const int npiv = min(PB, n - s0);
  const int tid = threadIdx.x;
  __shared__ T u[PB][PB];
  if (tid < ncol) {
    const int c = g0 + tid;
    for (int i = 0; i < npiv; ++i) {
      const int r = s0 + i, pr = P[r];
      if (pr != r) {
        const T t = M[(long)r * n + c];
        M[(long)r * n + c] = M[(long)pr * n + c];
        M[(long)pr * n + c] = t;
      }
    }
  }
  if (g0 < s0) return;
  if (tid < ncol) {
    const int c = g0 + tid;
    T x[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) x[i] = (i < npiv) ? M[(long)(s0 + i) * n + c] : T(0);
#pragma unroll
    for (int i = 1; i < PB; ++i) {
      if (i < npiv) {
        T a = x[i];
#pragma unroll
        for (int j = 0; j < i; ++j) a = fma(-M[(long)(s0 + i) * n + s0 + j], x[j], a);
        x[i] = a;
      }
    }
#pragma unroll
    for (int i = 0; i < PB; ++i)
      if (i < npiv) {
        M[(long)(s0 + i) * n + c] = x[i];
        u[i][tid] = x[i];
      }
  }
  __syncthreads();
  const int cc = tid % PB, rstep = 256 / PB;
  if (cc < ncol) {
    for (int r = s0 + npiv + tid / PB; r < n; r += rstep) {
      const T *Lr = M + (long)r * n + s0;
      T a = M[(long)r * n + g0 + cc];
#pragma unroll
      for (int j = 0; j < PB; ++j)
        if (j < npiv) a = fma(-Lr[j], u[j][cc], a);
      M[(long)r * n + g0 + cc] = a;
    }
  }
}

// the block's row swaps on the columns outside [j0, j0+jb), one thread per column
template <typename T>
__global__ void __launch_bounds__(256) k_lud_laswp(T *A, long ms, int n, int j0, int jb,
                                                   const int *piv) {
  const int q = blockIdx.y;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n - jb) return;
  const int c = t < j0 ? t : t + jb;
  T *M = A + (long)q * ms;
  const int *P = piv + (long)q * n;
  for (int i = j0; i < j0 + jb; ++i) {
    const int pr = P[i];
    if (pr != i) {
      const T x = M[(long)i * n + c];
      M[(long)i * n + c] = M[(long)pr * n + c];
      M[(long)pr * n + c] = x;
    }
  }
}

// U12 = L11^{-1} A12: L11 the unit-lower ib x ib block at (i0, i0) (ib <= 64),
// rows i0..i0+ib-1 of the columns c0..n-1
template <typename T>
__global__ void __launch_bounds__(256) k_lud_trsm(T *A, long ms, int n, int i0, int ib, int c0) {
  const int q = blockIdx.y;
  T *M = A + (long)q * ms;
  __shared__ T L[64][65];
  const int tid = threadIdx.x;
  for (int e = tid; e < ib * ib; e += 256) {
    const int i = e / ib, j = e % ib;
    L[i][j] = M[(long)(i0 + i) * n + i0 + j];
  }
  __syncthreads();
  const int c = c0 + blockIdx.x * 256 + tid;
  if (c >= n) return;
  T x[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) x[i] = (i < ib) ? M[(long)(i0 + i) * n + c] : T(0);
#pragma unroll
  for (int i = 1; i < 64; ++i) {
    if (i < ib) {
      T a = x[i];
#pragma unroll
      for (int j = 0; j < i; ++j) a = fma(-L[i][j], x[j], a);
      x[i] = a;
    }
  }
#pragma unroll
  for (int i = 0; i < 64; ++i)
    if (i < ib) M[(long)(i0 + i) * n + c] = x[i];
}

__global__ void k_lud_zero(int *v, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) v[i] = 0;
}

// NB: the block width, 64 or 128.  A block of 128 solves its U rows in two
// 64-row halves with a rank-64 update between them, and updates the trailing
// matrix with K = 128, which halves the trailing-matrix traffic of the
// factorisation (the update is HBM-bound at K = 64).
template <typename T, int R, int PB, int NB, bool FUSE = false>
static int lud_run(int nmat, int n, T *A, long ms, int *piv, int *info, hipStream_t s) {
  // the fused sub-panel kernel gathers the moved rows of at most 128 block columns
  static_assert(!FUSE || NB <= 128, "fused sub-panels take blocks of at most 128 columns");
  const long lda = n;
  const int prec = sizeof(T) == 8 ? 0 : 1;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = min(NB, n - j0);
    for (int s0 = j0; s0 < j0 + jb; s0 += PB) {
      hipLaunchKernelGGL((k_lud_panel<T, R, PB, FUSE>), dim3(nmat), dim3(1024), 0, s, A, ms, n,
                         s0, piv, info, j0, jb);
      if constexpr (FUSE) {
        // rows below the sub-panel, columns right of it: A -= L21 U12 (K = npiv)
        const int npiv = min(PB, n - s0), c1 = s0 + min(PB, j0 + jb - s0);
        const int mr = n - s0 - npiv, ncr = j0 + jb - c1;
        if (mr > 0 && ncr > 0) {
          int rc = smx_gemm_ex(3, prec, mr, ncr, npiv, -1.0, 1.0,
                               A + (long)(s0 + npiv) * lda + s0, lda, 1, ms, 0,
                               A + (long)s0 * lda + c1, lda, 1, ms, 0,
                               A + (long)(s0 + npiv) * lda + c1, lda, 1, ms, 0,
                               nmat, 1, nullptr, 1, nullptr, 0, 1, (void *)s);
          if (rc) return rc;
        }
      } else {
        hipLaunchKernelGGL((k_lud_block_update<T, PB>), dim3(cdiv(jb, PB), nmat), dim3(256), 0,
                           s, A, ms, n, j0, jb, s0, piv);
      }
    }
    if (n - jb > 0)
      hipLaunchKernelGGL(k_lud_laswp<T>, dim3(cdiv(n - jb, 256), nmat), dim3(256), 0, s, A, ms,
                         n, j0, jb, piv);
    const int c0 = j0 + jb, nr = n - c0;
    if (nr > 0) {
      const int h = min(64, jb);
      hipLaunchKernelGGL(k_lud_trsm<T>, dim3(cdiv(nr, 256), nmat), dim3(256), 0, s, A, ms, n,
                         j0, h, c0);
      if (jb > h) {
        // rows j0+64.. of U12 -= L21' (the block's lower-left part) U12's first 64 rows
        int rc = smx_gemm_ex(3, prec, jb - h, nr, h, -1.0, 1.0,
                             A + (long)(j0 + h) * lda + j0, lda, 1, ms, 0,
                             A + (long)j0 * lda + c0, lda, 1, ms, 0,
                             A + (long)(j0 + h) * lda + c0, lda, 1, ms, 0,
                             nmat, 1, nullptr, 1, nullptr, 0, 1, (void *)s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_lud_trsm<T>, dim3(cdiv(nr, 256), nmat), dim3(256), 0, s, A, ms,
                           n, j0 + h, jb - h, c0);
      }
      CHECK_LAUNCH("smx_dense_lu");
      int rc = smx_gemm_ex(3, prec, nr, nr, jb, -1.0, 1.0,
                           A + (long)c0 * lda + j0, lda, 1, ms, 0,
                           A + (long)j0 * lda + c0, lda, 1, ms, 0,
                           A + (long)c0 * lda + c0, lda, 1, ms, 0,
                           nmat, 1, nullptr, 1, nullptr, 0, 1, (void *)s);
      if (rc) return rc;
    }
  }
  CHECK_LAUNCH("smx_dense_lu");
  return 0;
}

// FP32 factor: 1 (default) the sub-panel kernel also applies the block's
// swaps and solves its U rows, and the rank-16 update is a matrix-core GEMM;
// 0 the separate block-update kernel
static int g_lud_fuse = 1;
extern "C" int smx_set_lud_fuse(int on) {
  g_lud_fuse = on ? 1 : 0;
  return 0;
}

extern "C" int smx_dense_lu(int prec, int nmat, int n, void *A, long mstride, int *piv,
                            int *info, void *stream) {
  if (nmat <= 0) return 0;
  if (prec != 0 && prec != 1) return fail_msg("smx_dense_lu: prec must be 0 (FP64) or 1 (FP32)");
  if (n <= 0 || n > (prec ? LUD_MAXN_F32 : LUD_MAXN) || mstride < (long)n * n)
    return fail_msg("smx_dense_lu: bad sizes (1 <= n <= 8192 FP64 / 32768 FP32, "
                    "mstride >= n*n)");
  hipStream_t s = STREAM(stream);
  hipLaunchKernelGGL(k_lud_zero, dim3(cdiv(nmat, 256)), dim3(256), 0, s, info, nmat);
  if (prec == 1) {
    float *a = (float *)A;
    // 16-column sub-panels: half the launches of the latency-bound panel
    // chain (a batch of a few matrices costs what its launches cost)
    if (g_lud_fuse) {
      if (n <= 1024) return lud_run<float, 1, 16, 128, true>(nmat, n, a, mstride, piv, info, s);
      if (n <= 2048) return lud_run<float, 2, 16, 128, true>(nmat, n, a, mstride, piv, info, s);
      if (n <= 4096) return lud_run<float, 4, 16, 128, true>(nmat, n, a, mstride, piv, info, s);
      if (n <= 8192) return lud_run<float, 8, 4, 128, true>(nmat, n, a, mstride, piv, info, s);
      if (n <= 16384) return lud_run<float, 16, 4, 128, true>(nmat, n, a, mstride, piv, info, s);
      return lud_run<float, 32, 2, 128, true>(nmat, n, a, mstride, piv, info, s);
    }
    if (n <= 1024) return lud_run<float, 1, 16, 128>(nmat, n, a, mstride, piv, info, s);
    if (n <= 2048) return lud_run<float, 2, 16, 128>(nmat, n, a, mstride, piv, info, s);
    if (n <= 4096) return lud_run<float, 4, 16, 128>(nmat, n, a, mstride, piv, info, s);
    if (n <= 8192) return lud_run<float, 8, 4, 128>(nmat, n, a, mstride, piv, info, s);
    if (n <= 16384) return lud_run<float, 16, 4, 128>(nmat, n, a, mstride, piv, info, s);
    return lud_run<float, 32, 2, 128>(nmat, n, a, mstride, piv, info, s);
  }
  double *a = (double *)A;
  if (n <= 1024) return lud_run<double, 1, 8, 64>(nmat, n, a, mstride, piv, info, s);
  if (n <= 2048) return lud_run<double, 2, 8, 64>(nmat, n, a, mstride, piv, info, s);
  if (n <= 4096) return lud_run<double, 4, 8, 64>(nmat, n, a, mstride, piv, info, s);
  return lud_run<double, 8, 4, 64>(nmat, n, a, mstride, piv, info, s);
}

// ---------------------------------------------------------------------------
// solve: y = P b, L y' = y (unit lower), U x = y'; one 1024-thread workgroup
// per matrix with the vector in LDS; 64-row blocks: a matrix-vector product
// against the solved part (four rows per wavefront at a time) and a
// wavefront-serial solve of the diagonal block staged in LDS.  T is the
// factor's type and the arithmetic type; TV the type of b and x.
// ---------------------------------------------------------------------------
// acc[u] += M[row u, j0:j1) . y[j0:j1) for four rows, FP32 factor: each lane
// takes four consecutive columns with one 16-byte load per row (j0, j1 and the
// row starts multiples of 4 floats), so a wavefront keeps four times the
// bytes in flight of a column-per-lane loop
__device__ __forceinline__ void lud_rowdots4(const float *const rows[4], const float *y, int j0,
                                             int j1, int lane, float acc[4]) {
  int j = j0 + 4 * lane;
  // two column chunks per pass: eight 16-byte loads in flight per lane
  for (; j + 256 < j1; j += 512) {
    float4 a[4], c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = *(const float4 *)(rows[u] + j);
      c[u] = *(const float4 *)(rows[u] + j + 256);
    }
    const float4 ya = *(const float4 *)(y + j), yc = *(const float4 *)(y + j + 256);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      acc[u] = fma(a[u].x, ya.x, acc[u]);
      acc[u] = fma(a[u].y, ya.y, acc[u]);
      acc[u] = fma(a[u].z, ya.z, acc[u]);
      acc[u] = fma(a[u].w, ya.w, acc[u]);
      acc[u] = fma(c[u].x, yc.x, acc[u]);
      acc[u] = fma(c[u].y, yc.y, acc[u]);
      acc[u] = fma(c[u].z, yc.z, acc[u]);
      acc[u] = fma(c[u].w, yc.w, acc[u]);
    }
  }
  if (j < j1) {
    const float4 ya = *(const float4 *)(y + j);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float4 a = *(const float4 *)(rows[u] + j);
      acc[u] = fma(a.x, ya.x, acc[u]);
      acc[u] = fma(a.y, ya.y, acc[u]);
      acc[u] = fma(a.z, ya.z, acc[u]);
      acc[u] = fma(a.w, ya.w, acc[u]);
    }
  }
}

// the 64 x 64 diagonal block of rows/columns [i0, i0 + bs): four entries per
// thread of the 1024, loaded to registers, then stored to LDS (ld 65)
template <typename T>
__device__ __forceinline__ void lud_db_load(const T *M, int n, int i0, int bs, int tid,
                                            T dbr[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = tid + 1024 * k, r = e >> 6, c = e & 63;
    dbr[k] = (r < bs && c < bs) ? M[(long)(i0 + r) * n + i0 + c] : T(0);
  }
}
template <typename T>
__device__ __forceinline__ void lud_db_store(T *Db, const T dbr[4], int bs, int tid) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = tid + 1024 * k, r = e >> 6, c = e & 63;
    if (r < bs && c < bs) Db[r * 65 + c] = dbr[k];
  }
}

// one wavefront solves the diagonal block held in LDS (Db, ld 65) for the
// 64 entries of yb: unit lower (forward) or upper (backward).  Lane l keeps
// row l of the block in registers, 16 columns at a time, so the serial
// column loop reads no LDS; lanes at or past bs carry unused values.
template <typename T>
__device__ __forceinline__ void lud_diag_lower(const T *Db, T *yb, int bs, int lane) {
  T yv = lane < bs ? yb[lane] : T(0);
#pragma unroll
  for (int t0 = 0; t0 < 64; t0 += 16) {
    T r[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) r[k] = Db[lane * 65 + t0 + k];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int t = t0 + k;
      if (t < bs) {
        const T yt = __shfl(yv, t, 64);
        if (lane > t) yv = fma(-r[k], yt, yv);
      }
    }
  }
  if (lane < bs) yb[lane] = yv;
}
template <typename T>
__device__ __forceinline__ void lud_diag_upper(const T *Db, T *yb, int bs, int lane) {
  T yv = lane < bs ? yb[lane] : T(0);
#pragma unroll
  for (int t0 = 48; t0 >= 0; t0 -= 16) {
    T r[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) r[k] = Db[lane * 65 + t0 + k];
#pragma unroll
    for (int k = 15; k >= 0; --k) {
      const int t = t0 + k;
      if (t < bs) {
        if (lane == t) yv = yv / r[k];
        const T yt = __shfl(yv, t, 64);
        if (lane < t) yv = fma(-r[k], yt, yv);
      }
    }
  }
  if (lane < bs) yb[lane] = yv;
}

template <typename T, typename TV>
__global__ void __launch_bounds__(1024) k_lud_solve(const T *A, long ms, int n,
                                                    const int *piv, const TV *b, TV *x,
                                                    const int *active,
                                                    unsigned long long *units) {
  extern __shared__ unsigned char smem_raw[];
  T *y = (T *)smem_raw;
  T *Db = y + n;                            // [64][65]
  const int q = blockIdx.x;
  if (active && !active[q]) return;
  if (units && threadIdx.x == 0) atomicAdd(units, 1ull);
  const T *M = A + (long)q * ms;
  // FP32 factors with rows of a multiple of 4 floats (and 16-byte aligned
  // matrices) take the 16-byte row loads
  const bool vec4 = sizeof(T) == 4 && (n % 4) == 0 && (ms % 4) == 0 &&
                    (((unsigned long)A) & 15) == 0;
  const int *P = piv + (long)q * n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int i = tid; i < n; i += 1024) y[i] = (T)b[(long)q * n + i];
  __syncthreads();
  if (tid == 0)
    for (int i = 0; i < n; ++i) {
      const int p = P[i];
      if (p != i && p >= 0 && p < n) {
        const T t = y[i];
        y[i] = y[p];
        y[p] = t;
      }
    }
  __syncthreads();
  const int nblk = (n + 63) / 64;
  for (int bk = 0; bk < nblk; ++bk) {
    const int i0 = bk * 64, bs = min(64, n - i0);
    // the diagonal block's loads go out before the row products, so their
    // latency hides behind them
    T dbr[4];
    lud_db_load(M, n, i0, bs, tid, dbr);
    // y[i0 + t] -= L[i0 + t, 0:i0] . y[0:i0]
    {
      T acc[4] = {T(0), T(0), T(0), T(0)};
      const T *rows[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) rows[u] = M + (long)(i0 + min(w + 16 * u, bs - 1)) * n;
      if constexpr (sizeof(T) == 4) {
        if (vec4) lud_rowdots4(rows, y, 0, i0, lane, acc);
      }
      if (!vec4)
        for (int j = lane; j < i0; j += 64) {
          const T yj = y[j];
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[u] = fma(rows[u][j], yj, acc[u]);
        }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const T a = wave_sum(acc[u]);
        if (lane == 0 && w + 16 * u < bs) y[i0 + w + 16 * u] -= a;
      }
    }
    lud_db_store(Db, dbr, bs, tid);
    __syncthreads();
    if (w == 0) lud_diag_lower(Db, y + i0, bs, lane);
    __syncthreads();
  }
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int i0 = bk * 64, bs = min(64, n - i0), j1 = i0 + bs;
    T dbr[4];
    lud_db_load(M, n, i0, bs, tid, dbr);
    {
      T acc[4] = {T(0), T(0), T(0), T(0)};
      const T *rows[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) rows[u] = M + (long)(i0 + min(w + 16 * u, bs - 1)) * n;
      if constexpr (sizeof(T) == 4) {
        if (vec4) lud_rowdots4(rows, y, j1, n, lane, acc);
      }
      if (!vec4)
        for (int j = j1 + lane; j < n; j += 64) {
          const T yj = y[j];
#pragma unroll
          for (int u = 0; u < 4; ++u) acc[u] = fma(rows[u][j], yj, acc[u]);
        }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const T a = wave_sum(acc[u]);
        if (lane == 0 && w + 16 * u < bs) y[i0 + w + 16 * u] -= a;
      }
    }
    lud_db_store(Db, dbr, bs, tid);
    __syncthreads();
    if (w == 0) lud_diag_upper(Db, y + i0, bs, lane);
    __syncthreads();
  }
  for (int i = tid; i < n; i += 1024) x[(long)q * n + i] = (TV)y[i];
}

// ---------------------------------------------------------------------------
// wide solve for a few matrices: the one-workgroup-per-matrix solve streams a
// whole factor through one CU, so with fewer matrices than CUs the triangular
// sweeps run block by block over many workgroups instead: x = P b (swaps
// replayed on an index array in LDS), then per 64-row block a one-wavefront
// diagonal solve and a right-looking update of every remaining row, spread
// over workgroups of 64 rows.  Vectors and accumulation in TV.
// ---------------------------------------------------------------------------
#define LUDW_CHUNK 2048
template <typename TV>
__global__ void __launch_bounds__(1024) k_ludw_perm(int n, const int *piv, const TV *b, TV *x,
                                                    const int *active,
                                                    unsigned long long *units) {
  extern __shared__ int perm[];                 // [n] + [LUDW_CHUNK]
  int *pc = perm + n;
  const int q = blockIdx.x;
  if (active && !active[q]) return;
  if (units && threadIdx.x == 0) atomicAdd(units, 1ull);
  const int *P = piv + (long)q * n;
  for (int i = threadIdx.x; i < n; i += 1024) perm[i] = i;
  for (int c0 = 0; c0 < n; c0 += LUDW_CHUNK) {
    const int c1 = min(n, c0 + LUDW_CHUNK);
    __syncthreads();
    for (int i = c0 + threadIdx.x; i < c1; i += 1024) pc[i - c0] = P[i];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = c0; i < c1; ++i) {
        const int p = pc[i - c0];
        if (p != i && p >= 0 && p < n) {
          const int t = perm[i];
          perm[i] = perm[p];
          perm[p] = t;
        }
      }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 1024) x[(long)q * n + i] = b[(long)q * n + perm[i]];
}

// diagonal block [i0, i0 + bs) of matrix blockIdx.x: unit lower (upper = 0)
// or upper with its diagonal (upper = 1), one wavefront
template <typename T, typename TV>
__global__ void __launch_bounds__(64) k_ludw_diag(const T *A, long ms, int n, TV *x, int i0,
                                                  int upper, const int *active) {
  __shared__ TV Db[64][65];
  const int q = blockIdx.x;
  if (active && !active[q]) return;
  const T *M = A + (long)q * ms;
  TV *y = x + (long)q * n;
  const int bs = min(64, n - i0), lane = threadIdx.x;
  for (int r = 0; r < bs; ++r)
    if (lane < bs) Db[r][lane] = (TV)M[(long)(i0 + r) * n + i0 + lane];
  __syncthreads();
  TV yv = lane < bs ? y[i0 + lane] : TV(0);
  if (!upper) {
    for (int t = 0; t < bs; ++t) {
      const TV yt = __shfl(yv, t, 64);
      if (lane > t && lane < bs) yv = fma(-Db[lane][t], yt, yv);
    }
  } else {
    for (int t = bs - 1; t >= 0; --t) {
      if (lane == t) yv = yv / Db[t][t];
      const TV yt = __shfl(yv, t, 64);
      if (lane < t) yv = fma(-Db[lane][t], yt, yv);
    }
  }
  if (lane < bs) y[i0 + lane] = yv;
}

// y[i] -= M[i, i0:i0+bs) . y[i0:i0+bs) for rows rlo <= i < rhi; 64 rows per
// workgroup (16 per wavefront), grid.y = matrix
template <typename T, typename TV>
__global__ void __launch_bounds__(256) k_ludw_upd(const T *A, long ms, int n, TV *x, int i0,
                                                  int bs, int rlo, int rhi, const int *active) {
  __shared__ TV yb[64];
  const int q = blockIdx.y;
  if (active && !active[q]) return;
  const T *M = A + (long)q * ms;
  TV *y = x + (long)q * n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid < bs) yb[tid] = y[i0 + tid];
  __syncthreads();
  const TV yl = lane < bs ? yb[lane] : TV(0);
  const int base = rlo + blockIdx.x * 64 + w * 16;
#pragma unroll 4
  for (int u = 0; u < 16; ++u) {
    const int i = base + u;
    if (i < rhi) {
      TV a = lane < bs ? (TV)M[(long)i * n + i0 + lane] * yl : TV(0);
      a = wave_sum(a);
      if (lane == 0) y[i] -= a;
    }
  }
}

template <typename T, typename TV>
static int ludw_solve(int nmat, int n, const T *LU, long ms, const int *piv, const TV *b, TV *x,
                      const int *active, hipStream_t st) {
  const size_t sh = (size_t)(n + LUDW_CHUNK) * sizeof(int);
  if (!lds_optin(k_ludw_perm<TV>, sh)) return fail_msg("smx_dense_lu_solve: n exceeds LDS");
  hipLaunchKernelGGL(k_ludw_perm<TV>, dim3(nmat), dim3(1024), sh, st, n, piv, b, x, active,
                     clk_units(3));
  const int nblk = (n + 63) / 64;
  for (int bk = 0; bk < nblk; ++bk) {
    const int i0 = bk * 64, bs = min(64, n - i0), r0 = i0 + bs;
    hipLaunchKernelGGL((k_ludw_diag<T, TV>), dim3(nmat), dim3(64), 0, st, LU, ms, n, x, i0, 0,
                       active);
    if (r0 < n)
      hipLaunchKernelGGL((k_ludw_upd<T, TV>), dim3(cdiv(n - r0, 64), nmat), dim3(256), 0, st,
                         LU, ms, n, x, i0, bs, r0, n, active);
  }
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int i0 = bk * 64, bs = min(64, n - i0);
    hipLaunchKernelGGL((k_ludw_diag<T, TV>), dim3(nmat), dim3(64), 0, st, LU, ms, n, x, i0, 1,
                       active);
    if (i0 > 0)
      hipLaunchKernelGGL((k_ludw_upd<T, TV>), dim3(cdiv(i0, 64), nmat), dim3(256), 0, st, LU,
                         ms, n, x, i0, bs, 0, i0, active);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// multi-workgroup solve (smx_dense_lu_solve_mw).  The batched solve above
// streams a whole factor through one CU, so when a masked FGMRES round leaves
// only some of the matrices active, most CUs idle.  Here each matrix's two
// sweeps are split into tasks of 64 rows: task t < nb is forward block t,
// task t >= nb backward block 2 nb - 1 - t (nb = ceil(n / 64)).  A task
// accumulates the products of its rows with the blocks already solved
// (left-looking), then solves its diagonal block in one wavefront.  It needs
// results of lower tasks of its own matrix only.
//
// Progress: a persistent grid (no more workgroups than are resident at once)
// takes tickets from a counter, one per task, a new one only after finishing
// the last, and works on (task, matrix) = (ticket / cnt, list[ticket % cnt])
// over the cnt active matrices.  Tasks wait only for lower tickets, all of
// which are held by running workgroups or done, so the lowest unfinished task
// can always run; every workgroup leaves once the tickets run out.  Each wait
// is also bounded (LUDM_SPIN polls), after which the solve sets the error
// word and goes on.
//
// Results pass between workgroups as 64-bit words (value bits | stage << 32)
// stored and loaded with agent-scope atomics, so no fences are needed: stage
// 1 is the forward result y, stage 2 the solution x, which overwrites y once
// every forward task has finished (the last forward task reads every y and
// the first backward task waits for it).  The workspace is cleared before
// each solve.  FP64 work types take two words per entry (low and high half).
// Arithmetic in the factor's type T, as in k_lud_solve.
// ---------------------------------------------------------------------------
#define LUDM_SPIN (1 << 22)
#define LUDM_HDR 64            // bytes: ticket, count, error word
#define LUDM_UNROLL 2          // column blocks loaded per wait in the products

static inline long ludm_list_bytes(int nmat) { return ((long)nmat * 4 + 255) / 256 * 256; }
static inline long ludm_ws_bytes(int prec, int nmat, int n) {
  return LUDM_HDR + ludm_list_bytes(nmat) + (long)nmat * n * (prec == 0 ? 16 : 8);
}

// perm_q[i] = the row of b that row i of P_q b takes: the swaps of piv
// replayed in order on an index array in LDS (one workgroup per matrix)
__global__ void __launch_bounds__(1024) k_lud_perm(int n, const int *piv, int *perm,
                                                   const int *active) {
  extern __shared__ int pidx[];                 // [n] + [LUDW_CHUNK]
  int *pc = pidx + n;
  const int q = blockIdx.x;
  if (active && !active[q]) return;
  const int *P = piv + (long)q * n;
  for (int i = threadIdx.x; i < n; i += 1024) pidx[i] = i;
  for (int c0 = 0; c0 < n; c0 += LUDW_CHUNK) {
    const int c1 = min(n, c0 + LUDW_CHUNK);
    __syncthreads();
    for (int i = c0 + threadIdx.x; i < c1; i += 1024) pc[i - c0] = P[i];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = c0; i < c1; ++i) {
        const int p = pc[i - c0];
        if (p != i && p >= 0 && p < n) {
          const int t = pidx[i];
          pidx[i] = pidx[p];
          pidx[p] = t;
        }
      }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 1024) perm[(long)q * n + i] = pidx[i];
}

// clear the workspace: header, the compact list of active matrices (block 0,
// one wavefront, in matrix order) and every stage word
__global__ void __launch_bounds__(256) k_ludm_prep(int nmat, long nwords, const int *active,
                                                   unsigned *hdr, int *list,
                                                   unsigned long long *tags,
                                                   unsigned long long *units) {
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int base = 0;
    for (int q0 = 0; q0 < nmat; q0 += 64) {
      const int q = q0 + lane;
      const bool on = q < nmat && (!active || active[q]);
      const unsigned long long m = __ballot(on);
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      if (on) list[base + before] = q;
      base += __popcll(m);
    }
    if (lane == 0) {
      hdr[0] = 0u;
      hdr[1] = (unsigned)base;
      hdr[2] = 0u;
      if (units) atomicAdd(units, (unsigned long long)base);
    }
  }
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < nwords; e += (long)gridDim.x * 256)
    tags[e] = 0ull;
}

template <typename T> struct LudmWord;
template <> struct LudmWord<float> {
  static constexpr int W = 1;
  static __device__ __forceinline__ void put(unsigned long long *tg, int i, float v,
                                             unsigned stage) {
    __hip_atomic_store(tg + i, ((unsigned long long)stage << 32) | __float_as_uint(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ bool get(unsigned long long *tg, int i, unsigned stage,
                                             float &v) {
    const unsigned long long w =
        __hip_atomic_load(tg + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __uint_as_float((unsigned)w);
    return (unsigned)(w >> 32) == stage;
  }
  static __device__ __forceinline__ float bcast(float v, int t) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), t));
  }
};
template <> struct LudmWord<double> {
  static constexpr int W = 2;
  static __device__ __forceinline__ void put(unsigned long long *tg, int i, double v,
                                             unsigned stage) {
    const unsigned long long s = (unsigned long long)stage << 32;
    __hip_atomic_store(tg + 2 * i, s | (unsigned)__double2loint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(tg + 2 * i + 1, s | (unsigned)__double2hiint(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  static __device__ __forceinline__ bool get(unsigned long long *tg, int i, unsigned stage,
                                             double &v) {
    const unsigned long long lo =
        __hip_atomic_load(tg + 2 * i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long hi =
        __hip_atomic_load(tg + 2 * i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v = __hiloint2double((int)(unsigned)hi, (int)(unsigned)lo);
    return (unsigned)(lo >> 32) == stage && (unsigned)(hi >> 32) == stage;
  }
  static __device__ __forceinline__ double bcast(double v, int t) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), t);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), t);
    return __hiloint2double(hi, lo);
  }
};

// entry i of the stage words once it holds ``stage`` (lanes with need = false
// take 0 and do not wait); one bounded wait per wavefront
template <typename T>
__device__ __forceinline__ T ludm_wait(unsigned long long *tg, int i, bool need,
                                       unsigned stage, unsigned *err) {
  T v = T(0);
  bool ok = !need;
  for (int it = 0;; ++it) {
    if (!ok) ok = LudmWord<T>::get(tg, i, stage, v);
    if (__all(ok)) break;
    if (it >= LUDM_SPIN) {
      if ((threadIdx.x & 63) == 0) atomicOr(err, 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return ok ? v : T(0);
}

// wavefront sum by DPP: row_shr 1, 2, 4, 8 sum each row of 16 lanes into its
// lane 15, row_bcast 15 and 31 add the rows into lane 63, which is read out
// (a fixed order; no LDS traffic, unlike the shuffles of wave_sum)
template <int CTRL, int RMASK>
__device__ __forceinline__ float ludm_dpp_add(float v) {
  return v + __int_as_float(
                 __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RMASK, 0xf, false));
}
template <int CTRL, int RMASK>
__device__ __forceinline__ double ludm_dpp_add(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, RMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, RMASK, 0xf, false);
  return v + __hiloint2double(hi, lo);
}
template <typename T>
__device__ __forceinline__ T ludm_wave_sum(T v) {
  v = ludm_dpp_add<0x111, 0xf>(v);   // row_shr:1
  v = ludm_dpp_add<0x112, 0xf>(v);   // row_shr:2
  v = ludm_dpp_add<0x114, 0xf>(v);   // row_shr:4
  v = ludm_dpp_add<0x118, 0xf>(v);   // row_shr:8
  v = ludm_dpp_add<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v = ludm_dpp_add<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
  return LudmWord<T>::bcast(v, 63);
}

// one task: ticket tk of the cnt active matrices.  LDS: Db the diagonal
// block, Dl the block of the last dependency (ld 65 each), sred the row sums.
// The four wavefronts multiply the rows with every solved block but the one
// solved last, which is usually ready long before; wavefront 0 alone then
// waits for that last block, multiplies it (a row per lane, from LDS) and
// solves the diagonal block, so one wavefront sits on the dependency chain.
template <typename T, typename TV>
__device__ __forceinline__ void ludm_task(const T *A, long ms, int n, const int *perm,
                                          const TV *b, TV *x, unsigned *hdr, const int *list,
                                          unsigned long long *tags, int cnt, int nb, int tk,
                                          T *Db, T *Dl, T *sred) {
  typedef LudmWord<T> Wd;
  // the wavefront index is uniform: row pointers below live in scalar registers
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int task = tk / cnt, q = list[tk - task * cnt];
  const bool fwd = task < nb;
  const int blk = fwd ? task : 2 * nb - 1 - task;
  const int i0 = blk * 64, bs = min(64, n - i0);
  const T *M = A + (long)q * ms;
  unsigned long long *tg = tags + (long)q * n * Wd::W;
  unsigned *err = hdr + 2;
  // solved column blocks this task multiplies: forward 0 .. blk-1 (y, stage
  // 1), backward nb-1 down to blk+1 (x, stage 2); the last of them is klast
  const int nk = fwd ? blk : nb - 1 - blk;
  const unsigned stage = fwd ? 1u : 2u;
  const int klast = fwd ? blk - 1 : blk + 1;

  // wavefront 0's right-hand side: (P b) rows of the block (forward)
  T rhs = T(0);
  if (fwd && w == 0 && lane < bs) rhs = (T)b[(long)q * n + perm[(long)q * n + i0 + lane]];
  // the diagonal block and the last dependency's block into LDS (a few rows
  // in flight per thread: the task's products follow and hide the rest)
#pragma unroll 2
  for (int j = 0; j < 16; ++j) {
    const int e = tid + 256 * j, r = e >> 6, c = e & 63;
    Db[r * 65 + c] = (r < bs && c < bs) ? M[(long)(i0 + r) * n + i0 + c] : T(0);
    const int cl = klast * 64 + c;
    Dl[r * 65 + c] = (nk > 0 && r < bs && cl < n) ? M[(long)(i0 + r) * n + cl] : T(0);
  }
  // products of rows i0 + 16 w + u with the solved blocks before klast
  T acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = T(0);
  const T *rows[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) rows[u] = M + (long)(i0 + min(16 * w + u, bs - 1)) * n;
  const int nkb = nk > 0 ? nk - 1 : 0;
  for (int s = 0; s < nkb; s += LUDM_UNROLL) {
    T a[LUDM_UNROLL][16];
    int col[LUDM_UNROLL];
    bool cok[LUDM_UNROLL];
#pragma unroll
    for (int v = 0; v < LUDM_UNROLL; ++v) {
      const int k = fwd ? s + v : nb - 1 - (s + v);
      col[v] = k * 64 + lane;
      cok[v] = (s + v < nkb) && col[v] < n;
      const int cc = cok[v] ? col[v] : 0;
#pragma unroll
      for (int u = 0; u < 16; ++u) a[v][u] = cok[v] ? rows[u][cc] : T(0);
    }
#pragma unroll
    for (int v = 0; v < LUDM_UNROLL; ++v) {
      if (s + v >= nkb) break;
      const T yk = ludm_wait<T>(tg, cok[v] ? col[v] : 0, cok[v], stage, err);
#pragma unroll
      for (int u = 0; u < 16; ++u) acc[u] = fma(a[v][u], yk, acc[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const T sv = ludm_wave_sum(acc[u]);
    if (lane == 0) sred[16 * w + u] = sv;
  }
  __syncthreads();
  if (w != 0) return;     // wavefront 0 alone finishes the block
  if (!fwd) rhs = ludm_wait<T>(tg, i0 + lane, lane < bs, 1u, err);
  T yv = lane < bs ? rhs - sred[lane] : T(0);
  if (nk > 0) {
    // the last dependency: y (or x) of block klast, then row lane of Dl
    // against it (four partial sums)
    const int cl = klast * 64 + lane;
    const T yl = ludm_wait<T>(tg, cl < n ? cl : 0, cl < n, stage, err);
    T p[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 1
    for (int c0 = 0; c0 < 64; c0 += 16) {
      T d[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) d[k] = Dl[lane * 65 + c0 + k];
#pragma unroll
      for (int k = 0; k < 16; ++k) p[k & 3] = fma(d[k], Wd::bcast(yl, c0 + k), p[k & 3]);
    }
    yv -= (p[0] + p[1]) + (p[2] + p[3]);
  }
  // the serial column steps carry no branches: every multiplier is loaded
  // unconditionally and masked (backward: scaled by the reciprocal diagonal)
  // by arithmetic beforehand
  if (fwd) {
    // unit lower diagonal block
#pragma unroll 1
    for (int t0 = 0; t0 < 64; t0 += 16) {
      T r[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const T dv = Db[lane * 65 + t0 + k];
        r[k] = dv * (lane > t0 + k ? T(1) : T(0));
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int t = t0 + k;
        if (t < bs) yv = fma(-r[k], Wd::bcast(yv, t), yv);
      }
    }
    if (lane < bs) Wd::put(tg, i0 + lane, yv, 1u);
  } else if constexpr (sizeof(T) == 4) {
    // upper diagonal block: row l divided by its diagonal, then a unit sweep
    const T dd = Db[lane * 65 + lane];
    const T rd = lane < bs ? T(1) / dd : T(0);
    yv *= rd;
#pragma unroll 1
    for (int t0 = 48; t0 >= 0; t0 -= 16) {
      T r[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const T dv = Db[lane * 65 + t0 + k];
        r[k] = dv * (lane < t0 + k ? rd : T(0));
      }
#pragma unroll
      for (int k = 15; k >= 0; --k) {
        const int t = t0 + k;
        if (t < bs) yv = fma(-r[k], Wd::bcast(yv, t), yv);
      }
    }
    if (lane < bs) {
      Wd::put(tg, i0 + lane, yv, 2u);
      x[(long)q * n + i0 + lane] = (TV)yv;
    }
  } else {
    // FP64: each entry divided by its diagonal when its column is reached
    // (the arithmetic of the batched solve; FP64 solves serve the defect
    // correction, whose stopping point follows the solve's rounding)
#pragma unroll 1
    for (int t0 = 48; t0 >= 0; t0 -= 16) {
      T r[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) r[k] = Db[lane * 65 + t0 + k];
#pragma unroll
      for (int k = 15; k >= 0; --k) {
        const int t = t0 + k;
        if (t < bs) {
          const T xt = Wd::bcast(yv, t) / Wd::bcast(r[k], t);
          yv = lane == t ? xt : (lane < t ? fma(-r[k], xt, yv) : yv);
        }
      }
    }
    if (lane < bs) {
      Wd::put(tg, i0 + lane, yv, 2u);
      x[(long)q * n + i0 + lane] = (TV)yv;
    }
  }
}

// persistent: at most as many workgroups as are resident at once, each
// taking tickets until none are left (a workgroup takes its next ticket only
// after finishing its task, so every ticket below the lowest unfinished one is
// done and that task can run)
template <typename T, typename TV>
__global__ void __launch_bounds__(256) k_ludm_solve(const T *A, long ms, int n,
                                                    const int *perm, const TV *b, TV *x,
                                                    unsigned *hdr, const int *list,
                                                    unsigned long long *tags) {
  __shared__ T Db[64 * 65], Dl[64 * 65];
  __shared__ T sred[64];
  __shared__ int s_tk;
  const int cnt = (int)hdr[1];
  const int nb = (n + 63) >> 6;
  const int total = cnt * 2 * nb;
  for (;;) {
    __syncthreads();        // the previous task's LDS reads are done
    if (threadIdx.x == 0) s_tk = (int)atomicAdd(&hdr[0], 1u);
    __syncthreads();
    const int tk = s_tk;
    if (tk >= total) break;
    ludm_task<T, TV>(A, ms, n, perm, b, x, hdr, list, tags, cnt, nb, tk, Db, Dl, sred);
  }
}

// workgroups of k_ludm_solve<T, TV> resident on the device at once
template <typename T, typename TV>
static int ludm_resident() {
  static int cached = 0;
  if (!cached) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ludm_solve<T, TV>, 256, 0) !=
            hipSuccess ||
        cus <= 0 || per <= 0)
      return 1024;
    cached = cus * per;
  }
  return cached;
}

extern "C" long smx_dense_lu_solve_mw_bytes(int prec, int nmat, int n) {
  if (nmat <= 0 || n <= 0 || prec < 0 || prec > 2) return 0;
  return ludm_ws_bytes(prec, nmat, n);
}

extern "C" int smx_dense_lu_perm(int nmat, int n, const int *piv, int *perm, const int *active,
                                 void *stream) {
  if (nmat <= 0) return 0;
  if (n <= 0 || n > LUD_MAXN_F32) return fail_msg("smx_dense_lu_perm: bad order");
  const size_t sh = (size_t)(n + LUDW_CHUNK) * sizeof(int);
  if (!lds_optin(k_lud_perm, sh)) return fail_msg("smx_dense_lu_perm: n exceeds LDS");
  hipLaunchKernelGGL(k_lud_perm, dim3(nmat), dim3(1024), sh, STREAM(stream), n, piv, perm,
                     active);
  CHECK_LAUNCH("smx_dense_lu_perm");
  return 0;
}

extern "C" int smx_dense_lu_solve_mw(int prec, int nmat, int n, const void *LU, long mstride,
                                     const int *perm, const void *b, void *x,
                                     const int *active, void *ws, long ws_bytes,
                                     void *stream) {
  if (nmat <= 0) return 0;
  if (prec < 0 || prec > 2)
    return fail_msg("smx_dense_lu_solve_mw: prec must be 0 (FP64), 1 (FP32) or 2 (FP32 "
                    "factor, FP64 vectors)");
  if (n <= 0 || n > (prec ? LUD_MAXN_F32 : LUD_MAXN) || mstride < (long)n * n)
    return fail_msg("smx_dense_lu_solve_mw: bad sizes (1 <= n <= 8192 FP64 / 32768 FP32, "
                    "mstride >= n*n)");
  if (!ws || ws_bytes < ludm_ws_bytes(prec, nmat, n))
    return fail_msg("smx_dense_lu_solve_mw: workspace smaller than "
                    "smx_dense_lu_solve_mw_bytes(prec, nmat, n)");
  const size_t vs = prec == 1 ? 4 : 8;
  const char *bb = (const char *)b, *xb = (const char *)x;
  if (bb < xb + (size_t)nmat * n * vs && xb < bb + (size_t)nmat * n * vs)
    return fail_msg("smx_dense_lu_solve_mw: b and x overlap");
  const int nb = (n + 63) / 64;
  if ((long)nmat * 2 * nb > 0x7fffffffL) return fail_msg("smx_dense_lu_solve_mw: too many tasks");
  hipStream_t st = STREAM(stream);
  unsigned *hdr = (unsigned *)ws;
  int *list = (int *)((char *)ws + LUDM_HDR);
  unsigned long long *tags =
      (unsigned long long *)((char *)ws + LUDM_HDR + ludm_list_bytes(nmat));
  const long nwords = (long)nmat * n * (prec == 0 ? 2 : 1);
  const bool clk = clk_on(3);
  if (clk) clk_begin(3, st);
  hipLaunchKernelGGL(k_ludm_prep, dim3(min(1024, cdiv(nwords, 256 * 8))), dim3(256), 0, st,
                     nmat, nwords, active, hdr, list, tags, clk_units(3));
  const int tasks = nmat * 2 * nb;
  if (prec == 0)
    hipLaunchKernelGGL((k_ludm_solve<double, double>),
                       dim3(min(tasks, ludm_resident<double, double>())), dim3(256), 0, st,
                       (const double *)LU, mstride, n, perm, (const double *)b, (double *)x,
                       hdr, list, tags);
  else if (prec == 1)
    hipLaunchKernelGGL((k_ludm_solve<float, float>),
                       dim3(min(tasks, ludm_resident<float, float>())), dim3(256), 0, st,
                       (const float *)LU, mstride, n, perm, (const float *)b, (float *)x, hdr,
                       list, tags);
  else
    hipLaunchKernelGGL((k_ludm_solve<float, double>),
                       dim3(min(tasks, ludm_resident<float, double>())), dim3(256), 0, st,
                       (const float *)LU, mstride, n, perm, (const double *)b, (double *)x,
                       hdr, list, tags);
  if (clk) clk_end(3, st);
  CHECK_LAUNCH("smx_dense_lu_solve_mw");
  return 0;
}

// matrices at most this many and order at least LUDW_MIN_N take the wide solve
#define LUDW_MAX_MATS 16
#define LUDW_MIN_N 1024

extern "C" int smx_dense_lu_solve(int prec, int nmat, int n, const void *LU, long mstride,
                                  const int *piv, const void *b, void *x, const int *active,
                                  void *stream) {
  if (nmat <= 0) return 0;
  if (prec < 0 || prec > 2)
    return fail_msg("smx_dense_lu_solve: prec must be 0 (FP64), 1 (FP32) or 2 (FP32 factor, "
                    "FP64 vectors)");
  if (n <= 0 || n > (prec ? LUD_MAXN_F32 : LUD_MAXN) || mstride < (long)n * n)
    return fail_msg("smx_dense_lu_solve: bad sizes (1 <= n <= 8192 FP64 / 32768 FP32, "
                    "mstride >= n*n)");
  const size_t es = prec == 0 ? 8 : 4;
  const size_t sh = (size_t)n * es + 64 * 65 * es;
  hipStream_t st = STREAM(stream);
  const bool clk = clk_on(3);
  if (clk) clk_begin(3, st);
  unsigned long long *units = clk_units(3);
  if (nmat <= LUDW_MAX_MATS && n >= LUDW_MIN_N && b != x) {
    int rc;
    if (prec == 0)
      rc = ludw_solve<double, double>(nmat, n, (const double *)LU, mstride, piv,
                                      (const double *)b, (double *)x, active, st);
    else if (prec == 1)
      rc = ludw_solve<float, float>(nmat, n, (const float *)LU, mstride, piv, (const float *)b,
                                    (float *)x, active, st);
    else
      rc = ludw_solve<float, double>(nmat, n, (const float *)LU, mstride, piv,
                                     (const double *)b, (double *)x, active, st);
    if (clk) clk_end(3, st);
    if (rc) return rc;
    CHECK_LAUNCH("smx_dense_lu_solve");
    return 0;
  }
  if (prec == 0) {
    if (!lds_optin(k_lud_solve<double, double>, sh))
      return fail_msg("smx_dense_lu_solve: vector exceeds LDS");
    hipLaunchKernelGGL((k_lud_solve<double, double>), dim3(nmat), dim3(1024), sh, st,
                       (const double *)LU, mstride, n, piv, (const double *)b, (double *)x,
                       active, units);
  } else if (prec == 1) {
    if (!lds_optin(k_lud_solve<float, float>, sh))
      return fail_msg("smx_dense_lu_solve: vector exceeds LDS");
    hipLaunchKernelGGL((k_lud_solve<float, float>), dim3(nmat), dim3(1024), sh, st,
                       (const float *)LU, mstride, n, piv, (const float *)b, (float *)x,
                       active, units);
  } else {
    if (!lds_optin(k_lud_solve<float, double>, sh))
      return fail_msg("smx_dense_lu_solve: vector exceeds LDS");
    hipLaunchKernelGGL((k_lud_solve<float, double>), dim3(nmat), dim3(1024), sh, st,
                       (const float *)LU, mstride, n, piv, (const double *)b, (double *)x,
                       active, units);
  }
  if (clk) clk_end(3, st);
  CHECK_LAUNCH("smx_dense_lu_solve");
  return 0;
}
