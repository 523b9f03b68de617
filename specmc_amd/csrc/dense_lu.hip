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
template <typename T, int R, int PB>
__global__ void __launch_bounds__(1024) k_lud_panel(T *A, long ms, int n, int s0, int *piv,
                                                    int *info) {
  const int q = blockIdx.x;
  T *M = A + (long)q * ms;
  int *P = piv + (long)q * n;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  __shared__ T rowp[PB], rowc[PB];
  __shared__ T rv[16];
  __shared__ int ri[16];
  __shared__ int s_p;
  T v[R][PB];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int r = s0 + tid + k * 1024;
#pragma unroll
    for (int c = 0; c < PB; ++c)
      v[k][c] = (r < n && s0 + c < n) ? M[(long)r * n + s0 + c] : T(0);
  }
#pragma unroll
  for (int c = 0; c < PB; ++c) {
    const int col = s0 + c;
    if (col >= n) break;
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
    for (int o = 32; o > 0; o >>= 1) {
      const T ob = __shfl_xor(best, o, 64);
      const int oi = __shfl_xor(bi, o, 64);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      rv[w] = best;
      ri[w] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      T b = rv[0];
      int i = ri[0];
      for (int u = 1; u < 16; ++u)
        if (rv[u] > b || (rv[u] == b && ri[u] < i)) {
          b = rv[u];
          i = ri[u];
        }
      if (i < col || i >= n) i = col;
      s_p = i;
      P[col] = i;
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
template <typename T, int R, int PB, int NB>
static int lud_run(int nmat, int n, T *A, long ms, int *piv, int *info, hipStream_t s) {
  const long lda = n;
  const int prec = sizeof(T) == 8 ? 0 : 1;
  for (int j0 = 0; j0 < n; j0 += NB) {
    const int jb = min(NB, n - j0);
    for (int s0 = j0; s0 < j0 + jb; s0 += PB) {
      hipLaunchKernelGGL((k_lud_panel<T, R, PB>), dim3(nmat), dim3(1024), 0, s, A, ms, n, s0,
                         piv, info);
      hipLaunchKernelGGL((k_lud_block_update<T, PB>), dim3(cdiv(jb, PB), nmat), dim3(256), 0, s,
                         A, ms, n, j0, jb, s0, piv);
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
