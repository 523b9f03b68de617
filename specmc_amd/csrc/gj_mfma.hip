// gj_mfma.hip -- FP64 Gauss-Jordan inversion of many small blocks, blocked
// by 16 columns, with the off-panel updates on the matrix cores
// (v_mfma_f64_16x16x4f64).  It serves smx_invert for FP64 blocks of order
// 16..128: the per-mode Helmholtz and BFBT blocks whose inverses the FGMRES
// preconditioners apply (inv(A) r instead of two triangular sweeps).
//
// Column-major N x N matrices, NT wavefronts per matrix (two matrices per
// workgroup, below), padded to NP = 16 NT with an identity block.  Wavefront w holds column
// block w (NP x 16) in the accumulator layout of v_mfma_f64_16x16x4f64:
// element i of tile t of lane l is row 16 t + l / 16 + 4 i, column
// 16 w + l % 16 (the FP64 layout measured on gfx950).
//
// No rows are swapped.  Step k factors column block k (the panel) by
// unblocked Gauss-Jordan in wavefront k, the panel rows spread one per lane:
// for column c the pivot p is the largest |a_rc| over the rows r not yet
// pivotal (ties, and values equal to 2^-44 relative, go to the first row);
// every other row subtracts f = a_rc / a_pc times row p and takes -f in
// column c; row p becomes a_pq / a_pc, with 1 / a_pc in column c.  Those 16 steps touch the other column blocks only through their 16
// pivot rows: with X the pivot rows of a block (in pivot order, before the
// step) and P the transformed panel, the block becomes (its non-pivot rows)
// + P X, a K = 16 product on the matrix cores.  After the last step, stored
// row R, column Q is entry (pinv[R], p[Q]) of the inverse, where p[c] is the
// pivot row of column c and pinv its inverse permutation.
//
// The inverse agrees with that of k_gj_invert_reg to roundoff, not bit for
// bit: the multipliers are formed per row and the off-panel entries take the
// 16 steps as one product.
#include "smx_common.h"

// the largest value over the wavefront: DPP row shifts 1, 2, 4, 8 carry
// each row's maximum into its lane 15, row broadcasts 15 and 31 carry them
// into lane 63, which every lane reads.  A lane whose shift source lies
// outside its row reads 0 (bound_ctrl); a lane of a row the broadcast leaves
// out keeps its own value.  Neither changes the maximum.
template <int CTRL, int RMASK>
__device__ __forceinline__ void max_u64_step(unsigned long long &b) {
  const int l0 = (int)(unsigned)b, h0 = (int)(unsigned)(b >> 32);
  const unsigned lo = __builtin_amdgcn_update_dpp(l0, l0, CTRL, RMASK, 0xf, true);
  const unsigned hi = __builtin_amdgcn_update_dpp(h0, h0, CTRL, RMASK, 0xf, true);
  const unsigned long long o = ((unsigned long long)hi << 32) | lo;
  b = o > b ? o : b;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long b) {
  max_u64_step<0x111, 0xf>(b);
  max_u64_step<0x112, 0xf>(b);
  max_u64_step<0x114, 0xf>(b);
  max_u64_step<0x118, 0xf>(b);
  max_u64_step<0x142, 0xa>(b);
  max_u64_step<0x143, 0xc>(b);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, 63);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return ((unsigned long long)hi << 32) | lo;
}

// Two matrices per workgroup (2 NT wavefronts, NT per matrix): wavefronts k
// and NT + k hold the panels of step k, and a workgroup's wavefronts go to
// the four SIMDs in turn, so for NT = 6 the two panels, the serial part of
// every step, run on different SIMDs (two one-matrix workgroups on a CU
// often share one: 2.13 ms for 8448 blocks of 96 against 3.16 ms with one
// such workgroup per CU, not half).  Each matrix has its own LDS; a half
// without a matrix (the last workgroup of an odd count, or an inactive run)
// skips the work but takes every barrier.
#define GJM_MPW 2
template <int NT>
__global__ void __launch_bounds__(64 * NT * GJM_MPW) k_gj_invert_mfma(double *Am, int nmat,
                                                                      int N, int per_run,
                                                                      const int *active) {
  constexpr int NP = 16 * NT, RPL = (NP + 63) / 64, PS = 17;
  static_assert(RPL <= 2, "panel rows: at most two per lane");
  constexpr int LDS_D = 2 * NP * PS + NT * 256;        // doubles of one matrix's region
  constexpr int LDS_I = 2 * NP + (3 * NP + 3) / 4;     // ints after them
  extern __shared__ double gjm_smem[];
  const int tid = threadIdx.x, lane = tid & 63, wg = tid >> 6;
  const int half = wg / NT, w = wg - half * NT;        // matrix slot, wavefront in it
  double *base = gjm_smem + half * (LDS_D + (LDS_I + 1) / 2);
  double *pn = base;                                   // [2][NP * PS] panels, row-major
  double *xs = pn + 2 * NP * PS;                       // [NT][16 * 16] pivot rows
  int *pv = (int *)(xs + NT * 256);                    // [NP] pivot row of each column
  int *pinv = pv + NP;                                 // [NP] column of each pivot row
  signed char *pos = (signed char *)(pinv + NP);       // [2][NP] place among the step's pivots
  unsigned char *used = (unsigned char *)(pos + 2 * NP);   // [NP] row already pivotal
  const int m = blockIdx.x * GJM_MPW + half;
  const bool live = m < nmat && !(active && !active[m / per_run]);
  double *Gm = Am + (long)(live ? m : 0) * N * N;
  const int ltid = tid - half * 64 * NT;               // thread index within the matrix
  const int lr = lane & 15, lg = lane >> 4;
  const int Q = 16 * w + lr;
  v4f64 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int R = 16 * t + lg + 4 * i;
      acc[t][i] = (live && R < N && Q < N) ? Gm[(long)Q * N + R] : (R == Q ? 1.0 : 0.0);
    }
  for (int e = ltid; e < NP; e += 64 * NT) used[e] = 0;
  __syncthreads();
  for (int k = 0; k < NT; ++k) {
    double *P = pn + (k & 1) * NP * PS;
    signed char *ps = pos + (k & 1) * NP;
    if (live && w == k) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) P[(16 * t + lg + 4 * i) * PS + lr] = acc[t][i];
      __builtin_amdgcn_wave_barrier();
      double v[RPL][16];
      bool us[RPL];
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int R = lane + 64 * r;
        us[r] = R >= NP || used[R];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[r][j] = R < NP ? P[R * PS + j] : 0.0;
      }
      int pj[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        // pivot: the largest key over the unused rows; the key is |a_rc|
        // with its low 8 mantissa bits replaced by 255 - r, so the largest
        // |a_rc| wins (to 2^-44 relative) and then the first row
        unsigned long long best = 0;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          double a = fabs(v[r][j]);
          a = a == a ? a : 0.0;   // a NaN never keeps an unused row from being chosen
          const unsigned long long key =
              ((unsigned long long)__double_as_longlong(a) & ~0xffull) |
              (unsigned long long)(255 - (lane + 64 * r));
          best = (!us[r] && key > best) ? key : best;
        }
        const int p = 255 - (int)(wave_max_u64(best) & 0xff);   // the same in every lane
        pj[j] = p;
        const int pl = p & 63, pr = p >> 6;
        // the pivot row, read out of its lane (uniform values); pr is uniform,
        // so only the slot that holds the row is read
        double rp[16];
        if (RPL == 1 || pr == 0) {
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) rp[jj] = readlane_t(v[0][jj], pl);
        } else {
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) {
            rp[jj] = readlane_t(v[1][jj], pl);
#pragma unroll
            for (int r = 2; r < RPL; ++r) {
              const double x = readlane_t(v[r][jj], pl);
              rp[jj] = pr == r ? x : rp[jj];
            }
          }
        }
        const double inv = 1.0 / rp[j];
        // every row: multiplier f = a_rc / a_pc, a_rq -= f a_pq, column c
        // takes -f; the pivot row is then rewritten as a_pq / a_pc with
        // 1 / a_pc in column c
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
          const double f = v[r][j] * inv;
#pragma unroll
          for (int jj = 0; jj < 16; ++jj) v[r][jj] = jj == j ? -f : fma(-f, rp[jj], v[r][jj]);
          us[r] = us[r] || lane + 64 * r == p;
        }
#pragma unroll
        for (int r = 0; r < RPL; ++r)
          if (pr == r && lane == pl) {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) v[r][jj] = jj == j ? inv : rp[jj] * inv;
          }
      }
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
        const int R = lane + 64 * r;
        if (R < NP) {
#pragma unroll
          for (int j = 0; j < 16; ++j) P[R * PS + j] = v[r][j];
          int b = -1;
#pragma unroll
          for (int j = 0; j < 16; ++j) b = R == pj[j] ? j : b;
          ps[R] = (signed char)b;
          if (us[r]) used[R] = 1;
        }
      }
      int pp = pj[0];
#pragma unroll
      for (int j = 1; j < 16; ++j) pp = lane == j ? pj[j] : pp;
      if (lane < 16) pv[16 * k + lane] = pp;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[t][i] = P[(16 * t + lg + 4 * i) * PS + lr];
    }
    __syncthreads();
    if (live && w != k) {
      // this block's pivot rows to X (zeroed in place), then + P X
      double *X = xs + w * 256;
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int b = ps[16 * t + lg + 4 * i];
          if (b >= 0) {
            X[b * 16 + lr] = acc[t][i];
            acc[t][i] = 0.0;
          }
        }
      __builtin_amdgcn_wave_barrier();
      double bx[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) bx[s] = X[(4 * s + lg) * 16 + lr];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s)
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(P[(16 * t + lr) * PS + 4 * s + lg],
                                                        bx[s], acc[t], 0, 0, 0);
    }
  }
  __syncthreads();
  if (live)
    for (int e = ltid; e < NP; e += 64 * NT) pinv[pv[e]] = e;
  __syncthreads();
  if (live && Q < N) {
    const long oq = (long)pv[Q] * N;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int R = 16 * t + lg + 4 * i;
        if (R < N) Gm[oq + pinv[R]] = acc[t][i];
      }
  }
}

template <int NT>
static size_t gjm_lds() {
  constexpr int NP = 16 * NT;
  constexpr int LDS_D = 2 * NP * 17 + NT * 256, LDS_I = 2 * NP + (3 * NP + 3) / 4;
  return (size_t)GJM_MPW * (LDS_D + (LDS_I + 1) / 2) * 8;
}

template <int NT>
static int gjm_launch(int nmat, int N, double *A, int per_run, const int *active,
                      hipStream_t st) {
  const size_t sh = gjm_lds<NT>();
  if (!lds_optin(k_gj_invert_mfma<NT>, sh)) return fail_msg("smx_invert: LDS opt-in failed");
  hipLaunchKernelGGL(k_gj_invert_mfma<NT>, dim3(cdiv(nmat, GJM_MPW)), dim3(64 * NT * GJM_MPW),
                     sh, st, A, nmat, N, per_run, active);
  return 0;
}

// FP64 inversion of nmat column-major N x N blocks, 16 <= N <= 128 (above,
// the panel rows and the column blocks no longer fit the registers together)
int gj_invert_mfma(int nmat, int N, double *A, int per_run, const int *active, hipStream_t st) {
  const int nt = (N + 15) / 16;
  switch (nt) {
    case 1: return gjm_launch<1>(nmat, N, A, per_run, active, st);
    case 2: return gjm_launch<2>(nmat, N, A, per_run, active, st);
    case 3: return gjm_launch<3>(nmat, N, A, per_run, active, st);
    case 4: return gjm_launch<4>(nmat, N, A, per_run, active, st);
    case 5: return gjm_launch<5>(nmat, N, A, per_run, active, st);
    case 6: return gjm_launch<6>(nmat, N, A, per_run, active, st);
    case 7: return gjm_launch<7>(nmat, N, A, per_run, active, st);
    case 8: return gjm_launch<8>(nmat, N, A, per_run, active, st);
  }
  return fail_msg("smx_invert: the matrix-core inversion takes 16 <= N <= 128");
}
