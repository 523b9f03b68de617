// smx_common.h -- helpers shared by the kernel sources of libspecmc_hip.so:
// error reporting behind smx_last_error() and block-wide reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

extern thread_local char g_err[512];

static inline int fail(const char *where, hipError_t e) {
  snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
  return (int)e == 0 ? -1 : (int)e;
}
static inline int fail_msg(const char *msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return -1;
}
#define CHECK_LAUNCH(name)                                                   \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess) return fail(name, _e);                             \
  } while (0)
#define STREAM(s) ((hipStream_t)(s))

// launch clock (specmc_kernels.hip): per-tag HIP events around a launch and
// the device counter of the units it processes (nullptr when off)
bool clk_on(int tag);
void clk_begin(int tag, hipStream_t st);
void clk_end(int tag, hipStream_t st);
unsigned long long *clk_units(int tag);

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// block reductions (wavefront shuffles, then LDS across wavefronts)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
// sum over the block; every thread receives the result
template <typename T>
__device__ T block_sum(T v, T *red) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}
// sum over the block with one barrier: successive calls alternate between
// the two halves of red2 (2 x 16 entries), so a call never overwrites the
// partials the previous call's readers may still be reading
template <typename T>
__device__ T block_sum_db(T v, T *red2, int parity) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  T *red = red2 + 16 * (parity & 1);
  v = wave_sum(v);
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}
template <typename T>
__device__ T block_max(T v, T *red) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = red[0];
  for (int i = 1; i < nw; ++i) s = fmax(s, red[i]);
  return s;
}
template <typename T>
__device__ T block_min(T v, T *red) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_min(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = red[0];
  for (int i = 1; i < nw; ++i) s = fmin(s, red[i]);
  return s;
}

// (best, bi) <- the first largest value over the wavefront: the largest
// value, the smallest index among equal values.  That is a total order, so
// any reduction order gives the same pair: DPP row shifts 1, 2, 4, 8 carry
// each row's choice into its lane 15, row broadcasts 15 and 31 carry them
// into lane 63, which is read out to every lane (no LDS traffic, unlike
// shuffles).  Lanes a DPP move leaves out see (-1, INT_MAX), which never
// wins, so candidates must be >= -1 (pivot searches use |a| or -1).
template <int CTRL, int RMASK>
__device__ __forceinline__ void argmax_step(double &best, int &bi) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(best), CTRL, RMASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)0xbff00000, __double2hiint(best), CTRL, RMASK,
                                             0xf, false);
  const int oi = __builtin_amdgcn_update_dpp(0x7fffffff, bi, CTRL, RMASK, 0xf, false);
  const double ov = __hiloint2double(hi, lo);
  if (ov > best || (ov == best && oi < bi)) {
    best = ov;
    bi = oi;
  }
}
template <int CTRL, int RMASK>
__device__ __forceinline__ void argmax_step(float &best, int &bi) {
  const float ov = __int_as_float(__builtin_amdgcn_update_dpp(
      (int)0xbf800000, __float_as_int(best), CTRL, RMASK, 0xf, false));
  const int oi = __builtin_amdgcn_update_dpp(0x7fffffff, bi, CTRL, RMASK, 0xf, false);
  if (ov > best || (ov == best && oi < bi)) {
    best = ov;
    bi = oi;
  }
}
__device__ __forceinline__ double readlane_t(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ float readlane_t(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <typename T>
__device__ __forceinline__ void wave_argmax_first(T &best, int &bi) {
  argmax_step<0x111, 0xf>(best, bi);   // row_shr:1
  argmax_step<0x112, 0xf>(best, bi);   // row_shr:2
  argmax_step<0x114, 0xf>(best, bi);   // row_shr:4
  argmax_step<0x118, 0xf>(best, bi);   // row_shr:8
  argmax_step<0x142, 0xa>(best, bi);   // row_bcast:15 into rows 1 and 3
  argmax_step<0x143, 0xc>(best, bi);   // row_bcast:31 into rows 2 and 3
  best = readlane_t(best, 63);
  bi = __builtin_amdgcn_readlane(bi, 63);
}

// Tile of this workgroup in the (N tiles x M tiles) plane of one batch entry.
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2: the launch order would put the N tiles of one row block on different
// XCDs, and each XCD would fetch that row block of A from HBM again.  The
// remap hands each XCD a contiguous range of tiles (x fastest), so the N
// tiles of a row block share one XCD and its L2.  A bijection for any
// tile count; it changes only which workgroup computes a tile.
__device__ __forceinline__ void xcd_tile(int &ti, int &tj) {
  const int gx = gridDim.x, T = gx * gridDim.y;
  const int L = blockIdx.x + gx * blockIdx.y;
  const int per = T >> 3, rem = T & 7, xcd = L & 7, idx = L >> 3;
  const int t = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
  ti = t / gx;
  tj = t - ti * gx;
}

// matrix-core operand types: v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32.
// Lane l supplies A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16];
// accumulator element r of lane l is C[row(l / 16, r)][l % 16].
typedef double v4f64 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<double> {
  typedef v4f64 V;
  static __device__ __forceinline__ V step(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator element r for lane-row group lk
  static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};
template <> struct MfmaAcc<float> {
  typedef v4f32 V;
  static __device__ __forceinline__ V step(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// dynamic LDS beyond 64 KB needs a per-kernel opt-in
static inline size_t lds_limit() {
  static size_t lim = 0;
  if (!lim) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        v > 0)
      lim = (size_t)v;
    else
      lim = 64 * 1024;
  }
  return lim;
}

// the attribute is set once per kernel for the largest size requested so
// far (the call is not a stream operation, and repeating it per launch would
// also repeat it inside graph captures)
template <typename K>
static inline bool lds_optin(K kern, size_t bytes) {
  if (bytes <= 64 * 1024) return true;
  if (bytes > lds_limit()) return false;
  static thread_local const void *seen[64];
  static thread_local size_t seen_bytes[64];
  static thread_local int nseen = 0;
  const void *k = (const void *)kern;
  for (int i = 0; i < nseen; ++i)
    if (seen[i] == k && seen_bytes[i] >= bytes) return true;
  if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
      hipSuccess)
    return false;
  for (int i = 0; i < nseen; ++i)
    if (seen[i] == k) {
      seen_bytes[i] = bytes;
      return true;
    }
  if (nseen < 64) {
    seen[nseen] = k;
    seen_bytes[nseen++] = bytes;
  }
  return true;
}

