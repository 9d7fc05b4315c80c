// gs_device.hpp -- device-side helpers shared by the HIP kernel files of
// the picker (geometry of a tile size, wave64 reductions, the CTF).
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "gs_fft.hpp"
#include "gs_kernels.hpp"

namespace gs {

constexpr int WG = 256;
constexpr float kRefPi = 3.141592654f;  // the reference's PI (GPU_func.cuh:29)

constexpr int pow2floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <int L>
struct Geo {
  static constexpr int M = L / 2;
  static constexpr int H = M + 1;
  static constexpr int HP = (H + 15) / 16 * 16;
  static constexpr int RB = clampi(pow2floor(8192 / M), 8, 32);
  static constexpr int CB = clampi(pow2floor(8192 / L), 8, 32);
  // correlation column blocks pair column k with column M - k: block cb holds
  // the pair units p = cb * CB/2 + u (u < CB/2) in slots u (column p) and
  // CB-1-u (column M - p), p = 0 .. M/2
  static constexpr int NCB = (M / 2 + 1 + CB / 2 - 1) / (CB / 2);
  static_assert(L % 4 == 0, "tile size divisible by 4");
};

// column of slot c of paired block cb
template <int L>
__device__ __forceinline__ int pair_col(int cb, int c) {
  constexpr int M = L / 2, CB = Geo<L>::CB;
  const int u = c < CB / 2 ? c : CB - 1 - c;
  const int p = cb * (CB / 2) + u;
  return c < CB / 2 ? p : M - p;
}
// slot offset of column k (0 <= k < M) in its paired block: (cb * L + y) * CB + c
template <int L>
__device__ __forceinline__ int pair_slot(int k, int y) {
  constexpr int M = L / 2, CB = Geo<L>::CB;
  const int p = min(k, M - k);
  const int u = p % (CB / 2), cb = p / (CB / 2);
  const int c = (k <= M / 2) ? u : CB - 1 - u;
  return (cb * L + y) * CB + c;
}

// streaming stores of the correlation intermediate (written once by pass 1,
// read once by pass 2), so they do not displace the spectra in L2
typedef float nt_float2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(float2* p, float2 v) {
  nt_float2 t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, reinterpret_cast<nt_float2*>(p));
}

// device-coherent accesses (agent-scope relaxed atomics: stores written
// through to the coherence point, loads that do not hit stale L2 lines), for
// an intermediate that a workgroup on another XCD consumes within the same
// launch
__device__ __forceinline__ void store_coh(float2* p, float2 v) {
  const unsigned long long u =
      ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), u, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float2 load_coh(const float2* p) {
  const unsigned long long u = __hip_atomic_load(
      reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return make_float2(__uint_as_float((unsigned)u), __uint_as_float((unsigned)(u >> 32)));
}

// value of lane (lane ^ (CB - 1)): the partner slot of a paired block when
// CB consecutive lanes hold the CB slots of one row
template <int CB>
__device__ __forceinline__ float xchg_f(float v) {
  if constexpr (CB == 8) {  // row_half_mirror
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, false));
  } else if constexpr (CB == 16) {  // row_mirror
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xF, 0xF, false));
  } else {
    return __shfl_xor(v, CB - 1, 64);
  }
}
template <int CB>
__device__ __forceinline__ float2 xchg(float2 v) {
  return make_float2(xchg_f<CB>(v.x), xchg_f<CB>(v.y));
}

// ---------------------------------------------------------------------------
// wave / block reductions (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// (value, index) max with lowest index on ties
__device__ __forceinline__ void vmax_combine(float& v, int& p, float v2, int p2) {
  if (v2 > v || (v2 == v && p2 < p)) { v = v2; p = p2; }
}
__device__ __forceinline__ void wave_max_idx(float& v, int& p) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    float v2 = __shfl_xor(v, o, 64);
    int p2 = __shfl_xor(p, o, 64);
    vmax_combine(v, p, v2, p2);
  }
}

// block-wide FP64 sum of two values; result valid in thread 0
__device__ __forceinline__ void block_sum2_d(double& a, double& b, double* scratch) {
  a = wave_sum_d(a);
  b = wave_sum_d(b);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) { scratch[2 * w] = a; scratch[2 * w + 1] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) { sa += scratch[2 * i]; sb += scratch[2 * i + 1]; }
    a = sa; b = sb;
  }
}


__device__ __forceinline__ float ctf_ast_f(int x1, int y1, int nx, int ny, const WeightArgs& w, int mode) {
  // CTF_AST, GPU_func.cu:292-315
  y1 = x1 > nx / 2 ? ny - y1 : y1;
  x1 = min(x1, nx - x1);
  const float dx = (float)min(x1, nx - x1);
  const float dy = (float)(y1 - ny / 2);
  const float ss = (dx * dx / (float)(nx * nx) + dy * dy / (float)(ny * ny)) / (w.apix * w.apix);
  const float ag = atan2f((float)(y1 - ny / 2), (float)x1);
  const float df_ast = 0.5f * (w.dfu + w.dfv + 2 * w.dfdiff * cosf(2 * (w.dfang * kRefPi / 180 - ag)));
  const float gamma = -2 * kRefPi * (w.cs * 2.5e6f * w.lambda * w.lambda * w.lambda * ss * ss +
                                     df_ast * 5000.0f * w.lambda * ss);
  const float v = sqrtf(1.0f - w.ampconst * w.ampconst) * sinf(gamma) + w.ampconst * cosf(gamma);
  if (mode == 0) return v > 0 ? 1.0f : -1.0f;
  if (mode == 2) return fabsf(v);
  return v;
}


// The refinement's rotation sample at output pixel (x, y): bilinear value of
// C2R pre-processing of a half-spectrum row:
// Z'[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) e^{+2 pi i k / L}
__device__ __forceinline__ float2 c2r_pre(float2 xk, float2 xmk, float2 wk_conj) {
  float2 s = make_float2(xk.x + xmk.x, xk.y - xmk.y);
  float2 d = make_float2(xk.x - xmk.x, xk.y + xmk.y);
  float2 t = cmul(d, wk_conj);
  return make_float2(s.x - t.y, s.y + t.x);
}

// src at the rotation of (x - L/2, y - L/2) by cs = (cos, sin), coordinates
// in double precision, the boundary rules of rotate_subIMG
// (GPU_func.cu:496-560); also the slopes d/dx2, d/dy2 and the mask factor of
// apply_mask (GPU_func.cu:229-247).
struct RotSampleD {
  float tb, dtx, dty, mask;
  float X, Y;
  bool in;
};
template <int L>
__device__ __forceinline__ RotSampleD rot_sample_d(const float* src, int x, int y, double2 cs,
                                                   float dmask) {
  RotSampleD o;
  const float xf = (float)(x - L / 2), yf = (float)(y - L / 2);
  o.X = xf;
  o.Y = yf;
  const float r = hypotf(xf, yf);
  const float dm2 = dmask * 0.5f, ehw = 6.0f;
  o.mask = r > dm2 + 2 * ehw ? 0.0f
                             : (r >= dm2 ? 0.5f * cosf(kRefPi * (r - dm2) / (2 * ehw)) + 0.5f : 1.0f);
  o.tb = o.dtx = o.dty = 0.0f;
  o.in = false;
  if (o.mask == 0.0f) return o;
  const double x2 = cs.x * xf + cs.y * yf + (double)(L / 2);
  const double y2 = -cs.y * xf + cs.x * yf + (double)(L / 2);
  if (x2 < 0 || x2 > L - 1.0 || y2 < 0 || y2 > L - 1.0) return o;
  o.in = true;
  const int ii = (int)floor(x2), jj = (int)floor(y2);
  long k0 = ii + (long)jj * L, k1 = k0 + 1, k2 = k0 + L + 1, k3 = k0 + L;
  if (ii == L - 1) { k1--; k2--; }
  if (jj == L - 1) { k2 -= L; k3 -= L; }
  const float t = (float)(x2 - ii), u = (float)(y2 - jj);
  const float tt = 1.0f - t, uu = 1.0f - u;
  const float a0 = src[k0], a1 = src[k1], a2 = src[k2], a3 = src[k3];
  o.tb = a0 * tt * uu + a1 * t * uu + a2 * t * u + a3 * tt * u;
  o.dtx = (a1 - a0) * uu + (a2 - a3) * u;
  o.dty = (a3 - a0) * tt + (a2 - a1) * t;
  return o;
}
// Bilinear adjoint as a gather: source pixel q = (qx, qy) collects g of every
// output pixel whose interpolation cell (rot_sample_d's arithmetic, boundary
// rules included) has q as a corner, times that corner's weight.  The
// outputs whose cell can touch q lie within 2 pixels of the inverse rotation
// of q: x - c = R^T (q - c).  An output pixel samples q's cells only if its
// rotated position lies within 1 of q on both axes, so within sqrt(2) of
// R^T (q - c) on each: the 4 x 4 window floor(c) - 1 .. floor(c) + 2 holds
// all of them (0.41 px to spare for the float rounding of c); the others
// carry weight 0.  Pixels beyond the mask's reach receive nothing.
template <int L>
__device__ __forceinline__ float rot_adj_gather(const float* g, int qx, int qy, double2 cs,
                                                float dmask) {
  const long i = (long)qy * L + qx;
  const float reach = dmask * 0.5f + 12.0f + 3.0f;
  const float qX = (float)(qx - L / 2), qY = (float)(qy - L / 2);
  float acc = 0.0f;
  if (qX * qX + qY * qY > reach * reach) return acc;
  const float cX = (float)(cs.x * qX - cs.y * qY), cY = (float)(cs.y * qX + cs.x * qY);
  const int x0 = (int)floorf(cX) + L / 2 - 1, y0 = (int)floorf(cY) + L / 2 - 1;
  for (int dy = 0; dy < 4; ++dy) {
    const int y = y0 + dy;
    if (y < 0 || y >= L) continue;
    for (int dx = 0; dx < 4; ++dx) {
      const int x = x0 + dx;
      if (x < 0 || x >= L) continue;
      const float gv = g[(long)y * L + x];
      if (gv == 0.0f) continue;
      const float xf = (float)(x - L / 2), yf = (float)(y - L / 2);
      // cheap FP32 pre-test: output x's cell corner is floor(x2), floor(y2),
      // and the cell holds q only if x2 lies in [qx - 1, qx + 1] (same in y;
      // the edge rules clamp inside that range); the FP32 coordinates differ
      // from the FP64 ones by at most about 1e-4 pixel (|x|, |y| <= L/2,
      // relative rounding 2^-24 per operation), far inside the 2e-3 margin
      const float x2f = (float)cs.x * xf + (float)cs.y * yf + (float)(L / 2);
      const float y2f = -(float)cs.y * xf + (float)cs.x * yf + (float)(L / 2);
      if (fabsf(x2f - (float)qx) > 1.002f || fabsf(y2f - (float)qy) > 1.002f) continue;
      const double x2 = cs.x * xf + cs.y * yf + (double)(L / 2);
      const double y2 = -cs.y * xf + cs.x * yf + (double)(L / 2);
      if (x2 < 0 || x2 > L - 1.0 || y2 < 0 || y2 > L - 1.0) continue;
      const int ii = (int)floor(x2), jj = (int)floor(y2);
      long k0 = ii + (long)jj * L, k1 = k0 + 1, k2 = k0 + L + 1, k3 = k0 + L;
      if (ii == L - 1) { k1--; k2--; }
      if (jj == L - 1) { k2 -= L; k3 -= L; }
      const float t = (float)(x2 - ii), u = (float)(y2 - jj);
      const float tt = 1.0f - t, uu = 1.0f - u;
      float wq = 0.0f;
      if (k0 == i) wq += tt * uu;
      if (k1 == i) wq += t * uu;
      if (k2 == i) wq += t * u;
      if (k3 == i) wq += tt * u;
      acc += gv * wq;
    }
  }
  return acc;
}

template <int L>
__device__ __forceinline__ float rot_value_d(const float* src, int x, int y, double2 cs,
                                             float dmask, float pre) {
  const RotSampleD o = rot_sample_d<L>(src, x, y, cs, dmask);
  return o.tb * pre * o.mask;
}

}  // namespace gs
