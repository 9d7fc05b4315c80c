// gs_plain.hip -- the stage-by-stage ("plain") GPU picker.
//
// Same arithmetic as the fused path, organised like the reference's
// pickPartcles_IMG_NORM loop (main.cu:669-716, 824-894): full complex L x L
// spectra in row-major order, one kernel per stage with every intermediate in
// HBM (rotated templates, the product CCG, the inverse transform in two
// passes), per-1024-element partial reductions copied to the host after every
// tile and finished there.  It exists as the baseline the batched/fused path
// is measured against; both must give identical pick lists.
#include <climits>
#include <cmath>

#include "gs_fft.hpp"
#include "gs_kernels.hpp"
#include "gs_sizes.hpp"

namespace gs {

constexpr int PWG = 256;
constexpr float kPiRef = 3.141592654f;  // GPU_func.cuh:29

// rotate_subIMG + apply_mask on real templates into complex images, and the
// per-1024-element sums of compute_sum_sqr (GPU_func.cu:363-407)
__global__ __launch_bounds__(PWG) void k_plain_rotate(const float* __restrict__ treal, int L,
                                                      float cose, float sine, float dmask,
                                                      float2* __restrict__ out,
                                                      double* __restrict__ part, int nblk) {
  __shared__ double red[2 * (PWG / 64)];
  const int b = blockIdx.y;
  const long L2 = (long)L * L;
  const long i0 = (long)blockIdx.x * 1024;
  const float* src = treal + b * L2;
  double a1 = 0, a2 = 0;
  for (int t = threadIdx.x; t < 1024; t += PWG) {
    const long i = i0 + t;
    if (i >= L2) break;
    const int x = (int)(i % L), y = (int)(i / L);
    const float xf = (float)(x - L / 2), yf = (float)(y - L / 2);
    const float r = hypotf(xf, yf);
    const float dm2 = dmask * 0.5f;
    float v = 0.0f;
    if (!(r > dm2 + 12.0f)) {
      const float x2 = __fadd_rn(__fadd_rn(__fmul_rn(cose, xf), __fmul_rn(sine, yf)), (float)(L / 2));
      const float y2 = __fadd_rn(__fadd_rn(__fmul_rn(-sine, xf), __fmul_rn(cose, yf)), (float)(L / 2));
      if (!(x2 < 0 || x2 > L - 1.0f || y2 < 0 || y2 > L - 1.0f)) {
        const int ii = (int)floorf(x2), jj = (int)floorf(y2);
        long k0 = ii + (long)jj * L, k1 = k0 + 1, k2 = k0 + L + 1, k3 = k0 + L;
        if (ii == L - 1) { k1--; k2--; }
        if (jj == L - 1) { k2 -= L; k3 -= L; }
        const float tx = x2 - (float)ii, u = y2 - (float)jj;
        const float tt = 1.0f - tx, uu = 1.0f - u;
        v = src[k0] * tt * uu + src[k1] * tx * uu + src[k2] * tx * u + src[k3] * tt * u;
      }
      if (r >= dm2) v *= 0.5f * cosf(kPiRef * (r - dm2) / 12.0f) + 0.5f;
    }
    out[b * L2 + i] = make_float2(v, 0.0f);
    a1 += v;
    a2 += (double)v * v;
  }
  for (int o = 32; o > 0; o >>= 1) { a1 += __shfl_xor(a1, o, 64); a2 += __shfl_xor(a2, o, 64); }
  if ((threadIdx.x & 63) == 0) { red[2 * (threadIdx.x >> 6)] = a1; red[2 * (threadIdx.x >> 6) + 1] = a2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s1 = 0, s2 = 0;
    for (int w = 0; w < PWG / 64; ++w) { s1 += red[2 * w]; s2 += red[2 * w + 1]; }
    part[2 * ((long)b * nblk + blockIdx.x)] = s1;
    part[2 * ((long)b * nblk + blockIdx.x) + 1] = s2;
  }
}

// substract_by_mean + divided_by_var (GPU_func.cu:424-443)
__global__ void k_plain_normalize(float2* __restrict__ data, long L2, const float* __restrict__ mean,
                                  const float* __restrict__ var) {
  const int b = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L2) return;
  float v = data[b * L2 + i].x;
  if (mean[b] != 0) v -= mean[b];
  if (var[b] != 0) v /= sqrtf(var[b]);
  data[b * L2 + i].x = v;
}

// batched complex transforms of rows / columns of L x L images, in place
template <int L, int DIR>
__global__ __launch_bounds__(PWG) void k_c2c_rows(float2* __restrict__ data, long nrows,
                                                  float scale, const float2* __restrict__ tw) {
  constexpr int NB = 8;
  __shared__ float2 lds[NB * exec_stride<L, NB, PWG, MAP_J>()];
  const long row0 = (long)blockIdx.x * NB;
  auto load = [&](int b, int i) -> float2 {
    const long r = row0 + b;
    return data[(r < nrows ? r : row0) * L + i];
  };
  auto store = [&](int b, int i, float2 v) {
    const long r = row0 + b;
    if (r < nrows) data[r * L + i] = cscale(v, scale);
  };
  fft_exec<L, NB, PWG, DIR, MAP_J, MAP_J>(lds, tw, load, store);
}

template <int L, int DIR>
__global__ __launch_bounds__(PWG) void k_c2c_cols(float2* __restrict__ data, float scale,
                                                  const float2* __restrict__ tw) {
  constexpr int NB = 8;
  __shared__ float2 lds[NB * exec_stride<L, NB, PWG, MAP_B>()];
  const int x0 = blockIdx.x * NB;
  float2* img = data + (long)blockIdx.y * L * L;
  auto load = [&](int c, int y) -> float2 {
    const int x = x0 + c;
    return img[(long)y * L + (x < L ? x : x0)];
  };
  auto store = [&](int c, int y, float2 v) {
    const int x = x0 + c;
    if (x < L) img[(long)y * L + x] = cscale(v, scale);
  };
  fft_exec<L, NB, PWG, DIR, MAP_B, MAP_B>(lds, tw, load, store);
}

// split_IMG (GPU_func.cu:564-581) into complex tiles, after mean removal
__global__ void k_plain_tiles(const float* __restrict__ img, int nx, int ny, int L,
                              const TileDesc* __restrict__ td, int tile0, const double* mean,
                              float2* __restrict__ out) {
  const int t = blockIdx.y;
  const long L2 = (long)L * L;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L2) return;
  const TileDesc d = td[tile0 + t];
  const int x = (int)(i % L) + d.ox, y = (int)(i / L) + d.oy;
  float v = 0.0f;
  if (x < nx && y < ny) v = (float)((double)img[(long)y * nx + x] - *mean);
  out[(long)t * L2 + i] = make_float2(v, 0.0f);
}

// compute_corner_CCG (GPU_func.cu:584-615)
__global__ void k_plain_ccg(const float2* __restrict__ tile, const float2* __restrict__ tspec,
                            int L, float sgn, float2* __restrict__ ccg) {
  const int b = blockIdx.y;
  const long L2 = (long)L * L;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L2) return;
  const int x = (int)(i % L), y = (int)(i / L);
  const float2 v = cmulc(tile[i], tspec[b * L2 + i]);
  ccg[b * L2 + i] = cscale(v, ((x + y) & 1) ? -sgn : sgn);
}

// get_peak_and_SUM (GPU_func.cu:667-708) over 1024-element runs of each map:
// (border-excluded max, its position, sum, sum of squares)
struct PeakPart {
  float peak;
  int pos;
  float s1, s2;
};

__global__ __launch_bounds__(256) void k_plain_peak(const float2* __restrict__ ccg, int L,
                                                    float lo, float hi, int nblk,
                                                    PeakPart* __restrict__ part) {
  __shared__ float rv[4], rs1[4], rs2[4];
  __shared__ int rp[4];
  const int b = blockIdx.y;
  const long L2 = (long)L * L;
  const long i0 = (long)blockIdx.x * 1024;
  float v = 0.0f, s1 = 0.0f, s2 = 0.0f;
  int p = INT_MAX;
  for (int t = threadIdx.x; t < 1024; t += 256) {
    const long i = i0 + t;
    if (i >= L2) break;
    const float c = ccg[b * L2 + i].x;
    s1 += c;
    s2 += c * c;
    const int x = (int)(i % L), y = (int)(i / L);
    const float m = ((float)x < lo || (float)x > hi || (float)y < lo || (float)y > hi) ? 0.0f : c;
    if (m > v || (m == v && (int)i < p)) { v = m; p = (int)i; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float v2 = __shfl_xor(v, o, 64);
    const int p2 = __shfl_xor(p, o, 64);
    if (v2 > v || (v2 == v && p2 < p)) { v = v2; p = p2; }
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { rv[w] = v; rp[w] = p; rs1[w] = s1; rs2[w] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      if (rv[k] > v || (rv[k] == v && rp[k] < p)) { v = rv[k]; p = rp[k]; }
    }
    PeakPart pp;
    pp.peak = v;
    pp.pos = p;
    pp.s1 = rs1[0] + rs1[1] + rs1[2] + rs1[3];
    pp.s2 = rs2[0] + rs2[1] + rs2[2] + rs2[3];
    part[(long)b * nblk + blockIdx.x] = pp;
  }
}

// ---------------------------------------------------------------------------
#define PDISPATCH(L, FN, ...) GS_DISPATCH(L, FN, __VA_ARGS__)

int plain_blocks(int L) { return (int)(((long)L * L + 1023) / 1024); }
int plain_part_bytes() { return (int)sizeof(PeakPart); }

void launch_plain_rotate(int L, const float* treal, int n, float cose, float sine, float dmask,
                         float2* out, double* part, hipStream_t s) {
  const int nblk = plain_blocks(L);
  hipLaunchKernelGGL(k_plain_rotate, dim3(nblk, n), dim3(PWG), 0, s, treal, L, cose, sine, dmask,
                     out, part, nblk);
}
void launch_plain_normalize(int L, float2* data, int n, const float* mean, const float* var,
                            hipStream_t s) {
  const long L2 = (long)L * L;
  hipLaunchKernelGGL(k_plain_normalize, dim3((unsigned)((L2 + 255) / 256), n), dim3(256), 0, s,
                     data, L2, mean, var);
}
template <int L>
static void c2c_2d_t(float2* data, int nimg, int dir, float scale, const float2* tw,
                     hipStream_t s) {
  const long nrows = (long)nimg * L;
  dim3 gr((unsigned)((nrows + 7) / 8));
  dim3 gc((L + 7) / 8, nimg);
  if (dir < 0) {
    hipLaunchKernelGGL((k_c2c_rows<L, -1>), gr, dim3(PWG), 0, s, data, nrows, 1.0f, tw);
    hipLaunchKernelGGL((k_c2c_cols<L, -1>), gc, dim3(PWG), 0, s, data, scale, tw);
  } else {
    hipLaunchKernelGGL((k_c2c_rows<L, +1>), gr, dim3(PWG), 0, s, data, nrows, 1.0f, tw);
    hipLaunchKernelGGL((k_c2c_cols<L, +1>), gc, dim3(PWG), 0, s, data, scale, tw);
  }
}
void launch_c2c_2d(int L, float2* data, int nimg, int dir, float scale, const float2* twL,
                   hipStream_t s) {
  PDISPATCH(L, c2c_2d_t, data, nimg, dir, scale, twL, s);
}
void launch_plain_tiles(int L, const float* img, int nx, int ny, const TileDesc* td, int tile0,
                        int ntiles, const double* mean, float2* out, hipStream_t s) {
  const long L2 = (long)L * L;
  hipLaunchKernelGGL(k_plain_tiles, dim3((unsigned)((L2 + 255) / 256), ntiles), dim3(256), 0, s,
                     img, nx, ny, L, td, tile0, mean, out);
}
void launch_plain_ccg(int L, const float2* tile, const float2* tspec, int n, float sgn,
                      float2* ccg, hipStream_t s) {
  const long L2 = (long)L * L;
  hipLaunchKernelGGL(k_plain_ccg, dim3((unsigned)((L2 + 255) / 256), n), dim3(256), 0, s, tile,
                     tspec, L, sgn, ccg);
}
void launch_plain_peak(int L, const float2* ccg, int n, float lo, float hi, void* part,
                       hipStream_t s) {
  const int nblk = plain_blocks(L);
  hipLaunchKernelGGL(k_plain_peak, dim3(nblk, n), dim3(256), 0, s, ccg, L, lo, hi, nblk,
                     static_cast<PeakPart*>(part));
}

}  // namespace gs
