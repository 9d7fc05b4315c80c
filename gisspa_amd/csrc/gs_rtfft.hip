// gs_rtfft.hip -- runtime-length FFTs and the whole-micrograph phase-flip
// preparation (init_d_image with phase_flip = 1, main.cu:548-586, and the
// per-tile amplitude normalisation of split_normalize_image, main.cu:614-653).
//
// Micrograph sizes are arbitrary (4092 = 2^2 3 11 31, 3838 = 2 19 101, ...),
// so these transforms take the radix list at run time: every Stockham pass is
// one launch in which each thread produces one output element as an R-term
// sum with table twiddles.  They run once per micrograph, next to ~10^7
// correlation maps, so simplicity wins over speed here.  Columns are handled
// by an LDS-tiled transpose and a second row pass.
#include <climits>
#include <cmath>

#include "gs_fft.hpp"
#include "gs_kernels.hpp"

namespace gs {

constexpr float kRefPiF = 3.141592654f;  // GPU_func.cuh:29

// One Stockham pass of radix R over `batch` sequences of length n:
// out[b][(j / Ns) Ns R + k + q Ns] = sum_r in[b][j + r n/R] w^{r (k + q Ns) n / (Ns R)}
__global__ void k_rt_pass(const float2* __restrict__ in, float2* __restrict__ out, int n, int R,
                          int Ns, long total, const float2* __restrict__ tw, int dir) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const long b = g / n;
  const int o = (int)(g - b * n);
  const int k = o % Ns;
  const int q = (o / Ns) % R;
  const int j = (o / (Ns * R)) * Ns + k;
  const int m = n / R;
  const int step = n / (Ns * R);
  const float2* src = in + b * n;
  float2 acc = make_float2(0.f, 0.f);
  const int e = k + q * Ns;
  for (int r = 0; r < R; ++r) {
    const long idx = ((long)r * e % (Ns * R)) * step;
    float2 w = tw[idx];
    if (dir > 0) w.y = -w.y;
    acc = cadd(acc, cmul(src[j + r * m], w));
  }
  out[g] = acc;
}

// out[x][y] = in[y][x] for an ny x nx array, 32 x 32 tiles through LDS
__global__ void k_transpose(const float2* __restrict__ in, float2* __restrict__ out, int nx,
                            int ny) {
  __shared__ float2 tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int i = threadIdx.y; i < 32; i += 8) {
    const int x = bx + threadIdx.x, y = by + i;
    if (x < nx && y < ny) tile[i][threadIdx.x] = in[(long)y * nx + x];
  }
  __syncthreads();
  for (int i = threadIdx.y; i < 32; i += 8) {
    const int y = by + threadIdx.x, x = bx + i;
    if (x < nx && y < ny) out[(long)x * ny + y] = tile[threadIdx.x][i];
  }
}

__global__ void k_real_to_complex(const float* __restrict__ in, float2* __restrict__ out, long n,
                                  float scale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(in[i] * scale, 0.f);
}

__global__ void k_complex_to_real(const float2* __restrict__ in, float* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i].x;
}

// radix list of n (4 first, then 2, 3, 5, 7, ..., the remaining prime)
static int rt_radices(int n, int* r) {
  int k = 0;
  while (n % 4 == 0) { r[k++] = 4; n /= 4; }
  while (n % 2 == 0) { r[k++] = 2; n /= 2; }
  for (int f = 3; n > 1 && k < 62; f += 2) {
    while (n % f == 0) { r[k++] = f; n /= f; }
    if ((long)f * f > n && n > 1) { r[k++] = n; n = 1; }
  }
  return k;
}

// `batch` length-n transforms in place in `data` (tmp: same size)
void rt_fft_rows(float2* data, float2* tmp, int n, long batch, const float2* tw, int dir,
                 hipStream_t s) {
  int rad[64];
  const int np = rt_radices(n, rad);
  const long total = batch * n;
  float2* src = data;
  float2* dst = tmp;
  int Ns = 1;
  for (int i = 0; i < np; ++i) {
    hipLaunchKernelGGL(k_rt_pass, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst,
                       n, rad[i], Ns, total, tw, dir);
    Ns *= rad[i];
    float2* t = src;
    src = dst;
    dst = t;
  }
  if (src != data)
    (void)hipMemcpyAsync(data, src, total * sizeof(float2), hipMemcpyDeviceToDevice, s);
}

void rt_transpose(const float2* in, float2* out, int nx, int ny, hipStream_t s) {
  dim3 grid((nx + 31) / 32, (ny + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, s, in, out, nx, ny);
}

// unnormalised 2-D transform of an ny x nx complex image, in place
void rt_fft_2d(float2* data, float2* tmp, int nx, int ny, const float2* twx, const float2* twy,
               int dir, hipStream_t s) {
  rt_fft_rows(data, tmp, nx, ny, twx, dir, s);
  rt_transpose(data, tmp, nx, ny, s);            // tmp: nx x ny (rows of length ny)
  rt_fft_rows(tmp, data, ny, nx, twy, dir, s);
  rt_transpose(tmp, data, ny, nx, s);
}

void launch_real_to_complex(const float* in, float2* out, long n, float scale, hipStream_t s) {
  hipLaunchKernelGGL(k_real_to_complex, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out,
                     n, scale);
}
void launch_complex_to_real(const float2* in, float* out, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_complex_to_real, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out,
                     n);
}

// ---------------------------------------------------------------------------
// phase-flip preparation of the whole micrograph
// ---------------------------------------------------------------------------
__device__ float ctf_ast_img(int x1, int y1, int nx, int ny, const WeightArgs& w, int mode) {
  // CTF_AST, GPU_func.cu:292-315
  y1 = x1 > nx / 2 ? ny - y1 : y1;
  x1 = min(x1, nx - x1);
  const float dx = (float)min(x1, nx - x1);
  const float dy = (float)(y1 - ny / 2);
  const float ss = (dx * dx / ((float)nx * nx) + dy * dy / ((float)ny * ny)) / (w.apix * w.apix);
  const float ag = atan2f((float)(y1 - ny / 2), (float)x1);
  const float df_ast = 0.5f * (w.dfu + w.dfv + 2 * w.dfdiff * cosf(2 * (w.dfang * kRefPiF / 180 - ag)));
  const float gamma = -2 * kRefPiF * (w.cs * 2.5e6f * w.lambda * w.lambda * w.lambda * ss * ss +
                                      df_ast * 5000.0f * w.lambda * ss);
  const float v = sqrtf(1.0f - w.ampconst * w.ampconst) * sinf(gamma) + w.ampconst * cosf(gamma);
  if (mode == 0) return v > 0 ? 1.0f : -1.0f;
  if (mode == 2) return fabsf(v);
  return v;
}

// do_phase_flip (GPU_func.cu:785-796)
__global__ void k_phase_flip(float2* __restrict__ F, int nx, int ny, WeightArgs w) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nx * ny) return;
  const int x = (int)(i % nx), y = (int)(i / nx);
  const float v = ctf_ast_img(x, (y + ny / 2) % ny, nx, ny, w, 0);
  F[i] = cscale(F[i], v);
}

// radial power of the half plane x <= nx/2 (SQRSum_by_circle mode 1,
// GPU_func.cu:98-128): one wave per bin over a host-built pixel list
__global__ __launch_bounds__(256) void k_img_radial(const float2* __restrict__ F,
                                                    const int* __restrict__ start,
                                                    const long* __restrict__ pix, int nbin,
                                                    double* __restrict__ ra) {
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (w >= nbin) return;
  double acc = 0;
  for (int i = start[w] + lane; i < start[w + 1]; i += 64) {
    const float2 v = F[pix[i]];
    acc += (double)v.x * v.x + (double)v.y * v.y;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) ra[w] = acc;
}

// whiten_filetr_weight_Img (GPU_func.cu:155-205) as a real factor on F, and
// set_0Hz_to_0_at_RI (GPU_func.cu:207-227)
__global__ void k_img_weight(float2* __restrict__ F, int nx, int ny, const double* __restrict__ ra,
                             const float* __restrict__ rb, WeightArgs w) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nx * ny) return;
  const int x = (int)(i % nx), y = (int)(i / nx);
  const float dx = (float)min(x, nx - x), dy = (float)min(y, ny - y);
  const int r = (int)floorf(hypotf(dx, dy) + 0.5f) - 1;
  const float ss = sqrtf((dx * dx / ((float)nx * nx) + dy * dy / ((float)ny * ny)) / (w.apix * w.apix));
  const int l = max(nx, ny);
  float v = 0.f, signal = 0.f, ncurve = 0.f, f = 1.0f;
  if (r < l / 2 && r >= 0) {
    v = ctf_ast_img(x, (y + ny / 2) % ny, nx, ny, w, 2);
    signal = expf(w.bf * ss * ss + w.bf2 * ss + w.bf3);
    ncurve = expf(w.a * ss * ss + w.b * ss + w.b2);
    f = f * sqrtf((signal * v * v + ncurve) / signal) / (float)sqrt(ra[r] / (double)rb[r]);
    if (r > (l * w.apix / 6)) f = f * expf(-100 * ss * ss);
  }
  const float hi = l * w.apix / w.highres, lo = l * w.apix / w.lowres;
  if (r < hi && r >= lo) {
  } else if (r >= hi && r < hi + 8) {
    f = f * (0.5f * cosf(kRefPiF * (r - hi) / (2 * 8)) + 0.5f);
  } else if (r >= (lo - 8) && r < lo && r >= 0) {
    f = f * (0.5f * cosf(kRefPiF * (lo - r) / (2 * 8)) + 0.5f);
  } else {
    f = 0.0f;
  }
  if (r < l / 2 && r >= 0) {
    signal /= (w.kk + 1);
    ncurve /= signal;
    f = f * sqrtf(1 / (ncurve + w.kk * v * v));
  }
  F[i] = (i == 0) ? make_float2(0.f, 0.f) : cscale(F[i], f);
}

void launch_phase_flip(float2* F, int nx, int ny, const WeightArgs& w, hipStream_t s) {
  const long n = (long)nx * ny;
  hipLaunchKernelGGL(k_phase_flip, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F, nx, ny, w);
}
void launch_img_radial(const float2* F, const int* start, const long* pix, int nbin, double* ra,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_img_radial, dim3((unsigned)((nbin * 64 + 255) / 256)), dim3(256), 0, s, F,
                     start, pix, nbin, ra);
}
void launch_img_weight(float2* F, int nx, int ny, const double* ra, const float* rb,
                       const WeightArgs& w, hipStream_t s) {
  const long n = (long)nx * ny;
  hipLaunchKernelGGL(k_img_weight, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, F, nx, ny,
                     ra, rb, w);
}

// ---------------------------------------------------------------------------
// per-tile amplitude normalisation on the [t][kx][ky] half spectra
// (compute_area_sum_ofSQR + normalize, main.cu:627-648): sum |F|^2 over the
// x <= l/2 half plane inside radius l/2, m = sqrt(sum) / count, F /= m
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tile_ringsum(const float2* __restrict__ spec, int L,
                                                      double* __restrict__ part) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  const int H = L / 2 + 1;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  double acc = 0;
  if (i < (long)H * L) {
    const int kx = (int)(i / L), ky = (int)(i % L);
    const float rr = hypotf((float)min(ky, L - ky), (float)min(kx, L - kx));
    const int r = (int)floorf(rr + 0.5f) - 1;
    if (r >= 0 && r < L / 2) {
      const float2 v = spec[(long)b * H * L + i];
      acc = (double)v.x * v.x + (double)v.y * v.y;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(long)b * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_tile_ampscale(const double* __restrict__ part, int nparts, int nb, double count,
                                float* __restrict__ scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double S = 0;
  for (int i = 0; i < nparts; ++i) S += part[(long)b * nparts + i];
  const double m = sqrt(S / (count * count));
  scale[b] = (m != 0) ? (float)(1.0 / m) : 1.0f;
}

__global__ void k_tile_apply(float2* __restrict__ spec, int L, const float* __restrict__ scale) {
  const int b = blockIdx.y;
  const int H = L / 2 + 1;
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < (long)H * L) spec[(long)b * H * L + i] = cscale(spec[(long)b * H * L + i], scale[b]);
}

void launch_tile_ampnorm(float2* spec, int L, int ntiles, double count, double* part,
                         float* scale, hipStream_t s) {
  const long tot = (long)(L / 2 + 1) * L;
  const int nparts = (int)((tot + 255) / 256);
  hipLaunchKernelGGL(k_tile_ringsum, dim3(nparts, ntiles), dim3(256), 0, s, spec, L, part);
  hipLaunchKernelGGL(k_tile_ampscale, dim3((ntiles + 63) / 64), dim3(64), 0, s, part, nparts,
                     ntiles, count, scale);
  hipLaunchKernelGGL(k_tile_apply, dim3(nparts, ntiles), dim3(256), 0, s, spec, L, scale);
}

int tile_ampnorm_parts(int L) { return (int)(((long)(L / 2 + 1) * L + 255) / 256); }

}  // namespace gs
