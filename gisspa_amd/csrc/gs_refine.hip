// gs_refine.hip -- the refinement objective of a picked candidate and its
// hand-written adjoint (no counterpart in the reference, which stops at the
// grid search of main.cu:1230-1257).
//
// For candidate b (tile t, view j, peak (x0, y0) in the tile) and
// theta = (psi degrees, dx, dy pixels, defocus um):
//   A      = W(k; defocus) B_j(k)        B_j: whitened template half spectrum
//   m      = sqrt(sum_half |A|^2) / count (amplitude normaliser, main.cu:444-452)
//   T      = C2R(A)                       real template before 1/m
//   Tb(x)  = bilinear(T) at the rotation of x by psi; R = Tb / m
//   Rm     = mask R, Rn = (Rm - mean) / sd over L^2
//   That   = FFT(Rn) / L^2 with the DC zeroed (as the picker's rotation path)
//   P      = I_t conj(That)               (the tile spectra carry (-1)^(kx+ky) sgn)
//   c      = Re sum_k P(k) e^{2 pi i k'.u / L},  u = (x0 + dx, y0 + dy)
//   sigma  = sqrt(sum_k |P(k)|^2),  score = c / sigma.
// The backward pass runs the chain in reverse: d score / d That (per
// frequency), C2R to d/dRn, the normalisation adjoint, the mask, the
// bilinear adjoint (a gather over the output pixels whose interpolation cell
// touches each source pixel, so no atomics) and its psi derivative, R2C to
// d/dA, and the derivative of the CTF weighting.  All sums are FP64.
#include "gs_device.hpp"
#include "gs_sizes.hpp"
#include "gs_fft.hpp"

namespace gs {

constexpr int RPARTS = 64;  // partial blocks per candidate in each reduction

// A cell's template weight f and its derivative fd by the defocus.
struct WeightD {
  float f, fd;
  int rbin;
};

// The template weighting of apply_weighting_function (GPU_func.cu:249-290,
// the arithmetic of k_tmpl_weight: low-pass factor, |CTF_AST| of
// GPU_func.cu:292-315, noise / signal curves) and its derivative by the
// defocus (um, positive underfocus: df_ast = -defocus + dfdiff cos(...)),
// split into a defocus-independent table per half-spectrum cell, so the
// refinement kernels evaluate the weight of a candidate's defocus with one
// sincos: with the template CTF's dfdiff / dfang fixed, gamma = 2 pi (g0 + g1
// defocus (um)), and 2 pi g1 is d gamma / d defocus.  The table holds the
// phase in revolutions, so the kernels reduce it by its nearest integer
// (exact) and take sin / cos of pi times twice the remainder, with no
// large-argument reduction.  Entry = (flp, ncurve, g0, g1); ncurve < 0 marks
// cells outside the |CTF| rings (weight = flp, slope 0, not in the
// normaliser's sums).
template <int L>
__global__ __launch_bounds__(WG) void k_ref_wtable(WeightArgs w, float4* __restrict__ tab) {
  constexpr int H = L / 2 + 1;
  const long i = (long)blockIdx.x * WG + threadIdx.x;
  if (i >= (long)H * L) return;
  const int kx = (int)(i / L), ky = (int)(i - (long)kx * L);
  const float hy = (float)min(ky, L - ky), hx = (float)min(kx, L - kx);
  const float rr = hypotf(hy, hx);
  const int rbin = (int)floorf(rr + 0.5f) - 1;
  const float hi = L * w.apix / w.highres, lo = L * w.apix / w.lowres;
  float flp = 1.0f;
  if (rbin < hi && rbin >= 0) {
  } else if (rbin >= hi && rbin < hi + 8) {
    flp = 0.5f * cosf(kRefPi * (rbin - hi) / (2 * 8)) + 0.5f;
  } else if (rbin >= (lo - 8) && rbin < lo && rbin >= 0) {
    flp = 0.5f * cosf(kRefPi * (lo - rbin) / (2 * 8)) + 0.5f;
  } else {
    flp = 0.0f;
  }
  float4 t = make_float4(flp, -1.0f, 0.0f, 0.0f);
  if (rbin < L / 2 && rbin >= 0) {
    int x1 = kx, y1 = (ky + L / 2) % L;
    y1 = x1 > L / 2 ? L - y1 : y1;
    x1 = min(x1, L - x1);
    const float dx = (float)min(x1, L - x1);
    const float dy = (float)(y1 - L / 2);
    const float sc = (dx * dx / (float)(L * L) + dy * dy / (float)(L * L)) / (w.apix * w.apix);
    const float ag = atan2f((float)(y1 - L / 2), (float)x1);
    // df_ast = -defocus + dfdiff cos(2 (dfang - ag)) (dfu + dfv = -2 defocus)
    const float ast = w.dfdiff * cosf(2 * (w.dfang * kRefPi / 180 - ag));
    // gamma / (2 pi), revolutions
    t.z = -(w.cs * 2.5e6f * w.lambda * w.lambda * w.lambda * sc * sc + ast * 5000.0f * w.lambda * sc);
    t.w = 5000.0f * w.lambda * sc;
    const float ss = rr * w.ds;
    const float signal = expf(w.bf * ss * ss + w.bf2 * ss + w.bf3) / (w.kk + 1);
    t.y = expf(w.a * ss * ss + w.b * ss + w.b2) / signal;
  }
  tab[i] = t;
}

// weight_deriv of one cell at a candidate's defocus from its table entry;
// sa = sqrt(1 - ampconst^2)
__device__ __forceinline__ WeightD weight_tab(float4 t, float defocus, const WeightArgs& w,
                                              float sa) {
  WeightD o;
  o.rbin = t.y < 0.0f ? -1 : 0;  // only the in-range test is used
  if (t.y < 0.0f) {
    o.f = t.x;
    o.fd = 0.0f;
    return o;
  }
  const float rev = t.z + t.w * defocus;  // gamma / (2 pi)
  const float fr = rev - rintf(rev);       // exact, |fr| <= 1/2
  float sg, cg;
  sincospif(2.0f * fr, &sg, &cg);
  const float v = sa * sg + w.ampconst * cg;
  const float dv = (sa * cg - w.ampconst * sg) * (2.0f * kRefPi * t.w);
  const float vv = fabsf(v), dvv = v >= 0 ? dv : -dv;
  const float q = t.y + w.kk * vv * vv;
  const float rq = rsqrtf(q);
  o.f = t.x * vv * rq;
  o.fd = t.x * t.y * rq * rq * rq * dvv;
  return o;
}

// block reduction of N doubles; valid in thread 0
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* scratch) {
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum_d(v[i]);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) scratch[w * N + i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double s = 0;
      for (int k = 0; k < WG / 64; ++k) s += scratch[k * N + i];
      v[i] = s;
    }
}

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------
// m = sqrt(S) / count and dm / d defocus = (sum |B|^2 f f') / (count sqrt(S))
__global__ void k_ref_norm_m(const double* __restrict__ part, int nb, double count,
                             RefScal* __restrict__ sc, float* __restrict__ pre, int np,
                             int stride) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double S = 0, SW = 0;
  for (int i = 0; i < np; ++i) {
    S += part[((long)b * stride + i) * 2];
    SW += part[((long)b * stride + i) * 2 + 1];
  }
  const double rs = sqrt(S);
  const double m = rs / count;
  sc[b].m = m != 0 ? m : 1.0;
  sc[b].dm = rs != 0 ? SW / (count * rs) : 0.0;
  pre[b] = (float)(1.0 / sc[b].m);
}

// mean and sd of the rotated masked template (the row partials of
// k_rows_r2c), and the spectrum scale 1/(L^2 sd) of the picker's k_rot_scale
__global__ void k_ref_rot_stats(const double* __restrict__ partials, int nparts, int nb,
                                double L2, RefScal* __restrict__ sc, float* __restrict__ bscale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s1 = 0, s2 = 0;
  for (int i = 0; i < nparts; ++i) {
    s1 += partials[2 * ((long)b * nparts + i)];
    s2 += partials[2 * ((long)b * nparts + i) + 1];
  }
  const double mean = s1 / L2, var = s2 / L2 - mean * mean;
  sc[b].mu = mean;
  sc[b].sd = var != 0 ? sqrt(var) : 1.0;
  bscale[b] = (float)(1.0 / (L2 * sc[b].sd));
}

// e^{2 pi i k' u / L} for the columns (k' = kx) and rows (signed ky)
template <int L>
__device__ __forceinline__ void phase_tables(float2* ex, float2* ey, double ux, double uy) {
  constexpr int H = L / 2 + 1;
  for (int k = threadIdx.x; k < H + L; k += WG) {
    const bool col = k < H;
    const int kk = col ? k : k - H;
    const double ks = col ? kk : (kk <= L / 2 ? kk : kk - L);
    const double a = fmod(2.0 * ks * (col ? ux : uy) / L, 2.0);
    double sn, cs;
    sincospi(a, &sn, &cs);
    (col ? ex[kk] : ey[kk]) = make_float2((float)cs, (float)sn);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// column kernels of both passes
// ---------------------------------------------------------------------------
// Column kernels with a computed load or a reducing store.  One workgroup
// per (column block of CB half-spectrum columns, candidate), the layout of
// k_cols_inv_T / k_cols_fwd; their partials use FSLOTS slots per candidate
// (one per column block).
constexpr int FSLOTS = 2 * RPARTS;
template <int L>
constexpr int ncol_blocks() { return (Geo<L>::H + Geo<L>::CB - 1) / Geo<L>::CB; }
// FFT schedule of the inverse column kernels that compute per-cell work in
// their loads (weighting, gradient): the first radix pass does that work,
// so at window 720 the 24 x 30 schedule (240 of 256 threads in the first
// pass) instead of the tuned 30 x 24 (192): 7,616 refinements 0.796 instead
// of 0.832 s
template <int L>
constexpr int inv_cols_variant() { return L == 720 ? 1 : 0; }

// Forward weighting fused into the inverse column pass: A = f(k; defocus) B_j
// is formed while loading (never stored), with the partials of sum |A|^2 and
// sum |B|^2 f f' (over the radial bins the normaliser counts), then the
// inverse column transform writes the transposed intermediate of the C2R,
// rows ylo .. yhi only (the rows the rotation reads).
template <int L>
__global__ __launch_bounds__(WG) void k_ref_weight_cols(const float2* __restrict__ white,
                                                        const RefCand* __restrict__ cand,
                                                        WeightArgs w0,
                                                        const float4* __restrict__ wtab,
                                                        int ylo, int yhi,
                                                        float2* __restrict__ Gout,
                                                        double* __restrict__ part,
                                                        const float2* __restrict__ twL) {
  using Gm = Geo<L>;
  constexpr int H = Gm::H, HP = Gm::HP, CB = Gm::CB;
  static_assert(ncol_blocks<L>() <= FSLOTS, "partial slots");
  __shared__ float lds[CB * exec_stride_split<L, CB, WG, MAP_B, inv_cols_variant<L>()>()];
  __shared__ double red[2 * (WG / 64)];
  const int b = blockIdx.y, kx0 = blockIdx.x * CB;
  const RefCand cd = cand[b];
  const float2* Bj = white + (long)cd.view * H * L;
  const int ncol = min(CB, H - kx0);
  const float sa = sqrtf(1.0f - w0.ampconst * w0.ampconst);
  double acc[2] = {0, 0};
  // A = f B formed in the first radix pass's loads, with the partial sums
  auto load = [&](int c, int ky) -> float2 {
    if (c >= ncol) return make_float2(0.f, 0.f);
    const int kx = kx0 + c;
    const WeightD wd = weight_tab(wtab[(long)kx * L + ky], cd.df, w0, sa);
    const float2 bv = Bj[(long)kx * L + ky];
    if (wd.rbin >= 0) {  // inside the |CTF| rings
      const double b2 = (double)bv.x * bv.x + (double)bv.y * bv.y;
      acc[0] += b2 * wd.f * wd.f;
      acc[1] += b2 * wd.f * wd.fd;
    }
    return cscale(bv, wd.f);
  };
  float2* gb = Gout + (long)b * L * HP;
  auto store = [&](int c, int y, float2 v) {
    if (c < ncol && y >= ylo && y <= yhi) gb[(long)y * HP + kx0 + c] = v;
  };
  fft_exec_split<L, CB, WG, +1, MAP_J, MAP_B, inv_cols_variant<L>()>(lds, twL, load, store);
  block_sum_n<2>(acc, red);
  if (threadIdx.x == 0) {
    part[((long)b * FSLOTS + blockIdx.x) * 2] = acc[0];
    part[((long)b * FSLOTS + blockIdx.x) * 2 + 1] = acc[1];
  }
}

// Backward: d score / d That(k) = g I with g = E / sigma - (c / sigma^3)
// conj(P), formed while loading; the DC is zeroed in the forward pass, so it
// carries no gradient.  The real gradient image is (1 / L^2) sum_half w
// Re(G e^{+2 pi i k.x / L}); the C2R computes exactly that once the kx = 0
// and kx = L/2 columns are Hermitian in ky, so those two columns are loaded
// as (G(ky) + conj G(-ky)) / 2.  The inverse column transform then writes the
// intermediate of the C2R that yields d/dRn.
// The normalisation adjoint needs sum_x g Rn with g = d/dRn; by Parseval it
// is sum_half w Re(G conj(That)) (w = 2 between the columns 0 and L/2, 1 on
// them), accumulated here from the values the C2R receives; sum_x g is
// Re G(0, 0) = 0, as the DC carries no gradient.  Rows ylo .. yhi of the
// intermediate are written (d/dRn is used inside the mask only).
template <int L>
__global__ __launch_bounds__(WG) void k_ref_grad_cols(const float2* __restrict__ tiles,
                                                      const float2* __restrict__ spec,
                                                      const RefCand* __restrict__ cand,
                                                      const RefScal* __restrict__ sc,
                                                      const int* __restrict__ skip,
                                                      int ylo, int yhi,
                                                      float2* __restrict__ Gout,
                                                      double* __restrict__ part,
                                                      const float2* __restrict__ twL) {
  using Gm = Geo<L>;
  constexpr int H = Gm::H, HP = Gm::HP, CB = Gm::CB;
  static_assert(ncol_blocks<L>() <= FSLOTS, "partial slots");
  __shared__ float lds[CB * exec_stride_split<L, CB, WG, MAP_B, inv_cols_variant<L>()>()];
  __shared__ float2 ex[H], ey[L];
  __shared__ double red[WG / 64];
  const int b = blockIdx.y, kx0 = blockIdx.x * CB;
  if (skip[b]) return;  // the whole workgroup
  const RefCand cd = cand[b];
  phase_tables<L>(ex, ey, (double)cd.x0 + cd.dx, (double)cd.y0 + cd.dy);  // (ends in a barrier)
  const float2* I = tiles + (long)cd.ptile * H * L;
  const float2* T = spec + (long)b * H * L;
  const double sg = sc[b].sigma;
  const float is = sg > 0 ? (float)(1.0 / sg) : 0.0f;
  const float cs3 = sg > 0 ? (float)(sc[b].c / (sg * sg * sg)) : 0.0f;
  // G = g I; Re(G conj(That)) = Re(g P) comes with it (the Parseval term)
  auto grad = [&](int kx, int ky, float* gp) -> float2 {
    if (kx == 0 && ky == 0) return make_float2(0.f, 0.f);
    const long i = (long)kx * L + ky;
    const float2 iv = I[i];
    const float2 P = cmulc(iv, T[i]);
    const float2 e = cmul(ex[kx], ey[ky]);
    const float2 g = make_float2(e.x * is - cs3 * P.x, e.y * is + cs3 * P.y);
    if (gp) *gp = fmaf(g.x, P.x, -g.y * P.y);
    return cmul(g, iv);
  };
  const int ncol = min(CB, H - kx0);
  // a thread's ~CB L / WG terms in FP32, the block in FP64; on the columns
  // kx = 0 and L/2 the symmetrised G gives the same sum over ky as G itself
  // (That is Hermitian there)
  float accf = 0.0f;
  auto load = [&](int c, int ky) -> float2 {
    if (c >= ncol) return make_float2(0.f, 0.f);
    const int kx = kx0 + c;
    float gp = 0.0f;
    float2 gv = grad(kx, ky, &gp);
    const bool edge = kx == 0 || kx == L / 2;
    if (edge) {
      const float2 gm = grad(kx, (L - ky) % L, nullptr);
      gv = make_float2(0.5f * (gv.x + gm.x), 0.5f * (gv.y - gm.y));
    }
    accf = fmaf(edge ? 1.0f : 2.0f, gp, accf);
    return gv;
  };
  float2* gb = Gout + (long)b * L * HP;
  auto store = [&](int c, int y, float2 v) {
    if (c < ncol && y >= ylo && y <= yhi) gb[(long)y * HP + kx0 + c] = v;
  };
  fft_exec_split<L, CB, WG, +1, MAP_J, MAP_B, inv_cols_variant<L>()>(lds, twL, load, store);
  double acc[1] = {(double)accf};
  block_sum_n<1>(acc, red);
  if (threadIdx.x == 0) part[(long)b * FSLOTS + blockIdx.x] = acc[0];
}

// mean of g = d/dRn (0: no gradient at DC) and of g Rn (the column partials
// of k_ref_grad_cols over L^2)
__global__ void k_ref_norm_fin(const double* __restrict__ part, int nb, int ncb, double L2,
                               RefScal* __restrict__ sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double gr = 0;
  for (int i = 0; i < ncb; ++i) gr += part[(long)b * FSLOTS + i];
  sc[b].gmean = 0.0;
  sc[b].grn = gr / L2;
}

// Backward: forward column transform of the row R2C of d/dT (d/dA), reduced
// on the spot to the defocus partial sum_half w Re(conj(d/dA) f' B) of
// d score / d defocus through A = f(defocus) B, so d/dA is never stored.
// The row transforms hold rows ylo .. yhi (d/dT is zero outside them).
template <int L>
__global__ __launch_bounds__(WG) void k_ref_cols_df(const float2* __restrict__ rows,
                                                    const float2* __restrict__ white,
                                                    const RefCand* __restrict__ cand,
                                                    WeightArgs w0,
                                                    const float4* __restrict__ wtab,
                                                    const int* __restrict__ skip,
                                                    int ylo, int yhi,
                                                    double* __restrict__ part,
                                                    const float2* __restrict__ twL) {
  using Gm = Geo<L>;
  constexpr int H = Gm::H, CB = Gm::CB;
  __shared__ float lds[CB * exec_stride_split<L, CB, WG, MAP_J>()];
  __shared__ double red[1 * (WG / 64)];
  const int b = blockIdx.y, kx0 = blockIdx.x * CB;
  if (skip[b]) return;  // the whole workgroup
  const RefCand cd = cand[b];
  const float2* base = rows + ((long)b * H + kx0) * L;
  const float2* Bj = white + (long)cd.view * H * L;
  const int ncol = min(CB, H - kx0);
  const float sa = sqrtf(1.0f - w0.ampconst * w0.ampconst);
  auto load = [&](int c, int ky) -> float2 {
    return c < ncol && ky >= ylo && ky <= yhi ? base[(long)c * L + ky] : make_float2(0.f, 0.f);
  };
  // the last radix pass hands each output d/dA(kx, ky) to the reduction
  double acc[1] = {0};
  auto store = [&](int c, int ky, float2 ga) {
    if (c >= ncol) return;
    const int kx = kx0 + c;
    const WeightD wd = weight_tab(wtab[(long)kx * L + ky], cd.df, w0, sa);
    if (wd.fd == 0.0f) return;
    const float2 bv = Bj[(long)kx * L + ky];
    const float wh = (kx == 0 || kx == L / 2) ? 1.0f : 2.0f;
    acc[0] += (double)(wh * wd.fd * (ga.x * bv.x + ga.y * bv.y));
  };
  // consecutive lanes hold consecutive rows of a column in the last pass, so
  // the table and spectrum reads of the reduction are contiguous
  fft_exec_split<L, CB, WG, -1, MAP_J, MAP_J>(lds, twL, load, store);
  block_sum_n<1>(acc, red);
  if (threadIdx.x == 0) part[(long)b * FSLOTS + blockIdx.x] = acc[0];
}

template <int L>
__global__ __launch_bounds__(WG) void k_ref_score(const float2* __restrict__ tiles,
                                                  const float2* __restrict__ spec,
                                                  const RefCand* __restrict__ cand,
                                                  double* __restrict__ part) {
  constexpr int H = L / 2 + 1;
  __shared__ float2 ex[H], ey[L];
  __shared__ double red[4 * (WG / 64)];
  const int b = blockIdx.y;
  const RefCand cd = cand[b];
  phase_tables<L>(ex, ey, (double)cd.x0 + cd.dx, (double)cd.y0 + cd.dy);
  const float2* I = tiles + (long)cd.ptile * H * L;
  const float2* T = spec + (long)b * H * L;
  double acc[4] = {0, 0, 0, 0};  // c, sum w|P|^2, sum w kx Re(i P E), sum w ky' Re(i P E)
  for (long i = (long)blockIdx.x * WG + threadIdx.x; i < (long)H * L; i += (long)RPARTS * WG) {
    const int kx = (int)(i / L), ky = (int)(i - (long)kx * L);
    const float2 P = cmulc(I[i], T[i]);
    const float w = (kx == 0 || kx == L / 2) ? 1.0f : 2.0f;
    const float2 pe = cmul(P, cmul(ex[kx], ey[ky]));
    const float kys = (float)(ky <= L / 2 ? ky : ky - L);
    acc[0] += (double)(w * pe.x);
    acc[1] += (double)(w * (P.x * P.x + P.y * P.y));
    acc[2] += (double)(-w * pe.y * (float)kx);
    acc[3] += (double)(-w * pe.y * kys);
  }
  block_sum_n<4>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) part[((long)b * RPARTS + blockIdx.x) * 4 + k] = acc[k];
}

// (and the skip flag of the adjoint: a trial whose score does not exceed its
// threshold is rejected by the refinement, so its gradient is never used)
__global__ void k_ref_score_fin(const double* __restrict__ part, int nb, int L,
                                RefScal* __restrict__ sc, const float* __restrict__ thr,
                                int* __restrict__ skip) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double a[4] = {0, 0, 0, 0};
  for (int i = 0; i < RPARTS; ++i)
    for (int k = 0; k < 4; ++k) a[k] += part[((long)b * RPARTS + i) * 4 + k];
  const double sigma = sqrt(a[1]);
  sc[b].c = a[0];
  sc[b].sigma = sigma;
  sc[b].score = sigma > 0 ? a[0] / sigma : 0.0;
  const double k2pi = 2.0 * M_PI / L;
  sc[b].gx = sigma > 0 ? k2pi * a[2] / sigma : 0.0;
  sc[b].gy = sigma > 0 ? k2pi * a[3] / sigma : 0.0;
  skip[b] = (thr && !((float)sc[b].score > thr[b])) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
// Row C2R of the gradient intermediate (g L^2 = d/dRn L^2, as
// k_rows_c2r_stream) whose store applies the normalisation and mask adjoint
// per pixel: d/dRm = (g - mean g - Rn mean(g Rn)) / sd, d/d(Tb p) = d/dRm
// mask, written as d/dTb; it also accumulates the sums for d/dpsi (bilinear
// slopes times d(x2, y2)/dpsi) and d/dm, one partial per row block (row
// block blockIdx.x + rb0; outside the launched blocks d/dTb is zero).
template <int L>
__global__ __launch_bounds__(WG) void k_ref_rows_c2r_grad(int rb0, const float2* __restrict__ Gin,
                                                          const float* __restrict__ Tun,
                                                          const double2* __restrict__ cs_b,
                                                          const float* __restrict__ pre,
                                                          const RefScal* __restrict__ sc,
                                                          const int* __restrict__ skip,
                                                          float dmask, float* __restrict__ gTb,
                                                          double* __restrict__ part,
                                                          const float2* __restrict__ twL,
                                                          const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, HP = G::HP, RB = G::RB;
  constexpr long L2 = (long)L * L;
  static_assert((L + RB - 1) / RB <= RPARTS, "partial slots");
  __shared__ float lds[RB * exec_stride_split<M, RB, WG, MAP_J>()];
  __shared__ double red[2 * (WG / 64)];
  const int b = blockIdx.y;
  const int y0 = (blockIdx.x + rb0) * RB;
  if (skip[b]) return;  // the whole workgroup
  const float2* gb = Gin + (long)b * L * HP;
  const double2 cs = cs_b[b];
  const float p = pre[b];
  const double mu = sc[b].mu, isd = 1.0 / sc[b].sd, gm = sc[b].gmean, grn = sc[b].grn;
  const float* src = Tun + b * L2;
  const float kdeg = kRefPi / 180.0f;
  auto load = [&](int r, int k) -> float2 {
    const float2* row = gb + (long)min(y0 + r, L - 1) * HP;
    float2 w = twL[k];
    w.y = -w.y;
    return c2r_pre(row[k], row[M - k], w);
  };
  float* ob = gTb + b * L2;
  double acc[2] = {0, 0};
  auto pixel = [&](int x, int y, float gval) -> float {
    const RotSampleD rs = rot_sample_d<L>(src, x, y, cs, dmask);
    if (!rs.in) return 0.0f;
    const double rn = ((double)(rs.tb * p * rs.mask) - mu) * isd;
    const double g = (double)gval / (double)L2;
    const double grm = (g - gm - rn * grn) * isd;  // d/dRm
    const double gv = grm * rs.mask;                // d/d(Tb p)
    const float out = (float)(gv * p);              // d/dTb
    const float dx2 = (float)(-cs.y * rs.X + cs.x * rs.Y) * kdeg;
    const float dy2 = (float)(-cs.x * rs.X - cs.y * rs.Y) * kdeg;
    acc[0] += (double)out * (double)(rs.dtx * dx2 + rs.dty * dy2);
    acc[1] += gv * (double)rs.tb;
    return out;
  };
  auto store = [&](int r, int n, float2 z) {
    const int y = y0 + r;
    if (y < L) {
      const float a = pixel(2 * n, y, z.x), c = pixel(2 * n + 1, y, z.y);
      *reinterpret_cast<float2*>(ob + (long)y * L + 2 * n) = make_float2(a, c);
    }
  };
  fft_exec_split<M, RB, WG, +1, MAP_J, MAP_J>(lds, twM, load, store);
  block_sum_n<2>(acc, red);
  if (threadIdx.x == 0) {
    part[((long)b * RPARTS + blockIdx.x) * 2] = acc[0];
    part[((long)b * RPARTS + blockIdx.x) * 2 + 1] = acc[1];
  }
}

__global__ void k_ref_rot_fin(const double* __restrict__ part, int nb, int np,
                              RefScal* __restrict__ sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double dp = 0, dpre = 0;
  for (int i = 0; i < np; ++i) {
    dp += part[((long)b * RPARTS + i) * 2];
    dpre += part[((long)b * RPARTS + i) * 2 + 1];
  }
  sc[b].dpsi = dp;
  // R = Tb / m: d/dm = -(sum d/d(Tb/m) Tb) / m^2
  sc[b].dmacc = -dpre / (sc[b].m * sc[b].m);
}

// the bilinear adjoint of d/dTb (rot_adj_gather) for every source pixel of
// rows ylo .. yhi (zero beyond the mask's reach, which the rows hold)
template <int L>
__global__ __launch_bounds__(WG) void k_ref_rot_adj(const float* __restrict__ gTb,
                                                    const double2* __restrict__ cs_b, float dmask,
                                                    const int* __restrict__ skip, int ylo, int yhi,
                                                    float* __restrict__ gT) {
  constexpr long L2 = (long)L * L;
  const int b = blockIdx.y;
  if (skip[b]) return;
  const double2 cs = cs_b[b];
  const float* gsrc = gTb + b * L2;
  const long iend = (long)(yhi + 1) * L;
  for (long i = (long)ylo * L + (long)blockIdx.x * WG + threadIdx.x; i < iend;
       i += (long)RPARTS * WG) {
    const int qy = (int)(i / L), qx = (int)(i - (long)qy * L);
    gT[b * L2 + i] = rot_adj_gather<L>(gsrc, qx, qy, cs, dmask);
  }
}

__global__ void k_ref_out(const double* __restrict__ part, int nb, RefScal* __restrict__ sc,
                          const int* __restrict__ skip, float* __restrict__ out, int np,
                          int stride) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  if (skip[b]) {  // the score only; no gradient was computed
    out[5 * b + 0] = (float)sc[b].score;
    out[5 * b + 1] = out[5 * b + 2] = out[5 * b + 3] = out[5 * b + 4] = 0.0f;
    return;
  }
  double d = 0;
  for (int i = 0; i < np; ++i) d += part[(long)b * stride + i];
  sc[b].ddf = d + sc[b].dmacc * sc[b].dm;
  out[5 * b + 0] = (float)sc[b].score;
  out[5 * b + 1] = (float)sc[b].dpsi;
  out[5 * b + 2] = (float)sc[b].gx;
  out[5 * b + 3] = (float)sc[b].gy;
  out[5 * b + 4] = (float)sc[b].ddf;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define RDISPATCH(L, FN, ...) GS_DISPATCH(L, FN, __VA_ARGS__)

template <int L>
static void ref_forward_t(const RefArgs& a, hipStream_t s) {
  const dim3 grid(RPARTS, a.nb), blk(WG);
  const int small = (a.nb + 63) / 64;
  constexpr int NCB = ncol_blocks<L>();
  // the masked template, its rotation source and the adjoint's images are
  // zero or unread outside a band of rows: the row transforms run on the
  // band, the column transforms write or read only its rows
  const RowBand band = rot_row_band(L, a.dmask);
  const int ylo = band.ylo(L), yhi = band.yhi(L);
  hipLaunchKernelGGL((k_ref_weight_cols<L>), dim3(NCB, a.nb), blk, 0, s, a.white, a.cand, a.w,
                     a.wtab, ylo, yhi, a.gtmp, a.part, a.twL);
  hipLaunchKernelGGL(k_ref_norm_m, dim3(small), dim3(64), 0, s, a.part, a.nb, a.count, a.sc,
                     a.pre, NCB, FSLOTS);
  launch_rows_c2r_stream(L, a.gtmp, a.nb, band, a.Tun, a.twL, a.twM, s);
  launch_rows_r2c_rotb(L, a.Tun, a.nb, a.cs, a.pre, a.dmask, band, a.spec, a.rot_part, a.twL,
                       a.twM, s);
  hipLaunchKernelGGL(k_ref_rot_stats, dim3(small), dim3(64), 0, s, a.rot_part, band.nrb, a.nb,
                     (double)L * L, a.sc, a.bscale);
  launch_cols_fwd(L, a.spec, a.nb, 1.0f, a.bscale, 1, a.twL, s, band);
  hipLaunchKernelGGL((k_ref_score<L>), grid, blk, 0, s, a.tiles, a.spec, a.cand, a.part);
  hipLaunchKernelGGL(k_ref_score_fin, dim3(small), dim3(64), 0, s, a.part, a.nb, L, a.sc, a.thr,
                     a.skip);
}

template <int L>
static void ref_backward_t(const RefArgs& a, hipStream_t s) {
  const dim3 grid(RPARTS, a.nb), blk(WG);
  const int small = (a.nb + 63) / 64;
  constexpr int NCB = ncol_blocks<L>();
  const RowBand band = rot_row_band(L, a.dmask);  // as in the forward pass
  const int ylo = band.ylo(L), yhi = band.yhi(L);
  // d/dThat (with the Parseval sum of the normalisation adjoint) -> the
  // transposed intermediate of the C2R
  hipLaunchKernelGGL((k_ref_grad_cols<L>), dim3(NCB, a.nb), blk, 0, s, a.tiles, a.spec, a.cand,
                     a.sc, a.skip, ylo, yhi, a.gtmp, a.part, a.twL);
  hipLaunchKernelGGL(k_ref_norm_fin, dim3(small), dim3(64), 0, s, a.part, a.nb, NCB,
                     (double)L * L, a.sc);
  // row C2R -> d/dRn -> d/dTb per pixel, d/dpsi and d/dm sums
  hipLaunchKernelGGL((k_ref_rows_c2r_grad<L>), dim3(band.nrb, a.nb), blk, 0, s, band.rb0, a.gtmp,
                     a.Tun, a.cs, a.pre, a.sc, a.skip, a.dmask, a.gTb, a.part, a.twL, a.twM);
  hipLaunchKernelGGL(k_ref_rot_fin, dim3(small), dim3(64), 0, s, a.part, a.nb, band.nrb, a.sc);
  // d/dT: the bilinear adjoint of d/dTb, gathered per source pixel (a
  // separate pass: gathered inside the row transform's loads it took 267 us
  // per batch of 64 at window 720 against 209 for the two kernels)
  hipLaunchKernelGGL((k_ref_rot_adj<L>), grid, blk, 0, s, a.gTb, a.cs, a.dmask, a.skip, ylo, yhi,
                     a.gT);
  // d/dT -> d/dA: unnormalised R2C
  launch_rows_r2c_img(L, a.gT, a.nb, band, a.A, a.twL, a.twM, s, a.skip);
  hipLaunchKernelGGL((k_ref_cols_df<L>), dim3(NCB, a.nb), blk, 0, s, a.A, a.white, a.cand, a.w,
                     a.wtab, a.skip, ylo, yhi, a.part, a.twL);
  hipLaunchKernelGGL(k_ref_out, dim3(small), dim3(64), 0, s, a.part, a.nb, a.sc, a.skip, a.out,
                     NCB, FSLOTS);
}

template <int L>
static void ref_wtable_t(const WeightArgs& w, float4* tab, hipStream_t s) {
  const long n = (long)(L / 2 + 1) * L;
  hipLaunchKernelGGL((k_ref_wtable<L>), dim3((unsigned)((n + WG - 1) / WG)), dim3(WG), 0, s, w,
                     tab);
}
void launch_ref_wtable(int L, const WeightArgs& w, float4* tab, hipStream_t s) {
  RDISPATCH(L, ref_wtable_t, w, tab, s);
}
void launch_ref_forward(int L, const RefArgs& a, hipStream_t s) { RDISPATCH(L, ref_forward_t, a, s); }
void launch_ref_backward(int L, const RefArgs& a, hipStream_t s) {
  RDISPATCH(L, ref_backward_t, a, s);
}
int ref_parts() { return RPARTS; }

}  // namespace gs
