// gs_kernels.hip -- HIP kernels of the MI355X GisSPA picker (gfx950, wave64).
//
// Reference (Autumn1998/GisSPA) correspondence is cited per kernel.  All
// reductions are deterministic: fixed thread-to-element assignment, fixed
// shuffle trees, FP64 where sums cross workgroups, no float atomics.
#include <climits>
#include <cmath>

#include "gs_device.hpp"
#include "gs_sizes.hpp"
#include "gs_fft.hpp"
#include "gs_kernels.hpp"

namespace gs {

// ---------------------------------------------------------------------------
// real-FFT pre/post processing
// ---------------------------------------------------------------------------
// X[k] from Z = DFT_M(x[2m] + i x[2m+1]):  X[k] = E + W^k O,
// E = (Z[k] + conj Z[M-k]) / 2, O = (Z[k] - conj Z[M-k]) / (2i), W = e^{-2 pi i / L}
__device__ __forceinline__ float2 r2c_post(float2 zk, float2 zmk, float2 wk) {
  float2 e = make_float2(0.5f * (zk.x + zmk.x), 0.5f * (zk.y - zmk.y));
  float2 d = make_float2(zk.x - zmk.x, zk.y + zmk.y);  // Z[k] - conj Z[M-k]
  float2 o = make_float2(0.5f * d.y, -0.5f * d.x);       // d / (2i)
  return cadd(e, cmul(wk, o));
}

// ---------------------------------------------------------------------------
// Row R2C: rows y0..y0+RB-1 of image b -> columns [b][kx][y] (transposed)
// ---------------------------------------------------------------------------
enum { SRC_PAD = 0, SRC_TILE = 1, SRC_ROT = 2, SRC_ROTB = 3, SRC_IMG = 4 };

struct RowSrc {
  // SRC_PAD
  const float* box;
  int bx, by;
  const float* scale;
  // SRC_TILE
  const float* img;
  int nx, ny;
  const int* offx;
  const int* offy;
  int bxn;
  const double* mean;
  // SRC_ROT (one angle for every b), SRC_ROTB (angle cs_b[b] = (cos, sin)
  // and a factor pre_b[b] applied to the interpolated value, per b)
  const float* treal;
  float cose, sine, dmask;
  double* partials;
  const double2* cs_b;  // SRC_ROTB: coordinates in double precision
  const float* pre_b;
  // SRC_IMG: real images [b][L][L]
  const float* imgs;
  // any source: rows of b with skip[b] != 0 are left alone (may be null)
  const int* skip;
  // first row block of the launch (row block blockIdx.x + rb0)
  int rb0;
};

template <int L, int SRC>
__device__ __forceinline__ float row_value(const RowSrc& s, int b, int y, int x, float aux) {
  if constexpr (SRC == SRC_PAD) {
    const int sx = (L - s.bx) / 2, sy = (L - s.by) / 2;
    const int i = x - sx, j = y - sy;
    if (i < 0 || i >= s.bx || j < 0 || j >= s.by) return 0.0f;
    return s.box[((long)b * s.by + j) * s.bx + i] * aux;
  } else if constexpr (SRC == SRC_TILE) {
    const int ix = b % s.bxn, iy = b / s.bxn;
    const int gx = s.offx[ix] + x, gy = s.offy[iy] + y;
    if (gx >= s.nx || gy >= s.ny) return 0.0f;
    return (float)((double)s.img[(long)gy * s.nx + gx] - (double)aux);
  } else if constexpr (SRC == SRC_IMG) {
    return s.imgs[((long)b * L + y) * L + x];

  } else if constexpr (SRC == SRC_ROTB) {
    // the refinement's rotation: continuous angle, double-precision
    // coordinates (the cells of the FP64 objective), same mask
    return rot_value_d<L>(s.treal + (long)b * L * L, x, y, s.cs_b[b], s.dmask, s.pre_b[b]);
  } else {
    // rotate_subIMG (GPU_func.cu:496-560) then apply_mask (GPU_func.cu:229-247)
    const float cose = s.cose;
    const float sine = s.sine;
    const float xf = (float)(x - L / 2), yf = (float)(y - L / 2);
    const float r = hypotf(xf, yf);
    const float dm2 = s.dmask * 0.5f;  // d_m = 1.5 * diameter, radius d_m / 2
    const float ehw = 6.0f;
    if (r > dm2 + 2 * ehw) return 0.0f;
    // the reference's float expression, rounded step by step (no FMA), so the
    // out-of-window test and the interpolation cell are those of
    // fl(fl(fl(c x) + fl(s y)) + L/2)
    const float x2 = __fadd_rn(__fadd_rn(__fmul_rn(cose, xf), __fmul_rn(sine, yf)), (float)(L / 2));
    const float y2 = __fadd_rn(__fadd_rn(__fmul_rn(-sine, xf), __fmul_rn(cose, yf)), (float)(L / 2));
    float v = 0.0f;
    if (!(x2 < 0 || x2 > L - 1.0f || y2 < 0 || y2 > L - 1.0f)) {
      const int ii = (int)floorf(x2), jj = (int)floorf(y2);
      long k0 = ii + (long)jj * L, k1 = k0 + 1, k2 = k0 + L + 1, k3 = k0 + L;
      if (ii == L - 1) { k1--; k2--; }
      if (jj == L - 1) { k2 -= L; k3 -= L; }
      const float t = x2 - (float)ii, u = y2 - (float)jj;
      const float tt = 1.0f - t, uu = 1.0f - u;
      const float* src = s.treal + (long)b * L * L;
      const float p0 = src[k0] * tt * uu, p1 = src[k1] * t * uu;
      const float p3 = src[k3] * tt * u, p2 = src[k2] * t * u;
      v = p0 + p1 + p2 + p3;
    }
    if (r >= dm2) v *= 0.5f * cosf(kRefPi * (r - dm2) / (2 * ehw)) + 0.5f;
    return v;
  }
}

template <int L, int SRC>
__global__ __launch_bounds__(WG) void k_rows_r2c(RowSrc s, float2* __restrict__ out,
                                                 const float2* __restrict__ twL,
                                                 const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, H = G::H, RB = G::RB;
  // the transform's split exchange, then the staging of its outputs (one
  // component at a time, rows M + 1 floats apart) for the R2C pairing of
  // Z[k] with Z[M - k]
  constexpr int ZS = M + 1;
  constexpr int LX = RB * exec_stride_split<M, RB, WG, MAP_J>();
  __shared__ float lds[LX > RB * ZS ? LX : RB * ZS];
  __shared__ double red[2 * (WG / 64)];
  const int b = blockIdx.y;
  const int y0 = (blockIdx.x + s.rb0) * RB;
  if (s.skip && s.skip[b]) return;  // the whole workgroup
  float aux = 0.0f;
  if constexpr (SRC == SRC_PAD) aux = s.scale[b];
  double acc1 = 0, acc2 = 0;
  // z[m] = (v(2m), v(2m+1)) of row y0 + r, formed in the first radix pass
  auto load = [&](int r, int m) -> float2 {
    const int y = y0 + r;
    float v0 = 0.0f, v1 = 0.0f;
    if (y < L) {
      if constexpr (SRC == SRC_TILE) {
        const int ix = b % s.bxn, iy = b / s.bxn;
        const int gy = s.offy[iy] + y;
        const int gx0 = s.offx[ix] + 2 * m;
        const double mu = *s.mean;
        if (gy < s.ny) {
          const float* row = s.img + (long)gy * s.nx;
          if (gx0 < s.nx) v0 = (float)((double)row[gx0] - mu);
          if (gx0 + 1 < s.nx) v1 = (float)((double)row[gx0 + 1] - mu);
        }
      } else {
        v0 = row_value<L, SRC>(s, b, y, 2 * m, aux);
        v1 = row_value<L, SRC>(s, b, y, 2 * m + 1, aux);
      }
    }
    if constexpr (SRC == SRC_ROT || SRC == SRC_ROTB) {
      acc1 += (double)v0 + (double)v1;
      acc2 += (double)v0 * v0 + (double)v1 * v1;
    }
    return make_float2(v0, v1);
  };
  constexpr Plan PL = split_plan(M, RB, WG, MAP_J, 0);
  constexpr int RL = PL.r[PL.count - 1];  // radix of the last pass
  constexpr int NSL = M / RL;             // its butterflies per row
  constexpr int TOTL = RB * NSL;
  constexpr int PERL = (TOTL + WG - 1) / WG;
  float2 zv[PERL][RL];
  auto store = [&](int, int, float2 z, int pr, int slot) { zv[pr][slot] = z; };
  fft_exec_split<M, RB, WG, -1, MAP_J, MAP_J>(lds, twM, load, store);
  if constexpr (SRC == SRC_ROT || SRC == SRC_ROTB) {
    block_sum2_d(acc1, acc2, red);  // (its barriers also end the transform's LDS reads)
    if (threadIdx.x == 0) {
      s.partials[2 * ((long)b * gridDim.x + blockIdx.x)] = acc1;
      s.partials[2 * ((long)b * gridDim.x + blockIdx.x) + 1] = acc2;
    }
  } else {
    __syncthreads();  // the transform's last LDS reads end before the staging
  }
  // output (row, n) of round pr, slot sl: row = g / NSL, n = g % NSL + sl NSL,
  // g = thread + pr WG (the last pass's butterfly mapping)
  auto stage = [&](auto comp) {
#pragma unroll
    for (int pr = 0; pr < PERL; ++pr) {
      const int g = threadIdx.x + pr * WG;
      if (TOTL % WG == 0 || g < TOTL) {
        const int r = g / NSL, j = g % NSL;
#pragma unroll
        for (int sl = 0; sl < RL; ++sl)
          lds[r * ZS + j + sl * NSL] = decltype(comp)::value ? zv[pr][sl].y : zv[pr][sl].x;
      }
    }
  };
  // post-process to X[k], k = 0..M, and write out[b][k][y0 + r]: item
  // g = k RB + r, so the RB rows of one k are consecutive threads
  constexpr int NPOST = (RB * H + WG - 1) / WG;
  float zre[NPOST][2];
  stage(std::integral_constant<int, 0>{});
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NPOST; ++i) {
    const int g = threadIdx.x + i * WG;
    if (g < RB * H) {
      const int k = g / RB, r = g - k * RB;
      zre[i][0] = lds[r * ZS + (k % M)];
      zre[i][1] = lds[r * ZS + ((M - k) % M)];
    }
  }
  __syncthreads();
  stage(std::integral_constant<int, 1>{});
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NPOST; ++i) {
    const int g = threadIdx.x + i * WG;
    if (g < RB * H) {
      const int k = g / RB, r = g - k * RB;
      const int y = y0 + r;
      if (y < L) {
        const float2 zk = make_float2(zre[i][0], lds[r * ZS + (k % M)]);
        const float2 zmk = make_float2(zre[i][1], lds[r * ZS + ((M - k) % M)]);
        out[((long)b * H + k) * L + y] = r2c_post(zk, zmk, twL[k]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Column transforms on [b][kx][y]
// ---------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(WG) void k_cols_fwd(float2* __restrict__ spec, float scale,
                                                 const float* __restrict__ bscale, int zero_dc,
                                                 int ylo, int yhi,
                                                 const float2* __restrict__ twL) {
  // in place: every input of the workgroup's columns is read in the first
  // radix pass, before the barrier that precedes the last pass's stores;
  // inputs outside rows ylo .. yhi are zero (not read)
  using G = Geo<L>;
  constexpr int H = G::H, CB = G::CB;
  __shared__ float lds[CB * exec_stride_split<L, CB, WG, MAP_J>()];
  const int b = blockIdx.y;
  const int kx0 = blockIdx.x * CB;
  float2* base = spec + ((long)b * H + kx0) * L;
  const int ncol = min(CB, H - kx0);
  const float sc = bscale ? bscale[b] : scale;
  auto load = [&](int c, int y) -> float2 {
    return c < ncol && y >= ylo && y <= yhi ? base[(long)c * L + y] : make_float2(0.f, 0.f);
  };
  auto store = [&](int c, int y, float2 v) {
    if (c >= ncol) return;
    v = cscale(v, sc);
    if (zero_dc && kx0 == 0 && c == 0 && y == 0) v = make_float2(0.f, 0.f);
    base[(long)c * L + y] = v;
  };
  // the last pass maps consecutive lanes to consecutive rows of a column, so
  // the stores of a wave cover contiguous runs of the column-major layout
  fft_exec_split<L, CB, WG, -1, MAP_J, MAP_J>(lds, twL, load, store);
}

// inverse columns, transposed write to G[b][y][kx] (pitch HP)
template <int L>
__global__ __launch_bounds__(WG) void k_cols_inv_T(const float2* __restrict__ spec,
                                                   float2* __restrict__ Gout,
                                                   const float2* __restrict__ twL) {
  using G = Geo<L>;
  constexpr int H = G::H, HP = G::HP, CB = G::CB;
  __shared__ float2 lds[CB * L];
  __shared__ float2 tw[L];
  const int b = blockIdx.y;
  const int kx0 = blockIdx.x * CB;
  load_twiddles<L, WG>(tw, twL);
  const float2* base = spec + ((long)b * H + kx0) * L;
  const int ncol = min(CB, H - kx0);
  for (int g = threadIdx.x; g < CB * L; g += WG) {
    const int c = g / L;
    lds[g] = (c < ncol) ? base[g] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<L, CB, WG, +1, L>(lds, tw);
  float2* gb = Gout + (long)b * L * HP;
  for (int g = threadIdx.x; g < CB * L; g += WG) {
    const int y = g / CB, c = g - y * CB;
    if (c < ncol) gb[(long)y * HP + kx0 + c] = lds[c * L + y];
  }
}

// row C2R of G[b][y][kx] into real out[b][y][x]
template <int L>
__global__ __launch_bounds__(WG) void k_rows_c2r(const float2* __restrict__ Gin,
                                                 float* __restrict__ out,
                                                 const float2* __restrict__ twL,
                                                 const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, H = G::H, HP = G::HP, RB = G::RB;
  __shared__ float2 lds[RB * H];
  __shared__ float2 tw[M];
  const int b = blockIdx.y;
  const int y0 = blockIdx.x * RB;
  load_twiddles<M, WG>(tw, twM);
  const float2* gb = Gin + (long)b * L * HP;
  for (int g = threadIdx.x; g < RB * H; g += WG) {
    const int r = g / H, k = g - r * H;
    const int y = y0 + r;
    lds[g] = (y < L) ? gb[(long)y * HP + k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  constexpr int NPAIR = M / 2 + 1;
  for (int g = threadIdx.x; g < RB * NPAIR; g += WG) {
    const int r = g / NPAIR, q = g - r * NPAIR;
    float2* row = lds + r * H;
    const int k1 = q, k2 = M - q;
    const float2 x1 = row[k1], x2 = row[k2];
    float2 w1 = twL[k1];
    w1.y = -w1.y;
    row[k1] = c2r_pre(x1, x2, w1);
    if (q > 0 && k2 != k1) {
      float2 w2 = twL[k2];
      w2.y = -w2.y;
      row[k2] = c2r_pre(x2, x1, w2);
    }
  }
  __syncthreads();
  lds_fft<M, RB, WG, +1, H>(lds, tw);
  float* ob = out + (long)b * L * L;
  for (int g = threadIdx.x; g < RB * M; g += WG) {
    const int r = g / M, n = g - r * M;
    const int y = y0 + r;
    if (y >= L) continue;
    const float2 z = lds[r * H + n];
    *reinterpret_cast<float2*>(ob + (long)y * L + 2 * n) = z;
  }
}

// The same row C2R streamed: each thread reads X[k] and X[M - k] of its
// inputs straight from the half spectra and pre-processes them on the way
// into the first transform pass (the executor of the correlation's pass 2),
// then writes its real pairs; no LDS staging of the spectra.  Row block
// blockIdx.x + rb0.
template <int L>
__global__ __launch_bounds__(WG) void k_rows_c2r_stream(const float2* __restrict__ Gin,
                                                        float* __restrict__ out, int rb0,
                                                        const float2* __restrict__ twL,
                                                        const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, HP = G::HP, RB = G::RB;
  __shared__ float lds[RB * exec_stride_split<M, RB, WG, MAP_J>()];
  const int b = blockIdx.y;
  const int y0 = (blockIdx.x + rb0) * RB;
  const float2* gb = Gin + (long)b * L * HP;
  auto load = [&](int r, int k) -> float2 {
    const float2* row = gb + (long)min(y0 + r, L - 1) * HP;
    float2 w = twL[k];
    w.y = -w.y;
    return c2r_pre(row[k], row[M - k], w);
  };
  float* ob = out + (long)b * L * L;
  auto store = [&](int r, int n, float2 z) {
    const int y = y0 + r;
    if (y < L) *reinterpret_cast<float2*>(ob + (long)y * L + 2 * n) = z;
  };
  fft_exec_split<M, RB, WG, +1, MAP_J, MAP_J>(lds, twM, load, store);
}

// ---------------------------------------------------------------------------
// Correlation (compute_corner_CCG GPU_func.cu:584-615, cuFFT inverse
// main.cu:848, get_peak_and_SUM GPU_func.cu:667-708, host loop main.cu:856-881)
// ---------------------------------------------------------------------------
// both passes exchange real and imaginary parts in turn (fft_exec_split):
// one float per element, sizes in float2 units.  The layout depends on the
// FFT schedule variant V (its radices), so each kernel sizes LDS for its own.
template <int L, int V = 0>
constexpr int corr_lds1() {
  return (Geo<L>::CB * exec_stride_split<L, Geo<L>::CB, WG, MAP_B, V>() + 1) / 2;
}
template <int L, int V = 0>
constexpr int corr_lds2() {
  return (Geo<L>::RB * exec_stride_split<L / 2, Geo<L>::RB, WG, MAP_J, V>() + 1) / 2;
}

// Rows of the intermediate that pass 2 reads: those of the row blocks
// blo / RB .. bhi / RB (the other blocks hold border pixels only)
template <int L>
__host__ __device__ __forceinline__ void corr_rows(int blo, int bhi, int* ylo, int* yhi) {
  constexpr int RB = Geo<L>::RB;
  *ylo = (blo / RB) * RB;
  *yhi = (bhi / RB) * RB + RB - 1;
}

// Pass 1 of map m, column block cb: P = tile * conj(template) on the paired
// columns, inverse column transforms, C2R pre-processing, blocked store.
template <int L, bool COH = false, int V = 0>
__device__ __forceinline__ void corr_pass1_body(float2* lds, float* red_q, int cb, int m, int gm,
                                                int ylo, int yhi,
                                                const float2* __restrict__ tiles,
                                                const float2* __restrict__ tspec,
                                                const MapRef* __restrict__ maps,
                                                float2* __restrict__ Gout,
                                                float2* __restrict__ psum,
                                                const float2* __restrict__ twL) {
  using G = Geo<L>;
  constexpr int M = G::M, CB = G::CB, NCB = G::NCB;
  const MapRef mr = maps[m];
  const float2* ib = tiles + (long)mr.tile * (M + 1) * L;
  const float2* tb = tspec + (long)mr.view * (M + 1) * L;
  // blocked intermediate G[m][cb][y][c]: every workgroup writes one
  // contiguous block, rows of CB slots at a time
  float2* gb = Gout + ((long)gm * NCB + cb) * L * CB;
  // P = tile * conj(template), read straight into the first radix pass (the
  // tile spectra carry the factor (-1)^(kx+ky) sgn of compute_corner_CCG,
  // applied once per micrograph).  Slots past the last pair unit read column
  // M and feed only unused outputs.
  //
  // The map's sums follow from Parseval on the way in: sum_u c(u)^2 =
  // L^2 sum_k |P(k)|^2 over the Hermitian-completed spectrum (columns
  // 0 < kx < L/2 count twice, a column in two slots or past M/2 not at all)
  // and sum_u c(u) = L^2 P(0,0).
  float sumsq = 0.0f;
  auto load = [&](int c, int y) -> float2 {
    const int u = c < CB / 2 ? c : CB - 1 - c;
    const int p = cb * (CB / 2) + u;
    const int kx = min(c < CB / 2 ? p : M - p, M);
    const float wgt = (p > M / 2 || (c >= CB / 2 && p == M / 2)) ? 0.0f
                      : ((kx == 0 || kx == M) ? 1.0f : 2.0f);
    const float2 v = cmulc(ib[kx * L + y], tb[kx * L + y]);
    sumsq = fmaf(wgt, fmaf(v.x, v.x, v.y * v.y), sumsq);
    return v;
  };
  // Last pass: the CB slots of one row y sit in CB consecutive lanes, so slot
  // c finds its partner column M - k in lane ^ (CB - 1).  Each lane then
  // writes the C2R pre-processed value
  //   Z'[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) e^{+2 pi i k / L}
  // of its own column k < M (the row transform in pass 2 is then a plain
  // M-point complex FFT of Z').
  // Every slot is written: the slots pass 2 never reads (column M, the
  // second copy of column M/2, units past M/2) only hold unused values.
  // Rows outside ylo..yhi are not read by pass 2 (their row blocks hold
  // border pixels only): output slot r of the last radix pass holds rows
  // r NS .. r NS + NS - 1 of every thread, so a slot wholly outside is
  // skipped by the whole workgroup (a uniform branch; rows of a slot that
  // straddles the bound are written).
  constexpr Plan PL1 = variant_plan(L, CB, WG, MAP_B, V);
  constexpr int NSL1 = L / PL1.r[PL1.count - 1];
  const int cl = threadIdx.x % CB;  // the slot this lane stores (MAP_B)
  float2 wk = twL[pair_col<L>(cb, cl)];  // column index <= M < L
  wk.y = -wk.y;
  auto store = [&](int c, int y, float2 v, int, int r) {
    if (r * NSL1 > yhi || r * NSL1 + NSL1 - 1 < ylo) return;
    const float2 pv = xchg<CB>(v);
    if constexpr (COH) store_coh(gb + y * CB + c, c2r_pre(v, pv, wk));
    else store_nt(gb + y * CB + c, c2r_pre(v, pv, wk));
  };
  // half the LDS of a float2 exchange: four workgroups per CU instead of three
  fft_exec_split<L, CB, WG, +1, MAP_J, MAP_B, V>(reinterpret_cast<float*>(lds), twL, load, store);
  sumsq = wave_sum_f(sumsq);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) red_q[wv] = sumsq;
  __syncthreads();
  if (threadIdx.x == 0) {
    float q = 0.0f;
    for (int i = 0; i < WG / 64; ++i) q += red_q[i];
    const float d = cb == 0 ? cmulc(ib[0], tb[0]).x : 0.0f;  // P(0, 0)
    psum[(long)m * NCB + cb] = make_float2(q, d);
  }
}

// XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so item
// w % 8 runs on XCD w % 8.  Each XCD owns the column blocks cb = xcd + 8 i of
// every map, so an XCD's L2 sees one slice of each spectrum.
template <int L>
__device__ __forceinline__ bool corr_pass1_item(int w, int nmaps, int* cb, int* m) {
  constexpr int NCB = Geo<L>::NCB, NCB8 = (NCB + 7) / 8;
  const int xcd = w & 7, k8 = w >> 3;
  *cb = xcd + 8 * (k8 % NCB8);
  *m = k8 / NCB8;
  return *cb < NCB && *m < nmaps;
}

template <int L, int V>
__global__ __launch_bounds__(WG) void k_corr_pass1(const float2* __restrict__ tiles,
                                                   const float2* __restrict__ tspec,
                                                   const MapRef* __restrict__ maps, int nmaps,
                                                   int ylo, int yhi,
                                                   float2* __restrict__ Gout,
                                                   float2* __restrict__ psum,
                                                   const float2* __restrict__ twL) {
  __shared__ float2 lds[corr_lds1<L, V>()];
  __shared__ float red_q[WG / 64];
  int cb, m;
  if (!corr_pass1_item<L>(blockIdx.x, nmaps, &cb, &m)) return;
  corr_pass1_body<L, false, V>(lds, red_q, cb, m, m, ylo, yhi, tiles, tspec, maps, Gout, psum,
                               twL);
}

// (-1)^(kx+ky) sgn on half spectra [t][kx][ky], in place
template <int L>
__global__ void k_checker_sign(float2* __restrict__ spec, long n, float sgn) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ky = (int)(i % L), kx = (int)((i / L) % (L / 2 + 1));
  spec[i] = cscale(spec[i], ((kx + ky) & 1) ? -sgn : sgn);
}

// per-map finalisation (host loop of main.cu:856-881 in FP64): combine the
// row-block partials in index order, so the result does not depend on which
// workgroup finished last
__device__ void corr_finalize_map(const CorrPartial* p, int nparts, const float2* ps, int ncb,
                                  double L2, double rc, MapStat* out) {
  float v = 0.0f;
  int pos = INT_MAX;
  for (int i = 0; i < nparts; ++i) vmax_combine(v, pos, p[i].vmax, p[i].pos);
  double q = 0;
  for (int i = 0; i < ncb; ++i) q += ps[i].x;
  const double s1 = L2 * (double)ps[0].y;
  const double s2 = L2 * q;
  const double peak = v;
  if (!(v > 0.0f)) pos = 0;
  const double ra = s1 - peak, rb = s2 - peak * peak;
  const double var = rb / rc - (ra / rc) * (ra / rc);
  const double sd = sqrt(var);
  MapStat st;
  st.peak = (float)peak;
  st.pos = pos;
  st.sd = (float)sd;
  st.score = (sd > 0) ? (float)(peak / sd) : 0.0f;
  st.sum = s1;
  st.sum2 = s2;
  *out = st;
}

// Pass 2 of map m, row block rb: row transforms of Z' and the border-
// excluded maximum with its position.
template <int L, bool COH = false, int V = 0>
__device__ __forceinline__ void corr_pass2_body(float2* lds, float* red_v, int* red_p, int m,
                                                int gm, int rb, const float2* __restrict__ Gin,
                                                int blo,
                                                int bhi, CorrPartial* __restrict__ part,
                                                int nparts, const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, RB = G::RB, CB = G::CB, NCB = G::NCB;
  const int y0 = rb * RB;
  const float2* gb = Gin + (long)gm * NCB * L * CB;
  constexpr Plan PL = variant_plan(M, RB, WG, MAP_J, V);
  constexpr int RL = PL.r[PL.count - 1];  // radix of the last pass
  constexpr int NSL = M / RL;             // its butterflies per transform
  if (rb < blo / RB || rb > bhi / RB) {
    // every pixel of the block is a border pixel (value 0, never a strict
    // maximum): the partial the transform would give, without it
    if (threadIdx.x == 0) {
      CorrPartial cp;
      cp.vmax = 0.0f; cp.pos = INT_MAX; cp.s1 = 0.0f; cp.s2 = 0.0f;
      part[(long)m * nparts + rb] = cp;
    }
    return;
  }
  // Z'[k] of row y, already pre-processed by pass 1 (rows past L only exist
  // when RB does not divide L; they are read clamped and never stored)
  auto load = [&](int r, int k) -> float2 {
    const float2* q = gb + pair_slot<L>(k, min(y0 + r, L - 1));
    if constexpr (COH) return load_coh(q);
    // plain loads: the two halves of a 64-byte slot row are read by
    // different loads, and a cached line serves both
    else return *q;
  };
  // The first radix pass reads elements k = j + q M1 (q = 0 .. R0-1) of a
  // row.  With M1 and M multiples of CB/2, the slot of k moves by a fixed
  // C1 bytes per step of q: forward for columns k <= M/2, backward for their
  // partners k > M/2.  So each thread computes two byte offsets once (q = 0
  // and q = R0-1) and every load is a buffer load whose step is a constant
  // scalar offset; only the one q whose columns straddle M/2 selects per
  // lane.  Reads past the map's intermediate return 0 (buffer range check).
  constexpr int R0 = PL.r[0], M1 = M / R0, HCB = CB / 2;
  constexpr bool AFFINE = !COH && PL.count > 1 && M1 % HCB == 0 && M % HCB == 0;
  constexpr int C1 = (M1 / HCB) * L * CB * 8;
  const __amdgpu_buffer_rsrc_t grs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float2*>(gb), 0, NCB * L * CB * 8, 0x00020000);
  int voff_lo = 0, voff_hi = 0;
  auto load_aff = [&](int r, int j, auto qc) -> float2 {
    constexpr int q = decltype(qc)::value;
    constexpr int KLO = q * M1, KHI = q * M1 + M1 - 1;
    if constexpr (q == 0) {
      const int y = min(y0 + r, L - 1);
      voff_lo = pair_slot<L>(j, y) * 8;
      voff_hi = pair_slot<L>(j + (R0 - 1) * M1, y) * 8;
    }
    int vo, so;
    if constexpr (KHI <= M / 2) {
      vo = voff_lo;
      so = q * C1;
    } else if constexpr (KLO > M / 2) {
      vo = voff_hi;
      so = (R0 - 1 - q) * C1;
    } else {
      vo = (j + KLO <= M / 2) ? voff_lo + q * C1 : voff_hi + (R0 - 1 - q) * C1;
      so = 0;
    }
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(grs, vo, so, 0));
  };
  // z'[n] = (c[y][2n], c[y][2n+1]).  Border pixels (x or y outside
  // [blo, bhi], the integer form of get_peak_and_SUM's d/4 test) count as 0.
  // The border tests are uniform: rows per workgroup, columns per output
  // slot.  Each thread visits its pixels in increasing linear index (round,
  // then output slot, then x).
  const unsigned brange = (unsigned)(bhi - blo);
  const bool rows_in = y0 >= blo && y0 + RB - 1 <= bhi;
  auto masked = [&](int r, int n, float2 z, int slot) -> float2 {
    float a = z.x, b = z.y;
    if (!rows_in) {
      if ((unsigned)(y0 + r - blo) > brange) { a = 0.0f; b = 0.0f; }
    }
    const int xlo = 2 * slot * NSL, xhi = 2 * (slot + 1) * NSL - 1;
    if (xlo < blo || xhi > bhi) {
      const int x0 = 2 * n;
      a = ((unsigned)(x0 - blo) <= brange) ? a : 0.0f;
      b = ((unsigned)(x0 + 1 - blo) <= brange) ? b : 0.0f;
    }
    return make_float2(a, b);
  };
  int pos = INT_MAX;
  float bmax = 0.0f;
  if constexpr (AFFINE) {
    // The maximum is taken with plain max operations; its position is found
    // afterwards, only by the threads that hold the workgroup's maximum:
    // their first equal pixel in visiting order is their lowest index.
    constexpr int PERL = (RB * NSL + WG - 1) / WG;  // rounds of the last pass
    float vals[PERL][RL][2];
#pragma unroll
    for (int pr = 0; pr < PERL; ++pr)
#pragma unroll
      for (int sl = 0; sl < RL; ++sl) vals[pr][sl][0] = vals[pr][sl][1] = 0.0f;
    auto store = [&](int r, int n, float2 z, int pr, int slot) {
      const float2 v = masked(r, n, z, slot);
      vals[pr][slot][0] = v.x;
      vals[pr][slot][1] = v.y;
    };
    fft_exec_split<M, RB, WG, +1, MAP_J, MAP_J, V>(reinterpret_cast<float*>(lds), twM, load_aff, store);
    float vmax = 0.0f;
#pragma unroll
    for (int pr = 0; pr < PERL; ++pr)
#pragma unroll
      for (int sl = 0; sl < RL; ++sl)
        vmax = fmaxf(vmax, fmaxf(vals[pr][sl][0], vals[pr][sl][1]));
    float wm = vmax;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) wm = fmaxf(wm, __shfl_xor(wm, o, 64));
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red_v[w] = wm;
    __syncthreads();
    bmax = red_v[0];
#pragma unroll
    for (int i = 1; i < WG / 64; ++i) bmax = fmaxf(bmax, red_v[i]);
    if (bmax > 0.0f && vmax == bmax) {
      int code = 0;
      // last to first, so the first equal pixel's code remains
#pragma unroll
      for (int pr = PERL - 1; pr >= 0; --pr)
#pragma unroll
        for (int sl = RL - 1; sl >= 0; --sl)
#pragma unroll
          for (int h = 1; h >= 0; --h)
            if (vals[pr][sl][h] == bmax) code = (pr * RL + sl) * 2 + h;
      const int pr = code / (2 * RL), slot = (code >> 1) % RL, half = code & 1;
      const int g = threadIdx.x + pr * WG;
      const int row = g / NSL, j = g % NSL;
      pos = (y0 + row) * L + 2 * (j + slot * NSL) + half;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pos = min(pos, __shfl_xor(pos, o, 64));
    if (lane == 0) red_p[w] = pos;
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int i = 0; i < WG / 64; ++i) pos = min(pos, red_p[i]);
    }
  } else {
    // the running maximum records a compile-time code of the winning
    // (round, slot, half); a strict '>' keeps the lowest index among equal
    // values, and the position is decoded once at the end
    float vmax = 0.0f;
    int vcode = -1;
    auto store = [&](int r, int n, float2 z, int pr, int slot) {
      const float2 v = masked(r, n, z, slot);
      const int code = (pr * RL + slot) * 2;
      if (v.x > vmax) { vmax = v.x; vcode = code; }
      if (v.y > vmax) { vmax = v.y; vcode = code + 1; }
    };
    fft_exec_split<M, RB, WG, +1, MAP_J, MAP_J, V>(reinterpret_cast<float*>(lds), twM, load, store);
    if (vcode >= 0) {
      const int pr = vcode / (2 * RL), slot = (vcode >> 1) % RL, half = vcode & 1;
      const int g = threadIdx.x + pr * WG;
      const int row = g / NSL, j = g % NSL;
      pos = (y0 + row) * L + 2 * (j + slot * NSL) + half;
    }
    wave_max_idx(vmax, pos);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red_v[w] = vmax; red_p[w] = pos; }
    __syncthreads();
    if (threadIdx.x == 0) {
      bmax = red_v[0];
      pos = red_p[0];
      for (int i = 1; i < WG / 64; ++i) vmax_combine(bmax, pos, red_v[i], red_p[i]);
    }
  }
  if (threadIdx.x == 0) {
    CorrPartial cp;
    cp.vmax = bmax; cp.pos = pos; cp.s1 = 0.0f; cp.s2 = 0.0f;
    part[(long)m * nparts + rb] = cp;
  }
}

template <int L, int V>
__global__ __launch_bounds__(WG) void k_corr_pass2(const float2* __restrict__ Gin, int blo,
                                                   int bhi, CorrPartial* __restrict__ part,
                                                   const float2* __restrict__ twM) {
  __shared__ float2 lds[corr_lds2<L, V>()];
  __shared__ float red_v[WG / 64];
  __shared__ int red_p[WG / 64];
  corr_pass2_body<L, false, V>(lds, red_v, red_p, blockIdx.y, blockIdx.y, blockIdx.x, Gin, blo,
                               bhi, part, gridDim.x, twM);
}

// Software-pipelined correlation: one launch runs pass 1 of batch k and
// pass 2 of batch k-1 (whose intermediate the previous launch completed).
// Groups of 8 consecutive workgroups (one per XCD) alternate between the two
// kinds, so every CU mixes the load-heavy pass 1 with the transform-heavy
// pass 2; groups past one kind's work go to the other kind.
template <int L>
__global__ __launch_bounds__(WG) void k_corr_fused(
    const float2* __restrict__ tiles, const float2* __restrict__ tspec,
    const MapRef* __restrict__ maps1, int nmaps1, float2* __restrict__ G1,
    float2* __restrict__ psum1, int ng1, const float2* __restrict__ G2, int nmaps2, int ng2,
    int blo, int bhi, CorrPartial* __restrict__ part2, int nparts,
    const float2* __restrict__ twL, const float2* __restrict__ twM) {
  constexpr int LDS = corr_lds1<L>() > corr_lds2<L>() ? corr_lds1<L>() : corr_lds2<L>();
  __shared__ float2 lds[LDS];
  __shared__ float red_f[WG / 64];
  __shared__ int red_i[WG / 64];
  const int g = blockIdx.x >> 3, x = blockIdx.x & 7;
  // the first 2 min(ng1, ng2) groups alternate, the rest belong to the kind
  // with more groups
  const int mn = min(ng1, ng2);
  int kind, gi;
  if (g < 2 * mn) {
    kind = g & 1;
    gi = g >> 1;
  } else {
    kind = ng1 > ng2 ? 0 : 1;
    gi = g - mn;
  }
  const int w = gi * 8 + x;
  if (kind == 0) {
    int cb, m, ylo, yhi;
    if (!corr_pass1_item<L>(w, nmaps1, &cb, &m)) return;
    corr_rows<L>(blo, bhi, &ylo, &yhi);
    corr_pass1_body<L>(lds, red_f, cb, m, m, ylo, yhi, tiles, tspec, maps1, G1, psum1, twL);
  } else {
    const int m = w / nparts, rb = w % nparts;
    if (m >= nmaps2) return;
    corr_pass2_body<L>(lds, red_f, red_i, m, m, rb, G2, blo, bhi, part2, nparts, twM);
  }
}

// Bounded wait of thread 0 until *ctr >= target.  The counter load is an
// agent-scope acquire and every thread then runs an agent-scope acquire
// fence, so the intermediate the producers wrote before their release is
// visible to the whole workgroup.  A wait that runs out (wall-clock budget
// of `ticks` counter ticks) sets *err and gives up, and once *err is set
// every later wait returns at once, so the launch always drains; the host
// then reports the error instead of results.
__device__ __forceinline__ void wait_count(const int* ctr, int target, int* err,
                                           long long ticks) {
  if (threadIdx.x == 0) {
    const unsigned long long t0 = wall_clock64();
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if ((long long)(wall_clock64() - t0) > ticks ||
          __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}
// every thread's stores are released at agent scope before the barrier, then
// thread 0 counts the workgroup (a release increment)
__device__ __forceinline__ void signal_count(int* ctr) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Whole-psi correlation in one launch.  Map m belongs to XCD m % 8: all its
// pass-1 and pass-2 workgroups have blockIdx = 8 i + m % 8, which the
// round-robin dispatch places on that XCD, so the XCDs never wait on each
// other.  XCD x walks rounds r = 0, 1, ...; round r holds pass 1 of map
// 8 r + x and pass 2 of map 8 (r - lag) + x, the two kinds alternating.  The
// intermediate lives in ring slots ((m / 8) % ring) * 8 + m % 8:
//  * pass 2 of map m waits until the NCB pass-1 blocks of m signalled
//    ready[m];
//  * pass 1 of map m waits until the NP pass-2 blocks of map m - 8 ring
//    signalled done[m - 8 ring].
// Every wait is on workgroups of lower index on the same XCD (lag < ring),
// which in-order dispatch has already placed, so the waits end; they are
// bounded anyway (err).  Intermediate accesses are device-coherent, which
// does not depend on that placement.
template <int L>
__global__ __launch_bounds__(WG) void k_corr_stream(
    const float2* __restrict__ tiles, const float2* __restrict__ tspec,
    const MapRef* __restrict__ maps, int nmaps, int lag, int ring, float2* __restrict__ Gring,
    float2* __restrict__ psum, int* __restrict__ ready, int* __restrict__ done, int blo, int bhi,
    CorrPartial* __restrict__ part, int* __restrict__ err, long long wait_ticks,
    const float2* __restrict__ twL, const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int NCB = G::NCB, NCB8 = (NCB + 7) / 8;
  constexpr int NP = (L + G::RB - 1) / G::RB;
  constexpr int S1 = 8 * NCB8, S2 = 8 * ((NP + 7) / 8);  // slots per round of each kind
  constexpr int SR = S1 + S2, MN = S1 < S2 ? S1 : S2;
  constexpr int LDS = corr_lds1<L>() > corr_lds2<L>() ? corr_lds1<L>() : corr_lds2<L>();
  __shared__ float2 lds[LDS];
  __shared__ float red_f[WG / 64];
  __shared__ int red_i[WG / 64];
  const int x = blockIdx.x & 7, i = blockIdx.x >> 3;
  const int r = i / SR, li = i % SR;
  int kind, k;
  if (li < 2 * MN) {
    kind = li & 1;
    k = li >> 1;
  } else {
    kind = S1 > S2 ? 0 : 1;
    k = li - MN;
  }
  if (kind == 0) {
    const int m = 8 * r + x, cb = k;
    if (m >= nmaps || cb >= NCB) return;
    if (m >= 8 * ring) wait_count(done + (m - 8 * ring), NP, err, wait_ticks);
    int ylo, yhi;
    corr_rows<L>(blo, bhi, &ylo, &yhi);
    corr_pass1_body<L, true>(lds, red_f, cb, m, (r % ring) * 8 + x, ylo, yhi, tiles, tspec, maps,
                             Gring,
                             psum, twL);
    signal_count(ready + m);
  } else {
    const int m = 8 * (r - lag) + x, rb = k;
    if (r < lag || m >= nmaps || rb >= NP) return;
    wait_count(ready + m, NCB, err, wait_ticks);
    corr_pass2_body<L, true>(lds, red_f, red_i, m, ((r - lag) % ring) * 8 + x, rb, Gring, blo,
                             bhi, part, NP, twM);
    signal_count(done + m);
  }
}

__global__ void k_fill_int(int* __restrict__ p, long n, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// one thread per map: combine the row-block partials of every map of a psi;
// map m (processing order) lands at out[tile * nviews + view]
__global__ void k_corr_finalize(const CorrPartial* __restrict__ part, int nparts,
                                const float2* __restrict__ psum, int ncb, int nmaps,
                                const MapRef* __restrict__ maps, int nviews, double L2,
                                MapStat* __restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmaps) return;
  const MapRef mr = maps[m];
  corr_finalize_map(part + (long)m * nparts, nparts, psum + (long)m * ncb, ncb, L2, L2 - 1.0,
                    out + (long)mr.tile * nviews + mr.view);
}

// ---------------------------------------------------------------------------
// norm_type 1 ("global" per-pixel normalisation, main.cu:1170-1229): the
// same row C2R as pass 2, then either accumulate c/L^2 and its square per
// tile pixel over all views x psi (add_CCG_to_sum, GPU_func.cu:618-633), or
// write the z-score map (update_CCG, GPU_func.cu:649-662).
// ---------------------------------------------------------------------------

// Accumulation.  The batch is tile-major: maps t * nv .. t * nv + nv - 1 are
// nv views of tile maps[t * nv].tile.  A workgroup transforms the rows of its
// row block for those nv maps in order and keeps the per-pixel sums of c/L^2
// and (c/L^2)^2 in registers: every output of the last transform pass lands
// in the same thread and the same (round, slot) for every map, so the
// accumulator of an output is a compile-time register.  The nv-map partial
// sums (FP32) are then added once to the tile's FP64 accumulators; batches
// run in stream order and each (tile, row block) belongs to one workgroup, so
// every accumulator sees a fixed order of additions.
// 512 threads per workgroup: each thread keeps the sums of fewer outputs
// (the accumulators, not the transform, set this kernel's registers)
constexpr int WGA = 2 * WG;
template <int L>
__global__ __launch_bounds__(WGA) void k_norm1_acc(const float2* __restrict__ Gin,
                                                  const MapRef* __restrict__ maps, int nv,
                                                  double* __restrict__ s1,
                                                  double* __restrict__ s2,
                                                  const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, RB = G::RB, CB = G::CB, NCB = G::NCB;
  constexpr long L2 = (long)L * L;
  constexpr Plan PL = split_plan(M, RB, WGA, MAP_J, 0);
  constexpr int RL = PL.r[PL.count - 1];   // radix of the last pass
  constexpr int NSL = M / RL;              // its butterflies per row
  constexpr int TOT = RB * NSL;
  constexpr int PER = (TOT + WGA - 1) / WGA;
  __shared__ float lds[RB * exec_stride_split<M, RB, WGA, MAP_J>()];
  const int y0 = blockIdx.x * RB;
  const int mfirst = blockIdx.y * nv;
  const int t = maps[mfirst].tile;
  const float inv = 1.0f / (float)L2;
  float acc1[PER][RL][2], acc2[PER][RL][2];
#pragma unroll
  for (int p = 0; p < PER; ++p)
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      acc1[p][r][0] = acc1[p][r][1] = 0.0f;
      acc2[p][r][0] = acc2[p][r][1] = 0.0f;
    }
  for (int i = 0; i < nv; ++i) {
    const float2* gb = Gin + (long)(mfirst + i) * NCB * L * CB;
    auto load = [&](int r, int k) -> float2 { return gb[pair_slot<L>(k, min(y0 + r, L - 1))]; };
    auto store = [&](int, int, float2 z, int p, int r) {
      const float a = z.x * inv, b = z.y * inv;
      acc1[p][r][0] += a;
      acc1[p][r][1] += b;
      acc2[p][r][0] += a * a;
      acc2[p][r][1] += b * b;
    };
    fft_exec_split<M, RB, WGA, +1, MAP_J, MAP_J>(lds, twM, load, store);
    __syncthreads();  // the last pass's LDS reads end before the next map's first writes
  }
  // output (row, n) of round p, slot r: row = g / NSL, n = g % NSL + r * NSL
  // with g = thread + p * WGA (the last pass's butterfly mapping)
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int g = threadIdx.x + p * WGA;
    const int y = y0 + g / NSL;
    if ((TOT % WGA != 0 && g >= TOT) || y >= L) continue;
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const long pix = t * L2 + (long)y * L + 2 * (g % NSL + r * NSL);
      s1[pix] += (double)acc1[p][r][0];
      s1[pix + 1] += (double)acc1[p][r][1];
      s2[pix] += (double)acc2[p][r][0];
      s2[pix + 1] += (double)acc2[p][r][1];
    }
  }
}

// Ordered key of a run candidate: larger z first, then the lower pixel index
// (vmax_combine's rule as one unsigned 64-bit maximum); +0 and -0 share a key.
__device__ __forceinline__ unsigned long long run_key(float v, int pos) {
  v = v == 0.0f ? 0.0f : v;
  const unsigned u = __float_as_uint(v);
  const unsigned o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
  return ((unsigned long long)o << 32) | (unsigned)(~pos);
}
__device__ __forceinline__ float run_key_value(unsigned long long k) {
  const unsigned o = (unsigned)(k >> 32);
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// z-scores of a batch (update_CCG, GPU_func.cu:649-662: (c/L^2 - mean) / sd,
// here times the stored 1/sd) reduced on the fly to
// the maximum of each run of 1024 pixels (get_peak_pos, GPU_func.cu:711-745)
// without writing the z-map: each thread keeps a running maximum of the run
// it is in (its outputs come in increasing pixel order) and folds it into
// the workgroup's per-run keys with an LDS atomic maximum when the run
// changes.  A run spans at most two row blocks (RB * L >= 1024); the row
// block holding its first pixel writes slot 0 of keys[m][run], the next one
// slot 1.
template <int L>
__global__ __launch_bounds__(WG) void k_norm1_zrun(const float2* __restrict__ Gin,
                                                   const MapRef* __restrict__ maps,
                                                   const float4* __restrict__ norm,
                                                   unsigned long long* __restrict__ keys,
                                                   const float2* __restrict__ twM) {
  using G = Geo<L>;
  constexpr int M = G::M, RB = G::RB, CB = G::CB, NCB = G::NCB;
  constexpr long L2 = (long)L * L;
  constexpr int NBLK = (int)((L2 + 1023) / 1024);
  constexpr int NRUN = RB * L / 1024 + 2;
  static_assert(RB * L >= 1024, "a run spans at most two row blocks");
  // real and imaginary parts exchanged in turn (half the LDS of a float2
  // exchange), like the correlation's pass 2
  __shared__ float lds[RB * exec_stride_split<M, RB, WG, MAP_J>()];
  __shared__ unsigned long long rk[NRUN];
  const int m = blockIdx.y;
  const int y0 = blockIdx.x * RB;
  if (y0 >= L) return;  // the grid's padding to a multiple of 8 row blocks
  const long p0 = (long)y0 * L, p1 = (long)min(y0 + RB, L) * L;
  const int k0 = (int)(p0 >> 10), k1 = (int)((p1 - 1) >> 10);
  for (int i = threadIdx.x; i <= k1 - k0; i += WG) rk[i] = 0ull;
  __syncthreads();
  const int t = maps[m].tile;
  const float2* gb = Gin + (long)m * NCB * L * CB;
  auto load = [&](int r, int k) -> float2 { return gb[pair_slot<L>(k, min(y0 + r, L - 1))]; };
  const float inv = 1.0f / (float)L2;
  int cur = -1, cp = 0;
  float cv = 0.0f;
  auto visit = [&](float v, int q) {
    const int k = q >> 10;
    if (k != cur) {
      if (cur >= 0) atomicMax(&rk[cur - k0], run_key(cv, cp));
      cur = k; cv = v; cp = q;
    } else if (v > cv || (v == cv && q < cp)) {
      cv = v; cp = q;
    }
  };
  auto store = [&](int r, int n, float2 z) {
    const int y = y0 + r;
    if (y >= L) return;
    const int pix = y * L + 2 * n;
    const float4 q = norm[(t * L2 + pix) >> 1];   // {mean, mean, 1/sd, 1/sd} of the pair
    visit((z.x * inv - q.x) * q.z, pix);
    visit((z.y * inv - q.y) * q.w, pix + 1);
  };
  fft_exec_split<M, RB, WG, +1, MAP_J, MAP_J>(lds, twM, load, store);
  if (cur >= 0) atomicMax(&rk[cur - k0], run_key(cv, cp));
  __syncthreads();
  for (int i = threadIdx.x; i <= k1 - k0; i += WG) {
    const int k = k0 + i;
    const int slot = ((long)k << 10) >= p0 ? 0 : 1;
    keys[((long)m * NBLK + k) * 2 + slot] = rk[i];
  }
}

// Per run: combine the slots of the row blocks it spans, then the threshold
// and centre bounds of the host loop (main.cu:796-819); records carry the
// map's canonical index mbase + view * ntiles + tile.
template <int L>
__global__ void k_norm1_runfin(const unsigned long long* __restrict__ keys, long long mbase,
                               int ntiles, const MapRef* __restrict__ maps,
                               const TileDesc* __restrict__ tdesc, float thres, float dmin,
                               unsigned long long* __restrict__ count, N1Rec* __restrict__ out,
                               long long cap) {
  constexpr long L2 = (long)L * L;
  constexpr int NBLK = (int)((L2 + 1023) / 1024);
  constexpr int RB = Geo<L>::RB;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (k >= NBLK) return;
  const long first = (long)k << 10, last = min(first + 1023, L2 - 1);
  unsigned long long key = keys[((long)m * NBLK + k) * 2];
  if ((first / L) / RB != (last / L) / RB) key = max(key, keys[((long)m * NBLK + k) * 2 + 1]);
  const float v = run_key_value(key);
  const int p = (int)(~(unsigned)key);
  if (!(v >= thres)) return;
  const TileDesc td = tdesc[maps[m].tile];
  const int cx = td.ox + p % L, cy = td.oy + p / L;
  if (cx >= dmin && cx < td.nx - dmin && cy >= dmin && cy < td.ny - dmin) {
    const unsigned long long i = atomicAdd(count, 1ull);
    if (i < (unsigned long long)cap) {
      N1Rec r;
      r.m = mbase + (long long)maps[m].view * ntiles + maps[m].tile;
      r.pos = p; r.z = v;
      out[i] = r;
    }
  }
}

// set_CCG_mean (GPU_func.cu:635-646): mean and standard deviation per pixel,
// stored per pixel pair as {mean, mean, 1/sd, 1/sd} (0 where sd is not
// positive: update_CCG's z is 0 there)
__global__ void k_norm1_stats(const double* __restrict__ s1, const double* __restrict__ s2,
                              long npairs, double total, float4* __restrict__ norm) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  float r[4];
  for (int h = 0; h < 2; ++h) {
    const double a = s1[2 * i + h] / total;
    const double sd = sqrt(s2[2 * i + h] / total - a * a);
    r[h] = (float)a;
    r[2 + h] = sd > 0 ? (float)(1.0 / sd) : 0.0f;
  }
  norm[i] = make_float4(r[0], r[1], r[2], r[3]);
}

// ---------------------------------------------------------------------------
// Per-psi normalisation scale: mean/var of the rotated masked template over
// L^2 (compute_sum_sqr + main.cu:686-699), divided_by_var, FFT scale 1/L^2.
// ---------------------------------------------------------------------------
__global__ void k_rot_scale(const double* __restrict__ partials, int nparts, int nb, double L2,
                            float* __restrict__ bscale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s1 = 0, s2 = 0;
  for (int i = 0; i < nparts; ++i) {
    s1 += partials[2 * ((long)b * nparts + i)];
    s2 += partials[2 * ((long)b * nparts + i) + 1];
  }
  const double mean = s1 / L2;
  const double var = s2 / L2 - mean * mean;
  double sc = 1.0 / L2;
  if (var != 0) sc /= sqrt(var);
  bscale[b] = (float)sc;
}

// ---------------------------------------------------------------------------
// Template preparation kernels (main.cu:311-474)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_box_sums(const float* __restrict__ src, int npix,
                                                 double* __restrict__ sums) {
  __shared__ double red[2 * (WG / 64)];
  const int b = blockIdx.x;
  const float* p = src + (long)b * npix;
  double a = 0, q = 0;
  for (int i = threadIdx.x; i < npix; i += WG) {
    const double v = p[i];
    a += v;
    q += v * v;
  }
  block_sum2_d(a, q, red);
  if (threadIdx.x == 0) { sums[2 * b] = a; sums[2 * b + 1] = q; }
}

__global__ void k_sigma_scale(const double* __restrict__ sums, int nb, double L2,
                              float* __restrict__ scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double mean = sums[2 * b] / L2;
  double sigma = sqrt(sums[2 * b + 1] / L2 - mean * mean);
  if (sigma < 0 || !isfinite(sigma)) sigma = 0;
  // scale_each (GPU_func.cu:85-94): skipped when |sigma| < 1e-8
  scale[b] = (sigma < 1e-8 && sigma > -1e-8) ? 1.0f : (float)(1.0 / sigma);
}

template <int L>
__global__ __launch_bounds__(WG) void k_radial_power(const float2* __restrict__ spec,
                                                     const int* __restrict__ bin_start,
                                                     const int* __restrict__ bin_pix,
                                                     double* __restrict__ ra) {
  using G = Geo<L>;
  constexpr int H = G::H, NB = L / 2;
  const int b = blockIdx.x;
  const float2* sp = spec + (long)b * H * L;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int r = w; r < NB; r += WG / 64) {
    double acc = 0;
    for (int i = bin_start[r] + lane; i < bin_start[r + 1]; i += 64) {
      const float2 v = sp[bin_pix[i]];
      acc += (double)v.x * v.x + (double)v.y * v.y;
    }
    acc = wave_sum_d(acc);
    if (lane == 0) ra[(long)b * NB + r] = acc;
  }
}


// whiten_Tmp (GPU_func.cu:130-152) + apply_weighting_function (249-290) +
// compute_area_sum_ofSQR partials (318-362); spectrum pixel (kx, ky) with
// kx <= L/2 is the reference's (x, y).
template <int L>
__global__ __launch_bounds__(WG) void k_tmpl_weight(float2* __restrict__ spec,
                                                    const double* __restrict__ ra,
                                                    const float* __restrict__ rbins, WeightArgs w,
                                                    double* __restrict__ partials, int weight) {
  using G = Geo<L>;
  constexpr int H = G::H, NB = L / 2;
  __shared__ double red[2 * (WG / 64)];
  const int b = blockIdx.y;
  const long i = (long)blockIdx.x * WG + threadIdx.x;
  double acc = 0, dummy = 0;
  if (i < (long)H * L) {
    const int kx = (int)(i / L), ky = (int)(i - (long)kx * L);
    float2 v = spec[(long)b * H * L + i];
    const float hy = (float)min(ky, L - ky), hx = (float)min(kx, L - kx);
    const float rr = hypotf(hy, hx);
    const int rbin = (int)floorf(rr + 0.5f) - 1;
    if (rbin < NB && rbin >= 0) {
      const double fb = ra[(long)b * NB + rbin] / (double)rbins[rbin];
      v = cscale(v, (float)(1.0 / sqrt(fb)));
    }
    // low pass and weighting (weight == 0: the whitened spectrum only, the
    // refinement's starting point; every thread still reaches the reduction)
    const int r_round = rbin;
    const float hi = L * w.apix / w.highres, lo = L * w.apix / w.lowres;
    float f = 1.0f;
    if (!weight) {
    } else if (r_round < hi && r_round >= 0) {
    } else if (r_round >= hi && r_round < hi + 8) {
      f = 0.5f * cosf(kRefPi * (r_round - hi) / (2 * 8)) + 0.5f;
    } else if (r_round >= (lo - 8) && r_round < lo && r_round >= 0) {
      f = 0.5f * cosf(kRefPi * (lo - r_round) / (2 * 8)) + 0.5f;
    } else {
      f = 0.0f;
    }
    const float ss = rr * w.ds;
    if (weight && r_round < L / 2 && r_round >= 0) {
      const float vv = ctf_ast_f(kx, (ky + L / 2) % L, L, L, w, 2);
      const float signal = expf(w.bf * ss * ss + w.bf2 * ss + w.bf3) / (w.kk + 1);
      const float ncurve = expf(w.a * ss * ss + w.b * ss + w.b2) / signal;
      f = f * vv * sqrtf(1 / (ncurve + w.kk * vv * vv));
    }
    v = cscale(v, f);
    spec[(long)b * H * L + i] = v;
    if (rbin < NB && rbin >= 0) acc = (double)v.x * v.x + (double)v.y * v.y;
  }
  block_sum2_d(acc, dummy, red);
  if (threadIdx.x == 0) partials[(long)b * gridDim.x + blockIdx.x] = acc;
}

// amplitude normaliser m = sqrt(S) / count per template (main.cu:444-452)
__global__ void k_tmpl_ampscale(const double* __restrict__ partials, int nparts, int nb,
                                double count, float* __restrict__ inv_m) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double S = 0;
  for (int i = 0; i < nparts; ++i) S += partials[(long)b * nparts + i];
  const double m = sqrt(S / (count * count));
  inv_m[b] = (m != 0) ? (float)(1.0 / m) : 1.0f;
}

template <int L>
__global__ __launch_bounds__(WG) void k_tmpl_ampnorm(float2* __restrict__ spec,
                                                     const float* __restrict__ inv_m) {
  using G = Geo<L>;
  constexpr int H = G::H;
  const int b = blockIdx.y;
  const long i = (long)blockIdx.x * WG + threadIdx.x;
  if (i < (long)H * L) spec[(long)b * H * L + i] = cscale(spec[(long)b * H * L + i], inv_m[b]);
}

// ---------------------------------------------------------------------------
// Micrograph statistics (init_d_image, main.cu:523-545) in FP64
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(WG) void k_img_sums(const float* __restrict__ img, long n,
                                                 double* __restrict__ partials) {
  __shared__ double red[2 * (WG / 64)];
  const long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long beg = blockIdx.x * chunk, end = min(n, beg + chunk);
  double a = 0, q = 0;
  for (long i = beg + threadIdx.x; i < end; i += WG) {
    const double v = img[i];
    a += v;
    q += v * v;
  }
  block_sum2_d(a, q, red);
  if (threadIdx.x == 0) { partials[2 * blockIdx.x] = a; partials[2 * blockIdx.x + 1] = q; }
}

// stats: [0] avg, [1] var (std), [2] up bound, [3] low bound, [4] clamp flag
__global__ void k_img_bounds(const double* __restrict__ partials, int nblocks, long n,
                             double* __restrict__ stats) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0, q = 0;
  for (int i = 0; i < nblocks; ++i) { a += partials[2 * i]; q += partials[2 * i + 1]; }
  const double avg = a / (double)n;
  const double s2 = q / (double)n;
  const double var = sqrt(s2 - avg * avg);
  stats[0] = avg;
  stats[1] = var;
  stats[2] = (double)(int)(avg + 6 * var);  // int truncation, main.cu:539-540
  stats[3] = (double)(int)(avg - 6 * var);
  stats[4] = (s2 > avg * avg) ? 1.0 : 0.0;
}

__global__ __launch_bounds__(WG) void k_img_clamp(float* __restrict__ img, long n, int invert,
                                                  const double* __restrict__ stats,
                                                  double* __restrict__ partials) {
  __shared__ double red[2 * (WG / 64)];
  const double avg = stats[0], up = stats[2], low = stats[3];
  const bool clamp = stats[4] != 0.0;
  const long chunk = (n + gridDim.x - 1) / gridDim.x;
  const long beg = blockIdx.x * chunk, end = min(n, beg + chunk);
  double a = 0, dummy = 0;
  for (long i = beg + threadIdx.x; i < end; i += WG) {
    float v = img[i];
    if (clamp && ((double)v > up || (double)v < low)) v = (float)avg;
    if (invert == 1) v = -v;
    img[i] = v;
    a += (double)v;
  }
  block_sum2_d(a, dummy, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

__global__ void k_img_mean(const double* __restrict__ partials, int nblocks, long n,
                           double* __restrict__ mean) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0;
  for (int i = 0; i < nblocks; ++i) a += partials[i];
  *mean = a / (double)n;
}

// ---------------------------------------------------------------------------
// host launchers, dispatched over the supported tile sizes
// ---------------------------------------------------------------------------

bool geometry_supported(int L) {
  switch (L) {
#define CASE(n) case n:
    GS_SIZES(CASE)
#undef CASE
    return true;
    default:
      return false;
  }
}

Geometry geometry_of(int L) {
  Geometry g{};
  switch (L) {
#define CASE(n)                                                                        \
  case n:                                                                              \
    g.L = n; g.M = Geo<n>::M; g.H = Geo<n>::H; g.HP = Geo<n>::HP; g.RB = Geo<n>::RB;   \
    g.CB = Geo<n>::CB;                                                                 \
    g.gmap = (long)Geo<n>::NCB * Geo<n>::CB * n;                                       \
    break;
    GS_SIZES(CASE)
#undef CASE
    default:
      break;
  }
  return g;
}

#define DISPATCH(L, FN, ...) GS_DISPATCH(L, FN, __VA_ARGS__)

template <int L>
static void rows_r2c_pad_t(const float* src, int bx, int by, const float* scale, int nb,
                           float2* out, const float2* twL, const float2* twM, hipStream_t s) {
  RowSrc rs{};
  rs.box = src; rs.bx = bx; rs.by = by; rs.scale = scale;
  dim3 grid((L + Geo<L>::RB - 1) / Geo<L>::RB, nb);
  hipLaunchKernelGGL((k_rows_r2c<L, SRC_PAD>), grid, dim3(WG), 0, s, rs, out, twL, twM);
}
void launch_rows_r2c_pad(int L, const float* src, int bx, int by, const float* scale, int nb,
                         float2* out, const float2* twL, const float2* twM, hipStream_t s) {
  DISPATCH(L, rows_r2c_pad_t, src, bx, by, scale, nb, out, twL, twM, s);
}

template <int L>
static void rows_r2c_tiles_t(const float* img, int nx, int ny, const int* offx, const int* offy,
                             int bxn, int ntiles, const double* mean, float2* out,
                             const float2* twL, const float2* twM, hipStream_t s) {
  RowSrc rs{};
  rs.img = img; rs.nx = nx; rs.ny = ny; rs.offx = offx; rs.offy = offy; rs.bxn = bxn;
  rs.mean = mean;
  dim3 grid((L + Geo<L>::RB - 1) / Geo<L>::RB, ntiles);
  hipLaunchKernelGGL((k_rows_r2c<L, SRC_TILE>), grid, dim3(WG), 0, s, rs, out, twL, twM);
}
void launch_rows_r2c_tiles(int L, const float* img, int nx, int ny, const int* offx,
                           const int* offy, int bxn, int ntiles, const double* mean,
                           float2* out, const float2* twL, const float2* twM, hipStream_t s) {
  DISPATCH(L, rows_r2c_tiles_t, img, nx, ny, offx, offy, bxn, ntiles, mean, out, twL, twM, s);
}

template <int L>
static void rows_r2c_rot_t(const float* treal, int nb, float cose, float sine, float dmask,
                           RowBand band, float2* out, double* partials, const float2* twL,
                           const float2* twM, hipStream_t s) {
  RowSrc rs{};
  rs.treal = treal; rs.cose = cose; rs.sine = sine; rs.dmask = dmask; rs.partials = partials;
  rs.rb0 = band.rb0;
  dim3 grid(band.nrb, nb);
  hipLaunchKernelGGL((k_rows_r2c<L, SRC_ROT>), grid, dim3(WG), 0, s, rs, out, twL, twM);
}
void launch_rows_r2c_rot(int L, const float* treal, int nb, float cose, float sine, float dmask,
                         RowBand band, float2* out, double* partials, const float2* twL,
                         const float2* twM, hipStream_t s) {
  DISPATCH(L, rows_r2c_rot_t, treal, nb, cose, sine, dmask, band, out, partials, twL, twM, s);
}

template <int L>
static void rows_r2c_rotb_t(const float* treal, int nb, const double2* cs_b, const float* pre_b,
                            float dmask, RowBand band, float2* out, double* partials,
                            const float2* twL, const float2* twM, hipStream_t s) {
  RowSrc rs{};
  rs.treal = treal; rs.cs_b = cs_b; rs.pre_b = pre_b; rs.dmask = dmask; rs.partials = partials;
  rs.rb0 = band.rb0;
  dim3 grid(band.nrb, nb);
  hipLaunchKernelGGL((k_rows_r2c<L, SRC_ROTB>), grid, dim3(WG), 0, s, rs, out, twL, twM);
}
void launch_rows_r2c_rotb(int L, const float* treal, int nb, const double2* cs_b, const float* pre_b,
                          float dmask, RowBand band, float2* out, double* partials,
                          const float2* twL, const float2* twM, hipStream_t s) {
  DISPATCH(L, rows_r2c_rotb_t, treal, nb, cs_b, pre_b, dmask, band, out, partials, twL, twM, s);
}

template <int L>
static void rows_r2c_img_t(const float* imgs, int nb, RowBand band, float2* out,
                           const float2* twL, const float2* twM, hipStream_t s,
                           const int* skip) {
  RowSrc rs{};
  rs.imgs = imgs;
  rs.skip = skip;
  rs.rb0 = band.rb0;
  dim3 grid(band.nrb, nb);
  hipLaunchKernelGGL((k_rows_r2c<L, SRC_IMG>), grid, dim3(WG), 0, s, rs, out, twL, twM);
}
void launch_rows_r2c_img(int L, const float* imgs, int nb, RowBand band, float2* out,
                         const float2* twL, const float2* twM, hipStream_t s, const int* skip) {
  DISPATCH(L, rows_r2c_img_t, imgs, nb, band, out, twL, twM, s, skip);
}

template <int L>
static void cols_fwd_t(float2* spec, int nb, float scale, const float* bscale, int zero_dc,
                       RowBand band, const float2* twL, hipStream_t s) {
  dim3 grid((Geo<L>::H + Geo<L>::CB - 1) / Geo<L>::CB, nb);
  hipLaunchKernelGGL((k_cols_fwd<L>), grid, dim3(WG), 0, s, spec, scale, bscale, zero_dc,
                     band.ylo(L), band.yhi(L), twL);
}
void launch_cols_fwd(int L, float2* spec, int nb, float scale, const float* bscale, int zero_dc,
                     const float2* twL, hipStream_t s, RowBand band) {
  if (band.nrb == 0) band = full_band(L);
  DISPATCH(L, cols_fwd_t, spec, nb, scale, bscale, zero_dc, band, twL, s);
}

void launch_rot_scale(int L, const double* partials, int nb, int nparts, float* bscale,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_rot_scale, dim3((nb + 255) / 256), dim3(256), 0, s, partials, nparts, nb,
                     (double)L * L, bscale);
}

RowBand full_band(int L) {
  const int RB = geometry_of(L).RB;
  RowBand b;
  b.RB = RB;
  b.rb0 = 0;
  b.nrb = (L + RB - 1) / RB;
  return b;
}

RowBand rot_row_band(int L, float dmask) {
  RowBand b = full_band(L);
  const float h = 0.5f * dmask;
  // NaN or a mask wider than the window: every row
  if (!(h < (float)L)) return b;
  const int hw = h >= 0.0f ? (int)std::ceil(h) + 20 : 20;
  const int y0 = std::max(0, L / 2 - hw), y1 = std::min(L - 1, L / 2 + hw);
  b.rb0 = y0 / b.RB;
  b.nrb = y1 / b.RB - b.rb0 + 1;
  return b;
}

template <int L>
static void cols_inv_T_t(const float2* spec, int nb, float2* G, const float2* twL, hipStream_t s) {
  dim3 grid((Geo<L>::H + Geo<L>::CB - 1) / Geo<L>::CB, nb);
  hipLaunchKernelGGL((k_cols_inv_T<L>), grid, dim3(WG), 0, s, spec, G, twL);
}
void launch_cols_inv_T(int L, const float2* spec, int nb, float2* G, const float2* twL,
                       hipStream_t s) {
  DISPATCH(L, cols_inv_T_t, spec, nb, G, twL, s);
}

template <int L>
static void rows_c2r_t(const float2* G, int nb, float* out, const float2* twL, const float2* twM,
                       hipStream_t s) {
  dim3 grid((L + Geo<L>::RB - 1) / Geo<L>::RB, nb);
  hipLaunchKernelGGL((k_rows_c2r<L>), grid, dim3(WG), 0, s, G, out, twL, twM);
}
template <int L>
static void rows_c2r_stream_t(const float2* G, int nb, RowBand band, float* out,
                              const float2* twL, const float2* twM, hipStream_t s) {
  dim3 grid(band.nrb, nb);
  hipLaunchKernelGGL((k_rows_c2r_stream<L>), grid, dim3(WG), 0, s, G, out, band.rb0, twL, twM);
}
void launch_rows_c2r_stream(int L, const float2* G, int nb, RowBand band, float* out,
                            const float2* twL, const float2* twM, hipStream_t s) {
  DISPATCH(L, rows_c2r_stream_t, G, nb, band, out, twL, twM, s);
}
void launch_rows_c2r(int L, const float2* G, int nb, float* out, const float2* twL,
                     const float2* twM, hipStream_t s) {
  DISPATCH(L, rows_c2r_t, G, nb, out, twL, twM, s);
}

// windows whose correlation passes are compiled in every schedule variant
// (the others run variant 0)
constexpr bool has_variants(int L) { return L == 512 || L == 720; }

template <int L>
static void corr_pass1_t(const float2* tiles, const float2* tspec, const MapRef* maps, int nmaps,
                         int ylo, int yhi, float2* G, float2* psum, const float2* twL, int v,
                         hipStream_t s) {
  constexpr int NCB8 = (Geo<L>::NCB + 7) / 8;
  dim3 grid(8 * NCB8 * nmaps);
#define P1(V_) hipLaunchKernelGGL((k_corr_pass1<L, V_>), grid, dim3(WG), 0, s, tiles, tspec, maps, \
                                  nmaps, ylo, yhi, G, psum, twL)
  if constexpr (has_variants(L)) {
    switch (v) {
      case 1: P1(1); return;
      case 2: P1(2); return;
      case 3: P1(3); return;
      default: break;
    }
  }
  P1(0);
#undef P1
}

template <int L>
static void corr_fused_t(const float2* tiles, const float2* tspec, const MapRef* maps1, int nmaps1,
                         float2* G1, float2* psum1, const float2* G2, int nmaps2, int blo,
                         int bhi, CorrPartial* part2, const float2* twL, const float2* twM,
                         hipStream_t s) {
  constexpr int NCB8 = (Geo<L>::NCB + 7) / 8;
  constexpr int NP = (L + Geo<L>::RB - 1) / Geo<L>::RB;
  const int ng1 = NCB8 * nmaps1;              // groups of 8 pass-1 items
  const int ng2 = (NP * nmaps2 + 7) / 8;      // groups of 8 pass-2 items
  const int groups = ng1 + ng2;
  if (groups == 0) return;
  hipLaunchKernelGGL((k_corr_fused<L>), dim3(8 * groups), dim3(WG), 0, s, tiles, tspec, maps1,
                     nmaps1, G1, psum1, ng1, G2, nmaps2, ng2, blo, bhi, part2, NP, twL, twM);
}
void launch_corr_pass1(int L, const float2* tiles, const float2* tspec, const MapRef* maps,
                       int nmaps, int ylo, int yhi, float2* G, float2* psum, const float2* twL,
                       hipStream_t s, int variant) {
  DISPATCH(L, corr_pass1_t, tiles, tspec, maps, nmaps, ylo, yhi, G, psum, twL, variant, s);
}

int corr_variants(int L) { return has_variants(L) ? kPlanVariants : 1; }
int corr_max_variants() { return kPlanVariants; }

void launch_checker_sign(int L, float2* spec, int ntiles, hipStream_t s) {
  const long n = (long)ntiles * (L / 2 + 1) * L;
  const float sgn = ((L / 2) % 2 == 1) ? -1.0f : 1.0f;  // compute_corner_CCG's `of` rule
  const unsigned nb = (unsigned)((n + 255) / 256);
#define CS(n_) case n_: hipLaunchKernelGGL((k_checker_sign<n_>), dim3(nb), dim3(256), 0, s, spec, n, sgn); break;
  switch (L) { GS_SIZES(CS) default: break; }
#undef CS
}

template <int L>
static void corr_pass2_t(const float2* G, int nmaps, int blo, int bhi, CorrPartial* part,
                         const float2* twL, const float2* twM, int v, hipStream_t s) {
  dim3 grid((L + Geo<L>::RB - 1) / Geo<L>::RB, nmaps);
#define P2(V_) hipLaunchKernelGGL((k_corr_pass2<L, V_>), grid, dim3(WG), 0, s, G, blo, bhi, part, twM)
  if constexpr (has_variants(L)) {
    switch (v) {
      case 1: P2(1); return;
      case 2: P2(2); return;
      case 3: P2(3); return;
      default: break;
    }
  }
  P2(0);
#undef P2
}
template <int L>
static void corr_stream_t(const float2* tiles, const float2* tspec, const MapRef* maps, int nmaps,
                          int lag, int ring, float2* Gring, float2* psum, int* ready, int* done,
                          int blo, int bhi, CorrPartial* part, int* err, long long wait_ticks,
                          const float2* twL, const float2* twM, hipStream_t s) {
  using G = Geo<L>;
  constexpr int NP = (L + G::RB - 1) / G::RB;
  constexpr int SR = 8 * ((G::NCB + 7) / 8) + 8 * ((NP + 7) / 8);
  const long nctr = 2l * nmaps;
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((nctr + 255) / 256)), dim3(256), 0, s, ready,
                     nctr, 0);
  const long rounds = (nmaps + 7) / 8 + lag;
  const long nwg = 8l * SR * rounds;
  hipLaunchKernelGGL((k_corr_stream<L>), dim3((unsigned)nwg), dim3(WG), 0, s, tiles, tspec, maps,
                     nmaps, lag, ring, Gring, psum, ready, done, blo, bhi, part, err, wait_ticks,
                     twL, twM);
}

// integer form of get_peak_and_SUM's exclusion: x < lo || x > hi
static void border_range(int L, float lo, float hi, int* blo, int* bhi) {
  *blo = (int)ceilf(lo);
  *bhi = (int)floorf(hi);
  if (*bhi < *blo) { *blo = L + 1; *bhi = L + 1; }  // nothing valid
}
void corr_store_rows(int L, float lo, float hi, int* ylo, int* yhi) {
  int blo, bhi;
  border_range(L, lo, hi, &blo, &bhi);
  GS_DISPATCH(L, corr_rows, blo, bhi, ylo, yhi);
}
void launch_corr_pass2(int L, const float2* G, int nmaps, float lo, float hi, CorrPartial* part,
                       const float2* twL, const float2* twM, hipStream_t s, int variant) {
  int blo, bhi;
  border_range(L, lo, hi, &blo, &bhi);
  DISPATCH(L, corr_pass2_t, G, nmaps, blo, bhi, part, twL, twM, variant, s);
}
void launch_fill_int(int* p, long n, int v, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}
void launch_corr_stream(int L, const float2* tiles, const float2* tspec, const MapRef* maps,
                        int nmaps, int lag, int ring, float2* Gring, float2* psum, int* counters,
                        float lo, float hi, CorrPartial* part, int* err, long long wait_ticks,
                        const float2* twL, const float2* twM, hipStream_t s) {
  int blo, bhi;
  border_range(L, lo, hi, &blo, &bhi);
  DISPATCH(L, corr_stream_t, tiles, tspec, maps, nmaps, lag, ring, Gring, psum, counters,
           counters + nmaps, blo, bhi, part, err, wait_ticks, twL, twM, s);
}
void launch_corr_fused(int L, const float2* tiles, const float2* tspec, const MapRef* maps1,
                       int nmaps1, float2* G1, float2* psum1, const float2* G2, int nmaps2,
                       float lo, float hi, CorrPartial* part2, const float2* twL,
                       const float2* twM, hipStream_t s) {
  int blo, bhi;
  border_range(L, lo, hi, &blo, &bhi);
  DISPATCH(L, corr_fused_t, tiles, tspec, maps1, nmaps1, G1, psum1, G2, nmaps2, blo, bhi, part2,
           twL, twM, s);
}

int corr_nparts(int L) { return (L + geometry_of(L).RB - 1) / geometry_of(L).RB; }

int corr_ncb(int L) {
  const Geometry g = geometry_of(L);
  return (int)(g.gmap / ((long)g.CB * g.L));
}

void launch_corr_finalize(int L, const CorrPartial* part, const float2* psum, int nmaps,
                          const MapRef* maps, int nviews, MapStat* out, hipStream_t s) {
  hipLaunchKernelGGL(k_corr_finalize, dim3((nmaps + 255) / 256), dim3(256), 0, s, part,
                     corr_nparts(L), psum, corr_ncb(L), nmaps, maps, nviews, (double)L * L, out);
}

void launch_box_sums(const float* src, int npix, int nb, double* sums, hipStream_t s) {
  hipLaunchKernelGGL(k_box_sums, dim3(nb), dim3(WG), 0, s, src, npix, sums);
}
void launch_sigma_scale(const double* sums, int nb, int L, float* scale, hipStream_t s) {
  hipLaunchKernelGGL(k_sigma_scale, dim3((nb + 255) / 256), dim3(256), 0, s, sums, nb,
                     (double)L * L, scale);
}

template <int L>
static void radial_power_t(const float2* spec, int nb, const int* bin_start, const int* bin_pix,
                           double* ra, hipStream_t s) {
  hipLaunchKernelGGL((k_radial_power<L>), dim3(nb), dim3(WG), 0, s, spec, bin_start, bin_pix, ra);
}
void launch_radial_power(int L, const float2* spec, int nb, const int* bin_start,
                         const int* bin_pix, double* ra, hipStream_t s) {
  DISPATCH(L, radial_power_t, spec, nb, bin_start, bin_pix, ra, s);
}

template <int L>
static void tmpl_weight_t(float2* spec, int nb, const double* ra, const float* rb,
                          const WeightArgs& w, double* partials, int weight, hipStream_t s) {
  const long tot = (long)Geo<L>::H * L;
  dim3 grid((unsigned)((tot + WG - 1) / WG), nb);
  hipLaunchKernelGGL((k_tmpl_weight<L>), grid, dim3(WG), 0, s, spec, ra, rb, w, partials, weight);
}
void launch_tmpl_weight(int L, float2* spec, int nb, const double* ra, const float* rb,
                        const WeightArgs& w, double* partials, int weight, hipStream_t s) {
  DISPATCH(L, tmpl_weight_t, spec, nb, ra, rb, w, partials, weight, s);
}

template <int L>
static void tmpl_ampnorm_t(float2* spec, int nb, const double* partials, double count,
                           float* inv_m, hipStream_t s) {
  const long tot = (long)Geo<L>::H * L;
  const int nparts = (int)((tot + WG - 1) / WG);
  hipLaunchKernelGGL(k_tmpl_ampscale, dim3((nb + 63) / 64), dim3(64), 0, s, partials, nparts, nb,
                     count, inv_m);
  hipLaunchKernelGGL((k_tmpl_ampnorm<L>), dim3(nparts, nb), dim3(WG), 0, s, spec, inv_m);
}
void launch_tmpl_ampnorm(int L, float2* spec, int nb, const double* partials, double count,
                         float* inv_m, hipStream_t s) {
  DISPATCH(L, tmpl_ampnorm_t, spec, nb, partials, count, inv_m, s);
}

int img_reduce_blocks(long n) {
  long b = (n + 65535) / 65536;
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (int)b;
}
void launch_img_sums(const float* img, long n, double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_img_sums, dim3(nblocks), dim3(WG), 0, s, img, n, partials);
}
void launch_img_bounds(const double* partials, int nblocks, long n, double* stats, hipStream_t s) {
  hipLaunchKernelGGL(k_img_bounds, dim3(1), dim3(64), 0, s, partials, nblocks, n, stats);
}
void launch_img_clamp(float* img, long n, int invert, const double* stats, double* partials,
                      int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(k_img_clamp, dim3(nblocks), dim3(WG), 0, s, img, n, invert, stats, partials);
}
void launch_img_mean(const double* partials, int nblocks, long n, double* mean, hipStream_t s) {
  hipLaunchKernelGGL(k_img_mean, dim3(1), dim3(64), 0, s, partials, nblocks, n, mean);
}

template <int L>
static void norm1_acc_t(const float2* G, const MapRef* maps, int ntiles, int nv, double* s1,
                        double* s2, const float2* twM, hipStream_t s) {
  dim3 grid((L + Geo<L>::RB - 1) / Geo<L>::RB, ntiles);
  hipLaunchKernelGGL(k_norm1_acc<L>, grid, dim3(WGA), 0, s, G, maps, nv, s1, s2, twM);
}
void launch_norm1_acc(int L, const float2* G, const MapRef* maps, int ntiles, int nv, double* s1,
                      double* s2, const float2* twM, hipStream_t s) {
  DISPATCH(L, norm1_acc_t, G, maps, ntiles, nv, s1, s2, twM, s);
}
template <int L>
static void norm1_zrun_t(const float2* G, const MapRef* maps, int nmaps, const void* norm,
                         unsigned long long* keys, const float2* twM, hipStream_t s) {
  // row blocks padded to a multiple of 8: workgroups are dealt to the 8 XCDs
  // round-robin, so XCD x then runs the row blocks x, x + 8, ... of every map
  // and its L2 holds only those rows of the tile's per-pixel mean and 1/sd,
  // which the tile's other views read again
  dim3 grid(((L + Geo<L>::RB - 1) / Geo<L>::RB + 7) / 8 * 8, nmaps);
  hipLaunchKernelGGL(k_norm1_zrun<L>, grid, dim3(WG), 0, s, G, maps,
                     static_cast<const float4*>(norm), keys, twM);
}
void launch_norm1_zrun(int L, const float2* G, const MapRef* maps, int nmaps, const void* norm,
                       void* keys, const float2* twM, hipStream_t s) {
  DISPATCH(L, norm1_zrun_t, G, maps, nmaps, norm, static_cast<unsigned long long*>(keys), twM, s);
}
template <int L>
static void norm1_runfin_t(const unsigned long long* keys, int nmaps, long long mbase,
                           int ntiles, const MapRef* maps, const TileDesc* tdesc, float thres,
                           float dmin, unsigned long long* count, N1Rec* out, long long cap,
                           hipStream_t s) {
  const int nblk = (int)(((long)L * L + 1023) / 1024);
  hipLaunchKernelGGL(k_norm1_runfin<L>, dim3((nblk + 255) / 256, nmaps), dim3(256), 0, s, keys,
                     mbase, ntiles, maps, tdesc, thres, dmin, count, out, cap);
}
void launch_norm1_runfin(int L, const void* keys, int nmaps, long long mbase, int ntiles,
                         const MapRef* maps, const TileDesc* tdesc, float thres, float dmin,
                         unsigned long long* count, void* out, long long cap, hipStream_t s) {
  DISPATCH(L, norm1_runfin_t, static_cast<const unsigned long long*>(keys), nmaps, mbase, ntiles,
           maps, tdesc, thres, dmin, count, static_cast<N1Rec*>(out), cap, s);
}
long norm1_key_bytes(int L) { return (((long)L * L + 1023) / 1024) * 2 * 8; }
void launch_norm1_stats(const double* s1, const double* s2, long n, double total, void* norm,
                        hipStream_t s) {
  const long npairs = n / 2;
  hipLaunchKernelGGL(k_norm1_stats, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0, s, s1,
                     s2, npairs, total, static_cast<float4*>(norm));
}

__global__ void k_select(const MapStat* __restrict__ st, long long nmaps, int T, int N, int L,
                         const TileDesc* __restrict__ tdesc, float thres, float dmin,
                         unsigned long long* __restrict__ count, SelRec* __restrict__ out,
                         long long cap) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nmaps) return;
  const MapStat s = st[m];
  if (!(s.score >= thres)) return;
  const int t = (int)((m % ((long long)T * N)) / N);
  const TileDesc td = tdesc[t];
  const float cx = (float)(td.ox + s.pos % L);
  const float cy = (float)(td.oy + s.pos / L);
  if (cx >= dmin && cx < td.nx - dmin && cy >= dmin && cy < td.ny - dmin) {
    const unsigned long long i = atomicAdd(count, 1ull);
    if (i < (unsigned long long)cap) {
      SelRec r;
      r.m = m; r.cx = cx; r.cy = cy; r.score = s.score; r.pad = 0;
      out[i] = r;
    }
  }
}

void launch_select(const MapStat* st, long long nmaps, int T, int N, int L, const TileDesc* tdesc,
                   float thres, float dmin, unsigned long long* count, SelRec* out,
                   long long cap, hipStream_t s) {
  if (nmaps <= 0) return;
  hipLaunchKernelGGL(k_select, dim3((unsigned)((nmaps + 255) / 256)), dim3(256), 0, s, st, nmaps, T,
                     N, L, tdesc, thres, dmin, count, out, cap);
}
int norm1_record_bytes() { return (int)sizeof(N1Rec); }

}  // namespace gs
