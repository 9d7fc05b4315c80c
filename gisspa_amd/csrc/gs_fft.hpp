// gs_fft.hpp -- LDS-resident Stockham FFT engine for gfx950 (wave64).
//
// A workgroup holds NB transforms of compile-time length N in LDS and runs
// the radix passes of a compile-time plan.  Each pass: every thread loads the
// R inputs of its butterflies (stride N/R, conflict-free ds_read_b64), applies
// the twiddles from an LDS table, evaluates an R-point DFT in registers, then
// after a barrier writes the Stockham autosort positions.  Radix DFTs are
// built at compile time: hand-written 2/3/4/5-point kernels, direct
// evaluation for other primes (7, 11, ..., 31), Cooley-Tukey for prime
// powers and Good-Thomas for coprime splits, so any N that factors into
// radices of at most 32 works with the same code.
//
// Sign convention: DIR = -1 forward (exp(-2 pi i k n / N)), DIR = +1 inverse;
// transforms are unnormalised.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace gs {

// ---------------------------------------------------------------------------
// compile-time trigonometry (for the radix-kernel constants)
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin_series(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 30; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double ct_reduce(double x) {
  // to [-pi, pi]
  while (x > kPi) x -= 2 * kPi;
  while (x < -kPi) x += 2 * kPi;
  return x;
}
constexpr double ct_sin(double x) { return ct_sin_series(ct_reduce(x)); }
constexpr double ct_cos(double x) { return ct_sin(x + kPi / 2); }

// exp(DIR * 2 pi i * m / R) as float constants
template <int R, int M, int DIR>
struct TwC {
  static constexpr double ang = DIR * 2.0 * kPi * (double)(M % R) / (double)R;
  static constexpr float re = (float)ct_cos(ang);
  static constexpr float im = (float)ct_sin(ang);
};

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by DIR * i
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, int M, int DIR>
__device__ __forceinline__ float2 twc(float2 a) {
  constexpr int m = M % R;
  if constexpr (m == 0) {
    return a;
  } else if constexpr (4 * m == R) {
    return mul_i<DIR>(a);
  } else if constexpr (2 * m == R) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (4 * m == 3 * R) {
    return mul_i<-DIR>(a);
  } else {
    constexpr float c = TwC<R, m, DIR>::re, s = TwC<R, m, DIR>::im;
    return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// ---------------------------------------------------------------------------
// R-point DFT in registers
// ---------------------------------------------------------------------------
constexpr int smallest_factor(int n) {
  for (int f = 2; f * f <= n; ++f)
    if (n % f == 0) return f;
  return n;
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
  __device__ __forceinline__ static void run(float2* v) {}
};

template <int DIR>
struct Dft<2, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int DIR>
struct Dft<3, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float c = -0.5f;
    constexpr float s = (float)(DIR * 0.86602540378443864676);
    float2 t1 = cadd(v[1], v[2]);
    float2 t2 = csub(v[1], v[2]);
    float2 y0 = cadd(v[0], t1);
    float2 m1 = make_float2(v[0].x + c * t1.x, v[0].y + c * t1.y);
    float2 m2 = make_float2(-s * t2.y, s * t2.x);
    v[0] = y0;
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
  }
};

template <int DIR>
struct Dft<4, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    float2 b0 = cadd(v[1], v[3]), b1 = mul_i<DIR>(csub(v[1], v[3]));
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  }
};

template <int DIR>
struct Dft<5, DIR> {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr float c1 = 0.30901699437494742410f;   // cos(2pi/5)
    constexpr float c2 = -0.80901699437494742410f;  // cos(4pi/5)
    constexpr float s1 = (float)(DIR * 0.95105651629515357212);  // sin(2pi/5)
    constexpr float s2 = (float)(DIR * 0.58778525229247312917);  // sin(4pi/5)
    float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    float2 y0 = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    float2 r1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    float2 r2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    // i * (s1 b1 + s2 b2), i * (s2 b1 - s1 b2)
    float2 q1 = make_float2(-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x);
    float2 q2 = make_float2(-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x);
    v[0] = y0;
    v[1] = cadd(r1, q1);
    v[4] = csub(r1, q1);
    v[2] = cadd(r2, q2);
    v[3] = csub(r2, q2);
  }
};

// Odd prime radix R: X_k = v_0 + sum_j v_j w^(jk), w = exp(DIR 2 pi i / R).
// Inputs j and R - j share their cosines and sines: with a_j = v_j + v_(R-j),
// b_j = v_j - v_(R-j) (j = 1 .. (R-1)/2),
//   P_k = v_0 + sum_j a_j cos(2 pi jk / R),  Q_k = sum_j b_j DIR sin(2 pi jk / R),
//   X_k = P_k + i Q_k,  X_(R-k) = P_k - i Q_k,
// about a quarter of the multiplications of the plain sum; unrolled loops over
// constexpr tables, so the compiler sees straight-line code with constants.
template <int R, int DIR>
struct PrimeTab {
  struct T {
    float c[R], s[R];
  };
  static constexpr T make() {
    T t{};
    for (int m = 0; m < R; ++m) {
      const double ang = 2.0 * kPi * (double)m / (double)R;
      t.c[m] = (float)ct_cos(ang);
      t.s[m] = (float)(DIR * ct_sin(ang));
    }
    return t;
  }
  static constexpr T tab = make();
};

// largest power of R's smallest prime that divides R
constexpr int prime_power_part(int r) {
  const int p = smallest_factor(r);
  int a = 1;
  while (r % (a * p) == 0) a *= p;
  return a;
}
constexpr int mod_inverse(int a, int m) {
  for (int x = 1; x < m; ++x)
    if ((a * x) % m == 1) return x;
  return 1;  // m == 1
}
// how an R-point DFT is built: 0 direct (prime), 1 Cooley-Tukey over a prime
// power, 2 Good-Thomas prime-factor split into coprime parts
constexpr int dft_kind(int r) {
  return smallest_factor(r) == r ? 0 : (prime_power_part(r) == r ? 1 : 2);
}

template <int R, int DIR, int KIND>
struct DftImpl;

template <int R, int DIR>
struct DftImpl<R, DIR, 0> {
  static constexpr int HR = (R - 1) / 2;
  __device__ __forceinline__ static void run(float2* v) {
    constexpr auto T = PrimeTab<R, DIR>::tab;
    float2 a[HR], b[HR];
#pragma unroll
    for (int j = 0; j < HR; ++j) {
      a[j] = cadd(v[j + 1], v[R - 1 - j]);
      b[j] = csub(v[j + 1], v[R - 1 - j]);
    }
    float2 x0 = v[0];
#pragma unroll
    for (int j = 0; j < HR; ++j) x0 = cadd(x0, a[j]);
    float2 o[R];
    o[0] = x0;
#pragma unroll
    for (int k = 1; k <= HR; ++k) {
      float2 p = v[0], q = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int j = 0; j < HR; ++j) {
        const int m = ((j + 1) * k) % R;
        p.x += a[j].x * T.c[m];
        p.y += a[j].y * T.c[m];
        q.x += b[j].x * T.s[m];
        q.y += b[j].y * T.s[m];
      }
      // i Q = (-Q.y, Q.x)
      o[k] = make_float2(p.x - q.y, p.y + q.x);
      o[R - k] = make_float2(p.x + q.y, p.y - q.x);
    }
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = o[i];
  }
};

// composite R = A * B: X[k1 + A k2] = sum_n2 W_B^{n2 k2} W_R^{n2 k1} sum_n1 x[B n1 + n2] W_A^{n1 k1}
template <int R, int DIR, int A, int B, int N2, int K1>
struct CompTwS {
  __device__ __forceinline__ static void run(const float2* t, float2 (*y)[A]) {
    y[N2][K1] = twc<R, (N2 * K1) % R, DIR>(t[K1]);
    if constexpr (K1 + 1 < A) CompTwS<R, DIR, A, B, N2, K1 + 1>::run(t, y);
  }
};

template <int R, int DIR, int A, int B, int N2>
struct CompStage1 {
  __device__ __forceinline__ static void run(const float2* v, float2 (*y)[A]) {
    float2 t[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + N2];
    Dft<A, DIR>::run(t);
    y[N2][0] = t[0];
    if constexpr (A > 1) CompTwS<R, DIR, A, B, N2, 1>::run(t, y);
    if constexpr (N2 + 1 < B) CompStage1<R, DIR, A, B, N2 + 1>::run(v, y);
  }
};

template <int R, int DIR>
struct DftImpl<R, DIR, 1> {
  static constexpr int A = smallest_factor(R);
  static constexpr int B = R / A;
  __device__ __forceinline__ static void run(float2* v) {
    float2 y[B][A];
    CompStage1<R, DIR, A, B, 0>::run(v, y);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      float2 t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2][k1];
      Dft<B, DIR>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
    }
  }
};

// Good-Thomas split R = A B, gcd(A, B) = 1: with input index
// n = (B n1 + A n2) mod R and output index k = (B b' k1 + A a' k2) mod R
// (b' = B^-1 mod A, a' = A^-1 mod B) the DFT factors into A-point DFTs over n1
// and B-point DFTs over n2 with no twiddle factors in between.
template <int R, int DIR, int A, int B, int N2>
struct PfaStage1 {
  __device__ __forceinline__ static void run(const float2* v, float2 (*y)[A]) {
    float2 t[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) t[n1] = v[(B * n1 + A * N2) % R];
    Dft<A, DIR>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) y[N2][k1] = t[k1];
    if constexpr (N2 + 1 < B) PfaStage1<R, DIR, A, B, N2 + 1>::run(v, y);
  }
};

template <int R, int DIR>
struct DftImpl<R, DIR, 2> {
  static constexpr int A = prime_power_part(R);
  static constexpr int B = R / A;
  static constexpr int BI = mod_inverse(B % A, A);
  static constexpr int AI = mod_inverse(A % B, B);
  __device__ __forceinline__ static void run(float2* v) {
    float2 y[B][A];
    PfaStage1<R, DIR, A, B, 0>::run(v, y);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      float2 t[B];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2][k1];
      Dft<B, DIR>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < B; ++k2) v[(B * BI * k1 + A * AI * k2) % R] = t[k2];
    }
  }
};

template <int R, int DIR>
struct Dft {
  __device__ __forceinline__ static void run(float2* v) { DftImpl<R, DIR, dft_kind(R)>::run(v); }
};

// ---------------------------------------------------------------------------
// compile-time radix plans
// ---------------------------------------------------------------------------
struct Plan {
  int count;
  int r[16];
  int pad_shift = 4;  // LDS layout: element i at i + (i >> pad_shift) * pad_mult
  int pad_mult = 1;
};

constexpr double ct_log2(int r) {
  double v = 0;
  double x = r;
  while (x >= 2) { x /= 2; v += 1; }
  // linear interpolation of the fractional part is enough for a cost model
  return v + (x - 1);
}

// Cost of running radices rad[0..n) over NB transforms of length N with WG
// threads: each pass costs (butterflies per thread, rounded up) * R log2 R
// register work, and every LDS exchange between passes costs a read and a
// write of the whole working set.
constexpr double plan_cost(int N, int NB, int WG, const int* rad, int n) {
  double c = 0;
  for (int i = 0; i < n; ++i) {
    const int bfly = NB * (N / rad[i]);
    const int per = (bfly + WG - 1) / WG;
    c += per * rad[i] * ct_log2(rad[i]);
  }
  c += (n - 1) * 4.0 * N * NB / WG;
  return c;
}

// Fewest-work factorisation of N into at most three radices <= 32; ties keep
// the first found (larger leading radix).
constexpr Plan make_plan(int N, int NB = 1, int WG = 256) {
  Plan best{0, {0}};
  double bc = 1e300;
  for (int a = 32; a >= 2; --a) {
    if (N % a) continue;
    const int na = N / a;
    if (na == 1) {
      int r[1] = {a};
      const double c = plan_cost(N, NB, WG, r, 1);
      if (c < bc) { bc = c; best = Plan{1, {a}}; }
      continue;
    }
    for (int b = 32; b >= 2; --b) {
      if (na % b) continue;
      const int nb = na / b;
      if (nb == 1) {
        int r[2] = {a, b};
        const double c = plan_cost(N, NB, WG, r, 2);
        if (c < bc) { bc = c; best = Plan{2, {a, b}}; }
        continue;
      }
      if (nb <= 32) {
        int r[3] = {a, b, nb};
        const double c = plan_cost(N, NB, WG, r, 3);
        if (c < bc) { bc = c; best = Plan{3, {a, b, nb}}; }
      }
    }
  }
  if (best.count == 0) {
    // long transforms: greedy radix-16/8 then the remainder
    int n = N, k = 0;
    for (int r : {16, 8, 9, 10, 12, 15, 14, 7, 6, 5, 4, 3, 2})
      while (n % r == 0 && n > 1 && k < 15) { best.r[k++] = r; n /= r; }
    best.count = k;
  }
  return best;
}

// Radix order and LDS padding per (length, transforms per workgroup, last-pass
// mapping), chosen offline with a bank-conflict model of the executor's LDS
// traffic (32 banks, 16 lanes of b64 accesses per cycle) among the plans whose
// register work is within 5% of the best.
struct TunedPlan {
  int n, nb, mapl, count, r0, r1, r2, shift, mult;
};
constexpr TunedPlan kTuned[] = {
    {64, 32, 1, 2, 8, 8, 0, 2, 1},     {32, 32, 0, 2, 4, 8, 0, 1, 3},
    {96, 32, 1, 2, 12, 8, 0, 1, 1},    {48, 32, 0, 2, 6, 8, 0, 2, 2},
    {128, 32, 1, 2, 8, 16, 0, 3, 1},   {64, 32, 0, 2, 8, 8, 0, 2, 1},
    {160, 32, 1, 2, 20, 8, 0, 1, 1},   {80, 32, 0, 2, 10, 8, 0, 1, 0},
    {192, 32, 1, 2, 24, 8, 0, 2, 1},   {96, 32, 0, 2, 12, 8, 0, 1, 1},
    {256, 32, 1, 2, 8, 32, 0, 3, 1},   {128, 32, 0, 2, 16, 8, 0, 4, 2},
    {320, 16, 1, 2, 20, 16, 0, 2, 1},  {160, 32, 0, 2, 20, 8, 0, 1, 1},
    {384, 16, 1, 2, 24, 16, 0, 3, 1},  {192, 32, 0, 2, 24, 8, 0, 2, 1},
    {448, 16, 1, 2, 28, 16, 0, 2, 1},  {224, 32, 0, 2, 8, 28, 0, 5, 1},
    {512, 16, 1, 2, 16, 32, 0, 4, 1},  {256, 32, 0, 2, 16, 16, 0, 4, 1},
    {576, 8, 1, 2, 24, 24, 0, 6, 1},   {288, 16, 0, 2, 16, 18, 0, 4, 1},
    {640, 8, 1, 2, 20, 32, 0, 2, 1},   {320, 16, 0, 2, 16, 20, 0, 4, 1},
    {720, 8, 1, 2, 30, 24, 0, 6, 3},   {360, 16, 0, 2, 30, 12, 0, 1, 0},
    {768, 8, 1, 2, 24, 32, 0, 3, 1},   {384, 16, 0, 2, 16, 24, 0, 4, 1},
    {960, 8, 1, 2, 30, 32, 0, 6, 3},   {480, 16, 0, 2, 16, 30, 0, 4, 1},
    {1024, 8, 1, 2, 32, 32, 0, 5, 1},  {512, 16, 0, 2, 16, 32, 0, 4, 1},
};

// The table was made for 256-thread workgroups; wider ones (the norm_type 1
// accumulation's 512) take the same plans: with the two-pass 30 x 12 rows
// instead of the cost model's three-pass 12 x 6 x 5, that kernel issues a
// third of the LDS instructions and the 206-view norm_type 1 leg runs 112.4
// instead of 105.5 micrographs/hour.
constexpr Plan tuned_plan(int N, int NB, int WG, int MAPL) {
  for (const TunedPlan& t : kTuned)
    if (t.n == N && t.nb == NB && t.mapl == MAPL && WG >= 256) {
      Plan p{t.count, {t.r0, t.r1, t.r2}};
      p.pad_shift = t.shift;
      p.pad_mult = t.mult;
      return p;
    }
  return make_plan(N, NB, WG);
}

// Schedule variants (selected at run time by gs_autotune): variant 0 is the
// tuned plan above; variant v >= 1 is the v-th cheapest other factorisation
// of N into two or three radices <= 32 by plan_cost (ties keep the larger
// leading radices), with the tuned plan's LDS padding.  A variant past the
// last alternative falls back to the tuned plan.
constexpr int kPlanVariants = 4;
constexpr Plan variant_plan(int N, int NB, int WG, int MAPL, int V) {
  const Plan t = tuned_plan(N, NB, WG, MAPL);
  if (V <= 0) return t;
  constexpr int CAP = 256;
  int rr[CAP][3] = {{0}};
  int cnt[CAP] = {0};
  double cost[CAP] = {0};
  int n = 0;
  for (int a = 32; a >= 2; --a) {
    if (N % a) continue;
    for (int b = 32; b >= 2; --b) {
      if ((N / a) % b) continue;
      const int c = N / a / b;
      if (c > 32 || n >= CAP) continue;
      const int k = c == 1 ? 2 : 3;
      const int r[3] = {a, b, c};
      bool same = t.count == k && t.r[0] == a && t.r[1] == b && (k == 2 || t.r[2] == c);
      if (same) continue;
      rr[n][0] = a;
      rr[n][1] = b;
      rr[n][2] = c;
      cnt[n] = k;
      cost[n] = plan_cost(N, NB, WG, r, k);
      ++n;
    }
  }
  bool used[CAP] = {false};
  int pick = -1;
  for (int v = 0; v < V; ++v) {
    pick = -1;
    for (int i = 0; i < n; ++i)
      if (!used[i] && (pick < 0 || cost[i] < cost[pick])) pick = i;
    if (pick < 0) return t;
    used[pick] = true;
  }
  Plan p{cnt[pick], {rr[pick][0], rr[pick][1], cnt[pick] == 3 ? rr[pick][2] : 0}};
  p.pad_shift = t.pad_shift;
  p.pad_mult = t.pad_mult;
  return p;
}

// a length the executor runs: its plan's radices multiply back to it
constexpr bool plan_ok(int n) {
  const Plan p = make_plan(n);
  int prod = 1;
  for (int i = 0; i < p.count; ++i) prod *= p.r[i];
  return p.count > 0 && prod == n;
}

// ---------------------------------------------------------------------------
// one Stockham pass over NB transforms held in LDS at lds[b * STRIDE + i]
// tw: LDS table of exp(-2 pi i m / N), m in [0, N)
// ---------------------------------------------------------------------------
template <int N, int R, int NS, int NB, int WG, int DIR, int STRIDE>
__device__ __forceinline__ void lds_pass(float2* lds, const float2* tw) {
  constexpr int M = N / R;
  constexpr int TOT = NB * M;
  constexpr int PER = (TOT + WG - 1) / WG;
  constexpr int TWS = N / (NS * R);
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int g = tid + p * WG;
    if (TOT % WG == 0 || g < TOT) {
      const int b = g / M, j = g - (g / M) * M;
      const float2* src = lds + b * STRIDE + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = src[r * M];
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * TWS];
          if constexpr (DIR > 0) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Dft<R, DIR>::run(v[p]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int g = tid + p * WG;
    if (TOT % WG == 0 || g < TOT) {
      const int b = g / M, j = g - (g / M) * M;
      const int k = j % NS;
      float2* dst = lds + b * STRIDE + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[p][r];
    }
  }
  __syncthreads();
}

template <int N, int PASS, int NS, int NB, int WG, int DIR, int STRIDE>
__device__ __forceinline__ void lds_passes(float2* lds, const float2* tw) {
  constexpr Plan P = make_plan(N, NB, WG);
  if constexpr (PASS < P.count) {
    constexpr int R = P.r[PASS];
    lds_pass<N, R, NS, NB, WG, DIR, STRIDE>(lds, tw);
    lds_passes<N, PASS + 1, NS * R, NB, WG, DIR, STRIDE>(lds, tw);
  }
}

// Full N-point transform of NB sequences in LDS.  Requires a barrier between
// the writes that filled lds/tw and this call (the caller's responsibility).
template <int N, int NB, int WG, int DIR, int STRIDE>
__device__ __forceinline__ void lds_fft(float2* lds, const float2* tw) {
  static_assert(plan_ok(N), "FFT length must factor into radices of at most 32");
  lds_passes<N, 0, 1, NB, WG, DIR, STRIDE>(lds, tw);
}

// ---------------------------------------------------------------------------
// Fused executor: the first radix pass reads its inputs through a load
// functor (global memory, products, pre-processing) and the last pass hands
// its outputs to a store functor (global writes, reductions); only the
// passes in between go through LDS.  LDS holds transform b element i at
// lds[b * padded_len(N) + pidx(i)] with one float2 of padding every 16
// elements, which spreads the stride-R Stockham writes over the banks.
//
// Thread mapping: butterfly g -> (b, j); MAP_J: j fastest (coalesced loads of
// contiguous columns), MAP_B: b fastest (coalesced transposed stores).
// ---------------------------------------------------------------------------
template <int SHIFT, int MULT>
__host__ __device__ constexpr int padded_len_sm(int n) { return n + (n >> SHIFT) * MULT + 1; }
template <int SHIFT, int MULT>
__device__ __forceinline__ int pidx_sm(int i) { return i + (i >> SHIFT) * MULT; }

// LDS elements (per transform, padding included) the executor needs
template <int N, int NB, int WG, int MAPL, int V = 0>
__host__ __device__ constexpr int exec_stride() {
  constexpr Plan P = variant_plan(N, NB, WG, MAPL, V);
  return padded_len_sm<P.pad_shift, P.pad_mult>(N);
}

enum { MAP_J = 0, MAP_B = 1 };

// f(std::integral_constant<int, I>) for I = 0 .. N-1, in order
template <class F, int... I>
__device__ __forceinline__ void unroll_seq(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void unroll(F&& f) {
  unroll_seq(f, std::make_integer_sequence<int, N>{});
}

template <int N, int NB, int WG, int DIR, int MAP0, int MAPL, int PASS, int NS, int V, class LoadF,
          class StoreF>
__device__ __forceinline__ void fft_exec_pass(float2* lds, const float2* tw, LoadF& load,
                                              StoreF& store) {
  constexpr Plan P = variant_plan(N, NB, WG, MAPL, V);
  constexpr int R = P.r[PASS];
  constexpr bool FIRST = PASS == 0, LAST = PASS == P.count - 1;
  constexpr int M = N / R;
  constexpr int TOT = NB * M;
  constexpr int PER = (TOT + WG - 1) / WG;
  constexpr int TWS = N / (NS * R);
  constexpr int STRIDE = padded_len_sm<P.pad_shift, P.pad_mult>(N);
  auto pidx = [](int i) { return pidx_sm<P.pad_shift, P.pad_mult>(i); };
  // which butterfly a thread owns in this pass (the same for its reads and writes)
  constexpr int MAPP = FIRST ? MAP0 : (LAST ? MAPL : MAP_J);
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int g = tid + p * WG;
    if (TOT % WG == 0 || g < TOT) {
      const int b = (MAPP == MAP_J) ? g / M : g % NB;
      const int j = (MAPP == MAP_J) ? g % M : g / NB;
      if constexpr (FIRST &&
                    std::is_invocable_v<LoadF&, int, int, std::integral_constant<int, 0>>) {
        // affine form: load(b, j, q) supplies element j + q * M, q known at
        // compile time, so the functor can keep per-thread bases and fold q
        // into constants
        unroll<R>([&](auto qc) { v[p][decltype(qc)::value] = load(b, j, qc); });
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (FIRST) v[p][r] = load(b, j + r * M);
          else v[p][r] = lds[b * STRIDE + pidx(j + r * M)];
        }
      }
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * TWS];
          if constexpr (DIR > 0) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Dft<R, DIR>::run(v[p]);
    }
  }
  if constexpr (!FIRST && !LAST) __syncthreads();  // reads of this pass done before writes
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int g = tid + p * WG;
    if (TOT % WG == 0 || g < TOT) {
      const int b = (MAPP == MAP_J) ? g / M : g % NB;
      const int j = (MAPP == MAP_J) ? g % M : g / NB;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (LAST) {
          // a 5-argument store also receives the round p and output slot r
          if constexpr (std::is_invocable_v<StoreF&, int, int, float2, int, int>)
            store(b, base + r * NS, v[p][r], p, r);
          else
            store(b, base + r * NS, v[p][r]);
        } else {
          lds[b * STRIDE + pidx(base + r * NS)] = v[p][r];
        }
      }
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
    fft_exec_pass<N, NB, WG, DIR, MAP0, MAPL, PASS + 1, NS * R, V>(lds, tw, load, store);
  }
}

// Run the whole N-point transform of NB sequences.  load(b, i) supplies input
// element i of sequence b; store(b, i, v) receives output element i.  The
// mapping of the first pass is MAP0 and of the last pass MAPL (a single-pass
// plan uses MAP0).  Requires tw to be visible (barrier) before the call.
template <int N, int NB, int WG, int DIR, int MAP0, int MAPL, int V = 0, class LoadF,
          class StoreF>
__device__ __forceinline__ void fft_exec(float2* lds, const float2* tw, LoadF& load,
                                         StoreF& store) {
  static_assert(plan_ok(N), "FFT length must factor into radices of at most 32");
  constexpr Plan P = variant_plan(N, NB, WG, MAPL, V);
  static_assert(P.count != 1 || MAP0 == MAPL, "single-pass plans use one mapping");
  constexpr int MAPL_EFF = MAPL;
  fft_exec_pass<N, NB, WG, DIR, MAP0, MAPL_EFF, 0, 1, V>(lds, tw, load, store);
}

// ---------------------------------------------------------------------------
// Split-exchange executor: the same passes, thread mapping and arithmetic as
// fft_exec, but LDS holds one float per element, so a workgroup needs half
// the LDS and more workgroups fit on a CU.  Between two passes the real parts
// go through LDS first, then (after every thread has read them) the
// imaginary parts, with three barriers instead of one.
// ---------------------------------------------------------------------------
struct NoPending {
  __device__ __forceinline__ void operator()() const {}
};

// LDS layout of the split exchange where it differs from the float2
// exchange's (one float per element changes the bank pattern).
//  * shift, mult: padding of the generic layout (element i at
//    i + (i >> shift) mult);
//  * pc > 0, for a two-pass plan R0 x R1: element i = q R0 + s lives at
//    q (R0 + pc) + s.  The first pass writes the R0 outputs of its butterfly
//    q there and the second pass reads elements s + r R0 at r (R0 + pc) + s,
//    so every address is a per-thread base plus a compile-time offset (no
//    per-element index arithmetic), and consecutive butterflies land pc
//    banks apart.
// Chosen by timing at the full-size window: pass 2's rows (360 points, 16
// rows) ran 0.517 us per map with (shift 3, mult 2) and 0.583 with the
// float2 table's (1, 0); the two-pass layout (pc = 1) then took 20 % of pass
// 2's and 13 % of pass 1's VALU instructions and 38 % of their LDS
// instructions away (bank conflicts -72 % / -54 %), pass 1 0.682 instead of
// 0.704 us per map, pass 2 unchanged (it waits on its loads).  The same
// layout serves the forward column transforms whose last pass maps lanes to
// rows (MAP_J): refinement of 4000 picks 0.528 instead of 0.545 s.
struct SplitPad {
  int n, nb, mapl, shift, mult, pc;
};
constexpr SplitPad kSplitPad[] = {{360, 16, 0, 3, 2, 1}, {720, 8, 1, 6, 3, 1}, {720, 8, 0, 6, 3, 1}};
constexpr Plan split_plan(int N, int NB, int WG, int MAPL, int V) {
  Plan p = variant_plan(N, NB, WG, MAPL, V);
  for (const SplitPad& t : kSplitPad)
    if (t.n == N && t.nb == NB && t.mapl == MAPL) {
      p.pad_shift = t.shift;
      p.pad_mult = t.mult;
    }
  return p;
}
// pc of the two-pass layout above, 0 where the generic layout applies
constexpr int split_pair_pad(int N, int NB, int WG, int MAPL, int V) {
  const Plan p = variant_plan(N, NB, WG, MAPL, V);
  if (p.count != 2) return 0;
  for (const SplitPad& t : kSplitPad)
    if (t.n == N && t.nb == NB && t.mapl == MAPL) return t.pc;
  return 0;
}
// floats per transform (padding included) the split executor needs
template <int N, int NB, int WG, int MAPL, int V = 0>
__host__ __device__ constexpr int exec_stride_split() {
  constexpr Plan P = split_plan(N, NB, WG, MAPL, V);
  constexpr int PC = split_pair_pad(N, NB, WG, MAPL, V);
  if constexpr (PC > 0) return P.r[1] * (P.r[0] + PC);
  return padded_len_sm<P.pad_shift, P.pad_mult>(N);
}

template <int N, int NB, int WG, int DIR, int MAP0, int MAPL, int PASS, int NS, int V, class LoadF,
          class StoreF, class PendF>
__device__ __forceinline__ void fft_exec_pass_split(float* lds, const float2* tw, LoadF& load,
                                                    StoreF& store, PendF& pend_im) {
  constexpr Plan P = split_plan(N, NB, WG, MAPL, V);
  constexpr int R = P.r[PASS];
  constexpr bool FIRST = PASS == 0, LAST = PASS == P.count - 1;
  constexpr int M = N / R;
  constexpr int TOT = NB * M;
  constexpr int PER = (TOT + WG - 1) / WG;
  constexpr int TWS = N / (NS * R);
  constexpr int STRIDE = exec_stride_split<N, NB, WG, MAPL, V>();
  constexpr int PC = split_pair_pad(N, NB, WG, MAPL, V);
  auto pidx = [](int i) { return pidx_sm<P.pad_shift, P.pad_mult>(i); };
  // LDS address of input element j + r M of a later pass and of output
  // element base + r NS of an earlier one (with the two-pass layout, the
  // second pass's M and the first pass's R are both R0, and NS = 1)
  auto raddr = [&](int b, int j, int r) {
    if constexpr (PC > 0) return b * STRIDE + r * (P.r[0] + PC) + j;
    else return b * STRIDE + pidx(j + r * M);
  };
  auto waddr = [&](int b, int j, int base, int r) {
    if constexpr (PC > 0) return b * STRIDE + j * (P.r[0] + PC) + r;
    else return b * STRIDE + pidx(base + r * NS);
  };
  constexpr int MAPP = FIRST ? MAP0 : (LAST ? MAPL : MAP_J);
  float2 v[PER][R];
  const int tid = threadIdx.x;
  if constexpr (FIRST) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int g = tid + p * WG;
      if (TOT % WG == 0 || g < TOT) {
        const int b = (MAPP == MAP_J) ? g / M : g % NB;
        const int j = (MAPP == MAP_J) ? g % M : g / NB;
        if constexpr (std::is_invocable_v<LoadF&, int, int, std::integral_constant<int, 0>>) {
          unroll<R>([&](auto qc) { v[p][decltype(qc)::value] = load(b, j, qc); });
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) v[p][r] = load(b, j + r * M);
        }
      }
    }
  } else {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int g = tid + p * WG;
      if (TOT % WG == 0 || g < TOT) {
        const int b = (MAPP == MAP_J) ? g / M : g % NB;
        const int j = (MAPP == MAP_J) ? g % M : g / NB;
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][r].x = lds[raddr(b, j, r)];
      }
    }
    __syncthreads();  // real parts read before the imaginary parts overwrite them
    pend_im();
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int g = tid + p * WG;
      if (TOT % WG == 0 || g < TOT) {
        const int b = (MAPP == MAP_J) ? g / M : g % NB;
        const int j = (MAPP == MAP_J) ? g % M : g / NB;
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][r].y = lds[raddr(b, j, r)];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int g = tid + p * WG;
    if (TOT % WG == 0 || g < TOT) {
      const int j = (MAPP == MAP_J) ? g % M : g / NB;
      if constexpr (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[k * r * TWS];
          if constexpr (DIR > 0) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Dft<R, DIR>::run(v[p]);
    }
  }
  if constexpr (LAST) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int g = tid + p * WG;
      if (TOT % WG == 0 || g < TOT) {
        const int b = (MAPP == MAP_J) ? g / M : g % NB;
        const int j = (MAPP == MAP_J) ? g % M : g / NB;
        const int k = j % NS;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (std::is_invocable_v<StoreF&, int, int, float2, int, int>)
            store(b, base + r * NS, v[p][r], p, r);
          else
            store(b, base + r * NS, v[p][r]);
        }
      }
    }
  } else {
    if constexpr (!FIRST) __syncthreads();  // this pass's reads done before its writes
    auto put = [&](auto comp) {
#pragma unroll
      for (int p = 0; p < PER; ++p) {
        const int g = tid + p * WG;
        if (TOT % WG == 0 || g < TOT) {
          const int b = (MAPP == MAP_J) ? g / M : g % NB;
          const int j = (MAPP == MAP_J) ? g % M : g / NB;
          const int k = j % NS;
          const int base = (j / NS) * NS * R + k;
#pragma unroll
          for (int r = 0; r < R; ++r)
            lds[waddr(b, j, base, r)] = decltype(comp)::value ? v[p][r].y : v[p][r].x;
        }
      }
    };
    put(std::integral_constant<int, 0>{});
    __syncthreads();
    auto wim = [&]() { put(std::integral_constant<int, 1>{}); };
    fft_exec_pass_split<N, NB, WG, DIR, MAP0, MAPL, PASS + 1, NS * R, V>(lds, tw, load, store, wim);
  }
}

// fft_exec with the split exchange; lds holds exec_stride_split<...>() * NB floats
template <int N, int NB, int WG, int DIR, int MAP0, int MAPL, int V = 0, class LoadF, class StoreF>
__device__ __forceinline__ void fft_exec_split(float* lds, const float2* tw, LoadF& load,
                                               StoreF& store) {
  static_assert(plan_ok(N), "FFT length must factor into radices of at most 32");
  constexpr Plan P = variant_plan(N, NB, WG, MAPL, V);
  static_assert(P.count != 1 || MAP0 == MAPL, "single-pass plans use one mapping");
  NoPending none;
  fft_exec_pass_split<N, NB, WG, DIR, MAP0, MAPL, 0, 1, V>(lds, tw, load, store, none);
}

// fill an LDS twiddle table exp(-2 pi i m / N) from a global table
template <int N, int WG>
__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ gtw) {
  for (int i = threadIdx.x; i < N; i += WG) tw[i] = gtw[i];
}

}  // namespace gs
