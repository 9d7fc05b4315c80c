// gs_fft.hpp -- LDS-resident Stockham FFT engine for gfx950 (wave64).
//
// A workgroup holds NB transforms of compile-time length N in LDS and runs
// the radix passes of a compile-time plan.  Each pass: every thread loads the
// R inputs of its butterflies (stride N/R, conflict-free ds_read_b64), applies
// the twiddles from an LDS table, evaluates an R-point DFT in registers, then
// after a barrier writes the Stockham autosort positions.  Radix DFTs are
// built at compile time from 2/3/4/5/7 kernels by Cooley-Tukey composition,
// so any N = 2^a 3^b 5^c 7^d works with the same code.
//
// Sign convention: DIR = -1 forward (exp(-2 pi i k n / N)), DIR = +1 inverse;
// transforms are unnormalised.
#pragma once
#include <hip/hip_runtime.h>

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

// generic prime radix: direct O(R^2) evaluation with compile-time constants
template <int R, int DIR, int Q, int RR>
struct DirectTerm {
  __device__ __forceinline__ static float2 acc(const float2* v) {
    float2 t = twc<R, (Q * RR) % R, DIR>(v[RR]);
    if constexpr (RR + 1 < R) {
      float2 u = DirectTerm<R, DIR, Q, RR + 1>::acc(v);
      return cadd(t, u);
    } else {
      return t;
    }
  }
};
template <int R, int DIR, int Q>
struct DirectOut {
  __device__ __forceinline__ static void run(const float2* v, float2* o) {
    o[Q] = DirectTerm<R, DIR, Q, 0>::acc(v);
    if constexpr (Q + 1 < R) DirectOut<R, DIR, Q + 1>::run(v, o);
  }
};

template <int R, int DIR, bool PRIME>
struct DftImpl;

template <int R, int DIR>
struct DftImpl<R, DIR, true> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 o[R];
    DirectOut<R, DIR, 0>::run(v, o);
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
struct DftImpl<R, DIR, false> {
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

template <int R, int DIR>
struct Dft {
  __device__ __forceinline__ static void run(float2* v) {
    DftImpl<R, DIR, smallest_factor(R) == R>::run(v);
  }
};

// ---------------------------------------------------------------------------
// compile-time radix plans
// ---------------------------------------------------------------------------
struct Plan {
  int count;
  int r[16];
};

constexpr Plan make_plan(int n) {
  Plan p{0, {0}};
  int a = 0, b = 0, c = 0, d = 0;
  while (n % 2 == 0) { n /= 2; ++a; }
  while (n % 3 == 0) { n /= 3; ++b; }
  while (n % 5 == 0) { n /= 5; ++c; }
  while (n % 7 == 0) { n /= 7; ++d; }
  // n must now be 1 (checked by the static_assert in the user)
  int rad[16] = {0};
  int k = 0;
  // odd radices first, pairing a factor 2 with each 5 and 7 (radix 10/14),
  // and 3*3 -> 9
  while (b >= 2) { rad[k++] = 9; b -= 2; }
  while (c > 0) {
    if (a > 0 && a % 3 != 0) { rad[k++] = 10; --a; } else rad[k++] = 5;
    --c;
  }
  while (d > 0) {
    if (a > 0 && a % 3 != 0) { rad[k++] = 14; --a; } else rad[k++] = 7;
    --d;
  }
  if (b == 1) {
    if (a >= 2 && a % 3 != 0) { rad[k++] = 12; a -= 2; }
    else if (a >= 1 && a % 3 != 0) { rad[k++] = 6; a -= 1; }
    else rad[k++] = 3;
  }
  // powers of two: 8s, with a 16 or 4/2 for the remainder
  while (a >= 3) {
    if (a == 4) { rad[k++] = 16; a -= 4; break; }
    rad[k++] = 8; a -= 3;
  }
  if (a == 2) rad[k++] = 4;
  if (a == 1) rad[k++] = 2;
  // order: largest power-of-two radix first (fewest conflicts on the first
  // write pass), then the rest
  for (int i = 0; i < k; ++i) p.r[i] = rad[k - 1 - i];
  p.count = k;
  return p;
}

constexpr bool plan_ok(int n) {
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  while (n % 7 == 0) n /= 7;
  return n == 1;
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
  constexpr Plan P = make_plan(N);
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
  static_assert(plan_ok(N), "FFT length must factor into 2, 3, 5, 7");
  lds_passes<N, 0, 1, NB, WG, DIR, STRIDE>(lds, tw);
}

// fill an LDS twiddle table exp(-2 pi i m / N) from a global table
template <int N, int WG>
__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ gtw) {
  for (int i = threadIdx.x; i < N; i += WG) tw[i] = gtw[i];
}

}  // namespace gs
