/*
 * gisspa_oracle.cpp -- FP64 CPU restatement of the GisSPA picking pipeline.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 *
 * Stage map (reference = Autumn1998/GisSPA):
 *   tile offsets ............ main.cu:228-250
 *   template prep ........... main.cu:182-223 (pad), 311-385 (edge normalise),
 *                             388-474 (whiten, weight, amplitude normalise);
 *                             GPU_func.cu:3-27, 85-94, 98-152, 249-290, 318-362, 409-422
 *   CTF ..................... GPU_func.cu:292-315
 *   micrograph prep ......... main.cu:523-586 (init_d_image), 588-655 (split +
 *                             phase-flip tile normalisation), 758-769 (tile FFT);
 *                             GPU_func.cu:155-227, 564-581, 785-796
 *   per-psi template ........ main.cu:669-716; GPU_func.cu:229-247, 363-407,
 *                             424-443, 496-560
 *   correlation/score ....... main.cu:824-894; GPU_func.cu:584-615, 667-708
 *   norm_type 1 ............. main.cu:718-756, 771-821, 1170-1229;
 *                             GPU_func.cu:618-662, 711-745
 *
 * Deliberate choices where the reference is undefined or nondeterministic:
 *   - em (ring mean) is 0: the reference reads d_templates before its H2D copy
 *     (main.cu:346 vs 357); equal to the ring mean whenever box <= l-4.
 *   - radial bins start from zero (main.cu:277-278 never clears them).
 *   - score = 0 when sd <= 0 (main.cu:875 leaves it uninitialised).
 *   - argmax ties: lowest linear index y*l+x wins (reference: reduction-tree
 *     order, GPU_func.cu:691, and first block, main.cu:859).
 *   - all sums are sequential FP64 instead of float atomics / float partials.
 *   - l is any even size; the reference's grids silently need l % 32 == 0.
 *
 * Pins: every function below is checked on the CPU against code it shares
 * nothing with (numpy's FFT, the textbook CTF, closed forms, brute force):
 * tests/test_oracle.py, tests/test_oracle_pins.py, tests/test_dedup_oracle.py,
 * tests/test_project_oracle.py.
 */
#include "gisspa_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef std::complex<double> cd;

/* The reference's own constant (GPU_func.cuh:29), used wherever its kernels
 * use PI.  FFT twiddles use the exact value. */
static const double REF_PI = 3.141592654;
static const int RA_SIZE = 5000;  /* GPU_func.cuh:31 */
static const double REF_EPS = 1e-8; /* GPU_func.cuh:28 */

/* ------------------------------------------------------------------------ */
/* FFT: Stockham autosort, mixed radix, direct DFT for prime factors.        */
/* sign = -1 forward, +1 inverse; unnormalised (cuFFT convention).          */
/* ------------------------------------------------------------------------ */
namespace {

struct Plan {
  int n;
  std::vector<int> radices;
  std::vector<cd> tw_fwd;                 /* exp(-2 pi i m / n) */
  std::map<int, std::vector<cd>> small;   /* per radix R: exp(-2 pi i m / R) */
};

std::mutex g_plan_mu;
std::map<int, std::shared_ptr<Plan>> g_plans;

std::shared_ptr<Plan> get_plan(int n) {
  std::lock_guard<std::mutex> lk(g_plan_mu);
  auto it = g_plans.find(n);
  if (it != g_plans.end()) return it->second;
  auto p = std::make_shared<Plan>();
  p->n = n;
  int m = n;
  while (m % 4 == 0) { p->radices.push_back(4); m /= 4; }
  while (m % 2 == 0) { p->radices.push_back(2); m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->radices.push_back(f); m /= f; }
    if (f * f > m && m > 1) { p->radices.push_back(m); m = 1; }
  }
  p->tw_fwd.resize(n);
  for (int k = 0; k < n; ++k) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    p->tw_fwd[k] = cd(std::cos(a), std::sin(a));
  }
  for (int R : p->radices) {
    std::vector<cd>& t = p->small[R];
    t.resize(R);
    for (int m = 0; m < R; ++m) {
      double a = -2.0 * M_PI * (double)m / (double)R;
      t[m] = cd(std::cos(a), std::sin(a));
    }
  }
  g_plans[n] = p;
  return p;
}

void dft_small(cd* v, int R, int sign, cd* tmp, const cd* tab) {
  if (R == 2) {
    cd a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
    return;
  }
  if (R == 4) {
    cd a0 = v[0] + v[2], a1 = v[0] - v[2];
    cd b0 = v[1] + v[3], b1 = v[1] - v[3];
    cd jb1 = (sign < 0) ? cd(b1.imag(), -b1.real()) : cd(-b1.imag(), b1.real());
    v[0] = a0 + b0; v[2] = a0 - b0; v[1] = a1 + jb1; v[3] = a1 - jb1;
    return;
  }
  for (int q = 0; q < R; ++q) {
    cd s = 0;
    for (int r = 0; r < R; ++r) {
      const cd w = tab[((long long)r * q) % R];
      s += v[r] * (sign < 0 ? w : std::conj(w));
    }
    tmp[q] = s;
  }
  for (int q = 0; q < R; ++q) v[q] = tmp[q];
}

void fft1d_strided(cd* data, int n, long stride, int sign, const Plan& pl,
                   std::vector<cd>& a, std::vector<cd>& b) {
  a.resize(n); b.resize(n);
  for (int i = 0; i < n; ++i) a[i] = data[(long)i * stride];
  cd* src = a.data();
  cd* dst = b.data();
  int Ns = 1;
  std::vector<cd> v, tmp;
  for (int R : pl.radices) {
    v.resize(R); tmp.resize(R);
    const cd* tab = pl.small.at(R).data();
    int m = n / R;
    for (int j = 0; j < m; ++j) {
      int k = j % Ns;
      int twstep = n / (Ns * R);
      for (int r = 0; r < R; ++r) {
        cd x = src[j + r * m];
        long idx = ((long)k * r * twstep) % n;
        cd w = pl.tw_fwd[idx];
        if (sign > 0) w = std::conj(w);
        v[r] = x * w;
      }
      dft_small(v.data(), R, sign, tmp.data(), tab);
      int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
    std::swap(src, dst);
    Ns *= R;
  }
  for (int i = 0; i < n; ++i) data[(long)i * stride] = src[i];
}

}  // namespace

extern "C" int or_fft1d(double* data, int n, int sign) {
  if (n <= 0) return -1;
  auto pl = get_plan(n);
  std::vector<cd> a, b;
  fft1d_strided(reinterpret_cast<cd*>(data), n, 1, sign, *pl, a, b);
  return 0;
}

/* Rows, then columns; each 1D transform is independent, so large images
 * (whole micrographs) split them over threads with identical results. */
static void fft2d(cd* d, int nx, int ny, int sign) {
  auto px = get_plan(nx);
  auto py = get_plan(ny);
  const bool par = (long)nx * ny >= (1L << 20);
#pragma omp parallel if (par)
  {
    std::vector<cd> a, b;
#pragma omp for schedule(static)
    for (int y = 0; y < ny; ++y) fft1d_strided(d + (long)y * nx, nx, 1, sign, *px, a, b);
#pragma omp for schedule(static)
    for (int x = 0; x < nx; ++x) fft1d_strided(d + x, ny, nx, sign, *py, a, b);
  }
}

extern "C" int or_fft2d(double* data, int nx, int ny, int sign) {
  if (nx <= 0 || ny <= 0) return -1;
  fft2d(reinterpret_cast<cd*>(data), nx, ny, sign);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Small helpers                                                            */
/* ------------------------------------------------------------------------ */

/* Radial bin used by SQRSum_by_circle / whiten_Tmp / compute_area_sum_ofSQR
 * (GPU_func.cu:119, 138, 329): floor(hypot(min(y,ny-y), min(x,nx-x)) + .5) - 1 */
static inline int radial_bin(int x, int y, int nx, int ny) {
  double dx = std::min(x, nx - x), dy = std::min(y, ny - y);
  return (int)std::floor(std::hypot(dy, dx) + 0.5) - 1;
}

static double lambda_of(double energy_kv) {  /* main.cu:154 */
  return 12.2639 / std::sqrt(energy_kv * 1000.0 + 0.97845 * energy_kv * energy_kv);
}

/* CTF_AST, GPU_func.cu:292-315.  Defocus values follow main.cu:150-153:
 * defocus is negated at parse (DataReader.cpp:378), dfu/dfv = defocus +- dfdiff. */
extern "C" double or_ctf_ast(int x1, int y1, int nx, int ny, const or_params* p,
                             const or_ctf* c, int mode) {
  double defocus = -c->defocus_um;
  double dfu = defocus + c->dfdiff, dfv = defocus - c->dfdiff;
  double lambda = lambda_of(p->energy);
  y1 = x1 > nx / 2 ? ny - y1 : y1;
  x1 = std::min(x1, nx - x1);
  double dx = std::min(x1, nx - x1);
  double dy = y1 - ny / 2;
  double ss = (dx * dx / ((double)nx * nx) + dy * dy / ((double)ny * ny)) / (p->apix * p->apix);
  double ag = std::atan2((double)(y1 - ny / 2), (double)x1);
  double df_ast = 0.5 * (dfu + dfv + 2 * c->dfdiff * std::cos(2 * (c->dfang * REF_PI / 180 - ag)));
  double gamma = -2 * REF_PI * (p->cs * 2.5e6 * lambda * lambda * lambda * ss * ss +
                                df_ast * 5000.0 * lambda * ss);
  double A = p->ampconst;
  double v = std::sqrt(1.0 - A * A) * std::sin(gamma) + A * std::cos(gamma);
  if (mode == 0) return v > 0 ? 1.0 : -1.0;
  if (mode == 2) return std::fabs(v);
  return v;
}

/* Tile offsets, main.cu:228-250. */
extern "C" int or_tile_offsets(int n, int l, int overlap, int* offs, int cap) {
  int interval = l - overlap;
  if (interval <= 0) return -1;
  int off = 0, cnt = 0;
  if (cnt < cap) offs[cnt] = 0;
  cnt++;
  while (off + l < n) {
    off += interval;
    if (off + l >= n) off = n - l;
    if (cnt < cap) offs[cnt] = off;
    cnt++;
  }
  return cnt;
}

extern "C" int or_num_psi(double phi_step) {
  int n = 0;
  for (float e = 0.0f; e < 360.0f; e += (float)phi_step) n++;
  return n;
}

extern "C" float or_psi_value(double phi_step, int index) {
  /* main.cu:1230 -- the loop variable is a float accumulator */
  float e = 0.0f;
  for (int i = 0; i < index; ++i) e += (float)phi_step;
  return e;
}

/* ------------------------------------------------------------------------ */
/* Template preparation                                                     */
/* ------------------------------------------------------------------------ */

/* Low-pass + |CTF| weighting factor of apply_weighting_function,
 * GPU_func.cu:249-290, for spectrum pixel (x, y) of an l x l template. */
static double template_weight(int x, int y, int l, const or_params* p, const or_ctf* c) {
  double r = std::hypot((double)std::min(y, l - y), (double)std::min(x, l - x));
  int r_round = (int)std::floor(r + 0.5) - 1;
  double f = 1.0;
  double hi = l * p->apix / p->highres, lo = l * p->apix / p->lowres;
  if (r_round < hi && r_round >= 0) {
  } else if (r_round >= hi && r_round < hi + 8) {
    f *= 0.5 * std::cos(REF_PI * (r_round - hi) / (2 * 8)) + 0.5;
  } else if (r_round >= (lo - 8) && r_round < lo && r_round >= 0) {
    f *= 0.5 * std::cos(REF_PI * (lo - r_round) / (2 * 8)) + 0.5;
  } else {
    f = 0;
  }
  double ds = 1.0 / (p->apix * p->l_config);
  double ss = r * ds;
  if (r_round < l / 2 && r_round >= 0) {
    double v = or_ctf_ast(x, (y + l / 2) % l, l, l, p, c, 2);
    double signal = std::exp(p->bfactor * ss * ss + p->bfactor2 * ss + p->bfactor3) / (p->kk + 1);
    double Ncurve = std::exp(p->a * ss * ss + p->b * ss + p->b2) / signal;
    f *= v * std::sqrt(1 / (Ncurve + p->kk * v * v));
  }
  return f;
}

/* Template J up to whitening (main.cu:182-430): pad into the centre, edge
 * normalise, FFT / l^2, divide by the radial amplitude (whiten_Tmp). */
static void template_whiten(const or_params* p, const float* tm, int bx, int by, cd* t) {
  const int l = p->l;
  const long L2 = (long)l * l;
  /* pad into the centre, main.cu:207-216 */
  for (long i = 0; i < L2; ++i) t[i] = 0;
  int sx = (l - bx) / 2, sy = (l - by) / 2;
  for (int j = 0; j < by; ++j)
    for (int i = 0; i < bx; ++i) t[(long)(sy + j) * l + sx + i] = cd(tm[(long)j * bx + i], 0);
  /* edge normalise, main.cu:366-383, GPU_func.cu:85-94 (em = 0) */
  double s1 = 0, s2 = 0;
  for (long i = 0; i < L2; ++i) { double v = t[i].real(); s1 += v; s2 += v * v; }
  double mean = s1 / (double)L2;
  double sigma = std::sqrt(s2 / (double)L2 - mean * mean);
  if (sigma < 0 || !std::isfinite(sigma)) sigma = 0;
  if (!(sigma < REF_EPS && sigma > -REF_EPS))
    for (long i = 0; i < L2; ++i) t[i] = cd((t[i].real() - 0.0) / sigma, t[i].imag());
  /* FFT, /l^2, main.cu:410-412 */
  fft2d(t, l, l, -1);
  for (long i = 0; i < L2; ++i) t[i] /= (double)L2;
  /* radial power, GPU_func.cu:98-128 (x <= l/2 half plane) */
  std::vector<double> ra(RA_SIZE, 0.0), rb(RA_SIZE, 0.0);
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      int r = radial_bin(x, y, l, l);
      if (r < l / 2 && r >= 0) { ra[r] += std::norm(t[(long)y * l + x]); rb[r] += 1; }
    }
  /* whiten, GPU_func.cu:130-152 */
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      int r = radial_bin(x, y, l, l);
      if (r < l / 2 && r >= 0) t[(long)y * l + x] /= std::sqrt(ra[r] / rb[r]);
    }
}

static int template_prep_impl(const or_params* p, const or_ctf* c, const float* tmpl, int ntmpl,
                              int bx, int by, double* out) {
  const int l = p->l;
  if (bx > l || by > l) return -2;
  const long L2 = (long)l * l;
  cd* T = reinterpret_cast<cd*>(out);
  for (int J = 0; J < ntmpl; ++J) {
    cd* t = T + J * L2;
    template_whiten(p, tmpl + (long)J * bx * by, bx, by, t);
    /* low pass + weighting, GPU_func.cu:249-290 */
    for (int y = 0; y < l; ++y)
      for (int x = 0; x < l; ++x) t[(long)y * l + x] *= template_weight(x, y, l, p, c);
    /* amplitude normaliser, GPU_func.cu:318-362 + main.cu:444-452 */
    double S = 0, cnt = 0;
    for (int y = 0; y < l; ++y)
      for (int x = 0; x <= l / 2; ++x) {
        int r = radial_bin(x, y, l, l);
        if (r < l / 2 && r >= 0) { S += std::norm(t[(long)y * l + x]); cnt += 1; }
      }
    double m = std::sqrt(S / (cnt * cnt));
    if (m != 0)
      for (long i = 0; i < L2; ++i) t[i] /= m;
    /* inverse FFT, main.cu:463 */
    fft2d(t, l, l, +1);
  }
  return 0;
}

extern "C" int or_template_prep(const or_params* p, const float* tmpl, int ntmpl, int bx, int by,
                                double* out) {
  /* the template CTF uses the first micrograph's defocus; callers that need
   * it use or_template_prep_ctf */
  or_ctf c{0, 0, 0};
  return template_prep_impl(p, &c, tmpl, ntmpl, bx, by, out);
}

extern "C" int or_template_prep_ctf(const or_params* p, const or_ctf* c, const float* tmpl,
                                    int ntmpl, int bx, int by, double* out) {
  return template_prep_impl(p, c, tmpl, ntmpl, bx, by, out);
}

/* ------------------------------------------------------------------------ */
/* Per-psi rotation + mask + normalisation + FFT, main.cu:669-716           */
/* ------------------------------------------------------------------------ */
static void rotate_one(const cd* in, cd* out, int l, float e) {
  /* rotate_subIMG, GPU_func.cu:496-560.  The source coordinates decide the
   * zero region (x2 outside [0, l-1]) and the interpolation cell, so they are
   * evaluated exactly as the reference's float expression: cose/sine rounded
   * to float (GPU_func.cu:498-499), then fl(fl(fl(c*x) + fl(s*y)) + l/2)
   * with no fused multiply-add.  The interpolation itself is FP64. */
  const float cose = (float)std::cos((double)e * REF_PI / 180);
  const float sine = (float)std::sin((double)e * REF_PI / 180);
  for (int j = 0; j < l; ++j)
    for (int i = 0; i < l; ++i) {
      const float y = (float)(j - l / 2), x = (float)(i - l / 2);
      const float ax = cose * x, bx = sine * y, ay = -sine * x, by = cose * y;
      const float sx = ax + bx, sy = ay + by;
      const double x2 = (float)(sx + (float)(l / 2));
      const double y2 = (float)(sy + (float)(l / 2));
      cd res = 0;
      if (!(x2 < 0 || x2 > l - 1.0 || y2 < 0 || y2 > l - 1.0)) {
        int ii = (int)std::floor(x2), jj = (int)std::floor(y2);
        long k0 = ii + (long)jj * l, k1 = k0 + 1, k2 = k0 + l + 1, k3 = k0 + l;
        if (ii == l - 1) { k1--; k2--; }
        if (jj == l - 1) { k2 -= l; k3 -= l; }
        double t = x2 - ii, u = y2 - jj, tt = 1.0 - t, uu = 1.0 - u;
        res = in[k0] * (tt * uu) + in[k1] * (t * uu) + in[k2] * (t * u) + in[k3] * (tt * u);
      }
      out[(long)j * l + i] = res;
    }
}

static void rotate_normalize_one(const or_params* p, const cd* in, cd* out, float e) {
  const int l = p->l;
  const long L2 = (long)l * l;
  rotate_one(in, out, l, e);
  /* apply_mask, GPU_func.cu:229-247 (real part only) */
  double dm = 1.5 * p->diameter, ehw = 6;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      double r = std::hypot((double)(x - l / 2), (double)(y - l / 2));
      cd& v = out[(long)y * l + x];
      if (r > dm / 2 + 2 * ehw) v = cd(0, v.imag());
      else if (r >= dm / 2) v = cd(v.real() * (0.5 * std::cos(REF_PI * (r - dm / 2) / (2 * ehw)) + 0.5), v.imag());
    }
  /* compute_sum_sqr + host mean/var, main.cu:682-699 */
  double s1 = 0, s2 = 0;
  for (long i = 0; i < L2; ++i) { double v = out[i].real(); s1 += v; s2 += v * v; }
  double mean = s1 / L2;
  double var = s2 / L2 - mean * mean;
  /* substract_by_mean / divided_by_var, GPU_func.cu:424-443 */
  if (mean != 0) for (long i = 0; i < L2; ++i) out[i] = cd(out[i].real() - mean, out[i].imag());
  if (var != 0) {
    double sd = std::sqrt(var);
    for (long i = 0; i < L2; ++i) out[i] = cd(out[i].real() / sd, out[i].imag());
  }
  fft2d(out, l, l, -1);
  for (long i = 0; i < L2; ++i) out[i] /= (double)L2;
}

extern "C" int or_rotate_normalize(const or_params* p, const double* tmpl_real, int n, float euler3,
                                   double* spec_out) {
  const long L2 = (long)p->l * p->l;
  const cd* in = reinterpret_cast<const cd*>(tmpl_real);
  cd* out = reinterpret_cast<cd*>(spec_out);
  for (int J = 0; J < n; ++J) rotate_normalize_one(p, in + J * L2, out + J * L2, euler3);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Micrograph preparation                                                   */
/* ------------------------------------------------------------------------ */

/* whiten_filetr_weight_Img, GPU_func.cu:155-205, as a real factor applied to
 * the amplitude of spectrum pixel (x, y). */
static double micrograph_weight(int x, int y, int nx, int ny, const double* ra, const double* rb,
                                const or_params* p, const or_ctf* c) {
  double dx = std::min(x, nx - x), dy = std::min(y, ny - y);
  int r = (int)std::floor(std::hypot(dx, dy) + 0.5) - 1;
  double ss = std::sqrt((dx * dx / ((double)nx * nx) + dy * dy / ((double)ny * ny)) / (p->apix * p->apix));
  int l = std::max(nx, ny);
  double v = 0, signal = 0, Ncurve = 0, f = 1.0;
  if (r < l / 2 && r >= 0) {
    v = or_ctf_ast(x, (y + ny / 2) % ny, nx, ny, p, c, 2);
    signal = std::exp(p->bfactor * ss * ss + p->bfactor2 * ss + p->bfactor3);
    Ncurve = std::exp(p->a * ss * ss + p->b * ss + p->b2);
    f = f * std::sqrt((signal * v * v + Ncurve) / signal) / std::sqrt(ra[r] / rb[r]);
    if (r > (l * p->apix / 6)) f = f * std::exp(-100 * ss * ss);
  }
  double hi = l * p->apix / p->highres, lo = l * p->apix / p->lowres;
  if (r < hi && r >= lo) {
  } else if (r >= hi && r < hi + 8) {
    f = f * (0.5 * std::cos(REF_PI * (r - hi) / (2 * 8)) + 0.5);
  } else if (r >= (lo - 8) && r < lo && r >= 0) {
    f = f * (0.5 * std::cos(REF_PI * (lo - r) / (2 * 8)) + 0.5);
  } else {
    f = 0;
  }
  if (r < l / 2 && r >= 0) {
    signal /= (p->kk + 1);
    Ncurve /= signal;
    f = f * std::sqrt(1 / (Ncurve + p->kk * v * v));
  }
  return f;
}

/* Amplitude normaliser over the x <= l/2 half plane inside radius l/2
 * (compute_area_sum_ofSQR + main.cu:642-643). */
static double amp_normaliser(const cd* t, int l) {
  double S = 0, cnt = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      int r = radial_bin(x, y, l, l);
      if (r < l / 2 && r >= 0) { S += std::norm(t[(long)y * l + x]); cnt += 1; }
    }
  return std::sqrt(S / (cnt * cnt));
}

extern "C" int or_micrograph_prep(const or_params* p, const or_ctf* c, const float* img, int nx,
                                  int ny, double* tile_spec, int tcap) {
  const int l = p->l;
  const long L2 = (long)l * l;
  const long NP = (long)nx * ny;
  int ov = p->overlap;
  std::vector<int> ox(nx / 1 + 2), oy(ny / 1 + 2);
  int bxn = or_tile_offsets(nx, l, ov, ox.data(), (int)ox.size());
  int byn = or_tile_offsets(ny, l, ov, oy.data(), (int)oy.size());
  if (bxn < 0 || byn < 0) return -1;
  if (bxn * byn > tcap) return -2;
  /* init_d_image, main.cu:523-586 */
  std::vector<double> h(NP);
  for (long i = 0; i < NP; ++i) h[i] = img[i];
  double s1 = 0, s2 = 0;
  for (long i = 0; i < NP; ++i) { s1 += h[i] / nx / ny; s2 += h[i] * h[i] / nx / ny; }
  double avg = s1;
  double var = std::sqrt(s2 - avg * avg);
  /* the reference truncates the bounds to int (main.cu:539-540) */
  int up_bound = (int)(avg + 6 * var);
  int low_bound = (int)(avg - 6 * var);
  if (s2 > avg * avg)
    for (long i = 0; i < NP; ++i)
      if (h[i] > up_bound || h[i] < low_bound) h[i] = avg;
  if (p->invert == 1)
    for (long i = 0; i < NP; ++i) h[i] = -h[i];
  std::vector<double> work;
  if (p->phase_flip == 1) {
    std::vector<cd> F(NP);
    for (long i = 0; i < NP; ++i) F[i] = h[i];
    fft2d(F.data(), nx, ny, -1);
    for (long i = 0; i < NP; ++i) F[i] /= (double)NP;
    /* do_phase_flip, GPU_func.cu:785-796 */
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) F[(long)y * nx + x] *= or_ctf_ast(x, (y + ny / 2) % ny, nx, ny, p, c, 0);
    /* SQRSum_by_circle mode 1, GPU_func.cu:98-128 */
    std::vector<double> ra(RA_SIZE, 0.0), rb(RA_SIZE, 0.0);
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x <= nx / 2; ++x) {
        int r = radial_bin(x, y, nx, ny);
        if (r < std::max(nx, ny) / 2 && r >= 0) { ra[r] += std::norm(F[(long)y * nx + x]); rb[r] += 1; }
      }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        F[(long)y * nx + x] *= micrograph_weight(x, y, nx, ny, ra.data(), rb.data(), p, c);
    F[0] = 0;  /* set_0Hz_to_0_at_RI */
    fft2d(F.data(), nx, ny, +1);
    for (long i = 0; i < NP; ++i) h[i] = F[i].real();
  } else {
    /* FFT, /n, DC := 0, IFFT, real part == subtract the mean */
    std::vector<cd> F(NP);
    for (long i = 0; i < NP; ++i) F[i] = h[i];
    fft2d(F.data(), nx, ny, -1);
    for (long i = 0; i < NP; ++i) F[i] /= (double)NP;
    F[0] = 0;
    fft2d(F.data(), nx, ny, +1);
    for (long i = 0; i < NP; ++i) h[i] = F[i].real();
  }
  /* split_IMG, GPU_func.cu:564-581 ; tile FFT main.cu:758-769 */
  cd* out = reinterpret_cast<cd*>(tile_spec);
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int iy = 0; iy < byn; ++iy)
    for (int ix = 0; ix < bxn; ++ix) {
      cd* t = out + (long)(ix + iy * bxn) * L2;
      for (int y = 0; y < l; ++y)
        for (int x = 0; x < l; ++x) {
          int gx = ox[ix] + x, gy = oy[iy] + y;
          t[(long)y * l + x] = (gx < nx && gy < ny) ? cd(h[(long)gy * nx + gx], 0) : cd(0, 0);
        }
      fft2d(t, l, l, -1);
      for (long i = 0; i < L2; ++i) t[i] /= (double)L2;
      if (p->phase_flip == 1) {
        /* per-tile amplitude normalisation, main.cu:617-653 */
        double m = amp_normaliser(t, l);
        if (m != 0) for (long i = 0; i < L2; ++i) t[i] /= m;
      }
    }
  return bxn * byn;
}

/* ------------------------------------------------------------------------ */
/* Correlation: compute_corner_CCG + IFFT                                   */
/* ------------------------------------------------------------------------ */
static void cc_map(const cd* img, const cd* tsp, int l, cd* ccg) {
  /* GPU_func.cu:584-615: IMG * conj(T), then the sign pattern that moves
   * zero lag to (l/2, l/2). */
  int of = (l / 2) % 2;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      long i = (long)y * l + x;
      cd v = img[i] * std::conj(tsp[i]);
      int st = (of == y % 2) ? 1 : 0;
      if ((x - st) % 2 == 0) v = -v;
      ccg[i] = v;
    }
  fft2d(ccg, l, l, +1);
}

/* get_peak_and_SUM (GPU_func.cu:667-708) + host loop (main.cu:856-881) */
static or_mapstat reduce_norm0(const cd* ccg, int l, double d_m) {
  or_mapstat s{};
  double peak = 0, sum = 0, sum2 = 0;
  int pos = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) {
      double v = ccg[(long)y * l + x].real();
      sum += v;
      sum2 += v * v;
      double m = v;
      if (x < d_m / 4 || x > l - d_m / 4 || y < d_m / 4 || y > l - d_m / 4) m = 0;
      if (m > peak) { peak = m; pos = y * l + x; }
    }
  double rc = (double)l * l - 1;
  double ra = sum - peak, rb = sum2 - peak * peak;
  double var = rb / rc - (ra / rc) * (ra / rc);
  double sd = std::sqrt(var);
  s.peak = peak; s.sum = sum; s.sum2 = sum2; s.pos = pos;
  s.score = (sd > 0) ? peak / sd : 0.0;
  return s;
}

extern "C" int or_cc_map(const or_params* p, const double* tile_spec, const double* tspec,
                         double* out) {
  const int l = p->l;
  std::vector<cd> ccg((long)l * l);
  cc_map(reinterpret_cast<const cd*>(tile_spec), reinterpret_cast<const cd*>(tspec), l, ccg.data());
  for (long i = 0; i < (long)l * l; ++i) out[i] = ccg[i].real();
  return 0;
}

extern "C" int or_rotate_bilinear(const double* in, int l, float e, double* out) {
  std::vector<cd> a((long)l * l), b((long)l * l);
  for (long i = 0; i < (long)l * l; ++i) a[i] = in[i];
  rotate_one(a.data(), b.data(), l, e);
  for (long i = 0; i < (long)l * l; ++i) out[i] = b[i].real();
  return 0;
}

extern "C" int or_correlate_norm0(const or_params* p, const double* tile_spec, const double* tspec,
                                  or_mapstat* out) {
  const int l = p->l;
  std::vector<cd> ccg((long)l * l);
  cc_map(reinterpret_cast<const cd*>(tile_spec), reinterpret_cast<const cd*>(tspec), l, ccg.data());
  *out = reduce_norm0(ccg.data(), l, p->diameter);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Whole pipeline                                                           */
/* ------------------------------------------------------------------------ */
static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* Threads of the oracle's parallel loops from now on (0: leave as is). */
extern "C" int or_set_threads(int nthreads) {
  set_threads(nthreads);
  return 0;
}

extern "C" int or_pick_norm0(const or_params* p, const or_ctf* c, const float* tmpl, int ntmpl,
                             int bx, int by, const float* img, int nx, int ny, int psi_begin,
                             int psi_end, or_mapstat* out, int nthreads) {
  set_threads(nthreads);
  const int l = p->l;
  const long L2 = (long)l * l;
  int tx[4096], ty[4096];
  int bxn = or_tile_offsets(nx, l, p->overlap, tx, 4096);
  int byn = or_tile_offsets(ny, l, p->overlap, ty, 4096);
  if (bxn <= 0 || byn <= 0) return -1;
  const int T = bxn * byn;
  std::vector<cd> treal((long)ntmpl * L2);
  int rc = template_prep_impl(p, c, tmpl, ntmpl, bx, by, reinterpret_cast<double*>(treal.data()));
  if (rc) return rc;
  std::vector<cd> tiles((long)T * L2);
  rc = or_micrograph_prep(p, c, img, nx, ny, reinterpret_cast<double*>(tiles.data()), T);
  if (rc < 0) return rc;
  std::vector<cd> tsp((long)ntmpl * L2);
  for (int ip = psi_begin; ip < psi_end; ++ip) {
    float e = or_psi_value(p->phi_step, ip);
#pragma omp parallel for schedule(dynamic)
    for (int J = 0; J < ntmpl; ++J) rotate_normalize_one(p, &treal[J * L2], &tsp[J * L2], e);
#pragma omp parallel
    {
      std::vector<cd> ccg(L2);
#pragma omp for schedule(dynamic) collapse(2)
      for (int t = 0; t < T; ++t)
        for (int J = 0; J < ntmpl; ++J) {
          cc_map(&tiles[t * L2], &tsp[J * L2], l, ccg.data());
          out[((long)(ip - psi_begin) * T + t) * ntmpl + J] = reduce_norm0(ccg.data(), l, p->diameter);
        }
    }
  }
  return T;
}

extern "C" int or_pick_norm1(const or_params* p, const or_ctf* c, const float* tmpl, int ntmpl,
                             int bx, int by, const float* img, int nx, int ny, or_pick1* out,
                             int cap, int nthreads) {
  set_threads(nthreads);
  const int l = p->l;
  const long L2 = (long)l * l;
  int tx[4096], ty[4096];
  int bxn = or_tile_offsets(nx, l, p->overlap, tx, 4096);
  int byn = or_tile_offsets(ny, l, p->overlap, ty, 4096);
  if (bxn <= 0 || byn <= 0) return -1;
  const int T = bxn * byn;
  const int P = or_num_psi(p->phi_step);
  std::vector<cd> treal((long)ntmpl * L2);
  int rc = template_prep_impl(p, c, tmpl, ntmpl, bx, by, reinterpret_cast<double*>(treal.data()));
  if (rc) return rc;
  std::vector<cd> tiles((long)T * L2);
  rc = or_micrograph_prep(p, c, img, nx, ny, reinterpret_cast<double*>(tiles.data()), T);
  if (rc < 0) return rc;
  std::vector<cd> tsp((long)ntmpl * L2);
  /* pass 1: per-pixel sums over all views x psi, main.cu:1176-1197 */
  std::vector<double> s1((long)T * L2, 0.0), s2((long)T * L2, 0.0);
  for (int ip = 0; ip < P; ++ip) {
    float e = or_psi_value(p->phi_step, ip);
#pragma omp parallel for schedule(dynamic)
    for (int J = 0; J < ntmpl; ++J) rotate_normalize_one(p, &treal[J * L2], &tsp[J * L2], e);
#pragma omp parallel
    {
      std::vector<cd> ccg(L2);
#pragma omp for schedule(dynamic)
      for (int t = 0; t < T; ++t)
        for (int J = 0; J < ntmpl; ++J) {
          cc_map(&tiles[t * L2], &tsp[J * L2], l, ccg.data());
          for (long i = 0; i < L2; ++i) {
            double cur = ccg[i].real() / (double)L2;
            s1[t * L2 + i] += cur;
            s2[t * L2 + i] += cur * cur;
          }
        }
    }
  }
  /* set_CCG_mean, GPU_func.cu:635-646 */
  double total_n = (double)ntmpl * P;
  for (long i = 0; i < (long)T * L2; ++i) {
    double avg = s1[i] / total_n;
    double var = std::sqrt(s2[i] / total_n - avg * avg);
    s1[i] = avg;
    s2[i] = var;
  }
  /* pass 2: z-score and per-1024-pixel-block maxima, main.cu:771-821 */
  int count = 0;
  const int BLK = 1024;
  const long nblk = (L2 + BLK - 1) / BLK;
  for (int ip = 0; ip < P; ++ip) {
    float e = or_psi_value(p->phi_step, ip);
    for (int J = 0; J < ntmpl; ++J) rotate_normalize_one(p, &treal[J * L2], &tsp[J * L2], e);
    std::vector<cd> ccg(L2);
    for (int iy = 0; iy < byn; ++iy)
      for (int ix = 0; ix < bxn; ++ix) {
        int t = ix + iy * bxn;
        for (int J = 0; J < ntmpl; ++J) {
          cc_map(&tiles[t * L2], &tsp[J * L2], l, ccg.data());
          for (long b = 0; b < nblk; ++b) {
            double best = -INFINITY;
            long bpos = b * BLK;
            for (long i = b * BLK; i < std::min(L2, (b + 1) * BLK); ++i) {
              double avg = s1[t * L2 + i], var = s2[t * L2 + i];
              double cur = ccg[i].real() / (double)L2;
              double z = var > 0 ? (cur - avg) / var : 0;
              if (z > best) { best = z; bpos = i; }
            }
            if (best >= p->threshold) {
              int cx = tx[ix] + (int)(bpos % l);
              int cy = ty[iy] + (int)(bpos / l);
              if (cx >= p->diameter && cx < nx - p->diameter && cy >= p->diameter &&
                  cy < ny - p->diameter) {
                if (count < cap) out[count] = or_pick1{ip, J, (double)cx, (double)cy, best};
                count++;
              }
            }
          }
        }
      }
  }
  return count;
}

extern "C" int or_select_norm0(const or_params* p, const or_mapstat* st, int npsi, int psi_begin,
                               int ntmpl, int nx, int ny, or_pick1* out, int cap) {
  const int l = p->l;
  int tx[4096], ty[4096];
  int bxn = or_tile_offsets(nx, l, p->overlap, tx, 4096);
  int byn = or_tile_offsets(ny, l, p->overlap, ty, 4096);
  if (bxn <= 0 || byn <= 0) return -1;
  const int T = bxn * byn;
  int count = 0;
  for (int ip = 0; ip < npsi; ++ip)
    for (int iy = 0; iy < byn; ++iy)
      for (int ix = 0; ix < bxn; ++ix) {
        int t = ix + iy * bxn;
        for (int J = 0; J < ntmpl; ++J) {
          const or_mapstat& s = st[((long)ip * T + t) * ntmpl + J];
          double cx = tx[ix] + s.pos % l, cy = ty[iy] + s.pos / l;
          if (s.score >= p->threshold && cx >= p->diameter && cx < nx - p->diameter &&
              cy >= p->diameter && cy < ny - p->diameter) {
            if (count < cap) out[count] = or_pick1{psi_begin + ip, J, cx, cy, s.score};
            count++;
          }
        }
      }
  return count;
}

/* All T x N maps of one psi (the correlation loop of main.cu:826-893), OpenMP
 * over maps; out[t * n + J]. */
extern "C" int or_correlate_many(const or_params* p, const double* tile_spec, int T,
                                 const double* tspec, int n, or_mapstat* out, int nthreads) {
  set_threads(nthreads);
  const int l = p->l;
  const long L2 = (long)l * l;
  const cd* tiles = reinterpret_cast<const cd*>(tile_spec);
  const cd* ts = reinterpret_cast<const cd*>(tspec);
#pragma omp parallel
  {
    std::vector<cd> ccg(L2);
#pragma omp for schedule(dynamic) collapse(2)
    for (int t = 0; t < T; ++t)
      for (int J = 0; J < n; ++J) {
        cc_map(&tiles[t * L2], &ts[J * L2], l, ccg.data());
        out[(long)t * n + J] = reduce_norm0(ccg.data(), l, p->diameter);
      }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Refinement objective                                                     */
/* ------------------------------------------------------------------------ */
/* The reference has no refinement; this is the objective this build's
 * adjoint differentiates, built from the picker's own stages so that at a
 * grid psi, zero shift and the template CTF's defocus it is the picker's
 * score up to the peak exclusion in sigma (main.cu:856-881 drops the peak
 * pixel; here sigma covers the whole map, so the objective is smooth):
 *   A(k)  = W(k; defocus) B(k), B = whitened template spectrum (template_whiten),
 *           W = template_weight with theta's defocus, dfdiff/dfang of c0;
 *   T     = Re IFFT(A / m), m = amplitude normaliser of A (main.cu:444-452);
 *   R(x)  = bilinear(T) at the rotation of x by psi (rotate_subIMG,
 *           GPU_func.cu:496-560, continuous angle, double coordinates);
 *   Rn    = (mask R - mean) / sd over l^2 (apply_mask, compute_sum_sqr,
 *           substract_by_mean, divided_by_var);
 *   Th    = FFT(Rn) / l^2;  P(k) = I(k) conj(Th(k)) (-1)^(kx+ky) sgn
 *           (compute_corner_CCG, GPU_func.cu:584-615);
 *   c(u)  = sum_{kx <= l/2} w Re(P(k) exp(2 pi i k'.u / l)), w = 1 on the
 *           columns kx = 0 and l/2 and 2 between them, k' = (kx, signed ky in
 *           (-l/2, l/2]), u = (x0 + dx, y0 + dy): the Hermitian-completed
 *           Fourier series of the CC map, equal to the inverse FFT of P at
 *           integer u and continued between pixels;
 *   sigma = sqrt(sum_k |P(k)|^2 - P(0)^2) (Parseval: the sd of c over the map);
 *   score = c(u) / sigma.
 * theta = {psi (degrees, the euler3 rotation), dx, dy (pixels), defocus (um)}. */
extern "C" int or_template_whitened(const or_params* p, const float* tmpl, int ntmpl, int bx,
                                    int by, double* out) {
  const int l = p->l;
  if (bx > l || by > l) return -2;
  const long L2 = (long)l * l;
  cd* T = reinterpret_cast<cd*>(out);
  for (int J = 0; J < ntmpl; ++J) template_whiten(p, tmpl + (long)J * bx * by, bx, by, T + J * L2);
  return 0;
}

extern "C" int or_refine_score(const or_params* p, const or_ctf* c0, const double* white,
                               const double* tile, double x0, double y0, const double* theta,
                               double* out) {
  const int l = p->l;
  const long L2 = (long)l * l;
  const cd* B = reinterpret_cast<const cd*>(white);
  const cd* I = reinterpret_cast<const cd*>(tile);
  std::vector<cd> A(L2);
  or_ctf c = *c0;
  c.defocus_um = theta[3];
  for (int y = 0; y < l; ++y)
    for (int x = 0; x < l; ++x) A[(long)y * l + x] = B[(long)y * l + x] * template_weight(x, y, l, p, &c);
  double S = 0, cnt = 0;
  for (int y = 0; y < l; ++y)
    for (int x = 0; x <= l / 2; ++x) {
      int r = radial_bin(x, y, l, l);
      if (r < l / 2 && r >= 0) { S += std::norm(A[(long)y * l + x]); cnt += 1; }
    }
  const double m = std::sqrt(S / (cnt * cnt));
  if (m != 0)
    for (long i = 0; i < L2; ++i) A[i] /= m;
  fft2d(A.data(), l, l, +1);
  /* rotation, continuous angle */
  const double ang = theta[0] * REF_PI / 180;
  const double cose = std::cos(ang), sine = std::sin(ang);
  std::vector<double> R(L2, 0.0);
  const double dm = 1.5 * p->diameter, ehw = 6;
  double s1 = 0, s2 = 0;
  for (int j = 0; j < l; ++j)
    for (int i = 0; i < l; ++i) {
      const double X = i - l / 2, Y = j - l / 2;
      const double x2 = cose * X + sine * Y + l / 2, y2 = -sine * X + cose * Y + l / 2;
      double v = 0;
      if (!(x2 < 0 || x2 > l - 1.0 || y2 < 0 || y2 > l - 1.0)) {
        int ii = (int)std::floor(x2), jj = (int)std::floor(y2);
        long k0 = ii + (long)jj * l, k1 = k0 + 1, k2 = k0 + l + 1, k3 = k0 + l;
        if (ii == l - 1) { k1--; k2--; }
        if (jj == l - 1) { k2 -= l; k3 -= l; }
        const double t = x2 - ii, u = y2 - jj, tt = 1 - t, uu = 1 - u;
        v = A[k0].real() * tt * uu + A[k1].real() * t * uu + A[k2].real() * t * u + A[k3].real() * tt * u;
      }
      const double r = std::hypot(X, Y);
      if (r > dm / 2 + 2 * ehw) v = 0;
      else if (r >= dm / 2) v *= 0.5 * std::cos(REF_PI * (r - dm / 2) / (2 * ehw)) + 0.5;
      R[(long)j * l + i] = v;
      s1 += v;
      s2 += v * v;
    }
  const double mean = s1 / L2, var = s2 / L2 - mean * mean;
  const double sd = var > 0 ? std::sqrt(var) : 1.0;
  std::vector<cd> Th(L2);
  for (long i = 0; i < L2; ++i) Th[i] = cd((R[i] - mean) / sd, 0);
  fft2d(Th.data(), l, l, -1);
  const double sgn = ((l / 2) % 2 == 1) ? -1.0 : 1.0;
  const double ux = x0 + theta[1], uy = y0 + theta[2];
  double cc = 0, sq = 0;
  for (int ky = 0; ky < l; ++ky) {
    const double kys = ky <= l / 2 ? ky : ky - l;
    for (int kx = 0; kx < l; ++kx) {
      const cd P = I[(long)ky * l + kx] * std::conj(Th[(long)ky * l + kx] / (double)L2) *
                   (((kx + ky) & 1) ? -sgn : sgn);
      sq += std::norm(P);
      if (kx > l / 2) continue;
      const double w = (kx == 0 || kx == l / 2) ? 1.0 : 2.0;
      const double ph = 2 * M_PI * (kx * ux + kys * uy) / l;
      cc += w * (P.real() * std::cos(ph) - P.imag() * std::sin(ph));
    }
  }
  const cd P0 = I[0] * std::conj(Th[0] / (double)L2) * sgn;
  const double sigma = std::sqrt(sq - P0.real() * P0.real());
  out[0] = sigma > 0 ? cc / sigma : 0;
  out[1] = cc;
  out[2] = sigma;
  return 0;
}

// ---------------------------------------------------------------------------
// Duplicate removal of a pick list
// (python_scripts/remove_repeat_particles_from_list.py:9-145), restated in
// C++ with the same double-precision operations in the same order (no FMA
// contraction; cos/sin/acos/sqrt/fmod from the C library, as CPython's math
// module uses).
// ---------------------------------------------------------------------------
namespace {

const double kScriptPi = 3.14159265359;  // remove_repeat_particles_from_list.py:13

// Euler matrix of one pick, laid out as m_o in the script (:63-71); the
// script's m (:73-81) is its transpose with identical element values.
void euler_matrix(double alt_deg, double az_deg, double phi_deg, double* e) {
  const double alt = alt_deg * kScriptPi / 180, az = az_deg * kScriptPi / 180,
               phi = phi_deg * kScriptPi / 180;
  e[0] = std::cos(phi) * std::cos(az) - std::sin(phi) * std::cos(alt) * std::sin(az);
  e[1] = std::cos(phi) * std::sin(az) + std::sin(phi) * std::cos(alt) * std::cos(az);
  e[2] = std::sin(phi) * std::sin(alt);
  e[3] = (-1) * std::sin(phi) * std::cos(az) - std::cos(phi) * std::cos(alt) * std::sin(az);
  e[4] = (-1) * std::sin(phi) * std::sin(az) + std::cos(phi) * std::cos(alt) * std::cos(az);
  e[5] = std::cos(phi) * std::sin(alt);
  e[6] = std::sin(alt) * std::sin(az);
  e[7] = (-1) * std::sin(alt) * std::cos(az);
  e[8] = std::cos(alt);
}

// Orientation part of the pair test for outer pick i (matrix ei) and inner
// pick j (matrix ej), :83-136.  Returns 0 when the script skips the pair
// (`continue`), else 1 with *dist the script's euler_dist.  The quaternion
// branches follow the script exactly: for a non-positive trace the
// x-dominant case (:102-109) is always overwritten by the y/z case below it.
int euler_pair(const double* ei, const double* ej, double* dist) {
  const double* m = ei;   // transposed access below: m[r*3+c] = ei[c*3+r]
  const double* mo = ej;
  double mt[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) mt[r * 3 + c] = m[c * 3 + r];
  double rm[9];
  rm[0] = mt[0] * mo[0] + mt[3] * mo[1] + mt[6] * mo[2];
  rm[1] = mo[0] * mt[1] + mo[1] * mt[4] + mo[2] * mt[7];
  rm[2] = mo[0] * mt[2] + mo[1] * mt[5] + mo[2] * mt[8];
  rm[3] = mo[3] * mt[0] + mo[4] * mt[3] + mo[5] * mt[6];
  rm[4] = mo[3] * mt[1] + mo[4] * mt[4] + mo[5] * mt[7];
  rm[5] = mo[3] * mt[2] + mo[4] * mt[5] + mo[5] * mt[8];
  rm[6] = mo[6] * mt[0] + mo[7] * mt[3] + mo[8] * mt[6];
  rm[7] = mo[6] * mt[1] + mo[7] * mt[4] + mo[8] * mt[7];
  rm[8] = mo[6] * mt[2] + mo[7] * mt[5] + mo[8] * mt[8];
  double q0 = 0, t;
  const double tr = rm[0] + rm[4] + rm[8];
  if (tr > 0) {
    t = std::sqrt(1 + rm[0] + rm[4] + rm[8]) * 2;  // left to right, as :94
    if (t == 0) return 0;
    q0 = 0.25 * t;
  }
  if (tr <= 0) {
    if (rm[0] >= rm[4] && rm[0] >= rm[8]) {
      t = std::sqrt(1 + rm[0] - rm[4] - rm[8]) * 2;
      if (t == 0) return 0;
      q0 = (rm[7] - rm[5]) / t;
    }
    if (rm[4] > rm[8]) {
      if (1 - rm[0] + rm[4] - rm[8] < 0) return 0;
      t = std::sqrt(1 - rm[0] + rm[4] - rm[8]) * 2;
      if (t == 0) return 0;
      q0 = (rm[2] - rm[6]) / t;
    } else {
      if (1 - rm[0] - rm[4] + rm[8] < 0) return 0;
      t = std::sqrt(1 - rm[0] - rm[4] + rm[8]) * 2;
      if (t == 0) return 0;
      q0 = (rm[3] - rm[1]) / t;
    }
  }
  if (q0 > 1 || q0 < -1) return 0;
  *dist = std::fabs(std::acos(q0) * 2 * 180 / kScriptPi);
  return 1;
}

bool euler_close(double euler_dist, double thr) {
  double dist2 = std::fmod(euler_dist, 180.0);  // Python % on non-negative floats
  if (dist2 > thr + 1) dist2 = 180 - dist2;
  return euler_dist < thr || dist2 < thr;
}

}  // namespace

extern "C" int or_euler_distance(double alt1, double az1, double phi1, double alt2, double az2,
                                 double phi2, double* dist) {
  double ei[9], ej[9];
  euler_matrix(alt1, az1, phi1, ei);
  euler_matrix(alt2, az2, phi2, ej);
  return euler_pair(ei, ej, dist);
}

extern "C" int or_remove_repeats(int n, const int* group, const double* x, const double* y,
                                 const double* alt, const double* az, const double* phi,
                                 const double* score, double center, double euler_thres,
                                 unsigned char* keep) {
  if (n < 0) return -1;
  std::vector<double> e((size_t)n * 9);
  for (int i = 0; i < n; ++i) euler_matrix(alt[i], az[i], phi[i], &e[(size_t)i * 9]);
  std::vector<unsigned char> a(n, 0);
  int lo = 0;
  while (lo < n) {  // groups: runs of equal micrograph (:25-36)
    int hi = lo + 1;
    while (hi < n && group[hi] == group[lo]) ++hi;
    for (int i = lo; i < hi; ++i) {  // :39-141
      if (a[i]) continue;
      for (int j = i + 1; j < hi; ++j) {
        if (a[j]) continue;
        const double dx = x[j] - x[i], dy = y[j] - y[i];
        const double dist = std::fabs(std::sqrt(dx * dx + dy * dy));
        if (!(dist < center)) continue;
        double ed;
        if (!euler_pair(&e[(size_t)i * 9], &e[(size_t)j * 9], &ed)) continue;
        if (euler_close(ed, euler_thres)) {
          if (score[i] > score[j] && a[j] == 0) a[j] = 1;
          if (score[i] <= score[j] && a[i] == 0) a[i] = 1;
        }
      }
    }
    lo = hi;
  }
  for (int i = 0; i < n; ++i) keep[i] = a[i] ? 0 : 1;
  return 0;
}

// ---------------------------------------------------------------------------
// Template projection of a density map (the step before the path: the
// reference takes templates from EMAN's project3d, protocol step 5, not
// shipped).  out[j](row, col) = sum_{w=-c}^{box-1-c} trilinear(vol,
// R_j^T (col-c, row-c, w) + c), c = box/2, R = Rz(phi) Rx(alt) Rz(az), zero
// outside the box; every sample of every ray, in FP64.
// ---------------------------------------------------------------------------
extern "C" int or_project(const float* vol, int n, const double* euler, int nviews, double* out) {
  if (!vol || !euler || !out || n < 2) return -1;
  const int c = n / 2;
  auto vox = [&](long x, long y, long z) -> double {
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return 0.0;
    return vol[(z * n + y) * n + x];
  };
  for (int j = 0; j < nviews; ++j) {
    const double d = M_PI / 180.0;
    const double alt = euler[3 * j] * d, az = euler[3 * j + 1] * d, phi = euler[3 * j + 2] * d;
    const double ca = std::cos(alt), sa = std::sin(alt), cz = std::cos(az), sz = std::sin(az),
                 cp = std::cos(phi), sp = std::sin(phi);
    const double R[9] = {cp * cz - ca * sz * sp,  cp * sz + ca * cz * sp,  sp * sa,
                         -sp * cz - ca * sz * cp, -sp * sz + ca * cz * cp, cp * sa,
                         sa * sz,                 -sa * cz,                ca};
    for (int row = 0; row < n; ++row)
      for (int col = 0; col < n; ++col) {
        const double u = col - c, v = row - c;
        double s = 0;
        for (int w = -c; w <= n - 1 - c; ++w) {
          const double x = u * R[0] + v * R[3] + w * R[6] + c;
          const double y = u * R[1] + v * R[4] + w * R[7] + c;
          const double z = u * R[2] + v * R[5] + w * R[8] + c;
          const double fx0 = std::floor(x), fy0 = std::floor(y), fz0 = std::floor(z);
          const long x0 = (long)fx0, y0 = (long)fy0, z0 = (long)fz0;
          const double fx = x - fx0, fy = y - fy0, fz = z - fz0;
          double acc = 0;
          for (int k = 0; k < 8; ++k) {
            const int ix = k & 1, iy = (k >> 1) & 1, iz = k >> 2;
            const double wgt = (ix ? fx : 1 - fx) * (iy ? fy : 1 - fy) * (iz ? fz : 1 - fz);
            acc += wgt * vox(x0 + ix, y0 + iy, z0 + iz);
          }
          s += acc;
        }
        out[((size_t)j * n + row) * n + col] = s;
      }
  }
  return 0;
}
