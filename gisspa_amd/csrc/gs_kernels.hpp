// gs_kernels.hpp -- host-visible launch interface of the picker's HIP kernels.
//
// Spectrum layout ("column-major half spectrum"): an L x L real image has an
// R2C half spectrum of H = L/2 + 1 columns kx; it is stored as [kx][ky], each
// column of L complex values contiguous (ky fastest).  Row-major intermediate
// of the 2-D inverse: [y][kx] with a padded row pitch HP.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

namespace gs {

// per-workgroup partial of the correlation reduction
struct CorrPartial {
  float vmax;   // max over valid (border-excluded) pixels with value > 0, else 0
  int pos;      // y*L + x of vmax (lowest index on ties), INT_MAX if none
  float s1;     // sum of all pixels of the rows handled
  float s2;     // sum of squares
};

// per-map result of the norm_type 0 reduction (main.cu:856-881 semantics)
struct MapStat {
  float score;
  int pos;
  float peak;
  float sd;
  double sum;
  double sum2;
};

struct MapRef {
  int tile;
  int view;
};

// where a tile of the (multi-micrograph) tile pool comes from
struct TileDesc {
  int mic;     // micrograph of the batch
  int ox, oy;  // offset of the tile in its micrograph
  int nx, ny;  // micrograph size
};

// one selected correlation map (norm_type 0)
struct SelRec {
  long long m;  // (psi - psi_begin) * T * N + tile * N + view
  float cx, cy, score, pad;
};

// geometry shared by host and device code
struct Geometry {
  int L, M, H, HP;
  int RB;  // rows per workgroup for M-point transforms
  int CB;  // columns per workgroup for L-point transforms
  long gmap;  // float2 elements of one map's correlation intermediate
};

bool geometry_supported(int L);
Geometry geometry_of(int L);

// ---- launches (all asynchronous on `stream`) ----

// A range of whole row blocks (RB rows each) of an L x L image: blocks
// rb0 .. rb0 + nrb - 1, rows ylo() .. yhi().
struct RowBand {
  int rb0 = 0, nrb = 0, RB = 0;
  int ylo(int) const { return rb0 * RB; }
  int yhi(int L) const { return std::min(L, (rb0 + nrb) * RB) - 1; }
};
RowBand full_band(int L);
// The rows of a rotated template masked at diameter dmask (the picker's
// apply_mask, GPU_func.cu:229-247: zero beyond radius dmask / 2 + 12 of the
// window centre) that can be non-zero or be read by the bilinear rotation and
// its adjoint gather (which reach at most 5 rows further out), with 20 rows
// of margin; every row for a NaN diameter or a mask wider than the window.
RowBand rot_row_band(int L, float dmask);

// Real rows -> half spectrum [b][kx][y] (unscaled), from padded templates:
// out[b] = R2C rows of pad(src[b], box bx x by centred) * scale[b]
void launch_rows_r2c_pad(int L, const float* src, int bx, int by, const float* scale, int nb,
                         float2* out, const float2* twL, const float2* twM, hipStream_t s);
// Real rows -> half spectrum from micrograph tiles: tile (ix, iy) of the
// image, minus `mean` (device scalar), zero outside the image.
void launch_rows_r2c_tiles(int L, const float* img, int nx, int ny, const int* offx,
                           const int* offy, int bxn, int ntiles, const double* mean,
                           float2* out, const float2* twL, const float2* twM, hipStream_t s);
// Real rows -> half spectrum from rotated + masked real templates, with
// per-(template, row block) partial sums of value and value^2 (band.nrb per
// template).  Only the rows of `band` are transformed and written: the band
// must hold every row the mask leaves non-zero (rot_row_band).
void launch_rows_r2c_rot(int L, const float* treal, int nb, float cose, float sine, float dmask,
                         RowBand band, float2* out, double* partials, const float2* twL,
                         const float2* twM, hipStream_t s);
// Same with one angle per b (cs_b[b] = (cos, sin), coordinates in double
// precision) and the interpolated value multiplied by pre_b[b].
void launch_rows_r2c_rotb(int L, const float* treal, int nb, const double2* cs_b, const float* pre_b,
                          float dmask, RowBand band, float2* out, double* partials,
                          const float2* twL, const float2* twM, hipStream_t s);
// Real rows -> half spectrum of real images imgs[b][L][L], the rows of `band`.
void launch_rows_r2c_img(int L, const float* imgs, int nb, RowBand band, float2* out,
                         const float2* twL, const float2* twM, hipStream_t s,
                         const int* skip = nullptr);

// In-place column transform of [b][kx][y]; multiplies by scale (global) or
// per-b scale from `bscale` (may be null); zero_dc zeroes element (0,0).
// Inputs outside the rows of `band` are taken as zero (every row when
// band.nrb == 0).
void launch_cols_fwd(int L, float2* spec, int nb, float scale, const float* bscale, int zero_dc,
                     const float2* twL, hipStream_t s, RowBand band = RowBand());
// Per-template normalisation scale 1/(L^2 sd) of the rotated templates from
// the nparts row partials of each (main.cu:682-707 semantics).
void launch_rot_scale(int L, const double* partials, int nb, int nparts, float* bscale,
                      hipStream_t s);
// Column inverse of [b][kx][y] written row-major to G[b][y][kx] (pitch HP).
void launch_cols_inv_T(int L, const float2* spec, int nb, float2* G, const float2* twL,
                       hipStream_t s);
// Row C2R of G[b][y][kx] into real images out[b][y][x].
void launch_rows_c2r(int L, const float2* G, int nb, float* out, const float2* twL,
                     const float2* twM, hipStream_t s);
// the same transform with the spectra read and pre-processed on the way into
// the first pass (no LDS staging); equal to rounding
// (the rows of `band` only)
void launch_rows_c2r_stream(int L, const float2* G, int nb, RowBand band, float* out,
                            const float2* twL, const float2* twM, hipStream_t s);

// Correlation pass 1: P = tile[t] * conj(tmpl[J]) (the tile spectra carry
// (-1)^(kx+ky) sgn), column inverse and C2R pre-processing, written to the
// paired-column blocks of G (Geometry::gmap elements per map).
// psum[m * corr_ncb(L) + cb] = (sum of weighted |P|^2, P(0,0).re) of the block
// `variant` selects the FFT schedule (radix order) of the pass: 0 is the
// tuned plan, 1 .. corr_variants(L) - 1 the alternatives gs_autotune times.
void launch_corr_pass1(int L, const float2* tiles, const float2* tspec, const MapRef* maps,
                       int nmaps, int ylo, int yhi, float2* G, float2* psum, const float2* twL,
                       hipStream_t s, int variant = 0);
// Rows ylo..yhi of the intermediate that the peak search of pass 2 reads for
// the border-exclusion bounds lo, hi (pass 1 may leave the others unwritten;
// norm_type 1 reads every row, 0 .. L-1)
void corr_store_rows(int L, float lo, float hi, int* ylo, int* yhi);
// schedule variants compiled for window L (1 = the tuned plan only), and
// the most any window has
int corr_variants(int L);
int corr_max_variants();
// one launch: pass 1 of maps1 (into G1/psum1) and pass 2 of the maps whose
// intermediate G2 a previous launch wrote (into part2); either count may be 0
void launch_corr_fused(int L, const float2* tiles, const float2* tspec, const MapRef* maps1,
                       int nmaps1, float2* G1, float2* psum1, const float2* G2, int nmaps2,
                       float lo, float hi, CorrPartial* part2, const float2* twL,
                       const float2* twM, hipStream_t s);
void launch_fill_int(int* p, long n, int v, hipStream_t s);
// one launch for all maps of a psi: pass 1 and pass 2 with in-kernel
// dependencies through a ring of `ring` intermediates (lag < ring);
// counters: 2 nmaps ints (zeroed by the launcher); *err set if a wait ran
// out (wait_ticks: budget of one wait in wall-clock counter ticks)
void launch_corr_stream(int L, const float2* tiles, const float2* tspec, const MapRef* maps,
                        int nmaps, int lag, int ring, float2* Gring, float2* psum, int* counters,
                        float lo, float hi, CorrPartial* part, int* err, long long wait_ticks,
                        const float2* twL, const float2* twM, hipStream_t s);
// multiply tile half spectra by (-1)^(kx+ky) sgn (compute_corner_CCG's factor)
void launch_checker_sign(int L, float2* spec, int ntiles, hipStream_t s);
int corr_ncb(int L);
// Correlation pass 2: row C2R of G and the border-excluded max / sums, one
// CorrPartial per (map, row block): part[m * corr_nparts(L) + block].
void launch_corr_pass2(int L, const float2* G, int nmaps, float dm4_lo, float dm4_hi,
                       CorrPartial* part, const float2* twL, const float2* twM, hipStream_t s,
                       int variant = 0);
int corr_nparts(int L);
// Combine the partials of nmaps maps (processing order `maps`) into
// out[tile * nviews + view] (score in FP64).
void launch_corr_finalize(int L, const CorrPartial* part, const float2* psum, int nmaps,
                          const MapRef* maps, int nviews, MapStat* out, hipStream_t s);

// ---- norm_type 1 ----
// s1/s2[tile][pixel] += sums of c / L^2 and (c / L^2)^2 over a tile-major
// batch of ntiles x nv maps (maps t * nv .. t * nv + nv - 1 share a tile)
void launch_norm1_acc(int L, const float2* G, const MapRef* maps, int ntiles, int nv, double* s1,
                      double* s2, const float2* twM, hipStream_t s);
// per pixel pair of n (even) pixels: {mean, mean, 1/sd, 1/sd} (float4)
void launch_norm1_stats(const double* s1, const double* s2, long n, double total, void* norm,
                        hipStream_t s);
// per-run maxima of the z-scores of a batch of maps, without z-maps:
// keys[m][run][2] (norm1_key_bytes(L) bytes per map)
void launch_norm1_zrun(int L, const float2* G, const MapRef* maps, int nmaps, const void* norm,
                       void* keys, const float2* twM, hipStream_t s);
long norm1_key_bytes(int L);
// records of 1024-pixel runs above threshold: m = mbase + view * ntiles +
// tile (canonical view-major map index), pos = y * L + x of the run maximum
// (its run is pos / 1024), z its z-score.  *count (64-bit) counts every
// record; the first cap are written, so a caller that reads count > cap
// grows the buffer and runs again.
struct N1Rec {
  long long m;
  int pos;
  float z;
};
void launch_norm1_runfin(int L, const void* keys, int nmaps, long long mbase, int ntiles,
                         const MapRef* maps, const TileDesc* tdesc, float thres, float dmin,
                         unsigned long long* count, void* out, long long cap, hipStream_t s);
// threshold + centre-bounds selection of norm_type 0 maps (main.cu:871-891);
// *count counts every selected map, the first cap records are written
void launch_select(const MapStat* st, long long nmaps, int T, int N, int L, const TileDesc* tdesc,
                   float thres, float dmin, unsigned long long* count, SelRec* out,
                   long long cap, hipStream_t s);
int norm1_record_bytes();

// ---- template preparation ----
// per-template sums of value and value^2 of the raw box (FP64)
void launch_box_sums(const float* src, int npix, int nb, double* sums, hipStream_t s);
// 1/sigma per template (sigma over the L^2 window, em = 0)
void launch_sigma_scale(const double* sums, int nb, int L, float* scale, hipStream_t s);
// radial power bins (deterministic CSR gather): ra[b][r] = sum |F|^2, r < L/2
void launch_radial_power(int L, const float2* spec, int nb, const int* bin_start,
                         const int* bin_pix, double* ra, hipStream_t s);
// whiten + low-pass + |CTF| weighting, in place on the half spectrum, and the
// per-(template, column block) partial sums of |F|^2 inside the radius.
struct WeightArgs {
  float apix, highres, lowres, kk, ds;
  float a, b, b2, bf, bf2, bf3;
  float dfu, dfv, dfdiff, dfang, lambda, cs, ampconst;
};
// whitening + (weight != 0) low pass and |CTF| weighting with the amplitude
// normaliser's partial sums; weight == 0 leaves the whitened spectrum
void launch_tmpl_weight(int L, float2* spec, int nb, const double* ra, const float* rb,
                        const WeightArgs& w, double* partials, int weight, hipStream_t s);
// divide every template by its amplitude normaliser (inv_m: nb floats scratch)
void launch_tmpl_ampnorm(int L, float2* spec, int nb, const double* partials, double count,
                         float* inv_m, hipStream_t s);

// ---- micrograph ----
void launch_img_sums(const float* img, long n, double* partials, int nblocks, hipStream_t s);
void launch_img_bounds(const double* partials, int nblocks, long n, double* stats, hipStream_t s);
void launch_img_clamp(float* img, long n, int invert, const double* stats, double* partials,
                      int nblocks, hipStream_t s);
void launch_img_mean(const double* partials, int nblocks, long n, double* mean, hipStream_t s);

int img_reduce_blocks(long n);

// ---- runtime-length FFTs and the phase-flip micrograph path (gs_rtfft.hip) ----
void rt_fft_2d(float2* data, float2* tmp, int nx, int ny, const float2* twx, const float2* twy,
               int dir, hipStream_t s);
void launch_real_to_complex(const float* in, float2* out, long n, float scale, hipStream_t s);
void launch_complex_to_real(const float2* in, float* out, long n, hipStream_t s);
void launch_phase_flip(float2* F, int nx, int ny, const WeightArgs& w, hipStream_t s);
void launch_img_radial(const float2* F, const int* start, const long* pix, int nbin, double* ra,
                       hipStream_t s);
void launch_img_weight(float2* F, int nx, int ny, const double* ra, const float* rb,
                       const WeightArgs& w, hipStream_t s);
void launch_tile_ampnorm(float2* spec, int L, int ntiles, double count, double* part,
                         float* scale, hipStream_t s);
int tile_ampnorm_parts(int L);

// ---- stage-by-stage ("plain") picker (gs_plain.hip) ----
struct PeakPartHost {
  float peak;
  int pos;
  float s1, s2;
};
int plain_blocks(int L);
void launch_plain_rotate(int L, const float* treal, int n, float cose, float sine, float dmask,
                         float2* out, double* part, hipStream_t s);
void launch_plain_normalize(int L, float2* data, int n, const float* mean, const float* var,
                            hipStream_t s);
void launch_c2c_2d(int L, float2* data, int nimg, int dir, float scale, const float2* twL,
                   hipStream_t s);
void launch_plain_tiles(int L, const float* img, int nx, int ny, const TileDesc* td, int tile0,
                        int ntiles, const double* mean, float2* out, hipStream_t s);
void launch_plain_ccg(int L, const float2* tile, const float2* tspec, int n, float sgn,
                      float2* ccg, hipStream_t s);
void launch_plain_peak(int L, const float2* ccg, int n, float lo, float hi, void* part,
                       hipStream_t s);

// ---- refinement objective and its adjoint (gs_refine.hip) ----
struct RefCand {
  int ptile, view;   // tile of the pool, view (template)
  float x0, y0;      // peak position in the tile
  float psi, dx, dy, df;  // psi degrees (euler3), shift (pixels), defocus (um)
};
struct RefScal {
  double m, dm, mu, sd, c, sigma, score, gx, gy, gmean, grn, dpsi, dmacc, ddf;
};
struct RefArgs {
  int nb;
  const float2* white;  // [N][H][L] whitened template half spectra
  const float2* tiles;  // tile pool
  const RefCand* cand;
  WeightArgs w;         // template CTF weighting (defocus replaced per candidate)
  const float4* wtab;   // [H][L] its defocus-independent part (launch_ref_wtable)
  double count;         // cells of the amplitude normaliser
  float dmask;          // 1.5 diameter
  const double2* cs;    // (cos, sin) of psi per candidate
  float2* A;            // [nb][H][L] scratch half spectra
  float2* gtmp;         // [nb][L][HP]
  float* Tun;           // [nb][L][L]
  float2* spec;         // [nb][H][L]
  double* rot_part;     // [nb][row blocks][2]
  float* bscale;        // [nb]
  float* pre;           // [nb]
  float* gTb;           // [nb][L][L]
  float* gT;            // [nb][L][L]
  double* part;         // [nb][ref_parts()][4]
  RefScal* sc;          // [nb]
  float* out;           // [nb][5]: score, d psi, d dx, d dy, d defocus
  const float* thr;     // [nb] or null: the adjoint runs only where the score
                        // (as float) exceeds thr (a refinement step's test)
  int* skip;            // [nb] work flags: set where the adjoint is skipped
  const float2* twL;
  const float2* twM;
};
// defocus-independent part of the refinement's CTF weighting per half-spectrum
// cell ([H][L] float4), built once per template set
void launch_ref_wtable(int L, const WeightArgs& w, float4* tab, hipStream_t s);
void launch_ref_forward(int L, const RefArgs& a, hipStream_t s);
void launch_ref_backward(int L, const RefArgs& a, hipStream_t s);
int ref_parts();

}  // namespace gs
