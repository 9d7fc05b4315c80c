/*
 * gisspa_oracle.h -- FP64 CPU restatement of the GisSPA picking pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in gisspa_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never calls it.
 *
 * Every stage follows the reference (Autumn1998/GisSPA) file:line cited at
 * its definition in gisspa_oracle.cpp.  Arithmetic is double precision with
 * sequential (deterministic) sums; where the reference uses float-specific
 * control values (the psi accumulator, integer-truncated bounds) the oracle
 * reproduces those values exactly.
 *
 * Complex arrays are interleaved (re, im) doubles, row-major [y][x], x fastest.
 */
#ifndef GISSPA_ORACLE_H
#define GISSPA_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the arithmetic fields of the reference's Parameters
 * (EMReader/DataReader.h:14-42).  All values are as read from the config. */
typedef struct or_params {
  double apix, phi_step, kk, energy, cs, highres, lowres, diameter;
  double threshold, ampconst;
  double a, b, b2, bfactor, bfactor2, bfactor3;
  int l;            /* window_size after adjustPaddingSize */
  int l_config;     /* window_size as configured (ds uses it, main.cu:155) */
  int overlap;      /* -1 = not given */
  int phase_flip;
  int invert;
  int norm_type;
} or_params;

/* Per-micrograph CTF values from the .lst (defocus in um, positive). */
typedef struct or_ctf {
  double defocus_um, dfdiff, dfang;
} or_ctf;

/* One candidate per (psi, tile, view) for norm_type 0. */
typedef struct or_mapstat {
  double peak, sum, sum2, score;
  int pos;          /* y*l + x of the peak inside the tile */
  int pad;
} or_mapstat;

/* ---- FFT (mixed radix, any length) ---- */
int or_fft1d(double* data, int n, int sign);                 /* in place */
int or_fft2d(double* data, int nx, int ny, int sign);        /* in place, unnormalised */

/* ---- individual stages ---- */
double or_ctf_ast(int x1, int y1, int nx, int ny, const or_params* p,
                  const or_ctf* c, int mode);
int or_tile_offsets(int n, int l, int overlap, int* offs, int cap);
int or_template_prep(const or_params* p, const float* tmpl, int ntmpl, int bx, int by,
                     double* out /* ntmpl*l*l complex */);
/* same, with the CTF used for the |CTF| weighting (the reference uses the
 * first micrograph's values, main.cu:1106 then 1133) */
int or_template_prep_ctf(const or_params* p, const or_ctf* c, const float* tmpl, int ntmpl,
                         int bx, int by, double* out);
int or_rotate_normalize(const or_params* p, const double* tmpl_real /* n*l*l complex */,
                        int n, float euler3, double* spec_out /* n*l*l complex */);
int or_micrograph_prep(const or_params* p, const or_ctf* c, const float* img, int nx, int ny,
                       double* tile_spec /* T*l*l complex */, int tcap);
/* real CC map of one (tile, template) pair (before the reduction) */
int or_cc_map(const or_params* p, const double* tile_spec, const double* tspec, double* out);
/* bilinear rotation of a real l x l image (rotate_subIMG) */
int or_rotate_bilinear(const double* in, int l, float e, double* out);
int or_correlate_norm0(const or_params* p, const double* tile_spec, const double* tspec,
                       or_mapstat* out);
int or_correlate_many(const or_params* p, const double* tile_spec, int T, const double* tspec,
                      int n, or_mapstat* out, int nthreads);
/* ---- whole pipeline ---- */
int or_pick_norm0(const or_params* p, const or_ctf* c, const float* tmpl, int ntmpl, int bx, int by,
                  const float* img, int nx, int ny, int psi_begin, int psi_end,
                  or_mapstat* out /* (psi_end-psi_begin)*T*ntmpl */, int nthreads);
/* norm_type 1 picks: writes up to cap (psi_index, J, cx, cy, score) records,
 * returns count (or <0 on error). */
typedef struct or_pick1 { int psi_index, view; double cx, cy, score; } or_pick1;
int or_pick_norm1(const or_params* p, const or_ctf* c, const float* tmpl, int ntmpl, int bx, int by,
                  const float* img, int nx, int ny, or_pick1* out, int cap, int nthreads);
/* Threshold + centre-bounds selection of norm_type 0 candidates
 * (main.cu:871-891), canonical order psi -> tile (y outer, x inner) -> view.
 * Returns the number of picks (records beyond cap are counted, not written). */
int or_select_norm0(const or_params* p, const or_mapstat* st, int npsi, int psi_begin, int ntmpl,
                    int nx, int ny, or_pick1* out, int cap);
int or_num_psi(double phi_step);
/* Refinement objective (no reference counterpart; defined in the .cpp):
 * whitened template spectra (l x l complex per template) and the score of
 * one candidate at theta = {psi deg, dx, dy, defocus um}; out = {score, c, sigma}. */
int or_template_whitened(const or_params* p, const float* tmpl, int ntmpl, int bx, int by,
                         double* out);
int or_refine_score(const or_params* p, const or_ctf* c0, const double* white,
                    const double* tile, double x0, double y0, const double* theta, double* out);
float or_psi_value(double phi_step, int index);

/* ---- duplicate removal of a pick list
 *      (python_scripts/remove_repeat_particles_from_list.py:9-145) ---- */
/* The script's orientation distance (degrees) between an outer pick 1 and
 * an inner pick 2; returns 0 where the script skips the pair. */
int or_euler_distance(double alt1, double az1, double phi1, double alt2, double az2,
                      double phi2, double* dist);
/* keep[i] = 0 for the picks the script drops.  Groups are runs of equal
 * `group` values (the script's runs of equal micrograph path). */
int or_remove_repeats(int n, const int* group, const double* x, const double* y,
                      const double* alt, const double* az, const double* phi,
                      const double* score, double center, double euler_thres,
                      unsigned char* keep);

/* ---- template projection (EMAN project3d's role, protocol step 5) ---- */
/* out[j][row][col] = sum over w of the trilinear value of vol ([z][y][x],
 * n^3) at R_j^T (col-c, row-c, w) + c, c = n/2, R = Rz(phi) Rx(alt) Rz(az)
 * from euler[3j..3j+2] = (alt, az, phi) degrees; zero outside the box. */
int or_project(const float* vol, int n, const double* euler, int nviews, double* out);

/* Threads of the oracle's parallel loops from now on (0 leaves them as is). */
int or_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
