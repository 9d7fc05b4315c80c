/*
 * gisspa.h -- C ABI of the MI355X GisSPA picker (libgisspa.so).
 *
 * The reference (Autumn1998/GisSPA) has no library or FFI: its picking path
 * is a set of free functions in main.cu called from main() with raw device
 * pointers (GPU_func.cuh:43-77).  This header replaces that host<->kernel
 * seam with a context API.  Each entry point names the reference code it
 * replaces.  Conventions:
 *   - every function returns GS_OK (0) or a negative gs_status; nothing
 *     calls exit() (the reference's CUDA_CALL/CUFFT_CALL exit(-1),
 *     GPU_func.cuh:33-41);
 *   - the caller owns host buffers; the context owns device memory, the
 *     stream and the FFT tables;
 *   - a context is used by one host thread at a time; several contexts may
 *     coexist (one per device or stream);
 *   - outputs are in a deterministic, canonical order.
 */
#ifndef GISSPA_H
#define GISSPA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gs_status {
  GS_OK = 0,
  GS_ERR_ARG = -1,          /* invalid argument */
  GS_ERR_UNSUPPORTED = -2,  /* e.g. a window size without a compiled kernel */
  GS_ERR_DEVICE = -3,       /* HIP runtime error */
  GS_ERR_NOMEM = -4,        /* device allocation failed */
  GS_ERR_STATE = -5,        /* call order (templates / micrograph not set) */
  GS_ERR_CAPACITY = -6,     /* output buffer too small (n_out holds the need) */
  GS_ERR_IO = -7            /* file missing, unreadable or in an unsupported format
                               (gs_io_error() holds the message) */
} gs_status;

/* Run parameters: the reference's Parameters (EMReader/DataReader.h:14-42)
 * with the same config-key meanings.  gs_params_default() fills the
 * reference's defaults. */
typedef struct gs_params {
  /* required keys */
  float apix;        /* angpix, A/pixel */
  float phi_step;    /* phistep, degrees */
  float kk;          /* kk */
  float energy;      /* energy, kV */
  float cs;          /* cs, mm */
  float highres;     /* Highres, A */
  float lowres;      /* Lowres, A */
  float diameter;    /* diameter, pixels */
  /* optional keys (reference defaults) */
  float threshold;   /* threshold = 7 */
  int window_size;   /* window_size = 320 */
  int phase_flip;    /* phase_flip = 0 */
  int overlap;       /* overlap = -1 (stride window_size + 1) */
  int norm_type;     /* norm_type = 0 */
  int invert;        /* invert = 0 */
  float ampconst;    /* ampconst = 0.07 (fixed in the reference, a key here) */
  /* fixed SNR model (DataReader.h:25) */
  float a, b, b2, bfactor, bfactor2, bfactor3;
} gs_params;

/* Micrograph CTF values as written in the micrograph .lst
 * (defocus in um, positive underfocus; dfdiff um; dfang degrees),
 * EMReader/DataReader.cpp:366-387. */
typedef struct gs_ctf {
  float defocus;
  float dfdiff;
  float dfang;
} gs_ctf;

/* One picked particle: a line of the reference's output (main.cu:985). */
typedef struct gs_particle {
  int psi_index;   /* index of the in-plane step (euler3 = psi_index * phistep) */
  int view;        /* Euler-list line J */
  int tile;        /* tile index ix + iy * bx within its micrograph */
  float psi;       /* printed psi = 360 - euler3, wrapped to [0, 360) */
  float cx, cy;    /* centre in micrograph pixels */
  float score;
  int mic;         /* micrograph of the batch (0 for gs_set_micrograph) */
} gs_particle;

/* Execution modes (gs_set_mode). */
enum {
  GS_MODE_FUSED = 0,  /* half-spectrum correlation, fused product/transform/reduction,
                         all micrographs of a batch in one tile pool (default) */
  GS_MODE_PLAIN = 1   /* stage-by-stage like the reference: full complex spectra, one
                         kernel per stage, host reduction after every tile */
};

/* A candidate for the refinement (gs_adjoint, gs_refine). */
typedef struct gs_peak {
  int mic, tile, view;  /* as in gs_particle */
  float cx, cy;         /* picked centre in micrograph pixels (tile offset + peak position) */
  float psi;            /* in-plane rotation of the template, degrees: euler3
                           (gs_particle.psi prints 360 - euler3) */
  float dx, dy;         /* sub-pixel shift of the centre, pixels */
  float defocus;        /* defocus of the template CTF weighting, um (positive underfocus) */
} gs_peak;

/* Refinement objective of a candidate and its gradient. */
typedef struct gs_grad {
  float score;
  float d_psi, d_dx, d_dy, d_defocus;  /* per degree, pixel, pixel, um */
} gs_grad;

/* Per (psi, tile, view) correlation result of norm_type 0
 * (main.cu:856-881): peak of the border-excluded CC map, its position
 * y*l + x in the tile, sums over the whole map, and the score. */
typedef struct gs_mapstat {
  float score;
  int pos;
  float peak;
  float sd;
  double sum;
  double sum2;
} gs_mapstat;

typedef struct gs_info {
  int l;           /* window (tile) size after adjustment */
  int n_views;
  int n_psi;
  int nx, ny;      /* first micrograph of the current batch */
  int tiles_x, tiles_y;
  int box_x, box_y;
  int n_mics;      /* micrographs in the current batch */
  int n_tiles;     /* tiles of all of them */
} gs_info;

/* Device time of the last calls, milliseconds (hipEvent). */
typedef struct gs_timing {
  double pick_total_ms;        /* whole last gs_pick */
  double micrograph_prep_ms;   /* last gs_set_micrograph(s): upload, prep, tile FFTs */
  /* With gs_set_profiling on, every correlation launch of the first psi of
   * a gs_pick is bracketed by events on the stream it runs on (see
   * gs_set_pipeline for what one launch holds). */
  double corr_ms;              /* summed duration of the sampled launches (pass-1
                                  launches under GS_PIPE_LANES) */
  double corr_pass2_ms;        /* GS_PIPE_LANES: summed pass-2 launch durations */
  double corr_psi_ms;          /* the sampled psi's whole correlation, first launch
                                  start to last launch end */
  long long corr_launches;
  long long corr_sampled_maps; /* CC maps those launches covered (both passes) */
  long long corr_maps;         /* CC maps evaluated by the last gs_pick */
  long long refine_evals;      /* candidate evaluations (objective + gradient) of the
                                  last gs_refine */
} gs_timing;

/* Fill p with the reference's defaults (DataReader.h:16-42); required keys
 * are set to -1 like the reference. */
void gs_params_default(gs_params* p);

/* Create a context on `device`.  Replaces main.cu:1071 (cudaSetDevice),
 * the stream of main.cu:274 and the plans of main.cu:293-307. */
int gs_create(const gs_params* p, int device, void** ctx);

/* Upload and prepare the templates (replaces readAndPaddingTemplate,
 * edgeNormalize and handleTemplate, main.cu:182-474).  host: n images of
 * bx x by floats, x fastest.  tctf: CTF used for the |CTF| weighting --
 * the reference uses the first micrograph's values (main.cu:1106, 1133).
 * If bx or by exceeds window_size the window is enlarged to the next
 * multiple of 32 (adjustPaddingSize, main.cu:158-178).  Windows compiled:
 * every multiple of 32 from 64 to 1024, and 720; any other size returns
 * GS_ERR_UNSUPPORTED (the reference's grids assume a multiple of 32). */
int gs_set_templates(void* ctx, const float* host, int n, int bx, int by, const gs_ctf* tctf);
/* Templates of different boxes in one set: template j is by[j] x bx[j]
 * floats at host[j].  The reference reads every template with its own
 * header and centres it in the window by that size (main.cu:182-223); the
 * window grows to the largest box when needed (main.cu:158-178).  Everything
 * after the padding is as for gs_set_templates; gs_get_info reports the
 * common canvas as the box. */
int gs_set_templates_mixed(void* ctx, int n, const float* const* host, const int* bx,
                           const int* by, const gs_ctf* tctf);

/* Upload and prepare a micrograph (replaces init_d_image,
 * split_normalize_image and fft_on_img, main.cu:523-655, 758-769).
 * host: ny rows of nx floats (any size; the reference requires every
 * micrograph to have the first one's size, main.cu:1108-1109). */
int gs_set_micrograph(void* ctx, const float* host, int nx, int ny, const gs_ctf* c);
/* Same, with the image already in device memory (left unmodified). */
int gs_set_micrograph_device(void* ctx, const float* dev, int nx, int ny, const gs_ctf* c);

/* Upload and prepare a batch of micrographs of any sizes (each with its own
 * CTF); their tiles form one pool that gs_pick correlates in the same
 * launches.  images[i] points to ny[i] rows of nx[i] floats, in host memory
 * (on_device = 0) or device memory (on_device = 1).  No reference
 * counterpart: the reference processes one micrograph per loop iteration
 * (main.cu:1149-1269). */
int gs_set_micrographs(void* ctx, int n, const float* const* images, const int* nx, const int* ny,
                       const gs_ctf* ctfs, int on_device);

/* GS_MODE_FUSED or GS_MODE_PLAIN (norm_type 0; plain mode is the
 * stage-by-stage baseline). */
int gs_set_mode(void* ctx, int mode);
/* Capture the per-psi launch sequence of the fused mode in HIP graphs
 * (built on first use, replayed afterwards; 0 turns them off). */
int gs_set_graphs(void* ctx, int enable);

/* Run the in-plane steps [psi_begin, psi_end) of the current micrograph
 * (replaces the psi loop of main.cu:1230-1257 with pickPartcles_IMG_NORM,
 * main.cu:824-894, for norm_type 0; main.cu:1170-1229 for norm_type 1).
 * Picks above the threshold with a centre at least `diameter` from the
 * edges are returned in the reference's order: psi, then tile (y outer,
 * x inner), then view (then block for norm_type 1).  If cap is too small,
 * returns GS_ERR_CAPACITY with *n_out set to the required count; the picks
 * stay in the context and gs_get_picks returns them without recomputing.
 * psi_end < 0 means all steps. */
int gs_pick(void* ctx, int psi_begin, int psi_end, gs_particle* out, size_t cap, size_t* n_out);
/* The picks of the last gs_pick (no reference counterpart: the reference
 * writes them to the output file as it goes, main.cu:1255). */
int gs_get_picks(void* ctx, gs_particle* out, size_t cap, size_t* n_out);

/* Per-map results of the last norm_type 0 gs_pick, [psi][tile][view]. */
int gs_get_mapstats(void* ctx, gs_mapstat* out, size_t cap, size_t* n_out);

/* Intermediate arrays for stage-level parity checks. */
int gs_get_template_real(void* ctx, float* out, size_t cap);            /* n*l*l */
int gs_get_tile_spectra(void* ctx, float* out, size_t cap);             /* T*(l/2+1)*l complex, [t][kx][ky] */
int gs_get_rotated_spectra(void* ctx, int psi_index, float* out, size_t cap); /* n*(l/2+1)*l complex */

/* Kernel timing probe: prepares the rotated templates of psi_index, then
 * launches the two correlation kernels `reps` times each on the context
 * stream over the first batch of maps, timed with hipEvents.  out_ms[0],
 * out_ms[1] = mean duration of one pass-1 / pass-2 launch; *maps_per_launch
 * = CC maps one launch covers.  No reference counterpart (measurement). */
int gs_time_kernels(void* ctx, int psi_index, int reps, double* out_ms, int* maps_per_launch);

/* Refinement objective and its gradient for n candidates of the current
 * micrograph batch (no reference counterpart; the reference stops at the grid
 * search of main.cu:1230-1257).  The objective is the picker's normalised
 * CC continued to non-grid parameters: the view's whitened template spectrum
 * weighted with the candidate's defocus (apply_weighting_function,
 * GPU_func.cu:249-290) and normalised (main.cu:444-452), back to real space,
 * rotated by psi with bilinear interpolation (rotate_subIMG), masked and
 * normalised (main.cu:669-716), correlated with the tile spectrum
 * (compute_corner_CCG), and evaluated at the shifted centre by the Fourier
 * series of the CC map, divided by the map's standard deviation (Parseval).
 * At a grid psi, zero shift and the picking defocus it equals the picker's
 * score except that the standard deviation keeps the peak pixel.  The
 * gradient comes from a hand-written adjoint of every stage (FFTs, bilinear
 * rotation, normalisation, CTF weighting). */
int gs_adjoint(void* ctx, const gs_peak* peaks, int n, gs_grad* out);
/* Local refinement of n candidates in place: up to `iters` steps of
 * gradient ascent on the objective of gs_adjoint (step length adapted per
 * candidate from a parabola through the start and the trial: its maximum,
 * within 1-2x the step after an accepted trial and 0.1-0.5x after a
 * rejected one; psi within one phistep, the shift within 1.5 pixels and the
 * defocus within 0.5 um of the start); a candidate stops once its step is
 * below 0.05 degree / pixel (0.005 um).  out (may be null) receives the final objective and
 * gradient. */
int gs_refine(void* ctx, gs_peak* peaks, int n, int iters, gs_grad* out);

/* One line of a pick list as the duplicate removal reads it
 * (python_scripts/remove_repeat_particles_from_list.py:41-46,52-61): the
 * micrograph it belongs to, centre=, euler= and score= of the line. */
typedef struct gs_pick_rec {
  int group;             /* micrograph: a group is a run of equal values */
  int pad;
  double x, y;           /* centre (pixels) */
  double alt, az, phi;   /* Euler angles (degrees), phi = the printed psi */
  double score;
} gs_pick_rec;

/* Duplicate removal of a pick list on `device`; replaces
 * python_scripts/remove_repeat_particles_from_list.py:9-145 (same result).
 * Within each group, walking the list in order, a pick that is still kept
 * is compared with every later kept pick whose centre lies closer than
 * `center` pixels and whose relative orientation (quaternion angle) lies
 * within `euler_thres` degrees of 0 or of 180 (the script's folding); of
 * the two the lower score is dropped (the earlier one on a tie), and the
 * earlier pick keeps dropping later ones even once it has been dropped
 * itself.  keep[i] = 1 for the lines the script writes; *n_kept (may be
 * null) = their number.  Needs no context; allocates and frees its own
 * device memory. */
int gs_remove_repeats(int device, const gs_pick_rec* recs, size_t n, double center,
                      double euler_thres, unsigned char* keep, size_t* n_kept);

/* gs_remove_repeats without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (serving loops, harnesses
 * that forbid hipMalloc).  Same arguments and the same keep flags as
 * gs_remove_repeats; every device buffer of the call is a slice of one
 * static device array of GS_NOALLOC_SCRATCH_MIB MiB per device (held from
 * the library's first use of that device until the process ends), host
 * staging a static 12 MiB region, work on the calling thread's default
 * stream.  Limits: n <= GS_NOALLOC_MAX_PICKS picks and at most
 * GS_NOALLOC_MAX_UNDECIDED pairs whose orientation test needs the host's
 * acos; the neighbour lists must fit what the per-pick buffers leave of the
 * array (about 3.6 million duplicate pairs at the maximum n, more for
 * fewer picks).  Beyond a limit the call returns GS_ERR_CAPACITY and keep is
 * unspecified.  All calls on one device share its array, which keeps the
 * last call's contents (each call writes every part it reads first); calls
 * that overlapped on one device would race, so calls in one process are
 * serialised by a lock and return only after their device work is done. */
#define GS_NOALLOC_SCRATCH_MIB 32
#define GS_NOALLOC_MAX_PICKS 65536
#define GS_NOALLOC_MAX_UNDECIDED 65536
int gs_remove_repeats_noalloc(int device, const gs_pick_rec* recs, size_t n, double center,
                              double euler_thres, unsigned char* keep, size_t* n_kept);

/* Projections of a density map at Euler angles: the templates of a picking
 * run, made by EMAN's project3d in the reference's protocol (step 5; not
 * part of the reference's code).  vol: box^3 float32 [z][y][x]; euler:
 * nviews x (alt, az, phi) degrees, R = Rz(phi) Rx(alt) Rz(az) (EMAN);
 * out: nviews x box x box float32,
 *   out[j][row][col] = sum_w trilinear(vol, R_j^T (col-c, row-c, w) + c),
 * c = box/2, w = -c .. box-1-c, voxels outside the box 0 (the volume rotated
 * into its box, summed along z).  Needs no context; allocates and frees its
 * own device memory. */
int gs_project(int device, const float* vol, int box, const float* euler, int nviews,
               float* out);

/* Result of gs_autotune. */
typedef struct gs_tune {
  int pass1_variant, pass2_variant;  /* FFT schedule (radix order) of the two
                                        correlation passes */
  int batch;                         /* CC maps per launch batch */
  int pipeline;                      /* GS_PIPE_* */
  int candidates;                    /* launch schedules timed */
  int tile_block;                    /* tiles per block of the map order (0: the
                                        cache-sized default) */
  double psi_ms;                     /* correlation of one psi at the choice */
  double pass1_us[4], pass2_us[4];   /* each FFT variant alone, us per CC map
                                        (0: not compiled for this window) */
} gs_tune;

/* Choose the correlation schedule on the device (no counterpart in the
 * reference, whose cuFFT plans are fixed, main.cu:256-308): the FFT schedule
 * of each correlation pass among the compiled variants (radix orders; timed
 * alone on one batch of in-plane step psi_index; the tuned variant 0 stays
 * unless another is more than 2 % faster), then the tile block (the
 * one in use or all tiles, when the tiles span more than one block), the
 * batch size (half, the current, double) and, unless gs_set_pipeline chose
 * one, the pipeline (GS_PIPE_SERIAL, GS_PIPE_LANES, GS_PIPE_LAUNCH), timed
 * on the whole correlation of that step (best of two runs each; the
 * schedule in use stays unless another is more than 1 % faster).  Applies
 * the choice to the context and reports it in out (may be null).  Needs templates and a
 * micrograph; replaces the last gs_pick's map statistics (its picks are
 * kept).  All schedules give the same scores within FP32 rounding (the
 * pipelines bit-identical). */
int gs_autotune(void* ctx, int psi_index, gs_tune* out);
/* Set the FFT schedule variants (0..3) of the correlation passes directly
 * (0 = the tuned plan; gs_autotune's choices).  Windows compiled without
 * variants (all but 512 and 720) run variant 0. */
int gs_set_fft_variants(void* ctx, int pass1, int pass2);

int gs_get_info(void* ctx, gs_info* info);
int gs_get_timing(void* ctx, gs_timing* t);
/* Sample per-launch correlation times in gs_pick (see gs_timing). */
int gs_set_profiling(void* ctx, int enable);
/* Number of CC-map evaluations one gs_pick launch batch holds (tuning; 0
 * restores the defaults).  norm_type 0: default about 1.5 GiB of
 * intermediate per stream.  norm_type 1: batches are groups of
 * max(1, maps_per_batch / tiles) views over all tiles; default about 8 GiB
 * of intermediate per stream. */
int gs_set_batch(void* ctx, int maps_per_batch);
/* Correlation schedule (norm_type 0, fused mode):
 *   GS_PIPE_STREAM: one launch per psi; map m runs on XCD m % 8, which
 *     processes its maps in rounds, pass 2 of the map `lag` rounds back
 *     beside pass 1 of the current one, the intermediates in a ring of `ring`
 *     slots per XCD;
 *   GS_PIPE_LAUNCH: one launch per batch (gs_set_batch), pass 1 of batch k
 *     beside pass 2 of batch k - 1;
 *   GS_PIPE_LANES: separate pass-1 and pass-2 launches per batch on two
 *     streams, alternate batches per stream;
 *   GS_PIPE_SERIAL: the same launches on the context stream only (default;
 *     one intermediate; each launch runs alone, so its duration is its own).
 * lag, ring <= 0 select the defaults (2, 2 lag); ring must exceed lag.
 * All schedules give bit-identical results. */
enum { GS_PIPE_STREAM = 0, GS_PIPE_LAUNCH = 1, GS_PIPE_LANES = 2, GS_PIPE_SERIAL = 3 };
int gs_set_pipeline(void* ctx, int mode, int lag, int ring);
/* Tiles per block of the norm_type 0 processing order: all views of a block
 * before the next block.  0 (the default) takes as many tiles as fit in
 * 192 MiB, so the block's tile spectra stay in the last-level cache (96 at
 * window 720: one 5760 x 4092 micrograph is one block). */
int gs_set_tile_block(void* ctx, int tiles);
/* Initial capacity, in records, of the device-side pick buffers (maps that
 * pass the threshold for norm_type 0, 24 bytes each; 1024-pixel runs for
 * norm_type 1, 16 bytes each).  A gs_pick that finds more grows the buffer
 * to the exact count and selects again, so the capacity changes only memory
 * and time, never the picks.  0 restores the default (4,194,304). */
int gs_set_record_capacity(void* ctx, long long records);
int gs_destroy(void* ctx);
const char* gs_strerror(int code);

/* ---- Image files (host only; no device is touched) ------------------------
 * emdata::readImage (EMReader/emdata_.cpp:174-205) reads templates and
 * micrographs from HDF5 files in two layouts, or else from MRC files:
 *   - EMAN1 ("old-style", a root attribute num_dataset, EMReader/emhdf_.cpp:
 *     103-122): image i is the root dataset "i", nx = dims[0], ny = dims[1],
 *     nz = dims[2] (emdata_.cpp:216-250, emhdf_.cpp:124-150);
 *   - EMAN2 ("new-style"): image i is the group /MDF/images/i, its size the
 *     attributes EMAN.nx/ny/nz, its pixels the dataset "image"
 *     (emdata_.cpp:330-380, emhdf2_.cpp:245-363);
 *   - MRC modes 0, 1 and 2, either byte order (emdata_.cpp:384-617; the
 *     reference also rejects modes above 4, emdata_.cpp:457), plain or
 *     compressed as .gz (zlib) or .bz / .bz2 (the system's libbz2; the
 *     reference pipes these through bzcat, emdata_.cpp:403-413).
 * The HDF5 files are parsed without libhdf5: superblocks 0-3, symbol-table
 * and compact groups, contiguous / compact / chunked datasets (deflate,
 * shuffle, fletcher32), integer and IEEE float elements, converted to float.
 * Errors return GS_ERR_IO with the message in gs_io_error() (per thread). */
const char* gs_io_error(void);
/* Number of images in a file (MRC sections, HDF images). */
int gs_image_count(const char* path, int* n);
/* Image `index` (index < 0: the whole file -- an MRC stack or volume, or an
 * HDF stack of equal-size 2D images as nz sections) as float, x fastest.
 * Sizes are always set; out may be NULL to query them; a cap (in floats)
 * below nx*ny*nz returns GS_ERR_CAPACITY. */
int gs_read_image(const char* path, int index, float* out, size_t cap, int* nx, int* ny, int* nz);
/* Any HDF5 dataset by path ("/MDF/images/0/image"), elements converted to
 * float in storage order; *rank and dims[0..rank-1] (slowest first, at most
 * max_rank entries) are always set; out may be NULL to query them. */
int gs_hdf5_read(const char* path, const char* object, float* out, size_t cap, int* rank,
                 long long* dims, int max_rank);
/* An attribute of an HDF5 object: numeric values into num (*n_num of them,
 * GS_ERR_CAPACITY beyond cap), a string into str (NUL-terminated,
 * GS_ERR_CAPACITY if it does not fit).  *is_string tells which. */
int gs_hdf5_attr(const char* path, const char* object, const char* name, double* num, size_t cap,
                 size_t* n_num, char* str, size_t str_cap, int* is_string);

#ifdef __cplusplus
}
#endif
#endif /* GISSPA_H */
