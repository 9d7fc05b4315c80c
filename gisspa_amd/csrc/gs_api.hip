// gs_api.cpp -- context and C ABI of libgisspa (see include/gisspa.h).
//
// Host orchestration of the picking path.  Replaces the free functions of the
// reference's main.cu (Autumn1998/GisSPA) with one context per device:
// buffers are allocated once, every stage runs asynchronously on the
// context's stream, and the only host synchronisations are the final copy of
// the picks (the reference synchronises and reduces on the host once per tile
// per psi, main.cu:852-866).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/gisspa.h"
#include "gs_host.hpp"
#include "gs_kernels.hpp"

using namespace gs;

namespace {

const double kRefPiD = 3.141592654;  // GPU_func.cuh:29

// One micrograph of the current batch.
struct MicSlot {
  int nx = 0, ny = 0;
  gs_ctf ctf{};
  std::vector<int> offx, offy;
  int tile0 = 0, ntiles = 0;
};

// Whole-image phase-flip state for one micrograph size.
struct FlipCache {
  DevBuf twx, twy, rbin_start, rbin_pix, rbin_count, rbin_ra;
  int nbin = 0;
};

struct Ctx {
  gs_params p{};
  int device = 0;
  hipStream_t stream = nullptr;
  int L = 0, L_config = 0;
  Geometry geo{};
  DevBuf twL, twM;
  int mode = GS_MODE_FUSED;
  bool use_graphs = false;
  // templates
  int N = 0, bx = 0, by = 0;
  gs_ctf tctf{};
  DevBuf treal, tspec, rot_part, bscale;
  DevBuf tbox;                            // the template boxes as uploaded
  DevBuf bin_start, bin_pix, bin_count;   // radial bins of the half plane
  int bins_L = 0;
  DevBuf twhite;                          // whitened spectra (refinement)
  bool white_ok = false;
  DevBuf r_wtab;  // refinement weighting table, valid with white_ok
  // refinement batch buffers
  int ref_cap = 0;
  DevBuf r_cand, r_cs, r_A, r_gtmp, r_Tun, r_spec, r_rot, r_bscale, r_pre, r_gTb, r_gT, r_part,
      r_sc, r_out, r_thr, r_skip;
  double half_count = 0;  // pixels of the x <= l/2 half plane inside radius l/2
  // micrograph batch: tile pool of every micrograph, tile descriptors
  std::vector<MicSlot> mics;
  bool have_mic = false;
  int T = 0;
  DevBuf tiles, tdesc;
  DevBuf img, img_part, img_stats, img_mean, offx_d, offy_d, zero_mean;
  DevBuf cimg, ctmp, tile_part, tile_scale;
  std::vector<std::pair<long long, std::unique_ptr<FlipCache>>> flips;
  // correlation: two intermediates, so that one launch runs pass 1 of a
  // batch into one while pass 2 of the previous batch reads the other
  static constexpr int NLANE = 2;
  int batch = 0;
  DevBuf G[NLANE];
  // GS_PIPE_STREAM: one launch per psi; each XCD runs pass 2 of its map
  // `lag` rounds back beside pass 1 of its current map, intermediates in a
  // ring of `ring` slots per XCD
  // default GS_PIPE_SERIAL: a pass-1 launch fills the GPU, so the two-lane
  // overlap measured no faster than one stream (179.3 vs 179.4 ms per psi of
  // the full-size workload), and one stream needs one intermediate
  int pipeline = GS_PIPE_SERIAL, lag = 2, ring = 4;  // lag, ring: rounds of 8 maps
  // FFT schedule variant of each correlation pass (GS_PIPE_LANES and
  // GS_PIPE_SERIAL launches)
  int var1 = 0, var2 = 0;
  DevBuf Gring, counters, err;
  // GS_PIPE_LANES: two streams, each running pass 1 then pass 2 of every
  // other batch
  hipStream_t lane_stream[NLANE] = {nullptr, nullptr};
  hipEvent_t ev_ready = nullptr, ev_lane[NLANE] = {nullptr, nullptr};
  DevBuf part;  // row-block partials of every map of one psi
  DevBuf psum;  // column-block Parseval sums of every map of one psi
  DevBuf maps, stats, sel, sel_count;
  DevBuf maps_blk;
  int tile_block = 0;  // tiles per block of the map order; 0: as many as fit in 192 MiB
  int maps_T = -1, maps_N = -1, maps_tb = -1, maps_nv = -1;
  int n1_nv = 1;   // views per norm_type 1 batch (maps order)
  bool batch_set = false;   // gs_set_batch chose the batch size
  bool pipeline_set = false;  // gs_set_pipeline chose the schedule (gs_autotune keeps it)
  DevBuf n1_s1, n1_s2, n1_norm, n1_keys, n1_rec;  // norm_type 1
  // plain mode buffers
  DevBuf pl_tspec, pl_tiles, pl_ccg, pl_part, pl_rotpart, pl_mean, pl_var;
  // per-psi graphs of the fused correlation (gs_set_graphs)
  std::vector<hipGraphExec_t> graphs;
  std::vector<const void*> graph_key;  // buffers the graphs were captured with
  // launch sampling (gs_set_profiling): three events per batch of one psi
  bool profiling = false, profile_now = false;
  std::vector<hipEvent_t> prof_ev;
  int prof_batches = 0;
  std::vector<gs_particle> picks;  // of the last gs_pick
  // initial capacity (records) of the device-side pick buffers: selected
  // maps of norm_type 0 (SelRec, 24 B) and run records of norm_type 1
  // (N1Rec, 16 B); a run that selects more grows the buffer and selects again
  long long rec_cap0 = 1ll << 22;
  long long wait_ticks = 0;  // in-kernel dependency wait budget (wall-clock ticks)
  int stats_psi_begin = 0, stats_npsi = 0;
  gs_timing timing{};
};

int tile_offsets(int n, int l, int overlap, std::vector<int>& offs) {
  // main.cu:228-250
  const int interval = l - overlap;
  if (interval <= 0) return GS_ERR_ARG;
  offs.clear();
  int off = 0;
  offs.push_back(0);
  while (off + l < n) {
    off += interval;
    if (off + l >= n) off = n - l;
    offs.push_back(off);
  }
  return GS_OK;
}

float psi_value(double phi_step, int index) {
  float e = 0.0f;  // float accumulator, main.cu:1230
  for (int i = 0; i < index; ++i) e += (float)phi_step;
  return e;
}

int num_psi(double phi_step) {
  int n = 0;
  for (float e = 0.0f; e < 360.0f; e += (float)phi_step) n++;
  return n;
}

int upload_twiddles(Ctx* c) {
  const int L = c->L, M = L / 2;
  std::vector<float2> tl(L), tm(M);
  for (int i = 0; i < L; ++i) {
    const double a = -2.0 * M_PI * i / L;
    tl[i] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  for (int i = 0; i < M; ++i) {
    const double a = -2.0 * M_PI * i / M;
    tm[i] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  CHK(c->twL.ensure(L * sizeof(float2)));
  CHK(c->twM.ensure(M * sizeof(float2)));
  HIPCHK(hipMemcpy(c->twL.p, tl.data(), L * sizeof(float2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->twM.p, tm.data(), M * sizeof(float2), hipMemcpyHostToDevice));
  return GS_OK;
}

int set_window(Ctx* c, int L) {
  if (!geometry_supported(L)) return GS_ERR_UNSUPPORTED;
  if (L != c->L) {
    c->L = L;
    c->geo = geometry_of(L);
    CHK(upload_twiddles(c));
  }
  return GS_OK;
}

WeightArgs weight_args(const Ctx* c, const gs_ctf& ctf) {
  WeightArgs w{};
  const gs_params& p = c->p;
  w.apix = p.apix; w.highres = p.highres; w.lowres = p.lowres; w.kk = p.kk;
  w.ds = (float)(1.0 / (p.apix * (double)c->L_config));  // main.cu:155 (before adjustment)
  w.a = p.a; w.b = p.b; w.b2 = p.b2; w.bf = p.bfactor; w.bf2 = p.bfactor2; w.bf3 = p.bfactor3;
  const float defocus = -ctf.defocus;  // negated at parse, DataReader.cpp:378
  w.dfu = defocus + ctf.dfdiff;
  w.dfv = defocus - ctf.dfdiff;
  w.dfdiff = ctf.dfdiff;
  w.dfang = ctf.dfang;
  w.lambda = (float)(12.2639 / std::sqrt(p.energy * 1000.0 + 0.97845 * p.energy * p.energy));
  w.cs = p.cs;
  w.ampconst = p.ampconst;
  return w;
}

// Template half spectra [N][kx][ky] from the device copy of the boxes
// (readAndPaddingTemplate, edgeNormalize, handleTemplate: main.cu:182-474):
// pad, sigma-normalise, FFT / L^2, whiten by the radial amplitude, and with
// weight != 0 the low pass + |CTF| weighting and the amplitude normaliser.
int template_spectra(Ctx* c, float2* spec, int weight) {
  const int L = c->L, H = c->geo.H;
  const int n = c->N, bx = c->bx, by = c->by;
  hipStream_t s = c->stream;
  const long L2 = (long)L * L;
  DevBuf sums, scale, ra, wpart, inv_m;
  CHK(sums.ensure((size_t)n * 2 * sizeof(double)));
  CHK(scale.ensure((size_t)n * sizeof(float)));
  launch_box_sums(c->tbox.as<float>(), bx * by, n, sums.as<double>(), s);
  launch_sigma_scale(sums.as<double>(), n, L, scale.as<float>(), s);
  launch_rows_r2c_pad(L, c->tbox.as<float>(), bx, by, scale.as<float>(), n, spec,
                      c->twL.as<float2>(), c->twM.as<float2>(), s);
  launch_cols_fwd(L, spec, n, (float)(1.0 / (double)L2), nullptr, 0, c->twL.as<float2>(), s);
  // radial bins of the half plane x <= l/2 (SQRSum_by_circle, GPU_func.cu:98-128)
  const int NBIN = L / 2;
  if (c->bins_L != L) {
    std::vector<std::vector<int>> bins(NBIN);
    for (int kx = 0; kx <= L / 2; ++kx)
      for (int ky = 0; ky < L; ++ky) {
        const double dy = std::min(ky, L - ky), dx = std::min(kx, L - kx);
        const int r = (int)std::floor(std::hypot(dy, dx) + 0.5) - 1;
        if (r >= 0 && r < NBIN) bins[r].push_back(kx * L + ky);
      }
    std::vector<int> bstart(NBIN + 1, 0), bpix;
    std::vector<float> rcount(NBIN);
    double count_total = 0;
    for (int r = 0; r < NBIN; ++r) {
      bstart[r + 1] = bstart[r] + (int)bins[r].size();
      bpix.insert(bpix.end(), bins[r].begin(), bins[r].end());
      rcount[r] = (float)bins[r].size();
      count_total += bins[r].size();
    }
    CHK(c->bin_start.ensure(bstart.size() * sizeof(int)));
    CHK(c->bin_pix.ensure(std::max<size_t>(1, bpix.size()) * sizeof(int)));
    CHK(c->bin_count.ensure(NBIN * sizeof(float)));
    HIPCHK(hipMemcpyAsync(c->bin_start.p, bstart.data(), bstart.size() * sizeof(int),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->bin_pix.p, bpix.data(), bpix.size() * sizeof(int),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->bin_count.p, rcount.data(), NBIN * sizeof(float),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));  // host tables are released on return
    c->half_count = count_total;
    c->bins_L = L;
  }
  CHK(ra.ensure((size_t)n * NBIN * sizeof(double)));
  launch_radial_power(L, spec, n, c->bin_start.as<int>(), c->bin_pix.as<int>(), ra.as<double>(),
                      s);
  const long tot = (long)H * L;
  const int nwp = (int)((tot + 255) / 256);
  CHK(wpart.ensure((size_t)n * nwp * sizeof(double)));
  CHK(inv_m.ensure((size_t)n * sizeof(float)));
  launch_tmpl_weight(L, spec, n, ra.as<double>(), c->bin_count.as<float>(), weight_args(c, c->tctf),
                     wpart.as<double>(), weight, s);
  if (weight)
    launch_tmpl_ampnorm(L, spec, n, wpart.as<double>(), c->half_count, inv_m.as<float>(), s);
  HIPCHK(hipStreamSynchronize(s));  // the scratch buffers are released on return
  HIPCHK(hipGetLastError());
  return GS_OK;
}

int prep_templates(Ctx* c, const float* host, int n, int bx, int by) {
  const int L = c->L, H = c->geo.H, HP = c->geo.HP;
  hipStream_t s = c->stream;
  const long L2 = (long)L * L;
  // the boxes stay on the device (the refinement rebuilds the whitened spectra)
  CHK(c->tbox.ensure((size_t)n * bx * by * sizeof(float)));
  HIPCHK(hipMemcpyAsync(c->tbox.p, host, (size_t)n * bx * by * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  c->white_ok = false;
  CHK(c->tspec.ensure((size_t)n * H * L * sizeof(float2)));
  float2* spec = c->tspec.as<float2>();
  CHK(template_spectra(c, spec, 1));
  // inverse 2-D transform to the real-space templates, in chunks
  DevBuf gtmp;
  CHK(c->treal.ensure((size_t)n * L2 * sizeof(float)));
  const int chunk = std::max(1, std::min(n, (int)((512l << 20) / ((long)L * HP * 8))));
  CHK(gtmp.ensure((size_t)chunk * L * HP * sizeof(float2)));
  for (int b0 = 0; b0 < n; b0 += chunk) {
    const int nb = std::min(chunk, n - b0);
    launch_cols_inv_T(L, spec + (size_t)b0 * H * L, nb, gtmp.as<float2>(), c->twL.as<float2>(), s);
    launch_rows_c2r(L, gtmp.as<float2>(), nb, c->treal.as<float>() + (size_t)b0 * L2,
                    c->twL.as<float2>(), c->twM.as<float2>(), s);
  }
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  return GS_OK;
}

// whitened template half spectra for the refinement, built on first use
int ensure_white(Ctx* c) {
  if (c->white_ok) return GS_OK;
  CHK(c->twhite.ensure((size_t)c->N * c->geo.H * c->L * sizeof(float2)));
  CHK(template_spectra(c, c->twhite.as<float2>(), 0));
  CHK(c->r_wtab.ensure((size_t)c->geo.H * c->L * sizeof(float4)));
  launch_ref_wtable(c->L, weight_args(c, c->tctf), c->r_wtab.as<float4>(), c->stream);
  HIPCHK(hipGetLastError());
  c->white_ok = true;
  return GS_OK;
}

int upload_table(DevBuf& buf, int n, hipStream_t s) {
  std::vector<float2> t(n);
  for (int i = 0; i < n; ++i) {
    const double a = -2.0 * M_PI * i / n;
    t[i] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  CHK(buf.ensure(n * sizeof(float2)));
  HIPCHK(hipMemcpyAsync(buf.p, t.data(), n * sizeof(float2), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return GS_OK;
}

FlipCache* flip_cache(Ctx* c, int nx, int ny) {
  const long long key = ((long long)nx << 32) | (unsigned)ny;
  for (auto& e : c->flips)
    if (e.first == key) return e.second.get();
  std::unique_ptr<FlipCache> fc(new FlipCache());
  hipStream_t s = c->stream;
  if (upload_table(fc->twx, nx, s) || upload_table(fc->twy, ny, s)) return nullptr;
  // SQRSum_by_circle mode 1 bins (GPU_func.cu:113-126): x <= nx/2,
  // 0 <= r < max(nx, ny)/2, r = floor(hypot(min(y,ny-y), min(x,nx-x)) + .5) - 1
  const int nbin = std::max(nx, ny) / 2;
  std::vector<std::vector<long>> bins(nbin);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x <= nx / 2; ++x) {
      const double dx = std::min(x, nx - x), dy = std::min(y, ny - y);
      const int r = (int)std::floor(std::hypot(dy, dx) + 0.5) - 1;
      if (r >= 0 && r < nbin) bins[r].push_back((long)y * nx + x);
    }
  std::vector<int> start(nbin + 1, 0);
  std::vector<long> pix;
  std::vector<float> cnt(nbin);
  for (int r = 0; r < nbin; ++r) {
    start[r + 1] = start[r] + (int)bins[r].size();
    pix.insert(pix.end(), bins[r].begin(), bins[r].end());
    cnt[r] = (float)bins[r].size();
  }
  if (fc->rbin_start.ensure(start.size() * sizeof(int)) ||
      fc->rbin_pix.ensure(std::max<size_t>(1, pix.size()) * sizeof(long)) ||
      fc->rbin_count.ensure(nbin * sizeof(float)) || fc->rbin_ra.ensure(nbin * sizeof(double)))
    return nullptr;
  if (hipMemcpyAsync(fc->rbin_start.p, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(fc->rbin_pix.p, pix.data(), pix.size() * sizeof(long), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(fc->rbin_count.p, cnt.data(), nbin * sizeof(float), hipMemcpyHostToDevice, s) ||
      hipStreamSynchronize(s))
    return nullptr;
  fc->nbin = nbin;
  c->flips.emplace_back(key, std::move(fc));
  return c->flips.back().second.get();
}

// Whole-micrograph phase flip, whitening, weighting and band pass, DC := 0
// (init_d_image with phase_flip = 1, main.cu:548-586), in place on c->img.
int phase_flip_image(Ctx* c, int nx, int ny, const gs_ctf& ctf) {
  hipStream_t s = c->stream;
  const long n = (long)nx * ny;
  CHK(c->cimg.ensure(n * sizeof(float2)));
  CHK(c->ctmp.ensure(n * sizeof(float2)));
  FlipCache* fc = flip_cache(c, nx, ny);
  if (!fc) return GS_ERR_DEVICE;
  const WeightArgs w = weight_args(c, ctf);
  // FFT of the image scaled by 1/n (cuFFT is unnormalised; main.cu:553-554)
  launch_real_to_complex(c->img.as<float>(), c->cimg.as<float2>(), n, (float)(1.0 / (double)n), s);
  rt_fft_2d(c->cimg.as<float2>(), c->ctmp.as<float2>(), nx, ny, fc->twx.as<float2>(),
            fc->twy.as<float2>(), -1, s);
  launch_phase_flip(c->cimg.as<float2>(), nx, ny, w, s);
  launch_img_radial(c->cimg.as<float2>(), fc->rbin_start.as<int>(), fc->rbin_pix.as<long>(),
                    fc->nbin, fc->rbin_ra.as<double>(), s);
  launch_img_weight(c->cimg.as<float2>(), nx, ny, fc->rbin_ra.as<double>(),
                    fc->rbin_count.as<float>(), w, s);
  rt_fft_2d(c->cimg.as<float2>(), c->ctmp.as<float2>(), nx, ny, fc->twx.as<float2>(),
            fc->twy.as<float2>(), +1, s);
  launch_complex_to_real(c->cimg.as<float2>(), c->img.as<float>(), n, s);
  return GS_OK;
}

// Prepare micrograph `mi` of the batch into its slice of the tile pool
// (init_d_image, split_normalize_image, fft_on_img: main.cu:523-655, 758-769).
int prep_micrograph(Ctx* c, int mi, const float* src, bool src_is_device) {
  hipStream_t s = c->stream;
  MicSlot& ms = c->mics[mi];
  const int nx = ms.nx, ny = ms.ny;
  const long n = (long)nx * ny;
  const int L = c->L;
  CHK(c->img.ensure(n * sizeof(float)));
  HIPCHK(hipMemcpyAsync(c->img.p, src, n * sizeof(float),
                        src_is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  const int nblk = img_reduce_blocks(n);
  CHK(c->img_part.ensure((size_t)nblk * 2 * sizeof(double)));
  CHK(c->img_stats.ensure(8 * sizeof(double)));
  CHK(c->img_mean.ensure(sizeof(double)));
  launch_img_sums(c->img.as<float>(), n, c->img_part.as<double>(), nblk, s);
  launch_img_bounds(c->img_part.as<double>(), nblk, n, c->img_stats.as<double>(), s);
  launch_img_clamp(c->img.as<float>(), n, c->p.invert, c->img_stats.as<double>(),
                   c->img_part.as<double>(), nblk, s);
  launch_img_mean(c->img_part.as<double>(), nblk, n, c->img_mean.as<double>(), s);
  const double* mean = c->img_mean.as<double>();
  if (c->p.phase_flip == 1) {
    CHK(phase_flip_image(c, nx, ny, ms.ctf));
    // the filtered image already has zero DC; split_IMG copies it as is
    if (!c->zero_mean.p) {
      CHK(c->zero_mean.ensure(sizeof(double)));
      HIPCHK(hipMemsetAsync(c->zero_mean.p, 0, sizeof(double), s));
    }
    mean = c->zero_mean.as<double>();
  }
  CHK(c->offx_d.ensure(ms.offx.size() * sizeof(int)));
  CHK(c->offy_d.ensure(ms.offy.size() * sizeof(int)));
  HIPCHK(hipMemcpyAsync(c->offx_d.p, ms.offx.data(), ms.offx.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->offy_d.p, ms.offy.data(), ms.offy.size() * sizeof(int), hipMemcpyHostToDevice, s));
  const int H = c->geo.H;
  float2* pool = c->tiles.as<float2>() + (size_t)ms.tile0 * H * L;
  launch_rows_r2c_tiles(L, c->img.as<float>(), nx, ny, c->offx_d.as<int>(), c->offy_d.as<int>(),
                        (int)ms.offx.size(), ms.ntiles, mean, pool, c->twL.as<float2>(),
                        c->twM.as<float2>(), s);
  launch_cols_fwd(L, pool, ms.ntiles, (float)(1.0 / ((double)L * L)), nullptr, 0,
                  c->twL.as<float2>(), s);
  if (c->p.phase_flip == 1) {
    // per-tile amplitude normalisation (main.cu:617-653)
    CHK(c->tile_part.ensure((size_t)ms.ntiles * tile_ampnorm_parts(L) * sizeof(double)));
    CHK(c->tile_scale.ensure((size_t)ms.ntiles * sizeof(float)));
    launch_tile_ampnorm(pool, L, ms.ntiles, c->half_count, c->tile_part.as<double>(),
                        c->tile_scale.as<float>(), s);
  }
  // the correlation's (-1)^(kx+ky) factor rides on the tile spectra
  launch_checker_sign(L, pool, ms.ntiles, s);
  if (c->mode == GS_MODE_PLAIN) {
    // plain mode keeps full complex row-major tile spectra (fft_on_img)
    const long L2 = (long)L * L;
    CHK(c->pl_tiles.ensure((size_t)c->T * L2 * sizeof(float2)));
    float2* pt = c->pl_tiles.as<float2>() + (size_t)ms.tile0 * L2;
    launch_plain_tiles(L, c->img.as<float>(), nx, ny, c->tdesc.as<TileDesc>(), ms.tile0,
                       ms.ntiles, mean, pt, s);
    launch_c2c_2d(L, pt, ms.ntiles, -1, (float)(1.0 / (double)L2), c->twL.as<float2>(), s);
  }
  // the caller may release the source image on return
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  return GS_OK;
}

// Lay out the tile pool of a batch and prepare every micrograph.
int set_batch(Ctx* c, int n, const float* const* images, const int* nx, const int* ny,
              const gs_ctf* ctfs, bool on_device) {
  const int L = c->L;
  std::vector<MicSlot> mics(n);
  std::vector<TileDesc> desc;
  int T = 0;
  for (int i = 0; i < n; ++i) {
    if (!images[i] || nx[i] <= 0 || ny[i] <= 0) return GS_ERR_ARG;
    MicSlot& ms = mics[i];
    ms.nx = nx[i];
    ms.ny = ny[i];
    ms.ctf = ctfs[i];
    CHK(tile_offsets(nx[i], L, c->p.overlap, ms.offx));
    CHK(tile_offsets(ny[i], L, c->p.overlap, ms.offy));
    ms.tile0 = T;
    ms.ntiles = (int)(ms.offx.size() * ms.offy.size());
    for (int iy = 0; iy < (int)ms.offy.size(); ++iy)
      for (int ix = 0; ix < (int)ms.offx.size(); ++ix)
        desc.push_back(TileDesc{i, ms.offx[ix], ms.offy[iy], nx[i], ny[i]});
    T += ms.ntiles;
  }
  c->mics = std::move(mics);
  c->T = T;
  c->have_mic = false;
  const int H = c->geo.H;
  CHK(c->tiles.ensure((size_t)T * H * L * sizeof(float2)));
  CHK(c->tdesc.ensure(desc.size() * sizeof(TileDesc)));
  HIPCHK(hipMemcpyAsync(c->tdesc.p, desc.data(), desc.size() * sizeof(TileDesc), hipMemcpyHostToDevice, c->stream));
  for (int i = 0; i < n; ++i) CHK(prep_micrograph(c, i, images[i], on_device));
  c->have_mic = true;
  return GS_OK;
}

int copy_out(Ctx* c, void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GS_OK;
}

void psi_rotation(const Ctx* c, int psi_index, float* cose, float* sine) {
  const float e = psi_value(c->p.phi_step, psi_index);
  // rotate_subIMG computes cos/sin in double and stores float (GPU_func.cu:498-499)
  *cose = (float)std::cos((double)e * kRefPiD / 180);
  *sine = (float)std::sin((double)e * kRefPiD / 180);
}

int rotate_templates(Ctx* c, int psi_index) {
  const int L = c->L;
  float cose, sine;
  psi_rotation(c, psi_index, &cose, &sine);
  const float dmask = 1.5f * c->p.diameter;
  // the masked templates are zero outside a band of rows around the centre:
  // only the band is transformed by rows, and the columns read only it
  const RowBand band = rot_row_band(L, dmask);
  CHK(c->rot_part.ensure((size_t)c->N * band.nrb * 2 * sizeof(double)));
  CHK(c->bscale.ensure((size_t)c->N * sizeof(float)));
  launch_rows_r2c_rot(L, c->treal.as<float>(), c->N, cose, sine, dmask, band,
                      c->tspec.as<float2>(), c->rot_part.as<double>(), c->twL.as<float2>(),
                      c->twM.as<float2>(), c->stream);
  launch_rot_scale(L, c->rot_part.as<double>(), c->N, band.nrb, c->bscale.as<float>(), c->stream);
  launch_cols_fwd(L, c->tspec.as<float2>(), c->N, 1.0f, c->bscale.as<float>(), 1,
                  c->twL.as<float2>(), c->stream, band);
  return GS_OK;
}

void drop_graphs(Ctx* c) {
  for (auto g : c->graphs)
    if (g) (void)hipGraphExecDestroy(g);
  c->graphs.clear();
  c->graph_key.clear();
}

// Profiling event i of the sampled psi, created on first use.
int prof_record(Ctx* c, size_t i, hipStream_t s) {
  while (c->prof_ev.size() <= i) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c->prof_ev.push_back(e);
  }
  HIPCHK(hipEventRecord(c->prof_ev[i], s));
  return GS_OK;
}

// a device buffer of at least n bytes
bool fits(const DevBuf& b, size_t n) { return b.p && b.bytes >= n; }

int ensure_corr_buffers(Ctx* c) {
  const int L = c->L;
  const long per_map = c->geo.gmap * (long)sizeof(float2);
  // default batch: about 1.5 GiB of intermediate per stream (759 maps at
  // L = 720); measured at the full workload, 189 maps 270, 384 286.6, 768 288.0
  // micrographs/hour (fewer launch boundaries where the two streams drain)
  if (c->batch <= 0)
    c->batch = (int)std::max(1l, std::min((long)c->T * c->N, (1536l << 20) / per_map));
  const int nb = std::min(c->batch, c->T * c->N);
  const int nrb = (L + c->geo.RB - 1) / c->geo.RB;
  // G[0] serves norm_type 1 and the kernel timing in every schedule; G[1]
  // the launch-per-batch and two-lane schedules
  CHK(c->G[0].ensure((size_t)nb * per_map));
  if (c->pipeline == GS_PIPE_STREAM) {
    CHK(c->Gring.ensure((size_t)8 * c->ring * per_map));
    CHK(c->counters.ensure((size_t)2 * c->T * c->N * sizeof(int)));
  } else if (c->pipeline != GS_PIPE_SERIAL) {
    CHK(c->G[1].ensure((size_t)nb * per_map));
  }
  CHK(c->err.ensure(sizeof(int)));
  CHK(c->part.ensure((size_t)c->T * c->N * nrb * sizeof(CorrPartial)));
  CHK(c->psum.ensure((size_t)c->T * c->N * corr_ncb(L) * sizeof(float2)));
  // Processing orders; results land in the canonical [tile][view] order
  // through the MapRef of each map.
  //  * maps (norm_type 1): groups of nv views (nv = batch / T), tile-major
  //    within a group (maps t * nv .. t * nv + nv - 1 are the group's views of
  //    tile t); one group is one batch;
  //  * maps_blk (norm_type 0): blocks of tb tiles, all views of a block
  //    before the next block, tiles fastest.  Within a block each template
  //    half spectrum is read from HBM once and stays in L2 across the block's
  //    tiles, while the block's tile spectra are re-read for every view from
  //    the 256 MB last-level cache, so the default block is as many tiles as
  //    fit in 192 MiB (96 at L = 720: a whole 5760 x 4092 micrograph, 88
  //    tiles, is one block; a batch of 16 micrographs, 1078 tiles, runs 2.09
  //    instead of 2.62 s per psi compared with one block of all tiles).
  const size_t nm = (size_t)c->T * c->N;
  const long tile_bytes = (long)(L / 2 + 1) * L * (long)sizeof(float2);
  const int tb = c->tile_block > 0 ? c->tile_block
                                   : (int)std::max(1l, (192l << 20) / tile_bytes);
  // norm_type 1 batches: nv views of every tile; by default about 8 GiB of
  // intermediate per stream (46 views at L = 720), so that the per-batch
  // flush of the FP64 accumulators is shared by many maps (206 views of the
  // full-size micrograph: 23 / 46 / 69 / 103 views per batch measured
  // 104.0 / 106.7 / 98.6 / 92.6 micrographs/hour)
  const long nb1 = c->batch_set ? nb : std::max(1l, (8192l << 20) / per_map);
  const int nv1 = (int)std::max(1l, std::min((long)c->N, nb1 / c->T));
  if (c->maps_T != c->T || c->maps_N != c->N || c->maps_tb != tb || c->maps_nv != nv1) {
    std::vector<MapRef> mr, mb;
    mr.reserve(nm);
    mb.reserve(nm);
    for (int j0 = 0; j0 < c->N; j0 += nv1)
      for (int t = 0; t < c->T; ++t)
        for (int j = j0; j < std::min(c->N, j0 + nv1); ++j) mr.push_back(MapRef{t, j});
    for (int t0 = 0; t0 < c->T; t0 += tb)
      for (int j = 0; j < c->N; ++j)
        for (int t = t0; t < std::min(c->T, t0 + tb); ++t) mb.push_back(MapRef{t, j});
    CHK(c->maps.ensure(nm * sizeof(MapRef)));
    CHK(c->maps_blk.ensure(nm * sizeof(MapRef)));
    HIPCHK(hipMemcpyAsync(c->maps.p, mr.data(), nm * sizeof(MapRef), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->maps_blk.p, mb.data(), nm * sizeof(MapRef), hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));  // mr, mb are released on return
    c->maps_T = c->T;
    c->maps_N = c->N;
    c->maps_tb = tb;
    c->maps_nv = nv1;
  }
  c->n1_nv = nv1;
  return GS_OK;
}

// Correlate all maps of one psi (the templates of this psi are in tspec):
// launch k runs pass 1 of batch k and pass 2 of batch k - 1 on the context
// stream (the stream order completes each intermediate before it is read);
// stats[tile][view] <- scores.
int correlate_psi(Ctx* c, MapStat* out) {
  const int L = c->L;
  const int nm = c->T * c->N;
  const int nb = std::min(c->batch, nm);
  if (nm <= 0 || nb <= 0) return GS_ERR_STATE;
  const int nbatch = (nm + nb - 1) / nb;
  const float lo = c->p.diameter / 4.0f;
  const float hi = (float)L - c->p.diameter / 4.0f;
  const int ncb = corr_ncb(L), np = corr_nparts(L);
  // every buffer the launches below address, at the size they address
  // (ensure_corr_buffers allocates what the schedule uses)
  const size_t per_map = (size_t)c->geo.gmap * sizeof(float2);
  const size_t half = (size_t)(L / 2 + 1) * L * sizeof(float2);
  if (!fits(c->tiles, (size_t)c->T * half) || !fits(c->tspec, (size_t)c->N * half) ||
      !fits(c->maps_blk, (size_t)nm * sizeof(MapRef)) ||
      !fits(c->part, (size_t)nm * np * sizeof(CorrPartial)) ||
      !fits(c->psum, (size_t)nm * ncb * sizeof(float2)) || !fits(c->err, sizeof(int)) ||
      !out)
    return GS_ERR_STATE;
  const bool serial = c->pipeline == GS_PIPE_SERIAL;
  if (c->pipeline == GS_PIPE_STREAM ? (!fits(c->Gring, (size_t)8 * c->ring * per_map) ||
                                       !fits(c->counters, (size_t)2 * nm * sizeof(int)))
                                    : (!fits(c->G[0], (size_t)nb * per_map) ||
                                       (!serial && !fits(c->G[1], (size_t)nb * per_map))))
    return GS_ERR_STATE;
  int ylo, yhi;  // rows of the intermediate the peak search reads
  corr_store_rows(L, lo, hi, &ylo, &yhi);
  hipStream_t s = c->stream;
  if (c->pipeline == GS_PIPE_STREAM) {
    if (c->profile_now) CHK(prof_record(c, 0, s));
    launch_corr_stream(L, c->tiles.as<float2>(), c->tspec.as<float2>(), c->maps_blk.as<MapRef>(),
                       nm, c->lag, c->ring, c->Gring.as<float2>(), c->psum.as<float2>(),
                       c->counters.as<int>(), lo, hi, c->part.as<CorrPartial>(),
                       c->err.as<int>(), c->wait_ticks, c->twL.as<float2>(), c->twM.as<float2>(),
                       s);
    if (c->profile_now) {
      CHK(prof_record(c, 1, s));
      c->prof_batches = 1;
    }
    launch_corr_finalize(L, c->part.as<CorrPartial>(), c->psum.as<float2>(), nm,
                         c->maps_blk.as<MapRef>(), c->N, out, s);
    return GS_OK;
  }
  if (c->pipeline == GS_PIPE_LANES || serial) {
    if (c->profile_now) CHK(prof_record(c, 3 * (size_t)nbatch, s));
    // the lanes start after the rotated templates of this psi are ready
    if (!serial) {
      HIPCHK(hipEventRecord(c->ev_ready, s));
      for (int i = 0; i < Ctx::NLANE; ++i)
        HIPCHK(hipStreamWaitEvent(c->lane_stream[i], c->ev_ready, 0));
    }
    for (int k = 0; k < nbatch; ++k) {
      const int m0 = k * nb, cnt = std::min(nb, nm - m0);
      const int lane = serial ? 0 : k % 2;
      hipStream_t ls = serial ? s : c->lane_stream[lane];
      if (c->profile_now) CHK(prof_record(c, 3 * (size_t)k, ls));
      launch_corr_pass1(L, c->tiles.as<float2>(), c->tspec.as<float2>(),
                        c->maps_blk.as<MapRef>() + m0, cnt, ylo, yhi, c->G[lane].as<float2>(),
                        c->psum.as<float2>() + (size_t)m0 * ncb, c->twL.as<float2>(), ls,
                        c->var1);
      if (c->profile_now) CHK(prof_record(c, 3 * (size_t)k + 1, ls));
      launch_corr_pass2(L, c->G[lane].as<float2>(), cnt, lo, hi,
                        c->part.as<CorrPartial>() + (size_t)m0 * np, c->twL.as<float2>(),
                        c->twM.as<float2>(), ls, c->var2);
      if (c->profile_now) CHK(prof_record(c, 3 * (size_t)k + 2, ls));
    }
    if (c->profile_now) c->prof_batches = nbatch;
    if (!serial) {
      for (int i = 0; i < Ctx::NLANE; ++i) {
        HIPCHK(hipEventRecord(c->ev_lane[i], c->lane_stream[i]));
        HIPCHK(hipStreamWaitEvent(s, c->ev_lane[i], 0));
      }
    }
    if (c->profile_now) CHK(prof_record(c, 3 * (size_t)nbatch + 1, s));
    launch_corr_finalize(L, c->part.as<CorrPartial>(), c->psum.as<float2>(), nm,
                         c->maps_blk.as<MapRef>(), c->N, out, s);
    return GS_OK;
  }
  for (int k = 0; k <= nbatch; ++k) {
    const int m1 = k * nb, cnt1 = k < nbatch ? std::min(nb, nm - m1) : 0;
    const int m2 = (k - 1) * nb, cnt2 = k >= 1 ? std::min(nb, nm - m2) : 0;
    if (c->profile_now) CHK(prof_record(c, 2 * (size_t)k, s));
    launch_corr_fused(L, c->tiles.as<float2>(), c->tspec.as<float2>(),
                      c->maps_blk.as<MapRef>() + (cnt1 ? m1 : 0), cnt1, c->G[k % 2].as<float2>(),
                      c->psum.as<float2>() + (size_t)(cnt1 ? m1 : 0) * ncb,
                      c->G[(k + 1) % 2].as<float2>(), cnt2, lo, hi,
                      c->part.as<CorrPartial>() + (size_t)(cnt2 ? m2 : 0) * np,
                      c->twL.as<float2>(), c->twM.as<float2>(), s);
    if (c->profile_now) CHK(prof_record(c, 2 * (size_t)k + 1, s));
  }
  if (c->profile_now) c->prof_batches = nbatch + 1;
  launch_corr_finalize(L, c->part.as<CorrPartial>(), c->psum.as<float2>(), nm,
                       c->maps_blk.as<MapRef>(), c->N, out, s);
  return GS_OK;
}

// One psi of the fused path: rotation + correlation, replayed from a
// captured graph when graphs are on.  The graph of a psi is keyed on every
// buffer it touches; any reallocation drops all graphs.
int fused_psi(Ctx* c, int ip, MapStat* out) {
  if (!c->use_graphs || c->profile_now) {
    CHK(rotate_templates(c, ip));
    return correlate_psi(c, out);
  }
  const std::vector<const void*> key = {c->tiles.p, c->tspec.p, c->treal.p, c->G[0].p, c->G[1].p,
                                        c->Gring.p, c->counters.p, c->err.p,
                                        (const void*)(long)c->pipeline, (const void*)(long)c->lag,
                                        (const void*)(long)c->ring,
                                        c->part.p, c->psum.p, c->maps_blk.p, c->rot_part.p,
                                        c->bscale.p, out, (const void*)(long)c->T,
                                        (const void*)(long)c->batch};
  const int P = num_psi(c->p.phi_step);
  if ((int)c->graphs.size() != P) {
    drop_graphs(c);
    c->graphs.assign(P, nullptr);
    c->graph_key.assign((size_t)P * key.size(), nullptr);
  }
  const void** gk = c->graph_key.data() + (size_t)ip * key.size();
  if (!c->graphs[ip] || !std::equal(key.begin(), key.end(), gk)) {
    if (c->graphs[ip]) (void)hipGraphExecDestroy(c->graphs[ip]);
    c->graphs[ip] = nullptr;
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = rotate_templates(c, ip);
    if (r == GS_OK) r = correlate_psi(c, out);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != GS_OK) return r;
    HIPCHK(e);
    hipGraphExec_t ge = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ie);
    c->graphs[ip] = ge;
    std::copy(key.begin(), key.end(), gk);
  }
  HIPCHK(hipGraphLaunch(c->graphs[ip], c->stream));
  return GS_OK;
}

// norm_type 1 (main.cu:1170-1229): pass 1 accumulates per-pixel sums of c/L^2
// and its square over every view x psi, pass 2 turns each map into z-scores
// and reports 1024-pixel runs whose maximum reaches the threshold.  Batches
// are groups of nv views over all tiles (tile-major, see ensure_corr_buffers):
// the accumulation sums each tile's nv maps in registers before touching the
// FP64 accumulators; the second pass reduces the z-scores to per-run maxima
// inside the row transform (no z-maps), and the maps of one tile are
// neighbours, so the tile's mean / sd maps are read from cache.  One stream;
// every accumulator sees a fixed order of additions.
int pick_norm1(Ctx* c, int psi_begin, int psi_end, std::vector<gs_particle>& picks) {
  const int L = c->L;
  const long L2 = (long)L * L;
  const int P = num_psi(c->p.phi_step);
  const int nm = c->T * c->N;
  const int nv1 = c->n1_nv;
  const long per_map = c->geo.gmap * (long)sizeof(float2);
  const int nb = c->T * nv1;   // maps of a full batch
  const int ncb = corr_ncb(L);
  const size_t npix = (size_t)c->T * L2;
  const long kb = norm1_key_bytes(L);
  if (nm <= 0 || nb <= 0) return GS_ERR_STATE;
  for (int i = 0; i < Ctx::NLANE; ++i) CHK(c->G[i].ensure((size_t)nb * per_map));
  CHK(c->psum.ensure((size_t)Ctx::NLANE * nb * ncb * sizeof(float2)));
  CHK(c->n1_s1.ensure(npix * sizeof(double)));
  CHK(c->n1_s2.ensure(npix * sizeof(double)));
  CHK(c->n1_norm.ensure(npix * 2 * sizeof(float)));
  CHK(c->n1_keys.ensure((size_t)Ctx::NLANE * nb * kb));
  CHK(c->n1_rec.ensure((size_t)std::max(1ll, c->rec_cap0) * sizeof(N1Rec)));
  CHK(c->sel_count.ensure(sizeof(unsigned long long)));
  const size_t half = (size_t)(L / 2 + 1) * L * sizeof(float2);
  if (!fits(c->maps, (size_t)nm * sizeof(MapRef)) || !fits(c->tiles, (size_t)c->T * half) ||
      !fits(c->tspec, (size_t)c->N * half) || !fits(c->tdesc, (size_t)c->T * sizeof(TileDesc)))
    return GS_ERR_STATE;
  hipStream_t s = c->stream;
  HIPCHK(hipMemsetAsync(c->n1_s1.p, 0, npix * sizeof(double), s));
  HIPCHK(hipMemsetAsync(c->n1_s2.p, 0, npix * sizeof(double), s));
  const MapRef* maps = c->maps.as<MapRef>();
  // Batches alternate between the two lane streams, each with its own
  // intermediate, partial sums and run keys, so that pass 1 of one batch runs
  // beside the row kernel of the previous one.  The accumulations run in
  // batch order: each waits for the previous batch's (ev_lane of the other
  // lane), so every FP64 accumulator sees the same order of additions as on
  // one stream.  The next psi's rotation waits for both lanes.
  auto lanes_start = [&]() -> int {
    HIPCHK(hipEventRecord(c->ev_ready, s));
    for (int i = 0; i < Ctx::NLANE; ++i)
      HIPCHK(hipStreamWaitEvent(c->lane_stream[i], c->ev_ready, 0));
    return GS_OK;
  };
  auto lanes_join = [&]() -> int {
    for (int i = 0; i < Ctx::NLANE; ++i) {
      HIPCHK(hipEventRecord(c->ev_lane[i], c->lane_stream[i]));
      HIPCHK(hipStreamWaitEvent(s, c->ev_lane[i], 0));
    }
    return GS_OK;
  };
  auto pass1 = [&](int m0, int cnt, int lane) {
    launch_corr_pass1(L, c->tiles.as<float2>(), c->tspec.as<float2>(), maps + m0, cnt, 0, L - 1,
                      c->G[lane].as<float2>(), c->psum.as<float2>() + (size_t)lane * nb * ncb,
                      c->twL.as<float2>(), c->lane_stream[lane]);
  };
  for (int ip = 0; ip < P; ++ip) {
    CHK(rotate_templates(c, ip));
    CHK(lanes_start());
    for (int m0 = 0, k = 0; m0 < nm; m0 += nb, ++k) {
      const int cnt = std::min(nb, nm - m0), lane = k % 2;
      hipStream_t ls = c->lane_stream[lane];
      pass1(m0, cnt, lane);
      if (k > 0) HIPCHK(hipStreamWaitEvent(ls, c->ev_lane[1 - lane], 0));
      launch_norm1_acc(L, c->G[lane].as<float2>(), maps + m0, c->T, cnt / c->T,
                       c->n1_s1.as<double>(), c->n1_s2.as<double>(), c->twM.as<float2>(), ls);
      HIPCHK(hipEventRecord(c->ev_lane[lane], ls));
    }
    CHK(lanes_join());
  }
  launch_norm1_stats(c->n1_s1.as<double>(), c->n1_s2.as<double>(), (long)npix, (double)c->N * P,
                     c->n1_norm.p, s);
  // second pass: every run record is counted, the first `cap` are written;
  // if more were found the record buffer grows and the pass runs again
  auto pass2 = [&](long long cap) -> int {
    HIPCHK(hipMemsetAsync(c->sel_count.p, 0, sizeof(unsigned long long), s));
    for (int ip = psi_begin; ip < psi_end; ++ip) {
      CHK(rotate_templates(c, ip));
      CHK(lanes_start());
      for (int m0 = 0, k = 0; m0 < nm; m0 += nb, ++k) {
        const int cnt = std::min(nb, nm - m0), lane = k % 2;
        hipStream_t ls = c->lane_stream[lane];
        void* keys = static_cast<char*>(c->n1_keys.p) + (size_t)lane * nb * kb;
        pass1(m0, cnt, lane);
        launch_norm1_zrun(L, c->G[lane].as<float2>(), maps + m0, cnt, c->n1_norm.p, keys,
                          c->twM.as<float2>(), ls);
        launch_norm1_runfin(L, keys, cnt, (long long)(ip - psi_begin) * nm, c->T, maps + m0,
                            c->tdesc.as<TileDesc>(), c->p.threshold, c->p.diameter,
                            c->sel_count.as<unsigned long long>(), c->n1_rec.p, cap, ls);
      }
      CHK(lanes_join());
    }
    return GS_OK;
  };
  auto read_count = [&](unsigned long long* n) -> int {
    HIPCHK(hipMemcpyAsync(n, c->sel_count.p, sizeof(*n), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(hipGetLastError());
    return GS_OK;
  };
  long long cap = (long long)(c->n1_rec.bytes / sizeof(N1Rec));
  unsigned long long count = 0;
  CHK(pass2(cap));
  CHK(read_count(&count));
  if (count > (unsigned long long)cap) {
    cap = (long long)count;
    CHK(c->n1_rec.ensure((size_t)cap * sizeof(N1Rec)));
    CHK(pass2(cap));
    unsigned long long again = 0;
    CHK(read_count(&again));
    if (again != count) return GS_ERR_DEVICE;  // the second pass is deterministic
  }
  std::vector<N1Rec> recs(count);
  if (count) CHK(copy_out(c, recs.data(), c->n1_rec.p, (size_t)count * sizeof(N1Rec)));
  // canonical order: micrograph, psi, tile, view, run (main.cu:784-819);
  // r.m = psi offset * nm + view * T + tile, the run is pos / 1024
  const long long nblk = (L2 + 1023) / 1024;
  auto key = [&](const N1Rec& r) {
    const long long ipl = r.m / nm, mm = r.m % nm;
    const long long tile = mm % c->T, view = mm / c->T;
    return ((ipl * c->T + tile) * c->N + view) * nblk + r.pos / 1024;
  };
  std::sort(recs.begin(), recs.end(), [&](const N1Rec& a, const N1Rec& b) { return key(a) < key(b); });
  picks.resize(count);
  for (size_t i = 0; i < count; ++i) {
    const N1Rec& r = recs[i];
    const int ip = (int)(r.m / nm) + psi_begin, mm = (int)(r.m % nm);
    const int t = mm % c->T;
    int mi = 0;
    while (mi + 1 < (int)c->mics.size() && c->mics[mi + 1].tile0 <= t) ++mi;
    const MicSlot& ms = c->mics[mi];
    const int lt = t - ms.tile0;
    gs_particle& pk = picks[i];
    pk.psi_index = ip;
    pk.tile = lt;
    pk.view = mm / c->T;
    pk.mic = mi;
    float cur_e = 360.0f - psi_value(c->p.phi_step, ip);
    if (cur_e >= 360) cur_e -= 360;
    pk.psi = cur_e;
    pk.cx = (float)(ms.offx[lt % ms.offx.size()] + r.pos % L);
    pk.cy = (float)(ms.offy[lt / ms.offx.size()] + r.pos / L);
    pk.score = r.z;
  }
  std::stable_sort(picks.begin(), picks.end(),
                   [](const gs_particle& a, const gs_particle& b) { return a.mic < b.mic; });
  return GS_OK;
}

// Stage-by-stage picker (main.cu:669-716, 824-894): per psi, rotation with a
// host-side mean/variance, a full complex FFT of every template; per tile,
// product, full complex inverse FFT, per-1024-element partials copied to the
// host and reduced there.  Scores land in stats[tile][view] like the fused
// path's.
int plain_psi(Ctx* c, int ip, MapStat* out) {
  const int L = c->L;
  const long L2 = (long)L * L;
  const int N = c->N, T = c->T;
  hipStream_t s = c->stream;
  const int nblk = plain_blocks(L);
  CHK(c->pl_tspec.ensure((size_t)N * L2 * sizeof(float2)));
  CHK(c->pl_ccg.ensure((size_t)N * L2 * sizeof(float2)));
  CHK(c->pl_rotpart.ensure((size_t)N * nblk * 2 * sizeof(double)));
  CHK(c->pl_part.ensure((size_t)N * nblk * sizeof(PeakPartHost)));
  CHK(c->pl_mean.ensure((size_t)N * sizeof(float)));
  CHK(c->pl_var.ensure((size_t)N * sizeof(float)));
  float cose, sine;
  psi_rotation(c, ip, &cose, &sine);
  launch_plain_rotate(L, c->treal.as<float>(), N, cose, sine, 1.5f * c->p.diameter,
                      c->pl_tspec.as<float2>(), c->pl_rotpart.as<double>(), s);
  std::vector<double> rp((size_t)N * nblk * 2);
  CHK(copy_out(c, rp.data(), c->pl_rotpart.p, rp.size() * sizeof(double)));
  std::vector<float> mean(N), var(N);
  for (int j = 0; j < N; ++j) {  // main.cu:688-699
    double s1 = 0, s2 = 0;
    for (int b = 0; b < nblk; ++b) { s1 += rp[2 * ((size_t)j * nblk + b)]; s2 += rp[2 * ((size_t)j * nblk + b) + 1]; }
    const double m = s1 / L2;
    mean[j] = (float)m;
    var[j] = (float)(s2 / L2 - m * m);
  }
  HIPCHK(hipMemcpyAsync(c->pl_mean.p, mean.data(), N * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->pl_var.p, var.data(), N * sizeof(float), hipMemcpyHostToDevice, s));
  launch_plain_normalize(L, c->pl_tspec.as<float2>(), N, c->pl_mean.as<float>(),
                         c->pl_var.as<float>(), s);
  launch_c2c_2d(L, c->pl_tspec.as<float2>(), N, -1, (float)(1.0 / (double)L2), c->twL.as<float2>(), s);
  const float sgn = ((L / 2) % 2 == 1) ? -1.0f : 1.0f;
  const float lo = c->p.diameter / 4.0f, hi = (float)L - c->p.diameter / 4.0f;
  std::vector<PeakPartHost> pp((size_t)N * nblk);
  for (int t = 0; t < T; ++t) {
    launch_plain_ccg(L, c->pl_tiles.as<float2>() + (size_t)t * L2, c->pl_tspec.as<float2>(), N,
                     sgn, c->pl_ccg.as<float2>(), s);
    launch_c2c_2d(L, c->pl_ccg.as<float2>(), N, +1, 1.0f, c->twL.as<float2>(), s);
    launch_plain_peak(L, c->pl_ccg.as<float2>(), N, lo, hi, c->pl_part.p, s);
    CHK(copy_out(c, pp.data(), c->pl_part.p, pp.size() * sizeof(PeakPartHost)));
    std::vector<MapStat> st(N);
    for (int j = 0; j < N; ++j) {  // main.cu:856-881, FP64 sums
      float peak = 0.0f;
      int pos = 0;
      double s1 = 0, s2 = 0;
      for (int b = 0; b < nblk; ++b) {
        const PeakPartHost& q = pp[(size_t)j * nblk + b];
        if (q.peak > peak || (q.peak == peak && peak > 0 && q.pos < pos)) { peak = q.peak; pos = q.pos; }
        s1 += q.s1;
        s2 += q.s2;
      }
      const double rc = (double)L2 - 1, ra = s1 - peak, rb = s2 - (double)peak * peak;
      const double sd = std::sqrt(rb / rc - (ra / rc) * (ra / rc));
      MapStat& m = st[j];
      m.peak = peak;
      m.pos = pos;
      m.sd = (float)sd;
      m.score = (sd > 0) ? (float)(peak / sd) : 0.0f;
      m.sum = s1;
      m.sum2 = s2;
    }
    HIPCHK(hipMemcpyAsync(out + (size_t)t * N, st.data(), N * sizeof(MapStat), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return GS_OK;
}

// ---------------------------------------------------------------------------
// refinement
// ---------------------------------------------------------------------------
// candidates per forward/backward launch sequence: as many as about 16 GiB
// of work buffers hold, 64 to 2048 (12.5 MB each at window 720: 1374).  On
// the top 10,000 picks of the full-size micrograph (168,271 evaluations)
// 64 / 128 / 256 / 512 / 1024 ran 60.3 / 62.9 / 66.5 / 68.6 / 69.5 k
// evaluations per second, and on the top 7,616 picks of 24 in-plane steps
// 512 / 1024 / 2048 took 0.797 / 0.785 / 0.777 s; more candidates per launch
// fill the GPU better and pay the per-kernel ramp and tail less often
int ref_batch(const Ctx* c) {
  const size_t L = c->L, H = c->geo.H, HP = c->geo.HP;
  const size_t per = 2 * H * L * sizeof(float2) + L * HP * sizeof(float2) +
                     3 * L * L * sizeof(float);
  return (int)std::min<size_t>(2048, std::max<size_t>(64, (16ull << 30) / per));
}

// work buffers of one batch of nb candidates (the per-candidate inputs and
// outputs of a whole call are sized by ref_eval)
int ref_ensure(Ctx* c, int nb) {
  if (c->ref_cap >= nb) return GS_OK;
  const size_t L = c->L, H = c->geo.H, HP = c->geo.HP;
  const size_t nrb = (L + c->geo.RB - 1) / c->geo.RB;
  CHK(c->r_A.ensure(nb * H * L * sizeof(float2)));
  CHK(c->r_gtmp.ensure(nb * L * HP * sizeof(float2)));
  CHK(c->r_Tun.ensure(nb * L * L * sizeof(float)));
  CHK(c->r_spec.ensure(nb * H * L * sizeof(float2)));
  CHK(c->r_rot.ensure(nb * nrb * 2 * sizeof(double)));
  CHK(c->r_bscale.ensure(nb * sizeof(float)));
  CHK(c->r_pre.ensure(nb * sizeof(float)));
  CHK(c->r_gTb.ensure(nb * L * L * sizeof(float)));
  CHK(c->r_gT.ensure(nb * L * L * sizeof(float)));
  CHK(c->r_part.ensure(nb * (size_t)ref_parts() * 4 * sizeof(double)));
  CHK(c->r_sc.ensure(nb * sizeof(RefScal)));
  CHK(c->r_skip.ensure(nb * sizeof(int)));
  c->ref_cap = nb;
  return GS_OK;
}

// candidates of the current batch in kernel form (checks every index)
int ref_cands(const Ctx* c, const gs_peak* pk, int n, std::vector<RefCand>& rc,
              std::vector<double2>& cs) {
  rc.resize(n);
  cs.resize(n);
  for (int i = 0; i < n; ++i) {
    const gs_peak& p = pk[i];
    if (p.mic < 0 || p.mic >= (int)c->mics.size() || p.view < 0 || p.view >= c->N) return GS_ERR_ARG;
    const MicSlot& ms = c->mics[p.mic];
    if (p.tile < 0 || p.tile >= ms.ntiles) return GS_ERR_ARG;
    const int bxn = (int)ms.offx.size();
    const float x0 = p.cx - (float)ms.offx[p.tile % bxn], y0 = p.cy - (float)ms.offy[p.tile / bxn];
    if (!(x0 >= 0 && x0 < c->L && y0 >= 0 && y0 < c->L)) return GS_ERR_ARG;
    if (!std::isfinite(p.psi) || !std::isfinite(p.dx) || !std::isfinite(p.dy) ||
        !std::isfinite(p.defocus))
      return GS_ERR_ARG;
    rc[i] = RefCand{ms.tile0 + p.tile, p.view, x0, y0, p.psi, p.dx, p.dy, p.defocus};
    const double a = (double)p.psi * kRefPiD / 180;  // rotate_subIMG's angle (GPU_func.cu:498-499)
    cs[i] = make_double2(std::cos(a), std::sin(a));
  }
  return GS_OK;
}

// objective and gradient of n candidates: out[5 i ..] = score, d psi, d dx,
// d dy, d defocus
// thr (may be null): a candidate whose score does not exceed thr[i] gets its
// score only (zero gradient); the refinement rejects such trial steps and
// never reads their gradient, so its adjoint is skipped on the device
int ref_eval(Ctx* c, const RefCand* rc, const double2* cs, int n, float* out,
             const float* thr = nullptr) {
  CHK(ensure_white(c));
  hipStream_t s = c->stream;
  const int rb = ref_batch(c);
  const int nb0 = std::min(n, rb);
  CHK(ref_ensure(c, nb0));
  // every candidate's inputs go over once and every output comes back once;
  // the batches run back to back on the stream (no host round trip between
  // them), each reusing the work buffers after the previous one in stream
  // order
  CHK(c->r_cand.ensure((size_t)n * sizeof(RefCand)));
  CHK(c->r_cs.ensure((size_t)n * sizeof(double2)));
  CHK(c->r_out.ensure((size_t)n * 5 * sizeof(float)));
  HIPCHK(hipMemcpyAsync(c->r_cand.p, rc, (size_t)n * sizeof(RefCand), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->r_cs.p, cs, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
  if (thr) {
    CHK(c->r_thr.ensure((size_t)n * sizeof(float)));
    HIPCHK(hipMemcpyAsync(c->r_thr.p, thr, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
  }
  for (int b0 = 0; b0 < n; b0 += rb) {
    const int nb = std::min(rb, n - b0);
    RefArgs a{};
    a.nb = nb;
    a.white = c->twhite.as<float2>();
    a.tiles = c->tiles.as<float2>();
    a.cand = c->r_cand.as<RefCand>() + b0;
    a.w = weight_args(c, c->tctf);
    a.wtab = c->r_wtab.as<float4>();
    a.count = c->half_count;
    a.dmask = 1.5f * c->p.diameter;
    a.cs = c->r_cs.as<double2>() + b0;
    a.A = c->r_A.as<float2>();
    a.gtmp = c->r_gtmp.as<float2>();
    a.Tun = c->r_Tun.as<float>();
    a.spec = c->r_spec.as<float2>();
    a.rot_part = c->r_rot.as<double>();
    a.bscale = c->r_bscale.as<float>();
    a.pre = c->r_pre.as<float>();
    a.gTb = c->r_gTb.as<float>();
    a.gT = c->r_gT.as<float>();
    a.part = c->r_part.as<double>();
    a.sc = c->r_sc.as<RefScal>();
    a.out = c->r_out.as<float>() + 5 * (size_t)b0;
    a.thr = thr ? c->r_thr.as<float>() + b0 : nullptr;
    a.skip = c->r_skip.as<int>();
    a.twL = c->twL.as<float2>();
    a.twM = c->twM.as<float2>();
    launch_ref_forward(c->L, a, s);
    launch_ref_backward(c->L, a, s);
  }
  CHK(copy_out(c, out, c->r_out.p, (size_t)n * 5 * sizeof(float)));  // synchronises
  HIPCHK(hipGetLastError());
  return GS_OK;
}

// Map the pool tile index t to (micrograph, tile within it).
void tile_owner(const Ctx* c, int t, int* mic, int* local) {
  int lo = 0, hi = (int)c->mics.size() - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (c->mics[mid].tile0 <= t) lo = mid; else hi = mid - 1;
  }
  *mic = lo;
  *local = t - c->mics[lo].tile0;
}

}  // namespace

extern "C" {

void gs_params_default(gs_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->apix = -1; p->phi_step = -1; p->kk = -1; p->energy = -1; p->cs = -1;
  p->highres = -1; p->lowres = -1; p->diameter = -1;
  p->threshold = 7; p->window_size = 320; p->phase_flip = 0; p->overlap = -1;
  p->norm_type = 0; p->invert = 0; p->ampconst = 0.07f;
  p->a = -9.32f; p->b = 2.65f; p->b2 = 0.01908f;
  p->bfactor = -78.7757f; p->bfactor2 = -12.9121f; p->bfactor3 = 1.28732f;
}

const char* gs_strerror(int code) {
  switch (code) {
    case GS_OK: return "ok";
    case GS_ERR_ARG: return "invalid argument";
    case GS_ERR_UNSUPPORTED: return "unsupported configuration";
    case GS_ERR_DEVICE: return "HIP runtime error";
    case GS_ERR_NOMEM: return "device memory allocation failed";
    case GS_ERR_STATE: return "templates or micrograph not set";
    case GS_ERR_CAPACITY: return "output buffer too small";
    case GS_ERR_IO: return "file missing, unreadable or in an unsupported format";
    default: return "unknown error";
  }
}

int gs_create(const gs_params* p, int device, void** out) {
  if (!p || !out) return GS_ERR_ARG;
  *out = nullptr;
  if (p->apix <= 0 || p->phi_step <= 0 || p->window_size <= 0 || p->diameter < 0 ||
      p->highres <= 0 || p->lowres <= 0)
    return GS_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    return GS_ERR_DEVICE;
  }
  HIPCHK(hipSetDevice(device));
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return GS_ERR_NOMEM;
  c->p = *p;
  c->device = device;
  c->L_config = p->window_size;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess;
  for (int i = 0; ok && i < Ctx::NLANE; ++i)
    ok = hipStreamCreateWithFlags(&c->lane_stream[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&c->ev_lane[i], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&c->ev_ready, hipEventDisableTiming) == hipSuccess;
  // in-kernel dependency waits give up after 2 s of the device's wall clock
  int rate_khz = 0;
  if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess ||
      rate_khz <= 0) {
    (void)hipGetLastError();
    rate_khz = 100000;  // the 100 MHz constant clock of CDNA parts
  }
  c->wait_ticks = 2000ll * rate_khz;
  if (!ok) {
    (void)hipGetLastError();
    delete c;
    return GS_ERR_DEVICE;
  }
  *out = c;
  return GS_OK;
}

int gs_destroy(void* ctx) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return GS_ERR_ARG;
  (void)hipSetDevice(c->device);
  (void)hipDeviceSynchronize();
  drop_graphs(c);
  for (auto e : c->prof_ev) (void)hipEventDestroy(e);
  for (int i = 0; i < Ctx::NLANE; ++i) {
    if (c->lane_stream[i]) (void)hipStreamDestroy(c->lane_stream[i]);
    if (c->ev_lane[i]) (void)hipEventDestroy(c->ev_lane[i]);
  }
  if (c->ev_ready) (void)hipEventDestroy(c->ev_ready);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return GS_OK;
}

int gs_set_templates(void* ctx, const float* host, int n, int bx, int by, const gs_ctf* tctf) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !host || n <= 0 || bx <= 0 || by <= 0 || !tctf) return GS_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  int L = c->p.window_size;
  if (L < bx || L < by) {  // adjustPaddingSize, main.cu:158-178
    L = 0;
    while (L < bx || L < by) L += 32;
  }
  CHK(set_window(c, L));
  drop_graphs(c);
  c->ref_cap = 0;  // refinement buffers follow the window
  c->N = n; c->bx = bx; c->by = by; c->tctf = *tctf;
  c->have_mic = false;  // tiles depend on the window
  return prep_templates(c, host, n, bx, by);
}

int gs_set_templates_mixed(void* ctx, int n, const float* const* host, const int* bx,
                           const int* by, const gs_ctf* tctf) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !host || !bx || !by || n <= 0 || !tctf) return GS_ERR_ARG;
  int B = 0;
  for (int j = 0; j < n; ++j) {
    if (!host[j] || bx[j] <= 0 || by[j] <= 0) return GS_ERR_ARG;
    B = std::max(B, std::max(bx[j], by[j]));
  }
  int L = c->p.window_size;
  if (L < B) {  // adjustPaddingSize (main.cu:158-178) on the largest box
    L = 0;
    while (L < B) L += 32;
  }
  // Canvas B with L - B even: then centring a template in B and B in L puts
  // it at ((L - bx) / 2, (L - by) / 2), the offset readAndPaddingTemplate
  // gives each template from its own header (main.cu:206-207).
  if ((L - B) % 2) ++B;
  std::vector<float> canvas((size_t)n * B * B, 0.0f);
  for (int j = 0; j < n; ++j) {
    const int ox = (B - bx[j]) / 2, oy = (B - by[j]) / 2;
    for (int y = 0; y < by[j]; ++y)
      std::copy(host[j] + (size_t)y * bx[j], host[j] + (size_t)(y + 1) * bx[j],
                canvas.begin() + ((size_t)j * B + oy + y) * B + ox);
  }
  return gs_set_templates(ctx, canvas.data(), n, B, B, tctf);
}

static int set_mics(Ctx* c, int n, const float* const* images, const int* nx, const int* ny,
                    const gs_ctf* ctfs, bool dev) {
  if (!c || n <= 0 || !images || !nx || !ny || !ctfs) return GS_ERR_ARG;
  if (c->N == 0) return GS_ERR_STATE;
  if (c->p.phase_flip != 0 && c->p.phase_flip != 1) return GS_ERR_UNSUPPORTED;
  if (c->mode == GS_MODE_PLAIN && c->p.phase_flip != 0) return GS_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());  // graphs and lanes may still read the pool
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  const int r = set_batch(c, n, images, nx, ny, ctfs, dev);
  float ms = 0;
  if (r == GS_OK && hipEventRecord(e1, c->stream) == hipSuccess &&
      hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
    c->timing.micrograph_prep_ms = ms;
  (void)hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return r;
}

int gs_set_micrograph(void* ctx, const float* host, int nx, int ny, const gs_ctf* ctf) {
  return set_mics(static_cast<Ctx*>(ctx), 1, &host, &nx, &ny, ctf, false);
}
int gs_set_micrograph_device(void* ctx, const float* dev, int nx, int ny, const gs_ctf* ctf) {
  return set_mics(static_cast<Ctx*>(ctx), 1, &dev, &nx, &ny, ctf, true);
}
int gs_set_micrographs(void* ctx, int n, const float* const* images, const int* nx, const int* ny,
                       const gs_ctf* ctfs, int on_device) {
  return set_mics(static_cast<Ctx*>(ctx), n, images, nx, ny, ctfs, on_device != 0);
}

int gs_set_mode(void* ctx, int mode) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || (mode != GS_MODE_FUSED && mode != GS_MODE_PLAIN)) return GS_ERR_ARG;
  if (mode == GS_MODE_PLAIN && (c->p.norm_type != 0 || c->p.phase_flip != 0))
    return GS_ERR_UNSUPPORTED;
  if (mode != c->mode) {
    c->mode = mode;
    c->have_mic = false;  // the two modes keep different tile spectra
  }
  return GS_OK;
}

int gs_set_profiling(void* ctx, int enable) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return GS_ERR_ARG;
  c->profiling = enable != 0;
  return GS_OK;
}

int gs_set_graphs(void* ctx, int enable) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return GS_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  c->use_graphs = enable != 0;
  if (!c->use_graphs) drop_graphs(c);
  return GS_OK;
}

int gs_set_pipeline(void* ctx, int mode, int lag, int ring) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || mode < GS_PIPE_STREAM || mode > GS_PIPE_SERIAL) return GS_ERR_ARG;
  if (lag <= 0) lag = 2;
  if (ring <= 0) ring = 2 * lag;
  if (ring <= lag) return GS_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  c->pipeline = mode;
  c->pipeline_set = true;
  c->lag = lag;
  c->ring = ring;
  drop_graphs(c);
  c->Gring.release();
  return GS_OK;
}

int gs_set_tile_block(void* ctx, int tiles) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || tiles < 0) return GS_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  c->tile_block = tiles;  // 0: the cache-sized default
  drop_graphs(c);
  return GS_OK;
}

int gs_set_record_capacity(void* ctx, long long records) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || records < 0) return GS_ERR_ARG;
  c->rec_cap0 = records ? records : (1ll << 22);
  return GS_OK;
}

int gs_set_batch(void* ctx, int maps_per_batch) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || maps_per_batch < 0) return GS_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipDeviceSynchronize());
  c->batch = maps_per_batch;
  c->batch_set = maps_per_batch > 0;
  drop_graphs(c);
  for (int i = 0; i < Ctx::NLANE; ++i) c->G[i].release();
  return GS_OK;
}

int gs_pick(void* ctx, int psi_begin, int psi_end, gs_particle* out, size_t cap, size_t* n_out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !n_out) return GS_ERR_ARG;
  *n_out = 0;
  c->picks.clear();  // a failed pick leaves no stale list for gs_get_picks
  if (c->N == 0 || !c->have_mic) return GS_ERR_STATE;
  if (c->p.norm_type != 0 && c->p.norm_type != 1) return GS_ERR_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->device));
  const int P = num_psi(c->p.phi_step);
  if (psi_end < 0 || psi_end > P) psi_end = P;
  if (psi_begin < 0 || psi_begin > psi_end) return GS_ERR_ARG;
  if (c->p.norm_type == 1) {
    CHK(ensure_corr_buffers(c));
    c->stats_npsi = 0;
    CHK(pick_norm1(c, psi_begin, psi_end, c->picks));
    return gs_get_picks(ctx, out, cap, n_out);
  }
  const int npsi = psi_end - psi_begin;
  const long long per_psi = (long long)c->T * c->N;
  CHK(ensure_corr_buffers(c));
  CHK(c->stats.ensure((size_t)std::max<long long>(1, npsi * per_psi) * sizeof(MapStat)));
  c->stats_psi_begin = psi_begin;
  c->stats_npsi = npsi;
  c->timing.corr_maps = npsi * per_psi;
  c->timing.corr_ms = 0;
  c->timing.corr_pass2_ms = 0;
  c->timing.corr_launches = 0;
  c->timing.corr_sampled_maps = 0;
  c->prof_batches = 0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  launch_fill_int(c->err.as<int>(), 1, 0, c->stream);
  for (int ip = psi_begin; ip < psi_end; ++ip) {
    MapStat* out = c->stats.as<MapStat>() + (size_t)(ip - psi_begin) * per_psi;
    c->profile_now = c->profiling && ip == psi_begin && c->mode == GS_MODE_FUSED;
    const int r = (c->mode == GS_MODE_PLAIN) ? plain_psi(c, ip, out) : fused_psi(c, ip, out);
    c->profile_now = false;
    CHK(r);
  }
  // device-side selection, then canonical order on the host: every selected
  // map is counted (64-bit), the first `cap` records are written; a larger
  // count grows the record buffer and selects again
  const long long nmaps = npsi * per_psi;
  CHK(c->sel.ensure((size_t)std::max(1ll, std::min(nmaps, c->rec_cap0)) * sizeof(SelRec)));
  CHK(c->sel_count.ensure(sizeof(unsigned long long)));
  auto select = [&](long long cap) -> int {
    HIPCHK(hipMemsetAsync(c->sel_count.p, 0, sizeof(unsigned long long), c->stream));
    launch_select(c->stats.as<MapStat>(), nmaps, c->T, c->N, c->L, c->tdesc.as<TileDesc>(),
                  c->p.threshold, c->p.diameter, c->sel_count.as<unsigned long long>(),
                  c->sel.as<SelRec>(), cap, c->stream);
    return GS_OK;
  };
  long long capsel = (long long)(c->sel.bytes / sizeof(SelRec));
  CHK(select(capsel));
  HIPCHK(hipEventRecord(e1, c->stream));
  unsigned long long count = 0;
  int err = 0;
  HIPCHK(hipMemcpyAsync(&count, c->sel_count.p, sizeof(count), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&err, c->err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipGetLastError());
  if (err) return GS_ERR_DEVICE;  // an in-kernel dependency wait ran out
  if (count > (unsigned long long)capsel) {
    if (count > (unsigned long long)nmaps) return GS_ERR_DEVICE;  // one record per map at most
    capsel = (long long)count;
    CHK(c->sel.ensure((size_t)capsel * sizeof(SelRec)));
    CHK(select(capsel));
    unsigned long long again = 0;
    HIPCHK(hipMemcpyAsync(&again, c->sel_count.p, sizeof(again), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (again != count) return GS_ERR_DEVICE;
  }
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  c->timing.pick_total_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const int evs = (c->pipeline == GS_PIPE_LANES || c->pipeline == GS_PIPE_SERIAL) ? 3 : 2;
  for (int ib = 0; ib < c->prof_batches; ++ib) {
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, c->prof_ev[evs * ib], c->prof_ev[evs * ib + 1]));
    c->timing.corr_ms += a;
    if (evs == 3) {
      HIPCHK(hipEventElapsedTime(&b, c->prof_ev[evs * ib + 1], c->prof_ev[evs * ib + 2]));
      c->timing.corr_pass2_ms += b;
    }
  }
  c->timing.corr_launches = c->prof_batches;
  if (c->prof_batches && evs == 3) {
    float w = 0;
    HIPCHK(hipEventElapsedTime(&w, c->prof_ev[3 * c->prof_batches],
                               c->prof_ev[3 * c->prof_batches + 1]));
    c->timing.corr_psi_ms = w;
  } else {
    c->timing.corr_psi_ms = c->timing.corr_ms;
  }
  if (c->prof_batches) c->timing.corr_sampled_maps = per_psi;
  std::vector<SelRec> recs(count);
  if (count)
    CHK(copy_out(c, recs.data(), c->sel.p, (size_t)count * sizeof(SelRec)));
  c->picks.resize(count);
  for (size_t i = 0; i < count; ++i) {
    const long long m = recs[i].m;
    const int ip = (int)(m / per_psi) + psi_begin;
    const long long rem = m % per_psi;
    gs_particle& pk = c->picks[i];
    pk.psi_index = ip;
    tile_owner(c, (int)(rem / c->N), &pk.mic, &pk.tile);
    pk.view = (int)(rem % c->N);
    float cur_e = 360.0f - psi_value(c->p.phi_step, ip);  // main.cu:1253-1254
    if (cur_e >= 360) cur_e -= 360;
    pk.psi = cur_e;
    pk.cx = recs[i].cx;
    pk.cy = recs[i].cy;
    pk.score = recs[i].score;
  }
  // one micrograph after the other, each in the reference's psi, tile, view order
  std::vector<long long> order(count);
  for (size_t i = 0; i < count; ++i) order[i] = (long long)i;
  std::sort(order.begin(), order.end(), [&](long long a, long long b) {
    if (c->picks[a].mic != c->picks[b].mic) return c->picks[a].mic < c->picks[b].mic;
    return recs[a].m < recs[b].m;
  });
  std::vector<gs_particle> sorted(count);
  for (size_t i = 0; i < count; ++i) sorted[i] = c->picks[order[i]];
  c->picks.swap(sorted);
  return gs_get_picks(ctx, out, cap, n_out);
}

int gs_get_picks(void* ctx, gs_particle* out, size_t cap, size_t* n_out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !n_out) return GS_ERR_ARG;
  *n_out = c->picks.size();
  if (c->picks.size() > cap || (!c->picks.empty() && !out)) return GS_ERR_CAPACITY;
  if (!c->picks.empty()) std::memcpy(out, c->picks.data(), c->picks.size() * sizeof(gs_particle));
  return GS_OK;
}

int gs_get_mapstats(void* ctx, gs_mapstat* out, size_t cap, size_t* n_out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !n_out) return GS_ERR_ARG;
  const size_t n = (size_t)c->stats_npsi * c->T * c->N;
  *n_out = n;
  if (cap < n || (n && !out)) return GS_ERR_CAPACITY;
  static_assert(sizeof(gs_mapstat) == sizeof(MapStat), "layout");
  HIPCHK(hipSetDevice(c->device));
  if (n) CHK(copy_out(c, out, c->stats.p, n * sizeof(MapStat)));
  return GS_OK;
}

int gs_get_template_real(void* ctx, float* out, size_t cap) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !out) return GS_ERR_ARG;
  const size_t n = (size_t)c->N * c->L * c->L;
  if (c->N == 0) return GS_ERR_STATE;
  if (cap < n) return GS_ERR_CAPACITY;
  HIPCHK(hipSetDevice(c->device));
  CHK(copy_out(c, out, c->treal.p, n * sizeof(float)));
  return GS_OK;
}

int gs_get_tile_spectra(void* ctx, float* out, size_t cap) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !out) return GS_ERR_ARG;
  if (!c->have_mic) return GS_ERR_STATE;
  const size_t n = (size_t)c->T * c->geo.H * c->L * 2;
  if (cap < n) return GS_ERR_CAPACITY;
  HIPCHK(hipSetDevice(c->device));
  CHK(copy_out(c, out, c->tiles.p, n * sizeof(float)));
  // undo the correlation sign the pool carries: plain tile spectra / L^2
  const int L = c->L, H = c->geo.H;
  const float sgn = ((L / 2) % 2 == 1) ? -1.0f : 1.0f;
  for (size_t i = 0; i < n / 2; ++i) {
    const int ky = (int)(i % L), kx = (int)((i / L) % H);
    if (((kx + ky) & 1) ? (sgn > 0) : (sgn < 0)) {
      out[2 * i] = -out[2 * i];
      out[2 * i + 1] = -out[2 * i + 1];
    }
  }
  return GS_OK;
}

int gs_get_rotated_spectra(void* ctx, int psi_index, float* out, size_t cap) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !out || psi_index < 0) return GS_ERR_ARG;
  if (c->N == 0) return GS_ERR_STATE;
  const size_t n = (size_t)c->N * c->geo.H * c->L * 2;
  if (cap < n) return GS_ERR_CAPACITY;
  HIPCHK(hipSetDevice(c->device));
  CHK(rotate_templates(c, psi_index));
  CHK(copy_out(c, out, c->tspec.p, n * sizeof(float)));
  return GS_OK;
}

int gs_time_kernels(void* ctx, int psi_index, int reps, double* out_ms, int* maps_per_launch) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !out_ms || !maps_per_launch || reps <= 0 || psi_index < 0) return GS_ERR_ARG;
  if (c->N == 0 || !c->have_mic) return GS_ERR_STATE;
  HIPCHK(hipSetDevice(c->device));
  CHK(ensure_corr_buffers(c));
  CHK(rotate_templates(c, psi_index));
  const int L = c->L;
  const int nb = std::min(c->batch, c->T * c->N);
  const float lo = c->p.diameter / 4.0f, hi = (float)L - c->p.diameter / 4.0f;
  hipEvent_t e[3];
  for (auto& x : e) HIPCHK(hipEventCreate(&x));
  HIPCHK(hipEventRecord(e[0], c->stream));
  CHK(c->stats.ensure((size_t)nb * sizeof(MapStat)));
  int ylo, yhi;
  corr_store_rows(L, lo, hi, &ylo, &yhi);
  for (int i = 0; i < reps; ++i)
    launch_corr_pass1(L, c->tiles.as<float2>(), c->tspec.as<float2>(), c->maps_blk.as<MapRef>(), nb,
                      ylo, yhi, c->G[0].as<float2>(), c->psum.as<float2>(), c->twL.as<float2>(),
                      c->stream, c->var1);
  HIPCHK(hipEventRecord(e[1], c->stream));
  for (int i = 0; i < reps; ++i)
    launch_corr_pass2(L, c->G[0].as<float2>(), nb, lo, hi, c->part.as<CorrPartial>(),
                      c->twL.as<float2>(), c->twM.as<float2>(), c->stream, c->var2);
  HIPCHK(hipEventRecord(e[2], c->stream));
  HIPCHK(hipEventSynchronize(e[2]));
  HIPCHK(hipGetLastError());
  float a = 0, b = 0;
  HIPCHK(hipEventElapsedTime(&a, e[0], e[1]));
  HIPCHK(hipEventElapsedTime(&b, e[1], e[2]));
  for (auto& x : e) (void)hipEventDestroy(x);
  out_ms[0] = a / reps;
  out_ms[1] = b / reps;
  *maps_per_launch = nb;
  return GS_OK;
}

int gs_adjoint(void* ctx, const gs_peak* peaks, int n, gs_grad* out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || n < 0 || (n > 0 && (!peaks || !out))) return GS_ERR_ARG;
  if (c->N == 0 || !c->have_mic) return GS_ERR_STATE;
  if (c->mode != GS_MODE_FUSED) return GS_ERR_UNSUPPORTED;  // needs the half-spectrum tile pool
  if (n == 0) return GS_OK;
  HIPCHK(hipSetDevice(c->device));
  std::vector<RefCand> rc;
  std::vector<double2> cs;
  CHK(ref_cands(c, peaks, n, rc, cs));
  std::vector<float> r((size_t)n * 5);
  CHK(ref_eval(c, rc.data(), cs.data(), n, r.data()));
  for (int i = 0; i < n; ++i)
    out[i] = gs_grad{r[5 * i], r[5 * i + 1], r[5 * i + 2], r[5 * i + 3], r[5 * i + 4]};
  return GS_OK;
}

int gs_refine(void* ctx, gs_peak* peaks, int n, int iters, gs_grad* out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || n < 0 || iters < 0 || (n > 0 && !peaks)) return GS_ERR_ARG;
  if (c->N == 0 || !c->have_mic) return GS_ERR_STATE;
  if (c->mode != GS_MODE_FUSED) return GS_ERR_UNSUPPORTED;
  if (n == 0) return GS_OK;
  HIPCHK(hipSetDevice(c->device));
  // scaled steepest ascent: a step moves `step` in units of (1 degree,
  // 1 pixel, 1 pixel, 0.1 um) along the scaled gradient.  The next step
  // length comes from the parabola f(s) = f0 + d s + q s^2 through the
  // start (objective f0, slope d = |scaled gradient|) and the trial f(s):
  // its maximum -d / (2 q), kept within [s, 2 s] after an accepted trial and
  // within [0.1 s, 0.5 s] after a rejected one (2 s / 0.5 s when the
  // parabola has no maximum).  On the top 7,616 picks of 24 in-plane steps
  // of the full-size micrograph: 96,191 evaluations and mean objective
  // 18.2324 against 120,477 and 18.2318 with steps x1.25 / x0.5.
  const double scl[4] = {1.0, 1.0, 1.0, 0.1};
  const double lim[4] = {c->p.phi_step, 1.5, 1.5, 0.5};
  std::vector<gs_peak> cur(peaks, peaks + n), start(peaks, peaks + n);
  std::vector<RefCand> rc;
  std::vector<double2> cs;
  std::vector<float> fc((size_t)n * 5), ft((size_t)n * 5);
  CHK(ref_cands(c, cur.data(), n, rc, cs));
  CHK(ref_eval(c, rc.data(), cs.data(), n, fc.data()));
  c->timing.refine_evals = n;
  std::vector<double> step(n, 0.5);
  // a candidate whose step has shrunk below kMinStep (0.05 degree / pixel /
  // 0.005 um) has converged and leaves the evaluated set.  On the top 10,000
  // picks of the full-size micrograph (20 steps, steps x1.25 / x0.5): 0.01 ->
  // 198,261 evaluations, mean objective 19.9422; 0.05 -> 159,967, 19.9417;
  // 0.1 -> 134,883, 19.9400 (start 18.3470).  With the parabola steps, on
  // the top 7,616 picks of 24 in-plane steps: 0.03 / 0.05 / 0.08 / 0.1 ->
  // 105,169 / 96,191 / 87,134 / 82,473 evaluations, mean objective 18.2328 /
  // 18.2324 / 18.2311 / 18.2303; from 0.08 on some candidates stop without
  // any gain
  constexpr double kMinStep = 0.05;
  std::vector<int> act(n);
  for (int i = 0; i < n; ++i) act[i] = i;
  std::vector<gs_peak> tact;
  std::vector<float> thr;
  for (int it = 0; it < iters && !act.empty(); ++it) {
    tact.resize(act.size());
    for (size_t a = 0; a < act.size(); ++a) {
      const int i = act[a];
      const float* g = &fc[5 * (size_t)i + 1];
      double nrm = 0;
      for (int k = 0; k < 4; ++k) nrm += (scl[k] * g[k]) * (scl[k] * g[k]);
      nrm = std::sqrt(nrm);
      double th[4] = {cur[i].psi, cur[i].dx, cur[i].dy, cur[i].defocus};
      const double t0[4] = {start[i].psi, start[i].dx, start[i].dy, start[i].defocus};
      for (int k = 0; k < 4; ++k) {
        if (nrm > 0) th[k] += step[i] * scl[k] * scl[k] * g[k] / nrm;
        th[k] = std::min(std::max(th[k], t0[k] - lim[k]), t0[k] + lim[k]);
      }
      tact[a] = cur[i];
      tact[a].psi = (float)th[0];
      tact[a].dx = (float)th[1];
      tact[a].dy = (float)th[2];
      tact[a].defocus = (float)th[3];
    }
    const int na = (int)act.size();
    CHK(ref_cands(c, tact.data(), na, rc, cs));
    // the step is kept only if the objective rises, so the gradient at a
    // trial point is needed only there
    thr.resize(na);
    for (int a = 0; a < na; ++a) thr[a] = fc[5 * (size_t)act[a]];
    CHK(ref_eval(c, rc.data(), cs.data(), na, ft.data(), thr.data()));
    size_t keep = 0;
    for (int a = 0; a < na; ++a) {
      const int i = act[a];
      const float* g = &fc[5 * (size_t)i + 1];
      double d = 0;
      for (int k = 0; k < 4; ++k) d += (scl[k] * g[k]) * (scl[k] * g[k]);
      d = std::sqrt(d);
      const double s0 = step[i], f0 = fc[5 * (size_t)i], fs = ft[5 * (size_t)a];
      const double q = (fs - f0 - d * s0) / (s0 * s0);
      const bool up = fs > f0;
      double sn = q < 0 ? -d / (2 * q) : (up ? 2.0 : 0.5) * s0;
      sn = up ? std::min(std::max(sn, s0), 2.0 * s0) : std::min(std::max(sn, 0.1 * s0), 0.5 * s0);
      if (up) {
        cur[i] = tact[a];
        std::copy(&ft[5 * (size_t)a], &ft[5 * (size_t)a] + 5, &fc[5 * (size_t)i]);
      }
      step[i] = sn;
      if (step[i] >= kMinStep) act[keep++] = i;
    }
    act.resize(keep);
    c->timing.refine_evals += na;
  }
  for (int i = 0; i < n; ++i) {
    peaks[i] = cur[i];
    if (out)
      out[i] = gs_grad{fc[5 * (size_t)i], fc[5 * (size_t)i + 1], fc[5 * (size_t)i + 2],
                       fc[5 * (size_t)i + 3], fc[5 * (size_t)i + 4]};
  }
  return GS_OK;
}

int gs_set_fft_variants(void* ctx, int pass1, int pass2) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c) return GS_ERR_ARG;
  // windows compiled without variants run variant 0 whatever is set
  const int nv = corr_max_variants();
  if (pass1 < 0 || pass2 < 0 || pass1 >= nv || pass2 >= nv) return GS_ERR_ARG;
  c->var1 = pass1;
  c->var2 = pass2;
  drop_graphs(c);  // captured with the old schedule
  return GS_OK;
}

int gs_autotune(void* ctx, int psi_index, gs_tune* out) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || psi_index < 0) return GS_ERR_ARG;
  if (c->N == 0 || !c->have_mic) return GS_ERR_STATE;
  if (c->mode != GS_MODE_FUSED || c->p.norm_type != 0) return GS_ERR_UNSUPPORTED;
  if (psi_index >= num_psi(c->p.phi_step)) return GS_ERR_ARG;
  HIPCHK(hipSetDevice(c->device));
  gs_tune t{};
  const int L = c->L, nv = corr_variants(L);
  const bool graphs = c->use_graphs, prof = c->profiling;
  c->use_graphs = false;
  // 1. FFT schedule of each pass, timed alone on one batch of this psi
  CHK(ensure_corr_buffers(c));
  CHK(rotate_templates(c, psi_index));
  const int nb = std::min(c->batch, c->T * c->N);
  const float lo = c->p.diameter / 4.0f, hi = (float)L - c->p.diameter / 4.0f;
  int ylo, yhi;
  corr_store_rows(L, lo, hi, &ylo, &yhi);
  hipEvent_t e[2];
  for (auto& x : e) HIPCHK(hipEventCreate(&x));
  auto timed = [&](bool pass1, int v, double* us) -> int {
    for (int rep = 0; rep < 4; ++rep) {  // the first launch warms up
      if (rep == 1) HIPCHK(hipEventRecord(e[0], c->stream));
      if (pass1)
        launch_corr_pass1(L, c->tiles.as<float2>(), c->tspec.as<float2>(),
                          c->maps_blk.as<MapRef>(), nb, ylo, yhi, c->G[0].as<float2>(),
                          c->psum.as<float2>(), c->twL.as<float2>(), c->stream, v);
      else
        launch_corr_pass2(L, c->G[0].as<float2>(), nb, lo, hi, c->part.as<CorrPartial>(),
                          c->twL.as<float2>(), c->twM.as<float2>(), c->stream, v);
    }
    HIPCHK(hipEventRecord(e[1], c->stream));
    HIPCHK(hipEventSynchronize(e[1]));
    HIPCHK(hipGetLastError());
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e[0], e[1]));
    *us = 1e3 * ms / 3.0 / nb;
    return GS_OK;
  };
  int best1 = 0, best2 = 0;
  for (int pass = 1; pass <= 2; ++pass)
    for (int v = 0; v < nv && v < 4; ++v) {
      double us = 0;
      CHK(timed(pass == 1, v, &us));
      double* row = pass == 1 ? t.pass1_us : t.pass2_us;
      row[v] = us;
      int& best = pass == 1 ? best1 : best2;
      if (us < row[best]) best = v;
    }
  // the tuned schedule (variant 0) stays unless another is clearly (2 %)
  // faster alone: a variant 0.4 % ahead in this timing once ran the whole
  // correlation 5 % slower (pass 2, 0.492 against 0.494 us per map alone,
  // 0.732 against 0.698 ms per launch in the psi's schedule)
  if (t.pass1_us[best1] > 0.98 * t.pass1_us[0]) best1 = 0;
  if (t.pass2_us[best2] > 0.98 * t.pass2_us[0]) best2 = 0;
  for (auto& x : e) (void)hipEventDestroy(x);
  c->var1 = best1;
  c->var2 = best2;
  // 2. launch schedule, timed on the whole correlation of the psi (gs_pick
  //    with launch sampling): the tile block (the cache-sized default or all
  //    tiles in one block), then batch and pipeline
  c->profiling = true;
  std::vector<gs_particle> keep_picks = c->picks;
  auto psi_time = [&](double* ms) -> int {
    *ms = 1e300;
    for (int rep = 0; rep < 3; ++rep) {  // the first run allocates and warms up
      size_t np = 0;
      const int r = gs_pick(ctx, psi_index, psi_index + 1, nullptr, 0, &np);
      if (r != GS_OK && r != GS_ERR_CAPACITY) return r;
      if (rep > 0) *ms = std::min(*ms, c->timing.corr_psi_ms);
    }
    t.candidates++;
    return GS_OK;
  };
  {
    const int tb0 = c->tile_block;
    const long tile_bytes = (long)(L / 2 + 1) * L * (long)sizeof(float2);
    const int tb_auto = (int)std::max(1l, (192l << 20) / tile_bytes);
    const int cur_tb = tb0 > 0 ? tb0 : tb_auto;
    if (c->T > cur_tb) {  // otherwise every block size gives the same order
      double ms_cur = 0, ms_all = 0;
      CHK(psi_time(&ms_cur));
      c->tile_block = c->T;
      CHK(psi_time(&ms_all));
      c->tile_block = ms_all < 0.99 * ms_cur ? c->T : tb0;
    }
  }
  const int b0 = c->batch;
  const int batches[3] = {std::max(1, b0 / 2), b0, 2 * b0};
  // a schedule chosen with gs_set_pipeline stays; otherwise the launch
  // schedules are candidates
  const int p0 = c->pipeline;
  const int npipe = c->pipeline_set ? 1 : 3;
  const int pipes[3] = {c->pipeline_set ? p0 : GS_PIPE_SERIAL, GS_PIPE_LANES, GS_PIPE_LAUNCH};
  double best_ms = 1e300, cur_ms = 1e300;
  int best_b = b0, best_p = p0;
  for (int bi = 0; bi < 3; ++bi)
    for (int pi = 0; pi < npipe; ++pi) {
      c->batch = batches[bi];
      c->pipeline = pipes[pi];
      double ms = 1e300;
      CHK(psi_time(&ms));
      if (batches[bi] == b0 && pipes[pi] == p0) cur_ms = ms;
      if (ms < best_ms) {
        best_ms = ms;
        best_b = batches[bi];
        best_p = pipes[pi];
      }
    }
  // the schedule in use stays unless another is clearly (1 %) faster: the
  // candidates often differ by less than the run-to-run spread
  if (cur_ms < 1e300 && best_ms > 0.99 * cur_ms) {
    best_ms = cur_ms;
    best_b = b0;
    best_p = p0;
  }
  c->picks = keep_picks;
  c->stats_npsi = 0;
  c->batch = best_b;
  c->pipeline = best_p;
  c->profiling = prof;
  c->use_graphs = graphs;
  drop_graphs(c);  // captured with the old schedule
  t.pass1_variant = best1;
  t.pass2_variant = best2;
  t.tile_block = c->tile_block;
  t.batch = best_b;
  t.pipeline = best_p;
  t.psi_ms = best_ms;
  if (out) *out = t;
  return GS_OK;
}

int gs_get_info(void* ctx, gs_info* info) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !info) return GS_ERR_ARG;
  info->l = c->L;
  info->n_views = c->N;
  info->n_psi = num_psi(c->p.phi_step);
  const MicSlot* ms = c->mics.empty() ? nullptr : &c->mics[0];
  info->nx = ms ? ms->nx : 0;
  info->ny = ms ? ms->ny : 0;
  info->tiles_x = ms ? (int)ms->offx.size() : 0;
  info->tiles_y = ms ? (int)ms->offy.size() : 0;
  info->n_mics = (int)c->mics.size();
  info->n_tiles = c->T;
  info->box_x = c->bx;
  info->box_y = c->by;
  return GS_OK;
}

int gs_get_timing(void* ctx, gs_timing* t) {
  Ctx* c = static_cast<Ctx*>(ctx);
  if (!c || !t) return GS_ERR_ARG;
  *t = c->timing;
  return GS_OK;
}

}  // extern "C"
