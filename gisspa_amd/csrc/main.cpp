// main.cpp -- `./main config_file`: the reference's command line
// (Autumn1998/GisSPA main.cu:1025-1299) over the libgisspa C ABI.
//
// Same config keys, Euler .lst, micrograph .lst and output lines.  Errors end
// the run with a message and a non-zero status instead of continuing on a
// 10x10 zero image (emdata_.cpp:184-203) or garbage defaults.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <future>
#include <string>
#include <vector>

#include "../../include/gisspa.h"
#include "gs_io.hpp"

using namespace gsio;

static int fail(const char* what, int code) {
  fprintf(stderr, "Error: %s: %s (%d)\n", what, gs_strerror(code), code);
  return 1;
}

// One group of consecutive micrographs read from disk: their lst entries and
// images, or the first error met (the lst code of readRawImage, or a file
// error message).
struct MicGroup {
  std::vector<MicEntry> es;
  std::vector<Image> ims;
  int lst_error = 0;
  std::string err;
};

static MicGroup load_group(const Config& cfg, int n0, int k) {
  MicGroup g;
  g.es.resize(k);
  g.ims.resize(k);
  for (int i = 0; i < k; ++i) {
    const int r = read_lst_entry(cfg.input, n0 + i, g.es[i]);
    if (r < 0) {
      g.lst_error = r;
      return g;
    }
    if (read_image(g.es[i].path, g.es[i].nn, g.ims[i], &g.err)) {
      if (g.err.empty()) g.err = "cannot read " + g.es[i].path;
      return g;
    }
  }
  return g;
}

int main(int argc, char** argv) {
  if (argc == 1 || (argv[1][0] == '-' && argv[1][1] == 'h')) {
    print_help(stdout);
    return 0;
  }
  Config cfg;
  if (!read_config(argv[1], cfg)) return 1;
  print_config(cfg, stdout);
  if (!cfg.errors.empty()) return 1;

  Euler euler;
  if (!read_euler(cfg.eulerf, euler)) return 1;
  const int N = (int)euler.alt.size();
  if (N == 0) {
    fprintf(stderr, "Error: no views in %s\n", cfg.eulerf.c_str());
    return 1;
  }

  create_outpath(cfg.output);
  FILE* fp = fopen(cfg.output.c_str(), "wb");
  if (!fp) {
    printf("Out file create failed! %s\n", cfg.output.c_str());
    printf("Please confirm the path exists!\n");
    return 1;
  }

  // first image: its CTF weights the templates (main.cu:1106)
  MicEntry first;
  int r = read_lst_entry(cfg.input, cfg.first, first);
  if (r < 0) {
    printf("Error %d occured in readRawImage!\n", r);
    return 1;
  }
  Image img;
  std::string err;
  if (read_image(first.path, first.nn, img, &err)) {
    fprintf(stderr, "Error: %s\n", err.c_str());
    return 1;
  }
  img = Image();

  // templates: the first N images of the template file (MRC stack, or HDF
  // with one size per image, main.cu:182-223)
  const int nsec = image_count(cfg.templ, &err);
  if (nsec < N) {
    if (nsec < 0) fprintf(stderr, "Error: %s\n", err.c_str());
    else fprintf(stderr, "Error: %s holds %d templates, the Euler file lists %d\n",
                 cfg.templ.c_str(), nsec, N);
    return 1;
  }
  std::vector<Image> tms;
  if (read_images(cfg.templ, N, tms, &err)) {
    fprintf(stderr, "Error: %s\n", err.c_str());
    return 1;
  }
  bool one_box = true;
  for (const Image& t : tms) one_box = one_box && t.nx == tms[0].nx && t.ny == tms[0].ny;

  void* ctx = nullptr;
  if ((r = gs_create(&cfg.p, cfg.device_id, &ctx)) != GS_OK) return fail("gs_create", r);
  gs_ctf tctf{first.defocus, first.dfdiff, first.dfang};
  if (one_box) {
    const int bx = tms[0].nx, by = tms[0].ny;
    std::vector<float> tm((size_t)N * bx * by);
    for (int j = 0; j < N; ++j)
      std::copy(tms[j].data.begin(), tms[j].data.end(), tm.begin() + (size_t)j * bx * by);
    tms.clear();
    if ((r = gs_set_templates(ctx, tm.data(), N, bx, by, &tctf)) != GS_OK)
      return fail("gs_set_templates", r);
  } else {
    std::vector<const float*> ptr(N);
    std::vector<int> bx(N), by(N);
    for (int j = 0; j < N; ++j) {
      ptr[j] = tms[j].data.data();
      bx[j] = tms[j].nx;
      by[j] = tms[j].ny;
    }
    if ((r = gs_set_templates_mixed(ctx, N, ptr.data(), bx.data(), by.data(), &tctf)) != GS_OK)
      return fail("gs_set_templates_mixed", r);
    tms.clear();
  }
  gs_info info;
  gs_get_info(ctx, &info);
  if (info.l != cfg.p.window_size)
    printf("Aujust padding size to %d\n", info.l);

  const auto t_start = std::chrono::steady_clock::now();
  // Micrographs go to the picker in groups of micrograph_batch (any sizes,
  // each with its own CTF): one search covers the group's tiles, so each
  // template rotation serves all of them.  Picks come back micrograph by
  // micrograph, and the lines and "finished" messages follow the order of
  // the reference's one-image loop (main.cu:1149-1269).
  // The next group is read from disk while the GPU searches the current one.
  std::vector<gs_particle> picks(1 << 16);
  auto group_size = [&](int n0) { return std::min(cfg.mic_batch, cfg.last - n0); };
  std::future<MicGroup> next;
  if (cfg.first < cfg.last)
    next = std::async(std::launch::async, load_group, std::cref(cfg), cfg.first,
                      group_size(cfg.first));
  for (int n0 = cfg.first; n0 < cfg.last; n0 += cfg.mic_batch) {
    const int k = group_size(n0);
    MicGroup g = next.get();
    if (g.lst_error < 0) {
      printf("Error %d occured in readRawImage!\n", g.lst_error);
      return 1;
    }
    if (!g.err.empty()) {
      fprintf(stderr, "Error: %s\n", g.err.c_str());
      return 1;
    }
    std::vector<const float*> ptr(k);
    std::vector<int> nxs(k), nys(k);
    std::vector<gs_ctf> ctfs(k);
    for (int i = 0; i < k; ++i) {
      ptr[i] = g.ims[i].data.data();
      nxs[i] = g.ims[i].nx;
      nys[i] = g.ims[i].ny;
      ctfs[i] = gs_ctf{g.es[i].defocus, g.es[i].dfdiff, g.es[i].dfang};
    }
    if ((r = gs_set_micrographs(ctx, k, ptr.data(), nxs.data(), nys.data(), ctfs.data(), 0)) !=
        GS_OK)
      return fail("gs_set_micrographs", r);
    g.ims.clear();   // copied to the device
    const int n1 = n0 + cfg.mic_batch;
    if (n1 < cfg.last)
      next = std::async(std::launch::async, load_group, std::cref(cfg), n1, group_size(n1));
    size_t np = 0;
    r = gs_pick(ctx, 0, -1, picks.data(), picks.size(), &np);
    if (r == GS_ERR_CAPACITY) {
      picks.resize(np);
      r = gs_get_picks(ctx, picks.data(), picks.size(), &np);
    }
    if (r != GS_OK) return fail("gs_pick", r);
    size_t i = 0;
    for (int m = 0; m < k; ++m) {
      const MicEntry& e = g.es[m];
      for (; i < np && picks[i].mic == m; ++i) {
        const gs_particle& pk = picks[i];
        write_pick_line(fp, e.nn, e.path, e, euler.alt[pk.view], euler.az[pk.view], pk.psi, pk.cx,
                        pk.cy, pk.score);
      }
      fflush(fp);
      printf("------------------------Image %d finished--------------------------------\n", n0 + m);
    }
    if (i != np) {
      fprintf(stderr, "Error: picks out of micrograph order\n");
      return 1;
    }
  }
  const double secs =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  printf("Total consumed time is: %f seconds\n", secs);
  gs_destroy(ctx);
  fclose(fp);
  return 0;
}
