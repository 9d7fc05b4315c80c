// project3d -- templates for a picking run: projections of a density map at
// the views of an Euler file, written as an MRC stack.
//
//   project3d <volume.mrc> <euler.lst> <templates.mrcs> [gpu_id]
//
// The reference's protocol makes its templates with EMAN's project3d
// (step 5) from the Euler files it ships (Euler_file/*.lst, "alt az phi"
// per line); this tool fills that role on the GPU (gs_project): line J of the
// Euler file gives template J, the order the picker's `eulerfile` and
// `template` keys expect.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gs_io.hpp"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "project3d <volume.mrc> <euler.lst> <templates.mrcs> [gpu_id]\n");
    return 1;
  }
  const int device = argc > 4 ? std::atoi(argv[4]) : 0;
  gsio::Image vol;
  std::string err;
  if (gsio::read_image(argv[1], -1, vol, &err) != 0) {
    std::fprintf(stderr, "%s: %s\n", argv[1], err.c_str());
    return 1;
  }
  if (vol.nx != vol.ny || vol.nx != vol.nz || vol.nx < 2) {
    std::fprintf(stderr, "%s: a cubic volume is needed (%d x %d x %d)\n", argv[1], vol.nx, vol.ny,
                 vol.nz);
    return 1;
  }
  gsio::Euler eu;
  if (!gsio::read_euler(argv[2], eu)) return 1;
  const int nv = (int)eu.alt.size(), n = vol.nx;
  if (nv == 0) {
    std::fprintf(stderr, "%s: no views\n", argv[2]);
    return 1;
  }
  std::vector<float> ang((size_t)nv * 3);
  for (int j = 0; j < nv; ++j) {
    ang[3 * j] = eu.alt[j];
    ang[3 * j + 1] = eu.az[j];
    ang[3 * j + 2] = eu.phi[j];
  }
  std::vector<float> out((size_t)nv * n * n);
  const int rc = gs_project(device, vol.data.data(), n, ang.data(), nv, out.data());
  if (rc != GS_OK) {
    std::fprintf(stderr, "gs_project: %s\n", gs_strerror(rc));
    return 1;
  }
  if (gsio::write_mrc(argv[3], out.data(), n, n, nv) != 0) {
    std::fprintf(stderr, "cannot write %s\n", argv[3]);
    return 1;
  }
  std::fprintf(stderr, "%d projections of %d^3 written to %s\n", nv, n, argv[3]);
  return 0;
}
