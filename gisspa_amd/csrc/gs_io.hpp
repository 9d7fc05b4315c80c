// gs_io.hpp -- host-side formats of the reference's drop-in surface:
// the key = value config file, the Euler .lst, the micrograph .lst (#LST /
// #LSX), MRC images/stacks and the particle-list output line.
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gisspa.h"

namespace gsio {

// The reference's Parameters (EMReader/DataReader.h:14-42): paths, loop range
// and device id on top of the numeric gs_params.
struct Config {
  std::string input, templ, eulerf;
  std::string output = "./outlst";
  int first = -1, last = -1;
  int device_id = 0;
  // micrographs per correlation batch (not a reference key; the reference
  // processes one per loop iteration, main.cu:1149)
  int mic_batch = 4;
  gs_params p{};
  std::vector<std::string> errors;    // missing required keys
  std::vector<std::string> unknown;   // "Undefined option" lines
};

// readConfig + parse (EMReader/DataReader.cpp:128-265).  Returns false if the
// file cannot be opened.
bool read_config(const std::string& path, Config& cfg);
void print_config(const Config& cfg, FILE* out);
void print_help(FILE* out);

struct Euler {
  std::vector<float> alt, az;
  std::vector<float> phi;  // third column (0 if absent); the picker ignores it
};
// readEulerData (DataReader.cpp:267-284): "alt az ..." per line, '#' comments.
bool read_euler(const std::string& path, Euler& e);

struct MicEntry {
  int nn = 0;              // index column of the lst line
  std::string path;        // with the lst directory prefix (main.cu:125-138)
  float defocus = 0, dfdiff = 0, dfang = 0;  // as written (defocus positive, um)
};
// readInLst_and_consturctPairs + parsePairs + addPrefix (DataReader.cpp:318-387,
// main.cu:125-156) for image n.  Returns 1 on success, <0 on error codes of
// the reference (-1 not found, -2 short #LSX, -3 unreadable).
int read_lst_entry(const std::string& lst, int n, MicEntry& out);

struct Image {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> data;
};
// readMRC (EMReader/emdata_.cpp:384-617): modes 0, 1 and 2 (modes above 4
// rejected as there, emdata_.cpp:457; complex modes 3 and 4 are not
// images), either byte order; section n of a stack (n < 0 or out of range:
// whole file).
int read_mrc(const std::string& path, int n, Image& img, std::string* err);
// number of sections
int mrc_sections(const std::string& path);
// emdata::readImage (EMReader/emdata_.cpp:174-205): an HDF5 file (EMAN1 or
// EMAN2 layout) or else an MRC file.  n >= 0 reads image n; n < 0 the whole
// file (an MRC stack or volume; an HDF stack of equal-size 2D images as nz
// sections).  Returns 0, or a negative value with *err set.
int read_image(const std::string& path, int n, Image& img, std::string* err);
// the first count images of a stack, each with its own size (HDF stacks may
// mix sizes; MRC stacks share one)
int read_images(const std::string& path, int count, std::vector<Image>& out, std::string* err);
// number of images (MRC sections, HDF datasets); negative on error
int image_count(const std::string& path, std::string* err);
// write a mode-2 MRC (for tools and tests)
int write_mrc(const std::string& path, const float* data, int nx, int ny, int nz);

// One output line, main.cu:985.
void write_pick_line(FILE* fp, int nn, const std::string& path, const MicEntry& m,
                     float euler1, float euler2, float psi, float cx, float cy, float score);

// create_outpath (main.cu:1006-1023): mkdir every directory prefix.
void create_outpath(const std::string& path);

}  // namespace gsio
