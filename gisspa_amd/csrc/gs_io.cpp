// gs_io.cpp -- config / Euler / lst / MRC / output formats of the reference's
// drop-in surface (Autumn1998/GisSPA EMReader/ and main.cu).
#include "gs_io.hpp"

#include "gs_hdf.hpp"

#include <dlfcn.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

namespace gsio {

namespace {

std::string rstrip(std::string s) {
  while (!s.empty() && (s.back() == '\r' || s.back() == '\n' || s.back() == ' ' || s.back() == '\t'))
    s.pop_back();
  return s;
}

// parse (DataReader.cpp:128-247): option = text up to the first ' ' or '=';
// value = the rest after skipping ' ' and '='.
void parse_line(const std::string& line, Config& c) {
  size_t i = 0;
  while (i < line.size() && line[i] != ' ' && line[i] != '=') ++i;
  const std::string option = line.substr(0, i);
  while (i < line.size() && (line[i] == ' ' || line[i] == '=')) ++i;
  const std::string value = rstrip(line.substr(i));
  const char* v = value.c_str();
  gs_params& p = c.p;
  if (option == "input") c.input = value;
  else if (option == "template") c.templ = value;
  else if (option == "eulerfile") c.eulerf = value;
  else if (option == "kk") p.kk = (float)atof(v);
  else if (option == "angpix") p.apix = (float)atof(v);
  else if (option == "phistep") p.phi_step = (float)atof(v);
  else if (option == "energy") p.energy = (float)atof(v);
  else if (option == "cs") p.cs = (float)atof(v);
  else if (option == "Highres") p.highres = (float)atof(v);
  else if (option == "Lowres") p.lowres = (float)atof(v);
  else if (option == "diameter") p.diameter = (float)atof(v);
  else if (option == "output") c.output = value;
  else if (option == "first") c.first = atoi(v);
  else if (option == "last") c.last = atoi(v);
  else if (option == "threshold") p.threshold = (float)atof(v);
  else if (option == "GPU_ID") c.device_id = atoi(v);
  else if (option == "window_size") p.window_size = atoi(v);
  else if (option == "phase_flip") p.phase_flip = atoi(v);
  else if (option == "overlap") p.overlap = atoi(v);
  else if (option == "norm_type") p.norm_type = atoi(v);
  else if (option == "invert") p.invert = atoi(v);
  else if (option == "ampconst") p.ampconst = (float)atof(v);  // not a reference key (fixed 0.07 there)
  else if (option == "micrograph_batch") c.mic_batch = atoi(v);  // not a reference key
  else {
    c.unknown.push_back(value);
    printf("Undefined option: %s .Abort.\n", value.c_str());
  }
}

bool is_machine_big_endian() {
  const uint32_t one = 1;
  return reinterpret_cast<const unsigned char*>(&one)[0] == 0;
}

void swap4(void* p) {
  unsigned char* c = static_cast<unsigned char*>(p);
  std::swap(c[0], c[3]);
  std::swap(c[1], c[2]);
}

}  // namespace

bool read_config(const std::string& path, Config& c) {
  gs_params_default(&c.p);
  std::ifstream in(path);
  if (!in) {
    printf("Error => Open config file failed : %s\n\n", path.c_str());
    return false;
  }
  std::string buf;
  while (std::getline(in, buf)) {
    buf = rstrip(buf);
    if (buf.empty() || buf[0] == '#') continue;
    parse_line(buf, c);
  }
  // checkRequestPara (DataReader.cpp:302-315)
  auto req = [&](bool bad, const char* msg) {
    if (bad) {
      c.errors.push_back(msg);
      printf("Error : %s\n", msg);
    }
  };
  req(c.input.empty(), "lst/star file is requested.");
  req(c.templ.empty(), "template file is requested.");
  req(c.eulerf.empty(), "euler file is requested.");
  req(c.p.apix < 0, "angpix is requested.");
  req(c.p.phi_step < 0, "phistep is requested.");
  req(c.p.kk < 0, "kk is requested.");
  req(c.p.energy < 0, "energy is requested.");
  req(c.p.cs < 0, "cs is requested.");
  req(c.p.highres < 0, "highres is requested.");
  req(c.p.lowres < 0, "lowres is requested.");
  req(c.p.diameter < 0, "diameter is requested.");
  req(c.mic_batch < 1, "micrograph_batch must be at least 1.");
  return true;
}

void print_config(const Config& c, FILE* o) {
  // printAllPara, DataReader.h:44-77
  const gs_params& p = c.p;
  fprintf(o, "Parameter requested:\n");
  fprintf(o, "--input            %s\n", c.input.c_str());
  fprintf(o, "--template         %s\n", c.templ.c_str());
  fprintf(o, "--eulerfile        %s\n", c.eulerf.c_str());
  fprintf(o, "--angpix           %f\n", p.apix);
  fprintf(o, "--phistep          %f\n", p.phi_step);
  fprintf(o, "--kk               %f\n", p.kk);
  fprintf(o, "--energy           %f\n", p.energy);
  fprintf(o, "--cs               %f\n", p.cs);
  fprintf(o, "--Highres          %f\n", p.highres);
  fprintf(o, "--Lowres           %f\n", p.lowres);
  fprintf(o, "--diameter         %f\n", p.diameter);
  fprintf(o, "Parameter optional:\n");
  fprintf(o, "--threshold        %f\n", p.threshold);
  fprintf(o, "--output           %s\n", c.output.c_str());
  fprintf(o, "--first            %d\n", c.first);
  fprintf(o, "--last             %d\n", c.last);
  fprintf(o, "--window_size      %d\n", p.window_size);
  fprintf(o, "--GPU_ID           %d\n", c.device_id);
  fprintf(o, "--phase_flip       %d\n", p.phase_flip);
  fprintf(o, "--overlap          %d\n", p.overlap);
  fprintf(o, "--norm_type        %d\n", p.norm_type);
  fprintf(o, "--invert           %d\n", p.invert);
  fprintf(o, "--ampconst         %f\n", p.ampconst);
  fprintf(o, "--micrograph_batch %d\n", c.mic_batch);
  fprintf(o, "Parameter fixed:\n");
  fprintf(o, "--a                %f\n", p.a);
  fprintf(o, "--b                %f\n", p.b);
  fprintf(o, "--b2               %f\n", p.b2);
  fprintf(o, "--bfactor          %f\n", p.bfactor);
  fprintf(o, "--bfactor2         %f\n", p.bfactor2);
  fprintf(o, "--bfactor3         %f\n\n", p.bfactor3);
}

void print_help(FILE* o) {
  // printHelpMsg, EMReader/util_func.cpp:78-110 (same keys)
  fprintf(o, "This program detect targets with orientations and tanslations.\n");
  fprintf(o, "How to use this problem:  \"./main config_file\"  \n");
  fprintf(o, "All parameters should be set on the config file.\n");
  fprintf(o, "(for .eg,  \"input  =  /Data/inputfile\" )\n");
  fprintf(o, "Requested Parameters:\n");
  fprintf(o, "input            = input filtered images lstfile with ctf information\n");
  fprintf(o, "template         = input 2D projections templates (MRC stack)\n");
  fprintf(o, "eulerfile        = euler file with euler values\n");
  fprintf(o, "angpix           = input pixel size in angstroms\n");
  fprintf(o, "phistep          = inplane rotation sampling\n");
  fprintf(o, "kk               = overlapping density parameter, default is 3.\n");
  fprintf(o, "energy           = accerlerating voltage in kV.\n");
  fprintf(o, "cs               = spherical aberration in mm.\n");
  fprintf(o, "Highres          = high resolution cut \n");
  fprintf(o, "Lowres           = low resolution cut\n");
  fprintf(o, "diameter         = target diameter in pixels\n");
  fprintf(o, "Optional Parameters:\n");
  fprintf(o, "norm_type        = CCG norm type, 0/ 1\n");
  fprintf(o, "invert           = 0 / 1, choose whether to invert contrast, 0 for skipping and 1 for invert\n");
  fprintf(o, "threshold        = cc threshold value, only output score beyond this value\n");
  fprintf(o, "output           = output lstfile filaname\n");
  fprintf(o, "first            = the first image id to process.\n");
  fprintf(o, "last             = the last image id to process.\n");
  fprintf(o, "window_size      = the window size which is splitted from raw IMG.\n");
  fprintf(o, "GPU_ID           = ID of GPU device.\n");
  fprintf(o, "phase_flip       = Whether use phase filp operation(1) or not(0).\n");
  fprintf(o, "overlap          = size of overlap between diff window.\n");
  fprintf(o, "ampconst         = amplitude contrast of the CTF (default 0.07).\n");
  fprintf(o, "micrograph_batch = micrographs searched together, sharing each template rotation\n");
  fprintf(o, "                   (default 4; output unchanged).\n");
}

bool read_euler(const std::string& path, Euler& e) {
  std::ifstream in(path);
  if (!in) {
    printf("Error => Open euler file failed : %s\n\n", path.c_str());
    return false;
  }
  std::string buf;
  while (std::getline(in, buf)) {
    if (buf.empty() || buf[0] == '#') continue;
    float a = 0, b = 0, c = 0;
    sscanf(buf.c_str(), "%f %f %f", &a, &b, &c);
    e.alt.push_back(a);
    e.az.push_back(b);
    e.phi.push_back(c);
  }
  return true;
}

int read_lst_entry(const std::string& lst, int n, MicEntry& out) {
  const int MAXLEN = 1024 + 10;
  std::vector<char> buf(MAXLEN + 1), t(MAXLEN + 1), u(MAXLEN + 1);
  FILE* in = fopen(lst.c_str(), "rb");
  if (!in) {
    printf("'%s' not found\n", lst.c_str());
    return -3;
  }
  int nn = 0;
  t[0] = 0;
  u[0] = 0;
  if (!fgets(buf.data(), MAXLEN, in)) {
    fclose(in);
    return -1;
  }
  if (strncmp(buf.data(), "#LSX", 4) == 0) {
    int step = 0;
    if (!fgets(buf.data(), MAXLEN, in) || !fgets(buf.data(), MAXLEN, in)) {
      fclose(in);
      return -1;
    }
    sscanf(buf.data() + 1, " %d", &step);
    if (step == 0) {
      fclose(in);
      return -1;
    }
    long tl = ftell(in);
    if (n > 0) fseek(in, (long)step * n, SEEK_CUR);
    if (!fgets(buf.data(), MAXLEN, in)) {
      fclose(in);
      return -2;
    }
    if ((ftell(in) - tl) % step != 0) {
      printf("Error in #LSX file detected. Please rerun lstfast.py\n");
      fclose(in);
      return -3;
    }
    sscanf(buf.data(), " %d %1034s %1034[^\n]", &nn, t.data(), u.data());
  } else {
    // the first line is a header and is skipped (DataReader.cpp:351-360)
    if (n < 0) n = 0;
    int i;
    for (i = 0; i <= n; i++) {
      u[0] = 0;
      if (!fgets(buf.data(), MAXLEN, in)) break;
      if (buf[0] == '#') i--;
      else sscanf(buf.data(), " %d %1034s %1034[^\n]", &nn, t.data(), u.data());
    }
    if (i <= n) {
      fclose(in);
      return -1;
    }
  }
  fclose(in);
  out.nn = nn;
  // addPrefix, main.cu:125-138: the lst file's directory
  const size_t slash = lst.rfind('/');
  std::string prefix = (slash == std::string::npos) ? "" : lst.substr(0, slash + 1);
  out.path = prefix + std::string(t.data());
  out.defocus = out.dfdiff = out.dfang = 0;
  // Split on whitespace + parsePairs (DataReader.cpp:366-387)
  std::istringstream ss(std::string(u.data()));
  std::string tok;
  while (ss >> tok) {
    if (tok.compare(0, 8, "defocus=") == 0) sscanf(tok.c_str() + 8, "%f", &out.defocus);
    else if (tok.compare(0, 7, "dfdiff=") == 0) sscanf(tok.c_str() + 7, "%f", &out.dfdiff);
    else if (tok.compare(0, 6, "dfang=") == 0) sscanf(tok.c_str() + 6, "%f", &out.dfang);
  }
  return 1;
}

namespace {
struct MrcHeader {
  int32_t w[56];
  char labels[10][80];
};
static_assert(sizeof(MrcHeader) == 1024, "MRC header is 1024 bytes");

bool ends_with(const std::string& s, const char* suffix) {
  const size_t n = strlen(suffix);
  return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

// bzip2 decompression through the system's libbz2 (loaded at run time; its
// public stream interface: bzlib.h of bzip2 1.0)
struct BzStream {
  char* next_in;
  unsigned int avail_in, total_in_lo32, total_in_hi32;
  char* next_out;
  unsigned int avail_out, total_out_lo32, total_out_hi32;
  void* state;
  void* (*bzalloc)(void*, int, int);
  void (*bzfree)(void*, void*);
  void* opaque;
};

bool bunzip2(const std::vector<char>& in, std::vector<char>& out) {
  void* lib = dlopen("libbz2.so.1", RTLD_NOW | RTLD_LOCAL);
  if (!lib) lib = dlopen("libbz2.so.1.0", RTLD_NOW | RTLD_LOCAL);
  if (!lib) return false;
  using InitF = int (*)(BzStream*, int, int);
  using StepF = int (*)(BzStream*);
  auto init = reinterpret_cast<InitF>(dlsym(lib, "BZ2_bzDecompressInit"));
  auto step = reinterpret_cast<StepF>(dlsym(lib, "BZ2_bzDecompress"));
  auto end = reinterpret_cast<StepF>(dlsym(lib, "BZ2_bzDecompressEnd"));
  bool ok = false;
  if (init && step && end) {
    const int kOk = 0, kStreamEnd = 4;
    size_t pos = 0;   // concatenated streams (as bzcat reads them)
    ok = true;
    while (ok && pos < in.size()) {
      BzStream z{};
      if (init(&z, 0, 0) != kOk) {
        ok = false;
        break;
      }
      z.next_in = const_cast<char*>(in.data()) + pos;
      z.avail_in = (unsigned)std::min<size_t>(in.size() - pos, 1u << 30);
      int r = kOk;
      char buf[1 << 16];
      while (r == kOk) {
        z.next_out = buf;
        z.avail_out = sizeof(buf);
        r = step(&z);
        out.insert(out.end(), buf, buf + (sizeof(buf) - z.avail_out));
        if (r == kOk && z.avail_in == 0 && z.avail_out != 0) r = -1;   // truncated
      }
      pos = (size_t)(z.next_in - in.data());
      end(&z);
      ok = r == kStreamEnd;
    }
  }
  dlclose(lib);
  return ok;
}

// Open an MRC file for reading.  readMRC (emdata_.cpp:403-413) meant to read
// .gz / .Z files through zcat and reads .bz / .bz2 files through bzcat; here
// .gz (zlib) and .bz / .bz2 (libbz2) files are decompressed into `mem` and
// read from memory, so stack sections can still be addressed.  (.Z is not
// read: the reference's zcat stream is replaced by a plain fopen before use,
// emdata_.cpp:403-413, so it does not read such files either.)
FILE* open_mrc_stream(const std::string& path, std::vector<char>& mem, std::string* err) {
  const bool gz = ends_with(path, ".gz");
  const bool bz = ends_with(path, ".bz") || ends_with(path, ".bz2");
  if (!gz && !bz) return fopen(path.c_str(), "rb");
  mem.clear();
  bool ok = false;
  if (gz) {
    gzFile g = gzopen(path.c_str(), "rb");
    if (g) {
      char buf[1 << 16];
      int k;
      while ((k = gzread(g, buf, sizeof(buf))) > 0) mem.insert(mem.end(), buf, buf + k);
      ok = k == 0;
      gzclose(g);
    }
  } else {
    std::ifstream f(path, std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    ok = f.good() || f.eof();
    ok = ok && !raw.empty() && bunzip2(raw, mem);
  }
  if (!ok || mem.empty()) {
    if (err) *err = "cannot decompress " + path;
    return nullptr;
  }
  return fmemopen(mem.data(), mem.size(), "rb");
}

int open_mrc(const std::string& path, FILE** fp, MrcHeader& h, bool& swap, std::string* err,
             std::vector<char>& mem) {
  *fp = open_mrc_stream(path, mem, err);
  if (!*fp) {
    if (err && err->empty()) *err = "Cannot open MRC file " + path;
    return -1;
  }
  if (fread(&h, sizeof(h), 1, *fp) != 1) {
    if (err) *err = "short MRC header " + path;
    fclose(*fp);
    return -1;
  }
  const unsigned char* m2 = reinterpret_cast<const unsigned char*>(&h);
  const bool ord = (m2[0] == 0 && m2[1] == 0);
  swap = ord ^ is_machine_big_endian();
  if (swap)
    for (int i = 0; i < 56; ++i) swap4(&h.w[i]);
  const int nx = h.w[0], ny = h.w[1], nz = h.w[2], mode = h.w[3], nlabl = h.w[55];
  if (ny > 0xffff || nz > 0xffff || mode > 4 || mode < 0 || nlabl > 11 || nx <= 0 || ny <= 0 ||
      nz <= 0) {
    if (err) *err = "Invalid MRC file " + path;
    fclose(*fp);
    return -1;
  }
  if (ny == 1) {
    if (err) *err = "This appears to be a 1d file. I only read 2d files.";
    fclose(*fp);
    return -4;
  }
  return 0;
}
}  // namespace

int mrc_sections(const std::string& path) {
  FILE* fp;
  MrcHeader h;
  bool sw;
  std::vector<char> mem;
  if (open_mrc(path, &fp, h, sw, nullptr, mem)) return -1;
  fclose(fp);
  return h.w[2];
}

int read_mrc(const std::string& path, int n, Image& img, std::string* err) {
  FILE* fp;
  MrcHeader h;
  bool swap;
  std::vector<char> mem;   // decompressed file, when compressed
  int r = open_mrc(path, &fp, h, swap, err, mem);
  if (r) return r;
  const int nx = h.w[0], ny = h.w[1], nzf = h.w[2], mode = h.w[3], nsym = h.w[23];
  if (mode == 3 || mode == 4) {
    if (err) *err = "complex MRC modes are not images";
    fclose(fp);
    return -1;
  }
  int nz, startz;
  if (n >= 0 && n < nzf) { nz = 1; startz = n; }
  else { nz = nzf; startz = 0; }
  const int bpp = (mode == 0) ? 1 : (mode == 1 ? 2 : 4);
  const size_t npix = (size_t)nx * ny * nz;
  if (fseeko(fp, (off_t)1024 + nsym + (off_t)startz * nx * ny * bpp, SEEK_SET) != 0) {
    if (err) *err = "Incomplete data read " + path;
    fclose(fp);
    return -1;
  }
  img.nx = nx; img.ny = ny; img.nz = nz;
  img.data.resize(npix);
  if (mode == 2 && !swap) {   // native floats: straight into the image
    const size_t got = fread(img.data.data(), sizeof(float), npix, fp);
    fclose(fp);
    if (got != npix) {
      if (err) *err = "Incomplete data read " + path;
      return -1;
    }
    return 0;
  }
  std::vector<unsigned char> raw(npix * bpp);
  if (fread(raw.data(), 1, raw.size(), fp) != raw.size()) {
    if (err) *err = "Incomplete data read " + path;
    fclose(fp);
    return -1;
  }
  fclose(fp);
  for (size_t i = 0; i < npix; ++i) {
    if (mode == 0) {
      img.data[i] = (float)((double)raw[i] / 100.0 - 1.28);  // emdata_.cpp:503, in double
    } else if (mode == 1) {
      unsigned char b[2] = {raw[2 * i], raw[2 * i + 1]};
      if (swap) std::swap(b[0], b[1]);
      int16_t s;
      memcpy(&s, b, 2);
      img.data[i] = (float)s;
    } else {
      unsigned char b[4] = {raw[4 * i], raw[4 * i + 1], raw[4 * i + 2], raw[4 * i + 3]};
      if (swap) swap4(b);
      memcpy(&img.data[i], b, 4);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// HDF5 images: the HDF branches of emdata::readImage (EMReader/emdata_.cpp:174-199)
// ---------------------------------------------------------------------------
namespace {

struct Hdf {
  gshdf::File f;
  uint64_t root = 0;
  bool old = false;   // EMAN1 layout
};

int hdf_open(const std::string& path, Hdf& h, std::string* err) {
  if (!h.f.open(path, err) || !h.f.find("/", &h.root, err)) return -1;
  // EMhdf::init_test (EMReader/emhdf_.cpp:103-122): a root attribute
  // num_dataset marks the old layout
  gshdf::Value v;
  std::string e;
  h.old = h.f.attribute(h.root, "num_dataset", &v, &e);
  if (!e.empty()) {
    if (err) *err = path + ": " + e;
    return -1;
  }
  return 0;
}

bool numeric_name(const std::string& s) {
  if (s.empty() || s.size() > 9) return false;
  for (char ch : s)
    if (ch < '0' || ch > '9') return false;
  return true;
}

// number of images: the old layout's num_dataset (EMReader/emhdf_.cpp:229-233
// iterates the root datasets), the new layout's members of /MDF/images
int hdf_count(Hdf& h, std::string* err) {
  std::vector<std::pair<std::string, uint64_t>> mem;
  uint64_t g = h.root;
  if (!h.old && !h.f.find("/MDF/images", &g, err)) return -1;
  if (!h.f.members(g, mem, err)) return -1;
  int n = 0;
  for (const auto& m : mem)
    if (numeric_name(m.first)) n = std::max(n, atoi(m.first.c_str()) + 1);
  return n;
}

int hdf_image(Hdf& h, int idx, Image& img, std::string* err) {
  if (idx < 0) idx = 0;   // readHDF / readHDF2: a negative index reads image 0
  uint64_t obj;
  std::vector<uint64_t> dims;
  if (h.old) {
    // readHDF (emdata_.cpp:216-250) with EMhdf::read_dims (emhdf_.cpp:124-150):
    // dataset "<idx>" under the root, nx = dims[0], ny = dims[1], nz = dims[2]
    if (!h.f.find("/" + std::to_string(idx), &obj, err)) return -1;
    if (!h.f.read(obj, img.data, &dims, err)) return -1;
    if (dims.size() != 2 && dims.size() != 3) {
      if (err) *err = "HDF image " + std::to_string(idx) + " is neither 2D nor 3D";
      return -1;
    }
    img.nx = (int)dims[0];
    img.ny = (int)dims[1];
    img.nz = dims.size() == 3 ? (int)dims[2] : 1;
    return 0;
  }
  // readHDF2 (emdata_.cpp:330-380) with EMhdf2::read_header / read_data
  // (emhdf2_.cpp:245-363): the size from the EMAN.nx/ny/nz attributes of
  // /MDF/images/<idx>, the pixels from its dataset "image"
  const std::string grp = "/MDF/images/" + std::to_string(idx);
  if (!h.f.find(grp, &obj, err)) return -1;
  int n[3] = {0, 0, 0};
  const char* names[3] = {"EMAN.nx", "EMAN.ny", "EMAN.nz"};
  for (int k = 0; k < 3; ++k) {
    gshdf::Value v;
    std::string e;
    if (h.f.attribute(obj, names[k], &v, &e) && !v.num.empty()) n[k] = (int)v.num[0];
    else if (!e.empty()) {
      if (err) *err = grp + ": " + e;
      return -1;
    }
  }
  if (n[2] > 1) { img.nx = n[0]; img.ny = n[1]; img.nz = n[2]; }
  else if (n[1] > 1) { img.nx = n[0]; img.ny = n[1]; img.nz = 1; }
  else if (n[0] > 1) { img.nx = n[0]; img.ny = 1; img.nz = 1; }
  else {
    if (err) *err = grp + ": no EMAN.nx/ny/nz size";
    return -1;
  }
  if (!h.f.find(grp + "/image", &obj, err) || !h.f.read(obj, img.data, &dims, err)) return -1;
  if (img.data.size() != (size_t)img.nx * img.ny * img.nz) {
    if (err) *err = grp + "/image holds " + std::to_string(img.data.size()) +
                    " values, its header says " + std::to_string(img.nx) + "x" +
                    std::to_string(img.ny) + "x" + std::to_string(img.nz);
    return -1;
  }
  return 0;
}

}  // namespace

int image_count(const std::string& path, std::string* err) {
  if (!gshdf::is_hdf5(path)) {
    const int n = mrc_sections(path);
    if (n < 0 && err) *err = "cannot read " + path;
    return n;
  }
  Hdf h;
  if (hdf_open(path, h, err)) return -1;
  return hdf_count(h, err);
}

int read_image(const std::string& path, int n, Image& img, std::string* err) {
  if (!gshdf::is_hdf5(path)) return read_mrc(path, n, img, err);
  Hdf h;
  if (hdf_open(path, h, err)) return -1;
  if (n >= 0) return hdf_image(h, n, img, err);
  // the whole file: one image, or a stack of equal-size 2D images
  const int cnt = hdf_count(h, err);
  if (cnt <= 0) {
    if (err && cnt == 0) *err = path + " holds no images";
    return -1;
  }
  if (hdf_image(h, 0, img, err)) return -1;
  if (cnt == 1) return 0;
  if (img.nz != 1) {
    if (err) *err = path + ": a stack of 3D images";
    return -1;
  }
  const size_t per = (size_t)img.nx * img.ny;
  img.data.resize(per * cnt);
  Image one;
  for (int k = 1; k < cnt; ++k) {
    if (hdf_image(h, k, one, err)) return -1;
    if (one.nx != img.nx || one.ny != img.ny || one.nz != 1) {
      if (err) *err = path + ": images of different sizes in one stack";
      return -1;
    }
    std::copy(one.data.begin(), one.data.end(), img.data.begin() + per * k);
  }
  img.nz = cnt;
  return 0;
}

int read_images(const std::string& path, int count, std::vector<Image>& out, std::string* err) {
  out.clear();
  if (!gshdf::is_hdf5(path)) {
    Image all;
    if (read_mrc(path, -1, all, err)) return -1;
    if (all.nz < count) {
      if (err) *err = path + " holds " + std::to_string(all.nz) + " images";
      return -1;
    }
    const size_t per = (size_t)all.nx * all.ny;
    out.resize(count);
    for (int k = 0; k < count; ++k) {
      out[k].nx = all.nx;
      out[k].ny = all.ny;
      out[k].nz = 1;
      out[k].data.assign(all.data.begin() + per * k, all.data.begin() + per * (k + 1));
    }
    return 0;
  }
  Hdf h;
  if (hdf_open(path, h, err)) return -1;
  out.resize(count);
  for (int k = 0; k < count; ++k)
    if (hdf_image(h, k, out[k], err)) return -1;
  return 0;
}

int write_mrc(const std::string& path, const float* data, int nx, int ny, int nz) {
  MrcHeader h;
  memset(&h, 0, sizeof(h));
  h.w[0] = nx; h.w[1] = ny; h.w[2] = nz; h.w[3] = 2;
  h.w[7] = nx; h.w[8] = ny; h.w[9] = nz;
  float cell[3] = {(float)nx, (float)ny, (float)nz};
  memcpy(&h.w[10], cell, sizeof(cell));
  float ang[3] = {90, 90, 90};
  memcpy(&h.w[13], ang, sizeof(ang));
  h.w[16] = 1; h.w[17] = 2; h.w[18] = 3;
  memcpy(&h.w[52], "MAP ", 4);
  h.w[53] = 0x00004444;
  FILE* fp = fopen(path.c_str(), "wb");
  if (!fp) return -1;
  fwrite(&h, sizeof(h), 1, fp);
  fwrite(data, sizeof(float), (size_t)nx * ny * nz, fp);
  fclose(fp);
  return 0;
}

void write_pick_line(FILE* fp, int nn, const std::string& path, const MicEntry& m, float euler1,
                     float euler2, float psi, float cx, float cy, float score) {
  // main.cu:985-987: defocus is printed as (-1) * (negated lst value)
  const float neg = -m.defocus;
  fprintf(fp, "%d\t%s\tdefocus=%f\tdfdiff=%f\tdfang=%f\teuler=%f,%f,%f\tcenter=%f,%f\tscore=%f\n",
          nn, path.c_str(), (-1) * neg, m.dfdiff, m.dfang, euler1, euler2, psi, cx, cy, score);
}

void create_outpath(const std::string& path) {
  for (size_t pos = 0; pos < path.size(); ++pos) {
    if (path[pos] != '/' || pos == 0) continue;
    mkdir(path.substr(0, pos).c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
  }
}

}  // namespace gsio
