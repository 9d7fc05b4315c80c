// remove_repeats -- duplicate removal of a GisSPA pick list on the GPU.
//
//   remove_repeats <input1> <number of images> <center thres> <euler thres> <output> [gpu_id]
//
// Same arguments, input and output as
// python_scripts/remove_repeat_particles_from_list.py (usage at :151-166):
// reads the picker's output lines (main.cu:985 format), groups them into
// runs of equal micrograph path (tab field 2), and writes the lines that
// survive, unchanged and in input order.  The pair test and the walk run in
// libgisspa (gs_remove_repeats).  Differences: a malformed line, more groups
// than <number of images>, or an empty input end the run with a message and
// status 1 (the script raises on each).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gisspa.h"

namespace {

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t a = 0;
  for (;;) {
    const size_t b = s.find(sep, a);
    out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return out;
}

// float() of a Python string: surrounding white space allowed, nothing else.
bool to_double(const std::string& s, double* v) {
  const char* p = s.c_str();
  char* end = nullptr;
  errno = 0;
  *v = std::strtod(p, &end);
  if (end == p) return false;
  while (*end == ' ' || *end == '\t' || *end == '\n' || *end == '\r') ++end;
  return *end == '\0';
}

// "key=a,b,..." -> component k of the value (:41-46)
bool field_value(const std::vector<std::string>& f, size_t idx, size_t k, double* v) {
  if (idx >= f.size()) return false;
  const auto kv = split(f[idx], '=');
  if (kv.size() < 2) return false;
  const auto parts = split(kv[1], ',');
  if (k >= parts.size()) return false;
  return to_double(parts[k], v);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "<input1> <number of images> <center> <euler_thres> <output> [gpu_id]\n");
    return 1;
  }
  const char* in_path = argv[1];
  const long num = std::strtol(argv[2], nullptr, 10);
  const double center = std::strtod(argv[3], nullptr);
  const double euler_thres = std::strtod(argv[4], nullptr);
  const char* out_path = argv[5];
  const int device = argc > 6 ? std::atoi(argv[6]) : 0;

  FILE* fi = std::fopen(in_path, "rb");
  if (!fi) {
    std::fprintf(stderr, "cannot open %s\n", in_path);
    return 1;
  }
  std::vector<std::string> lines;
  {
    std::string cur;
    int ch;
    while ((ch = std::fgetc(fi)) != EOF) {
      cur.push_back((char)ch);
      if (ch == '\n') {
        lines.push_back(cur);
        cur.clear();
      }
    }
    if (!cur.empty()) lines.push_back(cur);
    std::fclose(fi);
  }
  if (lines.empty()) {
    std::fprintf(stderr, "%s: no pick lines\n", in_path);
    return 1;
  }

  std::vector<gs_pick_rec> recs(lines.size());
  std::string group_path;
  int group = -1;
  for (size_t i = 0; i < lines.size(); ++i) {
    std::string body = lines[i];
    const auto f = split(body, '\t');
    gs_pick_rec& r = recs[i];
    std::memset(&r, 0, sizeof(r));
    if (f.size() < 8 || !field_value(f, 6, 0, &r.x) || !field_value(f, 6, 1, &r.y) ||
        !field_value(f, 5, 0, &r.alt) || !field_value(f, 5, 1, &r.az) ||
        !field_value(f, 5, 2, &r.phi) || !field_value(f, 7, 0, &r.score)) {
      std::fprintf(stderr, "%s:%zu: not a pick line\n", in_path, i + 1);
      return 1;
    }
    if (i == 0 || f[1] != group_path) {  // :22-36
      group_path = f[1];
      ++group;
    }
    r.group = group;
  }
  if (group + 1 > num) {
    std::fprintf(stderr, "%d micrograph groups in %s exceed <number of images> = %ld\n",
                 group + 1, in_path, num);
    return 1;
  }

  std::vector<unsigned char> keep(lines.size());
  size_t kept = 0;
  const int rc = gs_remove_repeats(device, recs.data(), recs.size(), center, euler_thres,
                                   keep.data(), &kept);
  if (rc != GS_OK) {
    std::fprintf(stderr, "gs_remove_repeats: %s\n", gs_strerror(rc));
    return 1;
  }
  FILE* fo = std::fopen(out_path, "wb");
  if (!fo) {
    std::fprintf(stderr, "cannot open %s\n", out_path);
    return 1;
  }
  for (size_t i = 0; i < lines.size(); ++i)
    if (keep[i]) std::fwrite(lines[i].data(), 1, lines[i].size(), fo);
  if (std::fclose(fo) != 0) {
    std::fprintf(stderr, "cannot write %s\n", out_path);
    return 1;
  }
  std::fprintf(stderr, "%zu of %zu picks kept\n", kept, lines.size());
  return 0;
}
