// gs_files.cpp -- C ABI of the image-file readers (host code; see gs_io.cpp
// for the MRC and EMAN HDF layouts and gs_hdf.cpp for the HDF5 parser).
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/gisspa.h"
#include "gs_hdf.hpp"
#include "gs_io.hpp"

namespace {
thread_local std::string g_err;

int io_fail(const std::string& m) {
  g_err = m;
  return GS_ERR_IO;
}

int gs_image_count_impl(const char* path, int* n) {
  if (!path || !n) return GS_ERR_ARG;
  std::string err;
  const int c = gsio::image_count(path, &err);
  if (c < 0) return io_fail(err.empty() ? std::string("cannot read ") + path : err);
  *n = c;
  return GS_OK;
}

int gs_read_image_impl(const char* path, int index, float* out, size_t cap, int* nx, int* ny,
                  int* nz) {
  if (!path || !nx || !ny || !nz) return GS_ERR_ARG;
  gsio::Image img;
  std::string err;
  if (gsio::read_image(path, index, img, &err))
    return io_fail(err.empty() ? std::string("cannot read ") + path : err);
  *nx = img.nx;
  *ny = img.ny;
  *nz = img.nz;
  if (!out) return GS_OK;
  if (cap < img.data.size()) return GS_ERR_CAPACITY;
  memcpy(out, img.data.data(), img.data.size() * sizeof(float));
  return GS_OK;
}

int gs_hdf5_read_impl(const char* path, const char* object, float* out, size_t cap, int* rank,
                 long long* dims, int max_rank) {
  if (!path || !object || !rank || (max_rank > 0 && !dims)) return GS_ERR_ARG;
  gshdf::File f;
  std::string err;
  uint64_t obj;
  std::vector<uint64_t> d;
  if (!f.open(path, &err) || !f.find(object, &obj, &err) || !f.shape(obj, &d, &err))
    return io_fail(err);
  *rank = (int)d.size();
  for (int k = 0; k < (int)d.size() && k < max_rank; ++k) dims[k] = (long long)d[k];
  if (!out) return GS_OK;
  std::vector<float> v;
  if (!f.read(obj, v, &d, &err)) return io_fail(err);
  if (cap < v.size()) return GS_ERR_CAPACITY;
  memcpy(out, v.data(), v.size() * sizeof(float));
  return GS_OK;
}

int gs_hdf5_attr_impl(const char* path, const char* object, const char* name, double* num, size_t cap,
                 size_t* n_num, char* str, size_t str_cap, int* is_string) {
  if (!path || !object || !name || !n_num || !is_string) return GS_ERR_ARG;
  gshdf::File f;
  std::string err;
  uint64_t obj;
  gshdf::Value v;
  if (!f.open(path, &err) || !f.find(object, &obj, &err)) return io_fail(err);
  if (!f.attribute(obj, name, &v, &err))
    return io_fail(err.empty() ? std::string("no attribute ") + name + " on " + object : err);
  *is_string = v.type.cls == 3;
  *n_num = v.num.size();
  if (*is_string) {
    if (!str || str_cap < v.str.size() + 1) return GS_ERR_CAPACITY;
    memcpy(str, v.str.c_str(), v.str.size() + 1);
    return GS_OK;
  }
  if (cap < v.num.size() || (!num && !v.num.empty())) return GS_ERR_CAPACITY;
  if (!v.num.empty()) memcpy(num, v.num.data(), v.num.size() * sizeof(double));
  return GS_OK;
}

}  // namespace

// a corrupt file can ask for absurd allocations: report them as I/O errors
#define GS_IO_GUARD(call)                          \
  try {                                            \
    return call;                                   \
  } catch (const std::exception& e) {              \
    return io_fail(std::string("reading failed: ") + e.what()); \
  }

extern "C" {

const char* gs_io_error(void) { return g_err.c_str(); }

int gs_image_count(const char* path, int* n) { GS_IO_GUARD(gs_image_count_impl(path, n)) }

int gs_read_image(const char* path, int index, float* out, size_t cap, int* nx, int* ny,
                  int* nz) {
  GS_IO_GUARD(gs_read_image_impl(path, index, out, cap, nx, ny, nz))
}

int gs_hdf5_read(const char* path, const char* object, float* out, size_t cap, int* rank,
                 long long* dims, int max_rank) {
  GS_IO_GUARD(gs_hdf5_read_impl(path, object, out, cap, rank, dims, max_rank))
}

int gs_hdf5_attr(const char* path, const char* object, const char* name, double* num, size_t cap,
                 size_t* n_num, char* str, size_t str_cap, int* is_string) {
  GS_IO_GUARD(gs_hdf5_attr_impl(path, object, name, num, cap, n_num, str, str_cap, is_string))
}

}  // extern "C"
