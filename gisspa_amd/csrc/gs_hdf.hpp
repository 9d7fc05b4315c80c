// gs_hdf.hpp -- read-only HDF5 access for image files (no libhdf5).
//
// The reference reads templates and micrographs through libhdf5 1.8
// (EMReader/emhdf_.cpp, EMReader/emhdf2_.cpp, dispatched by
// emdata::readImage, EMReader/emdata_.cpp:174-205).  This parser covers what
// those files hold, following the published HDF5 file format specification
// (version 3.0):
//   * superblock versions 0-3, a user block before it (base address);
//   * object headers version 1 and 2, with continuation blocks;
//   * groups stored as symbol tables (v1 B-tree + local heap) or as compact
//     link messages (dense link storage is reported as unsupported);
//   * attributes (message versions 1-3) of fixed-point, floating-point and
//     fixed-length string types, scalar or simple dataspaces;
//   * datasets with compact, contiguous or chunked layout (layout message
//     versions 1-3: v1 B-tree chunk index; version 4: contiguous, compact,
//     single-chunk and implicit indexes), filters deflate, shuffle and
//     fletcher32;
//   * element types: IEEE float 16/32/64 and 8-64 bit integers, either byte
//     order, converted to float.
#pragma once
#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace gshdf {

// True if the file starts with the HDF5 signature at 0, 512, 1024, ... (a
// user block may precede the superblock).
bool is_hdf5(const std::string& path);

struct Type {
  int cls = -1;        // 0 fixed-point, 1 floating point, 3 string
  int size = 0;        // bytes per element
  bool big = false;    // byte order
  bool sign = false;   // fixed-point signedness
};

struct Value {         // an attribute's value
  Type type;
  std::vector<uint64_t> dims;   // empty for a scalar
  std::vector<double> num;      // numeric elements
  std::string str;              // string value (padding removed)
};

class File {
 public:
  File() = default;
  ~File();
  File(const File&) = delete;
  File& operator=(const File&) = delete;

  bool open(const std::string& path, std::string* err);
  // object header address of a path ("/", "/MDF/images/0/image")
  bool find(const std::string& path, uint64_t* obj, std::string* err);
  // (name, object header address) of a group's members
  bool members(uint64_t group, std::vector<std::pair<std::string, uint64_t>>& out,
               std::string* err);
  // an attribute of an object; returns false with *err empty if absent
  bool attribute(uint64_t obj, const std::string& name, Value* out, std::string* err);
  // a dataset's shape (slowest dimension first)
  bool shape(uint64_t obj, std::vector<uint64_t>* dims, std::string* err);
  // a dataset's elements converted to float, in storage order (row-major)
  bool read(uint64_t obj, std::vector<float>& out, std::vector<uint64_t>* dims, std::string* err);

 private:
  struct Msg {
    int type;
    std::vector<uint8_t> data;
  };
  bool bytes(uint64_t addr, size_t n, std::vector<uint8_t>& out);
  bool header(uint64_t addr, std::vector<Msg>& msgs, std::string* err);
  struct Chunk {
    uint64_t addr;
    uint32_t size, mask;
    std::vector<uint64_t> off;   // element offsets, one per dimension (+ element size)
  };
  // level_want: the level the node must have (its parent's minus one; -1 at
  // the root)
  bool group_btree(uint64_t node, const std::vector<uint8_t>& heap,
                   std::vector<std::pair<std::string, uint64_t>>& out, int level_want,
                   std::string* err);
  bool chunk_btree(uint64_t node, int ndim, std::vector<Chunk>& chunks, int level_want,
                   std::string* err);

  std::map<uint64_t, std::vector<std::pair<std::string, uint64_t>>> members_cache_;
  FILE* fp_ = nullptr;
  uint64_t base_ = 0, size_ = 0, root_ = 0;
  int so_ = 8, sl_ = 8;   // sizes of offsets and lengths
};

}  // namespace gshdf
