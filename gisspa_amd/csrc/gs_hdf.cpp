// gs_hdf.cpp -- read-only HDF5 parser (file format specification 3.0).
#include "gs_hdf.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace gshdf {
namespace {

const uint8_t kSig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
const int kMaxDepth = 64;

// little-endian cursor over a byte range; any overrun clears ok
struct Cur {
  const uint8_t* p;
  const uint8_t* e;
  bool ok = true;
  Cur(const uint8_t* b, size_t n) : p(b), e(b + n) {}
  explicit Cur(const std::vector<uint8_t>& v) : p(v.data()), e(v.data() + v.size()) {}
  size_t left() const { return (size_t)(e - p); }
  uint64_t u(int n) {
    if (n < 0 || n > 8 || (size_t)n > left()) {
      ok = false;
      p = e;
      return 0;
    }
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v |= (uint64_t)p[i] << (8 * i);
    p += n;
    return v;
  }
  const uint8_t* take(size_t n) {
    if (n > left()) {
      ok = false;
      p = e;
      return nullptr;
    }
    const uint8_t* q = p;
    p += n;
    return q;
  }
};

bool fail(std::string* err, const std::string& m) {
  if (err) *err = m;
  return false;
}

bool undefined(uint64_t a, int so) {
  return so >= 8 ? a == ~0ull : a == ((1ull << (8 * so)) - 1);
}

size_t pad8(size_t n) { return (n + 7) & ~size_t(7); }

// Datatype message (type 0x0003)
bool decode_type(Cur& c, Type* t, std::string* err) {
  const int cv = (int)c.u(1);
  const uint32_t bits = (uint32_t)c.u(3);
  const uint32_t size = (uint32_t)c.u(4);
  if (!c.ok) return fail(err, "truncated datatype message");
  t->cls = cv & 15;
  t->size = (int)size;
  t->big = (bits & 1) != 0;
  t->sign = false;
  if (t->cls == 0) {
    t->sign = (bits & 8) != 0;
    c.u(2);
    c.u(2);   // bit offset, precision
    if (size != 1 && size != 2 && size != 4 && size != 8)
      return fail(err, "integer type of " + std::to_string(size) + " bytes");
  } else if (t->cls == 1) {
    if (bits & 0x40) return fail(err, "VAX floating-point byte order");
    const int off = (int)c.u(2), prec = (int)c.u(2);
    const int eloc = (int)c.u(1), esz = (int)c.u(1), mloc = (int)c.u(1), msz = (int)c.u(1);
    c.u(4);   // exponent bias
    const bool ieee = off == 0 && mloc == 0 && eloc == msz &&
                      ((size == 2 && prec == 16 && esz == 5 && msz == 10) ||
                       (size == 4 && prec == 32 && esz == 8 && msz == 23) ||
                       (size == 8 && prec == 64 && esz == 11 && msz == 52));
    if (!ieee) return fail(err, "floating-point type that is not IEEE half/single/double");
  } else if (t->cls != 3) {
    return fail(err, "unsupported datatype class " + std::to_string(t->cls));
  }
  if (!c.ok) return fail(err, "truncated datatype message");
  return true;
}

// Dataspace message (type 0x0001)
bool decode_space(Cur& c, int sl, std::vector<uint64_t>* dims, bool* null_space,
                  std::string* err) {
  const int ver = (int)c.u(1), rank = (int)c.u(1);
  c.u(1);   // flags (maximum dimensions / permutation follow; not needed)
  *null_space = false;
  if (ver == 1) {
    c.take(5);
  } else if (ver == 2) {
    *null_space = c.u(1) == 2;
  } else {
    return fail(err, "dataspace message version " + std::to_string(ver));
  }
  if (rank > 32) return fail(err, "dataspace rank " + std::to_string(rank));
  dims->assign(rank, 0);
  for (int i = 0; i < rank; ++i) (*dims)[i] = c.u(sl);
  if (!c.ok) return fail(err, "truncated dataspace message");
  return true;
}

// element count; ~0 when it exceeds 2^40 (a corrupt or absurd dataspace)
uint64_t count_of(const std::vector<uint64_t>& dims, bool null_space) {
  if (null_space) return 0;
  uint64_t n = 1;
  for (uint64_t d : dims) {
    if (d != 0 && n > (1ull << 40) / d) return ~0ull;
    n *= d;
  }
  return n;
}

// largest dataset read: 2^32 elements (16 GiB as float)
const uint64_t kMaxElements = 1ull << 32;

double half_to_double(uint16_t h) {
  const int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
  double v;
  if (e == 0) v = std::ldexp((double)m, -24);
  else if (e == 31) v = m ? NAN : INFINITY;
  else v = std::ldexp((double)(m | 1024), e - 25);
  return s ? -v : v;
}

// one element of type t at b, as double
double element(const uint8_t* b, const Type& t) {
  uint8_t le[8];
  for (int i = 0; i < t.size; ++i) le[i] = t.big ? b[t.size - 1 - i] : b[i];
  uint64_t v = 0;
  for (int i = 0; i < t.size; ++i) v |= (uint64_t)le[i] << (8 * i);
  if (t.cls == 1) {
    if (t.size == 4) {
      const uint32_t w = (uint32_t)v;
      float f;
      memcpy(&f, &w, 4);
      return f;
    }
    if (t.size == 8) {
      double d;
      memcpy(&d, &v, 8);
      return d;
    }
    return half_to_double((uint16_t)v);
  }
  if (t.sign && t.size < 8 && (v >> (8 * t.size - 1)) & 1) v |= ~0ull << (8 * t.size);
  return t.sign ? (double)(int64_t)v : (double)v;
}

void convert(const uint8_t* src, size_t n, const Type& t, float* dst) {
  if (t.cls == 1 && t.size == 4 && !t.big) {
    memcpy(dst, src, n * 4);
    return;
  }
  for (size_t i = 0; i < n; ++i) dst[i] = (float)element(src + i * t.size, t);
}

struct Filter {
  int id;
  std::vector<uint32_t> cd;
};

// Filter pipeline message (type 0x000B)
bool decode_filters(Cur& c, std::vector<Filter>& out, std::string* err) {
  const int ver = (int)c.u(1), nf = (int)c.u(1);
  if (ver == 1) c.take(6);
  else if (ver != 2) return fail(err, "filter pipeline version " + std::to_string(ver));
  out.clear();
  for (int i = 0; i < nf && c.ok; ++i) {
    Filter f;
    f.id = (int)c.u(2);
    const size_t nlen = (ver == 1 || f.id >= 256) ? (size_t)c.u(2) : 0;
    c.u(2);   // flags
    const int nv = (int)c.u(2);
    c.take(ver == 1 ? pad8(nlen) : nlen);
    for (int k = 0; k < nv; ++k) f.cd.push_back((uint32_t)c.u(4));
    if (ver == 1 && (nv & 1)) c.u(4);
    out.push_back(f);
  }
  if (!c.ok) return fail(err, "truncated filter pipeline message");
  return true;
}

// undo the filters of a chunk (pipeline order reversed), skipping those the
// chunk's filter mask excludes
bool unfilter(std::vector<uint8_t>& buf, const std::vector<Filter>& fl, uint32_t mask,
              size_t expect, int esize, std::string* err) {
  for (int i = (int)fl.size() - 1; i >= 0; --i) {
    if ((mask >> i) & 1) continue;
    const Filter& f = fl[i];
    if (f.id == 1) {   // deflate (zlib stream)
      std::vector<uint8_t> out(expect);
      uLongf n = (uLongf)expect;
      if (uncompress(out.data(), &n, buf.data(), (uLong)buf.size()) != Z_OK)
        return fail(err, "deflate chunk does not decompress");
      out.resize(n);
      buf.swap(out);
    } else if (f.id == 2) {   // shuffle: byte planes back to elements
      const size_t es = f.cd.empty() ? (size_t)esize : f.cd[0];
      if (es > 1) {
        const size_t n = buf.size() / es;
        std::vector<uint8_t> out(buf.size());
        for (size_t b = 0; b < es; ++b)
          for (size_t k = 0; k < n; ++k) out[k * es + b] = buf[b * n + k];
        std::copy(buf.begin() + n * es, buf.end(), out.begin() + n * es);
        buf.swap(out);
      }
    } else if (f.id == 3) {   // fletcher32: checksum appended
      if (buf.size() < 4) return fail(err, "fletcher32 chunk shorter than its checksum");
      buf.resize(buf.size() - 4);
    } else {
      return fail(err, "unsupported filter " + std::to_string(f.id));
    }
  }
  return true;
}

}  // namespace

bool is_hdf5(const std::string& path) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  uint8_t b[8];
  bool found = false;
  for (long off = 0; off <= (1l << 30) && !found; off = off ? off * 2 : 512) {
    if (fseeko(f, off, SEEK_SET) != 0 || fread(b, 1, 8, f) != 8) break;
    found = memcmp(b, kSig, 8) == 0;
  }
  fclose(f);
  return found;
}

File::~File() {
  if (fp_) fclose(fp_);
}

bool File::bytes(uint64_t addr, size_t n, std::vector<uint8_t>& out) {
  const uint64_t a = base_ + addr;
  if (undefined(addr, so_) || a > size_ || n > size_ - a) return false;
  out.resize(n);
  if (n == 0) return true;
  return fseeko(fp_, (off_t)a, SEEK_SET) == 0 && fread(out.data(), 1, n, fp_) == n;
}

bool File::open(const std::string& path, std::string* err) {
  fp_ = fopen(path.c_str(), "rb");
  if (!fp_) return fail(err, "cannot open " + path);
  fseeko(fp_, 0, SEEK_END);
  size_ = (uint64_t)ftello(fp_);
  uint64_t sb = ~0ull;
  uint8_t b[8];
  for (uint64_t off = 0; off + 8 <= size_; off = off ? off * 2 : 512) {
    if (fseeko(fp_, (off_t)off, SEEK_SET) != 0 || fread(b, 1, 8, fp_) != 8) break;
    if (memcmp(b, kSig, 8) == 0) {
      sb = off;
      break;
    }
  }
  if (sb == ~0ull) return fail(err, path + " is not an HDF5 file");
  std::vector<uint8_t> h(std::min<uint64_t>(128, size_ - sb));
  fseeko(fp_, (off_t)sb, SEEK_SET);
  if (fread(h.data(), 1, h.size(), fp_) != h.size() || h.size() < 24)
    return fail(err, path + ": truncated superblock");
  const int ver = h[8];
  uint64_t base = 0;
  if (ver == 0 || ver == 1) {
    so_ = h[13];
    sl_ = h[14];
    Cur c(h.data() + 16, h.size() - 16);
    c.u(2);
    c.u(2);
    c.u(4);   // group K values, consistency flags
    if (ver == 1) {
      c.u(2);
      c.u(2);
    }
    base = c.u(so_);
    c.u(so_);   // free-space info
    c.u(so_);   // end of file
    c.u(so_);   // driver info
    c.u(so_);   // root entry: link name offset
    root_ = c.u(so_);
    if (!c.ok) return fail(err, path + ": truncated superblock");
  } else if (ver == 2 || ver == 3) {
    so_ = h[9];
    sl_ = h[10];
    Cur c(h.data() + 12, h.size() - 12);
    base = c.u(so_);
    c.u(so_);   // superblock extension
    c.u(so_);   // end of file
    root_ = c.u(so_);
    if (!c.ok) return fail(err, path + ": truncated superblock");
  } else {
    return fail(err, path + ": superblock version " + std::to_string(ver));
  }
  if ((so_ != 2 && so_ != 4 && so_ != 8) || (sl_ != 2 && sl_ != 4 && sl_ != 8))
    return fail(err, path + ": unsupported offset/length sizes");
  base_ = base;
  return true;
}

bool File::header(uint64_t addr, std::vector<Msg>& msgs, std::string* err) {
  msgs.clear();
  std::vector<uint8_t> b;
  if (!bytes(addr, 16, b)) return fail(err, "object header outside the file");
  struct Block {
    uint64_t addr, len;
  };
  std::vector<Block> blocks;
  bool v2 = false;
  int oflags = 0;
  uint64_t nmsg = ~0ull;
  if (memcmp(b.data(), "OHDR", 4) == 0) {
    v2 = true;
    std::vector<uint8_t> p;
    if (!bytes(addr, 40, p) && !bytes(addr, 16, p)) return fail(err, "truncated object header");
    Cur c(p);
    c.take(4);
    if (c.u(1) != 2) return fail(err, "object header version");
    oflags = (int)c.u(1);
    if (oflags & 0x20) c.take(16);
    if (oflags & 0x10) c.take(4);
    const uint64_t len = c.u(1 << (oflags & 3));
    if (!c.ok) return fail(err, "truncated object header");
    blocks.push_back({addr + (uint64_t)(c.p - p.data()), len});
  } else if (b[0] == 1) {
    nmsg = b[2] | (uint64_t)b[3] << 8;
    const uint64_t len = b[8] | (uint64_t)b[9] << 8 | (uint64_t)b[10] << 16 |
                         (uint64_t)b[11] << 24;
    blocks.push_back({addr + 16, len});
  } else {
    return fail(err, "unknown object header version " + std::to_string(b[0]));
  }
  uint64_t seen = 0;
  for (size_t k = 0; k < blocks.size() && k < 4096; ++k) {
    std::vector<uint8_t> blk;
    if (!bytes(blocks[k].addr, blocks[k].len, blk)) return fail(err, "object header block outside the file");
    Cur c(blk);
    if (v2 && k > 0) {
      if (blk.size() < 8 || memcmp(blk.data(), "OCHK", 4) != 0)
        return fail(err, "object header continuation without its signature");
      c = Cur(blk.data() + 4, blk.size() - 8);
    }
    const size_t hdr = v2 ? ((oflags & 4) ? 6 : 4) : 8;
    while (c.left() >= hdr && seen < nmsg) {
      int type;
      size_t size;
      if (v2) {
        type = (int)c.u(1);
        size = (size_t)c.u(2);
        c.u(1);
        if (oflags & 4) c.u(2);
      } else {
        type = (int)c.u(2);
        size = (size_t)c.u(2);
        c.u(1);
        c.take(3);
      }
      const uint8_t* d = c.take(size);
      if (!c.ok) return fail(err, "object header message overruns its block");
      ++seen;
      if (type == 0x10) {   // continuation
        Cur q(d, size);
        const uint64_t off = q.u(so_), len = q.u(sl_);
        if (!q.ok) return fail(err, "truncated continuation message");
        blocks.push_back({off, len});
      } else if (type != 0) {
        msgs.push_back({type, std::vector<uint8_t>(d, d + size)});
      }
    }
  }
  return true;
}

bool File::group_btree(uint64_t node, const std::vector<uint8_t>& heap,
                       std::vector<std::pair<std::string, uint64_t>>& out, int level_want,
                       std::string* err) {
  std::vector<uint8_t> h;
  const size_t hs = 8 + 2 * (size_t)so_;
  if (!bytes(node, hs, h) || memcmp(h.data(), "TREE", 4) != 0 || h[4] != 0)
    return fail(err, "bad group B-tree node");
  const int level = h[5];
  // children sit exactly one level lower, so a malformed tree cannot loop
  if ((level_want >= 0 && level != level_want) || level > kMaxDepth)
    return fail(err, "group B-tree levels out of order");
  const size_t entries = h[6] | (size_t)h[7] << 8;
  std::vector<uint8_t> n;
  if (!bytes(node, hs + (entries + 1) * sl_ + entries * so_, n))
    return fail(err, "group B-tree node outside the file");
  Cur c(n.data() + hs, n.size() - hs);
  for (size_t i = 0; i < entries; ++i) {
    c.u(sl_);   // key (heap offset of the left-most name)
    const uint64_t child = c.u(so_);
    if (level > 0) {
      if (!group_btree(child, heap, out, level - 1, err)) return false;
      continue;
    }
    std::vector<uint8_t> s;
    if (!bytes(child, 8, s) || memcmp(s.data(), "SNOD", 4) != 0)
      return fail(err, "bad symbol table node");
    const size_t nsym = s[6] | (size_t)s[7] << 8;
    const size_t es = 2 * (size_t)so_ + 24;
    if (!bytes(child, 8 + nsym * es, s)) return fail(err, "symbol table node outside the file");
    Cur e(s.data() + 8, nsym * es);
    if (out.size() + nsym > (1u << 24)) return fail(err, "group with more than 2^24 members");
    for (size_t k = 0; k < nsym; ++k) {
      const uint64_t name_off = e.u(so_), obj = e.u(so_);
      e.take(24);
      if (name_off >= heap.size()) return fail(err, "link name outside the local heap");
      const char* nm = (const char*)heap.data() + name_off;
      out.emplace_back(std::string(nm, strnlen(nm, heap.size() - name_off)), obj);
    }
  }
  return true;
}

bool File::members(uint64_t group, std::vector<std::pair<std::string, uint64_t>>& out,
                   std::string* err) {
  out.clear();
  auto it = members_cache_.find(group);
  if (it != members_cache_.end()) {
    out = it->second;
    return true;
  }
  std::vector<Msg> msgs;
  if (!header(group, msgs, err)) return false;
  bool is_group = false;
  for (const Msg& m : msgs) {
    Cur c(m.data);
    if (m.type == 0x11) {   // symbol table: v1 B-tree + local heap
      is_group = true;
      const uint64_t bt = c.u(so_), hp = c.u(so_);
      std::vector<uint8_t> h;
      if (!c.ok || !bytes(hp, 8 + 2 * (size_t)sl_ + so_, h) || memcmp(h.data(), "HEAP", 4) != 0)
        return fail(err, "bad local heap");
      Cur q(h.data() + 8, h.size() - 8);
      const uint64_t dsize = q.u(sl_);
      q.u(sl_);
      const uint64_t daddr = q.u(so_);
      std::vector<uint8_t> heap;
      if (!bytes(daddr, dsize, heap)) return fail(err, "local heap data outside the file");
      if (!group_btree(bt, heap, out, -1, err)) return false;
    } else if (m.type == 0x06) {   // link (compact storage)
      is_group = true;
      if (c.u(1) != 1) return fail(err, "link message version");
      const int fl = (int)c.u(1);
      const int ltype = (fl & 8) ? (int)c.u(1) : 0;
      if (fl & 4) c.u(8);
      if (fl & 16) c.u(1);
      const size_t nlen = (size_t)c.u(1 << (fl & 3));
      const uint8_t* nm = c.take(nlen);
      if (!c.ok) return fail(err, "truncated link message");
      if (ltype != 0) continue;   // soft / external links are not followed
      const uint64_t obj = c.u(so_);
      if (!c.ok) return fail(err, "truncated link message");
      out.emplace_back(std::string((const char*)nm, nlen), obj);
    } else if (m.type == 0x02) {   // link info
      is_group = true;
      c.u(1);
      const int fl = (int)c.u(1);
      if (fl & 1) c.u(8);
      const uint64_t fheap = c.u(so_);
      if (c.ok && !undefined(fheap, so_))
        return fail(err, "group with dense link storage (fractal heap) is not supported");
    }
  }
  if (!is_group) return fail(err, "object is not a group");
  members_cache_[group] = out;
  return true;
}

bool File::find(const std::string& path, uint64_t* obj, std::string* err) {
  uint64_t cur = root_;
  size_t i = 0;
  while (i < path.size()) {
    while (i < path.size() && path[i] == '/') ++i;
    if (i >= path.size()) break;
    size_t j = path.find('/', i);
    if (j == std::string::npos) j = path.size();
    const std::string comp = path.substr(i, j - i);
    std::vector<std::pair<std::string, uint64_t>> mem;
    if (!members(cur, mem, err)) return false;
    bool hit = false;
    for (const auto& m : mem)
      if (m.first == comp) {
        cur = m.second;
        hit = true;
        break;
      }
    if (!hit) return fail(err, "no object " + path.substr(0, j));
    i = j;
  }
  *obj = cur;
  return true;
}

bool File::attribute(uint64_t obj, const std::string& name, Value* out, std::string* err) {
  if (err) err->clear();
  std::vector<Msg> msgs;
  if (!header(obj, msgs, err)) return false;
  for (const Msg& m : msgs) {
    if (m.type == 0x15) {   // attribute info: dense attribute storage
      Cur c(m.data);
      c.u(1);
      const int fl = (int)c.u(1);
      if (fl & 1) c.u(2);
      const uint64_t fheap = c.u(so_);
      if (c.ok && !undefined(fheap, so_))
        return fail(err, "dense attribute storage is not supported");
      continue;
    }
    if (m.type != 0x0C) continue;
    Cur c(m.data);
    const int ver = (int)c.u(1);
    const int fl = (int)c.u(1);
    const size_t nsz = (size_t)c.u(2), tsz = (size_t)c.u(2), ssz = (size_t)c.u(2);
    if (ver == 3) c.u(1);
    if (ver < 1 || ver > 3) return fail(err, "attribute message version " + std::to_string(ver));
    const bool padded = ver == 1;
    const uint8_t* nm = c.take(padded ? pad8(nsz) : nsz);
    const uint8_t* tp = c.take(padded ? pad8(tsz) : tsz);
    const uint8_t* sp = c.take(padded ? pad8(ssz) : ssz);
    if (!c.ok) return fail(err, "truncated attribute message");
    const std::string an((const char*)nm, strnlen((const char*)nm, nsz));
    if (an != name) continue;
    if (ver >= 2 && (fl & 3)) return fail(err, "shared attribute type/space is not supported");
    Cur tc(tp, tsz), sc(sp, ssz);
    if (!decode_type(tc, &out->type, err)) return false;
    bool null_space = false;
    if (!decode_space(sc, sl_, &out->dims, &null_space, err)) return false;
    const uint64_t n = count_of(out->dims, null_space);
    if (n > (1ull << 24)) return fail(err, "attribute too large");
    const uint8_t* data = c.take(n * out->type.size);
    if (!c.ok) return fail(err, "attribute data shorter than its type and space");
    out->num.clear();
    out->str.clear();
    if (out->type.cls == 3) {
      std::string s((const char*)data, n * out->type.size);
      s = s.substr(0, strnlen(s.c_str(), s.size()));
      while (!s.empty() && s.back() == ' ') s.pop_back();
      out->str = s;
    } else {
      for (uint64_t k = 0; k < n; ++k) out->num.push_back(element(data + k * out->type.size, out->type));
    }
    return true;
  }
  return false;
}

bool File::shape(uint64_t obj, std::vector<uint64_t>* dims, std::string* err) {
  std::vector<Msg> msgs;
  if (!header(obj, msgs, err)) return false;
  for (const Msg& m : msgs)
    if (m.type == 0x01) {
      Cur c(m.data);
      bool null_space = false;
      return decode_space(c, sl_, dims, &null_space, err);
    }
  return fail(err, "object is not a dataset (no dataspace)");
}

bool File::chunk_btree(uint64_t node, int ndim, std::vector<Chunk>& chunks, int level_want,
                       std::string* err) {
  std::vector<uint8_t> h;
  const size_t hs = 8 + 2 * (size_t)so_;
  if (!bytes(node, hs, h) || memcmp(h.data(), "TREE", 4) != 0 || h[4] != 1)
    return fail(err, "bad chunk B-tree node");
  const int level = h[5];
  if ((level_want >= 0 && level != level_want) || level > kMaxDepth)
    return fail(err, "chunk B-tree levels out of order");
  const size_t entries = h[6] | (size_t)h[7] << 8;
  const size_t ks = 8 + 8 * (size_t)ndim;
  std::vector<uint8_t> n;
  if (!bytes(node, hs + (entries + 1) * ks + entries * so_, n))
    return fail(err, "chunk B-tree node outside the file");
  Cur c(n.data() + hs, n.size() - hs);
  for (size_t i = 0; i < entries; ++i) {
    Chunk ch;
    ch.size = (uint32_t)c.u(4);
    ch.mask = (uint32_t)c.u(4);
    ch.off.resize(ndim);
    for (int d = 0; d < ndim; ++d) ch.off[d] = c.u(8);
    ch.addr = c.u(so_);
    if (!c.ok) return fail(err, "truncated chunk B-tree node");
    if (level > 0) {
      if (!chunk_btree(ch.addr, ndim, chunks, level - 1, err)) return false;
    } else {
      chunks.push_back(ch);
    }
  }
  return true;
}

bool File::read(uint64_t obj, std::vector<float>& out, std::vector<uint64_t>* dims_out,
                std::string* err) {
  std::vector<Msg> msgs;
  if (!header(obj, msgs, err)) return false;
  const Msg *ms = nullptr, *mt = nullptr, *ml = nullptr, *mf = nullptr;
  for (const Msg& m : msgs) {
    if (m.type == 0x01) ms = &m;
    else if (m.type == 0x03) mt = &m;
    else if (m.type == 0x08) ml = &m;
    else if (m.type == 0x0B) mf = &m;
  }
  if (!ms || !mt || !ml) return fail(err, "object is not a dataset");
  Type t;
  std::vector<uint64_t> dims;
  bool null_space = false;
  {
    Cur c(mt->data);
    if (!decode_type(c, &t, err)) return false;
    if (t.cls == 3) return fail(err, "string dataset");
    Cur s(ms->data);
    if (!decode_space(s, sl_, &dims, &null_space, err)) return false;
  }
  std::vector<Filter> filters;
  if (mf) {
    Cur c(mf->data);
    if (!decode_filters(c, filters, err)) return false;
  }
  const uint64_t n = count_of(dims, null_space);
  if (n > kMaxElements) return fail(err, "dataset too large");
  const int rank = (int)dims.size();

  Cur c(ml->data);
  const int ver = (int)c.u(1);
  int cls = -1;
  uint64_t addr = ~0ull, csize = 0;
  const uint8_t* compact = nullptr;
  std::vector<uint64_t> cdims;   // chunk dimensions (rank entries)
  int index = 0;                 // 0: v1 B-tree, 1: single chunk, 2: implicit
  uint64_t single_size = 0;
  uint32_t single_mask = 0;
  bool single_filtered = false;
  if (ver == 1 || ver == 2) {
    const int nd = (int)c.u(1);
    cls = (int)c.u(1);
    c.take(5);
    if (cls != 0) addr = c.u(so_);
    std::vector<uint64_t> d(nd);
    for (int k = 0; k < nd; ++k) d[k] = c.u(4);
    if (cls == 0) {
      csize = c.u(4);
      compact = c.take(csize);
    } else if (cls == 2) {
      if (nd != rank + 1) return fail(err, "chunk dimensionality does not match the dataspace");
      cdims.assign(d.begin(), d.begin() + rank);
    }
  } else if (ver == 3 || ver == 4) {
    cls = (int)c.u(1);
    if (cls == 0) {
      csize = c.u(2);
      compact = c.take(csize);
    } else if (cls == 1) {
      addr = c.u(so_);
      csize = c.u(sl_);
    } else if (cls == 2 && ver == 3) {
      const int nd = (int)c.u(1);
      addr = c.u(so_);
      if (nd != rank + 1) return fail(err, "chunk dimensionality does not match the dataspace");
      for (int k = 0; k < nd; ++k) {
        const uint64_t v = c.u(4);
        if (k < rank) cdims.push_back(v);
      }
    } else if (cls == 2) {
      const int fl = (int)c.u(1);
      const int nd = (int)c.u(1);
      const int enc = (int)c.u(1);
      if (nd != rank + 1) return fail(err, "chunk dimensionality does not match the dataspace");
      for (int k = 0; k < nd; ++k) {
        const uint64_t v = c.u(enc);
        if (k < rank) cdims.push_back(v);
      }
      const int it = (int)c.u(1);
      if (it == 1) {
        index = 1;
        single_filtered = (fl & 2) != 0;
        if (single_filtered) {
          single_size = c.u(sl_);
          single_mask = (uint32_t)c.u(4);
        }
        addr = c.u(so_);
      } else if (it == 2) {
        index = 2;
        addr = c.u(so_);
      } else {
        return fail(err, "chunk index type " + std::to_string(it) + " is not supported");
      }
    } else {
      return fail(err, "data layout class " + std::to_string(cls));
    }
  } else {
    return fail(err, "data layout message version " + std::to_string(ver));
  }
  if (!c.ok) return fail(err, "truncated data layout message");
  const size_t es = (size_t)t.size;
  if (cls == 1 && !undefined(addr, so_) && n * es > size_)
    return fail(err, "contiguous dataset larger than the file");
  out.assign(n, 0.0f);
  if (dims_out) *dims_out = dims;

  if (cls == 0) {   // compact
    if (csize < n * es) return fail(err, "compact data shorter than the dataset");
    convert(compact, n, t, out.data());
    return true;
  }
  if (cls == 1) {   // contiguous
    if (undefined(addr, so_) || n == 0) return true;   // never written: fill value 0
    if (t.cls == 1 && t.size == 4 && !t.big) {   // native floats: straight into the output
      const uint64_t a = base_ + addr;
      if (a > size_ || n * 4 > size_ - a || fseeko(fp_, (off_t)a, SEEK_SET) != 0 ||
          fread(out.data(), 4, n, fp_) != n)
        return fail(err, "dataset data outside the file");
      return true;
    }
    const uint64_t step = std::max<uint64_t>(1, (64ull << 20) / es);
    std::vector<uint8_t> buf;
    for (uint64_t k = 0; k < n; k += step) {
      const uint64_t m = std::min(step, n - k);
      if (!bytes(addr + k * es, m * es, buf)) return fail(err, "dataset data outside the file");
      convert(buf.data(), m, t, out.data() + k);
    }
    return true;
  }
  // chunked
  if (rank == 0) return fail(err, "chunked scalar dataset");
  uint64_t cn = 1;
  for (uint64_t d : cdims) {
    if (d == 0) return fail(err, "zero chunk dimension");
    if (cn > (1ull << 28) / d) return fail(err, "chunk too large");
    cn *= d;
  }
  std::vector<Chunk> chunks;
  if (index == 0) {
    if (!undefined(addr, so_) && !chunk_btree(addr, rank + 1, chunks, -1, err)) return false;
  } else if (index == 1) {
    if (!undefined(addr, so_))
      chunks.push_back({addr, single_filtered ? (uint32_t)single_size : (uint32_t)(cn * es),
                        single_filtered ? single_mask : 0u, std::vector<uint64_t>(rank + 1, 0)});
  } else {   // implicit: every chunk allocated, in chunk-grid order
    if (!undefined(addr, so_)) {
      std::vector<uint64_t> grid(rank);
      uint64_t total = 1;
      for (int d = 0; d < rank; ++d) {
        grid[d] = (dims[d] + cdims[d] - 1) / cdims[d];
        total *= grid[d];
      }
      for (uint64_t k = 0; k < total; ++k) {
        Chunk ch{addr + k * cn * es, (uint32_t)(cn * es), 0u, std::vector<uint64_t>(rank + 1, 0)};
        uint64_t r = k;
        for (int d = rank - 1; d >= 0; --d) {
          ch.off[d] = (r % grid[d]) * cdims[d];
          r /= grid[d];
        }
        chunks.push_back(ch);
      }
    }
  }
  std::vector<uint8_t> buf;
  std::vector<float> cv(cn);
  for (const Chunk& ch : chunks) {
    if (!bytes(ch.addr, ch.size, buf)) return fail(err, "chunk outside the file");
    if (!filters.empty() && !unfilter(buf, filters, ch.mask, cn * es, (int)es, err)) return false;
    if (buf.size() < cn * es) return fail(err, "chunk shorter than its dimensions");
    convert(buf.data(), cn, t, cv.data());
    // scatter the chunk's rows (last dimension contiguous) into the dataset
    const uint64_t rowlen = cdims[rank - 1];
    const uint64_t x0 = ch.off[rank - 1];
    if (x0 >= dims[rank - 1]) continue;
    const uint64_t run = std::min(rowlen, dims[rank - 1] - x0);
    const uint64_t nrows = cn / rowlen;
    for (uint64_t r = 0; r < nrows; ++r) {
      uint64_t rem = r, lin = 0, mul = 1;
      bool inside = true;
      std::vector<uint64_t> g(rank);
      for (int d = rank - 2; d >= 0; --d) {
        g[d] = ch.off[d] + rem % cdims[d];
        rem /= cdims[d];
        if (g[d] >= dims[d]) inside = false;
      }
      if (!inside) continue;
      g[rank - 1] = x0;
      for (int d = rank - 1; d >= 0; --d) {
        lin += g[d] * mul;
        mul *= dims[d];
      }
      std::copy(cv.begin() + r * rowlen, cv.begin() + r * rowlen + run, out.begin() + lin);
    }
  }
  return true;
}

}  // namespace gshdf
