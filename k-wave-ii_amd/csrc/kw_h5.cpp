// HDF5 subset: reader and writer (see kw_h5.h).  Structures follow the HDF5
// File Format Specification (superblock, object headers and their
// messages, version-1 B-trees, local heaps, symbol table nodes, filter
// pipeline); every read is bounds-checked against the file.
#include "kw_h5.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>

namespace kwh5 {

namespace {

constexpr uint64_t kUndef = ~0ull;
const uint8_t kSig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};

[[noreturn]] void fail(const std::string& m) { throw Error("HDF5: " + m); }

std::vector<std::string> split_path(const std::string& p) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : p) {
    if (ch == '/') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

uint64_t pad8(uint64_t n) { return (n + 7) & ~7ull; }

}  // namespace

// ===========================================================================
// reader
// ===========================================================================

struct Reader::Impl {
  const uint8_t* f = nullptr;
  size_t n = 0;
  int fd = -1;
  int so = 8, sl = 8;  // sizes of offsets and lengths
  uint64_t base = 0;
  uint64_t root = kUndef;

  struct Msg {
    uint16_t type;
    const uint8_t* d;
    size_t size;
  };
  struct Space {
    std::vector<uint64_t> dims;
  };
  struct DType {
    int cls = -1, size = 0;
    bool sign = false;
    int str_pad = 0;
  };
  struct Filter {
    int id;
    std::vector<uint32_t> cd;
  };

  ~Impl() {
    if (f) munmap((void*)f, n);
    if (fd >= 0) close(fd);
  }

  void need(uint64_t off, uint64_t len) const {
    if (off > n || len > n - off) fail("truncated or corrupt file (read past the end)");
  }
  uint64_t u(uint64_t off, int len) const {
    need(off, (uint64_t)len);
    uint64_t v = 0;
    for (int i = len - 1; i >= 0; --i) v = (v << 8) | f[off + i];
    return v;
  }
  static uint64_t ub(const uint8_t* p, int len) {
    uint64_t v = 0;
    for (int i = len - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
  }
  uint64_t addr(uint64_t a) const { return a == kUndef ? kUndef : a + base; }
  // bounds check of a message's field
  static void mneed(const Msg& m, size_t off, size_t len) {
    if (off > m.size || len > m.size - off) fail("corrupt object header message");
  }

  void open(const std::string& path) {
    fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) fail("cannot open " + path);
    struct stat st;
    if (fstat(fd, &st) != 0) fail("cannot stat " + path);
    n = (size_t)st.st_size;
    if (n < 48) fail(path + " is not an HDF5 file");
    void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) fail("cannot map " + path);
    f = (const uint8_t*)m;
    // the superblock is at 0, 512, 1024, ... (a user block may precede it)
    uint64_t sb = kUndef;
    for (uint64_t o = 0; o + 8 <= n; o = o == 0 ? 512 : o * 2)
      if (std::memcmp(f + o, kSig, 8) == 0) {
        sb = o;
        break;
      }
    if (sb == kUndef) fail(path + " is not an HDF5 file (no signature)");
    const int ver = (int)u(sb + 8, 1);
    if (ver == 0 || ver == 1) {
      so = (int)u(sb + 13, 1);
      sl = (int)u(sb + 14, 1);
      if ((so != 4 && so != 8) || (sl != 4 && sl != 8)) fail("unsupported offset size");
      uint64_t p = sb + 24 + (ver == 1 ? 4 : 0);
      base = u(p, so);
      p += 4 * (uint64_t)so;  // base, free space, end of file, driver block
      // root group symbol table entry: link name offset, object header
      root = addr(u(p + so, so));
    } else if (ver == 2 || ver == 3) {
      so = (int)u(sb + 9, 1);
      sl = (int)u(sb + 10, 1);
      if ((so != 4 && so != 8) || (sl != 4 && sl != 8)) fail("unsupported offset size");
      const uint64_t p = sb + 12;
      base = u(p, so);
      root = addr(u(p + 3 * (uint64_t)so, so));
    } else {
      fail("unsupported superblock version " + std::to_string(ver));
    }
  }

  // -- object headers -----------------------------------------------------
  void v1_messages(uint64_t p, uint64_t len, std::vector<Msg>& out, int& left) const {
    const uint64_t end = p + len;
    need(p, len);
    while (p + 8 <= end && left > 0) {
      const uint16_t type = (uint16_t)u(p, 2);
      const uint64_t sz = u(p + 2, 2);
      p += 8;
      if (p + sz > end) fail("object header message past its block");
      out.push_back(Msg{type, f + p, (size_t)sz});
      --left;
      p += sz;
    }
  }
  void v2_messages(uint64_t p, uint64_t end, bool crt_order, std::vector<Msg>& out) const {
    need(p, end - p);
    while (p + 4 <= end) {
      const uint16_t type = (uint16_t)u(p, 1);
      const uint64_t sz = u(p + 1, 2);
      const uint8_t fl = (uint8_t)u(p + 3, 1);
      (void)fl;
      p += 4 + (crt_order ? 2 : 0);
      if (p + sz > end) break;  // gap at the end of the chunk
      out.push_back(Msg{type, f + p, (size_t)sz});
      p += sz;
    }
  }
  std::vector<Msg> header(uint64_t a) const {
    std::vector<Msg> msgs;
    need(a, 16);
    if (std::memcmp(f + a, "OHDR", 4) == 0) {
      const int ver = (int)u(a + 4, 1);
      if (ver != 2) fail("unsupported object header version");
      const uint8_t fl = (uint8_t)u(a + 5, 1);
      uint64_t p = a + 6;
      if (fl & 0x20) p += 16;
      if (fl & 0x10) p += 4;
      const int csz = 1 << (fl & 3);
      const uint64_t len = u(p, csz);
      p += csz;
      const bool crt = fl & 0x04;
      v2_messages(p, p + len, crt, msgs);
      for (size_t i = 0; i < msgs.size(); ++i) {
        if (msgs[i].type != 0x10) continue;
        mneed(msgs[i], 0, so + sl);
        const uint64_t ca = addr(ub(msgs[i].d, so)), cl = ub(msgs[i].d + so, sl);
        need(ca, cl);
        if (cl < 8 || std::memcmp(f + ca, "OCHK", 4) != 0) fail("bad object header continuation");
        v2_messages(ca + 4, ca + cl - 4, crt, msgs);
      }
      return msgs;
    }
    const int ver = (int)u(a, 1);
    if (ver != 1) fail("unsupported object header version " + std::to_string(ver));
    int left = (int)u(a + 2, 2);
    const uint64_t len = u(a + 8, 4);
    v1_messages(a + 16, len, msgs, left);
    for (size_t i = 0; i < msgs.size() && left > 0; ++i) {
      if (msgs[i].type != 0x10) continue;
      mneed(msgs[i], 0, so + sl);
      v1_messages(addr(ub(msgs[i].d, so)), ub(msgs[i].d + so, sl), msgs, left);
    }
    return msgs;
  }
  static const Msg* find(const std::vector<Msg>& ms, uint16_t type) {
    for (const Msg& m : ms)
      if (m.type == type) return &m;
    return nullptr;
  }

  // -- groups -------------------------------------------------------------
  std::string heap_name(uint64_t heap, uint64_t off) const {
    need(heap, 32);
    if (std::memcmp(f + heap, "HEAP", 4) != 0) fail("bad local heap");
    const uint64_t dsize = u(heap + 8, sl);
    const uint64_t da = addr(u(heap + 8 + 2 * (uint64_t)sl, so));
    if (off >= dsize) fail("name offset outside the local heap");
    need(da, dsize);
    const char* s = (const char*)f + da + off;
    const size_t maxl = (size_t)(dsize - off);
    return std::string(s, strnlen(s, maxl));
  }
  void btree_group(uint64_t node, uint64_t heap, std::map<std::string, uint64_t>& out,
                   int depth) const {
    if (depth > 64) fail("group B-tree too deep");
    need(node, 24);
    if (std::memcmp(f + node, "TREE", 4) != 0) fail("bad group B-tree node");
    if (u(node + 4, 1) != 0) fail("not a group B-tree");
    const int level = (int)u(node + 5, 1);
    const int nent = (int)u(node + 6, 2);
    uint64_t p = node + 8 + 2 * (uint64_t)so;
    for (int i = 0; i < nent; ++i) {
      const uint64_t child = addr(u(p + sl, so));
      p += sl + so;
      if (level > 0) {
        btree_group(child, heap, out, depth + 1);
        continue;
      }
      need(child, 8);
      if (std::memcmp(f + child, "SNOD", 4) != 0) fail("bad symbol table node");
      const int ns = (int)u(child + 6, 2);
      for (int k = 0; k < ns; ++k) {
        const uint64_t e = child + 8 + (uint64_t)k * (2 * so + 24);
        out[heap_name(heap, u(e, so))] = addr(u(e + so, so));
      }
    }
  }
  std::map<std::string, uint64_t> members(uint64_t oh) const {
    std::map<std::string, uint64_t> out;
    const auto ms = header(oh);
    if (const Msg* st = find(ms, 0x11)) {
      mneed(*st, 0, 2 * so);
      btree_group(addr(ub(st->d, so)), addr(ub(st->d + so, so)), out, 0);
      return out;
    }
    for (const Msg& m : ms) {
      if (m.type == 0x02) {
        // link info: dense storage needs a fractal heap, not supported
        mneed(m, 0, 2);
        const uint8_t fl = m.d[1];
        const size_t o = 2 + ((fl & 1) ? 8 : 0);
        mneed(m, o, so);
        if (ub(m.d + o, so) != kUndef) fail("groups in dense link storage are not supported");
      }
      if (m.type != 0x06) continue;
      mneed(m, 0, 2);
      const uint8_t fl = m.d[1];
      size_t o = 2;
      int lt = 0;
      if (fl & 0x08) lt = m.d[o++];
      if (fl & 0x04) o += 8;
      if (fl & 0x10) o += 1;
      const int lsz = 1 << (fl & 3);
      mneed(m, o, lsz);
      const uint64_t nl = ub(m.d + o, lsz);
      o += lsz;
      mneed(m, o, nl);
      std::string name((const char*)m.d + o, (size_t)nl);
      o += nl;
      if (lt != 0) continue;  // soft or external link
      mneed(m, o, so);
      out[name] = addr(ub(m.d + o, so));
    }
    return out;
  }
  uint64_t lookup(const std::string& path, bool must = true) const {
    uint64_t oh = root;
    for (const std::string& c : split_path(path)) {
      const auto mem = members(oh);
      auto it = mem.find(c);
      if (it == mem.end()) {
        if (must) fail("no object '" + path + "'");
        return kUndef;
      }
      oh = it->second;
    }
    return oh;
  }

  // -- dataset messages ---------------------------------------------------
  Space space(const Msg& m) const {
    Space s;
    mneed(m, 0, 4);
    const int ver = m.d[0], rank = m.d[1];
    size_t o;
    if (ver == 1)
      o = 8;
    else if (ver == 2)
      o = 4;
    else
      fail("unsupported dataspace version");
    if (ver == 2 && m.d[3] == 2) return s;  // null dataspace
    mneed(m, o, (size_t)rank * sl);
    for (int i = 0; i < rank; ++i) s.dims.push_back(ub(m.d + o + (size_t)i * sl, sl));
    return s;
  }
  DType dtype(const Msg& m) const {
    DType t;
    mneed(m, 0, 8);
    t.cls = m.d[0] & 0x0f;
    const uint8_t b0 = m.d[1];
    t.size = (int)ub(m.d + 4, 4);
    if (t.cls == 0) {
      if (b0 & 1) fail("big-endian integers are not supported");
      t.sign = b0 & 0x08;
      if (t.size != 1 && t.size != 2 && t.size != 4 && t.size != 8) fail("unsupported integer size");
    } else if (t.cls == 1) {
      if (b0 & 0x41) fail("big-endian or VAX floating point is not supported");
      if (t.size != 4 && t.size != 8) fail("unsupported floating-point size");
    } else if (t.cls == 3) {
      t.str_pad = b0 & 0x0f;
    } else {
      fail("unsupported datatype class " + std::to_string(t.cls));
    }
    return t;
  }
  std::vector<Filter> filters(const Msg* m) const {
    std::vector<Filter> out;
    if (!m) return out;
    mneed(*m, 0, 2);
    const int ver = m->d[0], nf = m->d[1];
    size_t o = ver == 1 ? 8 : 2;
    for (int i = 0; i < nf; ++i) {
      Filter fl;
      mneed(*m, o, 2);
      fl.id = (int)ub(m->d + o, 2);
      o += 2;
      size_t nlen = 0;
      if (ver == 1 || fl.id >= 256) {
        mneed(*m, o, 2);
        nlen = ub(m->d + o, 2);
        o += 2;
      }
      mneed(*m, o, 4);
      const int nv = (int)ub(m->d + o + 2, 2);
      o += 4;
      o += ver == 1 ? pad8(nlen) : nlen;
      mneed(*m, o, (size_t)nv * 4);
      for (int k = 0; k < nv; ++k) fl.cd.push_back((uint32_t)ub(m->d + o + 4 * k, 4));
      o += (size_t)nv * 4;
      if (ver == 1 && (nv & 1)) o += 4;
      out.push_back(fl);
    }
    return out;
  }

  // undo the filters (in reverse order) of one chunk
  std::vector<uint8_t> unfilter(const uint8_t* p, size_t len, uint32_t mask,
                                const std::vector<Filter>& fs, size_t expect, int esize) const {
    std::vector<uint8_t> buf(p, p + len);
    for (int i = (int)fs.size() - 1; i >= 0; --i) {
      if (mask & (1u << i)) continue;
      const Filter& fl = fs[i];
      if (fl.id == 1) {
        std::vector<uint8_t> out(expect);
        uLongf dl = (uLongf)expect;
        if (uncompress(out.data(), &dl, buf.data(), (uLong)buf.size()) != Z_OK)
          fail("deflate: corrupt chunk");
        out.resize(dl);
        buf.swap(out);
      } else if (fl.id == 2) {
        const int es = fl.cd.empty() ? esize : (int)fl.cd[0];
        const size_t ne = buf.size() / es;
        std::vector<uint8_t> out(buf.size());
        for (int b = 0; b < es; ++b)
          for (size_t e = 0; e < ne; ++e) out[e * es + b] = buf[b * ne + e];
        std::copy(buf.begin() + ne * es, buf.end(), out.begin() + ne * es);
        buf.swap(out);
      } else if (fl.id == 3) {
        if (buf.size() < 4) fail("fletcher32: short chunk");
        buf.resize(buf.size() - 4);
      } else {
        fail("unsupported filter " + std::to_string(fl.id));
      }
    }
    if (buf.size() != expect) fail("chunk of unexpected size after the filters");
    return buf;
  }

  // raw bytes of a dataset in C order
  std::vector<uint8_t> raw(uint64_t oh, std::vector<uint64_t>& dims, DType& t) const {
    const auto ms = header(oh);
    const Msg* sm = find(ms, 0x01);
    const Msg* tm = find(ms, 0x03);
    const Msg* lm = find(ms, 0x08);
    if (!sm || !tm || !lm) fail("not a dataset");
    dims = space(*sm).dims;
    t = dtype(*tm);
    uint64_t ne = 1;
    for (uint64_t d : dims) ne *= d;
    const uint64_t nb = ne * (uint64_t)t.size;
    std::vector<uint8_t> out((size_t)nb, 0);
    const Msg& L = *lm;
    mneed(L, 0, 2);
    const int lver = L.d[0];
    int cls;
    size_t o;
    if (lver == 3) {
      cls = L.d[1];
      o = 2;
    } else if (lver == 1 || lver == 2) {
      cls = L.d[2];
      o = 8;
    } else {
      fail("unsupported layout version " + std::to_string(lver));
    }
    if (cls == 0) {  // compact
      if (lver != 3) fail("unsupported compact layout");
      mneed(L, o, 2);
      const uint64_t sz = ub(L.d + o, 2);
      mneed(L, o + 2, sz);
      if (sz < nb) fail("compact dataset too small");
      std::memcpy(out.data(), L.d + o + 2, (size_t)nb);
    } else if (cls == 1) {  // contiguous
      mneed(L, o, so);
      const uint64_t a = ub(L.d + o, so);
      if (a != kUndef) {
        need(addr(a), nb);
        std::memcpy(out.data(), f + addr(a), (size_t)nb);
      }
    } else if (cls == 2) {  // chunked, B-tree version 1 index
      int rk;
      uint64_t bt;
      if (lver == 3) {
        mneed(L, o, 1 + so);
        rk = L.d[o];
        bt = ub(L.d + o + 1, so);
        o += 1 + so;
      } else {
        rk = L.d[1];
        mneed(L, o, so);
        bt = ub(L.d + o, so);
        o += so;
      }
      mneed(L, o, 4 * (size_t)rk);
      std::vector<uint64_t> cd(rk);
      for (int i = 0; i < rk; ++i) cd[i] = ub(L.d + o + 4 * i, 4);
      if (rk != (int)dims.size() + 1) fail("chunk rank does not match the dataspace");
      const auto fs = filters(find(ms, 0x0B));
      uint64_t celems = 1;
      for (int i = 0; i + 1 < rk; ++i) celems *= cd[i];
      const size_t cbytes = (size_t)(celems * (uint64_t)t.size);
      if (bt != kUndef) read_chunks(addr(bt), dims, cd, fs, t.size, cbytes, out, 0);
    } else {
      fail("unsupported layout class");
    }
    return out;
  }

  void read_chunks(uint64_t node, const std::vector<uint64_t>& dims,
                   const std::vector<uint64_t>& cd, const std::vector<Filter>& fs, int es,
                   size_t cbytes, std::vector<uint8_t>& out, int depth) const {
    if (depth > 64) fail("chunk B-tree too deep");
    need(node, 24);
    if (std::memcmp(f + node, "TREE", 4) != 0) fail("bad chunk B-tree node");
    if (u(node + 4, 1) != 1) fail("not a chunk B-tree");
    const int level = (int)u(node + 5, 1);
    const int nent = (int)u(node + 6, 2);
    const int rk = (int)cd.size();
    const uint64_t ksz = 8 + 8 * (uint64_t)rk;
    uint64_t p = node + 8 + 2 * (uint64_t)so;
    const int nd = rk - 1;
    for (int i = 0; i < nent; ++i) {
      const uint64_t csize = u(p, 4);
      const uint32_t mask = (uint32_t)u(p + 4, 4);
      std::vector<uint64_t> off(nd);
      for (int d = 0; d < nd; ++d) off[d] = u(p + 8 + 8 * (uint64_t)d, 8);
      const uint64_t child = addr(u(p + ksz, so));
      p += ksz + so;
      if (level > 0) {
        read_chunks(child, dims, cd, fs, es, cbytes, out, depth + 1);
        continue;
      }
      need(child, csize);
      const std::vector<uint8_t> c = unfilter(f + child, (size_t)csize, mask, fs, cbytes, es);
      // copy the chunk's elements that lie inside the dataset
      std::vector<uint64_t> idx(nd, 0);
      const uint64_t run = std::min<uint64_t>(cd[nd - 1], dims[nd - 1] - std::min(dims[nd - 1], off[nd - 1]));
      for (;;) {
        bool inside = true;
        uint64_t dst = 0, src = 0;
        for (int d = 0; d < nd; ++d) {
          const uint64_t g = off[d] + idx[d];
          if (g >= dims[d]) inside = false;
          dst = dst * dims[d] + g;
          src = src * cd[d] + idx[d];
        }
        if (inside && run > 0)
          std::memcpy(out.data() + dst * es, c.data() + src * es, (size_t)(run * es));
        int d = nd - 2;
        for (; d >= 0; --d) {
          if (++idx[d] < cd[d]) break;
          idx[d] = 0;
        }
        if (d < 0) break;
      }
    }
  }

  // attributes
  bool attr_raw(uint64_t oh, const std::string& name, std::string* value) const {
    for (const Msg& m : header(oh)) {
      if (m.type != 0x0C) continue;
      mneed(m, 0, 8);
      const int ver = m.d[0];
      const uint64_t nsz = ub(m.d + 2, 2), tsz = ub(m.d + 4, 2), ssz = ub(m.d + 6, 2);
      size_t o = ver == 3 ? 9 : 8;
      auto step = [&](uint64_t s) { return ver == 1 ? pad8(s) : s; };
      mneed(m, o, nsz);
      std::string an((const char*)m.d + o, strnlen((const char*)m.d + o, (size_t)nsz));
      o += step(nsz);
      if (an != name) continue;
      if (!value) return true;
      mneed(m, o, tsz);
      Msg tm{0x03, m.d + o, (size_t)tsz};
      const DType t = dtype(tm);
      o += step(tsz);
      mneed(m, o, ssz);
      Msg smsg{0x01, m.d + o, (size_t)ssz};
      const Space s = space(smsg);
      o += step(ssz);
      if (t.cls != 3) fail("attribute '" + name + "' is not a string");
      uint64_t ne = 1;
      for (uint64_t d : s.dims) ne *= d;
      if (ne != 1) fail("attribute '" + name + "' is not a single string");
      mneed(m, o, t.size);
      std::string v((const char*)m.d + o, strnlen((const char*)m.d + o, (size_t)t.size));
      if (t.str_pad == 2)
        while (!v.empty() && v.back() == ' ') v.pop_back();
      *value = v;
      return true;
    }
    return false;
  }
};

Reader::Reader(const std::string& path) : p_(new Impl) { p_->open(path); }
Reader::~Reader() = default;

bool Reader::has(const std::string& path) const { return p_->lookup(path, false) != kUndef; }

std::vector<uint64_t> Reader::dims(const std::string& path) const {
  const auto ms = p_->header(p_->lookup(path));
  const Impl::Msg* sm = Impl::find(ms, 0x01);
  if (!sm) fail("'" + path + "' is not a dataset");
  return p_->space(*sm).dims;
}

Type Reader::type(const std::string& path) const {
  const auto ms = p_->header(p_->lookup(path));
  const Impl::Msg* tm = Impl::find(ms, 0x03);
  if (!tm) fail("'" + path + "' is not a dataset");
  const Impl::DType t = p_->dtype(*tm);
  Type r;
  r.size = t.size;
  r.kind = t.cls == 1 ? Kind::Float : t.cls == 3 ? Kind::String : (t.sign ? Kind::Int : Kind::UInt);
  return r;
}

std::vector<double> Reader::read_double(const std::string& path) const {
  std::vector<uint64_t> dims;
  Impl::DType t;
  const std::vector<uint8_t> r = p_->raw(p_->lookup(path), dims, t);
  const size_t ne = r.size() / (size_t)t.size;
  std::vector<double> out(ne);
  for (size_t i = 0; i < ne; ++i) {
    const uint8_t* e = r.data() + i * t.size;
    if (t.cls == 1) {
      if (t.size == 4) {
        float v;
        std::memcpy(&v, e, 4);
        out[i] = v;
      } else {
        std::memcpy(&out[i], e, 8);
      }
    } else if (t.cls == 0) {
      uint64_t v = Impl::ub(e, t.size);
      if (t.sign && t.size < 8 && (v >> (8 * t.size - 1))) v |= ~0ull << (8 * t.size);
      out[i] = t.sign ? (double)(int64_t)v : (double)v;
    } else {
      fail("'" + path + "' is not numeric");
    }
  }
  return out;
}

std::vector<uint64_t> Reader::read_u64(const std::string& path) const {
  std::vector<uint64_t> dims;
  Impl::DType t;
  const std::vector<uint8_t> r = p_->raw(p_->lookup(path), dims, t);
  const size_t ne = r.size() / (size_t)t.size;
  std::vector<uint64_t> out(ne);
  for (size_t i = 0; i < ne; ++i) {
    const uint8_t* e = r.data() + i * t.size;
    if (t.cls == 0) {
      uint64_t v = Impl::ub(e, t.size);
      if (t.sign && (v >> (8 * t.size - 1))) fail("'" + path + "' holds a negative value");
      out[i] = v;
    } else if (t.cls == 1) {
      double v;
      if (t.size == 4) {
        float x;
        std::memcpy(&x, e, 4);
        v = x;
      } else {
        std::memcpy(&v, e, 8);
      }
      if (!(v >= 0.0) || v != (double)(uint64_t)v) fail("'" + path + "' holds a non-integer value");
      out[i] = (uint64_t)v;
    } else {
      fail("'" + path + "' is not numeric");
    }
  }
  return out;
}

double Reader::scalar(const std::string& path) const {
  const std::vector<double> v = read_double(path);
  if (v.size() != 1) fail("'" + path + "' is not a scalar");
  return v[0];
}

std::string Reader::attr_string(const std::string& obj, const std::string& name) const {
  std::string v;
  if (!p_->attr_raw(p_->lookup(obj), name, &v)) fail("no attribute '" + name + "' on '" + obj + "'");
  return v;
}

bool Reader::has_attr(const std::string& obj, const std::string& name) const {
  const uint64_t oh = p_->lookup(obj, false);
  return oh != kUndef && p_->attr_raw(oh, name, nullptr);
}

std::vector<std::string> Reader::list(const std::string& group) const {
  std::vector<std::string> out;
  for (const auto& kv : p_->members(p_->lookup(group))) out.push_back(kv.first);
  return out;
}

// ===========================================================================
// writer
// ===========================================================================

Writer::Obj& Writer::ensure_group(const std::string& path) {
  auto parts = split_path(path);
  std::string cur;
  objs_[""].group = true;
  for (const std::string& c : parts) {
    const std::string parent = cur;
    cur = cur.empty() ? c : cur + "/" + c;
    Obj& o = objs_[cur];
    if (!o.group && o.data) fail("'" + cur + "' is a dataset");
    o.group = true;
    objs_[parent].children[c] = cur;
  }
  return objs_[cur];
}

void Writer::add_group(const std::string& path) { ensure_group(path); }

void Writer::add_data(const std::string& path, std::vector<uint64_t> dims, Kind k, int size,
                      const void* data) {
  auto parts = split_path(path);
  if (parts.empty()) fail("empty dataset name");
  std::string parent;
  for (size_t i = 0; i + 1 < parts.size(); ++i) parent = parent.empty() ? parts[i] : parent + "/" + parts[i];
  ensure_group(parent);
  const std::string full = parent.empty() ? parts.back() : parent + "/" + parts.back();
  if (objs_.count(full)) fail("'" + full + "' written twice");
  Obj& o = objs_[full];
  o.dims = std::move(dims);
  o.kind = k;
  o.size = size;
  o.data = data;
  objs_[parent].children[parts.back()] = full;
}

void Writer::add_f32(const std::string& p, std::vector<uint64_t> d, const float* x) {
  add_data(p, std::move(d), Kind::Float, 4, x);
}
void Writer::add_f64(const std::string& p, std::vector<uint64_t> d, const double* x) {
  add_data(p, std::move(d), Kind::Float, 8, x);
}
void Writer::add_u64(const std::string& p, std::vector<uint64_t> d, const uint64_t* x) {
  add_data(p, std::move(d), Kind::UInt, 8, x);
}

void Writer::attr(const std::string& obj, const std::string& name, const std::string& value) {
  const std::string key = split_path(obj).empty() ? "" : obj;
  if (key.empty()) objs_[""].group = true;
  auto it = objs_.find(key);
  if (it == objs_.end()) fail("attribute on unknown object '" + obj + "'");
  it->second.attrs.push_back({name, value});
}

namespace {

struct Buf {
  std::vector<uint8_t> b;
  void u(uint64_t v, int n) {
    for (int i = 0; i < n; ++i) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void bytes(const void* p, size_t n) {
    const uint8_t* q = (const uint8_t*)p;
    b.insert(b.end(), q, q + n);
  }
  void zeros(size_t n) { b.insert(b.end(), n, 0); }
  void align8() { zeros((size_t)(pad8(b.size()) - b.size())); }
};

// datatype message body
void put_dtype(Buf& o, Kind k, int size) {
  if (k == Kind::Float) {
    o.u(0x11, 1);  // class 1, version 1
    o.u(0x20, 1);  // little endian, implied mantissa MSB
    o.u(size == 4 ? 31 : 63, 1);  // sign bit
    o.u(0, 1);
    o.u((uint64_t)size, 4);
    o.u(0, 2);
    o.u((uint64_t)size * 8, 2);
    o.u(size == 4 ? 23 : 52, 1);  // exponent location
    o.u(size == 4 ? 8 : 11, 1);   // exponent size
    o.u(0, 1);                    // mantissa location
    o.u(size == 4 ? 23 : 52, 1);  // mantissa size
    o.u(size == 4 ? 127 : 1023, 4);
  } else if (k == Kind::UInt || k == Kind::Int) {
    o.u(0x10, 1);  // class 0, version 1
    o.u(k == Kind::Int ? 0x08 : 0x00, 1);
    o.u(0, 2);
    o.u((uint64_t)size, 4);
    o.u(0, 2);
    o.u((uint64_t)size * 8, 2);
  } else {
    o.u(0x13, 1);  // class 3, version 1
    o.u(0x00, 1);  // null terminated, ASCII
    o.u(0, 2);
    o.u((uint64_t)size, 4);
  }
}

void put_msg(Buf& o, uint16_t type, const Buf& body) {
  o.u(type, 2);
  o.u(pad8(body.b.size()), 2);
  o.u(0, 4);
  o.bytes(body.b.data(), body.b.size());
  o.align8();
}

Buf attr_msg(const std::string& name, const std::string& value) {
  Buf a;
  Buf t;
  put_dtype(t, Kind::String, (int)value.size() + 1);
  Buf s;  // scalar dataspace (version 1, rank 0)
  s.u(1, 1);
  s.zeros(7);
  a.u(1, 1);
  a.u(0, 1);
  a.u(name.size() + 1, 2);
  a.u(t.b.size(), 2);
  a.u(s.b.size(), 2);
  a.bytes(name.c_str(), name.size() + 1);
  a.align8();
  a.bytes(t.b.data(), t.b.size());
  a.align8();
  a.bytes(s.b.data(), s.b.size());
  a.align8();
  a.bytes(value.c_str(), value.size() + 1);
  return a;
}

}  // namespace

void Writer::write(const std::string& file) const {
  // file layout: superblock, then every object's metadata (object header;
  // for groups also local heap, heap data, B-tree node and symbol table
  // node), then the raw data of every dataset
  constexpr int kInternalK = 16;
  size_t maxn = 1;
  for (const auto& kv : objs_)
    if (kv.second.group) maxn = std::max(maxn, kv.second.children.size());
  const int leafK = (int)std::max<size_t>(4, (maxn + 1) / 2);
  const uint64_t snod_size = 8 + 2ull * leafK * 40;
  const uint64_t tree_size = 24 + 2ull * kInternalK * 8 + (2ull * kInternalK + 1) * 8;

  struct Lay {
    uint64_t oh = 0, oh_size = 0, heap = 0, heap_data = 0, heap_size = 0, tree = 0, snod = 0,
             data = 0, nbytes = 0;
    std::map<std::string, uint64_t> name_off;
  };
  std::map<std::string, Lay> lay;
  auto oh_body = [&](const Obj& o, const Lay& L) {
    Buf m;
    if (o.group) {
      Buf st;
      st.u(L.tree, 8);
      st.u(L.heap, 8);
      put_msg(m, 0x11, st);
    } else {
      Buf sp;  // dataspace version 1
      sp.u(1, 1);
      sp.u(o.dims.size(), 1);
      sp.u(0, 1);
      sp.zeros(5);
      for (uint64_t d : o.dims) sp.u(d, 8);
      put_msg(m, 0x01, sp);
      Buf dt;
      put_dtype(dt, o.kind, o.size);
      put_msg(m, 0x03, dt);
      Buf fv;  // fill value version 2: allocated early, written if set, undefined
      fv.u(2, 1);
      fv.u(1, 1);
      fv.u(2, 1);
      fv.u(0, 1);
      put_msg(m, 0x05, fv);
      Buf ly;  // layout version 3, contiguous
      ly.u(3, 1);
      ly.u(1, 1);
      ly.u(L.nbytes ? L.data : kUndef, 8);
      ly.u(L.nbytes, 8);
      put_msg(m, 0x08, ly);
    }
    for (const auto& a : o.attrs) put_msg(m, 0x0C, attr_msg(a.first, a.second));
    return m;
  };
  auto nmsgs = [](const Obj& o) { return (o.group ? 1 : 4) + (int)o.attrs.size(); };

  // pass 1: sizes and addresses
  uint64_t at = 96;
  for (const auto& kv : objs_) {
    const Obj& o = kv.second;
    Lay& L = lay[kv.first];
    if (!o.group) {
      uint64_t ne = 1;
      for (uint64_t d : o.dims) ne *= d;
      L.nbytes = ne * (uint64_t)o.size;
    }
    L.oh_size = oh_body(o, L).b.size();  // addresses do not change sizes
    L.oh = at;
    at += 16 + L.oh_size;
    if (o.group) {
      uint64_t hs = 8;  // offset 0: the empty name
      for (const auto& c : o.children) {
        L.name_off[c.first] = hs;
        hs += pad8(c.first.size() + 1);
      }
      L.heap_size = hs;
      L.heap = at;
      at += 32;
      L.heap_data = at;
      at += hs;
      L.tree = at;
      at += tree_size;
      L.snod = at;
      at += snod_size;
    }
  }
  for (auto& kv : lay) {
    if (objs_.at(kv.first).group || kv.second.nbytes == 0) continue;
    kv.second.data = at;
    at += pad8(kv.second.nbytes);
  }
  const uint64_t eof = at;

  // pass 2: metadata
  Buf out;
  out.b.reserve((size_t)std::min<uint64_t>(eof, 1 << 20));
  const Lay& R = lay.at("");
  out.bytes(kSig, 8);
  out.u(0, 1);  // superblock version
  out.u(0, 1);  // free-space version
  out.u(0, 1);  // root group symbol table entry version
  out.u(0, 1);
  out.u(0, 1);  // shared header message format version
  out.u(8, 1);  // size of offsets
  out.u(8, 1);  // size of lengths
  out.u(0, 1);
  out.u((uint64_t)leafK, 2);
  out.u(kInternalK, 2);
  out.u(0, 4);  // consistency flags
  out.u(0, 8);  // base address
  out.u(kUndef, 8);
  out.u(eof, 8);
  out.u(kUndef, 8);
  // root group symbol table entry, with its B-tree and heap cached
  out.u(0, 8);
  out.u(R.oh, 8);
  out.u(1, 4);
  out.u(0, 4);
  out.u(R.tree, 8);
  out.u(R.heap, 8);
  for (const auto& kv : objs_) {
    const Obj& o = kv.second;
    const Lay& L = lay.at(kv.first);
    if (out.b.size() != L.oh) fail("internal layout error");
    const Buf body = oh_body(o, L);
    out.u(1, 1);  // object header version 1
    out.u(0, 1);
    out.u((uint64_t)nmsgs(o), 2);
    out.u(1, 4);  // reference count
    out.u(body.b.size(), 4);
    out.u(0, 4);  // alignment
    out.bytes(body.b.data(), body.b.size());
    if (!o.group) continue;
    // local heap: header, then the names
    out.bytes("HEAP", 4);
    out.u(0, 1);
    out.zeros(3);
    out.u(L.heap_size, 8);
    out.u(1, 8);  // no free block (the HDF5 library's free-list terminator)
    out.u(L.heap_data, 8);
    out.zeros(8);
    for (const auto& c : o.children) {
      out.bytes(c.first.c_str(), c.first.size() + 1);
      out.align8();
    }
    // B-tree node (group nodes, level 0) with one child: the symbol table node
    out.bytes("TREE", 4);
    out.u(0, 1);
    out.u(0, 1);
    out.u(o.children.empty() ? 0 : 1, 2);
    out.u(kUndef, 8);
    out.u(kUndef, 8);
    const size_t tree_start = out.b.size() - 24;
    if (!o.children.empty()) {
      out.u(0, 8);
      out.u(L.snod, 8);
      out.u(L.name_off.at(o.children.rbegin()->first), 8);
    }
    out.zeros((size_t)(tree_start + tree_size - out.b.size()));
    // symbol table node: the members in name order
    out.bytes("SNOD", 4);
    out.u(1, 1);
    out.u(0, 1);
    out.u(o.children.size(), 2);
    const size_t snod_start = out.b.size() - 8;
    for (const auto& c : o.children) {
      const Obj& co = objs_.at(c.second);
      const Lay& CL = lay.at(c.second);
      out.u(L.name_off.at(c.first), 8);
      out.u(CL.oh, 8);
      out.u(co.group ? 1 : 0, 4);
      out.u(0, 4);
      out.u(co.group ? CL.tree : 0, 8);
      out.u(co.group ? CL.heap : 0, 8);
    }
    out.zeros((size_t)(snod_start + snod_size - out.b.size()));
  }
  FILE* fp = std::fopen(file.c_str(), "wb");
  if (!fp) fail("cannot create " + file);
  bool ok = std::fwrite(out.b.data(), 1, out.b.size(), fp) == out.b.size();
  uint64_t pos = out.b.size();
  static const uint8_t zeros8[8] = {0};
  for (const auto& kv : lay) {
    const Obj& o = objs_.at(kv.first);
    if (o.group || kv.second.nbytes == 0) continue;
    if (pos != kv.second.data) ok = false;
    ok = ok && std::fwrite(o.data, 1, (size_t)kv.second.nbytes, fp) == kv.second.nbytes;
    const uint64_t padn = pad8(kv.second.nbytes) - kv.second.nbytes;
    ok = ok && std::fwrite(zeros8, 1, (size_t)padn, fp) == padn;
    pos += pad8(kv.second.nbytes);
  }
  ok = (std::fclose(fp) == 0) && ok && pos == eof;
  if (!ok) fail("writing " + file + " failed");
}

}  // namespace kwh5
