// fsg_hdf5.hip -- FastSpec HDF5 line lists read through libhdf5 loaded at run time (host
// code only).
//
// The /Lines group of a list written by the reference's converters (python/vald2hdf.py:
// 763-842, python/mol2hdf.py) holds one dataset per column: W f8 contiguous (the full-column
// bound scan of ReadHDF, src/ioutils.f90:729-749), the other columns chunked (262,144 rows)
// with the shuffle and deflate filters; group attributes NAME, NL, WL_VAC, HFLAG, WMIN, WMAX.
// ReadHDF (src/ioutils.f90:545-1100) reads those with the HDF5 Fortran library; here the C
// library is loaded with dlopen (no link-time dependency: a host without it reads sidecars),
// its calls are serialised under one lock (the library is not built thread-safe), and the
// chunks of a column slice are read raw (H5Dread_chunk) and decoded -- inflate, then the
// inverse byte shuffle -- by the calling thread, so that the reader's worker threads decode
// their slices in parallel.  Datasets with other filters or layouts are read with a
// hyperslab H5Dread under the lock.
#include <dlfcn.h>
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fsg_ingest.h"

namespace fsg {

namespace {

// the parts of the HDF5 1.10+ C interface used here (H5public.h, H5Ipublic.h, ...)
typedef int64_t hid_t;
typedef int herr_t;
typedef int htri_t;
typedef unsigned long long hsize_t;
constexpr hid_t H5P_DEFAULT_ = 0;
constexpr unsigned H5F_ACC_RDONLY_ = 0u;
constexpr int H5T_INTEGER_ = 0, H5T_FLOAT_ = 1, H5T_STRING_ = 3;
constexpr int H5T_ORDER_LE_ = 0;
constexpr int H5D_CONTIGUOUS_ = 1, H5D_CHUNKED_ = 2;
constexpr int H5S_SELECT_SET_ = 0;
constexpr int H5Z_FILTER_DEFLATE_ = 1, H5Z_FILTER_SHUFFLE_ = 2;

struct H5Api {
  void *so = nullptr;
  herr_t (*open)();
  herr_t (*eset_auto2)(hid_t, void *, void *);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*gopen2)(hid_t, const char *, hid_t);
  herr_t (*gclose)(hid_t);
  htri_t (*lexists)(hid_t, const char *, hid_t);
  htri_t (*aexists)(hid_t, const char *);
  hid_t (*aopen)(hid_t, const char *, hid_t);
  hid_t (*aget_type)(hid_t);
  herr_t (*aread)(hid_t, hid_t, void *);
  herr_t (*aclose)(hid_t);
  hid_t (*dopen2)(hid_t, const char *, hid_t);
  herr_t (*dclose)(hid_t);
  hid_t (*dget_type)(hid_t);
  hid_t (*dget_space)(hid_t);
  hid_t (*dget_create_plist)(hid_t);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dread_chunk)(hid_t, hid_t, const hsize_t *, uint32_t *, void *);
  herr_t (*dget_chunk_storage_size)(hid_t, const hsize_t *, hsize_t *);
  int (*tget_class)(hid_t);
  size_t (*tget_size)(hid_t);
  int (*tget_order)(hid_t);
  htri_t (*tis_variable_str)(hid_t);
  herr_t (*tclose)(hid_t);
  int (*sget_simple_extent_ndims)(hid_t);
  int (*sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  int (*pget_layout)(hid_t);
  int (*pget_chunk)(hid_t, int, hsize_t *);
  int (*pget_nfilters)(hid_t);
  int (*pget_filter2)(hid_t, unsigned, unsigned *, size_t *, unsigned *, size_t, char *, unsigned *);
  herr_t (*pclose)(hid_t);
  herr_t (*free_memory)(void *);
  hid_t *native_int, *native_double;
};

// every call into the library, and the loading itself; recursive because an open list that
// is released inside a locked section (an error path) closes its handles under the lock
std::recursive_mutex g_h5_mu;
H5Api g_h5;
bool g_h5_tried = false, g_h5_ok = false;
std::string g_h5_why;

template <typename F>
bool sym(void *so, const char *name, F &f, bool required = true) {
  f = reinterpret_cast<F>(dlsym(so, name));
  if (!f && required) g_h5_why = std::string("libhdf5 lacks ") + name;
  return f != nullptr || !required;
}

// under g_h5_mu: loads libhdf5 once (FSG_HDF5_LIB, then the names a system or conda
// installation gives it)
bool h5_load() {
  if (g_h5_tried) return g_h5_ok;
  g_h5_tried = true;
  std::vector<std::string> cands;
  if (const char *e = std::getenv("FSG_HDF5_LIB")) cands.push_back(e);
  for (const char *c : {"libhdf5.so", "libhdf5.so.103", "libhdf5_serial.so", "libhdf5_serial.so.103",
                        "/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/hdf5/serial/libhdf5.so"})
    cands.push_back(c);
  void *so = nullptr;
  for (const auto &c : cands)
    if ((so = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!so) {
    g_h5_why = "no HDF5 library found (set FSG_HDF5_LIB to libhdf5.so)";
    return false;
  }
  H5Api &a = g_h5;
  a.so = so;
  bool ok = sym(so, "H5open", a.open) && sym(so, "H5Eset_auto2", a.eset_auto2) && sym(so, "H5Fopen", a.fopen) &&
            sym(so, "H5Fclose", a.fclose) && sym(so, "H5Gopen2", a.gopen2) && sym(so, "H5Gclose", a.gclose) &&
            sym(so, "H5Lexists", a.lexists) && sym(so, "H5Aexists", a.aexists) && sym(so, "H5Aopen", a.aopen) &&
            sym(so, "H5Aget_type", a.aget_type) && sym(so, "H5Aread", a.aread) && sym(so, "H5Aclose", a.aclose) &&
            sym(so, "H5Dopen2", a.dopen2) && sym(so, "H5Dclose", a.dclose) && sym(so, "H5Dget_type", a.dget_type) &&
            sym(so, "H5Dget_space", a.dget_space) && sym(so, "H5Dget_create_plist", a.dget_create_plist) &&
            sym(so, "H5Dread", a.dread) && sym(so, "H5Tget_class", a.tget_class) &&
            sym(so, "H5Tget_size", a.tget_size) && sym(so, "H5Tget_order", a.tget_order) &&
            sym(so, "H5Tis_variable_str", a.tis_variable_str) && sym(so, "H5Tclose", a.tclose) &&
            sym(so, "H5Sget_simple_extent_ndims", a.sget_simple_extent_ndims) &&
            sym(so, "H5Sget_simple_extent_dims", a.sget_simple_extent_dims) &&
            sym(so, "H5Sselect_hyperslab", a.sselect_hyperslab) && sym(so, "H5Screate_simple", a.screate_simple) &&
            sym(so, "H5Sclose", a.sclose) && sym(so, "H5Pget_layout", a.pget_layout) &&
            sym(so, "H5Pget_chunk", a.pget_chunk) && sym(so, "H5Pget_nfilters", a.pget_nfilters) &&
            sym(so, "H5Pget_filter2", a.pget_filter2) && sym(so, "H5Pclose", a.pclose) &&
            sym(so, "H5free_memory", a.free_memory);
  // direct chunk reads (1.10.3 and later); without them every read is a hyperslab read
  sym(so, "H5Dread_chunk", a.dread_chunk, false);
  sym(so, "H5Dget_chunk_storage_size", a.dget_chunk_storage_size, false);
  if (ok) {
    a.native_int = reinterpret_cast<hid_t *>(dlsym(so, "H5T_NATIVE_INT_g"));
    a.native_double = reinterpret_cast<hid_t *>(dlsym(so, "H5T_NATIVE_DOUBLE_g"));
    ok = a.native_int && a.native_double;
    if (!ok) g_h5_why = "libhdf5 lacks the native type identifiers";
  }
  if (ok) ok = a.open() >= 0;  // also sets the native type identifiers
  if (ok) a.eset_auto2(0, nullptr, nullptr);  // errors are reported here, not printed
  g_h5_ok = ok;
  return ok;
}

}  // namespace

// one open list: file, group and the datasets of its columns
struct H5List {
  hid_t file = -1, group = -1;
  struct Col {
    hid_t dset = -1;
    int64_t n = 0;
    int elem = 0;           // bytes per value
    int layout = 0;         // H5D_CONTIGUOUS / H5D_CHUNKED / other
    int64_t chunk = 0;      // rows per chunk
    std::vector<int> pipe;  // filters in write order, when all are shuffle/deflate
    bool direct = false;    // chunks decoded here (else hyperslab reads)
    hid_t mtype = -1;       // memory type = the file type (little-endian, stored bytes)
  };
  std::map<std::string, Col> cols;
  ~H5List() {
    std::lock_guard<std::recursive_mutex> lk(g_h5_mu);
    for (auto &kv : cols) {
      if (kv.second.mtype >= 0) g_h5.tclose(kv.second.mtype);
      if (kv.second.dset >= 0) g_h5.dclose(kv.second.dset);
    }
    if (group >= 0) g_h5.gclose(group);
    if (file >= 0) g_h5.fclose(file);
  }
};

namespace {

// under g_h5_mu: a scalar attribute of the group as a string (fixed or variable length)
bool attr_string(hid_t g, const char *name, std::string &out) {
  if (g_h5.aexists(g, name) <= 0) return false;
  hid_t a = g_h5.aopen(g, name, H5P_DEFAULT_);
  if (a < 0) return false;
  hid_t t = g_h5.aget_type(a);
  bool ok = false;
  if (t >= 0 && g_h5.tget_class(t) == H5T_STRING_) {
    if (g_h5.tis_variable_str(t) > 0) {
      char *p = nullptr;
      if (g_h5.aread(a, t, &p) >= 0 && p) {
        out = p;
        g_h5.free_memory(p);
        ok = true;
      }
    } else {
      std::vector<char> buf(g_h5.tget_size(t) + 1, '\0');
      if (g_h5.aread(a, t, buf.data()) >= 0) {
        out.assign(buf.data());
        ok = true;
      }
    }
  }
  if (t >= 0) g_h5.tclose(t);
  g_h5.aclose(a);
  while (!out.empty() && (out.back() == ' ' || out.back() == '\0')) out.pop_back();
  return ok;
}

template <typename T>
bool attr_number(hid_t g, const char *name, hid_t mtype, T &out) {
  if (g_h5.aexists(g, name) <= 0) return false;
  hid_t a = g_h5.aopen(g, name, H5P_DEFAULT_);
  if (a < 0) return false;
  const bool ok = g_h5.aread(a, mtype, &out) >= 0;
  g_h5.aclose(a);
  return ok;
}

// under g_h5_mu: opens column name and records its type, extent and storage
bool open_col(H5List &L, const std::string &name, char &kind, int &size, std::string &err) {
  H5List::Col c;
  c.dset = g_h5.dopen2(L.group, name.c_str(), H5P_DEFAULT_);
  if (c.dset < 0) {
    err = "ReadHDF: cannot open column " + name;
    return false;
  }
  c.mtype = g_h5.dget_type(c.dset);
  const int cls = g_h5.tget_class(c.mtype);
  c.elem = (int)g_h5.tget_size(c.mtype);
  if (cls == H5T_FLOAT_) kind = 'f';
  else if (cls == H5T_INTEGER_) kind = 'i';
  else if (cls == H5T_STRING_ && g_h5.tis_variable_str(c.mtype) <= 0) kind = 'S';
  else kind = 0;
  if (!kind || (kind != 'S' && g_h5.tget_order(c.mtype) != H5T_ORDER_LE_)) {
    err = "ReadHDF: unsupported type of column " + name;
    L.cols[name] = c;
    return false;
  }
  size = c.elem;
  hid_t sp = g_h5.dget_space(c.dset);
  hsize_t dims[1] = {0};
  if (sp < 0 || g_h5.sget_simple_extent_ndims(sp) != 1 || g_h5.sget_simple_extent_dims(sp, dims, nullptr) != 1) {
    if (sp >= 0) g_h5.sclose(sp);
    err = "ReadHDF: column " + name + " is not one-dimensional";
    L.cols[name] = c;
    return false;
  }
  g_h5.sclose(sp);
  c.n = (int64_t)dims[0];
  hid_t pl = g_h5.dget_create_plist(c.dset);
  c.layout = g_h5.pget_layout(pl);
  if (c.layout == H5D_CHUNKED_) {
    hsize_t ch[1] = {0};
    if (g_h5.pget_chunk(pl, 1, ch) == 1) c.chunk = (int64_t)ch[0];
    const int nf = g_h5.pget_nfilters(pl);
    bool known = c.chunk > 0;
    for (int f = 0; f < nf && known; f++) {
      unsigned flags = 0, cd[8], cfg = 0;
      size_t ncd = 8;
      char fname[64];
      const int id = g_h5.pget_filter2(pl, (unsigned)f, &flags, &ncd, cd, sizeof fname, fname, &cfg);
      if (id == H5Z_FILTER_DEFLATE_ || id == H5Z_FILTER_SHUFFLE_) c.pipe.push_back(id);
      else known = false;
    }
    c.direct = known && g_h5.dread_chunk && g_h5.dget_chunk_storage_size;
  }
  g_h5.pclose(pl);
  L.cols[name] = c;
  return true;
}

bool decode_chunk(const H5List::Col &c, uint32_t mask, std::vector<unsigned char> &raw,
                  std::vector<unsigned char> &tmp, std::vector<unsigned char> &out) {
  const size_t full = (size_t)c.chunk * (size_t)c.elem;
  std::vector<unsigned char> *cur = &raw, *next = &tmp;
  // the filters are undone in the reverse of their write order; mask bit f set: filter f
  // was not applied to this chunk
  for (int f = (int)c.pipe.size() - 1; f >= 0; f--) {
    if (mask & (1u << f)) continue;
    next->resize(full);
    if (c.pipe[f] == H5Z_FILTER_DEFLATE_) {
      uLongf dl = (uLongf)full;
      if (uncompress(next->data(), &dl, cur->data(), (uLong)cur->size()) != Z_OK || dl != (uLongf)full) return false;
    } else {  // byte shuffle: byte b of every value is stored in plane b
      if (cur->size() != full) return false;
      const size_t n = (size_t)c.chunk, s = (size_t)c.elem;
      const unsigned char *in = cur->data();
      unsigned char *o = next->data();
      for (size_t b = 0; b < s; b++)
        for (size_t i = 0; i < n; i++) o[i * s + b] = in[b * n + i];
    }
    std::swap(cur, next);
  }
  if (cur->size() != full) return false;
  out.swap(*cur);
  return true;
}

}  // namespace

bool hdf5_file(const char *path) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  unsigned char sig[8] = {0};
  const size_t r = std::fread(sig, 1, 8, f);
  std::fclose(f);
  static const unsigned char h5sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  return r == 8 && std::memcmp(sig, h5sig, 8) == 0;
}

int hdf5_open(const char *path, Sidecar &sc, std::string &err) {
  auto L = std::make_shared<H5List>();  // released after the lock below on every path
  std::lock_guard<std::recursive_mutex> lk(g_h5_mu);
  if (!h5_load()) {
    err = std::string("ReadHDF: ") + path + ": " + g_h5_why;
    return 1;
  }
  L->file = g_h5.fopen(path, H5F_ACC_RDONLY_, H5P_DEFAULT_);
  if (L->file < 0) {
    err = std::string("ReadHDF: h5fopen failed on ") + path;
    return 1;
  }
  L->group = g_h5.gopen2(L->file, "/Lines", H5P_DEFAULT_);
  if (L->group < 0) {
    err = "ReadHDF: group \"/Lines\" not found (expected columnar HDF5 format)";
    return 1;
  }
  sc.dir = path;
  // a multi-species (atomic) list has an EL column; a single-species (molecular) one does not
  sc.molecular = g_h5.lexists(L->group, "EL", H5P_DEFAULT_) <= 0;
  if (!attr_string(L->group, "NAME", sc.name)) {
    err = "ReadHDF: h5aopen_f(NAME) failed";
    return 1;
  }
  if (!sc.molecular && sc.name.compare(0, 4, "atom") != 0) {
    err = "ReadHDF: incorrect HDF5 file format";
    return 1;
  }
  int iv = 0;
  if (!attr_number(L->group, "WL_VAC", *g_h5.native_int, iv)) {
    err = "ReadHDF: h5aopen_f(WL_VAC) failed";
    return 1;
  }
  sc.wl_vac = iv;
  if (!sc.molecular && attr_number(L->group, "HFLAG", *g_h5.native_int, iv)) sc.hflag = iv;
  if (!attr_number(L->group, "WMIN", *g_h5.native_double, sc.wmin) ||
      !attr_number(L->group, "WMAX", *g_h5.native_double, sc.wmax)) {
    err = "ReadHDF: h5aopen_f(WMIN/WMAX) failed";
    return 1;
  }
  static const char *all[] = {"W", "EL", "ION", "E", "LGGF", "GRAD", "GQST", "GVDW",
                              "JLO", "JUP", "GLO", "GUP", "ISOT"};
  for (const char *cn : all) {
    if (g_h5.lexists(L->group, cn, H5P_DEFAULT_) <= 0) continue;
    SidecarCol col;
    int size = 0;
    if (!open_col(*L, cn, col.kind, size, err)) return 1;
    col.size = size;
    sc.cols[cn] = col;
  }
  if (!sc.cols.count("W")) {
    err = "ReadHDF: column W missing";
    return 1;
  }
  sc.nl = L->cols.at("W").n;
  for (const auto &kv : L->cols)
    if (kv.second.n != sc.nl) {
      err = "ReadHDF: column " + kv.first + " has another length than W";
      return 1;
    }
  sc.h5 = L;
  return 0;
}

int hdf5_read(const Sidecar &sc, const std::string &name, int64_t first, int64_t count, void *dst, std::string &err) {
  const H5List &L = *sc.h5;
  auto it = L.cols.find(name);
  if (it == L.cols.end()) {
    err = "ReadHDF: column " + name + " missing in " + sc.dir;
    return 1;
  }
  const H5List::Col &c = it->second;
  if (count <= 0) return 0;
  if (first < 0 || first + count > c.n) {
    err = "ReadHDF: rows beyond column " + name;
    return 1;
  }
  unsigned char *out = static_cast<unsigned char *>(dst);
  if (!c.direct) {  // contiguous (W) or an unknown pipeline: a hyperslab read
    std::lock_guard<std::recursive_mutex> lk(g_h5_mu);
    hid_t fs = g_h5.dget_space(c.dset);
    const hsize_t st[1] = {(hsize_t)first}, cnt[1] = {(hsize_t)count};
    hid_t ms = g_h5.screate_simple(1, cnt, nullptr);
    const bool ok = fs >= 0 && ms >= 0 && g_h5.sselect_hyperslab(fs, H5S_SELECT_SET_, st, nullptr, cnt, nullptr) >= 0 &&
                    g_h5.dread(c.dset, c.mtype, ms, fs, H5P_DEFAULT_, dst) >= 0;
    if (ms >= 0) g_h5.sclose(ms);
    if (fs >= 0) g_h5.sclose(fs);
    if (!ok) {
      err = "ReadHDF: reading column " + name + " failed";
      return 1;
    }
    return 0;
  }
  // chunk by chunk: the raw bytes under the lock, the decoding outside it
  std::vector<unsigned char> raw, tmp, plain;
  const int64_t c0 = first / c.chunk, c1 = (first + count - 1) / c.chunk;
  for (int64_t k = c0; k <= c1; k++) {
    const hsize_t off[1] = {(hsize_t)(k * c.chunk)};
    uint32_t mask = 0;
    {
      std::lock_guard<std::recursive_mutex> lk(g_h5_mu);
      hsize_t nb = 0;
      if (g_h5.dget_chunk_storage_size(c.dset, off, &nb) < 0 || nb == 0) {
        err = "ReadHDF: chunk of column " + name + " not stored";
        return 1;
      }
      raw.resize((size_t)nb);
      if (g_h5.dread_chunk(c.dset, H5P_DEFAULT_, off, &mask, raw.data()) < 0) {
        err = "ReadHDF: reading a chunk of column " + name + " failed";
        return 1;
      }
    }
    if (!decode_chunk(c, mask, raw, tmp, plain)) {
      err = "ReadHDF: a chunk of column " + name + " does not decode";
      return 1;
    }
    const int64_t a = std::max(first, k * c.chunk), b = std::min(first + count, (k + 1) * c.chunk);
    std::memcpy(out + (size_t)(a - first) * c.elem, plain.data() + (size_t)(a - k * c.chunk) * c.elem,
                (size_t)(b - a) * c.elem);
  }
  return 0;
}

}  // namespace fsg
