// HDF5 files in the reference's layout (nrgcore/include/utils/h5stream.hpp:244-478): one-dimensional
// datasets, a std::vector of vectors written as datasets name0, name1, ... (h5stream.hpp:295-305),
// groups created before their members, scalar attributes.  The HDF5 C library (1.10) is loaded at
// run time, so that libnrgk does not depend on it unless a file is written or read; the few entry
// points used are declared here from the library's public C API.
//
// Host code only: compiled by the host compiler.
#include "h5io.hpp"

#include <dlfcn.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

namespace nrgk {
namespace h5 {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
using hssize_t = long long;

namespace {
constexpr unsigned ACC_RDONLY = 0x0000u, ACC_RDWR = 0x0001u, ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0, E_DEFAULT = 0;
constexpr int S_SCALAR = 0;

struct Api {
  void* so = nullptr;
  herr_t (*open)();
  hid_t (*Fcreate)(const char*, unsigned, hid_t, hid_t);
  hid_t (*Fopen)(const char*, unsigned, hid_t);
  herr_t (*Fclose)(hid_t);
  hid_t (*Gcreate2)(hid_t, const char*, hid_t, hid_t, hid_t);
  herr_t (*Gclose)(hid_t);
  hid_t (*Screate_simple)(int, const hsize_t*, const hsize_t*);
  hid_t (*Screate)(int);
  herr_t (*Sclose)(hid_t);
  hssize_t (*Sget_simple_extent_npoints)(hid_t);
  hid_t (*Dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*Dopen2)(hid_t, const char*, hid_t);
  herr_t (*Dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
  herr_t (*Dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
  hid_t (*Dget_space)(hid_t);
  herr_t (*Dclose)(hid_t);
  hid_t (*Acreate_by_name)(hid_t, const char*, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*Awrite)(hid_t, hid_t, const void*);
  herr_t (*Aclose)(hid_t);
  htri_t (*Lexists)(hid_t, const char*, hid_t);
  herr_t (*Eset_auto2)(hid_t, void*, void*);
  hid_t* t_double = nullptr;
  hid_t* t_int = nullptr;
  hid_t* t_ulong = nullptr;
};

template <class T>
void sym(void* so, const char* name, T& f) {
  void* p = dlsym(so, name);
  if (!p) throw std::runtime_error(std::string("nrgk: HDF5 symbol ") + name + " not found");
  f = reinterpret_cast<T>(p);
}

const Api& api() {
  static Api a;
  static std::once_flag once;
  static std::string err;
  std::call_once(once, [] {
    const char* names[] = {"libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
    for (const char* n : names)
      if ((a.so = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!a.so) {
      err = "nrgk: the HDF5 library (libhdf5.so) is not available";
      return;
    }
    try {
      sym(a.so, "H5open", a.open);
      sym(a.so, "H5Fcreate", a.Fcreate);
      sym(a.so, "H5Fopen", a.Fopen);
      sym(a.so, "H5Fclose", a.Fclose);
      sym(a.so, "H5Gcreate2", a.Gcreate2);
      sym(a.so, "H5Gclose", a.Gclose);
      sym(a.so, "H5Screate_simple", a.Screate_simple);
      sym(a.so, "H5Screate", a.Screate);
      sym(a.so, "H5Sclose", a.Sclose);
      sym(a.so, "H5Sget_simple_extent_npoints", a.Sget_simple_extent_npoints);
      sym(a.so, "H5Dcreate2", a.Dcreate2);
      sym(a.so, "H5Dopen2", a.Dopen2);
      sym(a.so, "H5Dwrite", a.Dwrite);
      sym(a.so, "H5Dread", a.Dread);
      sym(a.so, "H5Dget_space", a.Dget_space);
      sym(a.so, "H5Dclose", a.Dclose);
      sym(a.so, "H5Acreate_by_name", a.Acreate_by_name);
      sym(a.so, "H5Awrite", a.Awrite);
      sym(a.so, "H5Aclose", a.Aclose);
      sym(a.so, "H5Lexists", a.Lexists);
      sym(a.so, "H5Eset_auto2", a.Eset_auto2);
      sym(a.so, "H5T_NATIVE_DOUBLE_g", a.t_double);
      sym(a.so, "H5T_NATIVE_INT_g", a.t_int);
      sym(a.so, "H5T_NATIVE_ULONG_g", a.t_ulong);
      if (a.open() < 0) throw std::runtime_error("nrgk: H5open failed");
      a.Eset_auto2(E_DEFAULT, nullptr, nullptr);  // errors are reported through return codes
    } catch (const std::exception& e) {
      err = e.what();
    }
  });
  if (!err.empty()) throw std::runtime_error(err);
  return a;
}

hid_t type_of(const Api& a, Type t) { return t == Type::F64 ? *a.t_double : t == Type::I32 ? *a.t_int : *a.t_ulong; }
}  // namespace

File::File(const std::string& path, bool write) : path_(path) {
  const Api& a = api();
  id_ = write ? a.Fcreate(path.c_str(), ACC_TRUNC, P_DEFAULT, P_DEFAULT) : a.Fopen(path.c_str(), ACC_RDONLY, P_DEFAULT);
  if (id_ < 0) throw std::runtime_error("nrgk: cannot " + std::string(write ? "create" : "open") + " HDF5 file " + path);
}
File::File(const std::string& path, Mode) : path_(path) {
  const Api& a = api();
  id_ = a.Fopen(path.c_str(), ACC_RDWR, P_DEFAULT);
  if (id_ < 0) throw std::runtime_error("nrgk: cannot open HDF5 file " + path + " for writing");
}
File::~File() {
  if (id_ >= 0) api().Fclose(id_);
}

void File::group(const std::string& name) {
  const Api& a = api();
  // every missing level of the path, as h5stream's createGroup does for its one level
  for (size_t p = 1; p <= name.size(); p++) {
    if (p != name.size() && name[p] != '/') continue;
    const std::string sub = name.substr(0, p);
    if (sub.empty() || sub == "/") continue;
    if (a.Lexists(id_, sub.c_str(), P_DEFAULT) > 0) continue;
    const hid_t g = a.Gcreate2(id_, sub.c_str(), P_DEFAULT, P_DEFAULT, P_DEFAULT);
    if (g < 0) throw std::runtime_error("nrgk: cannot create HDF5 group " + sub + " in " + path_);
    a.Gclose(g);
  }
}

void File::write(const std::string& name, Type t, const void* data, size_t n) {
  const Api& a = api();
  const hsize_t dims[1] = {(hsize_t)n};
  const hid_t sp = a.Screate_simple(1, dims, nullptr);
  const hid_t ds = a.Dcreate2(id_, name.c_str(), type_of(a, t), sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
  herr_t rc = ds < 0 ? -1 : 0;
  if (ds >= 0 && n) rc = a.Dwrite(ds, type_of(a, t), S_ALL, S_ALL, P_DEFAULT, data);  // empty: no write
  if (ds >= 0) a.Dclose(ds);
  a.Sclose(sp);
  if (rc < 0) throw std::runtime_error("Error! :: Unable to write datasetName " + name + " in  the file " + path_);
}

void File::attribute(const std::string& object, const std::string& name, unsigned long value) {
  const Api& a = api();
  group(object);
  // a scalar attribute on the group (h5stream's write_atr on the group of saveqOperator)
  const hid_t sp = a.Screate(S_SCALAR);
  const hid_t at = a.Acreate_by_name(id_, object.c_str(), name.c_str(), *a.t_ulong, sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
  herr_t rc = at < 0 ? -1 : a.Awrite(at, *a.t_ulong, &value);
  if (at >= 0) a.Aclose(at);
  a.Sclose(sp);
  if (rc < 0) throw std::runtime_error("nrgk: cannot write attribute " + name + " on " + object + " in " + path_);
}

size_t File::size(const std::string& name) const {
  const Api& a = api();
  if (a.Lexists(id_, name.c_str(), P_DEFAULT) <= 0) throw std::runtime_error("Error :: Unable to read dataset! " + name);
  const hid_t ds = a.Dopen2(id_, name.c_str(), P_DEFAULT);
  if (ds < 0) throw std::runtime_error("Error :: Unable to read dataset! " + name);
  const hid_t sp = a.Dget_space(ds);
  const hssize_t n = a.Sget_simple_extent_npoints(sp);
  a.Sclose(sp);
  a.Dclose(ds);
  return n < 0 ? 0 : (size_t)n;
}

void File::read(const std::string& name, Type t, void* data, size_t n) const {
  const Api& a = api();
  if (size(name) != n) throw std::runtime_error("nrgk: dataset " + name + " has another size");
  if (n == 0) return;
  const hid_t ds = a.Dopen2(id_, name.c_str(), P_DEFAULT);
  const herr_t rc = a.Dread(ds, type_of(a, t), S_ALL, S_ALL, P_DEFAULT, data);
  a.Dclose(ds);
  if (rc < 0) throw std::runtime_error("Error :: Unable to read dataset! " + name);
}

bool File::exists(const std::string& name) const {
  const Api& a = api();
  // every level must exist for H5Lexists on the full path
  for (size_t p = 1; p <= name.size(); p++) {
    if (p != name.size() && name[p] != '/') continue;
    const std::string sub = name.substr(0, p);
    if (sub.empty() || sub == "/") continue;
    if (a.Lexists(id_, sub.c_str(), P_DEFAULT) <= 0) return false;
  }
  return true;
}

}  // namespace h5
}  // namespace nrgk
