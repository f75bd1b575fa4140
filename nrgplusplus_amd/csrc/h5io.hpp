// Minimal HDF5 writer and reader in the reference's h5stream layout (see h5io.cc).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nrgk {
namespace h5 {

enum class Type { F64, I32, U64 };

class File {
 public:
  enum class Mode { ReadWrite };
  File(const std::string& path, bool write);
  File(const std::string& path, Mode);  // an existing file, for further datasets
  ~File();
  File(const File&) = delete;
  File& operator=(const File&) = delete;
  void group(const std::string& name);  // creates every missing level
  void write(const std::string& name, Type t, const void* data, size_t n);
  void attribute(const std::string& object, const std::string& name, unsigned long value);
  size_t size(const std::string& name) const;
  void read(const std::string& name, Type t, void* data, size_t n) const;
  bool exists(const std::string& name) const;

  void write(const std::string& name, const std::vector<double>& v) { write(name, Type::F64, v.data(), v.size()); }
  void write(const std::string& name, const std::vector<int>& v) { write(name, Type::I32, v.data(), v.size()); }
  void write(const std::string& name, const std::vector<unsigned long>& v) {
    write(name, Type::U64, v.data(), v.size());
  }

 private:
  std::string path_;
  int64_t id_ = -1;
};

}  // namespace h5
}  // namespace nrgk
