// A self-contained subset of the HDF5 file format for the native-engine
// file contract of kspaceFirstOrder3DC / kspaceFirstOrder2DC: the input
// files the reference writes (+kwave/+legacy/private/kspaceFirstOrder_
// saveToDisk.m with writeMatrix.m and writeAttributes.m) and the output
// files it reads back with h5read (kspaceFirstOrder3DC.m:571-777).
//
// Reader: superblock versions 0 and 1; object headers version 1 and 2 with
// continuation blocks; groups stored as symbol tables (B-tree version 1 +
// local heap) or as compact link messages; datasets with contiguous,
// compact or chunked storage (chunk index B-tree version 1) and the deflate,
// shuffle and Fletcher-32 filters; little-endian integer and floating-point
// data (converted to double or uint64 on read); fixed-length string
// attributes.  Anything else is reported as unsupported.
//
// Writer: superblock version 0, groups as symbol tables, datasets stored
// contiguously in float32, float64 or uint64, fixed-length string
// attributes: the layout the HDF5 library itself writes by default, which
// MATLAB's h5read reads.
//
// Dimensions are given in HDF5 (C) order, slowest first; a MATLAB array of
// size [Nx Ny Nz] is stored as dims {Nz, Ny, Nx} with x fastest in memory.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace kwh5 {

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

enum class Kind { Float, Int, UInt, String };

struct Type {
  Kind kind = Kind::Float;
  int size = 0;  // bytes per element
};

class Reader {
 public:
  explicit Reader(const std::string& path);
  ~Reader();
  bool has(const std::string& path) const;
  // dims (HDF5 order) and element type of a dataset
  std::vector<uint64_t> dims(const std::string& path) const;
  Type type(const std::string& path) const;
  // every element converted to double / uint64 (integers are range-checked)
  std::vector<double> read_double(const std::string& path) const;
  std::vector<uint64_t> read_u64(const std::string& path) const;
  double scalar(const std::string& path) const;
  // a string attribute of an object ("" for the root group); throws if absent
  std::string attr_string(const std::string& obj, const std::string& name) const;
  bool has_attr(const std::string& obj, const std::string& name) const;
  // names of the members of a group ("" for the root group), sorted
  std::vector<std::string> list(const std::string& group) const;

 private:
  struct Impl;
  std::unique_ptr<Impl> p_;
};

class Writer {
 public:
  // Datasets and groups are described first and written at once by
  // write(); the data pointers must stay valid until then.
  void add_group(const std::string& path);
  void add_f32(const std::string& path, std::vector<uint64_t> dims, const float* data);
  void add_f64(const std::string& path, std::vector<uint64_t> dims, const double* data);
  void add_u64(const std::string& path, std::vector<uint64_t> dims, const uint64_t* data);
  // string attribute of a dataset or group ("" for the root group)
  void attr(const std::string& obj, const std::string& name, const std::string& value);
  void write(const std::string& file) const;

 private:
  struct Obj {
    bool group = false;
    std::vector<uint64_t> dims;
    Kind kind = Kind::Float;
    int size = 0;
    const void* data = nullptr;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::map<std::string, std::string> children;  // name -> path
  };
  std::map<std::string, Obj> objs_;  // by path; "" is the root group
  Obj& ensure_group(const std::string& path);
  void add_data(const std::string& path, std::vector<uint64_t> dims, Kind k, int size,
                const void* data);
};

}  // namespace kwh5
