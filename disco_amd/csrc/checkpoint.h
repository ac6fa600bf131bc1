// Checkpoint container: named, typed, dense arrays in one file.
//
// A checkpoint holds the datasets of the reference's HDF5 checkpoint
// (Output/h5out.c:366-622: groups Grid, Data, Pars, Opts; dataset paths such
// as "Grid/r_jph" or "Data/Cells") with their shapes and element types.  The
// file is not HDF5 (libhdf5 is not part of this build): it is a flat
// sequence of records, each
//   u32 name length, name bytes, u32 type (0 f64, 1 i32, 2 bytes), u32 rank,
//   u64 dims[rank], u64 payload bytes, payload (native little-endian),
// after the header "DISCOCKP", u32 version (1), u32 record count.
// disco_amd/checkpoint.py reads it into a dict of NumPy arrays.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace disco {
namespace ckpt {

enum Type : uint32_t { F64 = 0, I32 = 1, BYTES = 2 };

struct Dataset {
    std::string name;
    uint32_t type = F64;
    std::vector<uint64_t> dims;
    std::vector<char> data;

    size_t count() const;
    const double* f64() const { return reinterpret_cast<const double*>(data.data()); }
    const int32_t* i32() const { return reinterpret_cast<const int32_t*>(data.data()); }
};

Dataset make_f64(const std::string& name, std::vector<uint64_t> dims, const double* v);
Dataset make_i32(const std::string& name, std::vector<uint64_t> dims, const int32_t* v);
Dataset make_str(const std::string& name, const std::string& s);

// 0 on success; *err holds the reason otherwise
int write_file(const std::string& path, const std::vector<Dataset>& sets, std::string* err);
int read_file(const std::string& path, std::vector<Dataset>* sets, std::string* err);
// the dataset of that name, or null
const Dataset* find(const std::vector<Dataset>& sets, const std::string& name);

}  // namespace ckpt
}  // namespace disco
