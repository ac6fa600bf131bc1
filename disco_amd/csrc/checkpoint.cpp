// Checkpoint container (checkpoint.h): writer and reader.
#include "checkpoint.h"

#include <cstdio>
#include <cstring>

namespace disco {
namespace ckpt {

static const char MAGIC[8] = {'D', 'I', 'S', 'C', 'O', 'C', 'K', 'P'};
static const uint32_t VERSION = 1;

size_t Dataset::count() const {
    size_t n = 1;
    for (uint64_t d : dims) n *= (size_t)d;
    return n;
}

static Dataset make(const std::string& name, uint32_t type, std::vector<uint64_t> dims, const void* v, size_t esz) {
    Dataset d;
    d.name = name;
    d.type = type;
    d.dims = std::move(dims);
    d.data.resize(d.count() * esz);
    if (!d.data.empty()) std::memcpy(d.data.data(), v, d.data.size());
    return d;
}

Dataset make_f64(const std::string& name, std::vector<uint64_t> dims, const double* v) {
    return make(name, F64, std::move(dims), v, sizeof(double));
}

Dataset make_i32(const std::string& name, std::vector<uint64_t> dims, const int32_t* v) {
    return make(name, I32, std::move(dims), v, sizeof(int32_t));
}

Dataset make_str(const std::string& name, const std::string& s) {
    return make(name, BYTES, {(uint64_t)s.size()}, s.data(), 1);
}

int write_file(const std::string& path, const std::vector<Dataset>& sets, std::string* err) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        *err = "cannot open " + path + " for writing";
        return -1;
    }
    bool ok = std::fwrite(MAGIC, 1, 8, f) == 8;
    uint32_t hdr[2] = {VERSION, (uint32_t)sets.size()};
    ok = ok && std::fwrite(hdr, sizeof hdr, 1, f) == 1;
    for (const Dataset& d : sets) {
        uint32_t nl = (uint32_t)d.name.size(), ty = d.type, rk = (uint32_t)d.dims.size();
        uint64_t nb = d.data.size();
        ok = ok && std::fwrite(&nl, 4, 1, f) == 1 && std::fwrite(d.name.data(), 1, nl, f) == nl &&
             std::fwrite(&ty, 4, 1, f) == 1 && std::fwrite(&rk, 4, 1, f) == 1 &&
             (rk == 0 || std::fwrite(d.dims.data(), 8, rk, f) == rk) && std::fwrite(&nb, 8, 1, f) == 1 &&
             (nb == 0 || std::fwrite(d.data.data(), 1, nb, f) == nb);
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        *err = "write error on " + path;
        return -1;
    }
    return 0;
}

int read_file(const std::string& path, std::vector<Dataset>* sets, std::string* err) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        *err = "cannot open " + path;
        return -1;
    }
    sets->clear();
    char magic[8];
    uint32_t hdr[2];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, MAGIC, 8) == 0 &&
              std::fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == VERSION;
    for (uint32_t s = 0; ok && s < hdr[1]; ++s) {
        Dataset d;
        uint32_t nl = 0, rk = 0;
        uint64_t nb = 0;
        ok = std::fread(&nl, 4, 1, f) == 1 && nl < 4096;
        if (ok) {
            d.name.resize(nl);
            ok = std::fread(&d.name[0], 1, nl, f) == nl && std::fread(&d.type, 4, 1, f) == 1 &&
                 std::fread(&rk, 4, 1, f) == 1 && rk <= 8 && d.type <= BYTES;
        }
        if (ok) {
            d.dims.resize(rk);
            ok = (rk == 0 || std::fread(d.dims.data(), 8, rk, f) == rk) && std::fread(&nb, 8, 1, f) == 1;
        }
        const size_t esz = d.type == F64 ? 8 : d.type == I32 ? 4 : 1;
        ok = ok && nb == d.count() * esz;
        if (ok) {
            d.data.resize(nb);
            ok = nb == 0 || std::fread(d.data.data(), 1, nb, f) == nb;
        }
        if (ok) sets->push_back(std::move(d));
    }
    std::fclose(f);
    if (!ok) {
        *err = path + " is not a readable checkpoint (version " + std::to_string(VERSION) + ")";
        return -1;
    }
    return 0;
}

const Dataset* find(const std::vector<Dataset>& sets, const std::string& name) {
    for (const Dataset& d : sets)
        if (d.name == name) return &d;
    return nullptr;
}

}  // namespace ckpt
}  // namespace disco
