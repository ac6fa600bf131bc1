// Test probe (built by tests/test_checkpoint_cpu.py with the host compiler):
// the library's checkpoint container (checkpoint.cpp).
//   ckpt_probe write PATH   writes three known datasets
//   ckpt_probe read PATH    prints "name type dims... : sum" per dataset
#include <cstdio>
#include <cstring>
#include <string>

#include "checkpoint.h"

using namespace disco::ckpt;

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    std::string err;
    if (std::strcmp(argv[1], "write") == 0) {
        const double a[6] = {1.5, -2.25, 3.0, 1e-300, 7.0, -0.0};
        const int32_t b[2] = {-7, 123456789};
        std::vector<Dataset> d = {make_f64("Data/Cells", {2, 3}, a), make_i32("Grid/Np", {1, 2}, b),
                                  make_str("Opts/HYDRO", "mhd")};
        if (write_file(argv[2], d, &err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            return 1;
        }
        return 0;
    }
    std::vector<Dataset> d;
    if (read_file(argv[2], &d, &err)) {
        std::fprintf(stderr, "%s\n", err.c_str());
        return 1;
    }
    for (const Dataset& s : d) {
        double sum = 0.0;
        for (size_t i = 0; i < s.count(); ++i)
            sum += s.type == F64 ? s.f64()[i] : s.type == I32 ? (double)s.i32()[i] : (double)(unsigned char)s.data[i];
        std::printf("%s %u", s.name.c_str(), s.type);
        for (uint64_t v : s.dims) std::printf(" %llu", (unsigned long long)v);
        std::printf(" : %.17g\n", sum);
    }
    return 0;
}
