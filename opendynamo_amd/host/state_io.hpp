// Solver-native state files (drunkseal/opendynamo src/io.rs:1-234):
//   <out_dir>/<prefix>_<f32|f64>_<label:06>.h5
//   group data: 1-D datasets rho rhou rhov rhow E, native state type,
//               var-major device layout (elem-major 64-point blocks)
//   root attrs: nx ny nz fr_order step (u64); t lx ly lz (f64); u_real (3 chars)
//   compression: none | gzip1..gzip9 (chunk 262144 + shuffle + deflate) |
//                blosc-lz4 (clevel 9) | blosc-zstd (clevel 5), byte shuffle
// The loader reads the datasets in their native type, so f64 resume is
// bit-exact (the reference narrows every plane to f32, io.rs:221).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace odh {

struct Compression {
    enum Kind { None, Gzip, BloscLz4, BloscZstd } kind = None;
    int level = 0;
};

Compression parse_compression(const std::string& s);  // io.rs:35-52, throws

// Writes the file and returns its path.  `data` holds 5 planes of
// nx*ny*nz*64 values of the type named by kind ("f32" / "f64").
std::string save_state_h5(uint64_t label, double t, const std::string& kind, const std::string& out_dir,
                          const std::string& prefix, const void* data, uint64_t nx, uint64_t ny, uint64_t nz,
                          double lx, double ly, double lz, Compression c);

struct LoadedState {
    uint64_t nx = 0, ny = 0, nz = 0, fr_order = 0, step = 0;
    double t = 0.0, lx = 0.0, ly = 0.0, lz = 0.0;
    std::string u_real;
    std::vector<unsigned char> data;  // 5 planes in the file's native type
};

// io.rs:163-234: validates mesh, fr_order, u_real, dataset presence and length.
LoadedState load_state_h5(const std::string& path, uint64_t nx, uint64_t ny, uint64_t nz,
                          const std::string& expect_kind);

// Header only (no validation) and the planes widened to f64, for tools.
LoadedState read_state_h5(const std::string& path);

}  // namespace odh
