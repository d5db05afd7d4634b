#include "state_io.hpp"

#include <blosc.h>
#include <hdf5.h>
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace odh {
namespace {

const char* kVars[5] = {"rho", "rhou", "rhov", "rhow", "E"};  // io.rs:13-15
constexpr uint64_t kChunkElems = 1u << 18;                     // io.rs:8-11
constexpr uint64_t kFrOrder = 3;

// The HDF5 library in use is not built thread-safe; IO threads serialise on this.
std::mutex& h5_mutex() {
    static std::mutex m;
    return m;
}

// ---- HDF5 blosc filter, registered id 32001 -------------------------------
// cd_values: [0] filter revision, [1] blosc format version, [2] type size,
// [3] chunk bytes, [4] compression level, [5] shuffle, [6] compressor code
// (0 blosclz, 1 lz4, 2 lz4hc, 3 snappy, 4 zlib, 5 zstd).  The chunk payload
// is a self-describing blosc frame, so any writer's files decode.
constexpr H5Z_filter_t kBlosc = 32001;
constexpr int kBloscThreads = 4;  // solver.rs:219-226

herr_t blosc_set_local(hid_t dcpl, hid_t type, hid_t) {
    unsigned flags = 0;
    size_t n = 8;
    unsigned v[8] = {0};
    if (H5Pget_filter_by_id2(dcpl, kBlosc, &flags, &n, v, 0, nullptr, nullptr) < 0) return -1;
    if (n < 7) n = 7;
    v[0] = 2;
    v[1] = 2;
    size_t ts = H5Tget_size(type);
    v[2] = (unsigned)ts;
    hsize_t dims[32];
    int nd = H5Pget_chunk(dcpl, 32, dims);
    if (nd < 0) return -1;
    size_t bytes = ts;
    for (int q = 0; q < nd; ++q) bytes *= dims[q];
    v[3] = (unsigned)bytes;
    return H5Pmodify_filter(dcpl, kBlosc, flags, n, v);
}

size_t blosc_filter(unsigned flags, size_t nv, const unsigned v[], size_t nbytes, size_t* buf_size, void** buf) {
    if (!(flags & H5Z_FLAG_REVERSE)) {
        size_t ts = nv > 2 ? v[2] : 1;
        int clevel = nv > 4 ? (int)v[4] : 5, shuffle = nv > 5 ? (int)v[5] : 1, code = nv > 6 ? (int)v[6] : 0;
        static const char* names[6] = {"blosclz", "lz4", "lz4hc", "snappy", "zlib", "zstd"};
        const char* comp = (code >= 0 && code < 6) ? names[code] : "blosclz";
        size_t cap = nbytes + BLOSC_MAX_OVERHEAD;
        void* out = std::malloc(cap);
        if (!out) return 0;
        int r = blosc_compress_ctx(clevel, shuffle, ts, nbytes, *buf, out, cap, comp, 0, kBloscThreads);
        if (r <= 0) {
            std::free(out);
            return 0;
        }
        std::free(*buf);
        *buf = out;
        *buf_size = cap;
        return (size_t)r;
    }
    size_t n = 0, cbytes = 0, blocksize = 0;
    blosc_cbuffer_sizes(*buf, &n, &cbytes, &blocksize);
    void* out = std::malloc(n ? n : 1);
    if (!out) return 0;
    int r = blosc_decompress_ctx(*buf, out, n, kBloscThreads);
    if (r <= 0) {
        std::free(out);
        return 0;
    }
    std::free(*buf);
    *buf = out;
    *buf_size = n;
    return (size_t)r;
}

void init_hdf5() {
    static bool done = false;
    if (done) return;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    if (H5Zfilter_avail(kBlosc) <= 0) {
        static const H5Z_class2_t cls = {H5Z_CLASS_T_VERS, kBlosc, 1, 1, "blosc", nullptr, blosc_set_local,
                                         (H5Z_func_t)blosc_filter};
        if (H5Zregister(&cls) < 0) throw std::runtime_error("cannot register the HDF5 blosc filter");
    }
    done = true;
}

struct H5Handle {  // closes an hid_t with the matching H5*close
    hid_t id;
    herr_t (*close)(hid_t);
    H5Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~H5Handle() {
        if (id >= 0) close(id);
    }
    H5Handle(const H5Handle&) = delete;
    H5Handle& operator=(const H5Handle&) = delete;
    operator hid_t() const { return id; }
};

void check(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error(what);
}

void mkdirs(const std::string& dir) {
    std::string cur;
    for (size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                throw std::runtime_error("cannot create directory " + cur);
        }
        if (i < dir.size()) cur += dir[i];
    }
}

void write_attr_u64(hid_t f, const char* name, uint64_t v) {
    H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
    H5Handle a(H5Acreate2(f, name, H5T_NATIVE_UINT64, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    check(a >= 0 && H5Awrite(a, H5T_NATIVE_UINT64, &v) >= 0, std::string("cannot write attr ") + name);
}
void write_attr_f64(hid_t f, const char* name, double v) {
    H5Handle sp(H5Screate(H5S_SCALAR), H5Sclose);
    H5Handle a(H5Acreate2(f, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    check(a >= 0 && H5Awrite(a, H5T_NATIVE_DOUBLE, &v) >= 0, std::string("cannot write attr ") + name);
}

uint64_t read_attr_u64(hid_t f, const char* name) {
    H5Handle a(H5Aopen(f, name, H5P_DEFAULT), H5Aclose);
    check(a >= 0, std::string("state file has no attribute ") + name);
    uint64_t v = 0;
    check(H5Aread(a, H5T_NATIVE_UINT64, &v) >= 0, std::string("cannot read attribute ") + name);
    return v;
}
double read_attr_f64(hid_t f, const char* name) {
    H5Handle a(H5Aopen(f, name, H5P_DEFAULT), H5Aclose);
    check(a >= 0, std::string("state file has no attribute ") + name);
    double v = 0;
    check(H5Aread(a, H5T_NATIVE_DOUBLE, &v) >= 0, std::string("cannot read attribute ") + name);
    return v;
}
std::string read_attr_str(hid_t f, const char* name) {
    H5Handle a(H5Aopen(f, name, H5P_DEFAULT), H5Aclose);
    check(a >= 0, std::string("state file has no attribute ") + name);
    H5Handle t(H5Aget_type(a), H5Tclose);
    if (H5Tis_variable_str(t) > 0) {
        char* s = nullptr;
        H5Handle mt(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(mt, H5T_VARIABLE);
        check(H5Aread(a, mt, &s) >= 0, "cannot read string attribute");
        std::string out = s ? s : "";
        H5free_memory(s);
        return out;
    }
    size_t n = H5Tget_size(t);
    std::string buf(n, '\0');
    check(H5Aread(a, t, &buf[0]) >= 0, "cannot read string attribute");
    size_t z = buf.find('\0');
    if (z != std::string::npos) buf.resize(z);
    while (!buf.empty() && buf.back() == ' ') buf.pop_back();
    return buf;
}

LoadedState read_file(const std::string& path, bool validate, uint64_t ex, uint64_t ey, uint64_t ez,
                      const std::string& expect_kind) {
    FILE* probe = std::fopen(path.c_str(), "rb");
    if (!probe) throw std::runtime_error("initial_state file not found: " + path);
    std::fclose(probe);
    std::lock_guard<std::mutex> lock(h5_mutex());
    init_hdf5();
    H5Handle f(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    check(f >= 0, "cannot open HDF5 file " + path);
    LoadedState s;
    s.nx = read_attr_u64(f, "nx");
    s.ny = read_attr_u64(f, "ny");
    s.nz = read_attr_u64(f, "nz");
    s.fr_order = read_attr_u64(f, "fr_order");
    s.step = read_attr_u64(f, "step");
    s.t = read_attr_f64(f, "t");
    s.lx = H5Aexists(f, "lx") > 0 ? read_attr_f64(f, "lx") : 0.0;
    s.ly = H5Aexists(f, "ly") > 0 ? read_attr_f64(f, "ly") : 0.0;
    s.lz = H5Aexists(f, "lz") > 0 ? read_attr_f64(f, "lz") : 0.0;
    s.u_real = read_attr_str(f, "u_real");
    if (validate) {
        if (s.nx != ex || s.ny != ey || s.nz != ez)
            throw std::runtime_error("initial_state mesh " + std::to_string(s.nx) + "x" + std::to_string(s.ny) + "x" +
                                     std::to_string(s.nz) + " does not match config " + std::to_string(ex) + "x" +
                                     std::to_string(ey) + "x" + std::to_string(ez) + "; refusing to load");
        if (s.fr_order != kFrOrder)
            throw std::runtime_error("initial_state fr_order " + std::to_string(s.fr_order) +
                                     " != configured FR_ORDER 3");
        if (s.u_real != expect_kind)
            throw std::runtime_error("initial_state u_real '" + s.u_real + "' does not match the configured pipeline ('" +
                                     expect_kind + "'); refusing to load");
    }
    const bool f64 = s.u_real == "f64";
    const size_t elem = f64 ? 8 : 4;
    const uint64_t n = s.nx * s.ny * s.nz * 64;
    H5Handle g(H5Gopen2(f, "data", H5P_DEFAULT), H5Gclose);
    check(g >= 0, "initial_state has no group /data");
    s.data.resize(5 * n * elem);
    for (int v = 0; v < 5; ++v) {
        if (H5Lexists(g, kVars[v], H5P_DEFAULT) <= 0)
            throw std::runtime_error(std::string("initial_state missing dataset /data/") + kVars[v]);
        H5Handle d(H5Dopen2(g, kVars[v], H5P_DEFAULT), H5Dclose);
        check(d >= 0, std::string("cannot open /data/") + kVars[v]);
        H5Handle sp(H5Dget_space(d), H5Sclose);
        hssize_t len = H5Sget_simple_extent_npoints(sp);
        if ((uint64_t)len != n)
            throw std::runtime_error(std::string("initial_state /data/") + kVars[v] + " has " + std::to_string(len) +
                                     " values, expected " + std::to_string(n));
        check(H5Dread(d, f64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                      s.data.data() + (size_t)v * n * elem) >= 0,
              std::string("cannot read /data/") + kVars[v]);
    }
    return s;
}

}  // namespace

Compression parse_compression(const std::string& s) {
    Compression c;
    if (s == "none") return c;
    if (s == "blosc-lz4") return {Compression::BloscLz4, 9};
    if (s == "blosc-zstd") return {Compression::BloscZstd, 5};
    for (int l = 1; l <= 9; ++l)
        if (s == "gzip" + std::to_string(l)) return {Compression::Gzip, l};
    throw std::runtime_error("invalid compression '" + s +
                             "': expected \"none\", \"gzip1\"..\"gzip9\", \"blosc-lz4\" or \"blosc-zstd\"");
}

std::string save_state_h5(uint64_t label, double t, const std::string& kind, const std::string& out_dir,
                          const std::string& prefix, const void* data, uint64_t nx, uint64_t ny, uint64_t nz,
                          double lx, double ly, double lz, Compression c) {
    if (kind != "f32" && kind != "f64") throw std::runtime_error("state kind must be f32 or f64");
    mkdirs(out_dir);
    char name[64];
    std::snprintf(name, sizeof(name), "_%s_%06llu.h5", kind.c_str(), (unsigned long long)label);
    const std::string path = out_dir + "/" + prefix + name;
    const bool f64 = kind == "f64";
    const size_t elem = f64 ? 8 : 4;
    const uint64_t n = nx * ny * nz * 64;
    std::lock_guard<std::mutex> lock(h5_mutex());
    init_hdf5();
    H5Handle f(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    check(f >= 0, "cannot create " + path);
    H5Handle g(H5Gcreate2(f, "data", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    check(g >= 0, "cannot create group data");
    hsize_t dims[1] = {n};
    H5Handle sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
    for (int v = 0; v < 5; ++v) {
        H5Handle dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
        if (c.kind != Compression::None) {
            hsize_t chunk[1] = {n < kChunkElems ? n : kChunkElems};
            check(H5Pset_chunk(dcpl, 1, chunk) >= 0, "H5Pset_chunk");
            if (c.kind == Compression::Gzip) {
                check(H5Pset_shuffle(dcpl) >= 0 && H5Pset_deflate(dcpl, (unsigned)c.level) >= 0, "deflate filter");
            } else {
                unsigned cd[7] = {2, 2, (unsigned)elem, 0, (unsigned)c.level, 1,
                                  c.kind == Compression::BloscLz4 ? 1u : 5u};
                check(H5Pset_filter(dcpl, kBlosc, H5Z_FLAG_OPTIONAL, 7, cd) >= 0, "blosc filter");
            }
        }
        hid_t mt = f64 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT;
        H5Handle d(H5Dcreate2(g, kVars[v], mt, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT), H5Dclose);
        check(d >= 0, std::string("cannot create dataset ") + kVars[v]);
        check(H5Dwrite(d, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, (const char*)data + (size_t)v * n * elem) >= 0,
              std::string("cannot write dataset ") + kVars[v]);
    }
    write_attr_u64(f, "nx", nx);
    write_attr_u64(f, "ny", ny);
    write_attr_u64(f, "nz", nz);
    write_attr_u64(f, "fr_order", kFrOrder);
    write_attr_u64(f, "step", label);
    write_attr_f64(f, "t", t);
    write_attr_f64(f, "lx", lx);
    write_attr_f64(f, "ly", ly);
    write_attr_f64(f, "lz", lz);
    {  // FixedAscii<3>
        H5Handle st(H5Tcopy(H5T_C_S1), H5Tclose);
        H5Tset_size(st, 3);
        H5Tset_strpad(st, H5T_STR_NULLPAD);
        H5Tset_cset(st, H5T_CSET_ASCII);
        H5Handle ssp(H5Screate(H5S_SCALAR), H5Sclose);
        H5Handle a(H5Acreate2(f, "u_real", st, ssp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
        check(a >= 0 && H5Awrite(a, st, kind.c_str()) >= 0, "cannot write attr u_real");
    }
    return path;
}

LoadedState load_state_h5(const std::string& path, uint64_t nx, uint64_t ny, uint64_t nz,
                          const std::string& expect_kind) {
    return read_file(path, true, nx, ny, nz, expect_kind);
}

LoadedState read_state_h5(const std::string& path) { return read_file(path, false, 0, 0, 0, ""); }

}  // namespace odh
