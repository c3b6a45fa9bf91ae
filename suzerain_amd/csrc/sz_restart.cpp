// Restart files: the reference's ESIO/HDF5 layout of a channel state
// (suzerain/support/field.cpp:51-200 save_coefficients/load_coefficients,
// support.cpp:340-360 and the definition_* savers for the scalars), written
// and read by a small HDF5 implementation of the subset those files use:
// a version 0 superblock, one root group held in a symbol table (one
// B-tree node, one symbol node, one local heap), version 1 object headers,
// contiguous little-endian datasets of float64, int32 or pairs of float64
// (the complex fields, an HDF5 array datatype of two doubles).  Reading also
// follows multi-level group B-trees and object header continuation blocks,
// as in the reference's own files.
//
// A field is stored row-major (Nz, Nx/2+1, Ny, 2): z wavenumbers in order
// (0, 1, ..., -1), the non-negative x wavenumbers, Ny B-spline
// coefficients; modes the channel does not retain are written as zero and
// skipped on reading.
#include "../../include/suzerain_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
void sz_internal_channel_spec(const sz_channel *h, sz_channel_spec *out);
hipStream_t sz_internal_channel_stream(sz_channel *h);
}

namespace {

constexpr uint64_t UNDEF = ~0ULL;
const unsigned char SIG[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};

enum Kind { F64 = 0, I32 = 1, C128 = 2 };

struct Dataset {
    std::string name;
    Kind kind;
    std::vector<uint64_t> dims;
    std::vector<unsigned char> bytes;
};

size_t elem_size(Kind k) { return k == F64 ? 8 : k == I32 ? 4 : 16; }

// ------------------------------------------------------------------ writer
struct Buf {
    std::vector<unsigned char> b;
    size_t size() const { return b.size(); }
    void u8(unsigned v) { b.push_back((unsigned char)v); }
    void u16(unsigned v) { u8(v & 0xff); u8((v >> 8) & 0xff); }
    void u32(uint32_t v) { for (int i = 0; i < 4; ++i) u8((v >> (8 * i)) & 0xff); }
    void u64(uint64_t v) { for (int i = 0; i < 8; ++i) u8((unsigned)((v >> (8 * i)) & 0xff)); }
    void zeros(size_t n) { b.insert(b.end(), n, 0); }
    void pad8() { while (b.size() % 8) u8(0); }
    void put64(size_t at, uint64_t v) { for (int i = 0; i < 8; ++i) b[at + i] = (unsigned char)(v >> (8 * i)); }
    void put32(size_t at, uint32_t v) { for (int i = 0; i < 4; ++i) b[at + i] = (unsigned char)(v >> (8 * i)); }
};

void float64_type(Buf &m)
{
    m.u8(0x11);                    // version 1, class 1 (floating point)
    m.u8(0x20);                    // little endian, mantissa msb implied
    m.u8(63);                      // sign bit location
    m.u8(0);
    m.u32(8);
    m.u16(0);                      // bit offset
    m.u16(64);                     // precision
    m.u8(52);                      // exponent location
    m.u8(11);                      // exponent size
    m.u8(0);                       // mantissa location
    m.u8(52);                      // mantissa size
    m.u32(1023);                   // exponent bias
}

Buf type_message(Kind k)
{
    Buf m;
    if (k == F64) {
        float64_type(m);
    } else if (k == I32) {
        m.u8(0x10);                // version 1, class 0 (fixed point)
        m.u8(0x08);                // little endian, signed
        m.u8(0);
        m.u8(0);
        m.u32(4);
        m.u16(0);
        m.u16(32);
    } else {
        m.u8(0x2a);                // version 2, class 10 (array)
        m.u8(0);
        m.u8(0);
        m.u8(0);
        m.u32(16);
        m.u8(1);                   // one dimension
        m.u8(0);
        m.u8(0);
        m.u8(0);
        m.u32(2);                  // of two
        m.u32(0);                  // permutation index
        float64_type(m);
    }
    m.pad8();
    return m;
}

void message(Buf &hdr, unsigned type, const Buf &body)
{
    hdr.u16(type);
    hdr.u16((unsigned)body.size());
    hdr.u8(0);
    hdr.zeros(3);
    hdr.b.insert(hdr.b.end(), body.b.begin(), body.b.end());
}

void write_file(const std::string &path, std::vector<Dataset> sets)
{
    std::sort(sets.begin(), sets.end(),
              [](const Dataset &a, const Dataset &b) { return a.name < b.name; });
    const unsigned leafK = (unsigned)std::max<size_t>(4, (sets.size() + 1) / 2);
    const unsigned internK = 16;
    Buf f;
    // superblock, root symbol table entry patched below
    f.b.insert(f.b.end(), SIG, SIG + 8);
    f.u8(0); f.u8(0); f.u8(0); f.u8(0); f.u8(0); f.u8(8); f.u8(8); f.u8(0);
    f.u16(leafK);
    f.u16(internK);
    f.u32(0);
    f.u64(0);                      // base address
    f.u64(UNDEF);                  // free space
    const size_t eof_at = f.size();
    f.u64(0);                      // end of file (patched)
    f.u64(UNDEF);                  // driver information
    const size_t root_at = f.size();
    f.u64(0);                      // link name offset
    f.u64(0);                      // object header (patched)
    f.u32(1);                      // cache type: symbol table
    f.u32(0);
    f.u64(0);                      // B-tree (patched)
    f.u64(0);                      // heap (patched)
    // local heap: "" then each name, 8-byte padded
    Buf names;
    names.zeros(8);
    std::vector<uint64_t> name_off;
    for (const Dataset &d : sets) {
        name_off.push_back(names.size());
        names.b.insert(names.b.end(), d.name.begin(), d.name.end());
        names.u8(0);
        names.pad8();
    }
    const size_t heap_at = f.size();
    f.b.insert(f.b.end(), {'H', 'E', 'A', 'P'});
    f.u8(0); f.zeros(3);
    f.u64(names.size());
    f.u64(UNDEF);                  // no free block
    f.u64(heap_at + 32);
    f.b.insert(f.b.end(), names.b.begin(), names.b.end());
    // datasets: object header then data
    std::vector<uint64_t> ohdr;
    for (const Dataset &d : sets) {
        Buf space, fill, layout;
        space.u8(1); space.u8((unsigned)d.dims.size()); space.u8(0); space.u8(0); space.u32(0);
        for (uint64_t n : d.dims) space.u64(n);
        fill.u8(2); fill.u8(1); fill.u8(2); fill.u8(0); fill.pad8();
        Buf hdr;
        const size_t data_size = d.bytes.size();
        layout.u8(3); layout.u8(1);
        const size_t addr_at = layout.size();
        layout.u64(0);
        layout.u64(data_size);
        layout.pad8();
        message(hdr, 0x0001, space);
        message(hdr, 0x0003, type_message(d.kind));
        message(hdr, 0x0005, fill);
        const size_t layout_body = hdr.size() + 8;
        message(hdr, 0x0008, layout);
        f.pad8();
        const size_t at = f.size();
        ohdr.push_back(at);
        f.u8(1); f.u8(0); f.u16(4); f.u32(1); f.u32((uint32_t)hdr.size()); f.u32(0);
        const size_t body_at = f.size();
        f.b.insert(f.b.end(), hdr.b.begin(), hdr.b.end());
        f.pad8();
        const size_t data_at = f.size();
        f.put64(body_at + layout_body + addr_at, data_at);
        f.b.insert(f.b.end(), d.bytes.begin(), d.bytes.end());
    }
    // symbol node with every entry (names ascending)
    f.pad8();
    const size_t snod_at = f.size();
    f.b.insert(f.b.end(), {'S', 'N', 'O', 'D'});
    f.u8(1); f.u8(0); f.u16((unsigned)sets.size());
    for (size_t i = 0; i < sets.size(); ++i) {
        f.u64(name_off[i]);
        f.u64(ohdr[i]);
        f.u32(0); f.u32(0);
        f.zeros(16);
    }
    f.zeros((2 * leafK - sets.size()) * 40);
    // group B-tree: one leaf-level node, keys "" and the last name
    const size_t tree_at = f.size();
    f.b.insert(f.b.end(), {'T', 'R', 'E', 'E'});
    f.u8(0); f.u8(0); f.u16(1);
    f.u64(UNDEF); f.u64(UNDEF);
    f.u64(0);
    f.u64(snod_at);
    f.u64(sets.empty() ? 0 : name_off.back());
    f.zeros((size_t)(2 * internK - 1) * 16);
    // root group object header: the symbol table message
    f.pad8();
    const size_t root_hdr = f.size();
    Buf st;
    st.u64(tree_at);
    st.u64(heap_at);
    Buf rh;
    message(rh, 0x0011, st);
    f.u8(1); f.u8(0); f.u16(1); f.u32(1); f.u32((uint32_t)rh.size()); f.u32(0);
    f.b.insert(f.b.end(), rh.b.begin(), rh.b.end());
    f.put64(root_at + 8, root_hdr);
    f.put64(root_at + 24, tree_at);
    f.put64(root_at + 32, heap_at);
    f.put64(eof_at, f.size());
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("cannot create " + path);
    const size_t w = std::fwrite(f.b.data(), 1, f.size(), fp);
    const int c = std::fclose(fp);
    if (w != f.size() || c != 0) throw std::runtime_error("write failed: " + path);
}

// ------------------------------------------------------------------ reader
struct Reader {
    std::vector<unsigned char> d;
    std::map<std::string, uint64_t> objects;

    explicit Reader(const std::string &path)
    {
        FILE *fp = std::fopen(path.c_str(), "rb");
        if (!fp) throw std::runtime_error("cannot open " + path);
        std::fseek(fp, 0, SEEK_END);
        const long n = std::ftell(fp);
        std::fseek(fp, 0, SEEK_SET);
        d.resize(n > 0 ? (size_t)n : 0);
        const size_t r = d.empty() ? 0 : std::fread(d.data(), 1, d.size(), fp);
        std::fclose(fp);
        if (r != d.size() || d.size() < 96 || std::memcmp(d.data(), SIG, 8) != 0)
            throw std::runtime_error(path + ": not an HDF5 file");
        if (d[8] != 0) throw std::runtime_error(path + ": superblock version not 0");
        if (d[13] != 8 || d[14] != 8) throw std::runtime_error(path + ": offsets not 8 bytes");
        const uint64_t root = u64(56 + 8);
        group(root);
    }
    void need(uint64_t at, uint64_t n) const
    {
        if (at > d.size() || n > d.size() - at) throw std::runtime_error("HDF5 read past end");
    }
    unsigned u8(uint64_t o) const { need(o, 1); return d[o]; }
    unsigned u16(uint64_t o) const { need(o, 2); return d[o] | (d[o + 1] << 8); }
    uint32_t u32(uint64_t o) const
    {
        need(o, 4);
        uint32_t v = 0;
        for (int i = 3; i >= 0; --i) v = (v << 8) | d[o + i];
        return v;
    }
    uint64_t u64(uint64_t o) const
    {
        need(o, 8);
        uint64_t v = 0;
        for (int i = 7; i >= 0; --i) v = (v << 8) | d[o + i];
        return v;
    }
    std::string cstr(uint64_t o) const
    {
        std::string s;
        while (true) {
            const unsigned c = u8(o++);
            if (!c) break;
            s.push_back((char)c);
        }
        return s;
    }
    struct Msg { unsigned type; uint64_t body, size; };
    std::vector<Msg> messages(uint64_t at) const
    {
        if (u8(at) != 1) throw std::runtime_error("object header version not 1");
        const unsigned n = u16(at + 2);
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{at + 16, u32(at + 8)}};
        std::vector<Msg> out;
        for (size_t bi = 0; bi < blocks.size() && out.size() < n; ++bi) {
            uint64_t o = blocks[bi].first;
            const uint64_t end = o + blocks[bi].second;
            while (o + 8 <= end && out.size() < n) {
                const Msg m{u16(o), o + 8, u16(o + 2)};
                if (m.type == 0x10) blocks.push_back({u64(m.body), u64(m.body + 8)});
                out.push_back(m);
                o = m.body + m.size;
            }
        }
        return out;
    }
    void tree(uint64_t node, uint64_t heapdata, int depth)
    {
        if (depth > 32 || std::memcmp(&d[node], "TREE", 4) != 0 || u8(node + 4) != 0)
            throw std::runtime_error("group B-tree node expected");
        const unsigned level = u8(node + 5), used = u16(node + 6);
        for (unsigned i = 0; i < used; ++i) {
            const uint64_t child = u64(node + 24 + 8 + 16 * i);
            if (level > 0) { tree(child, heapdata, depth + 1); continue; }
            need(child, 8);
            if (std::memcmp(&d[child], "SNOD", 4) != 0) throw std::runtime_error("symbol node expected");
            const unsigned ns = u16(child + 6);
            for (unsigned k = 0; k < ns; ++k) {
                const uint64_t e = child + 8 + 40 * k;
                objects[cstr(heapdata + u64(e))] = u64(e + 8);
            }
        }
    }
    void group(uint64_t at)
    {
        for (const Msg &m : messages(at)) {
            if (m.type != 0x11) continue;
            const uint64_t btree = u64(m.body), heap = u64(m.body + 8);
            need(heap, 32);
            if (std::memcmp(&d[heap], "HEAP", 4) != 0) throw std::runtime_error("local heap expected");
            tree(btree, u64(heap + 24), 0);
            return;
        }
        throw std::runtime_error("root group has no symbol table");
    }
    bool has(const std::string &name) const { return objects.count(name) != 0; }

    // Kind of the datatype message at o (float64, int32/64, pairs of float64).
    Kind kind(uint64_t o, size_t &size) const
    {
        const unsigned cls = u8(o) & 0x0f;
        size = u32(o + 4);
        if ((cls == 0 || cls == 1) && (u8(o + 1) & 1)) throw std::runtime_error("big-endian data");
        if (cls == 1 && size == 8) return F64;
        if (cls == 0 && (size == 4 || size == 8)) return I32;
        if (cls == 10) {
            const unsigned ver = u8(o) >> 4, nd = u8(o + 8);
            uint64_t q = o + 8 + (ver == 2 ? 4 : 1);
            if (nd != 1 || u32(q) != 2) throw std::runtime_error("array datatype not 2 values");
            q += 4 * nd * (ver == 2 ? 2 : 1);
            size_t bs;
            if (kind(q, bs) != F64) throw std::runtime_error("array of non-float64");
            return C128;
        }
        throw std::runtime_error("unsupported datatype");
    }

    Dataset read(const std::string &name) const
    {
        auto it = objects.find(name);
        if (it == objects.end()) throw std::runtime_error("no dataset " + name);
        Dataset ds;
        ds.name = name;
        bool have_space = false, have_type = false, have_layout = false;
        size_t esize = 0;
        uint64_t addr = UNDEF, size = 0;
        for (const Msg &m : messages(it->second)) {
            if (m.type == 0x01) {
                const unsigned ver = u8(m.body), nd = u8(m.body + 1);
                const uint64_t p = m.body + (ver == 1 ? 8 : 4);
                for (unsigned i = 0; i < nd; ++i) ds.dims.push_back(u64(p + 8 * i));
                have_space = true;
            } else if (m.type == 0x03) {
                ds.kind = kind(m.body, esize);
                have_type = true;
            } else if (m.type == 0x08) {
                if (u8(m.body) != 3 || u8(m.body + 1) != 1)
                    throw std::runtime_error(name + ": only contiguous layouts are read");
                addr = u64(m.body + 2);
                size = u64(m.body + 10);
                have_layout = true;
            }
        }
        if (!have_space || !have_type || !have_layout) throw std::runtime_error(name + ": not a dataset");
        uint64_t count = 1;
        for (uint64_t n : ds.dims) count *= n;
        if (size < count * esize) throw std::runtime_error(name + ": short data");
        if (addr == UNDEF) {
            ds.bytes.assign(count * elem_size(ds.kind), 0);
            return ds;
        }
        need(addr, count * esize);
        if (ds.kind == I32 && esize == 8) {          // 64-bit integers -> int32
            ds.bytes.resize(count * 4);
            for (uint64_t i = 0; i < count; ++i) {
                const int32_t v = (int32_t)(int64_t)u64(addr + 8 * i);
                std::memcpy(&ds.bytes[4 * i], &v, 4);
            }
        } else {
            ds.bytes.assign(d.begin() + addr, d.begin() + addr + count * esize);
        }
        return ds;
    }
    double scalar(const std::string &name) const
    {
        const Dataset ds = read(name);
        if (ds.bytes.size() < elem_size(ds.kind)) throw std::runtime_error(name + ": empty");
        if (ds.kind == I32) { int32_t v; std::memcpy(&v, ds.bytes.data(), 4); return v; }
        double v;
        std::memcpy(&v, ds.bytes.data(), 8);
        return v;
    }
    std::vector<double> vec(const std::string &name) const
    {
        const Dataset ds = read(name);
        if (ds.kind != F64) throw std::runtime_error(name + ": not float64");
        std::vector<double> v(ds.bytes.size() / 8);
        std::memcpy(v.data(), ds.bytes.data(), ds.bytes.size());
        return v;
    }
};

Dataset f64s(const std::string &name, const std::vector<double> &v)
{
    Dataset d{name, F64, {v.size()}, {}};
    d.bytes.resize(v.size() * 8);
    std::memcpy(d.bytes.data(), v.data(), d.bytes.size());
    return d;
}

Dataset f64(const std::string &name, double v) { return f64s(name, {v}); }

Dataset i32(const std::string &name, int v)
{
    Dataset d{name, I32, {1}, std::vector<unsigned char>(4)};
    std::memcpy(d.bytes.data(), &v, 4);
    return d;
}

int inorder(int i, int n) { return i < n / 2 + 1 ? i : i - n; }

int fail(char *err, int errlen, const std::string &m)
{
    if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
    return -1;
}

const char *FIELD_NAMES[5] = {"rho_E", "rho_u", "rho_v", "rho_w", "rho"};

}  // namespace

extern "C" {

// Internal (tests): write nsets datasets (kinds 0 float64, 1 int32, 2 pairs
// of float64) with ranks and row-major dims, or read one by name (dims and
// kind out; data copied when capacity bytes suffice).
int sz_internal_h5_write(const char *path, int nsets, const char *const *names,
                         const int *kinds, const int *ranks, const long long *dims,
                         const void *const *data, char *err, int errlen)
{
    try {
        std::vector<Dataset> sets;
        long long di = 0;
        for (int i = 0; i < nsets; ++i) {
            Dataset d{names[i], (Kind)kinds[i], {}, {}};
            uint64_t count = 1;
            for (int r = 0; r < ranks[i]; ++r) {
                d.dims.push_back((uint64_t)dims[di + r]);
                count *= (uint64_t)dims[di + r];
            }
            di += ranks[i];
            const size_t nb = count * elem_size(d.kind);
            d.bytes.assign((const unsigned char *)data[i], (const unsigned char *)data[i] + nb);
            sets.push_back(std::move(d));
        }
        write_file(path, std::move(sets));
    } catch (const std::exception &e) {
        return fail(err, errlen, e.what());
    }
    return 0;
}

int sz_internal_h5_read(const char *path, const char *name, int *kind, int *rank,
                        long long *dims, int maxrank, void *data, long long capacity,
                        char *err, int errlen)
{
    try {
        Reader r(path);
        const Dataset ds = r.read(name);
        if ((int)ds.dims.size() > maxrank) return fail(err, errlen, "rank too large");
        if (kind) *kind = (int)ds.kind;
        if (rank) *rank = (int)ds.dims.size();
        for (size_t i = 0; i < ds.dims.size(); ++i) dims[i] = (long long)ds.dims[i];
        if (data && capacity >= (long long)ds.bytes.size())
            std::memcpy(data, ds.bytes.data(), ds.bytes.size());
    } catch (const std::exception &e) {
        return fail(err, errlen, e.what());
    }
    return 0;
}

int sz_channel_save(sz_channel *h, int which, const char *path, char *err, int errlen)
{
    if (!h) return fail(err, errlen, "null channel");
    if (which != 0 && which != 1) return fail(err, errlen, "state must be 0 or 1");
    try {
        sz_channel_spec sp;
        sz_internal_channel_spec(h, &sp);
        int Xw = 0, nsys = 0;
        long long F = 0, Fp = 0;
        sz_channel_geometry(h, &Xw, &nsys, &F, &Fp);
        const int Ny = sp.Ny, Nz = sp.Nz, Fx = sp.Nx / 2 + 1;
        const int xkeep = (sp.Nx - 1) / 2 + 1, zabs = (sp.Nz - 1) / 2;
        std::vector<double> host((size_t)F * 5 * 2);
        hipStream_t st = sz_internal_channel_stream(h);
        if (hipMemcpyAsync(host.data(), sz_channel_state(h, which), host.size() * 8,
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return fail(err, errlen, "device copy failed");
        std::vector<Dataset> sets;
        for (int f = 0; f < 5; ++f) {
            Dataset d{FIELD_NAMES[f], C128, {(uint64_t)Nz, (uint64_t)Fx, (uint64_t)Ny}, {}};
            std::vector<double> v((size_t)Nz * Fx * Ny * 2, 0.0);
            for (int n = 0; n < Nz; ++n) {
                const int wz = inorder(n, Nz);
                if ((wz < 0 ? -wz : wz) > zabs) continue;
                const int cz = wz >= 0 ? wz : sp.dNz + wz;
                for (int m = 0; m < Fx && m < xkeep; ++m) {
                    const double *src = host.data() + 2 * ((size_t)f * F + ((size_t)cz * Xw + m) * Ny);
                    std::copy(src, src + 2 * Ny, v.begin() + 2 * (((size_t)n * Fx + m) * Ny));
                }
            }
            d.bytes.resize(v.size() * 8);
            std::memcpy(d.bytes.data(), v.data(), d.bytes.size());
            sets.push_back(std::move(d));
        }
        // the B-spline basis (load_bsplines / save_bsplines) and the grid
        sz_bsplineop *bop = static_cast<sz_bsplineop *>(sz_channel_buffer(h, 8));
        const int nbreak = Ny - sp.k + 2;
        std::vector<double> bp((size_t)nbreak), pts((size_t)Ny), wts((size_t)Ny);
        sz_bspline_breakpoints(Ny, sp.k, 0.0, sp.Ly, sp.htdelta, bp.data());
        for (int j = 0; j < Ny; ++j) pts[(size_t)j] = sz_bsplineop_collocation_point(bop, j);
        sz_bsplineop_integration_coefficients(bop, wts.data());
        std::vector<double> knots;
        for (int i = 0; i < sp.k; ++i) knots.push_back(bp.front());
        for (int i = 1; i < nbreak - 1; ++i) knots.push_back(bp[(size_t)i]);
        for (int i = 0; i < sp.k; ++i) knots.push_back(bp.back());
        const double twopi = 8 * std::atan(1.0);
        std::vector<double> kx((size_t)Fx), kz((size_t)Nz);
        for (int m = 0; m < Fx; ++m) kx[(size_t)m] = twopi / sp.Lx * m;
        for (int n = 0; n < Nz; ++n) kz[(size_t)n] = twopi / sp.Lz * inorder(n, Nz);
        sets.push_back(f64s("breakpoints_y", bp));
        sets.push_back(f64s("collocation_points_y", pts));
        sets.push_back(f64s("integration_weights", wts));
        sets.push_back(f64s("knots", knots));
        sets.push_back(f64s("kx", kx));
        sets.push_back(f64s("kz", kz));
        sets.push_back(i32("k", sp.k));
        sets.push_back(i32("Nx", sp.Nx));
        sets.push_back(i32("Ny", sp.Ny));
        sets.push_back(i32("Nz", sp.Nz));
        sets.push_back(f64("DAFx", (double)sp.dNx / sp.Nx));
        sets.push_back(f64("DAFz", (double)sp.dNz / sp.Nz));
        sets.push_back(f64("htdelta", sp.htdelta));
        sets.push_back(f64("Lx", sp.Lx));
        sets.push_back(f64("Ly", sp.Ly));
        sets.push_back(f64("Lz", sp.Lz));
        const sz_ns_scenario &s = sp.scenario;
        sets.push_back(f64("alpha", s.alpha));
        sets.push_back(f64("beta", s.beta));
        sets.push_back(f64("gamma", s.gamma));
        sets.push_back(f64("Ma", s.Ma));
        sets.push_back(f64("Pr", s.Pr));
        sets.push_back(f64("Re", s.Re));
        const double nan = std::nan("");
        sets.push_back(f64("bulk_rho", sp.constrained ? sp.bulk[4] : nan));
        sets.push_back(f64("bulk_rho_u", sp.constrained ? sp.bulk[1] : nan));
        sets.push_back(f64("bulk_rho_E", sp.constrained ? sp.bulk[0] : nan));
        const sz_isothermal_bc &bc = sp.bc;
        sets.push_back(f64("lower_T", bc.lower_T));
        sets.push_back(f64("lower_u", bc.lower_u));
        sets.push_back(f64("lower_v", bc.lower_v));
        sets.push_back(f64("lower_w", bc.lower_w));
        sets.push_back(f64("upper_T", bc.upper_T));
        sets.push_back(f64("upper_u", bc.upper_u));
        sets.push_back(f64("upper_v", bc.upper_v));
        sets.push_back(f64("upper_w", bc.upper_w));
        sets.push_back(f64("t", sz_channel_time(h)));
        write_file(path, std::move(sets));
    } catch (const std::exception &e) {
        return fail(err, errlen, std::string("sz_channel_save: ") + e.what());
    }
    return 0;
}

int sz_channel_load(sz_channel *h, int which, const char *path, char *err, int errlen)
{
    if (!h) return fail(err, errlen, "null channel");
    if (which != 0 && which != 1) return fail(err, errlen, "state must be 0 or 1");
    try {
        Reader r(path);
        sz_channel_spec sp;
        sz_internal_channel_spec(h, &sp);
        int Xw = 0, nsys = 0;
        long long F = 0, Fp = 0;
        sz_channel_geometry(h, &Xw, &nsys, &F, &Fp);
        const int Ny = sp.Ny;
        // the same B-spline basis (the reference also reprojects between
        // bases; not provided here)
        if ((int)r.scalar("k") != sp.k) throw std::runtime_error("B-spline order differs");
        const std::vector<double> fbp = r.vec("breakpoints_y");
        std::vector<double> bp((size_t)(Ny - sp.k + 2));
        sz_bspline_breakpoints(Ny, sp.k, 0.0, sp.Ly, sp.htdelta, bp.data());
        if (fbp.size() != bp.size()) throw std::runtime_error("B-spline basis differs");
        for (size_t i = 0; i < bp.size(); ++i)
            if (std::fabs(fbp[i] - bp[i]) > 1e-12 * sp.Ly)
                throw std::runtime_error("B-spline breakpoints differ");
        const int xkeep = (sp.Nx - 1) / 2 + 1, zabs = (sp.Nz - 1) / 2;
        std::vector<double> host((size_t)F * 5 * 2, 0.0);
        for (int f = 0; f < 5; ++f) {
            const Dataset ds = r.read(FIELD_NAMES[f]);
            if (ds.kind != C128 || ds.dims.size() != 3 || (int)ds.dims[2] != Ny)
                throw std::runtime_error(std::string(FIELD_NAMES[f]) + ": not (Nz, Nx/2+1, Ny) complex");
            const int Fz = (int)ds.dims[0], Fx = (int)ds.dims[1];
            // the file's retained modes (load_coefficients: the real
            // coefficient count of Fx complex ones keeps Fx's parity)
            const int Nfx = 2 * (Fx - 1) + (Fx & 1);
            const int fxabs = (Nfx - 1) / 2, fzabs = (Fz - 1) / 2;
            const double *v = reinterpret_cast<const double *>(ds.bytes.data());
            for (int n = 0; n < Fz; ++n) {
                const int wz = inorder(n, Fz), az = wz < 0 ? -wz : wz;
                if (az > fzabs || az > zabs) continue;
                const int cz = wz >= 0 ? wz : sp.dNz + wz;
                for (int m = 0; m < Fx; ++m) {
                    if (m > fxabs || m >= xkeep) continue;
                    const double *src = v + 2 * (((size_t)n * Fx + m) * Ny);
                    std::copy(src, src + 2 * Ny,
                              host.begin() + 2 * ((size_t)f * F + ((size_t)cz * Xw + m) * Ny));
                }
            }
        }
        hipStream_t st = sz_internal_channel_stream(h);
        if (hipMemcpyAsync(sz_channel_state(h, which), host.data(), host.size() * 8,
                           hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            throw std::runtime_error("device copy failed");
        if (r.has("t")) sz_channel_set_time(h, r.scalar("t"));
    } catch (const std::exception &e) {
        return fail(err, errlen, std::string("sz_channel_load: ") + e.what());
    }
    return 0;
}

// A channel specification from a restart file: grid, scenario, walls and
// bulk targets (NaN targets disabled); solver NULL (zgbsv).
int sz_restart_read_spec(const char *path, sz_channel_spec *out, char *err, int errlen)
{
    if (!out) return fail(err, errlen, "null spec");
    try {
        Reader r(path);
        sz_channel_spec sp{};
        sp.Nx = (int)r.scalar("Nx");
        sp.Ny = (int)r.scalar("Ny");
        sp.Nz = (int)r.scalar("Nz");
        sp.k = (int)r.scalar("k");
        auto dealiased = [](int N, double daf) {
            int d = (int)std::ceil(N * daf - 1e-9);
            return d < N ? N : d;
        };
        sp.dNx = dealiased(sp.Nx, r.has("DAFx") ? r.scalar("DAFx") : 1.5);
        sp.dNz = dealiased(sp.Nz, r.has("DAFz") ? r.scalar("DAFz") : 1.5);
        if (sp.dNx % 2) ++sp.dNx;              // even transform lengths
        sp.htdelta = r.scalar("htdelta");
        sp.Lx = r.scalar("Lx");
        sp.Ly = r.scalar("Ly");
        sp.Lz = r.scalar("Lz");
        sp.scenario.alpha = r.scalar("alpha");
        sp.scenario.beta = r.scalar("beta");
        sp.scenario.gamma = r.scalar("gamma");
        sp.scenario.Ma = r.scalar("Ma");
        sp.scenario.Pr = r.scalar("Pr");
        sp.scenario.Re = r.scalar("Re");
        sp.bc.enforce_lower = 1;
        sp.bc.enforce_upper = 1;
        sp.bc.lower_T = r.scalar("lower_T");
        sp.bc.lower_u = r.scalar("lower_u");
        sp.bc.lower_v = r.scalar("lower_v");
        sp.bc.lower_w = r.scalar("lower_w");
        sp.bc.upper_T = r.scalar("upper_T");
        sp.bc.upper_u = r.scalar("upper_u");
        sp.bc.upper_v = r.scalar("upper_v");
        sp.bc.upper_w = r.scalar("upper_w");
        sp.linearize = SZ_LINEARIZE_RHOME_XYZ;
        const double nan = std::nan("");
        const double brho = r.has("bulk_rho") ? r.scalar("bulk_rho") : nan;
        const double bu = r.has("bulk_rho_u") ? r.scalar("bulk_rho_u") : nan;
        const double be = r.has("bulk_rho_E") ? r.scalar("bulk_rho_E") : nan;
        sp.constrained = (brho == brho || bu == bu || be == be) ? 1 : 0;
        sp.bulk[0] = be;
        sp.bulk[1] = bu;
        sp.bulk[2] = nan;
        sp.bulk[3] = nan;
        sp.bulk[4] = brho;
        *out = sp;
    } catch (const std::exception &e) {
        return fail(err, errlen, std::string("sz_restart_read_spec: ") + e.what());
    }
    return 0;
}

}  // extern "C"
