// MATLAB level-5 MAT-file reader for the example instances (the reference's
// loaders, examples/main.cu:41-129 and examples/sparse_sharpe.cu:59-137,
// which read the graph `A` of test_instances/*.mat and the price matrix
// `data` of DataSets/*.mat).  Host-only C++17; link with -lz.
//
// Reads every top-level variable that is a real 2-D numeric or logical
// array, dense or sparse, stored plainly or in miCOMPRESSED (zlib) elements;
// values are converted to double and returned row-major.  Little-endian
// files only (header bytes 126-127 = "IM").
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

namespace gpuboscia {
namespace mat {

struct Matrix {
    std::string name;
    int rows = 0, cols = 0;
    std::vector<double> values;  // row-major: values[r * cols + c]
};

namespace detail {

enum : std::uint32_t {
    kInt8 = 1, kUInt8 = 2, kInt16 = 3, kUInt16 = 4, kInt32 = 5, kUInt32 = 6,
    kSingle = 7, kDouble = 9, kInt64 = 12, kUInt64 = 13, kMatrix = 14, kCompressed = 15,
};
enum : std::uint32_t { kClassSparse = 5 };

struct Element {
    std::uint32_t type = 0;
    std::uint32_t bytes = 0;
    const unsigned char* data = nullptr;
};

inline std::uint32_t u32(const unsigned char* p) {
    std::uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

// The next data element at p (advanced past it, with its 8-byte padding;
// miCOMPRESSED payloads are often written unpadded, so those are not padded).
inline Element next(const unsigned char*& p, const unsigned char* end) {
    if (end - p < 8) throw std::runtime_error("truncated MAT element");
    const std::uint32_t tag = u32(p);
    Element e;
    if (tag >> 16) {  // small data element: 4-byte tag, up to 4 bytes of data
        e.type = tag & 0xffffu;
        e.bytes = tag >> 16;
        e.data = p + 4;
        p += 8;
        return e;
    }
    e.type = tag;
    e.bytes = u32(p + 4);
    e.data = p + 8;
    const std::size_t stored = e.type == kCompressed ? e.bytes : ((std::size_t)e.bytes + 7u) & ~std::size_t(7u);
    if ((std::size_t)(end - e.data) < e.bytes) throw std::runtime_error("truncated MAT payload");
    p = (std::size_t)(end - e.data) < stored ? end : e.data + stored;
    return e;
}

inline std::size_t type_size(std::uint32_t t) {
    switch (t) {
        case kInt8: case kUInt8: return 1;
        case kInt16: case kUInt16: return 2;
        case kInt32: case kUInt32: case kSingle: return 4;
        case kDouble: case kInt64: case kUInt64: return 8;
        default: return 0;
    }
}

inline double number(const Element& e, std::size_t i) {
    const unsigned char* q = e.data + i * type_size(e.type);
    switch (e.type) {
        case kInt8: { std::int8_t v; std::memcpy(&v, q, 1); return v; }
        case kUInt8: return *q;
        case kInt16: { std::int16_t v; std::memcpy(&v, q, 2); return v; }
        case kUInt16: { std::uint16_t v; std::memcpy(&v, q, 2); return v; }
        case kInt32: { std::int32_t v; std::memcpy(&v, q, 4); return v; }
        case kUInt32: { std::uint32_t v; std::memcpy(&v, q, 4); return v; }
        case kSingle: { float v; std::memcpy(&v, q, 4); return v; }
        case kDouble: { double v; std::memcpy(&v, q, 8); return v; }
        case kInt64: { std::int64_t v; std::memcpy(&v, q, 8); return (double)v; }
        case kUInt64: { std::uint64_t v; std::memcpy(&v, q, 8); return (double)v; }
        default: throw std::runtime_error("unsupported MAT numeric type");
    }
}

inline std::size_t count(const Element& e) {
    const std::size_t s = type_size(e.type);
    if (!s) throw std::runtime_error("unsupported MAT numeric type");
    return e.bytes / s;
}

// One miMATRIX body; false for what this reader skips (cells, structs,
// complex or N-d arrays, objects).
inline bool parse_matrix(const unsigned char* p, const unsigned char* end, Matrix& out) {
    const Element flags = next(p, end);
    if (flags.bytes < 4) return false;
    const std::uint32_t f = u32(flags.data);
    const std::uint32_t cls = f & 0xffu;
    const bool complex = (f >> 11) & 1u;
    const Element dims = next(p, end);
    if (dims.type != kInt32 || dims.bytes != 8) return false;
    std::int32_t d[2];
    std::memcpy(d, dims.data, 8);
    const Element name = next(p, end);
    if (complex || cls < 5 || cls > 15 || d[0] < 0 || d[1] < 0) return false;  // numeric classes 6..15, sparse 5
    out.name.assign(reinterpret_cast<const char*>(name.data), name.bytes);
    out.rows = d[0];
    out.cols = d[1];
    out.values.assign((std::size_t)out.rows * out.cols, 0.0);
    if (cls == kClassSparse) {
        const Element ir = next(p, end);
        const Element jc = next(p, end);
        const Element pr = next(p, end);
        if (count(jc) < (std::size_t)out.cols + 1) throw std::runtime_error("invalid sparse MAT column index");
        const bool logical = pr.bytes == 0;
        for (int c = 0; c < out.cols; ++c) {
            const std::size_t a = (std::size_t)number(jc, c), b = (std::size_t)number(jc, c + 1);
            for (std::size_t k = a; k < b; ++k) {
                const std::size_t r = (std::size_t)number(ir, k);
                if (r >= (std::size_t)out.rows) throw std::runtime_error("invalid sparse MAT row index");
                out.values[r * out.cols + c] = logical ? 1.0 : number(pr, k);
            }
        }
        return true;
    }
    const Element re = next(p, end);
    if (count(re) != out.values.size()) throw std::runtime_error("MAT array size does not match its dimensions");
    for (int c = 0; c < out.cols; ++c)  // column-major on file
        for (int r = 0; r < out.rows; ++r)
            out.values[(std::size_t)r * out.cols + c] = number(re, (std::size_t)c * out.rows + r);
    return true;
}

inline std::vector<unsigned char> inflate_all(const Element& e) {
    std::vector<unsigned char> out(std::max<std::size_t>((std::size_t)e.bytes * 4u, 1024u));
    z_stream z{};
    if (inflateInit(&z) != Z_OK) throw std::runtime_error("cannot inflate MAT data");
    z.next_in = const_cast<Bytef*>(e.data);
    z.avail_in = e.bytes;
    int st = Z_OK;
    while (st == Z_OK) {
        if (z.total_out == out.size()) out.resize(out.size() * 2);
        z.next_out = out.data() + z.total_out;
        z.avail_out = (uInt)(out.size() - z.total_out);
        st = inflate(&z, Z_NO_FLUSH);
        if (st == Z_BUF_ERROR && z.avail_in == 0) break;
        if (st == Z_BUF_ERROR) st = Z_OK;
    }
    const std::size_t n = z.total_out;
    inflateEnd(&z);
    if (st != Z_STREAM_END) throw std::runtime_error("cannot inflate MAT data");
    out.resize(n);
    return out;
}

}  // namespace detail

// Every real 2-D numeric or logical variable of a MAT v5 file.
inline std::vector<Matrix> read_all(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open MAT file: " + path);
    const std::vector<unsigned char> file((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (file.size() < 136 || file[126] != 'I' || file[127] != 'M')
        throw std::runtime_error("only little-endian MAT v5 is supported");
    std::vector<Matrix> out;
    const unsigned char* p = file.data() + 128;
    const unsigned char* end = file.data() + file.size();
    while (end - p >= 8) {
        const detail::Element e = detail::next(p, end);
        Matrix m;
        if (e.type == detail::kCompressed) {
            const std::vector<unsigned char> body = detail::inflate_all(e);
            const unsigned char* q = body.data();
            const detail::Element inner = detail::next(q, body.data() + body.size());
            if (inner.type == detail::kMatrix && detail::parse_matrix(inner.data, inner.data + inner.bytes, m))
                out.push_back(std::move(m));
        } else if (e.type == detail::kMatrix) {
            if (detail::parse_matrix(e.data, e.data + e.bytes, m)) out.push_back(std::move(m));
        }
    }
    return out;
}

// The variable `name`; throws when the file has none.
inline Matrix read(const std::string& path, const std::string& name) {
    for (Matrix& m : read_all(path))
        if (m.name == name) return std::move(m);
    throw std::runtime_error("MAT file has no matrix named '" + name + "': " + path);
}

}  // namespace mat
}  // namespace gpuboscia
