// msgpack.h -- a small MessagePack writer and reader (host code) for the
// snapshot document: maps with string keys, unsigned integers, floats,
// strings and binary blobs, as nlohmann::json::to_msgpack writes tcnn's
// Trainer::serialize output (trainer.h:267-278). Multi-byte lengths and
// numbers are big-endian (MessagePack specification).
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace npmg {
namespace mp {

struct Writer {
    std::vector<uint8_t> out;

    void byte(uint8_t b) { out.push_back(b); }
    void be(uint64_t v, int bytes) {
        for (int i = bytes - 1; i >= 0; --i) out.push_back((uint8_t)(v >> (8 * i)));
    }
    void map(uint32_t n) {
        if (n < 16) byte((uint8_t)(0x80u | n));
        else if (n < 65536) { byte(0xde); be(n, 2); }
        else { byte(0xdf); be(n, 4); }
    }
    void str(const std::string& s) {
        const size_t n = s.size();
        if (n < 32) byte((uint8_t)(0xa0u | n));
        else if (n < 256) { byte(0xd9); be(n, 1); }
        else if (n < 65536) { byte(0xda); be(n, 2); }
        else { byte(0xdb); be(n, 4); }
        out.insert(out.end(), s.begin(), s.end());
    }
    void uint(uint64_t v) {
        if (v < 128) byte((uint8_t)v);
        else if (v < 256) { byte(0xcc); be(v, 1); }
        else if (v < 65536) { byte(0xcd); be(v, 2); }
        else if (v <= 0xffffffffull) { byte(0xce); be(v, 4); }
        else { byte(0xcf); be(v, 8); }
    }
    void f64(double d) {
        uint64_t u;
        std::memcpy(&u, &d, 8);
        byte(0xcb);
        be(u, 8);
    }
    // Header of a binary blob of n bytes; the caller appends the bytes.
    void bin_header(size_t n) {
        if (n < 256) { byte(0xc4); be(n, 1); }
        else if (n < 65536) { byte(0xc5); be(n, 2); }
        else { byte(0xc6); be(n, 4); }
    }
    void bin(const void* p, size_t n) {
        bin_header(n);
        const uint8_t* b = static_cast<const uint8_t*>(p);
        out.insert(out.end(), b, b + n);
    }
};

struct Value {
    enum Kind { NIL, BOOL, UINT, INT, F64, STR, BIN, MAP, ARR } kind = NIL;
    uint64_t u = 0;
    int64_t i = 0;
    double d = 0.0;
    std::string s;
    const uint8_t* b = nullptr;  // BIN: points into the parsed buffer
    size_t bn = 0;
    std::vector<std::pair<std::string, Value>> map;
    std::vector<Value> arr;

    const Value* get(const char* key) const {
        if (kind != MAP) return nullptr;
        for (const auto& kv : map)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool is_number() const { return kind == UINT || kind == INT || kind == F64; }
    double number() const { return kind == UINT ? (double)u : kind == INT ? (double)i : d; }
};

class Reader {
public:
    Reader(const uint8_t* p, size_t n) : p_(p), n_(n) {}

    Value parse() {
        Value v = value(0);
        if (pos_ != n_) throw std::runtime_error("snapshot: trailing bytes after the document");
        return v;
    }

private:
    const uint8_t* p_;
    size_t n_, pos_ = 0;

    void need(size_t k) const {
        if (k > n_ - pos_) throw std::runtime_error("snapshot: truncated document");
    }
    uint64_t be(int bytes) {
        need((size_t)bytes);
        uint64_t v = 0;
        for (int i = 0; i < bytes; ++i) v = (v << 8) | p_[pos_++];
        return v;
    }
    std::string text(size_t n) {
        need(n);
        std::string s(reinterpret_cast<const char*>(p_ + pos_), n);
        pos_ += n;
        return s;
    }
    Value blob(size_t n) {
        need(n);
        Value v;
        v.kind = Value::BIN;
        v.b = p_ + pos_;
        v.bn = n;
        pos_ += n;
        return v;
    }
    Value map_of(size_t n, int depth) {
        Value v;
        v.kind = Value::MAP;
        for (size_t k = 0; k < n; ++k) {
            Value key = value(depth + 1);
            if (key.kind != Value::STR) throw std::runtime_error("snapshot: map keys must be strings");
            Value val = value(depth + 1);
            v.map.emplace_back(std::move(key.s), std::move(val));
        }
        return v;
    }
    Value arr_of(size_t n, int depth) {
        Value v;
        v.kind = Value::ARR;
        for (size_t k = 0; k < n; ++k) v.arr.push_back(value(depth + 1));
        return v;
    }
    Value value(int depth) {
        if (depth > 32) throw std::runtime_error("snapshot: nesting too deep");
        need(1);
        const uint8_t t = p_[pos_++];
        Value v;
        if (t < 0x80) { v.kind = Value::UINT; v.u = t; return v; }
        if (t >= 0xe0) { v.kind = Value::INT; v.i = (int8_t)t; return v; }
        if ((t & 0xf0) == 0x80) return map_of(t & 0x0f, depth);
        if ((t & 0xf0) == 0x90) return arr_of(t & 0x0f, depth);
        if ((t & 0xe0) == 0xa0) { v.kind = Value::STR; v.s = text(t & 0x1f); return v; }
        switch (t) {
        case 0xc0: return v;
        case 0xc2: case 0xc3: v.kind = Value::BOOL; v.u = t == 0xc3; return v;
        case 0xc4: return blob(be(1));
        case 0xc5: return blob(be(2));
        case 0xc6: return blob(be(4));
        case 0xc7: { const size_t n = be(1); be(1); return blob(n); }  // ext: binary with a subtype
        case 0xc8: { const size_t n = be(2); be(1); return blob(n); }
        case 0xc9: { const size_t n = be(4); be(1); return blob(n); }
        case 0xca: { const uint32_t u = (uint32_t)be(4); float f; std::memcpy(&f, &u, 4); v.kind = Value::F64; v.d = f; return v; }
        case 0xcb: { const uint64_t u = be(8); std::memcpy(&v.d, &u, 8); v.kind = Value::F64; return v; }
        case 0xcc: v.kind = Value::UINT; v.u = be(1); return v;
        case 0xcd: v.kind = Value::UINT; v.u = be(2); return v;
        case 0xce: v.kind = Value::UINT; v.u = be(4); return v;
        case 0xcf: v.kind = Value::UINT; v.u = be(8); return v;
        case 0xd0: v.kind = Value::INT; v.i = (int8_t)be(1); return v;
        case 0xd1: v.kind = Value::INT; v.i = (int16_t)be(2); return v;
        case 0xd2: v.kind = Value::INT; v.i = (int32_t)be(4); return v;
        case 0xd3: v.kind = Value::INT; v.i = (int64_t)be(8); return v;
        case 0xd4: be(1); return blob(1);  // fixext 1..16
        case 0xd5: be(1); return blob(2);
        case 0xd6: be(1); return blob(4);
        case 0xd7: be(1); return blob(8);
        case 0xd8: be(1); return blob(16);
        case 0xd9: v.kind = Value::STR; v.s = text(be(1)); return v;
        case 0xda: v.kind = Value::STR; v.s = text(be(2)); return v;
        case 0xdb: v.kind = Value::STR; v.s = text(be(4)); return v;
        case 0xdc: return arr_of(be(2), depth);
        case 0xdd: return arr_of(be(4), depth);
        case 0xde: return map_of(be(2), depth);
        case 0xdf: return map_of(be(4), depth);
        default: throw std::runtime_error("snapshot: unsupported MessagePack type");
        }
    }
};

}  // namespace mp
}  // namespace npmg
