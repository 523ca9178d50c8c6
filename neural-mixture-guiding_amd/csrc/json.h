// json.h -- minimal JSON reader for configuration strings.
//
// Parses objects, arrays, strings (with the standard escapes), numbers,
// true/false/null. Throws std::runtime_error with a position on malformed
// input; the C ABI turns that into NPMG_EINVAL.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace npmg {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    bool contains(const std::string& k) const { return kind == Object && obj.count(k) != 0; }
    const Json& at(const std::string& k) const {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end()) throw std::runtime_error("missing key '" + k + "'");
        return it->second;
    }
    double number(const std::string& k, double def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        if (v.kind == Bool) return v.b ? 1.0 : 0.0;
        if (v.kind != Number) throw std::runtime_error("key '" + k + "' must be a number");
        return v.num;
    }
    std::string string(const std::string& k, const std::string& def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        if (v.kind != String) throw std::runtime_error("key '" + k + "' must be a string");
        return v.str;
    }
    bool boolean(const std::string& k, bool def) const {
        if (!contains(k)) return def;
        const Json& v = at(k);
        if (v.kind == Number) return v.num != 0;
        if (v.kind != Bool) throw std::runtime_error("key '" + k + "' must be a boolean");
        return v.b;
    }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& s) : s_(s) {}
    Json parse() {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(p_));
    }
    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\t' || s_[p_] == '\r')) ++p_;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    Json value() {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        const char c = s_[p_];
        Json v;
        if (c == '{') {
            v.kind = Json::Object;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
                std::string k = strv();
                ws();
                if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
                ++p_;
                v.obj[k] = value();
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == '}') { ++p_; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Json::Array;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == ']') { ++p_; return v; }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == ']') { ++p_; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Json::String;
            v.str = strv();
        } else if (lit("true")) {
            v.kind = Json::Bool; v.b = true;
        } else if (lit("false")) {
            v.kind = Json::Bool; v.b = false;
        } else if (lit("null")) {
            v.kind = Json::Null;
        } else {
            const char* b = s_.c_str() + p_;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("bad value");
            v.kind = Json::Number;
            p_ += (size_t)(e - b);
        }
        return v;
    }
    std::string strv() {
        ++p_;  // opening quote
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char c = s_[p_++];
            if (c == '\\') {
                if (p_ >= s_.size()) fail("bad escape");
                const char e = s_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > s_.size()) fail("bad \\u escape");
                    const unsigned cp = (unsigned)std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16);
                    p_ += 4;
                    out += cp < 0x80 ? (char)cp : '?';
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
};

inline Json parse_json(const std::string& s) { return JsonParser(s).parse(); }

}  // namespace npmg
