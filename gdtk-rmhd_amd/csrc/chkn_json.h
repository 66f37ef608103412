// chkn_json.h -- minimal JSON reader for chicken's config.json (objects, arrays, numbers,
// strings, booleans, null).  Numbers are parsed with strtod, as a conforming JSON reader
// would, so "%g"-rounded values in config.json are read exactly as the reference reads them.
#pragma once

#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace chkn {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;

    const Json& operator[](const std::string& key) const
    {
        if (kind != Object) throw std::runtime_error("JSON: not an object when reading '" + key + "'");
        auto it = obj.find(key);
        if (it == obj.end()) throw std::runtime_error("JSON: missing key '" + key + "'");
        return it->second;
    }
    const Json& operator[](size_t i) const
    {
        if (kind != Array || i >= arr.size()) throw std::runtime_error("JSON: bad array index");
        return arr[i];
    }
    bool has(const std::string& key) const { return kind == Object && obj.count(key) > 0; }
    double as_num() const
    {
        if (kind != Number) throw std::runtime_error("JSON: expected a number");
        return num;
    }
    int as_int() const { return (int)as_num(); }
    bool as_bool() const
    {
        if (kind != Bool) throw std::runtime_error("JSON: expected a boolean");
        return b;
    }
    const std::string& as_str() const
    {
        if (kind != String) throw std::runtime_error("JSON: expected a string");
        return str;
    }
    size_t size() const { return kind == Array ? arr.size() : 0; }
};

class JsonParser {
public:
    explicit JsonParser(const std::string& text) : s_(text), p_(0) {}
    Json parse()
    {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t p_;

    [[noreturn]] void fail(const char* what)
    {
        throw std::runtime_error(std::string("JSON parse error: ") + what + " at offset " +
                                 std::to_string(p_));
    }
    void ws()
    {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\t' || s_[p_] == '\r')) ++p_;
    }
    bool lit(const char* w)
    {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    Json value()
    {
        ws();
        if (p_ >= s_.size()) fail("unexpected end");
        char ch = s_[p_];
        Json v;
        if (ch == '{') {
            v.kind = Json::Object;
            ++p_;
            ws();
            if (p_ < s_.size() && s_[p_] == '}') { ++p_; return v; }
            for (;;) {
                ws();
                if (p_ >= s_.size() || s_[p_] != '"') fail("expected key");
                std::string k = string_lit();
                ws();
                if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
                ++p_;
                v.obj[k] = value();
                ws();
                if (p_ < s_.size() && s_[p_] == ',') { ++p_; continue; }
                if (p_ < s_.size() && s_[p_] == '}') { ++p_; break; }
                fail("expected ',' or '}'");
            }
        } else if (ch == '[') {
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
        } else if (ch == '"') {
            v.kind = Json::String;
            v.str = string_lit();
        } else if (lit("true")) {
            v.kind = Json::Bool; v.b = true;
        } else if (lit("false")) {
            v.kind = Json::Bool; v.b = false;
        } else if (lit("null")) {
            v.kind = Json::Null;
        } else {
            const char* start = s_.c_str() + p_;
            char* end = nullptr;
            double d = std::strtod(start, &end);
            if (end == start) fail("bad value");
            v.kind = Json::Number;
            v.num = d;
            p_ += (size_t)(end - start);
        }
        return v;
    }
    std::string string_lit()
    {
        ++p_;  // opening quote
        std::string out;
        while (p_ < s_.size() && s_[p_] != '"') {
            char ch = s_[p_++];
            if (ch == '\\') {
                if (p_ >= s_.size()) fail("bad escape");
                char e = s_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > s_.size()) fail("bad unicode escape");
                    unsigned code = (unsigned)std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16);
                    p_ += 4;
                    if (code < 0x80) out += (char)code;
                    else if (code < 0x800) { out += (char)(0xC0 | (code >> 6)); out += (char)(0x80 | (code & 0x3F)); }
                    else { out += (char)(0xE0 | (code >> 12)); out += (char)(0x80 | ((code >> 6) & 0x3F)); out += (char)(0x80 | (code & 0x3F)); }
                    break;
                }
                default: out += e; break;
                }
            } else {
                out += ch;
            }
        }
        if (p_ >= s_.size()) fail("unterminated string");
        ++p_;
        return out;
    }
};

}  // namespace chkn
