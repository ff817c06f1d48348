// json.hpp -- minimal JSON reader for the circuit wire format.
//
// The reference decodes `model_dump_json()` text with serde_json into the
// schema of src/types.rs:7-135.  This reader builds a small DOM; the schema
// checks (required fields, integer vs float, tags) are done by the decoder in
// circuit.hpp.  Numbers keep their source text so integers up to 2^64-1 are
// exact and floats are converted once with strtod (correctly rounded).
#pragma once
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dqh {

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    std::string text;  // number source text, or string value
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // insertion order kept

    const Json *get(const std::string &key) const {
        // serde keeps the last occurrence of a duplicated key for maps; for
        // struct fields it reports "duplicate field".  Circuits from
        // model_dump_json never repeat keys, so the first match is used.
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool is_integer() const {
        if (type != Number) return false;
        for (char c : text)
            if (c == '.' || c == 'e' || c == 'E') return false;
        return true;
    }
};

class JsonParser {
   public:
    explicit JsonParser(const std::string &s) : s_(s) {}

    Json parse() {
        Json v = value(0);
        ws();
        if (i_ != s_.size()) err("trailing characters");
        return v;
    }

   private:
    const std::string &s_;
    size_t i_ = 0;

    [[noreturn]] void err(const std::string &what) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); ++k) {
            if (s_[k] == '\n') {
                ++line;
                col = 1;
            } else {
                ++col;
            }
        }
        throw JsonError(what + " at line " + std::to_string(line) + " column " + std::to_string(col));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\n' || s_[i_] == '\r' || s_[i_] == '\t')) ++i_;
    }
    bool lit(const char *w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    Json value(int depth) {
        if (depth > 128) err("recursion limit exceeded");
        ws();
        if (i_ >= s_.size()) err("EOF while parsing a value");
        const char c = s_[i_];
        Json v;
        if (c == '{') {
            v.type = Json::Object;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') err("key must be a string");
                std::string key = str();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') err("expected `:`");
                ++i_;
                v.obj.emplace_back(std::move(key), value(depth + 1));
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    return v;
                }
                err("expected `,` or `}`");
            }
        }
        if (c == '[') {
            v.type = Json::Array;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    return v;
                }
                err("expected `,` or `]`");
            }
        }
        if (c == '"') {
            v.type = Json::String;
            v.text = str();
            return v;
        }
        if (lit("true")) {
            v.type = Json::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.type = Json::Bool;
            return v;
        }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            v.type = Json::Number;
            const size_t start = i_;
            if (s_[i_] == '-') ++i_;
            if (i_ >= s_.size() || !(s_[i_] >= '0' && s_[i_] <= '9')) err("invalid number");
            if (s_[i_] == '0') {
                ++i_;
            } else {
                while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
            }
            if (i_ < s_.size() && s_[i_] == '.') {
                ++i_;
                if (i_ >= s_.size() || !(s_[i_] >= '0' && s_[i_] <= '9')) err("invalid number");
                while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
            }
            if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
                ++i_;
                if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
                if (i_ >= s_.size() || !(s_[i_] >= '0' && s_[i_] <= '9')) err("invalid number");
                while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
            }
            v.text = s_.substr(start, i_ - start);
            return v;
        }
        err("expected value");
    }
    static void put_utf8(std::string &out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (i_ + 4 > s_.size()) err("EOF while parsing a string");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            const char h = s_[i_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= uint32_t(h - '0');
            else if (h >= 'a' && h <= 'f') v |= uint32_t(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= uint32_t(h - 'A' + 10);
            else err("invalid escape");
        }
        return v;
    }
    std::string str() {
        ++i_;  // opening quote
        std::string out;
        for (;;) {
            if (i_ >= s_.size()) err("EOF while parsing a string");
            const char c = s_[i_++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) err("control character while parsing a string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) err("EOF while parsing a string");
            const char e = s_[i_++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                uint32_t cp = hex4();
                if (cp >= 0xD800 && cp < 0xDC00) {
                    if (i_ + 2 <= s_.size() && s_[i_] == '\\' && s_[i_ + 1] == 'u') {
                        i_ += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) err("lone leading surrogate in hex escape");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else {
                        err("lone leading surrogate in hex escape");
                    }
                }
                put_utf8(out, cp);
                break;
            }
            default:
                err("invalid escape");
            }
        }
    }
};

}  // namespace dqh
