// newt_json.hpp — a small strict JSON reader for model files.
//
// Objects keep their members in order and reject duplicate keys; numbers are
// read as doubles (strtod, correctly rounded), which the model loader then
// rounds to f32 as newt's parser does (newt/src/json.rs).  Errors carry the
// byte offset of the offending text.
#pragma once

#include <cctype>
#include <cstdlib>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace newt {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool integral = false;  // written without fraction or exponent
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value* get(const std::string& key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct ParseError {
    size_t at;
    std::string what;
};

class Reader {
  public:
    explicit Reader(const char* text) : s_(text), n_(std::char_traits<char>::length(text)) {}

    Value document() {
        Value v = value(0);
        ws();
        if (i_ != n_) fail("trailing text after the document");
        return v;
    }

  private:
    const char* s_;
    size_t n_, i_ = 0;

    [[noreturn]] void fail(const std::string& what) { throw ParseError{i_, what}; }
    void ws() {
        while (i_ < n_ && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool lit(const char* w) {
        size_t k = 0;
        while (w[k]) {
            if (i_ + k >= n_ || s_[i_ + k] != w[k]) return false;
            ++k;
        }
        i_ += k;
        return true;
    }
    Value value(int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (i_ >= n_) fail("unexpected end of input");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            v.kind = Value::Object;
            ++i_;
            ws();
            if (i_ < n_ && s_[i_] == '}') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                if (i_ >= n_ || s_[i_] != '"') fail("expected a string key");
                std::string key = string();
                for (const auto& kv : v.obj)
                    if (kv.first == key) fail("duplicate key \"" + key + "\"");
                ws();
                if (i_ >= n_ || s_[i_] != ':') fail("expected ':'");
                ++i_;
                v.obj.emplace_back(key, value(depth + 1));
                ws();
                if (i_ < n_ && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < n_ && s_[i_] == '}') {
                    ++i_;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            ++i_;
            ws();
            if (i_ < n_ && s_[i_] == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (i_ < n_ && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < n_ && s_[i_] == ']') {
                    ++i_;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("unexpected character");
    }
    std::string string() {
        ++i_;  // opening quote
        std::string out;
        while (i_ < n_ && s_[i_] != '"') {
            char c = s_[i_++];
            if ((unsigned char)c < 0x20) fail("control character in a string");
            if (c == '\\') {
                if (i_ >= n_) fail("unterminated escape");
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
                        if (i_ + 4 > n_) fail("short \\u escape");
                        unsigned cp = (unsigned)std::strtoul(std::string(s_ + i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i_ >= n_) fail("unterminated string");
        ++i_;
        return out;
    }
    Value number() {
        const size_t start = i_;
        if (s_[i_] == '-') ++i_;
        if (i_ >= n_ || !std::isdigit((unsigned char)s_[i_])) fail("bad number");
        if (s_[i_] == '0') {
            ++i_;
        } else {
            while (i_ < n_ && std::isdigit((unsigned char)s_[i_])) ++i_;
        }
        bool integral = true;
        if (i_ < n_ && s_[i_] == '.') {
            integral = false;
            ++i_;
            if (i_ >= n_ || !std::isdigit((unsigned char)s_[i_])) fail("bad number");
            while (i_ < n_ && std::isdigit((unsigned char)s_[i_])) ++i_;
        }
        if (i_ < n_ && (s_[i_] == 'e' || s_[i_] == 'E')) {
            integral = false;
            ++i_;
            if (i_ < n_ && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
            if (i_ >= n_ || !std::isdigit((unsigned char)s_[i_])) fail("bad number");
            while (i_ < n_ && std::isdigit((unsigned char)s_[i_])) ++i_;
        }
        Value v;
        v.kind = Value::Number;
        v.integral = integral;
        v.num = std::strtod(std::string(s_ + start, i_ - start).c_str(), nullptr);
        return v;
    }
};

}  // namespace json
}  // namespace newt
