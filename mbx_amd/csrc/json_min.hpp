// A small JSON reader for the System configuration file: objects, arrays, numbers, strings,
// true/false/null.  Enough for the keys bblock::System reads; errors throw std::runtime_error.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mbxjson {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool has(const std::string& k) const { return kind == Object && obj.count(k); }
    const Value& at(const std::string& k) const {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end()) throw std::runtime_error("json: missing key '" + k + "'");
        return it->second;
    }
    double as_number() const {
        if (kind != Number) throw std::runtime_error("json: expected a number");
        return num;
    }
    const std::string& as_string() const {
        if (kind != String) throw std::runtime_error("json: expected a string");
        return str;
    }
};

class Parser {
   public:
    explicit Parser(const std::string& text) : s_(text) {}

    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) error("trailing characters");
        return v;
    }

   private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void error(const std::string& what) const {
        throw std::runtime_error("json: " + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    char peek() {
        ws();
        if (i_ >= s_.size()) error("unexpected end");
        return s_[i_];
    }
    void expect(char c) {
        if (peek() != c) error(std::string("expected '") + c + "'");
        ++i_;
    }
    bool literal(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    std::string string() {
        expect('"');
        std::string out;
        while (true) {
            if (i_ >= s_.size()) error("unterminated string");
            char c = s_[i_++];
            if (c == '"') break;
            if (c == '\\') {
                if (i_ >= s_.size()) error("bad escape");
                char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u':
                        if (i_ + 4 > s_.size()) error("bad \\u escape");
                        out += '?';
                        i_ += 4;
                        break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    Value value() {
        Value v;
        char c = peek();
        if (c == '{') {
            ++i_;
            v.kind = Value::Object;
            if (peek() == '}') {
                ++i_;
                return v;
            }
            while (true) {
                std::string k = string();
                expect(':');
                v.obj[k] = value();
                char d = peek();
                ++i_;
                if (d == '}') break;
                if (d != ',') error("expected ',' or '}'");
            }
        } else if (c == '[') {
            ++i_;
            v.kind = Value::Array;
            if (peek() == ']') {
                ++i_;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                char d = peek();
                ++i_;
                if (d == ']') break;
                if (d != ',') error("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (literal("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (literal("false")) {
            v.kind = Value::Bool;
        } else if (literal("null")) {
            v.kind = Value::Null;
        } else {
            const char* p = s_.c_str() + i_;
            char* end = nullptr;
            v.num = std::strtod(p, &end);
            if (end == p) error("bad value");
            v.kind = Value::Number;
            i_ += (size_t)(end - p);
        }
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace mbxjson
