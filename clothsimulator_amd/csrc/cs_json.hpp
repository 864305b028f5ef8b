// cs_json.hpp — a small JSON reader for scene and material files.
//
// Accepts standard JSON plus // line and /* block */ comments, which the
// reference's configuration files use (the reference reads them with
// jsoncpp's default, comment-tolerant reader).  Missing keys and out-of-range
// indices read as null, so optional fields are looked up the way the
// reference's parse helpers do (src/JsonHelper.cu:3-31: null -> default).
#pragma once
#include <cstdlib>
#include <stdexcept>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

namespace cs {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    static const Value& null_value() {
        static const Value v;
        return v;
    }
    bool is_null() const { return kind == Null; }
    bool is_array() const { return kind == Array; }
    bool is_object() const { return kind == Object; }
    bool is_string() const { return kind == String; }
    bool is_number() const { return kind == Number; }
    size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }
    const Value& operator[](const std::string& key) const {
        if (kind == Object)
            for (const auto& kv : obj)
                if (kv.first == key) return kv.second;
        return null_value();
    }
    const Value& operator[](size_t i) const { return kind == Array && i < arr.size() ? arr[i] : null_value(); }
    double as_number(const char* what) const {
        if (kind == Number) return num;
        if (kind == Bool) return b ? 1.0 : 0.0;
        throw std::runtime_error(std::string("expected a number for ") + what);
    }
    std::string as_string(const char* what) const {
        if (kind != String) throw std::runtime_error(std::string("expected a string for ") + what);
        return str;
    }
};

class Parser {
  public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        Value v = value();
        skip();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char* msg) const {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); k++) {
            if (s_[k] == '\n') { line++; col = 1; } else col++;
        }
        throw std::runtime_error(std::string("JSON: ") + msg + " at line " + std::to_string(line) + ", column " +
                                 std::to_string(col));
    }
    void skip() {
        for (;;) {
            while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') i_++;
            } else if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) fail("unterminated comment");
                i_ = e + 2;
            } else {
                return;
            }
        }
    }
    bool eat(char c) {
        skip();
        if (i_ < s_.size() && s_[i_] == c) { i_++; return true; }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail((std::string("expected '") + c + "'").c_str());
    }
    bool word(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
        return false;
    }
    std::string string_lit() {
        expect('"');
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
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
                        if (i_ + 4 > s_.size()) fail("bad \\u escape");
                        unsigned code = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (code < 0x80) out += (char)code;
                        else if (code < 0x800) { out += (char)(0xC0 | (code >> 6)); out += (char)(0x80 | (code & 63)); }
                        else { out += (char)(0xE0 | (code >> 12)); out += (char)(0x80 | ((code >> 6) & 63)); out += (char)(0x80 | (code & 63)); }
                        break;
                    }
                    default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        i_++;
        return out;
    }
    Value value() {
        skip();
        if (i_ >= s_.size()) fail("unexpected end of input");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            i_++;
            v.kind = Value::Object;
            if (eat('}')) return v;
            do {
                skip();
                std::string key = string_lit();
                expect(':');
                v.obj.emplace_back(std::move(key), value());
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            i_++;
            v.kind = Value::Array;
            if (eat(']')) return v;
            do v.arr.push_back(value());
            while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string_lit();
        } else if (word("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (word("false")) {
            v.kind = Value::Bool;
        } else if (word("null")) {
        } else if (c == '-' || c == '+' || c == '.' || (c >= '0' && c <= '9')) {
            const char* start = s_.c_str() + i_;
            char* end = nullptr;
            v.kind = Value::Number;
            v.num = std::strtod(start, &end);
            if (end == start) fail("bad number");
            i_ += (size_t)(end - start);
        } else {
            fail("unexpected character");
        }
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

// Serialises a value (keys in their parsed order, numbers with 17 significant digits so they
// read back exactly, two-space indentation).
inline void dump_to(const Value& v, std::string& out, int indent) {
    const std::string pad(2 * (indent + 1), ' '), end(2 * indent, ' ');
    switch (v.kind) {
        case Value::Null: out += "null"; break;
        case Value::Bool: out += v.b ? "true" : "false"; break;
        case Value::Number: {
            char buf[32];
            std::snprintf(buf, sizeof(buf), "%.17g", v.num);
            out += buf;
            break;
        }
        case Value::String: {
            out += '"';
            for (char c : v.str) {
                if (c == '"' || c == '\\') { out += '\\'; out += c; }
                else if ((unsigned char)c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof(buf), "\\u%04x", (unsigned)(unsigned char)c);
                    out += buf;
                } else out += c;
            }
            out += '"';
            break;
        }
        case Value::Array:
            out += "[";
            for (size_t i = 0; i < v.arr.size(); i++) {
                out += i ? ",\n" : "\n";
                out += pad;
                dump_to(v.arr[i], out, indent + 1);
            }
            out += v.arr.empty() ? "]" : "\n" + end + "]";
            break;
        case Value::Object:
            out += "{";
            for (size_t i = 0; i < v.obj.size(); i++) {
                out += i ? ",\n" : "\n";
                out += pad;
                Value key;
                key.kind = Value::String;
                key.str = v.obj[i].first;
                dump_to(key, out, indent + 1);
                out += ": ";
                dump_to(v.obj[i].second, out, indent + 1);
            }
            out += v.obj.empty() ? "}" : "\n" + end + "}";
            break;
    }
}

inline std::string dump(const Value& v) {
    std::string out;
    dump_to(v, out, 0);
    out += "\n";
    return out;
}

}  // namespace json
}  // namespace cs
