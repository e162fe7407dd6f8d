// mrh_json.hpp — a small JSON reader for scene files: the reference's scene
// files are JSON with // and /* */ comments (RayLib/SceneIO reads them with
// comments allowed), so comments are skipped and a trailing comma before a
// closing bracket is accepted.  Numbers are doubles; objects keep their keys
// in a map.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mrh {
namespace json {

struct Value;
using Object = std::map<std::string, Value>;
using Array = std::vector<Value>;

struct Value {
    enum Kind { NUL, BOOL, NUMBER, STRING, ARRAY, OBJECT } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::shared_ptr<Array> arr;
    std::shared_ptr<Object> obj;

    bool is_null() const { return kind == NUL; }
    bool is_number() const { return kind == NUMBER; }
    bool is_string() const { return kind == STRING; }
    bool is_array() const { return kind == ARRAY; }
    bool is_object() const { return kind == OBJECT; }
    const Array& array() const
    {
        if (kind != ARRAY) throw std::runtime_error("expected an array");
        return *arr;
    }
    const Object& object() const
    {
        if (kind != OBJECT) throw std::runtime_error("expected an object");
        return *obj;
    }
    double number() const
    {
        if (kind != NUMBER) throw std::runtime_error("expected a number");
        return num;
    }
    const std::string& string() const
    {
        if (kind != STRING) throw std::runtime_error("expected a string");
        return str;
    }
    bool boolean() const
    {
        if (kind != BOOL) throw std::runtime_error("expected true or false");
        return b;
    }
    // member lookup (null when absent or not an object)
    const Value* get(const std::string& key) const
    {
        if (kind != OBJECT) return nullptr;
        auto it = obj->find(key);
        return it == obj->end() ? nullptr : &it->second;
    }
};

class Parser {
public:
    explicit Parser(const std::string& text) : s(text) {}

    Value parse()
    {
        Value v = value();
        skip();
        if (p != s.size()) error("trailing characters");
        return v;
    }

private:
    const std::string& s;
    size_t p = 0;

    [[noreturn]] void error(const char* what) const
    {
        size_t line = 1;
        for (size_t i = 0; i < p && i < s.size(); i++) line += s[i] == '\n';
        throw std::runtime_error(std::string("JSON ") + what + " at line " + std::to_string(line));
    }

    void skip()
    {
        for (;;) {
            while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) p++;
            if (p + 1 < s.size() && s[p] == '/' && s[p + 1] == '/') {
                while (p < s.size() && s[p] != '\n') p++;
            } else if (p + 1 < s.size() && s[p] == '/' && s[p + 1] == '*') {
                const size_t e = s.find("*/", p + 2);
                if (e == std::string::npos) error("unterminated comment");
                p = e + 2;
            } else {
                return;
            }
        }
    }

    Value value()
    {
        skip();
        if (p >= s.size()) error("unexpected end");
        const char c = s[p];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            Value v;
            v.kind = Value::STRING;
            v.str = string();
            return v;
        }
        if (s.compare(p, 4, "true") == 0) {
            p += 4;
            Value v;
            v.kind = Value::BOOL;
            v.b = true;
            return v;
        }
        if (s.compare(p, 5, "false") == 0) {
            p += 5;
            Value v;
            v.kind = Value::BOOL;
            return v;
        }
        if (s.compare(p, 4, "null") == 0) {
            p += 4;
            return Value();
        }
        return number();
    }

    Value number()
    {
        const size_t start = p;
        if (p < s.size() && (s[p] == '-' || s[p] == '+')) p++;
        while (p < s.size() && ((s[p] >= '0' && s[p] <= '9') || s[p] == '.' || s[p] == 'e' || s[p] == 'E' ||
                                ((s[p] == '-' || s[p] == '+') && (s[p - 1] == 'e' || s[p - 1] == 'E'))))
            p++;
        if (p == start) error("unexpected character");
        Value v;
        v.kind = Value::NUMBER;
        try {
            size_t used = 0;
            v.num = std::stod(s.substr(start, p - start), &used);
            if (used != p - start) error("bad number");
        } catch (const std::logic_error&) {
            error("bad number");
        }
        return v;
    }

    std::string string()
    {
        p++;   // opening quote
        std::string out;
        while (p < s.size() && s[p] != '"') {
            char c = s[p++];
            if (c == '\\') {
                if (p >= s.size()) error("bad escape");
                const char e = s[p++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p + 4 > s.size()) error("bad escape");
                    const unsigned cp = (unsigned)std::stoul(s.substr(p, 4), nullptr, 16);
                    p += 4;
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
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p >= s.size()) error("unterminated string");
        p++;   // closing quote
        return out;
    }

    Value array()
    {
        p++;
        Value v;
        v.kind = Value::ARRAY;
        v.arr = std::make_shared<Array>();
        for (;;) {
            skip();
            if (p < s.size() && s[p] == ']') {
                p++;
                return v;
            }
            v.arr->push_back(value());
            skip();
            if (p < s.size() && s[p] == ',') {
                p++;
                continue;
            }
            if (p < s.size() && s[p] == ']') {
                p++;
                return v;
            }
            error("expected , or ]");
        }
    }

    Value object()
    {
        p++;
        Value v;
        v.kind = Value::OBJECT;
        v.obj = std::make_shared<Object>();
        for (;;) {
            skip();
            if (p < s.size() && s[p] == '}') {
                p++;
                return v;
            }
            if (p >= s.size() || s[p] != '"') error("expected a key");
            std::string k = string();
            skip();
            if (p >= s.size() || s[p] != ':') error("expected :");
            p++;
            (*v.obj)[k] = value();
            skip();
            if (p < s.size() && s[p] == ',') {
                p++;
                continue;
            }
            if (p < s.size() && s[p] == '}') {
                p++;
                return v;
            }
            error("expected , or }");
        }
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace mrh
