// json.h -- a small JSON reader for the command-line host (objects, arrays, numbers, strings,
// true/false/null). Numbers are kept as double; nested arrays are flattened on request.
#pragma once

#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sstmc_json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool has(const std::string& k) const { return kind == Object && obj.count(k); }
    const Value& at(const std::string& k) const {
        auto it = obj.find(k);
        if (kind != Object || it == obj.end()) throw std::runtime_error("missing field '" + k + "'");
        return it->second;
    }
    // Numbers of a scalar or of a (possibly nested) array, in order; nested arrays a[i][j][k]
    // are returned with the FIRST index fastest (MATLAB column order), with the shape in `shape`.
    std::vector<double> numbers(std::vector<size_t>* shape = nullptr) const;
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text) {}
    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char* what) {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
    }
    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) { ++i_; return true; }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail("unexpected character");
    }
    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        char c = s_[i_];
        Value v;
        if (c == '{') {
            ++i_;
            v.kind = Value::Object;
            if (eat('}')) return v;
            do {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected a key");
                std::string k = string();
                expect(':');
                v.obj[k] = value();
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            ++i_;
            v.kind = Value::Array;
            if (eat(']')) return v;
            do {
                v.arr.push_back(value());
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = Value::String;
            v.str = string();
        } else if (s_.compare(i_, 4, "true") == 0) {
            i_ += 4; v.kind = Value::Bool; v.b = true;
        } else if (s_.compare(i_, 5, "false") == 0) {
            i_ += 5; v.kind = Value::Bool; v.b = false;
        } else if (s_.compare(i_, 4, "null") == 0) {
            i_ += 4;
        } else {
            const char* begin = s_.c_str() + i_;
            char* end = nullptr;
            v.num = std::strtod(begin, &end);
            if (end == begin) fail("bad value");
            i_ += (size_t)(end - begin);
            v.kind = Value::Number;
        }
        return v;
    }
    std::string string() {
        ++i_;   // opening quote
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (i_ + 4 > s_.size()) fail("bad escape");
                    unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                    i_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) fail("unterminated string");
        ++i_;   // closing quote
        return out;
    }
};

inline void collect(const Value& v, size_t depth, std::vector<size_t>& shape, std::vector<const Value*>& leaves) {
    if (v.kind != Value::Array) {
        if (v.kind != Value::Number && v.kind != Value::Bool) throw std::runtime_error("expected numbers");
        if (depth != shape.size()) throw std::runtime_error("ragged numeric array");
        leaves.push_back(&v);
        return;
    }
    if (depth == shape.size()) shape.push_back(v.arr.size());
    else if (shape[depth] != v.arr.size()) throw std::runtime_error("ragged numeric array");
    for (const Value& e : v.arr) collect(e, depth + 1, shape, leaves);
}

inline std::vector<double> Value::numbers(std::vector<size_t>* shapeOut) const {
    std::vector<size_t> shape;
    std::vector<const Value*> leaves;   // row-major (last index fastest)
    collect(*this, 0, shape, leaves);
    std::vector<double> out(leaves.size());
    // row-major position -> column-major position (first index fastest)
    for (size_t r = 0; r < leaves.size(); ++r) {
        size_t rem = r, col = 0, stride = 1;
        std::vector<size_t> idx(shape.size());
        for (size_t d = shape.size(); d-- > 0;) {
            idx[d] = rem % shape[d];
            rem /= shape[d];
        }
        for (size_t d = 0; d < shape.size(); ++d) {
            col += idx[d] * stride;
            stride *= shape[d];
        }
        const Value* l = leaves[r];
        out[col] = l->kind == Bool ? (l->b ? 1.0 : 0.0) : l->num;
    }
    if (shapeOut) *shapeOut = shape;
    return out;
}

}  // namespace sstmc_json
