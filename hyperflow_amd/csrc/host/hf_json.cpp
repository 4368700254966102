// hf_json.cpp -- minimal JSON reader for HyperFlow case files.
#include "hf_json.hpp"

#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace hf_json {

class Parser {
public:
    explicit Parser(const std::string& t) : s(t) {}

    Value document()
    {
        Value v = value();
        ws();
        if (p != s.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s;
    size_t p = 0;

    [[noreturn]] void fail(const std::string& what) const
    {
        throw Error("JSON parse error at offset " + std::to_string(p) + ": " + what);
    }
    void ws()
    {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) p++;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if (s.compare(p, n, w) == 0) {
            p += n;
            return true;
        }
        return false;
    }

    Value value()
    {
        ws();
        if (p >= s.size()) fail("unexpected end");
        const char c = s[p];
        Value v;
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') {
            v.kind_ = Value::String;
            v.str_ = string();
            return v;
        }
        if (lit("true")) { v.kind_ = Value::Bool; v.bool_ = true; return v; }
        if (lit("false")) { v.kind_ = Value::Bool; v.bool_ = false; return v; }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail(std::string("unexpected character '") + c + "'");
    }

    Value number()
    {
        const size_t start = p;
        if (s[p] == '-') p++;
        while (p < s.size() && (std::isdigit((unsigned char)s[p]) || s[p] == '.' || s[p] == 'e' || s[p] == 'E' ||
                                s[p] == '+' || s[p] == '-'))
            p++;
        const std::string tok = s.substr(start, p - start);
        char* end = nullptr;
        errno = 0;
        const double d = std::strtod(tok.c_str(), &end);
        if (end != tok.c_str() + tok.size()) fail("bad number '" + tok + "'");
        Value v;
        v.kind_ = Value::Number;
        v.num_ = d;
        return v;
    }

    std::string string()
    {
        std::string out;
        p++;   // opening quote
        while (p < s.size() && s[p] != '"') {
            char c = s[p++];
            if (c == '\\') {
                if (p >= s.size()) fail("bad escape");
                const char e = s[p++];
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
                    if (p + 4 > s.size()) fail("bad \\u escape");
                    const unsigned cp = (unsigned)std::stoul(s.substr(p, 4), nullptr, 16);
                    p += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else {
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
        if (p >= s.size()) fail("unterminated string");
        p++;   // closing quote
        return out;
    }

    Value array()
    {
        Value v;
        v.kind_ = Value::Array;
        p++;
        ws();
        if (p < s.size() && s[p] == ']') { p++; return v; }
        for (;;) {
            v.items_.push_back(value());
            ws();
            if (p < s.size() && s[p] == ',') { p++; continue; }
            if (p < s.size() && s[p] == ']') { p++; return v; }
            fail("expected ',' or ']'");
        }
    }

    Value object()
    {
        Value v;
        v.kind_ = Value::Object;
        p++;
        ws();
        if (p < s.size() && s[p] == '}') { p++; return v; }
        for (;;) {
            ws();
            if (p >= s.size() || s[p] != '"') fail("expected key");
            std::string k = string();
            ws();
            if (p >= s.size() || s[p] != ':') fail("expected ':'");
            p++;
            Value item = value();
            auto it = v.index_.find(k);
            if (it != v.index_.end()) {
                v.items_[it->second] = item;          // last duplicate wins
            } else {
                v.index_[k] = v.items_.size();
                v.keys_.push_back(k);
                v.items_.push_back(item);
            }
            ws();
            if (p < s.size() && s[p] == ',') { p++; continue; }
            if (p < s.size() && s[p] == '}') { p++; return v; }
            fail("expected ',' or '}'");
        }
    }
};

Value Value::parse(const std::string& text) { return Parser(text).document(); }

Value Value::load(const std::string& path)
{
    std::ifstream f(path);
    if (!f) throw Error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

bool Value::has(const std::string& key) const { return kind_ == Object && index_.count(key) != 0; }

const Value& Value::operator[](const std::string& key) const
{
    if (kind_ != Object) throw Error("not an object (looking up '" + key + "')");
    auto it = index_.find(key);
    if (it == index_.end()) throw Error("missing key '" + key + "'");
    return items_[it->second];
}

const Value& Value::operator[](size_t i) const
{
    if (kind_ != Array || i >= items_.size()) throw Error("array index out of range");
    return items_[i];
}

size_t Value::size() const { return (kind_ == Array || kind_ == Object) ? items_.size() : 0; }

double Value::num() const
{
    if (kind_ != Number) throw Error("not a number");
    return num_;
}

int Value::integer() const { return (int)num(); }

const std::string& Value::str() const
{
    if (kind_ != String) throw Error("not a string");
    return str_;
}

bool Value::boolean() const
{
    if (kind_ != Bool) throw Error("not a boolean");
    return bool_;
}

std::vector<double> Value::numbers() const
{
    if (kind_ != Array) throw Error("not an array");
    std::vector<double> out;
    for (const Value& v : items_) out.push_back(v.num());
    return out;
}

}  // namespace hf_json
