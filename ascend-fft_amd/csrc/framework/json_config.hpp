// Strict JSON reader for the hardware profile and the design space.  Any
// malformation throws std::runtime_error, which Context::init turns into -4
// (the configuration contract of the reference's src/framework/json_config.hpp
// and butterfly.cpp:938-958).  Rejects: trailing bytes, duplicate keys,
// non-object roots, control characters in strings, leading zeros, empty
// exponents/fractions, nesting deeper than 64.
#pragma once
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace bfly {
namespace json {

struct Node {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    std::string scalar;                   // number text, string value, "true"/"false"
    std::vector<Node> items;              // Arr
    std::map<std::string, Node> fields;   // Obj

    bool has(const std::string& k) const { return kind == Obj && fields.count(k) != 0; }
    const Node& operator[](const std::string& k) const {
        if (kind != Obj) throw std::runtime_error("expected an object around field " + k);
        auto it = fields.find(k);
        if (it == fields.end()) throw std::runtime_error("missing field " + k);
        return it->second;
    }
    long long asInt() const {
        if (kind != Num) throw std::runtime_error("expected an integer");
        for (char c : scalar)
            if (c == '.' || c == 'e' || c == 'E') throw std::runtime_error("expected an integer, got " + scalar);
        char* end = nullptr;
        errno = 0;
        const long long v = std::strtoll(scalar.c_str(), &end, 10);
        if (errno || !end || *end) throw std::runtime_error("integer out of range: " + scalar);
        return v;
    }
    int asInt32() const {
        const long long v = asInt();
        if (v < -2147483648LL || v > 2147483647LL) throw std::runtime_error("int32 out of range");
        return (int)v;
    }
    double asDouble() const {
        if (kind != Num) throw std::runtime_error("expected a number");
        return std::strtod(scalar.c_str(), nullptr);
    }
    bool asBool() const {
        if (kind != Bool) throw std::runtime_error("expected a boolean");
        return scalar == "true";
    }
    const std::string& asString() const {
        if (kind != Str || scalar.empty()) throw std::runtime_error("expected a nonempty string");
        return scalar;
    }
    const std::vector<Node>& asArray() const {
        if (kind != Arr || items.empty()) throw std::runtime_error("expected a nonempty array");
        return items;
    }
    const std::map<std::string, Node>& asObject() const {
        if (kind != Obj || fields.empty()) throw std::runtime_error("expected a nonempty object");
        return fields;
    }
};

class Reader {
public:
    explicit Reader(const std::string& text) : s_(text) {}

    Node document() {
        Node root = value(0);
        ws();
        if (i_ != s_.size()) fail("trailing bytes after the document");
        if (root.kind != Node::Obj) fail("the document must be an object");
        return root;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char* why) const {
        throw std::runtime_error(std::string("JSON: ") + why + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
    }
    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            i_++;
            return true;
        }
        return false;
    }
    static void utf8(std::string& out, unsigned cp) {
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
    }
    std::string str() {
        ws();
        if (i_ >= s_.size() || s_[i_] != '"') fail("expected a string");
        i_++;
        std::string out;
        for (;;) {
            if (i_ >= s_.size()) fail("unterminated string");
            const unsigned char c = (unsigned char)s_[i_++];
            if (c == '"') return out;
            if (c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += (char)c;
                continue;
            }
            if (i_ >= s_.size()) fail("dangling escape");
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
                    if (i_ + 4 > s_.size()) fail("short \\u escape");
                    unsigned cp = 0;
                    for (int k = 0; k < 4; k++) {
                        const char h = s_[i_++];
                        if (!std::isxdigit((unsigned char)h)) fail("bad \\u escape");
                        cp = cp * 16 + (unsigned)(std::isdigit((unsigned char)h) ? h - '0'
                                                                                 : std::tolower(h) - 'a' + 10);
                    }
                    utf8(out, cp);
                } break;
                default: fail("unknown escape");
            }
        }
    }
    Node number() {
        const size_t start = i_;
        if (s_[i_] == '-') i_++;
        if (i_ >= s_.size() || !std::isdigit((unsigned char)s_[i_])) fail("bad number");
        if (s_[i_] == '0') {
            i_++;
        } else {
            while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) i_++;
        }
        if (i_ < s_.size() && s_[i_] == '.') {
            i_++;
            const size_t d = i_;
            while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) i_++;
            if (d == i_) fail("empty fraction");
        }
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            i_++;
            if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) i_++;
            const size_t d = i_;
            while (i_ < s_.size() && std::isdigit((unsigned char)s_[i_])) i_++;
            if (d == i_) fail("empty exponent");
        }
        Node n;
        n.kind = Node::Num;
        n.scalar = s_.substr(start, i_ - start);
        return n;
    }
    Node value(int depth) {
        if (depth > 64) fail("nesting too deep");
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        Node n;
        const char c = s_[i_];
        if (c == '{') {
            i_++;
            n.kind = Node::Obj;
            if (eat('}')) return n;
            do {
                std::string key = str();
                if (!eat(':')) fail("expected ':'");
                if (n.fields.count(key)) fail("duplicate key");
                n.fields.emplace(std::move(key), value(depth + 1));
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
            return n;
        }
        if (c == '[') {
            i_++;
            n.kind = Node::Arr;
            if (eat(']')) return n;
            do {
                n.items.push_back(value(depth + 1));
            } while (eat(','));
            if (!eat(']')) fail("expected ']'");
            return n;
        }
        if (c == '"') {
            n.kind = Node::Str;
            n.scalar = str();
            return n;
        }
        for (const char* lit : {"true", "false", "null"}) {
            const size_t L = std::char_traits<char>::length(lit);
            if (s_.compare(i_, L, lit) == 0) {
                i_ += L;
                n.kind = lit[0] == 'n' ? Node::Null : Node::Bool;
                n.scalar = lit;
                return n;
            }
        }
        return number();
    }
};

inline Node parse(const std::string& text) { return Reader(text).document(); }

}  // namespace json
}  // namespace bfly
