#include "json.hpp"

#include <cctype>
#include <charconv>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace eng {

bool Json::get_bool() const {
    if (type_ != Type::Bool) throw JsonError("not a boolean");
    return b_;
}
int64_t Json::get_int() const {
    if (type_ == Type::Int || type_ == Type::UInt) return i_;
    if (type_ == Type::Float) return static_cast<int64_t>(d_);
    throw JsonError("not a number");
}
uint64_t Json::get_uint() const {
    if (type_ == Type::UInt) return u_;
    if (type_ == Type::Int) return static_cast<uint64_t>(i_);
    if (type_ == Type::Float) return static_cast<uint64_t>(d_);
    throw JsonError("not a number");
}
double Json::get_double() const {
    if (type_ == Type::Float) return d_;
    if (type_ == Type::Int) return static_cast<double>(i_);
    if (type_ == Type::UInt) return static_cast<double>(u_);
    throw JsonError("not a number");
}
const std::string& Json::get_string() const {
    if (type_ != Type::String) throw JsonError("not a string");
    return s_;
}

bool Json::contains(const std::string& key) const {
    if (type_ != Type::Object) return false;
    for (const auto& kv : o_)
        if (kv.first == key) return true;
    return false;
}
Json& Json::operator[](const std::string& key) {
    if (type_ == Type::Null) type_ = Type::Object;
    if (type_ != Type::Object) throw JsonError("not an object");
    for (auto& kv : o_)
        if (kv.first == key) return kv.second;
    o_.emplace_back(key, Json());
    return o_.back().second;
}
const Json& Json::at(const std::string& key) const {
    if (type_ == Type::Object)
        for (const auto& kv : o_)
            if (kv.first == key) return kv.second;
    throw JsonError("missing key " + key);
}
void Json::erase(const std::string& key) {
    for (auto it = o_.begin(); it != o_.end(); ++it)
        if (it->first == key) {
            o_.erase(it);
            return;
        }
}
void Json::push_back(Json v) {
    if (type_ == Type::Null) type_ = Type::Array;
    if (type_ != Type::Array) throw JsonError("not an array");
    a_.push_back(std::move(v));
}

bool Json::operator==(const Json& o) const {
    if (is_number() && o.is_number()) {
        if (type_ == Type::Float || o.type_ == Type::Float) return get_double() == o.get_double();
        return i_ == o.i_ && (type_ == o.type_ || (i_ >= 0 && o.i_ >= 0));
    }
    if (type_ != o.type_) return false;
    switch (type_) {
        case Type::Null: return true;
        case Type::Bool: return b_ == o.b_;
        case Type::String: return s_ == o.s_;
        case Type::Array: return a_ == o.a_;
        case Type::Object: return o_ == o.o_;
        default: return false;
    }
}

// ---------------------------------------------------------------------------------------------
// number formatting
// ---------------------------------------------------------------------------------------------
std::string format_double(double v) {
    if (!std::isfinite(v)) return "null";
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    std::string s(buf, r.ptr);
    std::string out;
    size_t p = 0;
    if (s[0] == '-') {
        out += '-';
        p = 1;
    }
    const size_t e = s.find('e');
    std::string digits;
    for (size_t i = p; i < e; ++i)
        if (s[i] != '.') digits += s[i];
    const int exp10 = std::atoi(s.c_str() + e + 1);
    const int k = static_cast<int>(digits.size());
    const int n = exp10 + 1;  // value = 0.d1..dk * 10^n
    constexpr int kMinExp = -4, kMaxExp = 15;
    if (k <= n && n <= kMaxExp) {
        out += digits;
        out.append(static_cast<size_t>(n - k), '0');
        out += ".0";
    } else if (0 < n && n <= kMaxExp) {
        out += digits.substr(0, static_cast<size_t>(n));
        out += '.';
        out += digits.substr(static_cast<size_t>(n));
    } else if (kMinExp < n && n <= 0) {
        out += "0.";
        out.append(static_cast<size_t>(-n), '0');
        out += digits;
    } else {
        out += digits[0];
        if (k > 1) {
            out += '.';
            out += digits.substr(1);
        }
        out += 'e';
        const int x = n - 1;
        out += x < 0 ? '-' : '+';
        const int ax = x < 0 ? -x : x;
        if (ax < 10) out += '0';
        out += std::to_string(ax);
    }
    return out;
}

namespace {

void dump_string(std::string& out, const std::string& s) {
    out += '"';
    for (unsigned char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\b': out += "\\b"; break;
            case '\f': out += "\\f"; break;
            case '\n': out += "\\n"; break;
            case '\r': out += "\\r"; break;
            case '\t': out += "\\t"; break;
            default:
                if (c < 0x20) {
                    char b[8];
                    std::snprintf(b, sizeof(b), "\\u%04x", c);
                    out += b;
                } else {
                    out += static_cast<char>(c);
                }
        }
    }
    out += '"';
}

}  // namespace

void Json::dump_to(std::string& out, int indent, int level) const {
    const bool pretty = indent >= 0;
    const std::string pad = pretty ? std::string(static_cast<size_t>(indent * (level + 1)), ' ') : "";
    const std::string pad0 = pretty ? std::string(static_cast<size_t>(indent * level), ' ') : "";
    switch (type_) {
        case Type::Null: out += "null"; break;
        case Type::Bool: out += b_ ? "true" : "false"; break;
        case Type::Int: out += std::to_string(i_); break;
        case Type::UInt: out += std::to_string(u_); break;
        case Type::Float: out += format_double(d_); break;
        case Type::String: dump_string(out, s_); break;
        case Type::Array:
            if (a_.empty()) {
                out += "[]";
                break;
            }
            out += pretty ? "[\n" : "[";
            for (size_t i = 0; i < a_.size(); ++i) {
                out += pad;
                a_[i].dump_to(out, indent, level + 1);
                if (i + 1 < a_.size()) out += pretty ? ",\n" : ",";
            }
            out += pretty ? "\n" + pad0 + "]" : "]";
            break;
        case Type::Object:
            if (o_.empty()) {
                out += "{}";
                break;
            }
            out += pretty ? "{\n" : "{";
            for (size_t i = 0; i < o_.size(); ++i) {
                out += pad;
                dump_string(out, o_[i].first);
                out += pretty ? ": " : ":";
                o_[i].second.dump_to(out, indent, level + 1);
                if (i + 1 < o_.size()) out += pretty ? ",\n" : ",";
            }
            out += pretty ? "\n" + pad0 + "}" : "}";
            break;
    }
}

std::string Json::dump(int indent) const {
    std::string out;
    dump_to(out, indent, 0);
    return out;
}

// ---------------------------------------------------------------------------------------------
// parser
// ---------------------------------------------------------------------------------------------
namespace {

struct Parser {
    const std::string& s;
    size_t p = 0;
    bool comments;
    int line = 1;

    [[noreturn]] void fail(const std::string& what) const {
        throw JsonError("syntax error at line " + std::to_string(line) + ", byte " + std::to_string(p + 1) + ": " + what);
    }
    void skip() {
        for (;;) {
            while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) {
                if (s[p] == '\n') ++line;
                ++p;
            }
            if (comments && p + 1 < s.size() && s[p] == '/' && s[p + 1] == '/') {
                while (p < s.size() && s[p] != '\n') ++p;
                continue;
            }
            if (comments && p + 1 < s.size() && s[p] == '/' && s[p + 1] == '*') {
                const size_t e = s.find("*/", p + 2);
                if (e == std::string::npos) fail("unterminated comment");
                for (size_t i = p; i < e; ++i)
                    if (s[i] == '\n') ++line;
                p = e + 2;
                continue;
            }
            return;
        }
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (p + 4 > s.size()) fail("bad \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = s[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<uint32_t>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<uint32_t>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<uint32_t>(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string str() {
        if (s[p] != '"') fail("expected string");
        ++p;
        std::string out;
        while (p < s.size() && s[p] != '"') {
            char c = s[p++];
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size()) fail("bad escape");
            c = s[p++];
            switch (c) {
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
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (p + 2 > s.size() || s[p] != '\\' || s[p + 1] != 'u') fail("bad surrogate pair");
                        p += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        if (p >= s.size()) fail("unterminated string");
        ++p;
        return out;
    }
    Json number() {
        const size_t b = p;
        bool is_float = false;
        if (s[p] == '-') ++p;
        if (p >= s.size() || !std::isdigit(static_cast<unsigned char>(s[p]))) fail("bad number");
        if (s[p] == '0') {
            ++p;
        } else {
            while (p < s.size() && std::isdigit(static_cast<unsigned char>(s[p]))) ++p;
        }
        if (p < s.size() && s[p] == '.') {
            is_float = true;
            ++p;
            if (p >= s.size() || !std::isdigit(static_cast<unsigned char>(s[p]))) fail("bad number");
            while (p < s.size() && std::isdigit(static_cast<unsigned char>(s[p]))) ++p;
        }
        if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
            is_float = true;
            ++p;
            if (p < s.size() && (s[p] == '+' || s[p] == '-')) ++p;
            if (p >= s.size() || !std::isdigit(static_cast<unsigned char>(s[p]))) fail("bad number");
            while (p < s.size() && std::isdigit(static_cast<unsigned char>(s[p]))) ++p;
        }
        const std::string t = s.substr(b, p - b);
        if (!is_float) {
            if (t[0] == '-') {
                long long v = 0;
                const auto r = std::from_chars(t.data(), t.data() + t.size(), v);
                if (r.ec == std::errc()) return Json(v);
            } else {
                unsigned long long v = 0;
                const auto r = std::from_chars(t.data(), t.data() + t.size(), v);
                if (r.ec == std::errc()) return Json(v);
            }
        }
        return Json(std::strtod(t.c_str(), nullptr));
    }
    Json value(int depth) {
        if (depth > 512) fail("nesting too deep");
        skip();
        if (p >= s.size()) fail("unexpected end of input");
        const char c = s[p];
        if (c == '{') {
            ++p;
            Json o = Json::object();
            skip();
            if (p < s.size() && s[p] == '}') {
                ++p;
                return o;
            }
            for (;;) {
                skip();
                if (p >= s.size() || s[p] != '"') fail("expected object key");
                const std::string k = str();
                skip();
                if (p >= s.size() || s[p] != ':') fail("expected ':'");
                ++p;
                o[k] = value(depth + 1);
                skip();
                if (p < s.size() && s[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < s.size() && s[p] == '}') {
                    ++p;
                    return o;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            Json a = Json::array();
            skip();
            if (p < s.size() && s[p] == ']') {
                ++p;
                return a;
            }
            for (;;) {
                a.push_back(value(depth + 1));
                skip();
                if (p < s.size() && s[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < s.size() && s[p] == ']') {
                    ++p;
                    return a;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') return Json(str());
        if (s.compare(p, 4, "true") == 0) {
            p += 4;
            return Json(true);
        }
        if (s.compare(p, 5, "false") == 0) {
            p += 5;
            return Json(false);
        }
        if (s.compare(p, 4, "null") == 0) {
            p += 4;
            return Json();
        }
        if (c == '-' || std::isdigit(static_cast<unsigned char>(c))) return number();
        fail("unexpected character");
    }
};

}  // namespace

Json Json::parse(const std::string& text, bool allow_comments) {
    Parser ps{text, 0, allow_comments};
    Json v = ps.value(0);
    ps.skip();
    if (ps.p != text.size()) ps.fail("trailing characters");
    return v;
}

}  // namespace eng
