#include "json.hpp"

#include <cctype>
#include <cerrno>
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

// Length of the well-formed UTF-8 sequence (RFC 3629) that starts s[i], 0 when it is ill-formed or cut short.
size_t utf8_len(const std::string& s, size_t i) {
    const unsigned char c = static_cast<unsigned char>(s[i]);
    if (c < 0x80) return 1;
    int n = 0;
    unsigned lo = 0x80, hi = 0xBF;
    if (c >= 0xC2 && c <= 0xDF) n = 1;
    else if (c == 0xE0) n = 2, lo = 0xA0;
    else if ((c >= 0xE1 && c <= 0xEC) || c == 0xEE || c == 0xEF) n = 2;
    else if (c == 0xED) n = 2, hi = 0x9F;
    else if (c == 0xF0) n = 3, lo = 0x90;
    else if (c >= 0xF1 && c <= 0xF3) n = 3;
    else if (c == 0xF4) n = 3, hi = 0x8F;
    else return 0;
    for (int k = 1; k <= n; ++k) {
        if (i + static_cast<size_t>(k) >= s.size()) return 0;
        const unsigned char t = static_cast<unsigned char>(s[i + static_cast<size_t>(k)]);
        if (t < (k == 1 ? lo : 0x80u) || t > (k == 1 ? hi : 0xBFu)) return 0;
    }
    return static_cast<size_t>(n) + 1;
}

// Strings are written as UTF-8; an ill-formed sequence becomes U+FFFD, the way the reference's reports do it
// (nlohmann's error_handler_t::replace): a sequence cut short by a wrong byte is one U+FFFD and the wrong byte is
// read again as the start of the next character, a byte that cannot start a character is one U+FFFD.
void dump_string(std::string& out, const std::string& s) {
    std::string fixed;
    const std::string* src = &s;
    for (size_t i = 0; i < s.size(); ++i) {
        if (static_cast<unsigned char>(s[i]) < 0x80) continue;
        const size_t n = utf8_len(s, i);
        if (n > 0) {
            i += n - 1;
            continue;
        }
        // rebuild from here with replacements
        fixed = s.substr(0, i);
        size_t j = i;
        while (j < s.size()) {
            const size_t m = utf8_len(s, j);
            if (m > 0) {
                fixed.append(s, j, m);
                j += m;
                continue;
            }
            fixed += "\xEF\xBF\xBD";
            // skip the bytes that were a valid prefix of the broken sequence (at least the lead byte)
            size_t k = j + 1;
            const unsigned char c = static_cast<unsigned char>(s[j]);
            const bool lead = (c >= 0xC2 && c <= 0xF4);
            if (lead) {
                unsigned lo = 0x80, hi = 0xBF;
                if (c == 0xE0) lo = 0xA0;
                else if (c == 0xED) hi = 0x9F;
                else if (c == 0xF0) lo = 0x90;
                else if (c == 0xF4) hi = 0x8F;
                const size_t need = c <= 0xDF ? 1 : c <= 0xEF ? 2 : 3;
                for (size_t q = 1; q <= need && j + q < s.size(); ++q) {
                    const unsigned char t = static_cast<unsigned char>(s[j + q]);
                    if (t < (q == 1 ? lo : 0x80u) || t > (q == 1 ? hi : 0xBFu)) break;
                    k = j + q + 1;
                }
            }
            j = k;
        }
        src = &fixed;
        break;
    }
    out += '"';
    for (unsigned char c : *src) {
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
// The reference reads config files with nlohmann::json 3.11.3 (engine/third_party/json.hpp, MIT licence), and a
// syntax error reaches the user as that library's message text (--validate, batch_summary.csv).  This reader accepts
// the same language (RFC 8259, plus // and /* */ comments for config files) and reports errors with the same text:
// "[json.exception.parse_error.101] parse error at line L, column C: syntax error while parsing <context> -
// <what>; expected <token>", where the position counts the characters read so far (a read past the end counts as
// one), <what> names the unexpected token or, for a malformed one, the scanner's finding and the characters read
// since the last string or number began ("last read: '...'").  A number too large for a double is the library's
// out_of_range error 406 instead.
namespace {

enum class Tok {
    Uninitialized, True, False, Null, String, Unsigned, Integer, Float, BeginArray, BeginObject, EndArray, EndObject,
    NameSep, ValueSep, ParseError, End, LiteralOrValue
};

const char* tok_name(Tok t) {
    switch (t) {
        case Tok::Uninitialized: return "<uninitialized>";
        case Tok::True: return "true literal";
        case Tok::False: return "false literal";
        case Tok::Null: return "null literal";
        case Tok::String: return "string literal";
        case Tok::Unsigned:
        case Tok::Integer:
        case Tok::Float: return "number literal";
        case Tok::BeginArray: return "'['";
        case Tok::BeginObject: return "'{'";
        case Tok::EndArray: return "']'";
        case Tok::EndObject: return "'}'";
        case Tok::NameSep: return "':'";
        case Tok::ValueSep: return "','";
        case Tok::ParseError: return "<parse error>";
        case Tok::End: return "end of input";
        case Tok::LiteralOrValue: return "'[', '{', or a literal";
    }
    return "unknown token";
}

constexpr int kEof = -1;

class Scanner {
public:
    Scanner(const std::string& text, bool comments) : s_(text), comments_(comments) {}

    Tok scan() {
        if (read_total_ == 0 && !skip_bom()) {
            err_ = "invalid BOM; must be 0xEF 0xBB 0xBF if given";
            return Tok::ParseError;
        }
        skip_space();
        while (comments_ && cur_ == '/') {
            if (!comment()) return Tok::ParseError;
            skip_space();
        }
        switch (cur_) {
            case '[': return Tok::BeginArray;
            case ']': return Tok::EndArray;
            case '{': return Tok::BeginObject;
            case '}': return Tok::EndObject;
            case ':': return Tok::NameSep;
            case ',': return Tok::ValueSep;
            case 't': return literal("true", Tok::True);
            case 'f': return literal("false", Tok::False);
            case 'n': return literal("null", Tok::Null);
            case '"': return string();
            case '-': case '0': case '1': case '2': case '3': case '4': case '5': case '6': case '7': case '8': case '9':
                return number();
            case 0:
            case kEof: return Tok::End;
            default: err_ = "invalid literal"; return Tok::ParseError;
        }
    }

    // " at line L, column C" of the reference's messages
    std::string where() const {
        return " at line " + std::to_string(lines_ + 1) + ", column " + std::to_string(read_line_);
    }
    uint64_t read_total() const { return read_total_; }
    const char* error() const { return err_; }
    // characters read for the current token (since the last string / number began), control characters as <U+XXXX>
    std::string token_text() const {
        std::string out;
        for (const char c : tok_) {
            if (static_cast<unsigned char>(c) <= 0x1F) {
                char b[12];
                std::snprintf(b, sizeof(b), "<U+%.4X>", static_cast<unsigned>(static_cast<unsigned char>(c)));
                out += b;
            } else {
                out += c;
            }
        }
        return out;
    }
    std::string& text() { return buf_; }
    uint64_t u = 0;
    int64_t i = 0;
    double d = 0.0;

private:
    int get() {
        ++read_total_;
        ++read_line_;
        if (unget_) {
            unget_ = false;
        } else {
            cur_ = pos_ < s_.size() ? static_cast<unsigned char>(s_[pos_++]) : kEof;
        }
        if (cur_ != kEof) tok_.push_back(static_cast<char>(cur_));
        if (cur_ == '\n') {
            ++lines_;
            read_line_ = 0;
        }
        return cur_;
    }
    void unget() {
        unget_ = true;
        --read_total_;
        if (read_line_ == 0) {
            if (lines_ > 0) --lines_;
        } else {
            --read_line_;
        }
        if (cur_ != kEof) tok_.pop_back();
    }
    void begin_token() {
        buf_.clear();
        tok_.assign(1, static_cast<char>(cur_));
    }
    bool skip_bom() {
        if (get() == 0xEF) return get() == 0xBB && get() == 0xBF;
        unget();
        return true;
    }
    void skip_space() {
        do get();
        while (cur_ == ' ' || cur_ == '\t' || cur_ == '\n' || cur_ == '\r');
    }
    bool comment() {
        const int c = get();
        if (c == '/') {
            for (;;) {
                const int k = get();
                if (k == '\n' || k == '\r' || k == kEof || k == 0) return true;
            }
        }
        if (c == '*') {
            for (;;) {
                const int k = get();
                if (k == kEof || k == 0) {
                    err_ = "invalid comment; missing closing '*/'";
                    return false;
                }
                if (k == '*') {
                    if (get() == '/') return true;
                    unget();
                }
            }
        }
        err_ = "invalid comment; expecting '/' or '*' after '/'";
        return false;
    }
    Tok literal(const char* word, Tok t) {
        for (const char* p = word + 1; *p; ++p)
            if (get() != static_cast<unsigned char>(*p)) {
                err_ = "invalid literal";
                return Tok::ParseError;
            }
        return t;
    }
    int hex4() {
        int v = 0;
        for (int k = 0; k < 4; ++k) {
            const int c = get();
            v <<= 4;
            if (c >= '0' && c <= '9') v += c - '0';
            else if (c >= 'A' && c <= 'F') v += c - 'A' + 10;
            else if (c >= 'a' && c <= 'f') v += c - 'a' + 10;
            else return -1;
        }
        return v;
    }
    void put_utf8(uint32_t cp) {
        if (cp < 0x80) {
            buf_ += static_cast<char>(cp);
        } else if (cp <= 0x7FF) {
            buf_ += static_cast<char>(0xC0 | (cp >> 6));
            buf_ += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp <= 0xFFFF) {
            buf_ += static_cast<char>(0xE0 | (cp >> 12));
            buf_ += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            buf_ += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            buf_ += static_cast<char>(0xF0 | (cp >> 18));
            buf_ += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            buf_ += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            buf_ += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    // continuation bytes of a UTF-8 sequence: byte k must lie in [lo[k], hi[k]]
    bool tail(int n, const int* lo, const int* hi) {
        buf_ += static_cast<char>(cur_);
        for (int k = 0; k < n; ++k) {
            const int c = get();
            if (c < lo[k] || c > hi[k]) {
                err_ = "invalid string: ill-formed UTF-8 byte";
                return false;
            }
            buf_ += static_cast<char>(c);
        }
        return true;
    }
    Tok string() {
        begin_token();
        static const char* kNames[32] = {"NUL", "SOH", "STX", "ETX", "EOT", "ENQ", "ACK", "BEL", "BS", "HT", "LF",
                                         "VT", "FF", "CR", "SO", "SI", "DLE", "DC1", "DC2", "DC3", "DC4", "NAK",
                                         "SYN", "ETB", "CAN", "EM", "SUB", "ESC", "FS", "GS", "RS", "US"};
        for (;;) {
            const int c = get();
            if (c == kEof) {
                err_ = "invalid string: missing closing quote";
                return Tok::ParseError;
            }
            if (c == '"') return Tok::String;
            if (c == '\\') {
                const int e = get();
                switch (e) {
                    case '"': buf_ += '"'; break;
                    case '\\': buf_ += '\\'; break;
                    case '/': buf_ += '/'; break;
                    case 'b': buf_ += '\b'; break;
                    case 'f': buf_ += '\f'; break;
                    case 'n': buf_ += '\n'; break;
                    case 'r': buf_ += '\r'; break;
                    case 't': buf_ += '\t'; break;
                    case 'u': {
                        const int c1 = hex4();
                        if (c1 < 0) {
                            err_ = "invalid string: '\\u' must be followed by 4 hex digits";
                            return Tok::ParseError;
                        }
                        uint32_t cp = static_cast<uint32_t>(c1);
                        if (c1 >= 0xD800 && c1 <= 0xDBFF) {
                            if (get() == '\\' && get() == 'u') {
                                const int c2 = hex4();
                                if (c2 < 0) {
                                    err_ = "invalid string: '\\u' must be followed by 4 hex digits";
                                    return Tok::ParseError;
                                }
                                if (c2 < 0xDC00 || c2 > 0xDFFF) {
                                    err_ = "invalid string: surrogate U+D800..U+DBFF must be followed by U+DC00..U+DFFF";
                                    return Tok::ParseError;
                                }
                                cp = 0x10000u + ((static_cast<uint32_t>(c1) - 0xD800u) << 10) +
                                     (static_cast<uint32_t>(c2) - 0xDC00u);
                            } else {
                                err_ = "invalid string: surrogate U+D800..U+DBFF must be followed by U+DC00..U+DFFF";
                                return Tok::ParseError;
                            }
                        } else if (c1 >= 0xDC00 && c1 <= 0xDFFF) {
                            err_ = "invalid string: surrogate U+DC00..U+DFFF must follow U+D800..U+DBFF";
                            return Tok::ParseError;
                        }
                        put_utf8(cp);
                        break;
                    }
                    default: err_ = "invalid string: forbidden character after backslash"; return Tok::ParseError;
                }
                continue;
            }
            if (c < 0x20) {
                // "invalid string: control character U+00XX (NAME) must be escaped to \u00XX[ or \x]"
                static std::string msg[32];
                std::string& m = msg[c];
                if (m.empty()) {
                    char b[128];
                    const char* alt = c == 8 ? " or \\b" : c == 9 ? " or \\t" : c == 10 ? " or \\n" : c == 12 ? " or \\f"
                                                        : c == 13 ? " or \\r" : "";
                    std::snprintf(b, sizeof(b), "invalid string: control character U+%.4X (%s) must be escaped to \\u%.4X%s",
                                  static_cast<unsigned>(c), kNames[c], static_cast<unsigned>(c), alt);
                    m = b;
                }
                err_ = m.c_str();
                return Tok::ParseError;
            }
            if (c < 0x80) {
                buf_ += static_cast<char>(c);
                continue;
            }
            // UTF-8 sequences (RFC 3629, section 4)
            static const int any[3] = {0x80, 0x80, 0x80}, top[3] = {0xBF, 0xBF, 0xBF};
            bool ok;
            if (c >= 0xC2 && c <= 0xDF) {
                ok = tail(1, any, top);
            } else if (c == 0xE0) {
                const int lo[2] = {0xA0, 0x80};
                ok = tail(2, lo, top);
            } else if ((c >= 0xE1 && c <= 0xEC) || c == 0xEE || c == 0xEF) {
                ok = tail(2, any, top);
            } else if (c == 0xED) {
                const int hi[2] = {0x9F, 0xBF};
                ok = tail(2, any, hi);
            } else if (c == 0xF0) {
                const int lo[3] = {0x90, 0x80, 0x80};
                ok = tail(3, lo, top);
            } else if (c >= 0xF1 && c <= 0xF3) {
                ok = tail(3, any, top);
            } else if (c == 0xF4) {
                const int hi[3] = {0x8F, 0xBF, 0xBF};
                ok = tail(3, any, hi);
            } else {
                err_ = "invalid string: ill-formed UTF-8 byte";
                ok = false;
            }
            if (!ok) return Tok::ParseError;
        }
    }
    static bool digit(int c) { return c >= '0' && c <= '9'; }
    Tok number() {
        begin_token();
        Tok kind = Tok::Unsigned;
        buf_ += static_cast<char>(cur_);
        int c = cur_;
        if (c == '-') {
            kind = Tok::Integer;
            c = get();
            if (!digit(c)) {
                err_ = "invalid number; expected digit after '-'";
                return Tok::ParseError;
            }
            buf_ += static_cast<char>(c);
        }
        if (c != '0')
            while (digit(c = get())) buf_ += static_cast<char>(c);
        else
            c = get();
        if (c == '.') {
            kind = Tok::Float;
            buf_ += '.';
            c = get();
            if (!digit(c)) {
                err_ = "invalid number; expected digit after '.'";
                return Tok::ParseError;
            }
            buf_ += static_cast<char>(c);
            while (digit(c = get())) buf_ += static_cast<char>(c);
        }
        if (c == 'e' || c == 'E') {
            kind = Tok::Float;
            buf_ += static_cast<char>(c);
            c = get();
            if (c == '+' || c == '-') {
                buf_ += static_cast<char>(c);
                c = get();
                if (!digit(c)) {
                    err_ = "invalid number; expected digit after exponent sign";
                    return Tok::ParseError;
                }
            } else if (!digit(c)) {
                err_ = "invalid number; expected '+', '-', or digit after exponent";
                return Tok::ParseError;
            }
            buf_ += static_cast<char>(c);
            while (digit(c = get())) buf_ += static_cast<char>(c);
        }
        unget();
        // integers that do not fit their 64-bit type become doubles, like the library's strtoull / strtoll path
        if (kind == Tok::Unsigned) {
            errno = 0;
            const unsigned long long v = std::strtoull(buf_.c_str(), nullptr, 10);
            if (errno == 0) {
                u = v;
                return Tok::Unsigned;
            }
        } else if (kind == Tok::Integer) {
            errno = 0;
            const long long v = std::strtoll(buf_.c_str(), nullptr, 10);
            if (errno == 0) {
                i = v;
                return Tok::Integer;
            }
        }
        d = std::strtod(buf_.c_str(), nullptr);
        return Tok::Float;
    }

    const std::string& s_;
    bool comments_;
    size_t pos_ = 0;
    int cur_ = kEof;
    bool unget_ = false;
    uint64_t read_total_ = 0, read_line_ = 0, lines_ = 0;
    std::string tok_, buf_;
    const char* err_ = "";
};

class Reader {
public:
    Reader(const std::string& text, bool comments) : sc_(text, comments) { next(); }

    Json document() {
        Json v = value(0);
        if (next() != Tok::End) fail(Tok::End, "value");
        return v;
    }

private:
    Tok next() { return last_ = sc_.scan(); }
    [[noreturn]] void fail(Tok expected, const char* context) {
        std::string m = std::string("syntax error while parsing ") + context + " - ";
        if (last_ == Tok::ParseError)
            m += std::string(sc_.error()) + "; last read: '" + sc_.token_text() + "'";
        else
            m += std::string("unexpected ") + tok_name(last_);
        if (expected != Tok::Uninitialized) m += std::string("; expected ") + tok_name(expected);
        raise(m);
    }
    [[noreturn]] void raise(const std::string& what) {
        throw JsonError("[json.exception.parse_error.101] parse error" + sc_.where() + ": " + what);
    }
    // the value starting with the current token
    Json value(int depth) {
        if (depth > 100000) raise("nesting deeper than 100000 levels");
        switch (last_) {
            case Tok::BeginObject: {
                Json o = Json::object();
                if (next() == Tok::EndObject) return o;
                for (;;) {
                    if (last_ != Tok::String) fail(Tok::String, "object key");
                    const std::string key = sc_.text();
                    if (next() != Tok::NameSep) fail(Tok::NameSep, "object separator");
                    next();
                    o[key] = value(depth + 1);
                    if (next() == Tok::ValueSep) {
                        next();
                        continue;
                    }
                    if (last_ == Tok::EndObject) return o;
                    fail(Tok::EndObject, "object");
                }
            }
            case Tok::BeginArray: {
                Json a = Json::array();
                if (next() == Tok::EndArray) return a;
                for (;;) {
                    a.push_back(value(depth + 1));
                    if (next() == Tok::ValueSep) {
                        next();
                        continue;
                    }
                    if (last_ == Tok::EndArray) return a;
                    fail(Tok::EndArray, "array");
                }
            }
            case Tok::Float:
                if (!std::isfinite(sc_.d))
                    throw JsonRangeError("[json.exception.out_of_range.406] number overflow parsing '" + sc_.token_text() +
                                         "'");
                return Json(sc_.d);
            case Tok::True: return Json(true);
            case Tok::False: return Json(false);
            case Tok::Null: return Json();
            case Tok::Integer: return Json(static_cast<long long>(sc_.i));
            case Tok::Unsigned: return Json(static_cast<unsigned long long>(sc_.u));
            case Tok::String: return Json(sc_.text());
            case Tok::ParseError: fail(Tok::Uninitialized, "value");
            case Tok::End:
                if (sc_.read_total() == 1)
                    raise("attempting to parse an empty input; check that your input string or stream contains the "
                          "expected JSON");
                fail(Tok::LiteralOrValue, "value");
            default: fail(Tok::LiteralOrValue, "value");
        }
    }

    Scanner sc_;
    Tok last_ = Tok::Uninitialized;
};

}  // namespace

Json Json::parse(const std::string& text, bool allow_comments) {
    Reader r(text, allow_comments);
    return r.document();
}

}  // namespace eng
