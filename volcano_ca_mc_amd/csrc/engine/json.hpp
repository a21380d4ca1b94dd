// Minimal ordered JSON for the engine's config and report files.
//
// The reference writes config_used.json / stats.json / times.json with nlohmann::ordered_json::dump(2)
// (engine/src/io/ResultWriter.cpp:239-243).  This value type keeps insertion order, distinguishes
// signed / unsigned integers from doubles like that library does, and its dump(2) produces the same
// layout: two-space indentation, "key": value, integers verbatim, doubles as the shortest round-trip
// digits placed in fixed notation for decimal exponents in (-4, 15] (".0" appended to integral values)
// and in d.ddde+XX notation otherwise, non-ASCII UTF-8 passed through.
#pragma once

#include <cstdint>
#include <map>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace eng {

class JsonError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};
// A number literal too large for a double (not a syntax error: callers that treat syntax errors specially let it
// through, as the reference's JSON library does with its out_of_range error).
class JsonRangeError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Json {
public:
    enum class Type { Null, Bool, Int, UInt, Float, String, Array, Object };

    Json() = default;
    Json(std::nullptr_t) {}
    Json(bool b) : type_(Type::Bool), b_(b) {}
    Json(int v) : type_(v < 0 ? Type::Int : Type::UInt), i_(v), u_(static_cast<uint64_t>(v)) {}
    Json(long v) : type_(v < 0 ? Type::Int : Type::UInt), i_(v), u_(static_cast<uint64_t>(v)) {}
    Json(long long v) : type_(v < 0 ? Type::Int : Type::UInt), i_(v), u_(static_cast<uint64_t>(v)) {}
    Json(unsigned v) : type_(Type::UInt), i_(static_cast<int64_t>(v)), u_(v) {}
    Json(unsigned long v) : type_(Type::UInt), i_(static_cast<int64_t>(v)), u_(v) {}
    Json(unsigned long long v) : type_(Type::UInt), i_(static_cast<int64_t>(v)), u_(v) {}
    Json(double v) : type_(Type::Float), d_(v) {}
    Json(float v) : type_(Type::Float), d_(static_cast<double>(v)) {}
    Json(const char* s) : type_(Type::String), s_(s) {}
    Json(std::string s) : type_(Type::String), s_(std::move(s)) {}
    template <class T>
    Json(const std::vector<T>& v) : type_(Type::Array) {
        for (const T& x : v) a_.push_back(Json(x));
    }

    static Json object() {
        Json j;
        j.type_ = Type::Object;
        return j;
    }
    static Json array() {
        Json j;
        j.type_ = Type::Array;
        return j;
    }
    static Json array(std::initializer_list<Json> items) {
        Json j = array();
        for (const Json& x : items) j.a_.push_back(x);
        return j;
    }

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_bool() const { return type_ == Type::Bool; }
    bool is_object() const { return type_ == Type::Object; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_string() const { return type_ == Type::String; }
    bool is_number() const { return type_ == Type::Int || type_ == Type::UInt || type_ == Type::Float; }
    bool is_number_integer() const { return type_ == Type::Int || type_ == Type::UInt; }
    bool is_number_unsigned() const { return type_ == Type::UInt; }
    bool is_number_float() const { return type_ == Type::Float; }

    bool get_bool() const;
    int64_t get_int() const;
    uint64_t get_uint() const;
    double get_double() const;
    const std::string& get_string() const;

    // object access (insertion ordered)
    bool contains(const std::string& key) const;
    Json& operator[](const std::string& key);  // inserts null when missing (object or null only)
    const Json& at(const std::string& key) const;
    std::vector<std::pair<std::string, Json>>& items() { return o_; }
    const std::vector<std::pair<std::string, Json>>& items() const { return o_; }
    void erase(const std::string& key);

    // array access
    size_t size() const { return type_ == Type::Array ? a_.size() : (type_ == Type::Object ? o_.size() : 0); }
    Json& operator[](size_t i) { return a_.at(i); }
    const Json& operator[](size_t i) const { return a_.at(i); }
    void push_back(Json v);
    std::vector<Json>& elements() { return a_; }
    const std::vector<Json>& elements() const { return a_; }

    std::string dump(int indent = -1) const;
    static Json parse(const std::string& text, bool allow_comments = true);

    bool operator==(const Json& o) const;
    bool operator!=(const Json& o) const { return !(*this == o); }

private:
    void dump_to(std::string& out, int indent, int level) const;

    Type type_ = Type::Null;
    bool b_ = false;
    int64_t i_ = 0;
    uint64_t u_ = 0;
    double d_ = 0.0;
    std::string s_;
    std::vector<Json> a_;
    std::vector<std::pair<std::string, Json>> o_;
};

// Number formatting used by dump (exposed for the CSV / console writers and tests).
std::string format_double(double v);

}  // namespace eng
