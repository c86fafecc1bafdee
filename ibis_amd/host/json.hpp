// json.hpp -- a small JSON value and parser for the host layer.
//
// Enough JSON to read the reference's job files (config/config.json):
// null, booleans, numbers (as double), strings (with \" \\ \/ \b \f \n \r
// \t and \uXXXX escapes of the Basic Multilingual Plane), arrays and
// objects.  Objects are std::map, so they iterate in byte-wise sorted key
// order -- the order the reference's json objects iterate in, which fixes
// its boundary-tag (and ghost-cell) numbering.
#ifndef IBIS_HOST_JSON_HPP
#define IBIS_HOST_JSON_HPP

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ibis {

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    static Json parse(const std::string& text);
    static Json parse_file(const std::string& path);

    Type type() const { return type_; }
    bool is_null() const { return type_ == Type::Null; }
    bool is_object() const { return type_ == Type::Object; }
    bool is_array() const { return type_ == Type::Array; }
    bool is_number() const { return type_ == Type::Number; }
    bool is_string() const { return type_ == Type::String; }
    bool is_bool() const { return type_ == Type::Bool; }

    // typed access; throws std::runtime_error on a type mismatch
    double number() const;
    bool boolean() const;
    const std::string& string() const;
    const std::vector<Json>& array() const;
    const std::map<std::string, Json>& object() const;

    // object member (throws if absent) and presence test
    const Json& at(const std::string& key) const;
    bool contains(const std::string& key) const;
    // member as a number / bool / string, or `fallback` when absent
    double value(const std::string& key, double fallback) const;
    bool value(const std::string& key, bool fallback) const;
    std::string value(const std::string& key, const char* fallback) const;

private:
    friend class JsonParser;
    Type type_ = Type::Null;
    bool b_ = false;
    double n_ = 0.0;
    std::string s_;
    std::vector<Json> a_;
    std::map<std::string, Json> o_;
};

}  // namespace ibis

#endif
