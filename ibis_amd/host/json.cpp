// json.cpp -- recursive-descent JSON parser of json.hpp.
#include "json.hpp"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace ibis {

class JsonParser {
public:
    explicit JsonParser(const std::string& t) : t_(t) {}

    Json document() {
        Json v = value();
        ws();
        if (i_ != t_.size()) error("trailing characters");
        return v;
    }

private:
    const std::string& t_;
    size_t i_ = 0;

    [[noreturn]] void error(const char* what) const {
        throw std::runtime_error("JSON: " + std::string(what) + " at offset " +
                                 std::to_string(i_));
    }
    void ws() {
        while (i_ < t_.size() &&
               (t_[i_] == ' ' || t_[i_] == '\t' || t_[i_] == '\n' || t_[i_] == '\r'))
            i_++;
    }
    char peek() {
        ws();
        if (i_ >= t_.size()) error("unexpected end of input");
        return t_[i_];
    }
    void expect(char c) {
        if (peek() != c) error((std::string("expected '") + c + "'").c_str());
        i_++;
    }
    void literal(const char* word) {
        for (const char* p = word; *p; p++, i_++)
            if (i_ >= t_.size() || t_[i_] != *p) error("bad literal");
    }
    Json value() {
        Json v;
        const char c = peek();
        if (c == '{') {
            v.type_ = Json::Type::Object;
            i_++;
            if (peek() == '}') {
                i_++;
                return v;
            }
            for (;;) {
                if (peek() != '"') error("expected a key");
                std::string k = str();
                expect(':');
                v.o_[k] = value();
                if (peek() == ',') {
                    i_++;
                    continue;
                }
                expect('}');
                return v;
            }
        }
        if (c == '[') {
            v.type_ = Json::Type::Array;
            i_++;
            if (peek() == ']') {
                i_++;
                return v;
            }
            for (;;) {
                v.a_.push_back(value());
                if (peek() == ',') {
                    i_++;
                    continue;
                }
                expect(']');
                return v;
            }
        }
        if (c == '"') {
            v.type_ = Json::Type::String;
            v.s_ = str();
            return v;
        }
        if (c == 't') {
            literal("true");
            v.type_ = Json::Type::Bool;
            v.b_ = true;
            return v;
        }
        if (c == 'f') {
            literal("false");
            v.type_ = Json::Type::Bool;
            return v;
        }
        if (c == 'n') {
            literal("null");
            return v;
        }
        // number: strtod reads the JSON grammar (and a little more)
        const char* s = t_.c_str() + i_;
        char* end = nullptr;
        v.n_ = std::strtod(s, &end);
        if (end == s) error("bad value");
        i_ += (size_t)(end - s);
        v.type_ = Json::Type::Number;
        return v;
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
        expect('"');
        std::string out;
        while (i_ < t_.size() && t_[i_] != '"') {
            char c = t_[i_++];
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= t_.size()) error("bad escape");
            c = t_[i_++];
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
                    if (i_ + 4 > t_.size()) error("bad \\u escape");
                    const unsigned cp = (unsigned)std::strtoul(t_.substr(i_, 4).c_str(), nullptr, 16);
                    i_ += 4;
                    utf8(out, cp);
                    break;
                }
                default: error("bad escape");
            }
        }
        if (i_ >= t_.size()) error("unterminated string");
        i_++;
        return out;
    }
};

Json Json::parse(const std::string& text) { return JsonParser(text).document(); }

Json Json::parse_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

static std::runtime_error type_error(const char* want) {
    return std::runtime_error(std::string("JSON value is not ") + want);
}

double Json::number() const {
    if (type_ != Type::Number) throw type_error("a number");
    return n_;
}
bool Json::boolean() const {
    if (type_ != Type::Bool) throw type_error("a boolean");
    return b_;
}
const std::string& Json::string() const {
    if (type_ != Type::String) throw type_error("a string");
    return s_;
}
const std::vector<Json>& Json::array() const {
    if (type_ != Type::Array) throw type_error("an array");
    return a_;
}
const std::map<std::string, Json>& Json::object() const {
    if (type_ != Type::Object) throw type_error("an object");
    return o_;
}
const Json& Json::at(const std::string& key) const {
    const auto& o = object();
    auto it = o.find(key);
    if (it == o.end()) throw std::runtime_error("JSON: missing key '" + key + "'");
    return it->second;
}
bool Json::contains(const std::string& key) const {
    return type_ == Type::Object && o_.count(key) != 0;
}
double Json::value(const std::string& key, double fallback) const {
    return contains(key) ? at(key).number() : fallback;
}
bool Json::value(const std::string& key, bool fallback) const {
    return contains(key) ? at(key).boolean() : fallback;
}
std::string Json::value(const std::string& key, const char* fallback) const {
    return contains(key) ? at(key).string() : std::string(fallback);
}

}  // namespace ibis
