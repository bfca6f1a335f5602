#include "json.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace nmk {

class JsonParser {
public:
    explicit JsonParser(const std::string &s) : s_(s) {}

    Json parse_document() {
        skip_ws();
        Json v = parse_value(0);
        skip_ws();
        if (pos_ != s_.size()) fail("unexpected trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t pos_ = 0;

    [[noreturn]] void fail(const std::string &what) const {
        throw JsonError("[json.exception.parse_error] syntax error at byte " + std::to_string(pos_ + 1) + ": " + what);
    }
    void skip_ws() {
        while (pos_ < s_.size() && (s_[pos_] == ' ' || s_[pos_] == '\t' || s_[pos_] == '\n' || s_[pos_] == '\r'))
            pos_++;
    }
    char peek() const { return pos_ < s_.size() ? s_[pos_] : '\0'; }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        pos_++;
    }
    bool consume_word(const char *w) {
        size_t n = std::strlen(w);
        if (s_.compare(pos_, n, w) == 0) {
            pos_ += n;
            return true;
        }
        return false;
    }

    Json parse_value(int depth) {
        if (depth > 512) fail("nesting too deep");
        Json v;
        char c = peek();
        if (c == '{') {
            v.type_ = Json::Type::Object;
            pos_++;
            skip_ws();
            if (peek() == '}') {
                pos_++;
                return v;
            }
            for (;;) {
                skip_ws();
                if (peek() != '"') fail("expected object key");
                std::string key = parse_string();
                skip_ws();
                expect(':');
                skip_ws();
                Json item = parse_value(depth + 1);
                // duplicate keys: the last one wins
                bool replaced = false;
                for (auto &kv : v.obj_)
                    if (kv.first == key) {
                        kv.second = std::move(item);
                        replaced = true;
                        break;
                    }
                if (!replaced) v.obj_.emplace_back(std::move(key), std::move(item));
                skip_ws();
                if (peek() == ',') {
                    pos_++;
                    continue;
                }
                expect('}');
                return v;
            }
        }
        if (c == '[') {
            v.type_ = Json::Type::Array;
            pos_++;
            skip_ws();
            if (peek() == ']') {
                pos_++;
                return v;
            }
            for (;;) {
                skip_ws();
                v.arr_.push_back(parse_value(depth + 1));
                skip_ws();
                if (peek() == ',') {
                    pos_++;
                    continue;
                }
                expect(']');
                return v;
            }
        }
        if (c == '"') {
            v.type_ = Json::Type::String;
            v.str_ = parse_string();
            return v;
        }
        if (consume_word("true")) {
            v.type_ = Json::Type::Bool;
            v.b_ = true;
            return v;
        }
        if (consume_word("false")) {
            v.type_ = Json::Type::Bool;
            v.b_ = false;
            return v;
        }
        if (consume_word("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            v.type_ = Json::Type::Number;
            v.num_ = parse_number();
            return v;
        }
        fail("unexpected character");
    }

    double parse_number() {
        size_t start = pos_;
        if (peek() == '-') pos_++;
        if (peek() == '0') {
            pos_++;
        } else if (peek() >= '1' && peek() <= '9') {
            while (peek() >= '0' && peek() <= '9') pos_++;
        } else {
            fail("invalid number");
        }
        if (peek() == '.') {
            pos_++;
            if (!(peek() >= '0' && peek() <= '9')) fail("invalid number");
            while (peek() >= '0' && peek() <= '9') pos_++;
        }
        if (peek() == 'e' || peek() == 'E') {
            pos_++;
            if (peek() == '+' || peek() == '-') pos_++;
            if (!(peek() >= '0' && peek() <= '9')) fail("invalid number");
            while (peek() >= '0' && peek() <= '9') pos_++;
        }
        std::string tok = s_.substr(start, pos_ - start);
        return std::strtod(tok.c_str(), nullptr);
    }

    static void put_utf8(std::string &out, unsigned cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    unsigned parse_hex4() {
        if (pos_ + 4 > s_.size()) fail("truncated unicode escape");
        unsigned v = 0;
        for (int i = 0; i < 4; i++) {
            char h = s_[pos_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= unsigned(h - '0');
            else if (h >= 'a' && h <= 'f') v |= unsigned(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= unsigned(h - 'A' + 10);
            else fail("invalid unicode escape");
        }
        return v;
    }
    std::string parse_string() {
        expect('"');
        std::string out;
        for (;;) {
            if (pos_ >= s_.size()) fail("unterminated string");
            char c = s_[pos_++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (pos_ >= s_.size()) fail("unterminated escape");
            char e = s_[pos_++];
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
                    unsigned cp = parse_hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF && pos_ + 1 < s_.size() && s_[pos_] == '\\' && s_[pos_ + 1] == 'u') {
                        pos_ += 2;
                        unsigned lo = parse_hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("invalid escape");
            }
        }
        return out;
    }
};

Json Json::parse(const std::string &text) { return JsonParser(text).parse_document(); }

const char *Json::type_name() const {
    switch (type_) {
        case Type::Null: return "null";
        case Type::Bool: return "boolean";
        case Type::Number: return "number";
        case Type::String: return "string";
        case Type::Array: return "array";
        case Type::Object: return "object";
    }
    return "?";
}

const Json *Json::find(const std::string &key) const {
    if (type_ != Type::Object) return nullptr;
    for (const auto &kv : obj_)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

const Json &Json::at(const std::string &key) const {
    if (type_ != Type::Object)
        throw JsonError(std::string("[json.exception.type_error.304] cannot use at() with ") + type_name());
    const Json *v = find(key);
    if (!v) throw JsonError("[json.exception.out_of_range.403] key '" + key + "' not found");
    return *v;
}

bool Json::as_bool() const {
    if (type_ != Type::Bool)
        throw JsonError(std::string("[json.exception.type_error.302] type must be boolean, but is ") + type_name());
    return b_;
}

double Json::as_double() const {
    if (type_ != Type::Number)
        throw JsonError(std::string("[json.exception.type_error.302] type must be number, but is ") + type_name());
    return num_;
}

long long Json::as_int() const {
    double v = as_double();
    if (!std::isfinite(v) || std::fabs(v) > 9.2e18)
        throw JsonError("[json.exception.type_error.302] number out of integer range");
    return (long long)v;  // truncation, as a static_cast from a stored float
}

const std::string &Json::as_string() const {
    if (type_ != Type::String)
        throw JsonError(std::string("[json.exception.type_error.302] type must be string, but is ") + type_name());
    return str_;
}

const std::vector<Json> &Json::items() const {
    if (type_ != Type::Array)
        throw JsonError(std::string("[json.exception.type_error.302] type must be array, but is ") + type_name());
    return arr_;
}

std::vector<double> Json::as_double_vector() const {
    std::vector<double> v;
    for (const auto &e : items()) v.push_back(e.as_double());
    return v;
}

std::vector<long long> Json::as_int_vector() const {
    std::vector<long long> v;
    for (const auto &e : items()) v.push_back(e.as_int());
    return v;
}

}  // namespace nmk
