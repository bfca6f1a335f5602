// Minimal JSON reader for the bpm-solver settings string.
//
// The reference parses its settings with the vendored nlohmann::json
// (BeamPropagationSolver/src/UtilitiesLib/json.h) and reads values with
// `at(key).get<T>()` (SettingsLib/BaseSettings.h:22-54).  This reader
// supports the same document model (objects, arrays, strings, numbers,
// booleans, null) and the conversions that code path relies on:
// number -> int/double (truncating like a static_cast), anything else is a
// type error.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace nmk {

struct JsonError : std::runtime_error {
    explicit JsonError(const std::string &m) : std::runtime_error(m) {}
};

class Json {
public:
    enum class Type { Null, Bool, Number, String, Array, Object };

    Json() = default;
    static Json parse(const std::string &text);

    Type type() const { return type_; }
    const char *type_name() const;
    bool is_object() const { return type_ == Type::Object; }

    // object access: find returns nullptr when absent, at throws
    const Json *find(const std::string &key) const;
    const Json &at(const std::string &key) const;

    bool as_bool() const;
    double as_double() const;
    long long as_int() const;
    const std::string &as_string() const;
    std::vector<double> as_double_vector() const;
    std::vector<long long> as_int_vector() const;
    const std::vector<Json> &items() const;

private:
    Type type_ = Type::Null;
    bool b_ = false;
    double num_ = 0;
    std::string str_;
    std::vector<Json> arr_;
    std::vector<std::pair<std::string, Json>> obj_;

    friend class JsonParser;
};

}  // namespace nmk
