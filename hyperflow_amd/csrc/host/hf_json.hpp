// hf_json.hpp -- minimal JSON reader for HyperFlow case files.
//
// The reference reads its case with the vendored nlohmann/json
// (src/fvm_2d_cart/main.cpp:26-28).  Case files only use objects, arrays,
// numbers, strings and booleans, so this reader covers RFC 8259 values with
// numbers parsed by strtod (correctly rounded, like the reference's parser).
// Errors throw hf_json::Error with the byte offset.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hf_json {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

class Value {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };

    Value() = default;
    static Value parse(const std::string& text);
    static Value load(const std::string& path);

    Kind kind() const { return kind_; }
    bool is_object() const { return kind_ == Object; }
    bool is_array() const { return kind_ == Array; }
    bool has(const std::string& key) const;

    const Value& operator[](const std::string& key) const;   // throws if missing
    const Value& operator[](size_t i) const;
    size_t size() const;

    double num() const;
    int integer() const;
    const std::string& str() const;
    bool boolean() const;
    std::vector<double> numbers() const;                      // array of numbers

    const std::vector<std::string>& keys() const { return keys_; }

private:
    friend class Parser;
    Kind kind_ = Null;
    double num_ = 0.0;
    bool bool_ = false;
    std::string str_;
    std::vector<Value> items_;
    std::vector<std::string> keys_;                           // object keys in file order
    std::map<std::string, size_t> index_;
};

}  // namespace hf_json
