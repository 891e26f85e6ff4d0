// Minimal JSON value, parser and writer for the parameter file (the reference
// maps its structs with json_struct, dataset/parameters.hpp:19-84; this reader
// maps the same keys by hand).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace xpr {

struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;   // insertion order kept

    const Json *get(const std::string &key) const
    {
        for (const auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    static Json number(double v) { Json j; j.kind = Number; j.num = v; return j; }
    static Json boolean(bool v) { Json j; j.kind = Bool; j.b = v; return j; }
    static Json string(const std::string &v) { Json j; j.kind = String; j.str = v; return j; }
    static Json object() { Json j; j.kind = Object; return j; }
    static Json array() { Json j; j.kind = Array; return j; }
    void set(const std::string &k, Json v) { obj.emplace_back(k, std::move(v)); }
};

struct JsonError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

Json json_parse(const std::string &text);                 // throws JsonError
std::string json_dump(const Json &j, int indent = 4);

} // namespace xpr
