// toml_lite.hpp -- a small TOML reader for the simulation configuration files.
//
// Supports what those files use: [table] headers (dotted names allowed), key = value pairs
// with bare or quoted keys, basic "..." and literal '...' strings, integers (with _
// separators, +/- sign, 0x/0o/0b prefixes), floats (fraction, exponent, inf, nan), booleans,
// and # comments.  Arrays, inline tables and multi-line strings are rejected with an error
// naming the line.  Duplicate keys are errors, as in TOML.
#pragma once

#include <cstdint>
#include <map>
#include <optional>
#include <string>

namespace toml_lite {

struct Value {
    enum class Kind { String, Integer, Float, Boolean } kind{Kind::String};
    std::string s;
    std::int64_t i{};
    double f{};
    bool b{};
    int line{};
};

// table name -> key -> value ("" is the root table)
using Table = std::map<std::string, Value>;
using Document = std::map<std::string, Table>;

Document parse(const std::string &text, const std::string &source_name = "<string>");
Document parse_file(const std::string &path);

// Typed access.  A float key accepts an integer literal (converted); an integer key accepts
// only an integer literal.  Missing key -> nullopt; wrong type -> nullopt.
std::optional<double> get_double(const Table &t, const std::string &key);
std::optional<std::int64_t> get_int(const Table &t, const std::string &key);
std::optional<std::string> get_string(const Table &t, const std::string &key);
std::optional<bool> get_bool(const Table &t, const std::string &key);

}  // namespace toml_lite
