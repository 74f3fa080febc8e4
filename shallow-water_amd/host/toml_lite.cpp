// toml_lite.cpp -- see toml_lite.hpp.
#include "toml_lite.hpp"

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace toml_lite {
namespace {

[[noreturn]] void fail(const std::string &src, int line, const std::string &msg) {
    throw std::runtime_error(src + ":" + std::to_string(line) + ": " + msg);
}

std::string trim(const std::string &s) {
    std::size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

bool bare_key_char(char c) {
    return std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '-';
}

// A basic string starting at s[pos] == '"'; returns the decoded text, pos past the quote.
std::string basic_string(const std::string &s, std::size_t &pos, const std::string &src,
                         int line) {
    std::string out;
    ++pos;
    while (pos < s.size()) {
        const char c = s[pos++];
        if (c == '"') return out;
        if (c != '\\') {
            out += c;
            continue;
        }
        if (pos >= s.size()) break;
        const char e = s[pos++];
        switch (e) {
            case 'n': out += '\n'; break;
            case 't': out += '\t'; break;
            case 'r': out += '\r'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case 'u':
            case 'U': {
                const int n = e == 'u' ? 4 : 8;
                if (pos + n > s.size()) fail(src, line, "truncated unicode escape");
                const unsigned long cp = std::strtoul(s.substr(pos, n).c_str(), nullptr, 16);
                pos += n;
                // UTF-8 encode
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
                break;
            }
            default: fail(src, line, std::string("invalid escape \\") + e);
        }
    }
    fail(src, line, "unterminated string");
}

std::string literal_string(const std::string &s, std::size_t &pos, const std::string &src,
                           int line) {
    const std::size_t end = s.find('\'', pos + 1);
    if (end == std::string::npos) fail(src, line, "unterminated literal string");
    std::string out = s.substr(pos + 1, end - pos - 1);
    pos = end + 1;
    return out;
}

// Everything after position pos must be blank or a comment.
void expect_end(const std::string &s, std::size_t pos, const std::string &src, int line) {
    while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\r')) ++pos;
    if (pos < s.size() && s[pos] != '#') fail(src, line, "unexpected text after value");
}

Value scalar(const std::string &raw, const std::string &src, int line) {
    Value v;
    v.line = line;
    const std::string t = raw;
    if (t == "true" || t == "false") {
        v.kind = Value::Kind::Boolean;
        v.b = t == "true";
        return v;
    }
    if (t == "inf" || t == "+inf" || t == "-inf" || t == "nan" || t == "+nan" || t == "-nan") {
        v.kind = Value::Kind::Float;
        const bool neg = t[0] == '-';
        v.f = t.find("inf") != std::string::npos ? HUGE_VAL : std::nan("");
        if (neg) v.f = -v.f;
        return v;
    }
    std::string d;  // digits without '_'
    for (std::size_t k = 0; k < t.size(); ++k) {
        const char c = t[k];
        if (c == '_') {
            const bool ok = k > 0 && k + 1 < t.size() &&
                            std::isxdigit(static_cast<unsigned char>(t[k - 1])) &&
                            std::isxdigit(static_cast<unsigned char>(t[k + 1]));
            if (!ok) fail(src, line, "misplaced '_' in number '" + t + "'");
            continue;
        }
        d += c;
    }
    if (d.empty()) fail(src, line, "missing value");
    std::size_t start = (d[0] == '+' || d[0] == '-') ? 1 : 0;
    if (d.size() > start + 1 && d[start] == '0' &&
        (d[start + 1] == 'x' || d[start + 1] == 'o' || d[start + 1] == 'b')) {
        if (start != 0) fail(src, line, "signed prefixed integer '" + t + "'");
        const int base = d[1] == 'x' ? 16 : (d[1] == 'o' ? 8 : 2);
        errno = 0;
        char *end = nullptr;
        const unsigned long long u = std::strtoull(d.c_str() + 2, &end, base);
        if (errno || *end != '\0' || d.size() == 2) fail(src, line, "invalid integer '" + t + "'");
        v.kind = Value::Kind::Integer;
        v.i = static_cast<std::int64_t>(u);
        return v;
    }
    bool is_float = false;
    for (char c : d)
        if (c == '.' || c == 'e' || c == 'E') is_float = true;
    for (std::size_t k = start; k < d.size(); ++k) {
        const char c = d[k];
        if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '.' || c == 'e' || c == 'E' ||
              c == '+' || c == '-'))
            fail(src, line, "invalid value '" + t + "' (strings need quotes)");
    }
    errno = 0;
    char *end = nullptr;
    if (is_float) {
        v.kind = Value::Kind::Float;
        v.f = std::strtod(d.c_str(), &end);
    } else {
        v.kind = Value::Kind::Integer;
        v.i = std::strtoll(d.c_str(), &end, 10);
    }
    if (errno || *end != '\0') fail(src, line, "invalid number '" + t + "'");
    return v;
}

}  // namespace

Document parse(const std::string &text, const std::string &src) {
    Document doc;
    doc[""];
    std::string table;
    std::istringstream in(text);
    std::string raw;
    int line = 0;
    while (std::getline(in, raw)) {
        ++line;
        const std::string s = trim(raw);
        if (s.empty() || s[0] == '#') continue;
        if (s[0] == '[') {
            if (s.size() > 1 && s[1] == '[') fail(src, line, "arrays of tables are not supported");
            const std::size_t close = s.find(']');
            if (close == std::string::npos) fail(src, line, "unterminated table header");
            expect_end(s, close + 1, src, line);
            table = trim(s.substr(1, close - 1));
            if (table.empty()) fail(src, line, "empty table name");
            for (char c : table)
                if (!(bare_key_char(c) || c == '.')) fail(src, line, "invalid table name");
            if (doc.count(table) && !doc[table].empty())
                fail(src, line, "table [" + table + "] defined twice");
            doc[table];
            continue;
        }
        // key
        std::size_t pos = 0;
        std::string key;
        if (s[0] == '"') {
            key = basic_string(s, pos, src, line);
        } else if (s[0] == '\'') {
            key = literal_string(s, pos, src, line);
        } else {
            while (pos < s.size() && bare_key_char(s[pos])) ++pos;
            key = s.substr(0, pos);
            if (key.empty()) fail(src, line, "expected a key");
        }
        while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) ++pos;
        if (pos >= s.size() || s[pos] != '=') fail(src, line, "expected '=' after key '" + key + "'");
        ++pos;
        while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) ++pos;
        if (pos >= s.size()) fail(src, line, "missing value for key '" + key + "'");
        Value v;
        v.line = line;
        const char c = s[pos];
        if (c == '"') {
            if (s.compare(pos, 3, "\"\"\"") == 0) fail(src, line, "multi-line strings are not supported");
            v.kind = Value::Kind::String;
            v.s = basic_string(s, pos, src, line);
            expect_end(s, pos, src, line);
        } else if (c == '\'') {
            if (s.compare(pos, 3, "'''") == 0) fail(src, line, "multi-line strings are not supported");
            v.kind = Value::Kind::String;
            v.s = literal_string(s, pos, src, line);
            expect_end(s, pos, src, line);
        } else if (c == '[' || c == '{') {
            fail(src, line, "arrays and inline tables are not supported");
        } else {
            std::size_t end = s.find('#', pos);
            if (end == std::string::npos) end = s.size();
            v = scalar(trim(s.substr(pos, end - pos)), src, line);
        }
        Table &t = doc[table];
        if (t.count(key)) fail(src, line, "duplicate key '" + key + "'");
        t[key] = v;
    }
    return doc;
}

Document parse_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open '" + path + "'");
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), path);
}

std::optional<double> get_double(const Table &t, const std::string &key) {
    const auto it = t.find(key);
    if (it == t.end()) return std::nullopt;
    if (it->second.kind == Value::Kind::Float) return it->second.f;
    if (it->second.kind == Value::Kind::Integer) return static_cast<double>(it->second.i);
    return std::nullopt;
}

std::optional<std::int64_t> get_int(const Table &t, const std::string &key) {
    const auto it = t.find(key);
    if (it == t.end() || it->second.kind != Value::Kind::Integer) return std::nullopt;
    return it->second.i;
}

std::optional<std::string> get_string(const Table &t, const std::string &key) {
    const auto it = t.find(key);
    if (it == t.end() || it->second.kind != Value::Kind::String) return std::nullopt;
    return it->second.s;
}

std::optional<bool> get_bool(const Table &t, const std::string &key) {
    const auto it = t.find(key);
    if (it == t.end() || it->second.kind != Value::Kind::Boolean) return std::nullopt;
    return it->second.b;
}

}  // namespace toml_lite
