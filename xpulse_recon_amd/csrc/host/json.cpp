#include "json.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <sstream>

namespace xpr {
namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &t) : s(t) {}

    [[noreturn]] void error(const std::string &what)
    {
        size_t line = 1, col = 1;
        for (size_t k = 0; k < i && k < s.size(); ++k) {
            if (s[k] == '\n') { ++line; col = 1; } else { ++col; }
        }
        throw JsonError("JSON error at line " + std::to_string(line) + ", column " +
                        std::to_string(col) + ": " + what);
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool lit(const char *w)
    {
        size_t n = 0;
        while (w[n]) ++n;
        if (s.compare(i, n, w) == 0) { i += n; return true; }
        return false;
    }
    Json value()
    {
        ws();
        if (i >= s.size()) error("unexpected end of input");
        const char c = s[i];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') return Json::string(string());
        if (lit("true")) return Json::boolean(true);
        if (lit("false")) return Json::boolean(false);
        if (lit("null")) return Json();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        error(std::string("unexpected character '") + c + "'");
    }
    Json number()
    {
        const char *begin = s.c_str() + i;
        char *end = nullptr;
        const double v = std::strtod(begin, &end);
        if (end == begin) error("bad number");
        i += size_t(end - begin);
        return Json::number(v);
    }
    std::string string()
    {
        ++i;   // opening quote
        std::string out;
        while (true) {
            if (i >= s.size()) error("unterminated string");
            char c = s[i++];
            if (c == '"') break;
            if (c == '\\') {
                if (i >= s.size()) error("bad escape");
                char e = s[i++];
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
                    if (i + 4 > s.size()) error("bad \\u escape");
                    unsigned cp = std::stoul(s.substr(i, 4), nullptr, 16);
                    i += 4;
                    if (cp < 0x80) out += char(cp);
                    else if (cp < 0x800) { out += char(0xC0 | (cp >> 6)); out += char(0x80 | (cp & 0x3F)); }
                    else { out += char(0xE0 | (cp >> 12)); out += char(0x80 | ((cp >> 6) & 0x3F)); out += char(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: error("bad escape");
                }
            } else {
                out += c;
            }
        }
        return out;
    }
    Json array()
    {
        Json j = Json::array();
        ++i;
        ws();
        if (i < s.size() && s[i] == ']') { ++i; return j; }
        while (true) {
            j.arr.push_back(value());
            ws();
            if (i < s.size() && s[i] == ',') { ++i; continue; }
            if (i < s.size() && s[i] == ']') { ++i; return j; }
            error("expected ',' or ']'");
        }
    }
    Json object()
    {
        Json j = Json::object();
        ++i;
        ws();
        if (i < s.size() && s[i] == '}') { ++i; return j; }
        while (true) {
            ws();
            if (i >= s.size() || s[i] != '"') error("expected a key");
            std::string k = string();
            ws();
            if (i >= s.size() || s[i] != ':') error("expected ':'");
            ++i;
            j.obj.emplace_back(k, value());
            ws();
            if (i < s.size() && s[i] == ',') { ++i; continue; }
            if (i < s.size() && s[i] == '}') { ++i; return j; }
            error("expected ',' or '}'");
        }
    }
};

void dump(const Json &j, std::ostringstream &o, int indent, int depth)
{
    const std::string pad(size_t(indent * (depth + 1)), ' '), end(size_t(indent * depth), ' ');
    const char *nl = indent ? "\n" : "";
    switch (j.kind) {
    case Json::Null: o << "null"; break;
    case Json::Bool: o << (j.b ? "true" : "false"); break;
    case Json::Number: {
        if (std::isfinite(j.num) && j.num == std::floor(j.num) && std::fabs(j.num) < 1e15) {
            o << (long long)j.num;
        } else {
            char buf[64];
            std::snprintf(buf, sizeof buf, "%.9g", j.num);
            o << buf;
        }
        break;
    }
    case Json::String: {
        o << '"';
        for (char c : j.str) {
            if (c == '"' || c == '\\') o << '\\' << c;
            else if (c == '\n') o << "\\n";
            else if (c == '\t') o << "\\t";
            else o << c;
        }
        o << '"';
        break;
    }
    case Json::Array:
        o << '[';
        for (size_t k = 0; k < j.arr.size(); ++k) {
            if (k) o << ", ";
            dump(j.arr[k], o, 0, 0);
        }
        o << ']';
        break;
    case Json::Object:
        o << '{' << nl;
        for (size_t k = 0; k < j.obj.size(); ++k) {
            o << pad << '"' << j.obj[k].first << "\": ";
            dump(j.obj[k].second, o, indent, depth + 1);
            if (k + 1 < j.obj.size()) o << ',';
            o << nl;
        }
        o << end << '}';
        break;
    }
}

} // namespace

Json json_parse(const std::string &text)
{
    Parser p(text);
    Json j = p.value();
    p.ws();
    if (p.i != text.size()) p.error("trailing characters");
    return j;
}

std::string json_dump(const Json &j, int indent)
{
    std::ostringstream o;
    dump(j, o, indent, 0);
    return o.str();
}

} // namespace xpr
