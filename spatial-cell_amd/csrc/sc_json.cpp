#include "sc_json.h"

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace scj {
namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;
  std::string err;
  explicit Parser(const std::string& t) : s(t) {}

  void ws() { while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++; }
  bool fail(const char* m) { if (err.empty()) err = "offset " + std::to_string(i) + ": " + m; return false; }

  bool lit(const char* w) {
    size_t n = std::strlen(w);
    if (s.compare(i, n, w) != 0) return fail("bad literal");
    i += n;
    return true;
  }
  bool str(std::string& out) {
    if (i >= s.size() || s[i] != '"') return fail("expected string");
    i++;
    out.clear();
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\') {
        if (i >= s.size()) return fail("bad escape");
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
            if (i + 4 > s.size()) return fail("bad \\u escape");
            unsigned cp = (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
            i += 4;
            if (cp < 0x80) out += (char)cp;
            else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
            else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
            break;
          }
          default: return fail("bad escape");
        }
      } else {
        out += c;
      }
    }
    if (i >= s.size()) return fail("unterminated string");
    i++;
    return true;
  }
  bool value(Value& v) {
    ws();
    if (i >= s.size()) return fail("unexpected end");
    char c = s[i];
    if (c == '{') {
      v.kind = Value::Object;
      i++;
      ws();
      if (i < s.size() && s[i] == '}') { i++; return true; }
      for (;;) {
        ws();
        std::string k;
        if (!str(k)) return false;
        ws();
        if (i >= s.size() || s[i] != ':') return fail("expected ':'");
        i++;
        Value child;
        if (!value(child)) return false;
        v.obj.emplace_back(std::move(k), std::move(child));
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == '}') { i++; return true; }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Value::Array;
      i++;
      ws();
      if (i < s.size() && s[i] == ']') { i++; return true; }
      for (;;) {
        Value child;
        if (!value(child)) return false;
        v.arr.push_back(std::move(child));
        ws();
        if (i < s.size() && s[i] == ',') { i++; continue; }
        if (i < s.size() && s[i] == ']') { i++; return true; }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') { v.kind = Value::String; return str(v.str); }
    if (c == 't') { v.kind = Value::Bool; v.b = true; return lit("true"); }
    if (c == 'f') { v.kind = Value::Bool; v.b = false; return lit("false"); }
    if (c == 'n') { v.kind = Value::Null; return lit("null"); }
    if (c == '-' || (c >= '0' && c <= '9')) {
      const char* start = s.c_str() + i;
      char* end = nullptr;
      double d = std::strtod(start, &end);
      if (end == start) return fail("bad number");
      i += (size_t)(end - start);
      v.kind = Value::Number;
      v.num = d;
      return true;
    }
    return fail("unexpected character");
  }
};

}  // namespace

bool parse(const std::string& text, Value& out, std::string& err) {
  Parser p(text);
  out = Value();
  if (!p.value(out)) { err = p.err; return false; }
  p.ws();
  if (p.i != text.size()) { err = "offset " + std::to_string(p.i) + ": trailing characters"; return false; }
  return true;
}

bool read_file(const std::string& path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

}  // namespace scj
