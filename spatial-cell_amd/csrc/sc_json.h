// Minimal JSON reader for the config and setup documents (the reference uses
// jsoncpp: config.cuh:46-63, chemistry.cuh:14-28, interactions.cuh:21-50).
// Value semantics follow jsoncpp where the reference depends on them: a key
// that is present with value null reads as 0 / false through as_int/as_bool,
// and get(key, default) returns the default only when the key is absent.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace scj {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  bool has(const std::string& k) const {
    for (auto& kv : obj) if (kv.first == k) return true;
    return false;
  }
  const Value* find(const std::string& k) const {
    for (auto& kv : obj) if (kv.first == k) return &kv.second;
    return nullptr;
  }
  double as_double() const { return kind == Number ? num : (kind == Bool ? (b ? 1.0 : 0.0) : 0.0); }
  int as_int() const { return (int)as_double(); }
  float as_float() const { return (float)as_double(); }
  bool as_bool() const { return kind == Bool ? b : (kind == Number ? num != 0.0 : false); }
  double get_double(const std::string& k, double d) const { const Value* v = find(k); return v ? v->as_double() : d; }
  int get_int(const std::string& k, int d) const { const Value* v = find(k); return v ? v->as_int() : d; }
  float get_float(const std::string& k, float d) const { const Value* v = find(k); return v ? v->as_float() : d; }
  bool get_bool(const std::string& k, bool d) const { const Value* v = find(k); return v ? v->as_bool() : d; }
  std::string get_str(const std::string& k, const std::string& d) const {
    const Value* v = find(k);
    return (v && v->kind == String) ? v->str : d;
  }
  size_t size() const { return kind == Array ? arr.size() : (kind == Object ? obj.size() : 0); }
  const Value& operator[](size_t i) const { return arr[i]; }
};

// Parses text; on error returns false and sets err to "offset N: message".
bool parse(const std::string& text, Value& out, std::string& err);
bool read_file(const std::string& path, std::string& out);

}  // namespace scj
