// Minimal JSON reader for .nls scene files.
//
// The reference parses .nls with the vendored nlohmann/json 3.10.4
// (lib/json.hpp); the loader only needs objects, arrays, strings, numbers
// and literals, read once on the host, so this is a small recursive-descent
// parser of RFC 8259 text.  Numbers are read with strtod.
#pragma once
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace lamh {

struct JsonError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class Json {
 public:
  enum Kind { Null, Bool, Number, String, Array, Object };

  Json() = default;
  Kind kind() const { return kind_; }
  bool is_number() const { return kind_ == Number; }
  bool is_string() const { return kind_ == String; }
  bool is_array() const { return kind_ == Array; }
  bool is_object() const { return kind_ == Object; }

  double num() const {
    if (kind_ != Number) throw JsonError("expected a number");
    return num_;
  }
  const std::string &str() const {
    if (kind_ != String) throw JsonError("expected a string");
    return str_;
  }
  size_t size() const {
    if (kind_ == Array) return arr_.size();
    if (kind_ == Object) return obj_.size();
    return 0;
  }
  const Json &operator[](size_t i) const {
    if (kind_ != Array) throw JsonError("expected an array");
    if (i >= arr_.size()) throw JsonError("array index out of range");
    return arr_[i];
  }
  bool has(const std::string &k) const {
    return kind_ == Object && obj_.count(k) != 0;
  }
  // like nlohmann::json::at: throws when the key is absent
  const Json &at(const std::string &k) const {
    if (kind_ != Object) throw JsonError("expected an object at key '" + k + "'");
    auto it = obj_.find(k);
    if (it == obj_.end()) throw JsonError("missing key '" + k + "'");
    return it->second;
  }

  static Json parse(const std::string &text) {
    size_t pos = 0;
    Json v = parse_value(text, pos);
    skip_ws(text, pos);
    if (pos != text.size()) throw JsonError("trailing characters after JSON value");
    return v;
  }

 private:
  Kind kind_ = Null;
  bool b_ = false;
  double num_ = 0.0;
  std::string str_;
  std::vector<Json> arr_;
  std::map<std::string, Json> obj_;

  static void skip_ws(const std::string &s, size_t &p) {
    while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) p++;
  }
  static JsonError err(const std::string &s, size_t p, const char *what) {
    size_t line = 1;
    for (size_t i = 0; i < p && i < s.size(); i++) line += (s[i] == '\n');
    return JsonError(std::string(what) + " at line " + std::to_string(line));
  }
  static std::string parse_string(const std::string &s, size_t &p) {
    if (s[p] != '"') throw err(s, p, "expected '\"'");
    p++;
    std::string out;
    while (true) {
      if (p >= s.size()) throw err(s, p, "unterminated string");
      char c = s[p++];
      if (c == '"') break;
      if (c != '\\') { out.push_back(c); continue; }
      if (p >= s.size()) throw err(s, p, "bad escape");
      char e = s[p++];
      switch (e) {
        case '"': out.push_back('"'); break;
        case '\\': out.push_back('\\'); break;
        case '/': out.push_back('/'); break;
        case 'b': out.push_back('\b'); break;
        case 'f': out.push_back('\f'); break;
        case 'n': out.push_back('\n'); break;
        case 'r': out.push_back('\r'); break;
        case 't': out.push_back('\t'); break;
        case 'u': {
          if (p + 4 > s.size()) throw err(s, p, "bad \\u escape");
          unsigned cp = std::stoul(s.substr(p, 4), nullptr, 16);
          p += 4;
          // UTF-8 encode (BMP only; names in scene files are ASCII)
          if (cp < 0x80) out.push_back(char(cp));
          else if (cp < 0x800) { out.push_back(char(0xC0 | (cp >> 6))); out.push_back(char(0x80 | (cp & 0x3F))); }
          else { out.push_back(char(0xE0 | (cp >> 12))); out.push_back(char(0x80 | ((cp >> 6) & 0x3F))); out.push_back(char(0x80 | (cp & 0x3F))); }
          break;
        }
        default: throw err(s, p, "bad escape");
      }
    }
    return out;
  }
  static Json parse_value(const std::string &s, size_t &p) {
    skip_ws(s, p);
    if (p >= s.size()) throw err(s, p, "unexpected end of input");
    Json v;
    char c = s[p];
    if (c == '{') {
      v.kind_ = Object;
      p++;
      skip_ws(s, p);
      if (p < s.size() && s[p] == '}') { p++; return v; }
      while (true) {
        skip_ws(s, p);
        if (p >= s.size()) throw err(s, p, "unterminated object");
        std::string key = parse_string(s, p);
        skip_ws(s, p);
        if (p >= s.size() || s[p] != ':') throw err(s, p, "expected ':'");
        p++;
        v.obj_[key] = parse_value(s, p);
        skip_ws(s, p);
        if (p < s.size() && s[p] == ',') { p++; continue; }
        if (p < s.size() && s[p] == '}') { p++; break; }
        throw err(s, p, "expected ',' or '}'");
      }
    } else if (c == '[') {
      v.kind_ = Array;
      p++;
      skip_ws(s, p);
      if (p < s.size() && s[p] == ']') { p++; return v; }
      while (true) {
        v.arr_.push_back(parse_value(s, p));
        skip_ws(s, p);
        if (p < s.size() && s[p] == ',') { p++; continue; }
        if (p < s.size() && s[p] == ']') { p++; break; }
        throw err(s, p, "expected ',' or ']'");
      }
    } else if (c == '"') {
      v.kind_ = String;
      v.str_ = parse_string(s, p);
    } else if (s.compare(p, 4, "true") == 0) {
      v.kind_ = Bool; v.b_ = true; p += 4;
    } else if (s.compare(p, 5, "false") == 0) {
      v.kind_ = Bool; v.b_ = false; p += 5;
    } else if (s.compare(p, 4, "null") == 0) {
      v.kind_ = Null; p += 4;
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      const char *begin = s.c_str() + p;
      char *end = nullptr;
      v.kind_ = Number;
      v.num_ = std::strtod(begin, &end);
      if (end == begin) throw err(s, p, "bad number");
      p += size_t(end - begin);
    } else {
      throw err(s, p, "unexpected character");
    }
    return v;
  }
};

}  // namespace lamh
