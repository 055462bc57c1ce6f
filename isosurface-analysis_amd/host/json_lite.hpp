// Small JSON document model for the runtime modules' configuration files and
// the run-information file (meta.json).  Objects keep their key order, so a
// file read and written back keeps the layout of the types it describes
// (the field order of the reference's serialisation traits,
// json/src/JSONHandler.cpp:17-25).  Whitespace of the output is this file's
// own.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace jl {

struct Value {
  enum Type { Null, Bool, Number, String, Array, Object } type = Null;
  bool b = false;
  double num = 0.0;
  bool integral = false;  // written without a fraction
  std::string str;
  std::vector<Value> arr;
  std::vector<std::pair<std::string, Value>> obj;

  static Value object() { Value v; v.type = Object; return v; }
  static Value array() { Value v; v.type = Array; return v; }
  static Value of(const std::string &s) { Value v; v.type = String; v.str = s; return v; }
  static Value of(const char *s) { return of(std::string(s)); }
  static Value of_int(long long i) {
    Value v; v.type = Number; v.num = (double)i; v.integral = true; return v;
  }
  static Value of_float(double d) { Value v; v.type = Number; v.num = d; return v; }

  const Value *find(const std::string &k) const {
    for (auto &kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  Value &at(const std::string &k) {
    for (auto &kv : obj)
      if (kv.first == k) return kv.second;
    throw std::runtime_error("json: missing key '" + k + "'");
  }
  const Value &at(const std::string &k) const {
    const Value *v = find(k);
    if (!v) throw std::runtime_error("json: missing key '" + k + "'");
    return *v;
  }
  Value &set(const std::string &k, Value v) {
    for (auto &kv : obj)
      if (kv.first == k) { kv.second = std::move(v); return kv.second; }
    obj.emplace_back(k, std::move(v));
    return obj.back().second;
  }
  long long as_int() const {
    if (type != Number) throw std::runtime_error("json: number expected");
    return (long long)num;
  }
  double as_num() const {
    if (type != Number) throw std::runtime_error("json: number expected");
    return num;
  }
  const std::string &as_str() const {
    if (type != String) throw std::runtime_error("json: string expected");
    return str;
  }
};

namespace detail {

struct Parser {
  const std::string &s;
  size_t i = 0;
  explicit Parser(const std::string &src) : s(src) {}

  [[noreturn]] void fail(const char *what) {
    throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(i));
  }
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
  }
  bool eat(char c) {
    ws();
    if (i < s.size() && s[i] == c) { ++i; return true; }
    return false;
  }
  void expect(char c) {
    if (!eat(c)) fail("unexpected character");
  }
  static void put_utf8(std::string &o, unsigned cp) {
    if (cp < 0x80) o += (char)cp;
    else if (cp < 0x800) { o += (char)(0xC0 | (cp >> 6)); o += (char)(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) {
      o += (char)(0xE0 | (cp >> 12)); o += (char)(0x80 | ((cp >> 6) & 0x3F));
      o += (char)(0x80 | (cp & 0x3F));
    } else {
      o += (char)(0xF0 | (cp >> 18)); o += (char)(0x80 | ((cp >> 12) & 0x3F));
      o += (char)(0x80 | ((cp >> 6) & 0x3F)); o += (char)(0x80 | (cp & 0x3F));
    }
  }
  unsigned hex4() {
    if (i + 4 > s.size()) fail("short \\u escape");
    unsigned v = 0;
    for (int k = 0; k < 4; ++k) {
      char c = s[i++];
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
      else fail("bad \\u escape");
    }
    return v;
  }
  std::string string_body() {
    std::string o;
    while (true) {
      if (i >= s.size()) fail("unterminated string");
      char c = s[i++];
      if (c == '"') return o;
      if (c != '\\') { o += c; continue; }
      if (i >= s.size()) fail("unterminated escape");
      char e = s[i++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && i + 1 < s.size() && s[i] == '\\' && s[i + 1] == 'u') {
            i += 2;
            unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(o, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
  Value value() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    char c = s[i];
    Value v;
    if (c == '{') {
      ++i;
      v.type = Value::Object;
      if (eat('}')) return v;
      do {
        ws();
        if (i >= s.size() || s[i] != '"') fail("key expected");
        ++i;
        std::string k = string_body();
        expect(':');
        v.obj.emplace_back(std::move(k), value());
      } while (eat(','));
      expect('}');
    } else if (c == '[') {
      ++i;
      v.type = Value::Array;
      if (eat(']')) return v;
      do v.arr.push_back(value());
      while (eat(','));
      expect(']');
    } else if (c == '"') {
      ++i;
      v.type = Value::String;
      v.str = string_body();
    } else if (s.compare(i, 4, "true") == 0) {
      i += 4; v.type = Value::Bool; v.b = true;
    } else if (s.compare(i, 5, "false") == 0) {
      i += 5; v.type = Value::Bool;
    } else if (s.compare(i, 4, "null") == 0) {
      i += 4;
    } else {
      size_t j = i;
      if (j < s.size() && (s[j] == '-' || s[j] == '+')) ++j;
      bool frac = false;
      while (j < s.size() && ((s[j] >= '0' && s[j] <= '9') || s[j] == '.' || s[j] == 'e' ||
                              s[j] == 'E' || s[j] == '-' || s[j] == '+')) {
        if (s[j] == '.' || s[j] == 'e' || s[j] == 'E') frac = true;
        ++j;
      }
      if (j == i) fail("value expected");
      const std::string tok = s.substr(i, j - i);
      char *end = nullptr;
      v.type = Value::Number;
      v.num = std::strtod(tok.c_str(), &end);
      if (!end || *end) fail("bad number");
      v.integral = !frac;
      i = j;
    }
    return v;
  }
};

inline void quote(std::string &o, const std::string &s) {
  o += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      default:
        if (c < 0x20) {
          char b[8];
          std::snprintf(b, sizeof b, "\\u%04x", c);
          o += b;
        } else {
          o += (char)c;
        }
    }
  }
  o += '"';
}

inline void dump(std::string &o, const Value &v, int depth) {
  const std::string pad((size_t)(depth + 1) * 4, ' '), end((size_t)depth * 4, ' ');
  switch (v.type) {
    case Value::Null: o += "null"; break;
    case Value::Bool: o += v.b ? "true" : "false"; break;
    case Value::Number: {
      char b[40];
      if (v.integral && std::fabs(v.num) < 9.0e15)
        std::snprintf(b, sizeof b, "%lld", (long long)v.num);
      else
        std::snprintf(b, sizeof b, "%.9g", v.num);  // round-trips a float exactly
      o += b;
      break;
    }
    case Value::String: quote(o, v.str); break;
    case Value::Array:
      if (v.arr.empty()) { o += "[]"; break; }
      o += "[\n";
      for (size_t k = 0; k < v.arr.size(); ++k) {
        o += pad;
        dump(o, v.arr[k], depth + 1);
        o += k + 1 < v.arr.size() ? ",\n" : "\n";
      }
      o += end + "]";
      break;
    case Value::Object:
      if (v.obj.empty()) { o += "{}"; break; }
      o += "{\n";
      for (size_t k = 0; k < v.obj.size(); ++k) {
        o += pad;
        quote(o, v.obj[k].first);
        o += ": ";
        dump(o, v.obj[k].second, depth + 1);
        o += k + 1 < v.obj.size() ? ",\n" : "\n";
      }
      o += end + "}";
      break;
  }
}

}  // namespace detail

inline Value parse(const std::string &text) {
  detail::Parser p(text);
  Value v = p.value();
  p.ws();
  if (p.i != text.size()) p.fail("trailing characters");
  return v;
}

inline std::string dump(const Value &v) {
  std::string o;
  detail::dump(o, v, 0);
  o += '\n';
  return o;
}

}  // namespace jl
