// A small JSON reader for scene files: objects, arrays, strings, numbers,
// true/false/null.  Numbers keep the three kinds a scene file's readers
// distinguish: an unsigned integer (no sign, no fraction or exponent), a
// signed integer (leading minus), or a double parsed with strtod.  Reading a
// number as float converts that value once (integer -> float, or
// double -> float); a non-finite double becomes null.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fluidapp {
namespace json {

struct Value {
  enum class Kind { Null, Bool, Unsigned, Signed, Float, String, Array, Object };
  Kind kind = Kind::Null;
  bool b = false;
  uint64_t u = 0;
  int64_t i = 0;
  double d = 0.0;
  std::string s;
  std::vector<Value> items;                             // Array
  std::vector<std::pair<std::string, Value>> members;  // Object, in file order

  bool is_object() const { return kind == Kind::Object; }
  bool is_array() const { return kind == Kind::Array; }
  bool is_string() const { return kind == Kind::String; }
  bool is_number() const {
    return kind == Kind::Unsigned || kind == Kind::Signed || kind == Kind::Float;
  }
  // Member by key (the last one when a key repeats), or nullptr.
  const Value* find(const std::string& key) const {
    if (kind != Kind::Object) return nullptr;
    const Value* hit = nullptr;
    for (const auto& m : members)
      if (m.first == key) hit = &m.second;
    return hit;
  }
  bool has(const std::string& key) const { return find(key) != nullptr; }
  const Value& at(const std::string& key) const {
    const Value* v = find(key);
    if (!v) throw std::runtime_error("missing key '" + key + "'");
    return *v;
  }
  float as_float() const {
    switch (kind) {
      case Kind::Unsigned: return static_cast<float>(u);
      case Kind::Signed: return static_cast<float>(i);
      case Kind::Float: return static_cast<float>(d);
      default: throw std::runtime_error("type must be number");
    }
  }
  double as_double() const {
    switch (kind) {
      case Kind::Unsigned: return static_cast<double>(u);
      case Kind::Signed: return static_cast<double>(i);
      case Kind::Float: return d;
      default: throw std::runtime_error("type must be number");
    }
  }
  int as_int() const {
    switch (kind) {
      case Kind::Unsigned: return static_cast<int>(u);
      case Kind::Signed: return static_cast<int>(i);
      case Kind::Float: return static_cast<int>(d);
      default: throw std::runtime_error("type must be number");
    }
  }
  bool as_bool() const {
    if (kind != Kind::Bool) throw std::runtime_error("type must be boolean");
    return b;
  }
  const std::string& as_string() const {
    if (kind != Kind::String) throw std::runtime_error("type must be string");
    return s;
  }
};

class Parser {
 public:
  explicit Parser(const std::string& text) : t_(text) {}

  Value parse() {
    Value v = value();
    ws();
    if (p_ != t_.size()) fail("trailing characters");
    return v;
  }

 private:
  const std::string& t_;
  size_t p_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    size_t line = 1, col = 1;
    for (size_t k = 0; k < p_ && k < t_.size(); ++k) {
      if (t_[k] == '\n') {
        ++line;
        col = 1;
      } else {
        ++col;
      }
    }
    throw std::runtime_error(what + " at line " + std::to_string(line) + ", column " +
                             std::to_string(col));
  }
  void ws() {
    while (p_ < t_.size() &&
           (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r'))
      ++p_;
  }
  bool lit(const char* w) {
    size_t k = 0;
    while (w[k] && p_ + k < t_.size() && t_[p_ + k] == w[k]) ++k;
    if (w[k]) return false;
    p_ += k;
    return true;
  }
  Value value() {
    ws();
    if (p_ >= t_.size()) fail("unexpected end of input");
    const char c = t_[p_];
    Value v;
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') {
      v.kind = Value::Kind::String;
      v.s = string();
      return v;
    }
    if (lit("true")) {
      v.kind = Value::Kind::Bool;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.kind = Value::Kind::Bool;
      return v;
    }
    if (lit("null")) return v;
    if (c == '-' || (c >= '0' && c <= '9')) return number();
    fail(std::string("unexpected character '") + c + "'");
  }
  Value object() {
    Value v;
    v.kind = Value::Kind::Object;
    ++p_;
    ws();
    if (p_ < t_.size() && t_[p_] == '}') {
      ++p_;
      return v;
    }
    while (true) {
      ws();
      if (p_ >= t_.size() || t_[p_] != '"') fail("expected a key");
      std::string key = string();
      ws();
      if (p_ >= t_.size() || t_[p_] != ':') fail("expected ':'");
      ++p_;
      v.members.emplace_back(std::move(key), value());
      ws();
      if (p_ < t_.size() && t_[p_] == ',') {
        ++p_;
        continue;
      }
      if (p_ < t_.size() && t_[p_] == '}') {
        ++p_;
        return v;
      }
      fail("expected ',' or '}'");
    }
  }
  Value array() {
    Value v;
    v.kind = Value::Kind::Array;
    ++p_;
    ws();
    if (p_ < t_.size() && t_[p_] == ']') {
      ++p_;
      return v;
    }
    while (true) {
      v.items.push_back(value());
      ws();
      if (p_ < t_.size() && t_[p_] == ',') {
        ++p_;
        continue;
      }
      if (p_ < t_.size() && t_[p_] == ']') {
        ++p_;
        return v;
      }
      fail("expected ',' or ']'");
    }
  }
  static void put_utf8(std::string& out, unsigned cp) {
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
  }
  unsigned hex4() {
    if (p_ + 4 > t_.size()) fail("short \\u escape");
    unsigned cp = 0;
    for (int k = 0; k < 4; ++k) {
      const char h = t_[p_++];
      cp <<= 4;
      if (h >= '0' && h <= '9') cp |= static_cast<unsigned>(h - '0');
      else if (h >= 'a' && h <= 'f') cp |= static_cast<unsigned>(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') cp |= static_cast<unsigned>(h - 'A' + 10);
      else fail("bad \\u escape");
    }
    return cp;
  }
  std::string string() {
    ++p_;  // opening quote
    std::string out;
    while (true) {
      if (p_ >= t_.size()) fail("unterminated string");
      const char c = t_[p_++];
      if (c == '"') return out;
      if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= t_.size()) fail("unterminated escape");
      const char e = t_[p_++];
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
          unsigned cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00 && p_ + 1 < t_.size() && t_[p_] == '\\' &&
              t_[p_ + 1] == 'u') {
            p_ += 2;
            const unsigned lo = hex4();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
  }
  Value number() {
    const size_t b = p_;
    bool neg = false, frac = false;
    if (t_[p_] == '-') {
      neg = true;
      ++p_;
    }
    auto digits = [&]() {
      const size_t s = p_;
      while (p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9') ++p_;
      return p_ > s;
    };
    if (!digits()) fail("bad number");
    if (p_ < t_.size() && t_[p_] == '.') {
      frac = true;
      ++p_;
      if (!digits()) fail("bad number");
    }
    if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
      frac = true;
      ++p_;
      if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) ++p_;
      if (!digits()) fail("bad number");
    }
    const std::string tok = t_.substr(b, p_ - b);
    Value v;
    if (!frac) {
      errno = 0;
      char* end = nullptr;
      if (neg) {
        const long long x = std::strtoll(tok.c_str(), &end, 10);
        if (errno == 0 && end && *end == '\0') {
          v.kind = Value::Kind::Signed;
          v.i = x;
          return v;
        }
      } else {
        const unsigned long long x = std::strtoull(tok.c_str(), &end, 10);
        if (errno == 0 && end && *end == '\0') {
          v.kind = Value::Kind::Unsigned;
          v.u = x;
          return v;
        }
      }
    }
    // fractions, exponents and integers out of range: a double
    const double x = std::strtod(tok.c_str(), nullptr);
    if (!std::isfinite(x)) return v;  // null
    v.kind = Value::Kind::Float;
    v.d = x;
    return v;
  }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace json
}  // namespace fluidapp
