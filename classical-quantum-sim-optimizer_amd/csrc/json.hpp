// Minimal JSON reader for the circuit transport.
//
// The transport is the serde schema of src/types.rs:7-296 (qregs / cregs /
// instructions tagged by "kind"). Integer literals are kept exactly as
// 64-bit values so that usize operands and u64 condition values round-trip
// without passing through double; non-integer literals are parsed with
// strtod, which is correctly rounded.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dq {

struct JsonError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct JVal {
  enum Type { NUL, BOOL, NUM, STR, ARR, OBJ } type = NUL;
  bool b = false;
  // Numbers: is_int marks a literal without fraction/exponent.
  bool is_int = false, neg = false, overflow = false;
  uint64_t u = 0;  // magnitude of an integer literal
  double d = 0.0;
  std::string s;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;  // insertion order kept

  const JVal* get(const char* key) const {
    for (auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class JsonParser {
 public:
  JsonParser(const char* text, size_t len) : p_(text), end_(text + len) {}

  JVal parse() {
    JVal v = value(0);
    ws();
    if (p_ != end_) fail("trailing characters");
    return v;
  }

 private:
  const char* p_;
  const char* end_;

  [[noreturn]] void fail(const char* what) {
    throw JsonError(std::string("Circuit JSON parse error: ") + what);
  }
  void ws() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\r' || *p_ == '\t')) ++p_;
  }
  bool lit(const char* w) {
    size_t n = std::strlen(w);
    if ((size_t)(end_ - p_) >= n && std::memcmp(p_, w, n) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  JVal value(int depth) {
    if (depth > 256) fail("nesting too deep");
    ws();
    if (p_ >= end_) fail("unexpected end of input");
    JVal v;
    char c = *p_;
    if (c == '{') {
      v.type = JVal::OBJ;
      ++p_;
      ws();
      if (p_ < end_ && *p_ == '}') {
        ++p_;
        return v;
      }
      for (;;) {
        ws();
        if (p_ >= end_ || *p_ != '"') fail("expected object key");
        std::string key = str();
        ws();
        if (p_ >= end_ || *p_ != ':') fail("expected ':'");
        ++p_;
        v.obj.emplace_back(std::move(key), value(depth + 1));
        ws();
        if (p_ < end_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < end_ && *p_ == '}') {
          ++p_;
          break;
        }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      v.type = JVal::ARR;
      ++p_;
      ws();
      if (p_ < end_ && *p_ == ']') {
        ++p_;
        return v;
      }
      for (;;) {
        v.arr.push_back(value(depth + 1));
        ws();
        if (p_ < end_ && *p_ == ',') {
          ++p_;
          continue;
        }
        if (p_ < end_ && *p_ == ']') {
          ++p_;
          break;
        }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.type = JVal::STR;
      v.s = str();
    } else if (lit("true")) {
      v.type = JVal::BOOL;
      v.b = true;
    } else if (lit("false")) {
      v.type = JVal::BOOL;
      v.b = false;
    } else if (lit("null")) {
      v.type = JVal::NUL;
    } else if (lit("NaN") || lit("Infinity")) {
      // Python's json.dumps emits these for non-finite floats; serde_json
      // rejects them, and so does this reader.
      fail("non-finite number literal");
    } else if (c == '-' || (c >= '0' && c <= '9')) {
      num(v);
    } else {
      fail("unexpected character");
    }
    return v;
  }
  static int hexval(char c) {
    if (c >= '0' && c <= '9') return c - '0';
    if (c >= 'a' && c <= 'f') return c - 'a' + 10;
    if (c >= 'A' && c <= 'F') return c - 'A' + 10;
    return -1;
  }
  void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
      out += (char)cp;
    } else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  uint32_t hex4() {
    if (end_ - p_ < 4) fail("bad unicode escape");
    uint32_t cp = 0;
    for (int i = 0; i < 4; ++i) {
      int h = hexval(p_[i]);
      if (h < 0) fail("bad unicode escape");
      cp = cp * 16 + (uint32_t)h;
    }
    p_ += 4;
    return cp;
  }
  std::string str() {
    ++p_;  // opening quote
    std::string out;
    while (p_ < end_ && *p_ != '"') {
      char c = *p_++;
      if ((unsigned char)c < 0x20) fail("control character in string");
      if (c != '\\') {
        out += c;
        continue;
      }
      if (p_ >= end_) fail("bad escape");
      char e = *p_++;
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
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
              p_ += 2;
              uint32_t lo = hex4();
              if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
              cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
            } else {
              fail("lone surrogate");
            }
          }
          utf8(out, cp);
          break;
        }
        default: fail("bad escape");
      }
    }
    if (p_ >= end_) fail("unterminated string");
    ++p_;
    return out;
  }
  void num(JVal& v) {
    v.type = JVal::NUM;
    const char* start = p_;
    if (*p_ == '-') {
      v.neg = true;
      ++p_;
    }
    if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) fail("bad number");
    if (*p_ == '0' && p_ + 1 < end_ && p_[1] >= '0' && p_[1] <= '9') fail("leading zero");
    bool is_int = true;
    uint64_t mag = 0;
    bool over = false;
    while (p_ < end_ && *p_ >= '0' && *p_ <= '9') {
      uint64_t digit = (uint64_t)(*p_ - '0');
      if (mag > (UINT64_MAX - digit) / 10) over = true;
      else mag = mag * 10 + digit;
      ++p_;
    }
    if (p_ < end_ && *p_ == '.') {
      is_int = false;
      ++p_;
      if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) fail("bad fraction");
      while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
    }
    if (p_ < end_ && (*p_ == 'e' || *p_ == 'E')) {
      is_int = false;
      ++p_;
      if (p_ < end_ && (*p_ == '+' || *p_ == '-')) ++p_;
      if (p_ >= end_ || !(*p_ >= '0' && *p_ <= '9')) fail("bad exponent");
      while (p_ < end_ && *p_ >= '0' && *p_ <= '9') ++p_;
    }
    std::string text(start, p_);
    v.d = std::strtod(text.c_str(), nullptr);
    v.is_int = is_int;
    v.u = mag;
    v.overflow = over;
  }
};

}  // namespace dq
