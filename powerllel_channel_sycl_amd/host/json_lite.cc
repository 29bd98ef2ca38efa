#include "json_lite.hh"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace json_lite {
namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;

  [[noreturn]] void fail(const char* what) {
    throw ParseError(std::string("JSON: ") + what + " at offset " + std::to_string(i));
  }
  void ws() {
    while (i < s.size() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
  }
  bool lit(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s.compare(i, n, w) == 0) {
      i += n;
      return true;
    }
    return false;
  }
  std::string string() {
    if (s[i] != '"') fail("expected a string");
    ++i;
    std::string out;
    while (i < s.size() && s[i] != '"') {
      char c = s[i++];
      if (c == '\\') {
        if (i >= s.size()) fail("unterminated escape");
        char e = s[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i + 4 > s.size()) fail("short \\u escape");
            unsigned cp = std::strtoul(s.substr(i, 4).c_str(), nullptr, 16);
            i += 4;
            if (cp < 0x80) {
              out += static_cast<char>(cp);
            } else if (cp < 0x800) {
              out += static_cast<char>(0xC0 | (cp >> 6));
              out += static_cast<char>(0x80 | (cp & 0x3F));
            } else {
              out += static_cast<char>(0xE0 | (cp >> 12));
              out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
              out += static_cast<char>(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= s.size()) fail("unterminated string");
    ++i;
    return out;
  }
  Value value() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    Value v;
    char c = s[i];
    if (c == '{') {
      v.kind = Value::Object;
      ++i;
      ws();
      if (s[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        ws();
        std::string k = string();
        ws();
        if (s[i] != ':') fail("expected ':'");
        ++i;
        v.obj[k] = value();
        ws();
        if (s[i] == ',') {
          ++i;
          continue;
        }
        if (s[i] == '}') {
          ++i;
          return v;
        }
        fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Value::Array;
      ++i;
      ws();
      if (s[i] == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.arr.push_back(value());
        ws();
        if (s[i] == ',') {
          ++i;
          continue;
        }
        if (s[i] == ']') {
          ++i;
          return v;
        }
        fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = Value::String;
      v.str = string();
      return v;
    }
    if (lit("true")) {
      v.kind = Value::Bool;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.kind = Value::Bool;
      return v;
    }
    if (lit("null")) return v;
    const char* start = s.c_str() + i;
    char* end = nullptr;
    v.num = std::strtod(start, &end);
    if (end == start) fail("unexpected character");
    std::string tok(start, static_cast<size_t>(end - start));
    v.integral = tok.find_first_of(".eE") == std::string::npos;
    v.kind = Value::Number;
    i += static_cast<size_t>(end - start);
    return v;
  }
};

}  // namespace

Value parse(const std::string& text) {
  Parser p{text};
  Value v = p.value();
  p.ws();
  if (p.i != text.size()) p.fail("trailing characters");
  return v;
}

Value parse_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw ParseError("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

}  // namespace json_lite
