// json_lite.hh -- a small JSON reader for the run-parameter file (the
// reference reads it with RapidJSON, src/mod_parameters.cc:1-10, which this
// image does not carry).  Objects, arrays, strings, numbers, true/false/null.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace json_lite {

struct Value {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  bool b = false;
  double num = 0.0;
  bool integral = false;  // the literal had no fraction or exponent
  std::string str;
  std::vector<Value> arr;
  std::map<std::string, Value> obj;

  const Value* find(const std::string& key) const {
    if (kind != Object) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
  }
};

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

Value parse(const std::string& text);
Value parse_file(const std::string& path);

}  // namespace json_lite
