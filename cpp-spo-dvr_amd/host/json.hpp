// A small JSON reader for the run input files (objects, arrays, numbers, strings, booleans, null).
#pragma once

#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

namespace spo::json {

struct Value {
  enum class Kind { Null, Bool, Number, String, Array, Object } kind = Kind::Null;
  bool boolean = false;
  double number = 0.0;
  std::string text;  // string value, or the literal text of a number
  std::vector<Value> items;
  std::vector<std::pair<std::string, Value>> members;  // in file order

  const Value* find(const std::string& key) const;
};

// Parses a whole document; throws std::runtime_error with the byte offset on malformed input.
Value parse(const std::string& text);
Value parse_file(const std::string& path);

}  // namespace spo::json
