// Minimal YAML reader for the subset that Cantera mechanism files use: block mappings and sequences
// (including "- key: value" items and sequences at their parent key's indentation), flow sequences and
// mappings that may span lines, quoted scalars, comments and block scalars (|, >). Anchors, tags and
// multi-document streams are not supported; a file using them is rejected with a line number.
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace combustion::yaml {

struct Node {
  enum Kind { Null, Scalar, Sequence, Mapping } kind = Null;
  std::string scalar;
  std::vector<Node> items;                          // Sequence
  std::vector<std::pair<std::string, Node>> pairs;  // Mapping, in file order
  int line = 0;

  bool is_null() const { return kind == Null; }
  // Mapping lookup; returns a Null node when the key is absent (like yaml-rust's BadValue).
  const Node& operator[](const std::string& key) const;
  const Node& operator[](size_t i) const;
  size_t size() const { return kind == Sequence ? items.size() : kind == Mapping ? pairs.size() : 0; }
  const std::string& str() const;  // throws unless Scalar
  double f64() const;              // throws unless a numeric scalar
  long long i64() const;
};

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

Node parse(const std::string& text);

}  // namespace combustion::yaml
