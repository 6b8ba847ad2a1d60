#include "yaml.hpp"

#include <cerrno>
#include <cstdlib>

namespace combustion::yaml {

static const Node null_node{};

const Node& Node::operator[](const std::string& key) const {
  if (kind != Mapping) return null_node;
  for (auto& [k, v] : pairs)
    if (k == key) return v;
  return null_node;
}
const Node& Node::operator[](size_t i) const {
  if (kind != Sequence || i >= items.size()) return null_node;
  return items[i];
}
const std::string& Node::str() const {
  if (kind != Scalar) throw Error("line " + std::to_string(line) + ": expected a scalar");
  return scalar;
}
double Node::f64() const {
  const std::string& s = str();
  if (s == ".inf" || s == ".Inf" || s == "inf") return 1.0 / 0.0;
  char* end = nullptr;
  errno = 0;
  double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end != 0) throw Error("line " + std::to_string(line) + ": '" + s + "' is not a number");
  return v;
}
long long Node::i64() const {
  double v = f64();
  long long i = (long long)v;
  if ((double)i != v) throw Error("line " + std::to_string(line) + ": '" + scalar + "' is not an integer");
  return i;
}

namespace {

struct Line {
  int indent;
  std::string text;  // content after indentation, comment stripped, right-trimmed
  int number;        // 1-based line number
  size_t raw;        // index into raw lines
};

std::string rtrim(std::string s) {
  while (!s.empty() && (s.back() == ' ' || s.back() == '\t' || s.back() == '\r')) s.pop_back();
  return s;
}
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t')) a++;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) b--;
  return s.substr(a, b - a);
}

// Removes a trailing comment: '#' at the start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); i++) {
    char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
    } else if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return rtrim(s.substr(0, i));
    }
  }
  return rtrim(s);
}

class Parser {
 public:
  explicit Parser(const std::string& text) {
    size_t start = 0;
    while (start <= text.size()) {
      size_t end = text.find('\n', start);
      if (end == std::string::npos) end = text.size();
      raw_.push_back(text.substr(start, end - start));
      start = end + 1;
    }
    for (size_t r = 0; r < raw_.size(); r++) {
      const std::string& s = raw_[r];
      int indent = 0;
      while (indent < (int)s.size() && s[indent] == ' ') indent++;
      if (indent < (int)s.size() && s[indent] == '\t') fail((int)r + 1, "tab indentation");
      std::string t = strip_comment(s.substr(indent));
      if (t.empty()) continue;
      if (t == "---" && lines_.empty()) continue;
      if (t == "---" || t == "...") fail((int)r + 1, "multi-document streams are not supported");
      lines_.push_back({indent, t, (int)r + 1, r});
    }
  }

  Node document() {
    if (lines_.empty()) return Node{};
    Node n = block(lines_[0].indent);
    if (pos_ < lines_.size()) fail(lines_[pos_].number, "unexpected content");
    return n;
  }

 private:
  std::vector<std::string> raw_;
  std::vector<Line> lines_;
  size_t pos_ = 0;

  [[noreturn]] static void fail(int line, const std::string& what) { throw Error("line " + std::to_string(line) + ": " + what); }

  static bool is_seq_item(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

  // Position of the ':' that ends a mapping key on this line, or npos.
  static size_t key_colon(const std::string& t) {
    if (t.empty() || t[0] == '[' || t[0] == '{' || is_seq_item(t)) return std::string::npos;
    char quote = 0;
    for (size_t i = 0; i < t.size(); i++) {
      char c = t[i];
      if (quote) {
        if (c == quote) quote = 0;
        continue;
      }
      if ((c == '"' || c == '\'') && i == 0) {
        quote = c;
        continue;
      }
      if (c == ':' && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
    }
    return std::string::npos;
  }

  static std::string unquote(const std::string& s) {
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
      return s.substr(1, s.size() - 2);
    return s;
  }

  static Node make_scalar(const std::string& s, int line) {
    Node n;
    n.line = line;
    std::string t = trim(s);
    if (t.empty() || t == "~" || t == "null") return n;
    if (t[0] == '&' || t[0] == '*' || t[0] == '!') fail(line, "anchors, aliases and tags are not supported");
    n.kind = Node::Scalar;
    n.scalar = unquote(t);
    return n;
  }

  Node block(int indent) {
    const Line& l = lines_[pos_];
    if (l.indent != indent) fail(l.number, "bad indentation");
    if (is_seq_item(l.text)) return sequence(indent);
    if (key_colon(l.text) != std::string::npos) return mapping(indent);
    // A bare scalar or flow collection on its own line.
    int number = l.number;
    std::string text = l.text;
    pos_++;
    return inline_value(text, indent, number);
  }

  Node sequence(int indent) {
    Node seq;
    seq.kind = Node::Sequence;
    seq.line = lines_[pos_].number;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent && is_seq_item(lines_[pos_].text)) {
      Line& l = lines_[pos_];
      if (l.text == "-") {
        pos_++;
        if (pos_ < lines_.size() && lines_[pos_].indent > indent)
          seq.items.push_back(block(lines_[pos_].indent));
        else
          seq.items.push_back(Node{});
        continue;
      }
      size_t spaces = 1;
      while (spaces < l.text.size() && l.text[spaces] == ' ') spaces++;
      std::string rest = l.text.substr(spaces);
      if (key_colon(rest) != std::string::npos || is_seq_item(rest)) {
        // "- key: value" opens a mapping whose keys align with "key".
        l.indent = indent + (int)spaces;
        l.text = rest;
        seq.items.push_back(block(l.indent));
      } else {
        int number = l.number;
        pos_++;
        seq.items.push_back(inline_value(rest, indent, number));
      }
    }
    return seq;
  }

  Node mapping(int indent) {
    Node map;
    map.kind = Node::Mapping;
    map.line = lines_[pos_].number;
    while (pos_ < lines_.size() && lines_[pos_].indent == indent) {
      const Line& l = lines_[pos_];
      size_t c = key_colon(l.text);
      if (c == std::string::npos) fail(l.number, "expected 'key: value'");
      std::string key = unquote(trim(l.text.substr(0, c)));
      std::string rest = trim(l.text.substr(c + 1));
      int number = l.number;
      size_t raw_index = l.raw;
      pos_++;
      Node value;
      if (rest.empty()) {
        if (pos_ < lines_.size() &&
            (lines_[pos_].indent > indent || (lines_[pos_].indent == indent && is_seq_item(lines_[pos_].text))))
          value = block(lines_[pos_].indent);
        value.line = value.line ? value.line : number;
      } else if (rest[0] == '|' || rest[0] == '>') {
        value = block_scalar(rest[0] == '>', indent, raw_index, number);
      } else {
        value = inline_value(rest, indent, number);
      }
      for (auto& [k, v] : map.pairs)
        if (k == key) fail(number, "duplicate key '" + key + "'");
      map.pairs.emplace_back(key, std::move(value));
    }
    return map;
  }

  Node block_scalar(bool folded, int indent, size_t key_raw, int number) {
    Node n;
    n.kind = Node::Scalar;
    n.line = number;
    size_t r = key_raw + 1;
    std::string text;
    for (; r < raw_.size(); r++) {
      const std::string& s = raw_[r];
      int ind = 0;
      while (ind < (int)s.size() && s[ind] == ' ') ind++;
      bool blank = trim(s).empty();
      if (!blank && ind <= indent) break;
      if (!text.empty()) text += folded ? " " : "\n";
      text += trim(s);
    }
    while (pos_ < lines_.size() && lines_[pos_].raw < r) pos_++;
    n.scalar = text;
    return n;
  }

  // A value that starts on the current line: a flow collection (possibly continued on following,
  // more indented lines) or a plain scalar (possibly continued the same way).
  Node inline_value(std::string text, int parent_indent, int number) {
    if (text[0] == '[' || text[0] == '{') {
      while (!balanced(text)) {
        if (pos_ >= lines_.size()) fail(number, "unterminated flow collection");
        text += " " + lines_[pos_].text;
        pos_++;
      }
      size_t i = 0;
      Node n = flow(text, i, number);
      skip_spaces(text, i);
      if (i != text.size()) fail(number, "trailing characters after flow collection");
      return n;
    }
    while (pos_ < lines_.size() && lines_[pos_].indent > parent_indent && key_colon(lines_[pos_].text) == std::string::npos &&
           !is_seq_item(lines_[pos_].text)) {
      text += " " + lines_[pos_].text;
      pos_++;
    }
    return make_scalar(text, number);
  }

  static bool balanced(const std::string& s) {
    int depth = 0;
    char quote = 0;
    for (char c : s) {
      if (quote) {
        if (c == quote) quote = 0;
      } else if (c == '"' || c == '\'') {
        quote = c;
      } else if (c == '[' || c == '{') {
        depth++;
      } else if (c == ']' || c == '}') {
        depth--;
      }
    }
    return depth <= 0;
  }

  static void skip_spaces(const std::string& s, size_t& i) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) i++;
  }

  // Scalar inside a flow collection: ends at ',' or the closing bracket, or at ':' for keys.
  static std::string flow_scalar(const std::string& s, size_t& i, bool key, int number) {
    skip_spaces(s, i);
    if (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
      char q = s[i];
      size_t j = s.find(q, i + 1);
      if (j == std::string::npos) fail(number, "unterminated quote");
      std::string out = s.substr(i + 1, j - i - 1);
      i = j + 1;
      return out;
    }
    size_t start = i;
    while (i < s.size()) {
      char c = s[i];
      if (c == ',' || c == ']' || c == '}') break;
      if (key && c == ':' && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == ',' || s[i + 1] == '}')) break;
      i++;
    }
    return trim(s.substr(start, i - start));
  }

  Node flow(const std::string& s, size_t& i, int number) {
    skip_spaces(s, i);
    Node n;
    n.line = number;
    if (i < s.size() && s[i] == '[') {
      n.kind = Node::Sequence;
      i++;
      skip_spaces(s, i);
      if (i < s.size() && s[i] == ']') {
        i++;
        return n;
      }
      while (true) {
        skip_spaces(s, i);
        if (i < s.size() && (s[i] == '[' || s[i] == '{'))
          n.items.push_back(flow(s, i, number));
        else
          n.items.push_back(make_scalar(flow_scalar(s, i, false, number), number));
        skip_spaces(s, i);
        if (i >= s.size()) fail(number, "unterminated flow sequence");
        if (s[i] == ',') {
          i++;
          skip_spaces(s, i);
          if (i < s.size() && s[i] == ']') {  // trailing comma
            i++;
            return n;
          }
          continue;
        }
        if (s[i] == ']') {
          i++;
          return n;
        }
        fail(number, "expected ',' or ']'");
      }
    }
    if (i < s.size() && s[i] == '{') {
      n.kind = Node::Mapping;
      i++;
      skip_spaces(s, i);
      if (i < s.size() && s[i] == '}') {
        i++;
        return n;
      }
      while (true) {
        std::string key = flow_scalar(s, i, true, number);
        skip_spaces(s, i);
        Node value;
        if (i < s.size() && s[i] == ':') {
          i++;
          skip_spaces(s, i);
          if (i < s.size() && (s[i] == '[' || s[i] == '{'))
            value = flow(s, i, number);
          else
            value = make_scalar(flow_scalar(s, i, false, number), number);
        }
        n.pairs.emplace_back(key, std::move(value));
        skip_spaces(s, i);
        if (i >= s.size()) fail(number, "unterminated flow mapping");
        if (s[i] == ',') {
          i++;
          skip_spaces(s, i);
          if (i < s.size() && s[i] == '}') {
            i++;
            return n;
          }
          continue;
        }
        if (s[i] == '}') {
          i++;
          return n;
        }
        fail(number, "expected ',' or '}'");
      }
    }
    fail(number, "expected a flow collection");
  }
};

}  // namespace

Node parse(const std::string& text) {
  Parser p(text);
  return p.document();
}

}  // namespace combustion::yaml
