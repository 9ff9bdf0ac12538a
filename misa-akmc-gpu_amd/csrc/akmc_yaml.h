// Minimal YAML reader for the reference's config.yaml schema
// (examples/config.yaml): block maps by indentation, flow sequences
// ([a, b, c]), scalars, quoted strings and # comments.  Map keys keep their
// document order, which matters for the alloy list (the reference walks it
// cumulatively, src/lattice/lattice_types.cpp:9-19).
#pragma once
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace akmc {
namespace yaml {

struct Node {
  enum Kind { NONE, SCALAR, MAP, SEQ } kind = NONE;
  std::string scalar;
  std::vector<std::pair<std::string, Node>> map;
  std::vector<Node> seq;

  bool is_map() const { return kind == MAP; }
  bool is_seq() const { return kind == SEQ; }
  const Node* get(const std::string& k) const {
    for (const auto& kv : map)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
  long long as_int(long long def) const {
    if (kind != SCALAR || scalar.empty()) return def;
    char* end = nullptr;
    const double d = strtod(scalar.c_str(), &end);
    if (end == scalar.c_str()) return def;
    return (long long)d == d ? (long long)d : strtoll(scalar.c_str(), nullptr, 10);
  }
  double as_double(double def) const {
    if (kind != SCALAR || scalar.empty()) return def;
    char* end = nullptr;
    const double d = strtod(scalar.c_str(), &end);
    return end == scalar.c_str() ? def : d;
  }
  bool as_bool(bool def) const {
    if (kind != SCALAR) return def;
    if (scalar == "true" || scalar == "True" || scalar == "yes" || scalar == "on") return true;
    if (scalar == "false" || scalar == "False" || scalar == "no" || scalar == "off") return false;
    return def;
  }
  long long get_int(const std::string& k, long long def) const {
    const Node* n = get(k);
    return n ? n->as_int(def) : def;
  }
  double get_double(const std::string& k, double def) const {
    const Node* n = get(k);
    return n ? n->as_double(def) : def;
  }
  bool get_bool(const std::string& k, bool def) const {
    const Node* n = get(k);
    return n ? n->as_bool(def) : def;
  }
};

inline std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r");
  return s.substr(a, b - a + 1);
}

// Removes a trailing comment that is not inside quotes.
inline std::string strip_comment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    const char c = s[i];
    if (q) {
      if (c == q) q = 0;
    } else if (c == '"' || c == '\'') {
      q = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return s.substr(0, i);
    }
  }
  return s;
}

inline Node scalar_node(const std::string& raw) {
  Node n;
  std::string v = trim(raw);
  if (v.size() >= 2 && (v[0] == '"' || v[0] == '\'') && v.back() == v[0]) v = v.substr(1, v.size() - 2);
  n.kind = Node::SCALAR;
  n.scalar = v;
  return n;
}

inline bool parse_flow_seq(const std::string& v, Node* out) {
  if (v.size() < 2 || v[0] != '[' || v.back() != ']') return false;
  out->kind = Node::SEQ;
  std::string body = v.substr(1, v.size() - 2);
  std::stringstream ss(body);
  std::string item;
  while (std::getline(ss, item, ',')) {
    if (trim(item).empty()) continue;
    out->seq.push_back(scalar_node(item));
  }
  return true;
}

struct Line {
  int indent;
  std::string text;
  int lineno;
};

inline bool parse_block(const std::vector<Line>& lines, size_t* pos, int indent, Node* out, std::string* err) {
  out->kind = Node::MAP;
  while (*pos < lines.size()) {
    const Line& L = lines[*pos];
    if (L.indent < indent) return true;
    if (L.indent > indent) {
      *err = "unexpected indentation at line " + std::to_string(L.lineno);
      return false;
    }
    const size_t colon = L.text.find(':');
    if (colon == std::string::npos) {
      *err = "expected 'key: value' at line " + std::to_string(L.lineno);
      return false;
    }
    std::string key = trim(L.text.substr(0, colon));
    if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'')) key = key.substr(1, key.size() - 2);
    const std::string val = trim(L.text.substr(colon + 1));
    (*pos)++;
    Node child;
    if (val.empty()) {
      if (*pos < lines.size() && lines[*pos].indent > indent) {
        if (!parse_block(lines, pos, lines[*pos].indent, &child, err)) return false;
      }
    } else if (val[0] == '[') {
      if (!parse_flow_seq(val, &child)) {
        *err = "bad flow sequence at line " + std::to_string(L.lineno);
        return false;
      }
    } else {
      child = scalar_node(val);
    }
    out->map.emplace_back(key, std::move(child));
  }
  return true;
}

inline bool parse_string(const std::string& text, Node* root, std::string* err) {
  std::vector<Line> lines;
  std::stringstream ss(text);
  std::string raw;
  int no = 0;
  while (std::getline(ss, raw)) {
    no++;
    std::string s = strip_comment(raw);
    if (trim(s).empty()) continue;
    if (trim(s) == "---") continue;
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ind++;
    lines.push_back({ind, trim(s), no});
  }
  size_t pos = 0;
  if (lines.empty()) {
    root->kind = Node::MAP;
    return true;
  }
  if (!parse_block(lines, &pos, lines[0].indent, root, err)) return false;
  if (pos != lines.size()) {
    *err = "trailing content at line " + std::to_string(lines[pos].lineno);
    return false;
  }
  return true;
}

inline bool parse_file(const char* path, Node* root, std::string* err) {
  std::ifstream f(path);
  if (!f) {
    *err = "cannot open file";
    return false;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_string(ss.str(), root, err);
}

}  // namespace yaml
}  // namespace akmc
