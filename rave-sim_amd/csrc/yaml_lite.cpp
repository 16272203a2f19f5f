#include "yaml_lite.hpp"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "engine.hpp"

namespace rs {

namespace {

[[noreturn]] void fail(const std::string &m) { throw Error(2, "YAML: " + m); }

struct Line {
  int indent;
  std::string text;  // without indentation and comment
  int no;
};

// strip a trailing comment that is outside quotes (" #" or leading "#")
std::string strip_comment(const std::string &s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
  }
  return s;
}

std::string rtrim(const std::string &s) {
  size_t e = s.size();
  while (e > 0 && (s[e - 1] == ' ' || s[e - 1] == '\t' || s[e - 1] == '\r')) --e;
  return s.substr(0, e);
}
std::string trim(const std::string &s) {
  size_t b = 0;
  while (b < s.size() && (s[b] == ' ' || s[b] == '\t')) ++b;
  return rtrim(s.substr(b));
}

int bracket_balance(const std::string &s) {
  int bal = 0;
  bool sq = false, dq = false;
  for (char c : s) {
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (!sq && !dq) {
      if (c == '[' || c == '{') ++bal;
      else if (c == ']' || c == '}') --bal;
    }
  }
  return bal;
}

// position of the ':' that makes `s` a "key: value" entry, or npos
size_t key_colon(const std::string &s) {
  if (s.empty() || s[0] == '[' || s[0] == '{') return std::string::npos;
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == ':' && !sq && !dq && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
  }
  return std::string::npos;
}

// anchors (&name) and aliases (*name), as PyYAML writes them for objects
// that occur more than once (a material list shared by several elements)
using Anchors = std::map<std::string, YNode>;

bool name_end(char c) { return c == ' ' || c == '\t' || c == ',' || c == ']' || c == '}'; }

// remove a leading "&name" from t and return the name ("" when t has none)
std::string take_anchor(std::string &t) {
  if (t.empty() || t[0] != '&') return "";
  size_t e = 1;
  while (e < t.size() && !name_end(t[e])) ++e;
  const std::string name = t.substr(1, e - 1);
  if (name.empty()) fail("empty anchor name");
  t = trim(t.substr(e));
  return name;
}

YNode alias_node(const Anchors &a, const std::string &name) {
  const auto it = a.find(name);
  if (it == a.end()) fail("alias *" + name + " before its anchor");
  return it->second;
}

YNode scalar_node(const std::string &raw) {
  std::string s = trim(raw);
  // drop tags such as !!float or !!python/tuple
  while (s.size() > 1 && s[0] == '!') {
    size_t sp = s.find(' ');
    s = sp == std::string::npos ? std::string() : trim(s.substr(sp));
  }
  YNode n;
  if (s.empty() || s == "~" || s == "null" || s == "Null" || s == "NULL") return n;
  n.type = YNode::Scalar;
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\''))) {
    n.quoted = true;
    std::string body = s.substr(1, s.size() - 2);
    if (s.front() == '\'') {
      std::string o;
      for (size_t i = 0; i < body.size(); ++i) {
        if (body[i] == '\'' && i + 1 < body.size() && body[i + 1] == '\'') ++i;
        o += body[i];
      }
      n.s = o;
    } else {
      std::string o;
      for (size_t i = 0; i < body.size(); ++i) {
        if (body[i] == '\\' && i + 1 < body.size()) {
          const char e = body[++i];
          o += e == 'n' ? '\n' : e == 't' ? '\t' : e;
        } else {
          o += body[i];
        }
      }
      n.s = o;
    }
  } else {
    n.s = s;
  }
  return n;
}

struct FlowParser {
  const std::string &s;
  Anchors &anchors;
  size_t i = 0;
  FlowParser(const std::string &t, Anchors &a) : s(t), anchors(a) {}
  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
  }
  std::string name() {
    const size_t b = ++i;
    while (i < s.size() && !name_end(s[i])) ++i;
    if (i == b) fail("empty anchor or alias name: " + s);
    return s.substr(b, i - b);
  }
  YNode value() {
    ws();
    if (i < s.size() && s[i] == '*') return alias_node(anchors, name());
    if (i < s.size() && s[i] == '&') {
      const std::string a = name();
      ws();
      YNode n = (i < s.size() && s[i] != ',' && s[i] != ']' && s[i] != '}') ? value() : YNode{};
      anchors[a] = n;
      return n;
    }
    if (i < s.size() && s[i] == '[') return seq();
    if (i < s.size() && s[i] == '{') return map();
    return scalar();
  }
  YNode scalar() {
    ws();
    size_t b = i;
    if (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
      const char q = s[i++];
      while (i < s.size()) {
        if (s[i] == q) {
          if (q == '\'' && i + 1 < s.size() && s[i + 1] == '\'') { i += 2; continue; }
          break;
        }
        if (q == '"' && s[i] == '\\') ++i;
        ++i;
      }
      ++i;
      return scalar_node(s.substr(b, i - b));
    }
    while (i < s.size() && s[i] != ',' && s[i] != ']' && s[i] != '}') ++i;
    return scalar_node(s.substr(b, i - b));
  }
  YNode seq() {
    YNode n;
    n.type = YNode::Seq;
    ++i;
    ws();
    if (i < s.size() && s[i] == ']') { ++i; return n; }
    while (true) {
      n.seq.push_back(value());
      ws();
      if (i < s.size() && s[i] == ',') { ++i; ws(); if (i < s.size() && s[i] == ']') { ++i; return n; } continue; }
      if (i < s.size() && s[i] == ']') { ++i; return n; }
      fail("bad flow sequence: " + s);
    }
  }
  YNode map() {
    YNode n;
    n.type = YNode::Map;
    ++i;
    ws();
    if (i < s.size() && s[i] == '}') { ++i; return n; }
    while (true) {
      ws();
      size_t b = i;
      while (i < s.size() && s[i] != ':') ++i;
      if (i >= s.size()) fail("bad flow mapping: " + s);
      YNode key = scalar_node(s.substr(b, i - b));
      ++i;
      n.map.emplace_back(key.s, value());
      ws();
      if (i < s.size() && s[i] == ',') { ++i; continue; }
      if (i < s.size() && s[i] == '}') { ++i; return n; }
      fail("bad flow mapping: " + s);
    }
  }
};

YNode inline_value(const std::string &v, Anchors &anchors) {
  const std::string t = trim(v);
  if (!t.empty() && (t[0] == '[' || t[0] == '{' || t[0] == '&' || t[0] == '*')) {
    FlowParser p(t, anchors);
    YNode n = p.value();
    p.ws();
    if (p.i != t.size()) fail("trailing characters after flow collection: " + t);
    return n;
  }
  return scalar_node(t);
}

struct BlockParser {
  std::vector<Line> L;
  size_t pos = 0;
  Anchors anchors;

  bool is_dash(const std::string &t) const { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

  // join continuation lines of an unbalanced flow collection
  std::string take_flow(std::string v) {
    while (bracket_balance(v) > 0 && pos < L.size()) v += " " + L[pos++].text;
    return v;
  }

  YNode node(int min_indent) {
    if (pos >= L.size() || L[pos].indent < min_indent) return YNode{};
    const int ind = L[pos].indent;
    const std::string &t = L[pos].text;
    if (is_dash(t)) return sequence(ind);
    if (key_colon(t) != std::string::npos) return mapping(ind);
    ++pos;
    return inline_value(take_flow(t), anchors);
  }

  YNode sequence(int ind) {
    YNode n;
    n.type = YNode::Seq;
    while (pos < L.size() && L[pos].indent == ind && is_dash(L[pos].text)) {
      std::string rest = L[pos].text.substr(1);
      size_t k = 0;
      while (k < rest.size() && rest[k] == ' ') ++k;
      std::string item = rest.substr(k);
      const std::string anchor = take_anchor(item);  // "- &name" before the item
      k = rest.size() - item.size();
      if (item.empty()) {
        ++pos;
        if (pos < L.size() && L[pos].indent > ind) n.seq.push_back(node(L[pos].indent));
        else n.seq.push_back(YNode{});
      } else {
        // re-read the remainder of this line as a node at its own column
        L[pos].indent = ind + 1 + (int)k;
        L[pos].text = item;
        n.seq.push_back(node(L[pos].indent));
      }
      if (!anchor.empty()) anchors[anchor] = n.seq.back();
    }
    return n;
  }

  YNode mapping(int ind) {
    YNode n;
    n.type = YNode::Map;
    while (pos < L.size() && L[pos].indent == ind && !is_dash(L[pos].text)) {
      const std::string t = L[pos].text;
      const size_t c = key_colon(t);
      if (c == std::string::npos) fail("expected 'key:' at line " + std::to_string(L[pos].no));
      const std::string key = scalar_node(t.substr(0, c)).s;
      std::string val = c + 1 < t.size() ? trim(t.substr(c + 1)) : std::string();
      ++pos;
      // "key: &name" followed by a block node on the next lines
      std::string anchor;
      if (!val.empty() && val[0] == '&') {
        std::string rest = val;
        anchor = take_anchor(rest);
        if (rest.empty()) val.clear();
        else anchor.clear();  // an inline anchored value: the flow parser records it
      }
      if (!val.empty()) {
        n.map.emplace_back(key, inline_value(take_flow(val), anchors));
      } else if (pos < L.size() && L[pos].indent > ind) {
        n.map.emplace_back(key, node(L[pos].indent));
      } else if (pos < L.size() && L[pos].indent == ind && is_dash(L[pos].text)) {
        n.map.emplace_back(key, sequence(ind));
      } else {
        n.map.emplace_back(key, YNode{});
      }
      if (!anchor.empty()) anchors[anchor] = n.map.back().second;
    }
    return n;
  }
};

}  // namespace

bool YNode::has(const std::string &k) const {
  if (type != Map) return false;
  for (auto &kv : map)
    if (kv.first == k) return true;
  return false;
}
const YNode &YNode::operator[](const std::string &k) const {
  if (type == Map)
    for (auto &kv : map)
      if (kv.first == k) return kv.second;
  throw Error(2, "config: key '" + k + "' does not exist");
}
const YNode &YNode::operator[](size_t i) const {
  if (type != Seq || i >= seq.size()) throw Error(2, "config: index " + std::to_string(i) + " out of range");
  return seq[i];
}
double YNode::as_double(const std::string &what) const {
  if (type != Scalar || quoted) throw Error(2, "config: entry '" + what + "' must be a float");
  std::string t = s;
  if (t == ".inf" || t == ".Inf" || t == "+.inf") return HUGE_VAL;
  if (t == "-.inf" || t == "-.Inf") return -HUGE_VAL;
  if (t == ".nan" || t == ".NaN") return NAN;
  errno = 0;
  char *end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str() || *end != '\0') throw Error(2, "config: entry '" + what + "' must be a float, got '" + s + "'");
  return v;
}
long long YNode::as_int(const std::string &what) const {
  if (type != Scalar || quoted) throw Error(2, "config: entry '" + what + "' must be an int");
  char *end = nullptr;
  const long long v = std::strtoll(s.c_str(), &end, 10);
  if (end == s.c_str() || *end != '\0') throw Error(2, "config: entry '" + what + "' must be an int, got '" + s + "'");
  return v;
}
bool YNode::as_bool(const std::string &what) const {
  if (type == Scalar && !quoted) {
    if (s == "true" || s == "True" || s == "TRUE") return true;
    if (s == "false" || s == "False" || s == "FALSE") return false;
  }
  throw Error(2, "config: entry '" + what + "' must be a bool");
}
std::string YNode::as_string(const std::string &what) const {
  if (type != Scalar) throw Error(2, "config: entry '" + what + "' must be a string");
  return s;
}

YNode yaml_parse(const std::string &text) {
  BlockParser bp;
  std::istringstream in(text);
  std::string raw;
  int no = 0;
  while (std::getline(in, raw)) {
    ++no;
    if (raw.find('\t') != std::string::npos) {
      // tabs are not valid YAML indentation; treat them as spaces in content only
    }
    std::string s = rtrim(strip_comment(raw));
    int ind = 0;
    while (ind < (int)s.size() && s[ind] == ' ') ++ind;
    std::string t = s.substr(ind);
    if (t.empty() || t == "---" || t == "...") continue;
    if (t.rfind("%YAML", 0) == 0 || t.rfind("%TAG", 0) == 0) continue;
    bp.L.push_back({ind, t, no});
  }
  if (bp.L.empty()) return YNode{};
  YNode root = bp.node(0);
  if (bp.pos != bp.L.size()) fail("unparsed content at line " + std::to_string(bp.L[bp.pos].no));
  return root;
}

YNode yaml_load_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw Error(3, "cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return yaml_parse(ss.str());
}

}  // namespace rs
