// Minimal YAML reader for rave-sim sim-dir files (config.yaml, computed.yaml,
// %08d/subconfig.yaml) as written by big-wave (ruamel.yaml round trip,
// big-wave/config.py:191-200) or by PyYAML.  Supports block mappings and
// sequences (including compact "- - x" and "- key: v" forms), flow
// collections, plain / quoted scalars, comments.  Anchors, aliases,
// multi-line scalars and multiple documents are not supported.
#pragma once
#include <string>
#include <utility>
#include <vector>

namespace rs {

struct YNode {
  enum Type { Null, Scalar, Seq, Map };
  Type type = Null;
  std::string s;
  bool quoted = false;
  std::vector<YNode> seq;
  std::vector<std::pair<std::string, YNode>> map;

  bool has(const std::string &k) const;
  const YNode &operator[](const std::string &k) const;  // throws if missing
  const YNode &operator[](size_t i) const;               // throws if out of range
  size_t size() const { return type == Seq ? seq.size() : type == Map ? map.size() : 0; }
  bool is_null() const { return type == Null; }
  double as_double(const std::string &what) const;
  long long as_int(const std::string &what) const;
  bool as_bool(const std::string &what) const;
  std::string as_string(const std::string &what) const;
};

YNode yaml_parse(const std::string &text);
YNode yaml_load_file(const std::string &path);

}  // namespace rs
