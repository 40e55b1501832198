// yaml_lite: a small YAML-subset reader/writer, sufficient for the dedispersion
// config grammar (reference: configs/dedispersion/*.yml): block maps, block sequences
// (of scalars, flow maps, or block maps), flow sequences "[a, b]", flow maps
// "{k: v, ...}", plain or quoted scalars, '#' comments.
//
// Accessors track which keys were read, so that callers can reject unknown keys
// (reference semantics: src_lib/YamlFile.cpp, check_for_invalid_keys()).

#ifndef PIRATE_AMD_YAML_LITE_HPP
#define PIRATE_AMD_YAML_LITE_HPP

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace pirate_amd {

struct YamlNode {
    enum Kind { Null, Scalar, Seq, Map };
    Kind kind = Null;
    std::string scalar;                                      // if kind == Scalar
    std::vector<std::shared_ptr<YamlNode>> seq;              // if kind == Seq
    std::vector<std::pair<std::string, std::shared_ptr<YamlNode>>> map;  // if kind == Map
    mutable std::vector<bool> map_used;                      // parallel to 'map'

    static std::shared_ptr<YamlNode> parse(const std::string &text, const std::string &name);
};

// Read-only view of a node with a name (for error messages) and strict key tracking.
class YamlView {
public:
    YamlView(std::shared_ptr<YamlNode> node, std::string name);

    const std::string &name() const { return name_; }
    bool is_map() const { return node_->kind == YamlNode::Map; }
    bool is_seq() const { return node_->kind == YamlNode::Seq; }
    bool has_key(const std::string &k) const;
    YamlView operator[](const std::string &k) const;   // throws if missing
    YamlView operator[](long i) const;                 // sequence element
    long size() const;                                 // sequence length

    long as_long() const;
    double as_double() const;
    std::string as_string() const;
    std::vector<long> as_long_vector() const;
    std::vector<double> as_double_vector() const;

    long get_long(const std::string &k) const { return (*this)[k].as_long(); }
    long get_long(const std::string &k, long dflt) const { return has_key(k) ? get_long(k) : dflt; }
    double get_double(const std::string &k) const { return (*this)[k].as_double(); }
    std::string get_string(const std::string &k) const { return (*this)[k].as_string(); }

    // Throws if any key of this map was never accessed.
    void check_for_invalid_keys() const;

private:
    std::shared_ptr<YamlNode> node_;
    std::string name_;
};

}  // namespace pirate_amd

#endif
