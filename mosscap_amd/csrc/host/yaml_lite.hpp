// Minimal YAML reader for problem configuration files.
//
// Covers what a Mosscap configuration uses (MosscapConfig.hpp:62-123,
// SetupModules.hpp:9-262): nested block mappings by indentation, scalars,
// flow sequences "[a, b, c]", block sequences "- a", quoted strings and "#"
// comments. Anchors, multi-line scalars and flow mappings are not supported
// (parse errors name the line).
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mh::host {

struct YNode {
    enum Kind { Null, Scalar, Map, Seq } kind = Null;
    std::string scalar;
    std::map<std::string, std::shared_ptr<YNode>> map;
    std::vector<std::shared_ptr<YNode>> seq;

    bool is_null() const { return kind == Null; }
    // child lookup; returns nullptr when absent (like `if (f[head])`)
    const YNode* get(const std::string& key) const {
        if (kind != Map) return nullptr;
        auto it = map.find(key);
        return it == map.end() ? nullptr : it->second.get();
    }
};

std::shared_ptr<YNode> parse_yaml(const std::string& text);
std::shared_ptr<YNode> load_yaml_file(const std::string& path);

}  // namespace mh::host
