#include "yaml_lite.hpp"

#include <fstream>
#include <sstream>

namespace mh::host {

namespace {

struct Line {
    int indent;
    std::string text;  // stripped of indentation and comments
    int number;
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

std::string strip_comment(const std::string& s) {
    bool sq = false, dq = false;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (c == '\'' && !dq) sq = !sq;
        else if (c == '"' && !sq) dq = !dq;
        else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
    }
    return s;
}

std::string unquote(const std::string& s) {
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

std::shared_ptr<YNode> scalar_node(const std::string& v) {
    auto n = std::make_shared<YNode>();
    const std::string t = trim(v);
    if (t.empty() || t == "~" || t == "null") return n;
    if (t.front() == '[') {
        if (t.back() != ']') throw std::runtime_error("YAML: unterminated flow sequence: " + t);
        n->kind = YNode::Seq;
        std::string body = t.substr(1, t.size() - 2);
        std::stringstream ss(body);
        std::string item;
        while (std::getline(ss, item, ',')) {
            if (trim(item).empty()) continue;
            n->seq.push_back(scalar_node(item));
        }
        return n;
    }
    n->kind = YNode::Scalar;
    n->scalar = unquote(t);
    return n;
}

// parse lines[i..] at indentation > parent_indent
std::shared_ptr<YNode> parse_block(const std::vector<Line>& lines, size_t& i, int indent) {
    auto node = std::make_shared<YNode>();
    if (i >= lines.size()) return node;
    const bool seq = lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-";
    node->kind = seq ? YNode::Seq : YNode::Map;
    while (i < lines.size() && lines[i].indent == indent) {
        const Line& L = lines[i];
        if (seq) {
            if (L.text.rfind("-", 0) != 0)
                throw std::runtime_error("YAML line " + std::to_string(L.number) + ": mixed sequence and mapping");
            const std::string rest = trim(L.text.substr(1));
            ++i;
            if (rest.empty()) {
                node->seq.push_back(i < lines.size() && lines[i].indent > indent ? parse_block(lines, i, lines[i].indent)
                                                                                  : std::make_shared<YNode>());
            } else {
                node->seq.push_back(scalar_node(rest));
            }
            continue;
        }
        const size_t colon = L.text.find(':');
        if (colon == std::string::npos)
            throw std::runtime_error("YAML line " + std::to_string(L.number) + ": expected 'key: value'");
        const std::string key = unquote(trim(L.text.substr(0, colon)));
        const std::string rest = trim(L.text.substr(colon + 1));
        ++i;
        if (rest.empty()) {
            node->map[key] = (i < lines.size() && lines[i].indent > indent) ? parse_block(lines, i, lines[i].indent)
                                                                              : std::make_shared<YNode>();
        } else {
            node->map[key] = scalar_node(rest);
        }
    }
    if (i < lines.size() && lines[i].indent > indent)
        throw std::runtime_error("YAML line " + std::to_string(lines[i].number) + ": bad indentation");
    return node;
}

}  // namespace

std::shared_ptr<YNode> parse_yaml(const std::string& text) {
    std::vector<Line> lines;
    std::stringstream ss(text);
    std::string raw;
    int num = 0;
    while (std::getline(ss, raw)) {
        ++num;
        const std::string s = strip_comment(raw);
        if (trim(s).empty() || trim(s) == "---") continue;
        int ind = 0;
        while (ind < (int)s.size() && s[ind] == ' ') ++ind;
        if (ind < (int)s.size() && s[ind] == '\t')
            throw std::runtime_error("YAML line " + std::to_string(num) + ": tabs are not allowed for indentation");
        lines.push_back(Line{ind, trim(s), num});
    }
    size_t i = 0;
    if (lines.empty()) {
        auto n = std::make_shared<YNode>();
        n->kind = YNode::Map;
        return n;
    }
    auto root = parse_block(lines, i, lines[0].indent);
    if (i != lines.size())
        throw std::runtime_error("YAML line " + std::to_string(lines[i].number) + ": bad indentation");
    return root;
}

std::shared_ptr<YNode> load_yaml_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open config file " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse_yaml(ss.str());
}

}  // namespace mh::host
