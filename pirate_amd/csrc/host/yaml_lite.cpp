#include "yaml_lite.hpp"
#include "common.hpp"

#include <cerrno>
#include <cstdlib>
#include <sstream>

namespace pirate_amd {

using NodeP = std::shared_ptr<YamlNode>;

namespace {

struct Line {
    int indent;
    std::string text;   // comment-stripped, right-trimmed, indentation removed
    int lineno;
};

std::string rtrim(const std::string &s)
{
    size_t n = s.size();
    while (n > 0 && (s[n-1] == ' ' || s[n-1] == '\t' || s[n-1] == '\r'))
        n--;
    return s.substr(0, n);
}

std::string trim(const std::string &s)
{
    size_t i = 0;
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t'))
        i++;
    return rtrim(s.substr(i));
}

// Removes a '#' comment (one that starts the line or follows whitespace), ignoring
// '#' characters inside quotes.
std::string strip_comment(const std::string &s)
{
    char q = 0;
    for (size_t i = 0; i < s.size(); i++) {
        char c = s[i];
        if (q) {
            if (c == q)
                q = 0;
        }
        else if (c == '\'' || c == '"')
            q = c;
        else if (c == '#' && (i == 0 || s[i-1] == ' ' || s[i-1] == '\t'))
            return s.substr(0, i);
    }
    return s;
}

NodeP make_scalar(const std::string &s)
{
    auto n = std::make_shared<YamlNode>();
    n->kind = YamlNode::Scalar;
    std::string t = trim(s);
    if (t.size() >= 2 && ((t.front() == '\'' && t.back() == '\'') || (t.front() == '"' && t.back() == '"')))
        t = t.substr(1, t.size() - 2);
    n->scalar = t;
    return n;
}

// Finds the ':' that separates a mapping key from its value (outside brackets/quotes),
// requiring it to be followed by whitespace or end-of-string. Returns npos if none.
size_t find_key_colon(const std::string &s)
{
    int depth = 0;
    char q = 0;
    for (size_t i = 0; i < s.size(); i++) {
        char c = s[i];
        if (q) {
            if (c == q)
                q = 0;
            continue;
        }
        if (c == '\'' || c == '"')
            q = c;
        else if (c == '[' || c == '{')
            depth++;
        else if (c == ']' || c == '}')
            depth--;
        else if (c == ':' && depth == 0 && (i+1 == s.size() || s[i+1] == ' ' || s[i+1] == '\t'))
            return i;
    }
    return std::string::npos;
}

// ---- flow-style parser ----

struct FlowParser {
    const std::string &s;
    size_t pos = 0;
    std::string where;

    FlowParser(const std::string &s_, const std::string &w) : s(s_), where(w) { }

    void skip_ws() { while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) pos++; }

    [[noreturn]] void fail(const std::string &msg)
    {
        throw pdd_error("yaml: " + where + ": " + msg + " in '" + s + "'");
    }

    NodeP parse_value()
    {
        skip_ws();
        if (pos >= s.size())
            fail("unexpected end of flow value");
        if (s[pos] == '[')
            return parse_seq();
        if (s[pos] == '{')
            return parse_map();
        return parse_scalar_token();
    }

    NodeP parse_scalar_token()
    {
        skip_ws();
        size_t start = pos;
        char q = 0;
        if (pos < s.size() && (s[pos] == '\'' || s[pos] == '"')) {
            q = s[pos++];
            while (pos < s.size() && s[pos] != q)
                pos++;
            if (pos >= s.size())
                fail("unterminated quote");
            pos++;
            return make_scalar(s.substr(start, pos - start));
        }
        while (pos < s.size() && s[pos] != ',' && s[pos] != ']' && s[pos] != '}')
            pos++;
        return make_scalar(s.substr(start, pos - start));
    }

    NodeP parse_seq()
    {
        auto n = std::make_shared<YamlNode>();
        n->kind = YamlNode::Seq;
        pos++;  // '['
        skip_ws();
        if (pos < s.size() && s[pos] == ']') {
            pos++;
            return n;
        }
        for (;;) {
            n->seq.push_back(parse_value());
            skip_ws();
            if (pos >= s.size())
                fail("unterminated '['");
            if (s[pos] == ',') {
                pos++;
                continue;
            }
            if (s[pos] == ']') {
                pos++;
                return n;
            }
            fail("expected ',' or ']'");
        }
    }

    NodeP parse_map()
    {
        auto n = std::make_shared<YamlNode>();
        n->kind = YamlNode::Map;
        pos++;  // '{'
        skip_ws();
        if (pos < s.size() && s[pos] == '}') {
            pos++;
            return n;
        }
        for (;;) {
            skip_ws();
            size_t start = pos;
            while (pos < s.size() && s[pos] != ':')
                pos++;
            if (pos >= s.size())
                fail("expected ':' in flow map");
            std::string key = trim(s.substr(start, pos - start));
            pos++;  // ':'
            NodeP v = parse_value();
            n->map.emplace_back(key, v);
            skip_ws();
            if (pos >= s.size())
                fail("unterminated '{'");
            if (s[pos] == ',') {
                pos++;
                continue;
            }
            if (s[pos] == '}') {
                pos++;
                return n;
            }
            fail("expected ',' or '}'");
        }
    }
};

NodeP parse_inline(const std::string &text, const std::string &where)
{
    std::string t = trim(text);
    if (t.empty()) {
        auto n = std::make_shared<YamlNode>();
        n->kind = YamlNode::Null;
        return n;
    }
    if (t[0] == '[' || t[0] == '{') {
        FlowParser fp(t, where);
        NodeP n = fp.parse_value();
        fp.skip_ws();
        if (fp.pos != t.size())
            fp.fail("trailing characters after flow value");
        return n;
    }
    return make_scalar(t);
}

// ---- block-style parser ----

struct BlockParser {
    std::vector<Line> lines;
    size_t i = 0;
    std::string name;

    [[noreturn]] void fail(const Line &l, const std::string &msg)
    {
        std::ostringstream ss;
        ss << "yaml: " << name << " line " << l.lineno << ": " << msg;
        throw pdd_error(ss.str());
    }

    // Parses a block node whose lines have indentation exactly 'indent'.
    NodeP parse_block(int indent)
    {
        PDD_CHECK(i < lines.size(), "yaml: internal error");
        if (lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-")
            return parse_seq(indent);
        return parse_map(indent);
    }

    NodeP parse_seq(int indent)
    {
        auto n = std::make_shared<YamlNode>();
        n->kind = YamlNode::Seq;
        while (i < lines.size() && lines[i].indent == indent &&
               (lines[i].text.rfind("- ", 0) == 0 || lines[i].text == "-")) {
            Line l = lines[i];
            std::string rest = (l.text == "-") ? std::string() : l.text.substr(2);
            std::string rt = trim(rest);
            size_t lead = rest.find_first_not_of(' ');
            int child_indent = indent + 2 + int((lead == std::string::npos) ? 0 : lead);
            if (rt.empty()) {
                i++;
                if (i < lines.size() && lines[i].indent > indent)
                    n->seq.push_back(parse_block(lines[i].indent));
                else
                    n->seq.push_back(parse_inline("", name));
                continue;
            }
            size_t c = find_key_colon(rt);
            if (c != std::string::npos && rt[0] != '[' && rt[0] != '{') {
                // "- key: value" starts a block map; rewrite this line as a map line at the
                // item's indentation and parse the map.
                lines[i].indent = child_indent;
                lines[i].text = rt;
                n->seq.push_back(parse_map(child_indent));
                continue;
            }
            n->seq.push_back(parse_inline(rt, name));
            i++;
        }
        return n;
    }

    NodeP parse_map(int indent)
    {
        auto n = std::make_shared<YamlNode>();
        n->kind = YamlNode::Map;
        while (i < lines.size() && lines[i].indent == indent) {
            const Line l = lines[i];
            if (l.text.rfind("- ", 0) == 0)
                fail(l, "unexpected sequence item in mapping");
            size_t c = find_key_colon(l.text);
            if (c == std::string::npos)
                fail(l, "expected 'key: value'");
            std::string key = trim(l.text.substr(0, c));
            std::string val = trim(l.text.substr(c + 1));
            for (auto &kv : n->map)
                if (kv.first == key)
                    fail(l, "duplicate key '" + key + "'");
            i++;
            if (val.empty()) {
                if (i < lines.size() && lines[i].indent > indent)
                    n->map.emplace_back(key, parse_block(lines[i].indent));
                else if (i < lines.size() && lines[i].indent == indent &&
                         lines[i].text.rfind("- ", 0) == 0)
                    n->map.emplace_back(key, parse_seq(indent));   // "key:\n- a\n- b" (same indent)
                else
                    n->map.emplace_back(key, parse_inline("", name));
            }
            else
                n->map.emplace_back(key, parse_inline(val, name + ":" + key));
        }
        if (i < lines.size() && lines[i].indent > indent)
            fail(lines[i], "unexpected indentation");
        return n;
    }
};

}  // anonymous namespace


NodeP YamlNode::parse(const std::string &text, const std::string &name)
{
    BlockParser bp;
    bp.name = name;

    std::istringstream in(text);
    std::string raw;
    int lineno = 0;
    while (std::getline(in, raw)) {
        lineno++;
        for (char c : raw)
            if (c == '\t')
                throw pdd_error("yaml: " + name + ": tab characters are not supported (line " + std::to_string(lineno) + ")");
        std::string s = rtrim(strip_comment(raw));
        if (trim(s).empty())
            continue;
        if (trim(s) == "---")
            continue;
        int ind = 0;
        while (ind < int(s.size()) && s[ind] == ' ')
            ind++;
        // A flow collection may continue over several lines: join until brackets balance.
        std::string body = s.substr(ind);
        auto depth_of = [](const std::string &x) {
            int d = 0;
            char q = 0;
            for (char c : x) {
                if (q) { if (c == q) q = 0; continue; }
                if (c == '\'' || c == '"') q = c;
                else if (c == '[' || c == '{') d++;
                else if (c == ']' || c == '}') d--;
            }
            return d;
        };
        int start_line = lineno;
        while (depth_of(body) > 0 && std::getline(in, raw)) {
            lineno++;
            body += " " + trim(rtrim(strip_comment(raw)));
        }
        bp.lines.push_back({ind, body, start_line});
    }

    if (bp.lines.empty()) {
        auto n = std::make_shared<YamlNode>();
        n->kind = YamlNode::Map;
        return n;
    }

    NodeP root = bp.parse_block(bp.lines[0].indent);
    if (bp.i != bp.lines.size())
        bp.fail(bp.lines[bp.i], "unexpected content (bad indentation?)");
    return root;
}


// ---------------------------------------------------------------------------------------------


YamlView::YamlView(std::shared_ptr<YamlNode> node, std::string name) :
    node_(std::move(node)), name_(std::move(name))
{
    if (node_->kind == YamlNode::Map && node_->map_used.size() != node_->map.size())
        node_->map_used.assign(node_->map.size(), false);
}

bool YamlView::has_key(const std::string &k) const
{
    if (node_->kind != YamlNode::Map)
        return false;
    for (auto &kv : node_->map)
        if (kv.first == k)
            return true;
    return false;
}

YamlView YamlView::operator[](const std::string &k) const
{
    if (node_->kind != YamlNode::Map)
        throw pdd_error("yaml: " + name_ + " is not a mapping (looking up key '" + k + "')");
    for (size_t j = 0; j < node_->map.size(); j++) {
        if (node_->map[j].first == k) {
            node_->map_used[j] = true;
            return YamlView(node_->map[j].second, name_ + "." + k);
        }
    }
    throw pdd_error("yaml: " + name_ + ": missing required key '" + k + "'");
}

YamlView YamlView::operator[](long i) const
{
    if (node_->kind != YamlNode::Seq)
        throw pdd_error("yaml: " + name_ + " is not a sequence");
    if (i < 0 || i >= long(node_->seq.size()))
        throw pdd_error("yaml: " + name_ + ": index out of range");
    return YamlView(node_->seq[i], name_ + "[" + std::to_string(i) + "]");
}

long YamlView::size() const
{
    if (node_->kind != YamlNode::Seq)
        throw pdd_error("yaml: " + name_ + " is not a sequence");
    return node_->seq.size();
}

std::string YamlView::as_string() const
{
    if (node_->kind != YamlNode::Scalar)
        throw pdd_error("yaml: " + name_ + ": expected a scalar");
    return node_->scalar;
}

long YamlView::as_long() const
{
    std::string s = as_string();
    errno = 0;
    char *end = nullptr;
    long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || errno != 0 || *end != 0)
        throw pdd_error("yaml: " + name_ + ": expected an integer, got '" + s + "'");
    return v;
}

double YamlView::as_double() const
{
    std::string s = as_string();
    errno = 0;
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (s.empty() || errno != 0 || *end != 0)
        throw pdd_error("yaml: " + name_ + ": expected a number, got '" + s + "'");
    return v;
}

std::vector<long> YamlView::as_long_vector() const
{
    std::vector<long> ret;
    for (long j = 0; j < size(); j++)
        ret.push_back((*this)[j].as_long());
    return ret;
}

std::vector<double> YamlView::as_double_vector() const
{
    std::vector<double> ret;
    for (long j = 0; j < size(); j++)
        ret.push_back((*this)[j].as_double());
    return ret;
}

void YamlView::check_for_invalid_keys() const
{
    if (node_->kind != YamlNode::Map)
        return;
    for (size_t j = 0; j < node_->map.size(); j++)
        if (!node_->map_used[j])
            throw pdd_error("yaml: " + name_ + ": invalid key '" + node_->map[j].first + "'");
}

}  // namespace pirate_amd
