// yaml_lite.cpp -- the YAML subset reader (see yaml_lite.hpp).
#include "yaml_lite.hpp"

#include <cerrno>
#include <climits>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace quantkos::yaml {

namespace {

const Node &undefined_node() {
    static const Node n;
    return n;
}

std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

[[noreturn]] void syntax(int line, const std::string &what) {
    throw ParserException("yaml-cpp: error at line " + std::to_string(line) + ": " + what);
}

// Removes a trailing comment ('#' at line start or after a blank), outside quotes.
std::string strip_comment(const std::string &s) {
    char q = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (q) {
            if (c == q) q = 0;
            continue;
        }
        if (c == '"' || c == '\'') {
            q = c;
        } else if (c == '#' && (i == 0 || s[i - 1] == ' ')) {
            return s.substr(0, i);
        }
    }
    return s;
}

std::string unquote(const std::string &v, int line) {
    if (v.size() >= 2 && v.front() == '"' && v.back() == '"') {
        std::string out;
        for (size_t i = 1; i + 1 < v.size(); ++i) {
            if (v[i] == '\\' && i + 2 < v.size()) {
                const char n = v[++i];
                out += n == 'n' ? '\n' : n == 't' ? '\t' : n;
            } else {
                out += v[i];
            }
        }
        return out;
    }
    if (v.size() >= 2 && v.front() == '\'' && v.back() == '\'') {
        std::string out;
        for (size_t i = 1; i + 1 < v.size(); ++i) {
            out += v[i];
            if (v[i] == '\'' && i + 2 < v.size() && v[i + 1] == '\'') ++i;
        }
        return out;
    }
    if (!v.empty() && (v.front() == '"' || v.front() == '\'')) syntax(line, "unterminated quoted scalar");
    return v;
}

}  // namespace

struct Builder {
    static std::shared_ptr<Node> map_node(int line) {
        auto n = std::make_shared<Node>();
        n->defined_ = true;
        n->is_map_ = true;
        n->line_ = line;
        return n;
    }
    static std::shared_ptr<Node> scalar_node(const std::string &v, int line) {
        auto n = std::make_shared<Node>();
        n->defined_ = true;
        n->is_map_ = false;
        n->scalar_ = v;
        n->line_ = line;
        return n;
    }
    static void add(Node &parent, const std::string &key, std::shared_ptr<Node> child, int line) {
        if (parent.children_.count(key)) syntax(line, "duplicate key '" + key + "'");
        parent.children_[key] = std::move(child);
        parent.order_.push_back(key);
    }
    static std::shared_ptr<Node> get(Node &parent, const std::string &key) { return parent.children_[key]; }
    static void make_map(Node &n) {
        n.is_map_ = true;
        n.scalar_.clear();
    }
};

const Node &Node::operator[](const std::string &key) const {
    if (!defined_ || !is_map_) return undefined_node();
    auto it = children_.find(key);
    return it == children_.end() ? undefined_node() : *it->second;
}

Node Node::Load(const std::string &text) {
    auto root = Builder::map_node(0);
    struct Level {
        int indent;
        std::shared_ptr<Node> node;
    };
    std::vector<Level> stack{{-1, root}};
    std::shared_ptr<Node> last_child;   // node of the previous key (may receive children)
    int last_indent = -1;
    bool last_empty = false;
    std::istringstream in(text);
    std::string raw;
    int line = 0;
    while (std::getline(in, raw)) {
        ++line;
        std::string s = strip_comment(raw);
        if (trim(s).empty()) continue;
        size_t ind = 0;
        while (ind < s.size() && s[ind] == ' ') ++ind;
        if (ind < s.size() && s[ind] == '\t') syntax(line, "tab character in indentation");
        const std::string body = trim(s.substr(ind));
        if (body == "---" || body == "...") continue;
        if (body[0] == '-' && (body.size() == 1 || body[1] == ' ')) syntax(line, "sequences are not supported");
        if (body[0] == '[' || body[0] == '{' || body[0] == '&' || body[0] == '*' || body[0] == '!')
            syntax(line, "unsupported YAML construct");
        // key: the first ':' followed by a blank or the end of line, outside quotes
        size_t colon = std::string::npos;
        char q = 0;
        for (size_t i = 0; i < body.size(); ++i) {
            const char c = body[i];
            if (q) {
                if (c == q) q = 0;
                continue;
            }
            if (c == '"' || c == '\'') q = c;
            else if (c == ':' && (i + 1 == body.size() || body[i + 1] == ' ')) {
                colon = i;
                break;
            }
        }
        if (colon == std::string::npos) syntax(line, "expected 'key: value'");
        const std::string key = unquote(trim(body.substr(0, colon)), line);
        const std::string val = trim(body.substr(colon + 1));
        if (key.empty()) syntax(line, "empty key");
        if (!val.empty() && (val[0] == '[' || val[0] == '{' || val[0] == '|' || val[0] == '>' || val[0] == '&' ||
                             val[0] == '*' || val[0] == '!'))
            syntax(line, "unsupported YAML value for key '" + key + "'");
        const int indent = (int)ind;
        if (indent > last_indent && last_child) {
            if (!last_empty) syntax(line, "bad indentation under a scalar value");
            Builder::make_map(*last_child);
            stack.push_back({indent, last_child});
        } else {
            while (stack.size() > 1 && stack.back().indent > indent) stack.pop_back();
            if (stack.back().indent != indent && stack.size() > 1) syntax(line, "inconsistent indentation");
            if (stack.size() == 1 && stack.back().indent == -1) stack.back().indent = indent;
            if (stack.size() == 1 && stack.back().indent != indent) syntax(line, "inconsistent indentation");
        }
        auto child = Builder::scalar_node(unquote(val, line), line);
        Builder::add(*stack.back().node, key, child, line);
        last_child = child;
        last_indent = indent;
        last_empty = val.empty();
    }
    Node out = *root;
    return out;
}

Node Node::LoadFile(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw ParserException("yaml-cpp: bad file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return Load(ss.str());
}

namespace {
[[noreturn]] void bad(const Node &n, const char *type) {
    throw BadConversion("yaml-cpp: error at line " + std::to_string(n.line()) + ": bad conversion of '" +
                        n.Scalar() + "' to " + type);
}
}  // namespace

template <> double Node::as<double>() const {
    if (!IsScalar() || scalar_.empty()) bad(*this, "double");
    errno = 0;
    char *end = nullptr;
    const double v = std::strtod(scalar_.c_str(), &end);
    if (*end != '\0' || errno == ERANGE) bad(*this, "double");
    return v;
}

template <> float Node::as<float>() const { return (float)as<double>(); }

template <> long long Node::as<long long>() const {
    if (!IsScalar() || scalar_.empty()) bad(*this, "integer");
    errno = 0;
    char *end = nullptr;
    const long long v = std::strtoll(scalar_.c_str(), &end, 10);
    if (*end != '\0' || errno == ERANGE) bad(*this, "integer");
    return v;
}

template <> int Node::as<int>() const {
    const long long v = as<long long>();
    if (v < INT_MIN || v > INT_MAX) bad(*this, "int");
    return (int)v;
}

template <> unsigned long long Node::as<unsigned long long>() const {
    if (!IsScalar() || scalar_.empty() || scalar_[0] == '-') bad(*this, "unsigned integer");
    errno = 0;
    char *end = nullptr;
    const unsigned long long v = std::strtoull(scalar_.c_str(), &end, 10);
    if (*end != '\0' || errno == ERANGE) bad(*this, "unsigned integer");
    return v;
}

template <> unsigned long Node::as<unsigned long>() const { return (unsigned long)as<unsigned long long>(); }

template <> bool Node::as<bool>() const {
    // The YAML 1.1 boolean words that yaml-cpp accepts.
    static const char *T[] = {"y", "Y", "yes", "Yes", "YES", "true", "True", "TRUE", "on", "On", "ON"};
    static const char *F[] = {"n", "N", "no", "No", "NO", "false", "False", "FALSE", "off", "Off", "OFF"};
    if (IsScalar()) {
        for (const char *t : T)
            if (scalar_ == t) return true;
        for (const char *f : F)
            if (scalar_ == f) return false;
    }
    bad(*this, "bool");
}

template <> std::string Node::as<std::string>() const {
    if (!IsScalar()) bad(*this, "string");
    return scalar_;
}

}  // namespace quantkos::yaml
