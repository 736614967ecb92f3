// yaml_lite.hpp -- a YAML subset reader for QuantKos configuration files.
//
// The reference reads its configuration with yaml-cpp 0.8.0
// (src/IO/ConfigParser.cpp:51-56); that library is not part of this build.
// The files it reads use only: nested block mappings by indentation, plain or
// quoted scalars, '#' comments, blank lines, and "key : value" with optional
// spaces before the colon.  This reader accepts exactly that subset and
// rejects the rest (tabs, flow collections, sequences, anchors, multi-line
// scalars) with a syntax error, as yaml-cpp would for malformed input.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace quantkos::yaml {

struct ParserException : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct BadConversion : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Node {
public:
    Node() = default;

    // Child lookup; a missing key gives an undefined node (operator bool false).
    const Node &operator[](const std::string &key) const;
    bool IsDefined() const { return defined_; }
    bool IsMap() const { return defined_ && is_map_; }
    bool IsScalar() const { return defined_ && !is_map_; }
    explicit operator bool() const { return defined_; }
    const std::string &Scalar() const { return scalar_; }
    const std::vector<std::string> &Keys() const { return order_; }
    int line() const { return line_; }

    template <typename T>
    T as() const;

    // Parses a whole document.
    static Node Load(const std::string &text);
    static Node LoadFile(const std::string &path);

private:
    bool defined_ = false;
    bool is_map_ = false;
    std::string scalar_;
    int line_ = 0;
    std::map<std::string, std::shared_ptr<Node>> children_;
    std::vector<std::string> order_;
    friend struct Builder;
};

template <> double Node::as<double>() const;
template <> float Node::as<float>() const;
template <> int Node::as<int>() const;
template <> long long Node::as<long long>() const;
template <> unsigned long Node::as<unsigned long>() const;
template <> unsigned long long Node::as<unsigned long long>() const;
template <> bool Node::as<bool>() const;
template <> std::string Node::as<std::string>() const;

}  // namespace quantkos::yaml
