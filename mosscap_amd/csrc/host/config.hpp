// Configuration access with the reference's semantics (MosscapConfig.hpp:62-123):
// dotted-path lookup with a default, conversion errors as std::runtime_error
// naming the key, the target type and the value, and case-insensitive enum
// lookup in a name table.
#pragma once

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "yaml_lite.hpp"

namespace mh::host {

template <class T>
constexpr const char* type_label() {
    if constexpr (std::is_same_v<T, bool>) return "bool";
    else if constexpr (std::is_same_v<T, int>) return "int";
    else if constexpr (std::is_same_v<T, long long>) return "long long";
    else if constexpr (std::is_same_v<T, float>) return "float";
    else if constexpr (std::is_same_v<T, double>) return "double";
    else return "std::string";
}

// Scalar conversion following yaml-cpp's rules for the types used by the
// configuration: integers must be whole decimal numbers, reals accept any
// strtod form plus .inf/.nan, bools accept true/false/yes/no/on/off/y/n.
template <class T>
bool convert_scalar(const std::string& s, T& out) {
    if constexpr (std::is_same_v<T, std::string>) {
        out = s;
        return true;
    } else if constexpr (std::is_same_v<T, bool>) {
        std::string l = s;
        std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (l == "true" || l == "yes" || l == "on" || l == "y") {
            out = true;
            return true;
        }
        if (l == "false" || l == "no" || l == "off" || l == "n") {
            out = false;
            return true;
        }
        return false;
    } else if constexpr (std::is_integral_v<T>) {
        if (s.empty()) return false;
        char* end = nullptr;
        errno = 0;
        const long long v = std::strtoll(s.c_str(), &end, 10);
        if (errno != 0 || *end != '\0') return false;
        if constexpr (std::is_same_v<T, int>) {
            if (v < -2147483648LL || v > 2147483647LL) return false;
        }
        out = (T)v;
        return true;
    } else {
        std::string l = s;
        std::transform(l.begin(), l.end(), l.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        if (l == ".inf" || l == "+.inf") {
            out = (T)INFINITY;
            return true;
        }
        if (l == "-.inf") {
            out = (T)-INFINITY;
            return true;
        }
        if (l == ".nan") {
            out = (T)NAN;
            return true;
        }
        if (s.empty()) return false;
        char* end = nullptr;
        const double v = std::strtod(s.c_str(), &end);
        if (*end != '\0') return false;
        out = (T)v;
        return true;
    }
}

// get_or (MosscapConfig.hpp:62-88): the value at "a.b.c", or `default_val`
// when any component is absent.
template <class T>
T get_or(const YNode* f, const std::string& id, const T& default_val) {
    if (!f) return default_val;
    const size_t pos = id.find('.');
    if (pos != std::string::npos) {
        const YNode* child = f->get(id.substr(0, pos));
        if (child && !child->is_null()) return get_or(child, id.substr(pos + 1), default_val);
        return default_val;
    }
    const YNode* v = f->get(id);
    if (!v || v->is_null()) return default_val;
    T out{};
    if (v->kind != YNode::Scalar || !convert_scalar(v->scalar, out)) {
        throw std::runtime_error("Unable to convert key \"" + id + "\" to type \"" + type_label<T>() +
                                 "\". Has value: \"" + (v->kind == YNode::Scalar ? v->scalar : std::string("<non-scalar>")) +
                                 "\".");
    }
    return out;
}

template <class T>
T get_or(const std::shared_ptr<YNode>& f, const std::string& id, const T& default_val) {
    return get_or<T>(f.get(), id, default_val);
}

// Child node at a dotted path, or nullptr.
inline const YNode* find_node(const YNode* f, const std::string& id) {
    const YNode* n = f;
    size_t start = 0;
    while (n) {
        const size_t pos = id.find('.', start);
        n = n->get(id.substr(start, pos == std::string::npos ? std::string::npos : pos - start));
        if (pos == std::string::npos) break;
        start = pos + 1;
    }
    return n;
}

// find_associated_enum (MosscapConfig.hpp:108-123): lower-cases the name and
// returns its index in `names`.
template <class T>
T find_associated_enum(const char* const* names, int num_names, std::string to_compare, const char* type_name) {
    std::transform(to_compare.begin(), to_compare.end(), to_compare.begin(),
                   [](unsigned char c) { return (char)std::tolower(c); });
    for (int i = 0; i < num_names; ++i)
        if (to_compare == names[i]) return T(i);
    throw std::runtime_error(std::string("Unable to find associated key in \"") + type_name + "\" for \"" +
                             to_compare + "\".");
}

// Name tables, in enum order (Reconstruct.hpp:14-31, Riemann.hpp:15-19,
// TimeStepping.hpp:15-19, State.hpp:71-79).
inline constexpr const char* ReconstructionName[] = {"fog", "muscl", "ppm", "weno5z"};
inline constexpr const char* SlopeLimiterName[] = {"vanleer", "monotonizedcentral", "minmod"};
inline constexpr const char* RiemannSolverName[] = {"rusanov", "hll", "hllc"};
inline constexpr const char* TimeStepName[] = {"rk2", "ssprk3", "ssprk4"};
inline constexpr const char* BoundaryTypeName[] = {"wall",     "periodic", "symmetric", "symmetricoutflowdiode",
                                                   "zerograd", "constant", "userfn"};

}  // namespace mh::host
