// parameters.h -- string-keyed run parameters with typed access.
//
// Same interface and error behaviour as include/weavess/parameters.h:12-57
// of the reference: set<T> stores the streamed text, get<T> throws
// std::invalid_argument for a missing key and std::runtime_error when the
// stored text does not convert completely to T.
#pragma once

#include <sstream>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <unordered_map>

namespace weavess {

class Parameters {
public:
    template <typename T>
    void set(const std::string &name, const T &val) {
        std::ostringstream os;
        os << val;
        params_[name] = os.str();
    }

    template <typename T>
    T get(const std::string &name) const {
        auto it = params_.find(name);
        if (it == params_.end()) throw std::invalid_argument("Invalid paramter name : " + name + ".");
        std::istringstream is(it->second);
        T v;
        if (!(is >> v) || !is.eof())
            throw std::runtime_error("Fail to convert value" + it->second + " to type: " + typeid(v).name());
        return v;
    }

    bool has(const std::string &name) const { return params_.count(name) != 0; }

    std::string toString() const {
        std::string s;
        for (const auto &p : params_) s += p.first + ":" + p.second + " ";
        return s;
    }

private:
    std::unordered_map<std::string, std::string> params_;
};

// Strings are stored verbatim (a stream read would stop at whitespace).
template <>
inline std::string Parameters::get<std::string>(const std::string &name) const {
    auto it = params_.find(name);
    if (it == params_.end()) throw std::invalid_argument("Invalid paramter name : " + name + ".");
    return it->second;
}

}  // namespace weavess
