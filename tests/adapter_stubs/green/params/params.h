// Stub: green-params' define / operator[] / as<T> surface.  Reading a name that was never defined
// throws, so the test fails if the adapter reads a flag nobody defines.
#pragma once
#include <any>
#include <map>
#include <stdexcept>
#include <string>

namespace green::params {
  class value {
  public:
    explicit value(const std::any* v, std::string name) : v_(v), name_(std::move(name)) {}
    template <typename T>
    T as() const {
      if (!v_) throw std::out_of_range("undefined parameter " + name_);
      return std::any_cast<T>(*v_);
    }
    template <typename T>
    operator T() const { return as<T>(); }

  private:
    const std::any* v_;
    std::string     name_;
  };

  class params {
  public:
    template <typename T>
    void define(const std::string& name, const std::string&, T def) { v_[name] = def; }
    template <typename T>
    void set(const std::string& name, T val) {
      if (!v_.count(name)) throw std::out_of_range("set of undefined parameter " + name);
      v_[name] = val;
    }
    bool  defined(const std::string& name) const { return v_.count(name) != 0; }
    value operator[](const std::string& name) const {
      auto it = v_.find(name);
      return value(it == v_.end() ? nullptr : &it->second, name);
    }

  private:
    std::map<std::string, std::any> v_;
  };
}  // namespace green::params
