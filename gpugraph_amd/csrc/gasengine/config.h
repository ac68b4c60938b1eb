// Config: the typed parameter database the reference's drivers and engine
// constructors take (Algorithms/SSSP/config.h:70-200, config.cpp:26-118).
//
// Same interface: parameters are registered once per process with a type,
// a description and a default (registerParameter<T>, called from the
// driver's registerParameters(), which the first Config constructed runs),
// read with getParameter<T>, written with setParameter<T> or parsed from
// "name=value name=value" strings (parseParameterString, the -p flag) and
// files of such lines (parseFile, the -c flag).  int, float, double and bool
// values parse from text.
//
// Error behaviour differs on purpose: the reference prints a message and
// then dereferences a missing entry or rethrows with no active exception
// (std::terminate); here an unknown name, a type mismatch or an unparsable
// value prints the same kind of message and throws std::invalid_argument.
#pragma once

#include <any>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <typeinfo>

// Registers the driver's parameters (the reference's register.h); defined by
// the program that constructs a Config.
void registerParameters();

class Config {
 public:
  Config() {
    if (!registered()) {
      registered() = true;
      registerParameters();
    }
  }

  template <typename Type>
  static void registerParameter(std::string name, std::string description, Type default_value) {
    descriptions()[name] = Description{&typeid(Type), name, description, std::any(default_value)};
  }

  template <typename Type>
  Type getParameter(std::string name) {
    const Description &d = lookup(name, typeid(Type), "getParameter");
    auto it = params_.find(name);
    return std::any_cast<Type>(it == params_.end() ? d.default_value : it->second);
  }

  template <typename Type>
  void setParameter(std::string name, Type value) {
    lookup(name, typeid(Type), "setParameter");
    params_[name] = std::any(value);
  }

  // "name=value" tokens separated by spaces or commas.
  void parseParameterString(const char *str) {
    std::string s(str ? str : "");
    size_t i = 0;
    while (i < s.size()) {
      while (i < s.size() && (s[i] == ' ' || s[i] == ',' || s[i] == '\t' || s[i] == '\r')) i++;
      size_t j = i;
      while (j < s.size() && s[j] != ' ' && s[j] != ',' && s[j] != '\t' && s[j] != '\r') j++;
      if (j > i) setFromText(s.substr(i, j - i));
      i = j;
    }
  }

  // One parameter string per line.
  void parseFile(const char *filename) {
    std::ifstream in(filename);
    if (!in) {
      std::printf("Error opening file '%s'\n", filename);
      throw std::invalid_argument(std::string("cannot open config file ") + filename);
    }
    std::string line;
    while (std::getline(in, line)) parseParameterString(line.c_str());
  }

  static void printOptions() {
    for (auto &kv : descriptions()) std::cout << "           " << kv.first << ": " << kv.second.description << "\n";
  }

  void printConfig() {
    for (auto &kv : descriptions()) {
      auto it = params_.find(kv.first);
      const std::any &v = it == params_.end() ? kv.second.default_value : it->second;
      std::cout << kv.first << "=" << text(v) << "\n";
    }
  }

 private:
  struct Description {
    const std::type_info *type = nullptr;
    std::string name, description;
    std::any default_value;
  };

  static std::map<std::string, Description> &descriptions() {
    static std::map<std::string, Description> d;
    return d;
  }
  static bool &registered() {
    static bool r = false;
    return r;
  }

  const Description &lookup(const std::string &name, const std::type_info &t, const char *what) {
    auto it = descriptions().find(name);
    if (it == descriptions().end()) {
      std::printf("%s error: '%s' not found\n", what, name.c_str());
      throw std::invalid_argument("parameter '" + name + "' is not registered");
    }
    if (*it->second.type != t) {
      std::printf("%s error: '%s' type miss match\n", what, name.c_str());
      throw std::invalid_argument("parameter '" + name + "' has another type");
    }
    return it->second;
  }

  void setFromText(const std::string &tok) {
    const size_t eq = tok.find('=');
    const std::string name = tok.substr(0, eq);
    const std::string value = eq == std::string::npos ? std::string() : tok.substr(eq + 1);
    auto it = descriptions().find(name);
    if (it == descriptions().end()) {
      std::printf("Variable '%s' not registered\n", name.c_str());
      throw std::invalid_argument("parameter '" + name + "' is not registered");
    }
    const std::type_info &t = *it->second.type;
    char *end = nullptr;
    const char *c = value.c_str();
    if (t == typeid(int)) {
      long v = std::strtol(c, &end, 0);
      check(end, c, name);
      params_[name] = std::any((int)v);
    } else if (t == typeid(float)) {
      float v = std::strtof(c, &end);
      check(end, c, name);
      params_[name] = std::any(v);
    } else if (t == typeid(double)) {
      double v = std::strtod(c, &end);
      check(end, c, name);
      params_[name] = std::any(v);
    } else if (t == typeid(bool)) {
      if (value == "1" || value == "true")
        params_[name] = std::any(true);
      else if (value == "0" || value == "false")
        params_[name] = std::any(false);
      else
        check(c, c, name);
    } else if (t == typeid(std::string)) {
      params_[name] = std::any(value);
    } else {
      std::printf("getValue is not implemented for the datatype of variable '%s'\n", name.c_str());
      throw std::invalid_argument("parameter '" + name + "' has a type text cannot set");
    }
  }

  static void check(const char *end, const char *begin, const std::string &name) {
    if (end == begin || *end != '\0') {
      std::printf("cannot parse the value of '%s'\n", name.c_str());
      throw std::invalid_argument("bad value for parameter '" + name + "'");
    }
  }

  static std::string text(const std::any &v) {
    if (v.type() == typeid(int)) return std::to_string(std::any_cast<int>(v));
    if (v.type() == typeid(float)) return std::to_string(std::any_cast<float>(v));
    if (v.type() == typeid(double)) return std::to_string(std::any_cast<double>(v));
    if (v.type() == typeid(bool)) return std::any_cast<bool>(v) ? "1" : "0";
    if (v.type() == typeid(std::string)) return std::any_cast<std::string>(v);
    return "?";
  }

  std::map<std::string, std::any> params_;
};
