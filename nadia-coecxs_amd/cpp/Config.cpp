// Config.cpp -- NADIA's key/value configuration reader (NADIA/src/Config.c++:22-161).
// Host-only; builds with g++ alone.
#include "Config.h"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

namespace {

// whitespace-delimited words of s
std::vector<std::string> words(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream in(s);
  std::string w;
  while (in >> w) out.push_back(w);
  return out;
}

}  // namespace

// NADIA/src/Config.c++:22-66
Config::Config(std::string file_name) : status(FAILURE) {
  std::ifstream file(file_name.c_str());
  if (!file.is_open()) {
    std::cout << "Could not open the file " << file_name << std::endl;
    return;
  }
  std::string line;
  while (std::getline(file, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    size_t sep = line.find('=');
    if (sep == std::string::npos) sep = line.find(' ');
    if (sep == std::string::npos) continue;
    const std::vector<std::string> k = words(line.substr(0, sep));
    // std::map::insert keeps an existing entry: the first occurrence wins
    mapping.insert(MapType::value_type(k.empty() ? std::string() : k[0], line.substr(sep + 1)));
  }
  status = SUCCESS;
}

const std::string* Config::lookup(const std::string& key) {
  MapType::const_iterator it = mapping.find(key);
  if (it == mapping.end()) {
    std::cout << "Key:" << key << " not found in config file" << std::endl;
    status = FAILURE;
    return 0;
  }
  return &it->second;
}

// NADIA/src/Config.c++:70-82
std::string Config::getString(std::string key) {
  const std::string* v = lookup(key);
  if (!v) return "";
  const std::vector<std::string> w = words(*v);
  return w.empty() ? std::string() : w[0];
}

// NADIA/src/Config.c++:84-107
std::list<std::string>* Config::getStringList(std::string key) {
  const std::string* v = lookup(key);
  if (!v) return 0;
  const std::vector<std::string> w = words(*v);
  if (w.empty()) {
    std::cout << "Key:" << key << " has no entries in the config file" << std::endl;
    status = FAILURE;
    return 0;
  }
  return new std::list<std::string>(w.begin(), w.end());
}

// NADIA/src/Config.c++:109-130
std::list<int>* Config::getIntList(std::string key) {
  const std::string* v = lookup(key);
  if (!v) return 0;
  const std::vector<std::string> w = words(*v);
  if (w.empty()) {
    std::cout << "Key:" << key << " has no entries in the config file" << std::endl;
    status = FAILURE;
    return 0;
  }
  std::list<int>* out = new std::list<int>;
  for (size_t i = 0; i < w.size(); i++) out->push_back(std::atoi(w[i].c_str()));
  return out;
}

// NADIA/src/Config.c++:132-143
double Config::getDouble(std::string key) {
  const std::string* v = lookup(key);
  return v ? std::atof(v->c_str()) : 0;
}

// NADIA/src/Config.c++:145-156
int Config::getInt(std::string key) {
  const std::string* v = lookup(key);
  return v ? std::atoi(v->c_str()) : 0;
}
