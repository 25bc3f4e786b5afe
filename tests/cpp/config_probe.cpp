// config_probe.cpp -- prints what Config (include/nadia/Config.h) returns for
// each query, for tests/test_config.py.  Queries: s:<key> getString,
// d:<key> getDouble, i:<key> getInt, S:<key> getStringList,
// I:<key> getIntList; "status" prints getStatus().
#include <cstdio>
#include <iostream>
#include <list>
#include <sstream>
#include <string>

#include "Config.h"

int main(int argc, char* argv[]) {
  Config c(argv[1]);
  for (int a = 2; a < argc; a++) {
    const std::string q = argv[a];
    if (q == "status") {
      std::cout << "status " << c.getStatus() << std::endl;
      continue;
    }
    const char kind = q[0];
    const std::string key = q.substr(2);
    // the call prints its own messages first; the answer line follows
    std::ostringstream r;
    if (kind == 's') {
      r << " [" << c.getString(key) << "]";
    } else if (kind == 'd') {
      char b[64];
      std::snprintf(b, sizeof b, " %.17g", c.getDouble(key));
      r << b;
    } else if (kind == 'i') {
      r << " " << c.getInt(key);
    } else if (kind == 'S') {
      std::list<std::string>* l = c.getStringList(key);
      if (!l) r << " null";
      else
        for (std::list<std::string>::iterator it = l->begin(); it != l->end(); ++it) r << " [" << *it << "]";
      delete l;
    } else if (kind == 'I') {
      std::list<int>* l = c.getIntList(key);
      if (!l) r << " null";
      else
        for (std::list<int>::iterator it = l->begin(); it != l->end(); ++it) r << " " << *it;
      delete l;
    }
    std::cout << "= " << kind << " " << key << " ->" << r.str() << std::endl;
  }
  return 0;
}
