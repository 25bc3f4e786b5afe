// Config.h -- the key/value configuration file reader of the NADIA drop-in API.
//
// Interface and behaviour of NADIA's Config (NADIA/include/Config.h,
// NADIA/src/Config.c++:22-161), the reader behind the reconstruction tool:
//   * '#' starts a comment;
//   * a line holds "key = value" or, without '=', "key value" (split at the
//     first space); the key is the first whitespace-delimited word before the
//     separator, the value everything after it;
//   * the first occurrence of a key wins;
//   * getString returns the first word of the value; getStringList /
//     getIntList every word (new lists, owned by the caller); getDouble /
//     getInt convert the whole value with atof / atoi;
//   * a missing key prints "Key:<key> not found in config file", sets the
//     status to FAILURE and returns "" / 0 / a null list; an empty list
//     prints "Key:<key> has no entries in the config file" likewise.
// Host code only; it does not touch the device.
#ifndef NADIA_AMD_CONFIG_H
#define NADIA_AMD_CONFIG_H

#include <list>
#include <map>
#include <string>

#ifndef FAILURE
#define FAILURE 0
#endif
#ifndef SUCCESS
#define SUCCESS 1
#endif

typedef std::map<std::string, std::string> MapType;

class Config {
 private:
  MapType mapping;
  int status;

  // the stored value of `key`, or null after reporting it missing
  const std::string* lookup(const std::string& key);

 public:
  Config(std::string file_name);

  std::string getString(std::string key);
  double getDouble(std::string key);
  int getInt(std::string key);
  std::list<int>* getIntList(std::string key);
  std::list<std::string>* getStringList(std::string key);

  int getStatus() { return status; }
};

#endif
