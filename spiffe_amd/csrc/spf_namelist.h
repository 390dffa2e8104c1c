// spf_namelist.h -- reader for the namelist text spiffe's decks and boundary
// files are written in: groups "&name key=value, key=value ... &end" (a bare
// '&' also ends a group, and '$' may stand for '&'), values
// numbers, quoted strings or bare words (a key may take a comma-separated
// list), '!' starting a comment that runs to the end of the line.  Group and
// key names compare case-insensitively.  Plain C++, no HIP.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace spf {

struct NamelistItem {
  std::string key;                  // lower case
  std::vector<std::string> values;  // as written (quotes removed)
  std::vector<bool> quoted;
};

struct Namelist {
  std::string group;                // lower case, without '&'
  std::vector<NamelistItem> items;
  int line = 0;                     // line of the group's '&'

  const NamelistItem *find(const char *key) const;
  // typed reads: return false (value untouched) when the key is absent and set
  // err when the value does not parse
  bool get(const char *key, double *v, std::string &err) const;
  bool get(const char *key, int64_t *v, std::string &err) const;
  bool get(const char *key, int32_t *v, std::string &err) const;
  bool get(const char *key, int16_t *v, std::string &err) const;
  bool get(const char *key, std::string *v, std::string &err) const;
  // keys present that are not in `known` (a null-terminated list)
  std::string unknown_keys(const char *const *known) const;
};

// Parses every group of `text`; returns 0 or nonzero with a message.
int parse_namelists(const std::string &text, std::vector<Namelist> &out, std::string &err);

// Number conversions used for namelist values: a double as strtod reads it
// (Fortran 'd' exponents accepted), an integer as strtol reads its leading part.
bool parse_double(const std::string &s, double *v);
bool parse_long(const std::string &s, int64_t *v);

}  // namespace spf
