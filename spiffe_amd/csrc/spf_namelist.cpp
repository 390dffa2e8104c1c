// spf_namelist.cpp -- namelist text reader (see spf_namelist.h).
#include "spf_namelist.h"

#include <ctype.h>
#include <errno.h>
#include <stdlib.h>
#include <string.h>

namespace spf {

static std::string lower(std::string s) {
  for (auto &c : s) c = (char)tolower((unsigned char)c);
  return s;
}

bool parse_double(const std::string &s0, double *v) {
  std::string s = s0;
  for (auto &c : s)
    if (c == 'd' || c == 'D') c = 'e';
  const char *p = s.c_str();
  char *end = nullptr;
  errno = 0;
  double x = strtod(p, &end);
  if (end == p) return false;
  while (*end && isspace((unsigned char)*end)) end++;
  if (*end) return false;
  *v = x;
  return true;
}

bool parse_long(const std::string &s, int64_t *v) {
  const char *p = s.c_str();
  char *end = nullptr;
  long long x = strtoll(p, &end, 10);
  if (end == p) return false;
  *v = (int64_t)x;
  return true;
}

const NamelistItem *Namelist::find(const char *key) const {
  for (const auto &it : items)
    if (it.key == key) return &it;
  return nullptr;
}

bool Namelist::get(const char *key, double *v, std::string &err) const {
  const NamelistItem *it = find(key);
  if (!it) return false;
  if (it->values.size() != 1 || !parse_double(it->values[0], v)) {
    err = "&" + group + ": bad value for " + key;
    return false;
  }
  return true;
}

bool Namelist::get(const char *key, int64_t *v, std::string &err) const {
  const NamelistItem *it = find(key);
  if (!it) return false;
  if (it->values.size() != 1 || !parse_long(it->values[0], v)) {
    err = "&" + group + ": bad value for " + key;
    return false;
  }
  return true;
}

bool Namelist::get(const char *key, int32_t *v, std::string &err) const {
  int64_t x;
  if (!get(key, &x, err)) return false;
  *v = (int32_t)x;
  return true;
}

bool Namelist::get(const char *key, int16_t *v, std::string &err) const {
  int64_t x;
  if (!get(key, &x, err)) return false;
  *v = (int16_t)x;
  return true;
}

bool Namelist::get(const char *key, std::string *v, std::string &err) const {
  const NamelistItem *it = find(key);
  if (!it) return false;
  if (it->values.size() != 1) {
    err = "&" + group + ": bad value for " + key;
    return false;
  }
  *v = it->values[0];
  return true;
}

std::string Namelist::unknown_keys(const char *const *known) const {
  std::string out;
  for (const auto &it : items) {
    bool ok = false;
    for (const char *const *k = known; *k; k++)
      if (it.key == *k) ok = true;
    if (!ok) out += (out.empty() ? "" : ", ") + it.key;
  }
  return out;
}

namespace {

enum TokKind { T_GROUP, T_END, T_WORD, T_STRING, T_EQ, T_COMMA };

struct Tok {
  TokKind kind;
  std::string text;
  int line;
};

int tokenize(const std::string &s, std::vector<Tok> &toks, std::string &err) {
  size_t i = 0, n = s.size();
  int line = 1;
  while (i < n) {
    char c = s[i];
    if (c == '\n') {
      line++;
      i++;
    } else if (isspace((unsigned char)c)) {
      i++;
    } else if (c == '!') {
      while (i < n && s[i] != '\n') i++;
    } else if (c == '&' || c == '$') {
      size_t j = i + 1;
      while (j < n && (isalnum((unsigned char)s[j]) || s[j] == '_')) j++;
      std::string name = lower(s.substr(i + 1, j - i - 1));
      // a bare '&' (or '$') closes a group as "&end" does
      toks.push_back({(name.empty() || name == "end") ? T_END : T_GROUP, name, line});
      i = j;
    } else if (c == '"') {
      std::string v;
      size_t j = i + 1;
      while (j < n && s[j] != '"') {
        if (s[j] == '\\' && j + 1 < n) j++;
        if (s[j] == '\n') line++;
        v += s[j++];
      }
      if (j >= n) return err = "line " + std::to_string(line) + ": unterminated string", 1;
      toks.push_back({T_STRING, v, line});
      i = j + 1;
    } else if (c == '=') {
      toks.push_back({T_EQ, "=", line});
      i++;
    } else if (c == ',') {
      toks.push_back({T_COMMA, ",", line});
      i++;
    } else {
      size_t j = i;
      while (j < n && !isspace((unsigned char)s[j]) && s[j] != '=' && s[j] != ',' && s[j] != '&' &&
             s[j] != '$' && s[j] != '!' && s[j] != '"')
        j++;
      toks.push_back({T_WORD, s.substr(i, j - i), line});
      i = j;
    }
  }
  return 0;
}

}  // namespace

int parse_namelists(const std::string &text, std::vector<Namelist> &out, std::string &err) {
  std::vector<Tok> t;
  if (tokenize(text, t, err)) return 1;
  size_t i = 0;
  while (i < t.size()) {
    if (t[i].kind != T_GROUP) return err = "line " + std::to_string(t[i].line) + ": text outside a namelist group", 1;
    Namelist nl;
    nl.group = t[i].text;
    nl.line = t[i].line;
    i++;
    for (;;) {
      if (i >= t.size()) return err = "&" + nl.group + " (line " + std::to_string(nl.line) + ") has no &end", 1;
      if (t[i].kind == T_END) {
        i++;
        break;
      }
      if (t[i].kind == T_COMMA) {
        i++;
        continue;
      }
      if (t[i].kind != T_WORD || i + 1 >= t.size() || t[i + 1].kind != T_EQ)
        return err = "line " + std::to_string(t[i].line) + ": expected name=value in &" + nl.group, 1;
      NamelistItem it;
      it.key = lower(t[i].text);
      i += 2;
      // values: one or more, comma separated, up to the next "name =" or &end
      for (;;) {
        if (i >= t.size() || (t[i].kind != T_WORD && t[i].kind != T_STRING))
          return err = "line " + std::to_string(i < t.size() ? t[i].line : nl.line) + ": missing value for " +
                       it.key,
                 1;
        it.values.push_back(t[i].text);
        it.quoted.push_back(t[i].kind == T_STRING);
        i++;
        if (i < t.size() && t[i].kind == T_COMMA) {
          size_t j = i + 1;
          if (j < t.size() && (t[j].kind == T_WORD || t[j].kind == T_STRING) &&
              !(j + 1 < t.size() && t[j + 1].kind == T_EQ)) {
            i = j;
            continue;
          }
        }
        break;
      }
      nl.items.push_back(it);
    }
    out.push_back(nl);
  }
  return 0;
}

}  // namespace spf
