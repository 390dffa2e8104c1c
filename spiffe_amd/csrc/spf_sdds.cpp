// spf_sdds.cpp -- SDDS reader and writer for the subset in spf_sdds.h.
#include "spf_sdds.h"

#include <ctype.h>
#include <math.h>
#include <string.h>

#include "spf_namelist.h"

namespace spf {

namespace {

bool type_of(const std::string &t, SddsType *out) {
  if (t == "double") *out = SDDS_T_DOUBLE;
  else if (t == "float") *out = SDDS_T_FLOAT;
  else if (t == "long" || t == "ulong") *out = SDDS_T_LONG;
  else if (t == "short" || t == "ushort") *out = SDDS_T_SHORT;
  else if (t == "character") *out = SDDS_T_CHARACTER;
  else if (t == "string") *out = SDDS_T_STRING;
  else return false;
  return true;
}

const char *type_name(SddsType t) {
  switch (t) {
    case SDDS_T_DOUBLE: return "double";
    case SDDS_T_FLOAT: return "float";
    case SDDS_T_LONG: return "long";
    case SDDS_T_SHORT: return "short";
    case SDDS_T_CHARACTER: return "character";
    default: return "string";
  }
}

std::string quote(const std::string &s) {
  bool need = s.empty();
  for (char c : s)
    if (isspace((unsigned char)c) || c == ',' || c == '&' || c == '"' || c == '!' || c == '$') need = true;
  if (!need) return s;
  std::string q = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') q += '\\';
    q += c;
  }
  return q + "\"";
}

// tokens of one ASCII data line: whitespace separated, double quotes group
std::vector<std::string> split_line(const std::string &line) {
  std::vector<std::string> out;
  size_t i = 0, n = line.size();
  while (i < n) {
    while (i < n && isspace((unsigned char)line[i])) i++;
    if (i >= n) break;
    std::string tok;
    if (line[i] == '"') {
      i++;
      while (i < n && line[i] != '"') {
        if (line[i] == '\\' && i + 1 < n) i++;
        tok += line[i++];
      }
      i++;
    } else {
      while (i < n && !isspace((unsigned char)line[i])) tok += line[i++];
    }
    out.push_back(tok);
  }
  return out;
}

bool read_value(const std::string &tok, SddsType t, SddsColumn &c) {
  if (t == SDDS_T_STRING || t == SDDS_T_CHARACTER) {
    c.str.push_back(tok);
    return true;
  }
  double v;
  if (!parse_double(tok, &v)) return false;
  c.num.push_back(v);
  return true;
}

struct Cursor {
  const std::string &b;
  size_t p;
  bool take(void *dst, size_t n) {
    if (p + n > b.size()) return false;
    memcpy(dst, b.data() + p, n);
    p += n;
    return true;
  }
};

bool read_binary_value(Cursor &cur, SddsType t, double *num, std::string *str) {
  switch (t) {
    case SDDS_T_DOUBLE: {
      double v;
      if (!cur.take(&v, 8)) return false;
      *num = v;
      return true;
    }
    case SDDS_T_FLOAT: {
      float v;
      if (!cur.take(&v, 4)) return false;
      *num = v;
      return true;
    }
    case SDDS_T_LONG: {
      int32_t v;
      if (!cur.take(&v, 4)) return false;
      *num = v;
      return true;
    }
    case SDDS_T_SHORT: {
      int16_t v;
      if (!cur.take(&v, 2)) return false;
      *num = v;
      return true;
    }
    case SDDS_T_CHARACTER: {
      char v;
      if (!cur.take(&v, 1)) return false;
      *str = std::string(1, v);
      return true;
    }
    default: {
      int32_t len;
      if (!cur.take(&len, 4) || len < 0 || cur.p + (size_t)len > cur.b.size()) return false;
      str->assign(cur.b.data() + cur.p, (size_t)len);
      cur.p += (size_t)len;
      return true;
    }
  }
}

}  // namespace

int SddsFile::param_index(const char *name) const {
  for (size_t i = 0; i < params.size(); i++)
    if (params[i].name == name) return (int)i;
  return -1;
}

int SddsFile::column_index(const char *name) const {
  for (size_t i = 0; i < columns.size(); i++)
    if (columns[i].name == name) return (int)i;
  return -1;
}

int sdds_read(const std::string &path, SddsFile &out, std::string &err) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) return err = "cannot open " + path, 1;
  std::string buf;
  char chunk[1 << 16];
  size_t n;
  while ((n = fread(chunk, 1, sizeof(chunk), fp)) > 0) buf.append(chunk, n);
  fclose(fp);
  if (buf.compare(0, 5, "SDDS1") != 0 && buf.compare(0, 5, "SDDS2") != 0 && buf.compare(0, 5, "SDDS3") != 0 &&
      buf.compare(0, 5, "SDDS4") != 0 && buf.compare(0, 5, "SDDS5") != 0)
    return err = path + " is not an SDDS file", 1;
  // header: lines up to and including the one that closes &data
  size_t pos = buf.find('\n');
  if (pos == std::string::npos) return err = path + ": no header", 1;
  pos++;
  std::string header;
  bool in_data = false;
  size_t data_start = std::string::npos;
  while (pos < buf.size()) {
    size_t e = buf.find('\n', pos);
    if (e == std::string::npos) e = buf.size();
    std::string line = buf.substr(pos, e - pos);
    pos = e + 1;
    size_t k = line.find_first_not_of(" \t\r");
    if (k != std::string::npos && line[k] == '!') continue;
    header += line + "\n";
    if (line.find("&data") != std::string::npos) in_data = true;
    if (in_data && line.find("&end") != std::string::npos) {
      data_start = pos;
      break;
    }
  }
  if (data_start == std::string::npos) return err = path + ": no &data command", 1;
  std::vector<Namelist> groups;
  if (parse_namelists(header, groups, err)) return err = path + ": " + err, 1;
  bool no_row_counts = false, column_major = false;
  for (const auto &g : groups) {
    std::string e2;
    if (g.group == "description") {
      g.get("text", &out.description, e2);
    } else if (g.group == "parameter" || g.group == "column") {
      SddsField f;
      std::string t = "double";
      g.get("name", &f.name, e2);
      g.get("units", &f.units, e2);
      g.get("description", &f.description, e2);
      g.get("type", &t, e2);
      if (!type_of(t, &f.type)) return err = path + ": unsupported type " + t, 1;
      if (g.group == "parameter") {
        g.get("fixed_value", &f.fixed_value, e2);
        if (g.find("fixed_value") && f.fixed_value.empty()) f.fixed_value = " ";
        out.params.push_back(f);
      } else {
        out.columns.push_back(f);
      }
    } else if (g.group == "data") {
      std::string mode = "binary";
      int64_t v = 0;
      g.get("mode", &mode, e2);
      out.binary = mode == "binary";
      if (!out.binary && mode != "ascii") return err = path + ": unknown data mode " + mode, 1;
      if (g.get("no_row_counts", &v, e2)) no_row_counts = v != 0;
      v = 0;
      if (g.get("column_major_order", &v, e2)) column_major = v != 0;
      v = 0;
      if (g.get("additional_header_lines", &v, e2) && v) return err = path + ": additional_header_lines unsupported", 1;
    } else if (g.group == "array" || g.group == "include" || g.group == "associate") {
      return err = path + ": &" + g.group + " is not supported", 1;
    }
  }
  auto start_page = [&](SddsPage &pg) {
    pg.cols.assign(out.columns.size(), SddsColumn());
    pg.par_str.assign(out.params.size(), "");
    pg.par_num.assign(out.params.size(), NAN);
    for (size_t i = 0; i < out.params.size(); i++)
      if (!out.params[i].fixed_value.empty()) {
        pg.par_str[i] = out.params[i].fixed_value;
        double v;
        if (parse_double(pg.par_str[i], &v)) pg.par_num[i] = v;
      }
  };
  if (out.binary) {
    Cursor cur{buf, data_start};
    while (cur.b.size() - cur.p >= 4) {
      int32_t rows;
      cur.take(&rows, 4);
      if (rows < 0) return err = path + ": negative row count", 1;
      SddsPage pg;
      start_page(pg);
      pg.rows = rows;
      for (size_t i = 0; i < out.params.size(); i++) {
        if (!out.params[i].fixed_value.empty()) continue;
        double v = NAN;
        std::string s;
        if (!read_binary_value(cur, out.params[i].type, &v, &s)) return err = path + ": truncated page", 1;
        pg.par_num[i] = v;
        pg.par_str[i] = s.empty() && !isnan(v) ? std::to_string(v) : s;
      }
      auto put = [&](size_t j) -> bool {
        double v = NAN;
        std::string s;
        if (!read_binary_value(cur, out.columns[j].type, &v, &s)) return false;
        if (out.columns[j].type == SDDS_T_STRING || out.columns[j].type == SDDS_T_CHARACTER)
          pg.cols[j].str.push_back(s);
        else
          pg.cols[j].num.push_back(v);
        return true;
      };
      if (column_major) {
        for (size_t j = 0; j < out.columns.size(); j++)
          for (int32_t r = 0; r < rows; r++)
            if (!put(j)) return err = path + ": truncated page", 1;
      } else {
        for (int32_t r = 0; r < rows; r++)
          for (size_t j = 0; j < out.columns.size(); j++)
            if (!put(j)) return err = path + ": truncated page", 1;
      }
      out.pages.push_back(std::move(pg));
    }
    return 0;
  }
  // ASCII: parameters one per line, then the row count (unless
  // no_row_counts), then one row per line; a blank line ends a page written
  // without row counts
  std::vector<std::string> lines;
  {
    size_t p = data_start;
    while (p < buf.size()) {
      size_t e = buf.find('\n', p);
      if (e == std::string::npos) e = buf.size();
      std::string l = buf.substr(p, e - p);
      if (!l.empty() && l.back() == '\r') l.pop_back();
      p = e + 1;
      size_t k = l.find_first_not_of(" \t");
      if (k != std::string::npos && l[k] == '!') continue;
      lines.push_back(l);
    }
  }
  size_t li = 0;
  auto blank = [](const std::string &l) { return l.find_first_not_of(" \t") == std::string::npos; };
  for (;;) {
    while (li < lines.size() && blank(lines[li])) li++;
    if (li >= lines.size()) break;
    SddsPage pg;
    start_page(pg);
    for (size_t i = 0; i < out.params.size(); i++) {
      if (!out.params[i].fixed_value.empty()) continue;
      if (li >= lines.size()) return err = path + ": truncated page", 1;
      std::string l = lines[li++];
      std::vector<std::string> t = split_line(l);
      std::string v = t.empty() ? "" : (out.params[i].type == SDDS_T_STRING ? (t.size() == 1 ? t[0] : l) : t[0]);
      pg.par_str[i] = v;
      double d;
      if (out.params[i].type != SDDS_T_STRING && out.params[i].type != SDDS_T_CHARACTER) {
        if (!parse_double(v, &d)) return err = path + ": bad parameter value " + v, 1;
        pg.par_num[i] = d;
      }
    }
    int64_t rows = -1;
    if (!no_row_counts) {
      while (li < lines.size() && blank(lines[li])) li++;
      if (li >= lines.size()) return err = path + ": missing row count", 1;
      if (!parse_long(split_line(lines[li++])[0], &rows) || rows < 0) return err = path + ": bad row count", 1;
    }
    int64_t r = 0;
    while ((rows < 0 || r < rows) && li < lines.size()) {
      if (blank(lines[li])) {
        if (rows < 0) break;
        li++;
        continue;
      }
      std::vector<std::string> t = split_line(lines[li++]);
      if (t.size() < out.columns.size()) return err = path + ": short row", 1;
      for (size_t j = 0; j < out.columns.size(); j++)
        if (!read_value(t[j], out.columns[j].type, pg.cols[j])) return err = path + ": bad value " + t[j], 1;
      r++;
    }
    if (rows >= 0 && r < rows) return err = path + ": truncated page", 1;
    pg.rows = r;
    out.pages.push_back(std::move(pg));
  }
  return 0;
}

SddsWriter::~SddsWriter() { close(); }

void SddsWriter::close() {
  if (fp_) fclose(fp_);
  fp_ = nullptr;
}

int SddsWriter::open(const std::string &path, const std::string &description, const std::vector<SddsField> &params,
                     const std::vector<SddsField> &columns, bool binary, std::string &err) {
  close();
  fp_ = fopen(path.c_str(), "wb");
  if (!fp_) return err = "cannot write " + path, 1;
  binary_ = binary;
  params_ = params;
  columns_ = columns;
  fputs("SDDS1\n", fp_);
  if (binary) {
    uint16_t one = 1;
    fputs(*(const uint8_t *)&one ? "!# little-endian\n" : "!# big-endian\n", fp_);
  }
  if (!description.empty()) fprintf(fp_, "&description text=%s, &end\n", quote(description).c_str());
  for (const auto &p : params) {
    fprintf(fp_, "&parameter name=%s, ", quote(p.name).c_str());
    if (!p.units.empty()) fprintf(fp_, "units=%s, ", quote(p.units).c_str());
    if (!p.description.empty()) fprintf(fp_, "description=%s, ", quote(p.description).c_str());
    fprintf(fp_, "type=%s, ", type_name(p.type));
    if (!p.fixed_value.empty()) fprintf(fp_, "fixed_value=%s, ", quote(p.fixed_value).c_str());
    fputs("&end\n", fp_);
  }
  for (const auto &c : columns) {
    fprintf(fp_, "&column name=%s, ", quote(c.name).c_str());
    if (!c.units.empty()) fprintf(fp_, "units=%s, ", quote(c.units).c_str());
    if (!c.description.empty()) fprintf(fp_, "description=%s, ", quote(c.description).c_str());
    fprintf(fp_, "type=%s, &end\n", type_name(c.type));
  }
  fprintf(fp_, "&data mode=%s, &end\n", binary ? "binary" : "ascii");
  return ferror(fp_) ? (err = "write error on " + path, 1) : 0;
}

int SddsWriter::write_page(const std::vector<std::string> &par, int64_t rows, const std::vector<const void *> &cols,
                           std::string &err) {
  if (!fp_) return err = "SDDS file not open", 1;
  if (cols.size() != columns_.size()) return err = "column count mismatch", 1;
  if (rows > INT32_MAX) return err = "too many rows for one page", 1;
  size_t k = 0;
  std::vector<std::string> pv;
  for (const auto &p : params_)
    if (p.fixed_value.empty()) {
      if (k >= par.size()) return err = "missing parameter value for " + p.name, 1;
      pv.push_back(par[k++]);
    }
  if (binary_) {
    int32_t r = (int32_t)rows;
    fwrite(&r, 4, 1, fp_);
    size_t q = 0;
    for (const auto &p : params_) {
      if (!p.fixed_value.empty()) continue;
      const std::string &s = pv[q++];
      double v = strtod(s.c_str(), nullptr);
      switch (p.type) {
        case SDDS_T_DOUBLE: fwrite(&v, 8, 1, fp_); break;
        case SDDS_T_FLOAT: { float f = (float)v; fwrite(&f, 4, 1, fp_); break; }
        case SDDS_T_LONG: { int32_t i = (int32_t)strtoll(s.c_str(), nullptr, 10); fwrite(&i, 4, 1, fp_); break; }
        case SDDS_T_SHORT: { int16_t i = (int16_t)strtol(s.c_str(), nullptr, 10); fwrite(&i, 2, 1, fp_); break; }
        case SDDS_T_CHARACTER: { char ch = s.empty() ? ' ' : s[0]; fwrite(&ch, 1, 1, fp_); break; }
        default: { int32_t len = (int32_t)s.size(); fwrite(&len, 4, 1, fp_); fwrite(s.data(), 1, s.size(), fp_); }
      }
    }
    for (int64_t i = 0; i < rows; i++)
      for (size_t j = 0; j < columns_.size(); j++) {
        if (columns_[j].type == SDDS_T_STRING) {
          const std::string &s = ((const std::string *)cols[j])[i];
          int32_t len = (int32_t)s.size();
          fwrite(&len, 4, 1, fp_);
          fwrite(s.data(), 1, s.size(), fp_);
          continue;
        }
        double v = ((const double *)cols[j])[i];
        switch (columns_[j].type) {
          case SDDS_T_DOUBLE: fwrite(&v, 8, 1, fp_); break;
          case SDDS_T_FLOAT: { float f = (float)v; fwrite(&f, 4, 1, fp_); break; }
          case SDDS_T_LONG: { int32_t x = (int32_t)v; fwrite(&x, 4, 1, fp_); break; }
          case SDDS_T_SHORT: { int16_t x = (int16_t)v; fwrite(&x, 2, 1, fp_); break; }
          default: { char ch = (char)v; fwrite(&ch, 1, 1, fp_); }
        }
      }
  } else {
    for (const auto &s : pv) fprintf(fp_, "%s\n", quote(s).c_str());
    fprintf(fp_, "%lld\n", (long long)rows);
    for (int64_t i = 0; i < rows; i++) {
      for (size_t j = 0; j < columns_.size(); j++) {
        if (j) fputc(' ', fp_);
        if (columns_[j].type == SDDS_T_STRING) {
          fputs(quote(((const std::string *)cols[j])[i]).c_str(), fp_);
          continue;
        }
        double v = ((const double *)cols[j])[i];
        switch (columns_[j].type) {
          case SDDS_T_DOUBLE: fprintf(fp_, "%.17g", v); break;
          case SDDS_T_FLOAT: fprintf(fp_, "%.9g", (double)(float)v); break;
          case SDDS_T_LONG: fprintf(fp_, "%d", (int32_t)v); break;
          case SDDS_T_SHORT: fprintf(fp_, "%d", (int)(int16_t)v); break;
          default: fputc((char)v, fp_);
        }
      }
      fputc('\n', fp_);
    }
  }
  fflush(fp_);
  return ferror(fp_) ? (err = "write error", 1) : 0;
}

}  // namespace spf

extern "C" {

static int sdds_fail(char *err, int64_t len, const std::string &m) {
  if (err && len > 0) snprintf(err, (size_t)len, "%s", m.c_str());
  return 1;
}

int spf_sdds_info(const char *path, int64_t *n_pages, int64_t *n_rows, char *err, int64_t err_len) {
  spf::SddsFile f;
  std::string e;
  if (!path || spf::sdds_read(path, f, e)) return sdds_fail(err, err_len, path ? e : "null path");
  if (n_pages) *n_pages = (int64_t)f.pages.size();
  if (n_rows) *n_rows = f.pages.empty() ? 0 : f.pages[0].rows;
  return 0;
}

int spf_sdds_column(const char *path, int64_t page, const char *column, double *out, int64_t max, int64_t *n,
                    char *err, int64_t err_len) {
  spf::SddsFile f;
  std::string e;
  if (!path || !column || spf::sdds_read(path, f, e)) return sdds_fail(err, err_len, path ? e : "null path");
  if (page < 0 || page >= (int64_t)f.pages.size()) return sdds_fail(err, err_len, "no such page");
  int j = f.column_index(column);
  if (j < 0) return sdds_fail(err, err_len, std::string("no column ") + column);
  const auto &c = f.pages[page].cols[j];
  if (c.num.size() != (size_t)f.pages[page].rows) return sdds_fail(err, err_len, "column is not numeric");
  if (n) *n = f.pages[page].rows;
  for (int64_t i = 0; i < f.pages[page].rows && i < max; i++) out[i] = c.num[i];
  return 0;
}

int spf_sdds_parameter(const char *path, int64_t page, const char *name, double *value, char *err,
                       int64_t err_len) {
  spf::SddsFile f;
  std::string e;
  if (!path || !name || spf::sdds_read(path, f, e)) return sdds_fail(err, err_len, path ? e : "null path");
  if (page < 0 || page >= (int64_t)f.pages.size()) return sdds_fail(err, err_len, "no such page");
  int j = f.param_index(name);
  if (j < 0) return sdds_fail(err, err_len, std::string("no parameter ") + name);
  if (value) *value = f.pages[page].par_num[j];
  return 0;
}

}  // extern "C"
