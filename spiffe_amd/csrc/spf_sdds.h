// spf_sdds.h -- the subset of the SDDS file protocol spiffe's files use:
// SDDS1 headers of &description, &parameter, &column and &data lines; ASCII
// pages (with or without row counts) and binary pages in this machine's byte
// order (row count, non-fixed parameter values, then the rows, each row
// holding its columns in order); types double, float, long (32-bit), short,
// character and string.  Arrays and compressed files are not supported.
// Plain C++, no HIP.
#pragma once
#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

namespace spf {

enum SddsType { SDDS_T_DOUBLE, SDDS_T_FLOAT, SDDS_T_LONG, SDDS_T_SHORT, SDDS_T_CHARACTER, SDDS_T_STRING };

struct SddsField {
  std::string name, units, description;
  SddsType type = SDDS_T_DOUBLE;
  std::string fixed_value;   // parameters only; non-empty: stored in the header
};

// A value of a parameter or one column of a page, held as text for strings
// and characters and as doubles otherwise.
struct SddsColumn {
  std::vector<double> num;
  std::vector<std::string> str;
};

struct SddsPage {
  std::vector<std::string> par_str;   // every parameter as text (fixed ones too)
  std::vector<double> par_num;        // numeric value, NaN for strings
  std::vector<SddsColumn> cols;
  int64_t rows = 0;
};

struct SddsFile {
  std::string description;
  std::vector<SddsField> params, columns;
  bool binary = false;
  std::vector<SddsPage> pages;
  int param_index(const char *name) const;
  int column_index(const char *name) const;
};

// Reads a whole file; returns 0 or nonzero with a message.
int sdds_read(const std::string &path, SddsFile &out, std::string &err);

// Writes pages as they come.  Column data is passed column by column; values
// are converted to each column's type when written.
class SddsWriter {
 public:
  ~SddsWriter();
  int open(const std::string &path, const std::string &description, const std::vector<SddsField> &params,
           const std::vector<SddsField> &columns, bool binary, std::string &err);
  // par: one text value per non-fixed parameter (numbers formatted %.17g);
  // cols[j] points at `rows` values of column j: double for numeric types,
  // const std::string for strings
  int write_page(const std::vector<std::string> &par, int64_t rows, const std::vector<const void *> &cols,
                 std::string &err);
  void close();
  bool is_open() const { return fp_ != nullptr; }

 private:
  FILE *fp_ = nullptr;
  bool binary_ = true;
  std::vector<SddsField> params_, columns_;
};

}  // namespace spf
