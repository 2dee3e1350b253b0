// flags.h — command-line flags with gflags 2.2 syntax.
//
// The reference declares its flags with gflags (src/app.cuh:57-70, all
// commented out in the snapshot) and the library is not vendored into this
// build, so this is a small parser that accepts the same syntax
// (third-tool-lib/gflags/src/gflags.cc:740-752, 763-819, 1030-1110):
//   -f=v  --f=v  -f v  --f v
//   bool: -f  --f  -nof  --nof  -f=true|false|1|0|t|f|y|n|yes|no (any case)
//   '-' and '_' interchangeable in names
//   ints in decimal or 0x hex
//   --flagfile=path (one flag per line, '#' comments), '--' ends flags
// Parsed flags are removed from argv; positional arguments stay.
#ifndef SCAPH_HOST_FLAGS_H
#define SCAPH_HOST_FLAGS_H

#include <cstdint>
#include <string>
#include <vector>

namespace scaph {

enum class FlagType { kInt32, kInt64, kBool, kDouble, kString };

struct FlagInfo {
  std::string name;
  FlagType type;
  void* storage;
  std::string help;
  std::string default_text;
};

class FlagRegistry {
 public:
  static FlagRegistry& global();
  void add(const std::string& name, FlagType type, void* storage, const std::string& help);
  const FlagInfo* find(const std::string& name) const;
  // Sets a flag from its text value; returns false with *err filled on error.
  bool set(const std::string& name, const std::string& value, std::string* err);
  std::string usage() const;
  const std::vector<FlagInfo>& all() const { return flags_; }

 private:
  std::vector<FlagInfo> flags_;
};

// Parses and removes recognised flags.  Returns false (and fills *err) on
// an unknown flag or a malformed value; --help sets *help_requested.
bool parse_command_line_flags(int* argc, char*** argv, std::string* err,
                              bool* help_requested = nullptr);
// Same over a vector (used by the C-ABI engine constructor).
bool parse_flag_vector(const std::vector<std::string>& args, std::vector<std::string>* positional,
                       std::string* err, bool* help_requested = nullptr);

bool parse_bool_text(const std::string& s, bool* out);
bool parse_int_text(const std::string& s, int64_t* out);

}  // namespace scaph

#endif
