// flags.cpp — gflags-syntax parser (see flags.h for the accepted forms).
#include "flags.h"

#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <sstream>

namespace scaph {

namespace {

std::string normalise(std::string s) {
  for (char& c : s)
    if (c == '-') c = '_';
  return s;
}

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

std::string value_text(const FlagInfo& f) {
  std::ostringstream o;
  switch (f.type) {
    case FlagType::kInt32: o << *static_cast<int32_t*>(f.storage); break;
    case FlagType::kInt64: o << *static_cast<int64_t*>(f.storage); break;
    case FlagType::kBool: o << (*static_cast<bool*>(f.storage) ? "true" : "false"); break;
    case FlagType::kDouble: o << *static_cast<double*>(f.storage); break;
    case FlagType::kString: o << '"' << *static_cast<std::string*>(f.storage) << '"'; break;
  }
  return o.str();
}

}  // namespace

bool parse_bool_text(const std::string& s, bool* out) {
  const std::string l = lower(s);
  if (l == "1" || l == "t" || l == "true" || l == "y" || l == "yes") {
    *out = true;
    return true;
  }
  if (l == "0" || l == "f" || l == "false" || l == "n" || l == "no") {
    *out = false;
    return true;
  }
  return false;
}

bool parse_int_text(const std::string& s, int64_t* out) {
  if (s.empty()) return false;
  const char* p = s.c_str();
  char* end = nullptr;
  errno = 0;
  int base = 10;
  const char* q = p;
  if (*q == '+' || *q == '-') ++q;
  if (q[0] == '0' && (q[1] == 'x' || q[1] == 'X')) base = 16;
  long long v = std::strtoll(p, &end, base);
  if (errno != 0 || end == p || *end != '\0') return false;
  *out = v;
  return true;
}

FlagRegistry& FlagRegistry::global() {
  static FlagRegistry r;
  return r;
}

void FlagRegistry::add(const std::string& name, FlagType type, void* storage,
                       const std::string& help) {
  FlagInfo f{normalise(name), type, storage, help, ""};
  f.default_text = value_text(f);
  for (auto& g : flags_)
    if (g.name == f.name) {
      g = f;
      return;
    }
  flags_.push_back(f);
}

const FlagInfo* FlagRegistry::find(const std::string& name) const {
  const std::string n = normalise(name);
  for (const auto& f : flags_)
    if (f.name == n) return &f;
  return nullptr;
}

bool FlagRegistry::set(const std::string& name, const std::string& value, std::string* err) {
  const FlagInfo* f = find(name);
  if (!f) {
    if (err) *err = "ERROR: unknown command line flag '" + name + "'";
    return false;
  }
  auto bad = [&](const char* what) {
    if (err) *err = "ERROR: illegal value '" + value + "' specified for " + what + " flag '" + name + "'";
    return false;
  };
  switch (f->type) {
    case FlagType::kInt32: {
      int64_t v;
      if (!parse_int_text(value, &v) || v < INT32_MIN || v > INT32_MAX) return bad("int32");
      *static_cast<int32_t*>(f->storage) = (int32_t)v;
      return true;
    }
    case FlagType::kInt64: {
      int64_t v;
      if (!parse_int_text(value, &v)) return bad("int64");
      *static_cast<int64_t*>(f->storage) = v;
      return true;
    }
    case FlagType::kBool: {
      bool v;
      if (!parse_bool_text(value, &v)) return bad("bool");
      *static_cast<bool*>(f->storage) = v;
      return true;
    }
    case FlagType::kDouble: {
      char* end = nullptr;
      errno = 0;
      double v = std::strtod(value.c_str(), &end);
      if (value.empty() || errno != 0 || *end != '\0') return bad("double");
      *static_cast<double*>(f->storage) = v;
      return true;
    }
    case FlagType::kString:
      *static_cast<std::string*>(f->storage) = value;
      return true;
  }
  return false;
}

std::string FlagRegistry::usage() const {
  std::ostringstream o;
  o << "Flags:\n";
  for (const auto& f : flags_)
    o << "  -" << f.name << " (" << f.help << ") default: " << f.default_text << "\n";
  return o.str();
}

namespace {

bool read_flagfile(const std::string& path, std::vector<std::string>* out, std::string* err) {
  std::ifstream in(path);
  if (!in) {
    if (err) *err = "ERROR: can't open flagfile '" + path + "'";
    return false;
  }
  std::string line;
  while (std::getline(in, line)) {
    size_t a = line.find_first_not_of(" \t\r");
    if (a == std::string::npos) continue;
    size_t b = line.find_last_not_of(" \t\r");
    std::string t = line.substr(a, b - a + 1);
    if (t.empty() || t[0] == '#') continue;
    out->push_back(t);
  }
  return true;
}

bool parse_impl(std::vector<std::string> args, std::vector<std::string>* positional,
                std::string* err, bool* help, int depth) {
  FlagRegistry& reg = FlagRegistry::global();
  for (size_t i = 0; i < args.size(); ++i) {
    const std::string& a = args[i];
    if (a == "--") {
      for (size_t j = i + 1; j < args.size(); ++j) positional->push_back(args[j]);
      return true;
    }
    if (a.size() < 2 || a[0] != '-') {
      positional->push_back(a);
      continue;
    }
    std::string body = a.substr(a[1] == '-' ? 2 : 1);
    std::string name = body, value;
    bool has_value = false;
    const size_t eq = body.find('=');
    if (eq != std::string::npos) {
      name = body.substr(0, eq);
      value = body.substr(eq + 1);
      has_value = true;
    }
    const std::string nn = normalise(name);
    if (nn == "help" || nn == "helpfull" || nn == "helpshort") {
      if (help) *help = true;
      continue;
    }
    if (nn == "flagfile") {
      if (!has_value) {
        if (i + 1 >= args.size()) {
          if (err) *err = "ERROR: flag '" + name + "' is missing its argument";
          return false;
        }
        value = args[++i];
      }
      if (depth > 8) {
        if (err) *err = "ERROR: flagfile nesting too deep";
        return false;
      }
      std::vector<std::string> sub;
      if (!read_flagfile(value, &sub, err)) return false;
      if (!parse_impl(sub, positional, err, help, depth + 1)) return false;
      continue;
    }
    const FlagInfo* f = reg.find(nn);
    if (!f && nn.size() > 2 && nn.compare(0, 2, "no") == 0) {
      const FlagInfo* g = reg.find(nn.substr(2));
      if (g && g->type == FlagType::kBool && !has_value) {
        *static_cast<bool*>(g->storage) = false;
        continue;
      }
    }
    if (!f) {
      if (err) *err = "ERROR: unknown command line flag '" + name + "'";
      return false;
    }
    if (!has_value) {
      if (f->type == FlagType::kBool) {
        value = "true";
      } else {
        if (i + 1 >= args.size()) {
          if (err) *err = "ERROR: flag '" + name + "' is missing its argument";
          return false;
        }
        value = args[++i];
      }
    }
    if (!reg.set(nn, value, err)) return false;
  }
  return true;
}

}  // namespace

bool parse_flag_vector(const std::vector<std::string>& args, std::vector<std::string>* positional,
                       std::string* err, bool* help_requested) {
  if (help_requested) *help_requested = false;
  return parse_impl(args, positional, err, help_requested, 0);
}

bool parse_command_line_flags(int* argc, char*** argv, std::string* err, bool* help_requested) {
  std::vector<std::string> args;
  for (int i = 1; i < *argc; ++i) args.push_back((*argv)[i]);
  std::vector<std::string> pos;
  if (!parse_flag_vector(args, &pos, err, help_requested)) return false;
  // rewrite argv in place: program name + positionals (positionals that
  // came from a flagfile are kept alive in a static deque)
  static std::deque<std::string> keep;
  int k = 1;
  for (const auto& p : pos) {
    if (k >= *argc) break;
    char* ptr = nullptr;
    for (int j = 1; j < *argc; ++j)
      if (p == (*argv)[j]) {
        ptr = (*argv)[j];
        break;
      }
    if (!ptr) {
      keep.push_back(p);
      ptr = const_cast<char*>(keep.back().c_str());
    }
    (*argv)[k++] = ptr;
  }
  *argc = k;
  return true;
}

}  // namespace scaph
