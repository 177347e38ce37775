// Minimal command-line parser with the conventions of the reference tools
// (lib/cmdline/cmdline.h as configured by src/util/cmdline_types.h and
// cmdline_hooks.cpp):
//  - long options `--name value` / `--name=value`, short `-x value`, flags,
//    repeatable counted flags (-v / -vv);
//  - integers accept k/m/g suffixes (x1e3/1e6/1e9), vectors are comma lists;
//  - `-c file` and <input>.cfg files hold one `--name=value` per line, `#`
//    starts a comment; options given on the command line win, options marked
//    not-from-file are ignored in files.
#pragma once

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace strip_cli {

enum class Kind { Flag, Int, Double, Doubles, Path };

struct Option {
  std::string name;
  char short_name;
  Kind kind;
  std::string def;       // default value as text ("" = none)
  std::string help;
  bool not_from_file = false;
  // state
  bool given = false;
  bool from_cmdline = false;
  int count = 0;
  std::string value;
};

inline long parse_count_int(const std::string& s) {
  char* end = nullptr;
  long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str()) throw std::runtime_error("not an integer: " + s);
  if (*end) {
    if (end[1]) throw std::runtime_error("not an integer: " + s);
    switch (*end) {
      case 'k': case 'K': v *= 1000; break;
      case 'm': case 'M': v *= 1000000; break;
      case 'g': case 'G': v *= 1000000000; break;
      default: throw std::runtime_error("not an integer: " + s);
    }
  }
  return v;
}

class Parser {
 public:
  void add(const std::string& name, char short_name, Kind kind,
           const std::string& def, const std::string& help,
           bool not_from_file = false) {
    Option o;
    o.name = name;
    o.short_name = short_name;
    o.kind = kind;
    o.def = def;
    o.help = help;
    o.not_from_file = not_from_file;
    index_[name] = options_.size();
    options_.push_back(o);
  }

  void parse(int argc, char** argv) {
    std::vector<std::string> args(argv + 1, argv + argc);
    parse_args(args, true);
  }

  /// Loads `--name=value` lines; command-line options take precedence.
  void load_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open input config file: " + path);
    std::vector<std::string> args;
    std::string line;
    while (std::getline(in, line)) {
      while (!line.empty() && (line.back() == '\r' || line.back() == ' '))
        line.pop_back();
      if (line.empty() || line[0] == '#') continue;
      args.push_back(line);
    }
    parse_args(args, false);
  }

  bool exist(const std::string& name) const { return opt(name).given; }
  int count(const std::string& name) const { return opt(name).count; }

  std::string text(const std::string& name) const {
    const Option& o = opt(name);
    return o.given ? o.value : o.def;
  }
  long get_int(const std::string& name) const {
    const std::string t = text(name);
    return t.empty() ? 0 : parse_count_int(t);
  }
  double get_double(const std::string& name) const {
    const std::string t = text(name);
    return t.empty() ? 0.0 : std::stod(t);
  }
  std::vector<double> get_doubles(const std::string& name) const {
    std::vector<double> out;
    std::stringstream ss(text(name));
    std::string item;
    while (std::getline(ss, item, ','))
      if (!item.empty()) out.push_back(std::stod(item));
    return out;
  }
  void set(const std::string& name, const std::string& value) {
    Option& o = opt(name);
    o.value = value;
    o.given = true;
  }

  const std::vector<std::string>& rest() const { return rest_; }

  std::string usage(const std::string& program) const {
    std::ostringstream os;
    os << "usage: " << program << " [options] ...\n";
    for (const Option& o : options_) {
      os << "  ";
      if (o.short_name) os << "-" << o.short_name << ", ";
      else os << "    ";
      os << "--" << o.name;
      if (o.kind != Kind::Flag) os << "=<" << kind_name(o.kind) << ">";
      os << "  " << o.help;
      if (!o.def.empty()) os << " [=" << o.def << "]";
      os << "\n";
    }
    return os.str();
  }

  /// The options that were given, as `--name=value` lines (a .cfg file).
  std::string config_text(const std::vector<std::string>& keys) const {
    std::ostringstream os;
    for (const std::string& k : keys) {
      const Option& o = opt(k);
      if (o.kind == Kind::Flag) {
        if (o.given) os << "--" << o.name << "\n";
      } else if (!text(k).empty()) {
        os << "--" << o.name << "=" << text(k) << "\n";
      }
    }
    return os.str();
  }

 private:
  static const char* kind_name(Kind k) {
    switch (k) {
      case Kind::Int: return "int";
      case Kind::Double: return "float";
      case Kind::Doubles: return "float,...";
      case Kind::Path: return "file";
      default: return "";
    }
  }

  Option& opt(const std::string& name) {
    auto it = index_.find(name);
    if (it == index_.end()) throw std::runtime_error("unknown option " + name);
    return options_[it->second];
  }
  const Option& opt(const std::string& name) const {
    auto it = index_.find(name);
    if (it == index_.end()) throw std::runtime_error("unknown option " + name);
    return options_[it->second];
  }

  Option* by_short(char c) {
    for (Option& o : options_)
      if (o.short_name == c) return &o;
    return nullptr;
  }

  void assign(Option& o, const std::string& value, bool cmdline) {
    if (!cmdline && (o.not_from_file || o.from_cmdline)) return;
    if (o.kind == Kind::Flag) {
      o.count += 1;
      o.given = true;
    } else {
      if (o.kind == Kind::Int) parse_count_int(value);
      if (o.kind == Kind::Double) std::stod(value);
      o.value = value;
      o.given = true;
    }
    if (cmdline) o.from_cmdline = true;
  }

  void parse_args(const std::vector<std::string>& args, bool cmdline) {
    for (size_t i = 0; i < args.size(); ++i) {
      const std::string& a = args[i];
      if (a.size() > 2 && a[0] == '-' && a[1] == '-') {
        std::string name = a.substr(2), value;
        const size_t eq = name.find('=');
        const bool has_eq = eq != std::string::npos;
        if (has_eq) {
          value = name.substr(eq + 1);
          name = name.substr(0, eq);
        }
        Option& o = opt(name);
        if (o.kind != Kind::Flag && !has_eq) {
          if (i + 1 >= args.size())
            throw std::runtime_error("option --" + name + " needs a value");
          value = args[++i];
        }
        if (name == "config") {
          assign(o, value, cmdline);
          load_file(value);
          continue;
        }
        assign(o, value, cmdline);
      } else if (a.size() > 1 && a[0] == '-' && !std::isdigit((unsigned char)a[1])) {
        // -v, -vv, -i 10, -i10
        for (size_t k = 1; k < a.size(); ++k) {
          Option* o = by_short(a[k]);
          if (!o) throw std::runtime_error(std::string("unknown option -") + a[k]);
          if (o->kind == Kind::Flag) {
            assign(*o, "", cmdline);
            continue;
          }
          std::string value = a.substr(k + 1);
          if (value.empty()) {
            if (i + 1 >= args.size())
              throw std::runtime_error(std::string("option -") + a[k] +
                                       " needs a value");
            value = args[++i];
          }
          if (o->name == "config") {
            assign(*o, value, cmdline);
            load_file(value);
          } else {
            assign(*o, value, cmdline);
          }
          break;
        }
      } else {
        if (cmdline) rest_.push_back(a);
      }
    }
  }

  std::vector<Option> options_;
  std::map<std::string, size_t> index_;
  std::vector<std::string> rest_;
};

/// Path helpers of src/util/cmdline_types.h:12-84.
struct path {
  static size_t sep(const std::string& s) { return s.find_last_of("\\/"); }
  static size_t ext_pos(const std::string& s) {
    size_t e = s.find_last_of('.');
    size_t p = sep(s);
    if (p != std::string::npos && e != std::string::npos && p > e)
      e = std::string::npos;
    return e;
  }
  static std::string wo_ext(const std::string& s) {
    return s.substr(0, ext_pos(s));
  }
  static std::string ext(const std::string& s) {
    size_t e = ext_pos(s);
    return e == std::string::npos ? std::string() : s.substr(e);
  }
  static std::string wo_path(const std::string& s) {
    size_t p = sep(s);
    return p == std::string::npos ? s : s.substr(p + 1);
  }
  /// `<base>_<index>` zero-padded to the digits of max (add_index).
  static std::string add_index(const std::string& s, int index, int max) {
    int digits = 1;
    for (int m = 10; m <= max; m *= 10) ++digits;
    std::ostringstream os;
    os << s << "_";
    os.width(digits);
    os.fill('0');
    os << index;
    return os.str();
  }
};

}  // namespace strip_cli
