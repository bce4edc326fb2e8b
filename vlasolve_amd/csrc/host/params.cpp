// Parameter parsing with the semantics of the reference's paramsParser
// (src/paramsParser.hxx:19-420).
#include <cctype>
#include <fstream>
#include <iostream>

#include "vls_host.hpp"

namespace vlh {

int verbose = 1;

namespace {

std::vector<std::string> split(const std::string &s, const std::string &delims) {
  std::vector<std::string> out;
  size_t pos = 0;
  while (true) {
    size_t a = s.find_first_not_of(delims, pos);
    if (a == std::string::npos) break;
    size_t b = s.find_first_of(delims, a);
    out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
    if (b == std::string::npos) break;
    pos = b;
  }
  return out;
}

// istringstream extraction as in paramsParser::from_string; a failed
// extraction gives the value-initialised T (C++11 stream semantics)
template <class T>
T fromString(const std::string &s) {
  std::istringstream iss(s);
  T t{};
  if ((iss >> t).fail()) return T{};
  return t;
}
template <>
std::string fromString<std::string>(const std::string &s) {
  std::istringstream iss(s);
  std::string t;
  iss >> t;
  return t;
}

template <class T>
T fromEmpty() {
  return T(true);
}
template <>
std::string fromEmpty<std::string>() {
  return std::string(1, '\x01');
}

template <class T>
void logDefault(const std::string &cat, const std::string &key, int which, const T &def) {
  if (verbose < 1) return;
  if (which < 0)
    std::cout << "Default value assigned to " << cat << "." << key << " : '" << def << "'.\n";
  else
    std::cout << "Default value assigned to " << cat << "." << key << "[" << which << "] : '"
              << def << "'.\n";
}

}  // namespace

Params::Params(int argc, const char *const *argv) {
  int i0 = 0;
  std::string fname = "params.ini";
  if (argc > i0) i0++;  // program name
  if (argc > i0 && argv[i0][0] != '-') fname = argv[i0++];
  int n = i0;
  while (n < argc) {
    if (argv[n][0] != '-')
      throw Failure(std::string("ERROR parsing command line.\n  AT: ") + argv[n]);
    std::string key(&argv[n++][1]);
    std::vector<std::string> value;
    // a following token is a value unless it looks like a flag; "-2" and
    // "-0.5" are values (paramsParser.hxx:186)
    while (n < argc && (argv[n][0] != '-' || isdigit((unsigned char)argv[n][1])))
      value.push_back(argv[n++]);
    insert(key, value);
  }
  if (parseFile(fname)) fileName_ = fname;
  else if (fname != "params.ini")
    throw Failure("ERROR in paramsParser:could not read parameter file '" + fname + "'.");
}

void Params::insert(const std::string &rawKey, const std::vector<std::string> &v) {
  std::string cat = "default", key = rawKey;
  size_t p = rawKey.find('.');
  if (p != std::string::npos) {
    if (p != rawKey.rfind('.') || p == 0 || p == rawKey.size() - 1)
      throw Failure("ERROR in paramsParseT.\n  ill formated keyword '" + rawKey + "'.");
    cat = rawKey.substr(0, p);
    key = rawKey.substr(p + 1);
  }
  Value &val = cats_[cat][key];
  val.val.push_back(v);
  val.wasRead.push_back(false);
}

bool Params::parseFile(const std::string &fname) {
  std::ifstream in(fname.c_str());
  if (!in) return false;
  std::string line;
  while (std::getline(in, line)) {
    size_t h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    std::vector<std::string> tok = split(line, " =:{}[],\t\r");
    if (tok.empty()) continue;
    insert(tok[0], std::vector<std::string>(tok.begin() + 1, tok.end()));
  }
  return true;
}

const std::vector<std::string> *Params::lookup(const std::string &key, const std::string &cat,
                                               int which, const Value **vout) const {
  auto c = cats_.find(cat);
  if (c == cats_.end()) return nullptr;
  auto k = c->second.find(key);
  if (k == c->second.end()) return nullptr;
  const Value &v = k->second;
  if (which >= 0 && (int)v.val[0].size() <= which) return nullptr;
  *vout = &v;
  return &v.val[0];
}

template <class T>
T Params::get(const std::string &key, const std::string &cat, const T &def, int which) const {
  const Value *v = nullptr;
  const std::vector<std::string> *vals = lookup(key, cat, which, &v);
  if (!vals) {
    logDefault(cat, key, which, def);
    return def;
  }
  int w = which < 0 ? 0 : which;
  T result = vals->empty() ? fromEmpty<T>() : fromString<T>((*vals)[w]);
  v->wasRead[0] = true;
  if (verbose >= 2) {
    if (which < 0) std::cout << "New value assigned to " << cat << "." << key << " : '" << result << "'.\n";
    else std::cout << "New value assigned to " << cat << "." << key << "[" << which << "] : '" << result << "'.\n";
  }
  return result;
}

template int Params::get<int>(const std::string &, const std::string &, const int &, int) const;
template long Params::get<long>(const std::string &, const std::string &, const long &, int) const;
template double Params::get<double>(const std::string &, const std::string &, const double &,
                                    int) const;
template bool Params::get<bool>(const std::string &, const std::string &, const bool &, int) const;
template std::string Params::get<std::string>(const std::string &, const std::string &,
                                              const std::string &, int) const;

bool Params::isDefined(const std::string &key, const std::string &cat) const {
  auto c = cats_.find(cat);
  if (c == cats_.end()) return false;
  return c->second.count(key) > 0;
}

std::string Params::reportUnused() const {
  std::ostringstream os;
  bool found = false;
  for (auto &c : cats_)
    for (auto &k : c.second)
      for (size_t j = 0; j < k.second.wasRead.size(); j++) {
        if (k.second.wasRead[j]) continue;
        if (!found) {
          os << "WARNING! The following parameter definitions were ignored:\n";
          found = true;
        }
        os << "   " << c.first << "." << k.first << " = ";
        for (auto &s : k.second.val[j]) os << s << " ";
        if (k.second.val[j].empty()) os << "true";
        os << "\n";
      }
  return os.str();
}

}  // namespace vlh
