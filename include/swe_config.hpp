// .ini configuration and non-dimensional scaling, with the reference's class
// names and signatures so drivers written against it compile unchanged:
//   Parser            (docs/ConfigParser_8h_source.html: Parser(filename),
//                      double Get(section, property) const)
//   DimensionManager  (docs/DimensionManager_8h_source.html: Scale<Scales>,
//                      Unscale<Scales>)
// Header-only, host C++17; nothing here touches the device.
//
// File format: one "key = value" per line under "[section]" headers; blank
// lines and lines starting with ';' or '#' are skipped; keys, values and
// section names are trimmed.  Get parses the value as a double and throws
// ParserError for a missing section or key or a value that is not a number.
//
// Scales (the solver works with g = 1, include/SpaceDisc.h:14): height h0,
// length l0, velocity c0 = sqrt(g h0), source (1 / time) c0 / l0, time
// l0 / c0.  DimensionManager reads h0 and l0 from [Dimensions] and g from
// the same section (default 1.0, the reference's constant in
// DimensionManager.cpp; 9.81 turns SI inputs into solver units).
#ifndef SWE_CONFIG_HPP
#define SWE_CONFIG_HPP

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>

struct ParserError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DimensionError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct Parser {
  using HashMap = std::unordered_map<std::string, std::unordered_map<std::string, std::string>>;

  explicit Parser(const std::string& filename) {
    std::ifstream in(filename);
    if (!in) throw ParserError("cannot open " + filename);
    Read(in, filename);
  }
  // configuration text given directly (no file)
  static Parser FromString(const std::string& text) { return Parser(text, 0); }

  double Get(const std::string& section, const std::string& property) const {
    auto s = m_ini.find(section);
    if (s == m_ini.end()) throw ParserError("no section [" + section + "]");
    auto p = s->second.find(property);
    if (p == s->second.end()) throw ParserError("no key " + property + " in [" + section + "]");
    const std::string& v = p->second;
    char* end = nullptr;
    double x = std::strtod(v.c_str(), &end);
    if (v.empty() || end != v.c_str() + v.size())
      throw ParserError("[" + section + "] " + property + " = '" + v + "' is not a number");
    return x;
  }
  bool Has(const std::string& section, const std::string& property) const {
    auto s = m_ini.find(section);
    return s != m_ini.end() && s->second.count(property) > 0;
  }

 private:
  Parser(const std::string& text, int) {
    std::istringstream in(text);
    Read(in, "<string>");
  }
  static std::string Trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
  }
  void Read(std::istream& in, const std::string& name) {
    std::string line, section;
    int no = 0;
    while (std::getline(in, line)) {
      ++no;
      line = Trim(line);
      if (line.empty() || line[0] == ';' || line[0] == '#') continue;
      if (line.front() == '[') {
        if (line.back() != ']') throw ParserError(name + ":" + std::to_string(no) + ": bad section");
        section = Trim(line.substr(1, line.size() - 2));
        m_ini[section];
        continue;
      }
      size_t eq = line.find('=');
      if (eq == std::string::npos) throw ParserError(name + ":" + std::to_string(no) + ": expected key = value");
      m_ini[section][Trim(line.substr(0, eq))] = Trim(line.substr(eq + 1));
    }
  }
  HashMap m_ini;
};

enum class Scales { height, length, velocity, source, time };

struct DimensionManager {
  explicit DimensionManager(const Parser& parser)
      : m_h0(parser.Get("Dimensions", "h0")),
        m_l0(parser.Get("Dimensions", "l0")),
        m_g(parser.Has("Dimensions", "g") ? parser.Get("Dimensions", "g") : 1.0) {
    if (!(m_h0 > 0.) || !(m_l0 > 0.) || !(m_g > 0.))
      throw DimensionError("h0, l0 and g must be positive");
    m_c0 = std::sqrt(m_g * m_h0);
  }

  template <Scales T>
  double Scale(double x) const {
    switch (T) {
      case Scales::height: return x * m_h0;
      case Scales::length: return x * m_l0;
      case Scales::velocity: return x * m_c0;
      case Scales::source: return x * m_c0 / m_l0;
      case Scales::time: return x * m_l0 / m_c0;
    }
    throw DimensionError("invalid dimension type in scale function");
  }

  template <Scales T>
  double Unscale(double x) const {
    switch (T) {
      case Scales::height: return x / m_h0;
      case Scales::length: return x / m_l0;
      case Scales::velocity: return x / m_c0;
      case Scales::source: return x * m_l0 / m_c0;
      case Scales::time: return x * m_c0 / m_l0;
    }
    throw DimensionError("invalid dimension type in unscale function");
  }

 private:
  double m_h0, m_l0, m_g, m_c0;
};

#endif  // SWE_CONFIG_HPP
