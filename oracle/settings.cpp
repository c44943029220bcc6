// CPU oracle (test infrastructure only): the .rc settings format, restated
// from libs/core/settings.cpp, with the defaults of
// solvers/elliptic/src/ellipticSettings.cpp:38-103, libs/mesh/meshSettings.cpp:32-91,
// libs/core/platformSettings.cpp:31-56 and libs/parAlmond/parAlmondSettings.cpp:33-60.
#include <cctype>
#include <sstream>
#include <stdexcept>

#include "oracle.hpp"

namespace oracle {

void Settings::set(const std::string& name, const std::string& val) {
  if (!kv.count(name)) order.push_back(name);
  kv[name] = val;
}
bool Settings::has(const std::string& name) const { return kv.count(name) > 0; }
std::string Settings::get(const std::string& name) const {
  auto it = kv.find(name);
  if (it == kv.end()) throw std::runtime_error("Unable to find setting: [" + name + "]");
  return it->second;
}
// setting_t::compareVal (settings.cpp:73-75): substring match.
bool Settings::compare(const std::string& name, const std::string& token) const {
  return get(name).find(token) != std::string::npos;
}
int Settings::geti(const std::string& name) const { return std::stoi(get(name)); }
double Settings::getd(const std::string& name) const { return std::stod(get(name)); }

/* settings_t::readSettingsFromFile (settings.cpp:134-214): '#' ends a line,
 * "[NAME]" opens a setting, every non-space character after it is the value. */
void Settings::parseText(const std::string& text) {
  std::istringstream in(text);
  std::string line, name, val;
  while (std::getline(in, line)) {
    const int size = (int)line.size();
    for (int i = 0; i < size; i++) {
      char c = line[i];
      if (c == '#') break;
      if (c == '[') {
        if (name.length() && val.length()) {
          set(name, val);
          name.clear();
          val.clear();
        }
        name = "";
        val = "";
        i++;
        while (i < size && line[i] != ']') name += line[i++];
      } else {
        if (isspace((unsigned char)c)) continue;
        val += c;
      }
    }
    if (name.length() && val.length()) {
      set(name, val);
      name.clear();
      val.clear();
    }
  }
}

Settings Settings::defaults() {
  Settings s;
  // platform
  s.set("FORMAT", "2.0");
  s.set("THREAD MODEL", "Serial");
  s.set("PLATFORM NUMBER", "0");
  s.set("DEVICE NUMBER", "0");
  // mesh
  s.set("MESH FILE", "BOX");
  s.set("MESH DIMENSION", "3");
  s.set("ELEMENT TYPE", "12");
  s.set("ELEMENT MAP", "ISOPARAMETRIC");
  s.set("BOX DIMX", "10");
  s.set("BOX DIMY", "10");
  s.set("BOX DIMZ", "10");
  s.set("BOX NX", "10");
  s.set("BOX NY", "10");
  s.set("BOX NZ", "10");
  s.set("BOX GLOBAL NX", "0");
  s.set("BOX GLOBAL NY", "0");
  s.set("BOX GLOBAL NZ", "0");
  s.set("BOX BOUNDARY FLAG", "1");
  s.set("POLYNOMIAL DEGREE", "4");
  s.set("BOX CURVED MAP AMPLITUDE", "0");
  // elliptic run + solver
  s.set("DATA FILE", "data/ellipticSine2D.h");
  s.set("LAMBDA", "1.0");
  s.set("OUTPUT TO FILE", "FALSE");
  s.set("OUTPUT FILE NAME", "elliptic");
  s.set("DISCRETIZATION", "CONTINUOUS");
  s.set("LINEAR SOLVER", "PCG");
  s.set("LINEAR SOLVER STOPPING CRITERION", "ABS/REL-INITRESID");
  s.set("PRECONDITIONER", "NONE");
  s.set("MULTIGRID COARSENING", "HALFDOFS");
  s.set("MULTIGRID SMOOTHER", "CHEBYSHEV");
  s.set("MULTIGRID CHEBYSHEV DEGREE", "2");
  s.set("VERBOSE", "FALSE");
  // parAlmond
  s.set("PARALMOND CYCLE", "VCYCLE");
  s.set("PARALMOND STRENGTH", "SYMMETRIC");
  s.set("PARALMOND AGGREGATION", "SMOOTHED");
  s.set("PARALMOND SMOOTHER", "CHEBYSHEV");
  s.set("PARALMOND CHEBYSHEV DEGREE", "2");
  return s;
}

}  // namespace oracle
