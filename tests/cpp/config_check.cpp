// Reads a configuration file with include/swe_config.hpp and prints the
// values the reference's drivers and tests read, for tests/test_config.py.
#include <cstdio>
#include <swe_config.hpp>

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  try {
    Parser par(argv[1]);
    DimensionManager dim(par);
    std::printf("%.17g %.17g %.17g %.17g %.17g\n", dim.Unscale<Scales::source>(par.Get("Common", "cor")),
                dim.Unscale<Scales::height>(par.Get("Thacker", "H0")), dim.Scale<Scales::time>(2.0),
                dim.Scale<Scales::velocity>(1.5), dim.Unscale<Scales::length>(par.Get("Ball", "x0")));
    try {
      par.Get("Common", "missing");
      return 3;
    } catch (const ParserError&) {
    }
    try {
      par.Get("Ball", "label");
      return 4;
    } catch (const ParserError&) {
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
