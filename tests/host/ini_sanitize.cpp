// The driver's INI parser (shearing-rt_amd/cli/ini.cpp) under AddressSanitizer
// and UndefinedBehaviorSanitizer (built and run by
// tests/test_host_sanitizers.py): the reference's sample configuration, every
// typed lookup with its default path, and malformed inputs that must raise
// IniError.  Exits 0 when every case behaves.
#include <cstdio>
#include <string>

#include "../../shearing-rt_amd/cli/ini.hpp"

using shrt_cli::Ini;
using shrt_cli::IniError;

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const Ini pt = Ini::read(argv[1]);
    std::string warn;
    const int nx = pt.entry_int("simulation", "nx", 512, &warn);
    const double dt = pt.entry_double("simulation", "time step", 1e-3, &warn);
    const float nu = pt.entry_float("problem", "nu", 1e-3f, &warn);
    const bool noise = pt.entry_bool("problem", "initial noise", true, &warn);
    const int missing = pt.entry_int("nowhere", "nothing", 42, &warn);
    const bool badbool = pt.entry_bool("simulation", "nx", false, &warn);
    std::printf("nx %d dt %g nu %g noise %d missing %d badbool %d\n", nx, dt, (double)nu, (int)noise, missing,
                (int)badbool);
    int bad = 0;
    // (a repeated section is an error only when the first one holds keys:
    // read_ini drops an empty section before it looks for duplicates)
    const char* malformed[] = {"[sec\nkey = 1\n", "[a]\nno equals sign\n", "[a]\n = 1\n", "[a]\nk = 1\nk = 2\n",
                               "[a]\nk = 1\n[a]\n"};
    for (const char* text : malformed) {
        try {
            Ini::parse(text, "case");
            ++bad;
        } catch (const IniError&) {
        }
    }
    const Ini ok = Ini::parse("top = 1\n; comment\n# comment\n[e]\n[s]\nkey with spaces =  v  \nempty =\n[e]\n",
                              "case");
    std::string v;
    if (!ok.get("s", "key with spaces", &v) || v != "v" || !ok.get("s", "empty", &v) || !v.empty()) ++bad;
    return bad == 0 && missing == 42 ? 0 : 1;
}
