// ini.cpp -- see ini.hpp.
#include "ini.hpp"

#include <cstdio>
#include <fstream>
#include <locale>
#include <sstream>

namespace shrt_cli {
namespace {

std::string trim(const std::string& s) {
    const char* ws = " \t\r\n\f\v";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return std::string();
    const size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

// Full-string stream conversion (ptree's stream_translator): the extraction
// must succeed and leave nothing but whitespace.
template <typename T>
bool convert(const std::string& text, T* out) {
    std::istringstream is(text);
    is.imbue(std::locale::classic());
    T v{};
    is >> v;
    if (is.fail()) return false;
    is >> std::ws;
    if (!is.eof()) return false;
    *out = v;
    return true;
}

bool convert_bool(const std::string& text, bool* out) {
    {
        std::istringstream is(text);
        is.imbue(std::locale::classic());
        bool v = false;
        is >> v;
        if (!is.fail()) {
            is >> std::ws;
            if (is.eof()) { *out = v; return true; }
        }
    }
    std::istringstream is(text);
    is.imbue(std::locale::classic());
    bool v = false;
    is >> std::boolalpha >> v;
    if (is.fail()) return false;
    is >> std::ws;
    if (!is.eof()) return false;
    *out = v;
    return true;
}

void warn_missing(const std::string& section, const std::string& key, std::string* warn) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "WARNING: readEntry: Key \"%s\" in section [%s] not found. Using default.\n",
                  key.c_str(), section.c_str());
    std::fputs(buf, stdout);
    if (warn) *warn += buf;
}

}  // namespace

Ini Ini::read(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw IniError(path + ": cannot open file");
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str(), path);
}

Ini Ini::parse(const std::string& text, const std::string& name) {
    Ini ini;
    std::istringstream in(text);
    std::string raw;
    auto& top = ini.sections_[""];
    std::map<std::string, std::string>* cur = &top;
    std::string cur_name;
    bool in_section = false;
    int lineno = 0;
    while (std::getline(in, raw)) {
        ++lineno;
        const std::string line = trim(raw);
        auto err = [&](const std::string& why) {
            return IniError(name + "(" + std::to_string(lineno) + "): " + why);
        };
        if (line.empty() || line[0] == ';' || line[0] == '#') continue;
        if (line[0] == '[') {
            // an empty previous section is dropped again, as read_ini does
            if (in_section && cur->empty()) ini.sections_.erase(cur_name);
            const size_t end = line.find(']');
            if (end == std::string::npos) throw err("unmatched '['");
            const std::string sec = trim(line.substr(1, end - 1));
            // section names share the top level with keys outside any section
            if ((sec.empty() ? false : ini.sections_.count(sec) > 0) || top.count(sec)) throw err("duplicate section name");
            cur = &ini.sections_[sec];
            cur_name = sec;
            in_section = true;
            continue;
        }
        const size_t eq = line.find('=');
        if (eq == std::string::npos) throw err("'=' character not found in line");
        if (eq == 0) throw err("key expected");
        const std::string key = trim(line.substr(0, eq));
        const std::string value = trim(line.substr(eq + 1));
        if (cur->count(key) || (!in_section && ini.sections_.count(key) && key != "")) throw err("duplicate key name");
        (*cur)[key] = value;
    }
    return ini;
}

bool Ini::get(const std::string& section, const std::string& key, std::string* value) const {
    const auto s = sections_.find(section);
    if (s == sections_.end()) return false;
    const auto k = s->second.find(key);
    if (k == s->second.end()) return false;
    *value = k->second;
    return true;
}

int Ini::entry_int(const std::string& section, const std::string& key, int def, std::string* warn) const {
    std::string v;
    int out = 0;
    if (get(section, key, &v) && convert(v, &out)) return out;
    warn_missing(section, key, warn);
    return def;
}

double Ini::entry_double(const std::string& section, const std::string& key, double def, std::string* warn) const {
    std::string v;
    double out = 0;
    if (get(section, key, &v) && convert(v, &out)) return out;
    warn_missing(section, key, warn);
    return def;
}

float Ini::entry_float(const std::string& section, const std::string& key, float def, std::string* warn) const {
    std::string v;
    float out = 0;
    if (get(section, key, &v) && convert(v, &out)) return out;
    warn_missing(section, key, warn);
    return def;
}

bool Ini::entry_bool(const std::string& section, const std::string& key, bool def, std::string* warn) const {
    std::string v;
    bool out = false;
    if (get(section, key, &v) && convert_bool(v, &out)) return out;
    warn_missing(section, key, warn);
    return def;
}

}  // namespace shrt_cli
