// ini.hpp -- the configuration file of the shrt driver: the INI dialect of
// boost::property_tree::read_ini and the typed lookups of readEntry, which the
// reference's input_data uses (file_access.cu:123-188).
//
// Dialect (boost/property_tree/ini_parser.hpp, Boost 1.x):
//   * every line is trimmed of leading and trailing whitespace;
//   * empty lines and lines whose first character is ';' or '#' are comments;
//   * "[name]" opens a section (the trimmed text between the brackets);
//   * "key = value" (split at the first '='; key and value trimmed, the key may
//     hold spaces, the value may be empty) adds a key to the current section,
//     or to the top level before the first section;
//   * a line that is none of these, an empty key, a duplicate key in a section
//     or a duplicate section is a parse error ("<file>(<line>): <reason>").
// Values are not unquoted and trailing comments are part of the value.
//
// Typed reads (readEntry, file_access.cu:167-188): the value must convert in
// full (trailing whitespace allowed); int and floating values follow C++ stream
// extraction, bool accepts "0"/"1" or "true"/"false" (boolalpha, lower case).
// A missing key or a value that does not convert prints
//   WARNING: readEntry: Key "<name>" in section [<section>] not found. Using default.
// and yields the default, as the reference does for both cases.
#pragma once

#include <map>
#include <stdexcept>
#include <string>

namespace shrt_cli {

struct IniError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Ini {
public:
    // Parses `path`; throws IniError on a missing file or a malformed line.
    static Ini read(const std::string& path);
    static Ini parse(const std::string& text, const std::string& name);

    // Raw value of section.key; false when absent.
    bool get(const std::string& section, const std::string& key, std::string* value) const;

    // readEntry: typed value or the default with the reference's warning on
    // stdout.  `warn` receives the warning lines as well (may be null).
    int entry_int(const std::string& section, const std::string& key, int def, std::string* warn) const;
    double entry_double(const std::string& section, const std::string& key, double def, std::string* warn) const;
    float entry_float(const std::string& section, const std::string& key, float def, std::string* warn) const;
    bool entry_bool(const std::string& section, const std::string& key, bool def, std::string* warn) const;

private:
    std::map<std::string, std::map<std::string, std::string>> sections_;
};

}  // namespace shrt_cli
