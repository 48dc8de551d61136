// Input-deck reader of the quarz_amd host.  It reads the reference's deck format (src/Config.hpp:1-11; the
// reader it replaces is src/Config.cpp): "key = value" statements, flat or in nested blocks that prefix their
// keys ("xi { step = 0.01 }"), named blocks that also join a list ("bunch driver {" joins beams), ';' and newlines
// as separators, '#' comments, numbers with units converted with the reference density units.n0, and a key set
// twice is an error.
#pragma once

#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace qz {

class Config {
public:
    Config() = default;
    // text of a deck (not a file name)
    void parse(const std::string& text);
    // "key=value" lines (command-line overrides, src/main.cpp:82-93)
    void overrides(const std::string& text);
    void set(const std::string& key, const std::string& value) { kv_[key] = value; }

    bool has(const std::string& key) const;
    double number(const std::string& text, const std::string& what) const;
    double n0_cm3() const;

    double get_double(const std::string& key) const;
    double get_double(const std::string& key, double def) const;
    int get_int(const std::string& key) const;
    int get_int(const std::string& key, int def) const;
    bool get_bool(const std::string& key, bool def) const;
    std::string get_string(const std::string& key) const;
    std::string get_string(const std::string& key, const std::string& def) const;
    std::vector<std::string> get_list(const std::string& key) const;
    std::vector<std::pair<double, double>> get_pairs(const std::string& key) const;

    std::vector<std::string> unused_keys() const;
    std::string dump() const;   // "key = value" lines in key order

private:
    const std::string& raw(const std::string& key) const;
    std::map<std::string, std::string> kv_;
    mutable std::set<std::string> used_;
};

}  // namespace qz
