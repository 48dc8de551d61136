// Input-deck reader of the quarz_amd host: the format of the reference (src/Config.hpp:1-11,
// src/Config.cpp:48-236): "key = value" statements, flat or in blocks, ';' separators, '#' comments,
// named blocks (bunch/beam/species/diag), units with units.n0, a key given twice is an error.
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

private:
    const std::string& raw(const std::string& key) const;
    std::map<std::string, std::string> kv_;
    mutable std::set<std::string> used_;
};

}  // namespace qz
