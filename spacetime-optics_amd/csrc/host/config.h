/* config.h - shot .ini configuration with bhrt's grammar (common/config.hpp:1-38 of the reference):
   "[section]" headers, "key = value" lines, '#' comments, a trailing '\' continues the value on
   the next line, a top-level "include = base.ini" loads a base file first (later keys override),
   keys are addressed as "section.key", command-line "--set section.key=value" overrides. */
#pragma once
#include <map>
#include <string>
#include <vector>

namespace bh {

std::string trim(const std::string& s);

class Config {
public:
    bool load(const std::string& path, std::string* err);
    void set(const std::string& key, const std::string& value) { kv_[key] = value; }
    bool has(const std::string& key) const { return kv_.count(key) != 0; }
    std::string str(const std::string& key, const std::string& def = "") const;
    double num(const std::string& key, double def) const;
    int integer(const std::string& key, int def) const { return (int)num(key, def); }
    bool flag(const std::string& key, bool def) const;
    std::vector<std::string> list(const std::string& key, char sep = ',') const;
    const std::map<std::string, std::string>& all() const { return kv_; }
    std::string dir() const { return dir_; }
    // Applies "key=value" (one --set argument). Returns false if there is no '='.
    bool apply_set(const std::string& kv);

private:
    const std::string* find(const std::string& key) const;   // nullptr when the key is not set
    std::map<std::string, std::string> kv_;
    std::string dir_ = ".";
};

}  // namespace bh
