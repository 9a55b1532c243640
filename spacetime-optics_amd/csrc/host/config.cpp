#include "config.h"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace bh {

std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

static std::string strip_comment(const std::string& line) {
    size_t h = line.find('#');
    return trim(h == std::string::npos ? line : line.substr(0, h));
}

// Grammar of common/config.cpp:14-63
bool Config::load(const std::string& path, std::string* err) {
    std::ifstream in(path);
    if (!in) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    size_t slash = path.rfind('/');
    const std::string here = slash == std::string::npos ? "." : path.substr(0, slash);
    dir_ = here;
    std::string raw, section;
    int lineno = 0;
    while (std::getline(in, raw)) {
        ++lineno;
        std::string line = strip_comment(raw);
        // continuation: a trailing '\' joins the next line (kept as-is at end of file)
        while (!line.empty() && line.back() == '\\' && std::getline(in, raw)) {
            ++lineno;
            line.pop_back();
            line = trim(line + " " + strip_comment(raw));
        }
        if (line.empty()) continue;
        const std::string where = path + ":" + std::to_string(lineno);
        if (line[0] == '[') {
            if (line.back() != ']') {
                if (err) *err = where + ": bad section header";
                return false;
            }
            section = trim(line.substr(1, line.size() - 2));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) {
            if (err) *err = where + ": expected key = value";
            return false;
        }
        std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (section.empty() && key == "include") {
            std::string base = (!val.empty() && val[0] == '/') ? val : here + "/" + val;
            if (!load(base, err)) return false;
            dir_ = here;
            continue;
        }
        kv_[section.empty() ? key : section + "." + key] = val;
    }
    return true;
}

std::string Config::str(const std::string& key, const std::string& def) const {
    auto it = kv_.find(key);
    return it == kv_.end() ? def : it->second;
}

double Config::num(const std::string& key, double def) const {
    auto it = kv_.find(key);
    if (it == kv_.end() || it->second.empty()) return def;
    const char* s = it->second.c_str();
    char* end = nullptr;
    double v = std::strtod(s, &end);
    return end == s ? def : v;
}

bool Config::flag(const std::string& key, bool def) const {
    auto it = kv_.find(key);
    if (it == kv_.end()) return def;
    const std::string& v = it->second;
    if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
    if (v == "0" || v == "false" || v == "no" || v == "off") return false;
    return def;
}

std::vector<std::string> Config::list(const std::string& key, char sep) const {
    std::vector<std::string> out;
    std::stringstream ss(str(key));
    std::string item;
    while (std::getline(ss, item, sep)) {
        item = trim(item);
        if (!item.empty()) out.push_back(item);
    }
    return out;
}

bool Config::apply_set(const std::string& kv) {
    size_t eq = kv.find('=');
    if (eq == std::string::npos) return false;
    set(trim(kv.substr(0, eq)), trim(kv.substr(eq + 1)));
    return true;
}

}  // namespace bh
