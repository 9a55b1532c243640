#include "config.h"

#include <cstdlib>
#include <fstream>

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

const std::string* Config::find(const std::string& key) const {
    const auto it = kv_.find(key);
    return it == kv_.end() ? nullptr : &it->second;
}

std::string Config::str(const std::string& key, const std::string& def) const {
    const std::string* v = find(key);
    return v ? *v : def;
}

// The number a value starts with (strtod's prefix rule: trailing text is ignored); the default for
// a missing key or a value that does not start with a number (the empty value included).
double Config::num(const std::string& key, double def) const {
    const std::string* v = find(key);
    if (!v) return def;
    const char* begin = v->c_str();
    char* stop = nullptr;
    const double x = std::strtod(begin, &stop);
    return stop == begin ? def : x;
}

// The words bhrt's shot files use for switches (exact, lower case); anything else keeps the default
bool Config::flag(const std::string& key, bool def) const {
    static const char* const on_words[] = {"1", "true", "yes", "on"};
    static const char* const off_words[] = {"0", "false", "no", "off"};
    const std::string* v = find(key);
    if (!v) return def;
    for (const char* w : on_words)
        if (*v == w) return true;
    for (const char* w : off_words)
        if (*v == w) return false;
    return def;
}

// Items between separators, trimmed; empty items are dropped
std::vector<std::string> Config::list(const std::string& key, char sep) const {
    std::vector<std::string> items;
    const std::string s = str(key);
    for (size_t from = 0; from <= s.size();) {
        size_t to = s.find(sep, from);
        if (to == std::string::npos) to = s.size();
        std::string item = trim(s.substr(from, to - from));
        if (!item.empty()) items.push_back(std::move(item));
        from = to + 1;
    }
    return items;
}

bool Config::apply_set(const std::string& kv) {
    size_t eq = kv.find('=');
    if (eq == std::string::npos) return false;
    set(trim(kv.substr(0, eq)), trim(kv.substr(eq + 1)));
    return true;
}

}  // namespace bh
