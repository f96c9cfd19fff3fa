#include "foamDict.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace afarhost {

namespace {

std::vector<std::string> tokenize(const std::string& t) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < t.size()) {
        const char c = t[i];
        if (isspace((unsigned char)c)) {
            ++i;
        } else if (t.compare(i, 2, "//") == 0) {
            while (i < t.size() && t[i] != '\n') ++i;
        } else if (t.compare(i, 2, "/*") == 0) {
            const size_t e = t.find("*/", i + 2);
            i = e == std::string::npos ? t.size() : e + 2;
        } else if (c == '{' || c == '}' || c == ';' || c == '(' || c == ')') {
            out.push_back(std::string(1, c));
            ++i;
        } else if (c == '"') {
            const size_t e = t.find('"', i + 1);
            if (e == std::string::npos) throw dictError("unterminated string");
            out.push_back(t.substr(i + 1, e - i - 1));
            i = e + 1;
        } else {
            size_t j = i;
            while (j < t.size() && !isspace((unsigned char)t[j]) && t[j] != '{' && t[j] != '}' && t[j] != ';' &&
                   t[j] != '(' && t[j] != ')' && t.compare(j, 2, "//") != 0 && t.compare(j, 2, "/*") != 0)
                ++j;
            out.push_back(t.substr(i, j - i));
            i = j;
        }
    }
    return out;
}

}  // namespace

struct parser {
    const std::vector<std::string>& tk;
    size_t pos = 0;
    void parse(dictionary& d, bool nested) {
        while (pos < tk.size()) {
            if (tk[pos] == "}") {
                if (!nested) throw dictError("unbalanced '}'");
                ++pos;
                return;
            }
            const std::string key = tk[pos++];
            if (pos >= tk.size()) throw dictError("entry '" + key + "' without value");
            if (tk[pos] == "{") {
                ++pos;
                auto sub = std::make_shared<dictionary>();
                parse(*sub, true);
                d.dicts_[key] = sub;
                d.order_.push_back(key);
                continue;
            }
            std::vector<std::string> val;
            int depth = 0;
            while (pos < tk.size() && !(tk[pos] == ";" && depth == 0)) {
                if (tk[pos] == "(") ++depth;
                if (tk[pos] == ")") --depth;
                val.push_back(tk[pos++]);
            }
            if (pos >= tk.size()) throw dictError("entry '" + key + "' is not terminated by ';'");
            ++pos;
            d.entries_[key] = val;
            d.order_.push_back(key);
        }
        if (nested) throw dictError("missing '}'");
    }
};

dictionary dictionary::fromString(const std::string& text) {
    auto tk = tokenize(text);
    dictionary d;
    parser p{tk};
    p.parse(d, false);
    return d;
}

dictionary dictionary::fromFile(const std::string& path) {
    std::ifstream is(path);
    if (!is) throw dictError("cannot open " + path);
    std::stringstream ss;
    ss << is.rdbuf();
    return fromString(ss.str());
}

const dictionary* dictionary::find_dict(const std::string& path, std::string& leaf) const {
    const dictionary* d = this;
    size_t b = 0;
    for (;;) {
        const size_t e = path.find('.', b);
        if (e == std::string::npos) {
            leaf = path.substr(b);
            return d;
        }
        auto it = d->dicts_.find(path.substr(b, e - b));
        if (it == d->dicts_.end()) return nullptr;
        d = it->second.get();
        b = e + 1;
    }
}

bool dictionary::found(const std::string& path) const {
    std::string leaf;
    const dictionary* d = find_dict(path, leaf);
    return d && (d->entries_.count(leaf) || d->dicts_.count(leaf));
}

const std::vector<std::string>& dictionary::tokens(const std::string& path) const {
    std::string leaf;
    const dictionary* d = find_dict(path, leaf);
    if (!d || !d->entries_.count(leaf)) throw dictError("keyword " + path + " is undefined");
    return d->entries_.at(leaf);
}

std::string dictionary::lookupWord(const std::string& path) const {
    const auto& t = tokens(path);
    if (t.size() != 1) throw dictError("keyword " + path + " is not a single word");
    return t[0];
}

double dictionary::lookupScalar(const std::string& path) const {
    const std::string w = lookupWord(path);
    char* e = nullptr;
    const double v = strtod(w.c_str(), &e);
    if (e == w.c_str() || *e) throw dictError("keyword " + path + " is not a number: " + w);
    return v;
}

long dictionary::lookupLabel(const std::string& path) const {
    const std::string w = lookupWord(path);
    char* e = nullptr;
    const long v = strtol(w.c_str(), &e, 10);
    if (e == w.c_str() || *e) throw dictError("keyword " + path + " is not an integer: " + w);
    return v;
}

bool dictionary::lookupBool(const std::string& path) const {
    const std::string w = lookupWord(path);
    if (w == "true" || w == "on" || w == "yes" || w == "1") return true;
    if (w == "false" || w == "off" || w == "no" || w == "0") return false;
    throw dictError("keyword " + path + " is not a switch: " + w);
}

double dictionary::lookupOrDefault(const std::string& path, double def) const {
    return found(path) ? lookupScalar(path) : def;
}
long dictionary::lookupOrDefault(const std::string& path, long def) const {
    return found(path) ? lookupLabel(path) : def;
}
bool dictionary::lookupOrDefault(const std::string& path, bool def) const {
    return found(path) ? lookupBool(path) : def;
}

const dictionary& dictionary::subDict(const std::string& path) const {
    std::string leaf;
    const dictionary* d = find_dict(path, leaf);
    if (!d || !d->dicts_.count(leaf)) throw dictError("sub-dictionary " + path + " is undefined");
    return *d->dicts_.at(leaf);
}

}  // namespace afarhost
