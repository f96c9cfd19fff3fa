// Minimal reader of OpenFOAM dictionary files (the subset the solver reads:
// nested sub-dictionaries, `key value;` entries, // and /* */ comments, the
// FoamFile header).  Lookups use dotted paths ("CUDA.boundary.patch0").
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace afarhost {

class dictionary {
   public:
    static dictionary fromFile(const std::string& path);
    static dictionary fromString(const std::string& text);

    bool found(const std::string& path) const;
    // The tokens of an entry (everything between the key and ';').
    const std::vector<std::string>& tokens(const std::string& path) const;
    std::string lookupWord(const std::string& path) const;
    double lookupScalar(const std::string& path) const;
    long lookupLabel(const std::string& path) const;
    bool lookupBool(const std::string& path) const;
    double lookupOrDefault(const std::string& path, double def) const;
    long lookupOrDefault(const std::string& path, long def) const;
    bool lookupOrDefault(const std::string& path, bool def) const;
    const dictionary& subDict(const std::string& path) const;
    std::vector<std::string> keys() const { return order_; }

   private:
    const dictionary* find_dict(const std::string& path, std::string& leaf) const;
    std::map<std::string, std::vector<std::string>> entries_;
    std::map<std::string, std::shared_ptr<dictionary>> dicts_;
    std::vector<std::string> order_;
    friend struct parser;
};

struct dictError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

}  // namespace afarhost
