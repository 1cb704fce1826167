// String-keyed configuration tree: properties, child trees, string arrays
// and arrays of child trees, each remembering insertion order.  Same data
// model and accessor semantics as the reference's PropertyConfig
// (util/property_config.h:26-114): get_bool is `value == "true"`, numeric
// getters parse with stoi/stof, children and array members are deep copies.
#ifndef SG_PROPS_H
#define SG_PROPS_H

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace sg {

class Props {
  public:
    Props() = default;
    Props(const Props &other) { copy_from(other); }
    Props &operator=(const Props &other) {
        if (this != &other) { clear(); copy_from(other); }
        return *this;
    }
    Props(std::initializer_list<std::pair<std::string, std::string>> kv) {
        for (auto &p : kv) set(p.first, p.second);
    }

    // properties
    bool has(const std::string &k) const { return props_.count(k) > 0; }
    std::string get(const std::string &k, const std::string &def = "") const {
        auto it = props_.find(k);
        return it == props_.end() ? def : it->second;
    }
    const char *get_c_str(const std::string &k) const {
        return props_.at(k).c_str();
    }
    Props &set(const std::string &k, const std::string &v);
    int get_int(const std::string &k, int def) const;
    float get_float(const std::string &k, float def) const;
    bool get_bool(const std::string &k, bool def) const {
        return has(k) ? get(k) == "true" : def;
    }
    const std::vector<std::string> &keys() const { return keys_; }

    // children
    bool has_child(const std::string &k) const { return children_.count(k) > 0; }
    Props *get_child(const std::string &k) const {
        auto it = children_.find(k);
        return it == children_.end() ? nullptr : it->second.get();
    }
    Props &set_child(const std::string &k, const Props &child);
    const std::vector<std::string> &child_keys() const { return child_keys_; }

    // string arrays
    bool has_array(const std::string &k) const { return arrays_.count(k) > 0; }
    std::vector<std::string> get_array(const std::string &k) const {
        auto it = arrays_.find(k);
        return it == arrays_.end() ? std::vector<std::string>() : it->second;
    }
    Props &add_to_array(const std::string &k, const std::string &v);
    const std::vector<std::string> &array_keys() const { return array_keys_; }

    // child arrays
    bool has_child_array(const std::string &k) const {
        return child_arrays_.count(k) > 0;
    }
    std::vector<Props *> get_child_array(const std::string &k) const;
    Props &add_to_child_array(const std::string &k, const Props &child);
    const std::vector<std::string> &child_array_keys() const {
        return child_array_keys_;
    }

  private:
    void clear();
    void copy_from(const Props &o);

    std::map<std::string, std::string> props_;
    std::vector<std::string> keys_;
    std::map<std::string, std::unique_ptr<Props>> children_;
    std::vector<std::string> child_keys_;
    std::map<std::string, std::vector<std::string>> arrays_;
    std::vector<std::string> array_keys_;
    std::map<std::string, std::vector<std::unique_ptr<Props>>> child_arrays_;
    std::vector<std::string> child_array_keys_;
};

// JSON <-> Props (builder.cpp:15-67 semantics: string values, objects,
// arrays of strings or objects; other JSON values are ignored on read).
Props parse_json(const std::string &text);
std::string write_json(const Props &p);

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

}  // namespace sg

#endif
