#include "props.h"

#include <cctype>
#include <sstream>

namespace sg {

Props &Props::set(const std::string &k, const std::string &v) {
    if (!props_.count(k)) keys_.push_back(k);
    props_[k] = v;
    return *this;
}

int Props::get_int(const std::string &k, int def) const {
    if (!has(k)) return def;
    try {
        return std::stoi(get(k));
    } catch (...) {
        throw Error("Could not convert property \"" + k + "\" (" + get(k) +
                    ") to integer!");
    }
}

float Props::get_float(const std::string &k, float def) const {
    if (!has(k)) return def;
    try {
        return std::stof(get(k));
    } catch (...) {
        throw Error("Could not convert property \"" + k + "\" (" + get(k) +
                    ") to float!");
    }
}

Props &Props::set_child(const std::string &k, const Props &child) {
    if (!children_.count(k)) child_keys_.push_back(k);
    children_[k].reset(new Props(child));
    return *this;
}

Props &Props::add_to_array(const std::string &k, const std::string &v) {
    if (!arrays_.count(k)) array_keys_.push_back(k);
    arrays_[k].push_back(v);
    return *this;
}

std::vector<Props *> Props::get_child_array(const std::string &k) const {
    std::vector<Props *> out;
    auto it = child_arrays_.find(k);
    if (it != child_arrays_.end())
        for (auto &p : it->second) out.push_back(p.get());
    return out;
}

Props &Props::add_to_child_array(const std::string &k, const Props &child) {
    if (!child_arrays_.count(k)) child_array_keys_.push_back(k);
    child_arrays_[k].emplace_back(new Props(child));
    return *this;
}

void Props::clear() {
    props_.clear(); keys_.clear();
    children_.clear(); child_keys_.clear();
    arrays_.clear(); array_keys_.clear();
    child_arrays_.clear(); child_array_keys_.clear();
}

void Props::copy_from(const Props &o) {
    props_ = o.props_;
    keys_ = o.keys_;
    for (auto &k : o.child_keys_) {
        children_[k].reset(new Props(*o.children_.at(k)));
        child_keys_.push_back(k);
    }
    arrays_ = o.arrays_;
    array_keys_ = o.array_keys_;
    for (auto &k : o.child_array_keys_) {
        auto &dst = child_arrays_[k];
        for (auto &p : o.child_arrays_.at(k)) dst.emplace_back(new Props(*p));
        child_array_keys_.push_back(k);
    }
}

// ---------------------------------------------------------------------------
// Minimal JSON reader.  Object keys are visited in sorted order, as the
// reference's JSON library stores objects in a std::map.

namespace {

struct Json {
    enum Kind { NUL, STR, OTHER, OBJ, ARR } kind = NUL;
    std::string str;
    std::map<std::string, Json> obj;
    std::vector<Json> arr;
};

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &t) : s(t) {}

    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    [[noreturn]] void fail(const char *what) {
        throw Error(std::string("JSON parse error: ") + what + " at offset " +
                    std::to_string(i));
    }
    char peek() { ws(); if (i >= s.size()) fail("unexpected end"); return s[i]; }

    std::string string_lit() {
        if (s[i] != '"') fail("expected string");
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                char e = s[i++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (i + 4 > s.size()) fail("bad unicode escape");
                        unsigned cp = std::stoul(s.substr(i, 4), nullptr, 16);
                        i += 4;
                        if (cp < 0x80) out += (char)cp;
                        else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }

    Json value() {
        Json v;
        char c = peek();
        if (c == '{') {
            v.kind = Json::OBJ;
            ++i;
            if (peek() == '}') { ++i; return v; }
            while (true) {
                ws();
                std::string k = string_lit();
                if (peek() != ':') fail("expected ':'");
                ++i;
                v.obj[k] = value();
                char d = peek();
                ++i;
                if (d == '}') break;
                if (d != ',') fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = Json::ARR;
            ++i;
            if (peek() == ']') { ++i; return v; }
            while (true) {
                v.arr.push_back(value());
                char d = peek();
                ++i;
                if (d == ']') break;
                if (d != ',') fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = Json::STR;
            v.str = string_lit();
        } else {
            // number, true, false, null: skipped like the reference does
            v.kind = Json::OTHER;
            while (i < s.size() && s[i] != ',' && s[i] != '}' && s[i] != ']' &&
                   !std::isspace((unsigned char)s[i]))
                ++i;
        }
        return v;
    }
};

void to_props(const Json &o, Props &p) {
    for (auto &kv : o.obj) {
        const Json &v = kv.second;
        if (v.kind == Json::STR) {
            p.set(kv.first, v.str);
        } else if (v.kind == Json::OBJ) {
            Props child;
            to_props(v, child);
            p.set_child(kv.first, child);
        } else if (v.kind == Json::ARR) {
            for (auto &item : v.arr) {
                if (item.kind == Json::OBJ) {
                    Props child;
                    to_props(item, child);
                    p.add_to_child_array(kv.first, child);
                } else if (item.kind == Json::STR) {
                    p.add_to_array(kv.first, item.str);
                }
            }
        }
    }
}

std::string quote(const std::string &s) {
    std::string out = "\"";
    for (char c : s) {
        switch (c) {
            case '"': out += "\\\""; break;
            case '\\': out += "\\\\"; break;
            case '\n': out += "\\n"; break;
            case '\t': out += "\\t"; break;
            default: out += c;
        }
    }
    return out + "\"";
}

void write(const Props &p, std::ostringstream &os, int indent) {
    std::string pad(indent + 4, ' '), end(indent, ' ');
    std::vector<std::string> items;
    for (auto &k : p.keys()) {
        std::string v = p.get(k);
        if (!v.empty()) items.push_back(pad + quote(k) + ": " + quote(v));
    }
    for (auto &k : p.child_keys()) {
        std::ostringstream c;
        write(*p.get_child(k), c, indent + 4);
        items.push_back(pad + quote(k) + ": " + c.str());
    }
    for (auto &k : p.array_keys()) {
        auto arr = p.get_array(k);
        if (arr.empty()) continue;
        std::string a = pad + quote(k) + ": [";
        for (size_t j = 0; j < arr.size(); ++j)
            a += (j ? ", " : "") + quote(arr[j]);
        items.push_back(a + "]");
    }
    for (auto &k : p.child_array_keys()) {
        auto arr = p.get_child_array(k);
        if (arr.empty()) continue;
        std::string a = pad + quote(k) + ": [\n";
        for (size_t j = 0; j < arr.size(); ++j) {
            std::ostringstream c;
            write(*arr[j], c, indent + 8);
            a += std::string(indent + 8, ' ') + c.str() +
                 (j + 1 < arr.size() ? ",\n" : "\n");
        }
        items.push_back(a + pad + "]");
    }
    os << "{\n";
    for (size_t j = 0; j < items.size(); ++j)
        os << items[j] << (j + 1 < items.size() ? ",\n" : "\n");
    os << end << "}";
}

}  // namespace

Props parse_json(const std::string &text) {
    Parser p(text);
    Json root = p.value();
    if (root.kind != Json::OBJ) throw Error("JSON root must be an object");
    Props out;
    to_props(root, out);
    return out;
}

std::string write_json(const Props &p) {
    std::ostringstream os;
    write(p, os, 0);
    os << "\n";
    return os.str();
}

}  // namespace sg
