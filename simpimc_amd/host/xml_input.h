// xml_input.h -- minimal reader for simpimc's XML input files.
//
// Mirrors the attribute access the reference makes through its Input class
// (GetAttribute<T>(name, default), GetChild, GetChildList, GetName; see e.g.
// data_structures/path_class.h:25-40): elements with attributes and children;
// text content, comments, processing instructions and DOCTYPEs are skipped.
// Attribute names are case sensitive, as in the reference.
#pragma once

#include <cstdlib>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace spimc_host {

class Input {
   public:
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<Input> children;

    static Input Load(const std::string& path) {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("cannot open input file " + path);
        std::stringstream ss;
        ss << f.rdbuf();
        std::string s = ss.str();
        size_t pos = 0;
        Input root;
        bool found = false;
        while (true) {
            skip_misc(s, pos);
            if (pos >= s.size()) break;
            if (s[pos] != '<') {
                ++pos;
                continue;
            }
            root = parse_element(s, pos);
            found = true;
            break;
        }
        if (!found) throw std::runtime_error("no root element in " + path);
        return root;
    }

    const std::string& GetName() const { return name; }

    bool Has(const std::string& key) const {
        for (auto& a : attrs)
            if (a.first == key) return true;
        return false;
    }

    template <class T>
    T GetAttribute(const std::string& key) const {
        for (auto& a : attrs)
            if (a.first == key) return convert<T>(a.second, key);
        throw std::runtime_error("missing attribute '" + key + "' on <" + name + ">");
    }

    template <class T>
    T GetAttribute(const std::string& key, const T& def) const {
        for (auto& a : attrs)
            if (a.first == key) return convert<T>(a.second, key);
        return def;
    }

    const Input& GetChild(const std::string& n) const {
        for (auto& c : children)
            if (c.name == n) return c;
        throw std::runtime_error("missing element <" + n + "> in <" + name + ">");
    }

    bool HasChild(const std::string& n) const {
        for (auto& c : children)
            if (c.name == n) return true;
        return false;
    }

    std::vector<Input> GetChildList(const std::string& n) const {
        std::vector<Input> out;
        for (auto& c : children)
            if (c.name == n) out.push_back(c);
        return out;
    }

    const std::vector<Input>& GetChildList() const { return children; }

   private:
    template <class T>
    static T convert(const std::string& v, const std::string& key);

    static void skip_ws(const std::string& s, size_t& p) {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }

    static void skip_misc(const std::string& s, size_t& p) {
        while (true) {
            skip_ws(s, p);
            if (s.compare(p, 4, "<!--") == 0) {
                size_t e = s.find("-->", p);
                p = (e == std::string::npos) ? s.size() : e + 3;
            } else if (s.compare(p, 2, "<?") == 0 || s.compare(p, 2, "<!") == 0) {
                size_t e = s.find('>', p);
                p = (e == std::string::npos) ? s.size() : e + 1;
            } else {
                return;
            }
        }
    }

    static std::string read_name(const std::string& s, size_t& p) {
        size_t b = p;
        while (p < s.size() && !isspace((unsigned char)s[p]) && s[p] != '>' && s[p] != '/' && s[p] != '=') ++p;
        return s.substr(b, p - b);
    }

    static Input parse_element(const std::string& s, size_t& p) {
        Input e;
        ++p;  // '<'
        e.name = read_name(s, p);
        while (true) {
            skip_ws(s, p);
            if (p >= s.size()) throw std::runtime_error("unterminated element <" + e.name + ">");
            if (s[p] == '/') {
                p += 2;  // "/>"
                return e;
            }
            if (s[p] == '>') {
                ++p;
                break;
            }
            std::string key = read_name(s, p);
            skip_ws(s, p);
            if (p >= s.size() || s[p] != '=') throw std::runtime_error("bad attribute '" + key + "' on <" + e.name + ">");
            ++p;
            skip_ws(s, p);
            char q = s[p++];
            size_t end = s.find(q, p);
            if (end == std::string::npos) throw std::runtime_error("unterminated attribute value");
            e.attrs.emplace_back(key, s.substr(p, end - p));
            p = end + 1;
        }
        while (true) {
            skip_misc(s, p);
            if (p >= s.size()) throw std::runtime_error("unterminated element <" + e.name + ">");
            if (s.compare(p, 2, "</") == 0) {
                size_t end = s.find('>', p);
                p = end + 1;
                return e;
            }
            if (s[p] == '<') {
                e.children.push_back(parse_element(s, p));
            } else {
                ++p;  // text content
            }
        }
    }
};

template <>
inline std::string Input::convert<std::string>(const std::string& v, const std::string&) {
    return v;
}
template <>
inline double Input::convert<double>(const std::string& v, const std::string& key) {
    char* end = nullptr;
    double x = std::strtod(v.c_str(), &end);
    if (end == v.c_str()) throw std::runtime_error("attribute '" + key + "' is not a number: " + v);
    return x;
}
template <>
inline int Input::convert<int>(const std::string& v, const std::string& key) {
    return (int)convert<double>(v, key);
}
template <>
inline uint32_t Input::convert<uint32_t>(const std::string& v, const std::string& key) {
    return (uint32_t)convert<double>(v, key);
}
template <>
inline bool Input::convert<bool>(const std::string& v, const std::string& key) {
    if (v == "true" || v == "True") return true;
    if (v == "false" || v == "False") return false;
    return convert<double>(v, key) != 0.;
}

}  // namespace spimc_host
