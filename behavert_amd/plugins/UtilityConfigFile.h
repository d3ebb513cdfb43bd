// UtilityConfigFile.h -- INI-style configuration with BehaveRT's semantics.
//
// Mirrors BehaveRT::UtilityConfigFile (Core/src/UtilityConfigFile.cpp:69-160) and
// BehaveRT::StringConverter (Core/src/UtilityStringConverter.cpp:86-130):
//   * "[Section]" lines open a section (one per plug-in name()); "#" and "@"
//     start comments;
//   * "key=value", "key:value" or "key<TAB>value", whitespace trimmed;
//   * a missing key reads as "" which parses to 0 / false (no exception);
//   * parseFloat3 / parseUint3 read "x y z" and give 0,0,0 unless there are three.
// Unlike the reference, section headers are recognised with or without a
// trailing '\r' (the reference assumed CRLF files).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace BehaveRT {

class UtilityConfigFile {
public:
    UtilityConfigFile() {}
    explicit UtilityConfigFile(const std::string& filename) { load(filename); }

    bool load(const std::string& filename, const std::string& separators = "\t:=") {
        m_Settings.clear();
        std::ifstream in(filename.c_str(), std::ios::in | std::ios::binary);
        if (!in) return false;
        std::string section, line;
        while (std::getline(in, line)) {
            if (!line.empty() && line[line.size() - 1] == '\r') line.erase(line.size() - 1);
            if (line.empty() || line[0] == '#' || line[0] == '@') continue;
            if (line[0] == '[' && line[line.size() - 1] == ']') {
                section = line.substr(1, line.size() - 2);
                continue;
            }
            std::string::size_type sep = line.find_first_of(separators);
            if (sep == std::string::npos) continue;
            std::string key = trim(line.substr(0, sep));
            std::string::size_type vpos = line.find_first_not_of(separators, sep);
            std::string val = vpos == std::string::npos ? "" : trim(line.substr(vpos));
            m_Settings[section].insert(std::make_pair(key, val));
        }
        return true;
    }

    void set(const std::string& section, const std::string& key, const std::string& value) {
        m_Settings[section].erase(key);
        m_Settings[section].insert(std::make_pair(key, value));
    }

    std::string getSetting(const std::string& key, const std::string& section = "",
                           const std::string& defaultValue = "") const {
        auto s = m_Settings.find(section);
        if (s == m_Settings.end()) return defaultValue;
        auto k = s->second.find(key);
        return k == s->second.end() ? std::string() : k->second;
    }

private:
    static std::string trim(const std::string& s) {
        const char* ws = " \t\r\n";
        std::string::size_type a = s.find_first_not_of(ws);
        if (a == std::string::npos) return "";
        std::string::size_type b = s.find_last_not_of(ws);
        return s.substr(a, b - a + 1);
    }
    std::map<std::string, std::multimap<std::string, std::string>> m_Settings;
};

struct StringConverter {
    static float parseFloat(const std::string& v) { return v.empty() ? 0.0f : (float)std::atof(v.c_str()); }
    static int parseInt(const std::string& v) { return v.empty() ? 0 : std::atoi(v.c_str()); }
    static unsigned int parseUnsignedInt(const std::string& v) {
        return v.empty() ? 0u : (unsigned int)std::strtoul(v.c_str(), nullptr, 10);
    }
    static bool parseBool(const std::string& v) {
        return v.compare(0, 4, "true") == 0 || v.compare(0, 3, "yes") == 0 || v.compare(0, 1, "1") == 0;
    }
    static std::vector<std::string> split(const std::string& v) {
        std::vector<std::string> out;
        std::istringstream ss(v);
        std::string t;
        while (ss >> t) out.push_back(t);
        return out;
    }
    static float3 parseFloat3(const std::string& v) {
        std::vector<std::string> p = split(v);
        if (p.size() != 3) return make_float3(0, 0, 0);
        return make_float3(parseFloat(p[0]), parseFloat(p[1]), parseFloat(p[2]));
    }
    static uint3 parseUint3(const std::string& v) {
        std::vector<std::string> p = split(v);
        if (p.size() != 3) return make_uint3(0, 0, 0);
        return make_uint3(parseUnsignedInt(p[0]), parseUnsignedInt(p[1]), parseUnsignedInt(p[2]));
    }
};

}  // namespace BehaveRT
