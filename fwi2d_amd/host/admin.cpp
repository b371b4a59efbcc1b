// admin.cpp -- see admin.hpp.
#include "admin.hpp"

#include <cctype>

#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>

namespace fwi {

namespace {

bool find_parameter(const std::string &key, const std::string &file, std::string &out)
{
    std::ifstream in(file);
    if (!in.is_open()) throw std::runtime_error("Error: " + file + " could not be opened!");
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#' || line[0] == ' ' || line[0] == '\t') continue;
        if (line.compare(0, key.size(), key) != 0) continue;
        // the key must be the whole first token
        const size_t after = key.size();
        if (after < line.size() && line[after] != ' ' && line[after] != '\t' && line[after] != '=')
            continue;
        const size_t eq = line.find('=', after);
        if (eq == std::string::npos) continue;
        std::string v;
        for (size_t i = eq + 1; i < line.size(); ++i) {
            if (line[i] == '#') break;
            v += line[i];
        }
        v.erase(std::remove_if(v.begin(), v.end(),
                               [](char c) { return c == ' ' || c == '\t' || c == '\r'; }),
                v.end());
        if (!v.empty() && (v.front() == '"' || v.front() == '['))
            v.erase(std::remove_if(v.begin(), v.end(),
                                   [](char c) { return c == '"' || c == '[' || c == ']'; }),
                    v.end());
        out = v;
        return true;
    }
    return false;
}

}  // namespace

std::string catch_parameter(const std::string &key, const std::string &file)
{
    std::string v;
    if (!find_parameter(key, file, v) || v.empty())
        throw std::runtime_error("Error: parameter '" + key + "' is missing in " + file);
    return v;
}

std::string catch_parameter(const std::string &key, const std::string &file,
                            const std::string &fallback)
{
    std::string v;
    if (!find_parameter(key, file, v) || v.empty()) return fallback;
    return v;
}

int param_int(const std::string &key, const std::string &file)
{
    return std::stoi(catch_parameter(key, file));
}

float param_float(const std::string &key, const std::string &file)
{
    return std::stof(catch_parameter(key, file));
}

int param_int(const std::string &key, const std::string &file, int fallback)
{
    std::string v = catch_parameter(key, file, "");
    return v.empty() ? fallback : std::stoi(v);
}

float param_float(const std::string &key, const std::string &file, float fallback)
{
    std::string v = catch_parameter(key, file, "");
    return v.empty() ? fallback : std::stof(v);
}

uint64_t param_u64(const std::string &key, const std::string &file, uint64_t fallback)
{
    std::string v = catch_parameter(key, file, "");
    return v.empty() ? fallback : (uint64_t)std::stoull(v, nullptr, 0);
}

bool param_bool(const std::string &key, const std::string &file, bool fallback)
{
    std::string v = catch_parameter(key, file, "");
    if (v.empty()) return fallback;
    for (char &ch : v) ch = (char)std::tolower((unsigned char)ch);
    if (v == "true") return true;
    if (v == "false") return false;
    throw std::runtime_error("Error: " + key + " must be true or false, not " + v);
}

void import_binary_float(const std::string &path, float *array, long n)
{
    std::ifstream f(path, std::ios::in | std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Error: " + path + " could not be opened!");
    f.read(reinterpret_cast<char *>(array), n * (long)sizeof(float));
    if (f.gcount() != n * (long)sizeof(float))
        throw std::runtime_error("Error: " + path + " holds fewer than " + std::to_string(n) +
                                 " float32 values");
}

void export_binary_float(const std::string &path, const float *array, long n)
{
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Error: " + path + " could not be opened!");
    f.write(reinterpret_cast<const char *>(array), n * (long)sizeof(float));
    std::cout << "Binary file " << path << " was successfully written." << std::endl;
}

std::vector<std::string> import_text_file(const std::string &path)
{
    std::ifstream f(path);
    if (!f.is_open()) throw std::runtime_error("Error: " + path + " could not be opened!");
    std::vector<std::string> out;
    std::string line;
    while (std::getline(f, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        out.push_back(line);
    }
    return out;
}

std::vector<std::string> split(const std::string &s, char delimiter)
{
    std::vector<std::string> tokens;
    std::istringstream in(s);
    std::string t;
    while (std::getline(in, t, delimiter)) tokens.push_back(t);
    return tokens;
}

void Geometry::set_parameters(const std::string &parameters)
{
    auto read = [](const std::string &path, std::vector<float> &x, std::vector<float> &z) {
        for (const std::string &line : import_text_file(path)) {
            std::vector<std::string> t = split(line, ',');
            if (t.size() < 2) throw std::runtime_error("Error: bad line '" + line + "' in " + path);
            x.push_back(std::stof(t[0]));
            z.push_back(std::stof(t[1]));
        }
    };
    read(catch_parameter("SPS", parameters), xsrc, zsrc);
    read(catch_parameter("RPS", parameters), xrec, zrec);
    nsrc = (int)xsrc.size();
    nrec = (int)xrec.size();
    if (nsrc == 0 || nrec == 0) throw std::runtime_error("Error: empty SPS or RPS file");
}

}  // namespace fwi
