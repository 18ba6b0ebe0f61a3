// text_util.cpp -- see text_util.h
#include "text_util.h"

#include <ctype.h>
#include <stdarg.h>
#include <stdio.h>
#include <unistd.h>

#include <stdexcept>

namespace ldpcp_host {

std::vector<std::string> split(const std::string& s, const std::string& mask)
{
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (mask.find(c) != std::string::npos) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

std::string remove_comment(const std::string& s, const std::string& symbols)
{
    const size_t pos = s.find_first_of(symbols);
    return pos == std::string::npos ? s : s.substr(0, pos);
}

bool is_blank(const std::string& s) { return s.find_first_not_of(" \t") == std::string::npos; }

std::string replace_all(const std::string& s, const std::string& word, const std::string& with)
{
    if (word.empty()) return s;
    std::string out;
    size_t i = 0;
    while (true) {
        const size_t p = s.find(word, i);
        if (p == std::string::npos) break;
        out.append(s, i, p - i);
        out += with;
        i = p + word.size();
    }
    out.append(s, i, std::string::npos);
    return out;
}

bool iequal_words(const std::string& a, const std::string& b)
{
    const auto wa = split(a), wb = split(b);
    if (wa.size() != wb.size()) return false;
    for (size_t i = 0; i < wa.size(); ++i) {
        if (wa[i].size() != wb[i].size()) return false;
        for (size_t j = 0; j < wa[i].size(); ++j)
            if (tolower((unsigned char)wa[i][j]) != tolower((unsigned char)wb[i][j])) return false;
    }
    return true;
}

std::string join(const std::vector<std::string>& w, size_t first)
{
    std::string out;
    for (size_t i = first; i < w.size(); ++i) {
        out += w[i];
        if (i + 1 != w.size()) out += " ";
    }
    return out;
}

bool ends_with(const std::string& s, const std::string& suffix)
{
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

std::string format_int(const char* fmt, int v)
{
    char buf[256];
    snprintf(buf, sizeof(buf), fmt, v);
    return buf;
}

double to_double(const std::string& s)
{
    try {
        return std::stod(s);
    } catch (const std::invalid_argument&) {
        fprintf(stderr, "Error: Invalid string or overflow %s\n", s.c_str());
        return 0;
    }
}

int to_int(const std::string& s)
{
    try {
        return std::stoi(s);
    } catch (const std::invalid_argument&) {
        fprintf(stderr, "Error: Invalid string or overflow %s\n", s.c_str());
        return 0;
    }
}

unsigned long to_ulong(const std::string& s)
{
    try {
        return std::stoul(s);
    } catch (const std::invalid_argument&) {
        fprintf(stderr, "Error: Invalid string or overflow %s\n", s.c_str());
        return 0;
    }
}

bool is_int(const std::string& s)
{
    try {
        (void)std::stoi(s);
        return true;
    } catch (const std::exception&) {
        return false;
    }
}

static bool is_double(const std::string& s)
{
    try {
        (void)std::stod(s);
        return true;
    } catch (const std::exception&) {
        return false;
    }
}

bool to_boolean(const std::string& s)
{
    if (s == "TRUE" || s == "True" || s == "true") return true;
    if (is_int(s) && std::stoi(s) != 0) return true;
    if (is_double(s) && std::stod(s) != 0) return true;
    return false;
}

const char* bool_text(bool b) { return b ? "true" : "false"; }

std::string bytes_readable(double bytes)
{
    static const char* unit[5] = {"B", "K", "M", "G", "T"};
    int k = 0;
    while (bytes > 1024.0 && k < 4) {
        bytes /= 1024.0;
        k++;
    }
    char buf[64];
    snprintf(buf, sizeof(buf), "%13.7f %s", bytes, unit[k]);
    return buf;
}

std::string time_readable(double seconds, const char* fmt)
{
    double out = seconds;
    const char* unit = "s";
    double in = out / 60.0;
    if (in > 1.0) {
        out = in;
        unit = "m";
        in = out / 60.0;
        if (in > 1.0) {
            out = in;
            unit = "h";
            in = out / 24.0;
            if (in > 1.0) {
                out = in;
                unit = "d";
            }
        }
    }
    char buf[128];
    snprintf(buf, sizeof(buf), fmt, out, unit);
    return buf;
}

bool file_exists(const std::string& path) { return access(path.c_str(), F_OK) != -1; }

void backup_existing(const std::string& path)
{
    if (!file_exists(path)) return;
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
    const std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    for (int n = 1;; ++n) {
        const std::string bk = dir + "#" + name + "_bk." + std::to_string(n) + "#";
        if (!file_exists(bk)) {
            rename(path.c_str(), bk.c_str());
            return;
        }
    }
}

}  // namespace ldpcp_host
