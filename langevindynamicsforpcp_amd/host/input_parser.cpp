// input_parser.cpp -- see input_parser.h; behaviour follows inputParser.cpp:45-520.
#include "input_parser.h"

#include <stdio.h>
#include <stdlib.h>

#include <fstream>

#include "text_util.h"

namespace ldpcp_host {

static const char* io_text(InputType t)
{
    switch (t) {
    case InputType::Input: return "Input";
    case InputType::Output: return "Output";
    default: return "Parameter";
    }
}

void InputParser::AddOption(const std::string& key, const std::string& help, const std::string& def,
                            NumberRequirement req, int nvalues, OptionType type, InputType io)
{
    OptionDef d;
    d.key = key;
    d.help = help;
    d.def.text = def;
    d.def.words = split(def);
    d.req = req;
    d.nvalues = nvalues;
    d.type = type;
    d.io = io;
    defs_.push_back(d);
}

void InputParser::PrintHelp() const
{
    fprintf(stderr, "command usage\n");
    fprintf(stderr, "    Optional   %-5s %-14s %8s %1d %-20s Default: none\n", "-i", "--input", "", 1,
            "Input file name");
    fprintf(stderr, "    Optional   %-5s %-14s %8s %1d %-20s Default: none\n", "-o", "--output", "", 1,
            "Input file name to be generated");
    fprintf(stderr, "    Optional   %-5s %-14s %8s %1d %-20s Default: \n", "-p", "--print", "", 0,
            "Print input file content");
    fprintf(stderr, "    Optional   %-5s %-14s %8s %1d %-20s Default: 1\n", "", "--ranks", "", 1,
            "Emulated MPI ranks");
    fprintf(stderr, "    Optional   %-5s %-14s %8s %1d %-20s Default: 8\n", "", "--simd-width", "", 1,
            "Trajectories per rank");
}

bool InputParser::Parse(int argc, char** argv, bool print)
{
    std::string in_file, out_file;
    bool has_i = false, has_o = false, has_p = false;
    const char* env_ranks = getenv("LDPCP_RANKS");
    if (!env_ranks) env_ranks = getenv("OMPI_COMM_WORLD_SIZE");
    if (!env_ranks) env_ranks = getenv("PMI_SIZE");
    if (env_ranks) ranks = atoi(env_ranks) > 0 ? atoi(env_ranks) : 1;
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        auto next = [&](std::string& dst) {
            if (k + 1 >= argc) return false;
            dst = argv[++k];
            return true;
        };
        std::string v;
        if (a == "-i" || a == "--input") {
            if (!next(in_file)) { PrintHelp(); return false; }
            has_i = true;
        } else if (a == "-o" || a == "--output") {
            if (!next(out_file)) { PrintHelp(); return false; }
            has_o = true;
        } else if (a == "-p" || a == "--print") {
            has_p = true;
        } else if (a == "--ranks" && next(v)) {
            ranks = atoi(v.c_str()) > 0 ? atoi(v.c_str()) : 1;
        } else if (a == "--simd-width" && next(v)) {
            simd_width = atoi(v.c_str()) > 0 ? atoi(v.c_str()) : 1;
        } else if (a == "--steps-per-launch" && next(v)) {
            steps_per_launch = strtoul(v.c_str(), nullptr, 10);
        } else {
            if (print) PrintHelp();
            return false;
        }
    }
    if (!CheckDefinedEntry(print)) return false;
    if (has_i && has_o) {
        if (print) fprintf(stderr, "Error (InputParser): only one of -i and -o can be specified\n");
        return false;
    }
    if (has_i) {
        input_file_ = in_file;
        if (!ReadInputFile(in_file, print, false)) return false;
        if (has_p) {
            if (print) PrintAllValues(stderr);
            return false;
        }
        return true;
    }
    if (has_o) {
        PrintEmptyInputFile(out_file, print);
        return false;
    }
    if (print) PrintHelp();
    return false;
}

std::vector<std::string> InputParser::GetOptionValue(const std::string& key, int index) const
{
    auto it = index_.find(key);
    if (it == index_.end()) {
        fprintf(stderr, "Error (InputParser): undefined option (%s)\n", key.c_str());
        return {};
    }
    return values_[it->second][index].words;
}

size_t InputParser::GetNumberOfOptionValue(const std::string& key) const
{
    auto it = index_.find(key);
    if (it == index_.end()) {
        fprintf(stderr, "Error (InputParser): undefined option (%s)\n", key.c_str());
        return 0;
    }
    return values_[it->second].size();
}

bool InputParser::GetOptionSpecified(const std::string& key) const
{
    auto it = index_.find(key);
    if (it == index_.end()) {
        fprintf(stderr, "Error (InputParser): undefined option (%s)\n", key.c_str());
        return false;
    }
    return specified_[it->second];
}

void InputParser::PrintAllValues(FILE* f) const
{
    for (size_t i = 0; i < values_.size(); ++i)
        for (size_t j = 0; j < values_[i].size(); ++j)
            fprintf(f, "%-20s %3d %s\n", defs_[i].key.c_str(), (int)j, values_[i][j].text.c_str());
}

bool InputParser::ReadInputFile(const std::string& file, bool print, bool output_mode)
{
    std::vector<std::string> content;
    if (!OpenInput(file, &content, print)) return false;
    return ProcessInputLines(content, print, output_mode);
}

bool InputParser::PrintEmptyInputFile(const std::string& file, bool print)
{
    (void)print;
    bool reuse = false;
    if (file_exists(file)) reuse = ReadInputFile(file, false, true);
    backup_existing(file);
    FILE* f = fopen(file.c_str(), "w");
    if (!f) {
        fprintf(stderr, "Error (InputParser): Can't open empty input file (%s)\n", file.c_str());
        return false;
    }
    for (size_t i = 0; i < defs_.size(); ++i) {
        fprintf(f, "# [%s] %s\n", io_text(defs_[i].io), defs_[i].help.c_str());
        if (reuse && i < values_.size() && !values_[i].empty()) {
            for (const auto& v : values_[i]) fprintf(f, "%-20s %s\n", defs_[i].key.c_str(), v.text.c_str());
        } else {
            fprintf(f, "%-20s %s\n", defs_[i].key.c_str(), defs_[i].def.text.c_str());
        }
    }
    fclose(f);
    return true;
}

bool InputParser::CheckDefinedEntry(bool print) const
{
    bool ok = true;
    for (const auto& d : defs_) {
        const bool bad = (d.req == NumberRequirement::Equal && (int)d.def.words.size() != d.nvalues) ||
                         (d.req == NumberRequirement::EqualOrGreaterThan && (int)d.def.words.size() < d.nvalues);
        if (bad) {
            if (print)
                fprintf(stderr, "Error (InputParser): (%s) requires %d values but got %zd default values\n",
                        d.key.c_str(), d.nvalues, d.def.words.size());
            ok = false;
        }
    }
    return ok;
}

// `a = b` with exactly one word on each side (inputParser.cpp:383-401)
static bool is_assignment(const std::string& line, std::string& lhs, std::string& rhs)
{
    if (line.find('=') == std::string::npos) return false;
    const auto parts = split(line, " \t=");
    if (parts.size() != 2) return false;
    const auto l = split(parts[0]), r = split(parts[1]);
    if (l.size() != 1 || r.size() != 1) return false;
    lhs = l[0];
    rhs = r[0];
    return true;
}

void InputParser::ExpandLoops(const std::vector<int>& start, const std::vector<int>& end, const std::vector<int>& inc,
                              const std::vector<std::string>& fmt, const std::vector<std::string>& sym, int level,
                              const std::string& line, std::vector<std::string>* content)
{
    const int n = (int)start.size();
    for (int i = start[level]; i < end[level]; i += inc[level]) {
        const std::string l = replace_all(line, sym[level], format_int(fmt[level].c_str(), i));
        if (level == n - 1) content->push_back(l);
        else ExpandLoops(start, end, inc, fmt, sym, level + 1, l, content);
        if (inc[level] <= 0) break;
    }
}

bool InputParser::OpenInput(const std::string& file, std::vector<std::string>* content, bool print)
{
    std::ifstream in(file);
    if (!in) {
        if (print) fprintf(stderr, "Error (InputParser): can't open input file (%s)\n", input_file_.c_str());
        return false;
    }
    std::vector<std::string> lines;
    bool append = false;
    std::string raw;
    while (std::getline(in, raw)) {
        std::string line = remove_comment(raw, "#");
        if (append) lines.back() += line;
        else lines.push_back(line);
        if (ends_with(line, "\\")) {
            lines.back().resize(lines.back().size() - 1);
            append = true;
        } else {
            append = false;
        }
    }
    bool ok = true;
    for (std::string line : lines) {
        if (is_blank(line)) continue;
        for (const auto& d : defines_) line = replace_all(line, d.first, d.second);
        for (const auto& v : variables_) line = replace_all(line, v.first, v.second);
        const auto w = split(line);
        std::string lhs, rhs;
        if (is_assignment(line, lhs, rhs)) {
            variables_[lhs] = rhs;
        } else if (iequal_words(w[0], "include")) {
            if (w.size() >= 2) {
                for (size_t k = 1; k < w.size() && ok; ++k) ok = OpenInput(w[k], content, print);
            } else {
                if (print)
                    fprintf(stderr, "Error (InputParser): %s requires at least 1 argument, but now %d\n", w[0].c_str(),
                            (int)w.size() - 1);
                ok = false;
            }
        } else if (iequal_words(w[0], "define")) {
            if (w.size() == 3) {
                defines_[w[1]] = w[2];
            } else {
                if (print)
                    fprintf(stderr, "Error (InputParser): %s requires at least 1 argument, but now %d\n", w[0].c_str(),
                            (int)w.size() - 1);
                ok = false;
            }
        } else if (iequal_words(w[0], "for")) {
            if (w.size() >= 7) {
                const int s = to_int(w[1]), e = to_int(w[2]), st = to_int(w[3]);
                const std::string cmd = join(w, 6);
                for (int j = s; j < e; j += st) {
                    content->push_back(replace_all(cmd, w[5], format_int(w[4].c_str(), j)));
                    if (st <= 0) break;
                }
            }
        } else if (w[0].size() >= 3 && iequal_words(w[0].substr(0, 3), "for")) {
            const std::string n_text = w[0].substr(3);
            if (is_int(n_text)) {
                const int n = to_int(n_text);
                if ((int)w.size() < n * 5 + 1 || n <= 0) {
                    if (print) fprintf(stderr, "Error (InputParser): %s needs %d loop words\n", w[0].c_str(), n * 5);
                    ok = false;
                } else {
                    std::vector<int> s(n), e(n), st(n);
                    std::vector<std::string> f(n), sy(n);
                    for (int k = 0; k < n; ++k) {
                        s[k] = to_int(w[k * 5 + 1]);
                        e[k] = to_int(w[k * 5 + 2]);
                        st[k] = to_int(w[k * 5 + 3]);
                        f[k] = w[k * 5 + 4];
                        sy[k] = w[k * 5 + 5];
                    }
                    ExpandLoops(s, e, st, f, sy, 0, join(w, n * 5 + 1), content);
                }
            } else if (print) {
                fprintf(stderr, "Error (InputParser): for is not followed by an integer (%s)\n", n_text.c_str());
            }
        } else {
            content->push_back(line);
        }
        if (!ok) break;
    }
    return ok;
}

bool InputParser::ProcessInputLines(const std::vector<std::string>& content, bool print, bool output_mode)
{
    for (size_t i = 0; i < defs_.size(); ++i) index_[defs_[i].key] = (int)i;
    values_.assign(defs_.size(), {});
    specified_.assign(defs_.size(), false);
    for (const auto& line : content) {
        if (is_blank(line)) continue;
        const auto w = split(line);
        auto it = index_.find(w[0]);
        if (it == index_.end()) {
            if (print) fprintf(stderr, "Error (InputParser): undefined option (%s)\n", w[0].c_str());
            if (!output_mode) return false;
            continue;
        }
        const int k = it->second;
        specified_[k] = true;
        const std::string value = join(w, 1);
        if (defs_[k].type == OptionType::MustAndOnce || defs_[k].type == OptionType::CanBeLeftOut) {
            if (values_[k].empty()) {
                if (!AddInputValue(k, value, print) && !output_mode) return false;
            } else {
                if (print) fprintf(stderr, "Error (InputParser): (%s) can be specified only once\n", w[0].c_str());
                if (!output_mode) return false;
            }
        } else {
            if (!AddInputValue(k, value, print) && !output_mode) return false;
        }
    }
    for (size_t i = 0; i < defs_.size(); ++i) {
        if ((defs_[i].type == OptionType::MustAndOnce || defs_[i].type == OptionType::Multiple) && values_[i].empty()) {
            if (print) fprintf(stderr, "Error (InputParser): (%s) must be specified at least once\n", defs_[i].key.c_str());
            if (!output_mode) return false;
        }
    }
    for (size_t i = 0; i < defs_.size(); ++i)
        if ((defs_[i].type == OptionType::CanBeLeftOut || defs_[i].type == OptionType::Whatever) && values_[i].empty())
            AddInputValue((int)i, defs_[i].def.text, false);
    return true;
}

bool InputParser::AddInputValue(int k, const std::string& value, bool print)
{
    const auto w = split(value);
    const OptionDef& d = defs_[k];
    if (d.req == NumberRequirement::Equal && (int)w.size() != d.nvalues) {
        if (print)
            fprintf(stderr, "Error (InputParser): (%s) requires exactly %d values but got %zd (%s)\n", d.key.c_str(),
                    d.nvalues, w.size(), value.c_str());
        return false;
    }
    if (d.req == NumberRequirement::EqualOrGreaterThan && (int)w.size() < d.nvalues) {
        if (print)
            fprintf(stderr, "Error (InputParser): (%s) requires no less than %d values but got %zd\n", d.key.c_str(),
                    d.nvalues, w.size());
        return false;
    }
    values_[k].push_back(OptionValue{value, w});
    return true;
}

}  // namespace ldpcp_host
