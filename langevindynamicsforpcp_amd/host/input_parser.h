// input_parser.h -- the reference's input-file format and -i/-o/-p command
// line (INPUTPARSER::InputParser, inputParser.h:7-118, inputParser.cpp).
//
// File grammar: one `KEY value ...` per line; `#` starts a comment; a line
// ending in `\` continues on the next; `a = b` defines a text variable that
// is substituted in later lines; `DEFINE name value` likewise; `INCLUDE f...`
// reads other files; `FOR start end step format symbol line...` and
// `FORn` (n nested loops, 5 words each) expand lines.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace ldpcp_host {

enum class OptionType { Whatever = -2, CanBeLeftOut = -1, MustAndOnce = 0, Multiple = 1 };
enum class NumberRequirement { Whatever = -1, Equal = 0, EqualOrGreaterThan = 1 };
enum class InputType { Input = 0, Output = 1, Parameter = 2 };

struct OptionValue {
    std::string text;
    std::vector<std::string> words;
};

struct OptionDef {
    std::string key, help;
    OptionValue def;
    NumberRequirement req;
    int nvalues;
    OptionType type;
    InputType io;
};

class InputParser {
public:
    void AddOption(const std::string& key, const std::string& help, const std::string& def,
                   NumberRequirement req = NumberRequirement::EqualOrGreaterThan, int nvalues = 1,
                   OptionType type = OptionType::MustAndOnce, InputType io = InputType::Parameter);
    // -i file: read; -o file: write a commented template (re-using values of an
    // existing file, which is backed up first) and return false; -p: print the
    // parsed values and return false.  Returns true when the run should proceed.
    bool Parse(int argc, char** argv, bool print = true);
    std::vector<std::string> GetOptionValue(const std::string& key, int index = 0) const;
    size_t GetNumberOfOptionValue(const std::string& key) const;
    bool GetOptionSpecified(const std::string& key) const;
    // options beyond the reference's -i/-o/-p: the emulated ranks
    int ranks = 1;             // --ranks N: emulated MPI ranks (default: PMI/OMPI world size or 1)
    int simd_width = 8;        // --simd-width W: trajectories per rank per trial
    unsigned long steps_per_launch = 1ul << 20;

private:
    bool ReadInputFile(const std::string& file, bool print, bool output_mode);
    bool PrintEmptyInputFile(const std::string& file, bool print);
    bool CheckDefinedEntry(bool print) const;
    bool OpenInput(const std::string& file, std::vector<std::string>* content, bool print);
    bool ProcessInputLines(const std::vector<std::string>& content, bool print, bool output_mode);
    bool AddInputValue(int index, const std::string& value, bool print);
    void ExpandLoops(const std::vector<int>& start, const std::vector<int>& end, const std::vector<int>& inc,
                     const std::vector<std::string>& fmt, const std::vector<std::string>& sym, int level,
                     const std::string& line, std::vector<std::string>* content);
    void PrintAllValues(FILE* f) const;
    void PrintHelp() const;

    std::vector<OptionDef> defs_;
    std::vector<std::vector<OptionValue>> values_;
    std::vector<bool> specified_;
    std::map<std::string, int> index_;
    std::map<std::string, std::string> variables_, defines_;
    std::string input_file_;
};

}  // namespace ldpcp_host
