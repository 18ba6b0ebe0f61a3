// text_util.h -- string helpers with the parsing semantics of the reference's
// `text` class (text.cpp): whitespace splitting, comment stripping, value
// conversions, and the "#name_bk.N#" backup naming of MOLUTILITY.cpp:6-22.
#pragma once
#include <string>
#include <vector>

namespace ldpcp_host {

// split on any character of `mask`, dropping empty fields (text::split, text.cpp:287-308)
std::vector<std::string> split(const std::string& s, const std::string& mask = " \t");
// cut at the first character of `symbols` (text::parseRemoveComment, text.cpp:416-420)
std::string remove_comment(const std::string& s, const std::string& symbols);
// only spaces and tabs (text::isEmpty, text.cpp:594-596)
bool is_blank(const std::string& s);
std::string replace_all(const std::string& s, const std::string& word, const std::string& with);
bool iequal_words(const std::string& a, const std::string& b);
// words[first..] joined by single spaces (text::FromTextArray, text.cpp:862-873)
std::string join(const std::vector<std::string>& w, size_t first);
bool ends_with(const std::string& s, const std::string& suffix);
std::string format_int(const char* fmt, int v);

// value conversions (text.cpp:215-285): std::stod / stoi / stoul semantics,
// "Error: Invalid string or overflow" and 0 on malformed input
double to_double(const std::string& s);
int to_int(const std::string& s);
unsigned long to_ulong(const std::string& s);
bool is_int(const std::string& s);
// TRUE/True/true or any non-zero number (text::to_boolean, text.cpp:278-285)
bool to_boolean(const std::string& s);
// "true" / "false" (text::text(bool), text.cpp:15-17)
const char* bool_text(bool b);
// "%13.7f %s" with B/K/M/G/T units (text::FromByte2Readable, text.cpp:893-904)
std::string bytes_readable(double bytes);
// seconds -> "%f %s" in s/m/h/d (clsSystemTime::Print_Time_Diff, clsTime.cpp:28-45)
std::string time_readable(double seconds, const char* fmt = "%f %s");

bool file_exists(const std::string& path);
// if `path` exists, rename it to "<dir>/#<name>_bk.N#" with the first free N
void backup_existing(const std::string& path);

}  // namespace ldpcp_host
