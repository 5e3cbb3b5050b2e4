// Command-line plumbing of the xenodon CLI: the argument parser
// (src/core/arg_parse.{h,cpp}), the headless configuration grammar
// (src/core/Parser.cpp, src/core/Config.h, HeadlessConfig.cpp), the logger
// (src/core/Logger.cpp) and the few fmt-style formats the CLI prints.
#pragma once

#include <cstdint>
#include <cstdio>
#include <fstream>
#include <functional>
#include <istream>
#include <memory>
#include <string>
#include <vector>

#include "volume_io.hpp"

namespace xdcli {

// printf-style std::string.
std::string sformat(const char* f, ...) __attribute__((format(printf, 1, 2)));

// fmt 5.3's "{}" of a double: printf "%g" (six significant digits).
std::string fmt_double(double v);

// fmt's "{:n}" of an integer: digits grouped by three with ','.
std::string fmt_grouped(uint64_t v);

// Formats an output path template with the frame number as its only
// argument: "{}" / "{0}" / "{:d}", fill-align-width specs such as "{:0>3}",
// "{:>5}", "{:<4}", "{:^6}" and "{{" / "}}" escapes.  Anything else throws
// Error("Failed to format output filename: ...").
std::string format_output_path(const std::string& fmt, uint64_t frame);

// ---- logger -------------------------------------------------------------

class Logger {
    bool console_ = false;
    std::unique_ptr<std::ofstream> file_;

public:
    void add_console() { console_ = true; }
    void add_file(const std::string& path);
    // "[HH:MM:SS] message" to every sink
    void log(const std::string& msg);
};

// ---- arguments ------------------------------------------------------------

struct ParseError : Error {
    explicit ParseError(const std::string& m) : Error(m) {}
};

struct Flag {
    bool* var;
    std::string long_arg;
    int short_arg = -1;
    bool seen = false;
};

struct Parameter {
    std::function<bool(const std::string&)> action;
    std::string value_name;
    std::string long_arg;
    int short_arg = -1;
    bool seen = false;
};

struct Positional {
    std::function<bool(const std::string&)> action;
    std::string name;
};

struct Command {
    std::vector<Flag> flags;
    std::vector<Parameter> parameters;
    std::vector<Positional> positional;
};

// args::parse (src/core/arg_parse.cpp:6-81), same error messages.
void parse_args(const std::vector<std::string>& args, Command& cmd);

// args::parse_float: only digits, '.' and '-', and strtod must consume it all.
bool parse_float(const std::string& s, double& out);
bool parse_int(const std::string& s, int64_t& out);

// ---- headless configuration ------------------------------------------------

struct HeadlessDevice {
    uint32_t vkindex = 0;
    int32_t offset[2] = {0, 0};
    uint32_t extent[2] = {0, 0};
};

// HeadlessConfig from text:
//   device { vkindex = N  offset = (x, y)  extent = (w, h) }  (repeated)
// The file is parsed in the reference's multiline mode, where a '#' is an
// error ("Unexpected character '#'"), not a comment (src/core/Parser.cpp:12-25,
// src/core/Config.h:140).  Errors read "Parse error at line L, col C: ..." or
// "Configuration error: ...".
std::vector<HeadlessDevice> parse_headless_config(std::istream& in);

}  // namespace xdcli
