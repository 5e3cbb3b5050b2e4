#include "cli_util.hpp"

#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cstdlib>
#include <ctime>
#include <iostream>

namespace xdcli {

std::string sformat(const char* f, ...) {
    va_list ap;
    va_start(ap, f);
    va_list ap2;
    va_copy(ap2, ap);
    const int n = vsnprintf(nullptr, 0, f, ap);
    va_end(ap);
    std::string s(n > 0 ? (size_t)n : 0, '\0');
    if (n > 0) vsnprintf(&s[0], (size_t)n + 1, f, ap2);
    va_end(ap2);
    return s;
}

std::string fmt_double(double v) { return sformat("%g", v); }

std::string fmt_grouped(uint64_t v) {
    std::string d = std::to_string(v), out;
    for (size_t i = 0; i < d.size(); ++i) {
        if (i && (d.size() - i) % 3 == 0) out += ',';
        out += d[i];
    }
    return out;
}

std::string format_output_path(const std::string& f, uint64_t frame) {
    auto bad = [](const std::string& why) { return Error("Failed to format output filename: " + why); };
    std::string out;
    for (size_t i = 0; i < f.size(); ++i) {
        const char c = f[i];
        if (c == '}') {
            if (i + 1 < f.size() && f[i + 1] == '}') {
                out += '}';
                ++i;
                continue;
            }
            throw bad("unmatched '}' in format string");
        }
        if (c != '{') {
            out += c;
            continue;
        }
        if (i + 1 < f.size() && f[i + 1] == '{') {
            out += '{';
            ++i;
            continue;
        }
        const size_t close = f.find('}', i);
        if (close == std::string::npos) throw bad("missing '}' in format string");
        const std::string field = f.substr(i + 1, close - i - 1);
        i = close;
        const size_t colon = field.find(':');
        const std::string id = field.substr(0, colon);
        if (!id.empty() && id != "0") throw bad("argument index out of range");
        std::string spec = colon == std::string::npos ? "" : field.substr(colon + 1);
        // [[fill]align][0][width][d]
        char fill = ' ', align = '>';
        size_t k = 0;
        auto is_align = [](char a) { return a == '<' || a == '>' || a == '^'; };
        if (spec.size() >= 2 && is_align(spec[1])) {
            fill = spec[0];
            align = spec[1];
            k = 2;
        } else if (!spec.empty() && is_align(spec[0])) {
            align = spec[0];
            k = 1;
        } else if (!spec.empty() && spec[0] == '0') {
            fill = '0';
            k = 1;
        }
        size_t width = 0;
        while (k < spec.size() && isdigit((unsigned char)spec[k])) width = width * 10 + (size_t)(spec[k++] - '0');
        if (k < spec.size() && spec[k] == 'd') ++k;
        if (k != spec.size()) throw bad("invalid format specifier");
        std::string num = std::to_string(frame);
        if (num.size() < width) {
            const size_t pad = width - num.size();
            if (align == '<')
                num += std::string(pad, fill);
            else if (align == '^')
                num = std::string(pad / 2, fill) + num + std::string(pad - pad / 2, fill);
            else
                num = std::string(pad, fill) + num;
        }
        out += num;
    }
    return out;
}

void Logger::add_file(const std::string& path) {
    file_ = std::make_unique<std::ofstream>(path);
}

void Logger::log(const std::string& msg) {
    std::time_t t = std::time(nullptr);
    char ts[32];
    std::strftime(ts, sizeof ts, "[%H:%M:%S] ", std::localtime(&t));
    const std::string line = ts + msg;
    if (console_) std::cout << line << '\n' << std::flush;
    if (file_ && *file_) *file_ << line << '\n' << std::flush;
}

void parse_args(const std::vector<std::string>& args, Command& cmd) {
    auto find_flag = [&](const std::string& a) -> Flag* {
        for (auto& f : cmd.flags)
            if (a == f.long_arg || (a.size() == 2 && a[1] == f.short_arg)) return &f;
        return nullptr;
    };
    auto find_param = [&](const std::string& a) -> Parameter* {
        for (auto& p : cmd.parameters)
            if (a == p.long_arg || (a.size() == 2 && a[1] == p.short_arg)) return &p;
        return nullptr;
    };
    auto duplicate = [](const char* type, const std::string& l, int s) {
        if (s != -1) return ParseError(sformat("Duplicate specification of %s %s/-%c", type, l.c_str(), (char)s));
        return ParseError(sformat("Duplicate specification of %s %s", type, l.c_str()));
    };
    size_t pos_seen = 0;
    for (size_t i = 0; i < args.size(); ++i) {
        const std::string& a = args[i];
        if (!a.empty() && a[0] == '-') {
            if (Flag* f = find_flag(a)) {
                if (f->seen) throw duplicate("flag", f->long_arg, f->short_arg);
                f->seen = true;
                *f->var = true;
            } else if (Parameter* p = find_param(a)) {
                ++i;
                if (p->seen) throw duplicate("parameter", p->long_arg, p->short_arg);
                if (i == args.size())
                    throw ParseError(sformat("Parameter %s expects argument <%s>", a.c_str(), p->value_name.c_str()));
                p->seen = true;
                if (!p->action(args[i]))
                    throw ParseError(
                        sformat("Invalid value for <%s> of parameter %s", p->value_name.c_str(), a.c_str()));
            } else {
                throw ParseError(sformat("Unrecognized option %s", a.c_str()));
            }
        } else if (pos_seen == cmd.positional.size()) {
            throw ParseError(sformat("Unexpected positional argument '%s'", a.c_str()));
        } else {
            auto& p = cmd.positional[pos_seen];
            if (!p.action(a))
                throw ParseError(sformat("Invalid value for positional argument <%s>", p.name.c_str()));
            ++pos_seen;
        }
    }
    if (pos_seen != cmd.positional.size())
        throw ParseError(
            sformat("Missing required positional argument <%s>", cmd.positional[pos_seen].name.c_str()));
}

bool parse_float(const std::string& s, double& out) {
    for (char c : s)
        if (!isdigit((unsigned char)c) && c != '.' && c != '-') return false;
    if (s.empty()) return false;
    char* end = nullptr;
    const double v = strtod(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    out = v;
    return true;
}

bool parse_int(const std::string& s, int64_t& out) {
    if (s.empty()) return false;
    size_t i = s[0] == '-' ? 1 : 0;
    if (i == s.size()) return false;
    for (size_t k = i; k < s.size(); ++k)
        if (!isdigit((unsigned char)s[k])) return false;
    errno = 0;
    char* end = nullptr;
    const long long v = strtoll(s.c_str(), &end, 10);
    if (errno == ERANGE || end != s.c_str() + s.size()) return false;
    out = v;
    return true;
}

// ---- headless configuration ------------------------------------------------

namespace {

class ConfigParser {
    std::istream& in_;
    size_t line_ = 1, col_ = 1;

public:
    explicit ConfigParser(std::istream& in) : in_(in) {}

    Error error(const std::string& what) const {
        return Error(sformat("Parse error at line %zu, col %zu: %s", line_, col_, what.c_str()));
    }
    // Multiline mode of src/core/Parser.cpp:12-25: a '#' seen by peek() is an
    // error there (the single-line comment skip is for non-multiline input,
    // which headless configs never are); consume() reads it as a character.
    int peek() {
        const int c = in_.peek();
        if (c == '#') throw error("Unexpected character '#'");
        return c;
    }
    int consume() {
        if (in_.peek() == EOF) throw error("Unexpected end of input");
        const int c = in_.get();
        if (c == '\n') {
            ++line_;
            col_ = 0;
        }
        ++col_;
        return c;
    }
    void expect(int e) {
        const int c = consume();
        if (c != e) throw error(sformat("Expected character '%c', found '%c'", (char)e, (char)c));
    }
    void optws() {
        while (isspace(peek())) consume();
    }
    std::string key() {
        const int c = consume();
        if (!isalpha(c)) throw error(sformat("Expected alphabetic key, found '%c'", (char)c));
        std::string k(1, (char)c);
        while (isalpha(peek())) k += (char)consume();
        return k;
    }
    uint64_t number() {
        int c = consume();
        if (c < '0' || c > '9') throw error(sformat("Expected numeric character, found '%c'", (char)c));
        uint64_t v = (uint64_t)(c - '0');
        for (c = peek(); c >= '0' && c <= '9'; c = peek()) {
            v = v * 10 + (uint64_t)(c - '0');
            consume();
        }
        return v;
    }
    void pair(uint64_t& a, uint64_t& b) {
        expect('(');
        optws();
        a = number();
        optws();
        expect(',');
        optws();
        b = number();
        optws();
        expect(')');
    }
};

Error config_error(const std::string& what) { return Error("Configuration error: " + what); }

// The body of a braced or top-level block: `key <item>` entries separated
// by whitespace; `item` parses the value of a known key and returns false
// for an unknown one.
void parse_block(ConfigParser& p, const std::function<bool(const std::string&)>& item) {
    p.optws();
    int c = p.peek();
    while (isalpha(c)) {
        const std::string k = p.key();
        p.optws();
        if (!item(k)) throw p.error("Unexpected key '" + k + "'");
        c = p.peek();
        if (!isspace(c)) break;
        p.optws();
        c = p.peek();
    }
    p.optws();
}

}  // namespace

std::vector<HeadlessDevice> parse_headless_config(std::istream& in) {
    ConfigParser p(in);
    std::vector<HeadlessDevice> devices;
    parse_block(p, [&](const std::string& k) {
        if (k != "device") return false;
        p.expect('{');
        bool has_index = false, has_offset = false, has_extent = false;
        HeadlessDevice d;
        parse_block(p, [&](const std::string& kk) {
            if (kk != "vkindex" && kk != "offset" && kk != "extent") return false;
            bool& seen = kk == "vkindex" ? has_index : kk == "offset" ? has_offset : has_extent;
            p.expect('=');
            p.optws();
            if (kk == "vkindex") {
                const uint64_t v = p.number();
                if (seen) throw config_error("Ambiguous key 'vkindex'");
                d.vkindex = (uint32_t)v;
            } else {
                uint64_t a, b;
                p.pair(a, b);
                if (seen) throw config_error("Ambiguous key '" + kk + "'");
                if (kk == "offset") {
                    d.offset[0] = (int32_t)a;
                    d.offset[1] = (int32_t)b;
                } else {
                    d.extent[0] = (uint32_t)a;
                    d.extent[1] = (uint32_t)b;
                }
            }
            seen = true;
            return true;
        });
        if (!has_index) throw config_error("Missing key 'vkindex'");
        if (!has_offset) throw config_error("Missing key 'offset'");
        if (!has_extent) throw config_error("Missing key 'extent'");
        p.expect('}');
        devices.push_back(d);
        return true;
    });
    if (devices.empty()) throw config_error("At least one device entry is required");
    const int c = p.peek();
    if (c > 0) throw p.error(sformat("Expected end of input, found '%c'", (char)c));
    return devices;
}

}  // namespace xdcli
