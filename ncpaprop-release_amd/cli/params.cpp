// Option parsing and validation; see params.h.  Behaviour follows
// src/common/parameterset.cpp:322-625 (parsing) and :1230-2035 (tests and their messages).
#include "params.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace ncpacli {

namespace {

std::string trim(const std::string &s) {
    const char *ws = " \t\n\r";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

std::string set_string(const std::vector<std::string> &v) {
    std::ostringstream oss;
    oss << "{ ";
    for (size_t i = 0; i < v.size(); ++i) oss << (i ? ", " : "") << v[i];
    oss << " }";
    return oss.str();
}

std::string num(double v) { return std::to_string(v); }

enum { T_REQ, T_REQ_IF, T_RADIO, T_GT, T_GE, T_LT, T_LE, T_IGT, T_IGE, T_SET };

}  // namespace

Options::Options(std::string comments, std::string delims) : comments_(std::move(comments)), delims_(std::move(delims)) {}

void Options::add(const std::string &key, Kind kind, bool valid) {
    Option o;
    o.key = key;
    o.kind = kind;
    o.valid = valid;
    opts_.push_back(o);
}

void Options::flag(const std::string &key, bool default_set) {
    add(key, Kind::Flag, true);
    opts_.back().found = default_set;   // FlagParameter(key, true) starts out set (parameterset.cpp:1130-1134)
}
void Options::integer(const std::string &key, int def) {
    add(key, Kind::Int, true);
    opts_.back().i = def;
}
void Options::integer(const std::string &key) { add(key, Kind::Int, false); }
void Options::real(const std::string &key, double def) {
    add(key, Kind::Float, true);
    opts_.back().f = def;
}
void Options::real(const std::string &key) { add(key, Kind::Float, false); }
void Options::string(const std::string &key, const std::string &def) {
    add(key, Kind::String, true);
    opts_.back().s = def;
}
void Options::string(const std::string &key) { add(key, Kind::String, false); }

void Options::required(const std::string &key) { tests_.push_back({T_REQ, key, {}, 0.0}); }
void Options::required_if(const std::string &key, const std::string &other) {
    tests_.push_back({T_REQ_IF, key, {other}, 0.0});
}
void Options::radio(const std::string &name, const std::vector<std::string> &keys) {
    tests_.push_back({T_RADIO, name, keys, 0.0});
}
void Options::real_gt(const std::string &key, double v) { tests_.push_back({T_GT, key, {}, v}); }
void Options::real_ge(const std::string &key, double v) { tests_.push_back({T_GE, key, {}, v}); }
void Options::real_lt(const std::string &key, double v) { tests_.push_back({T_LT, key, {}, v}); }
void Options::real_le(const std::string &key, double v) { tests_.push_back({T_LE, key, {}, v}); }
void Options::int_gt(const std::string &key, int v) { tests_.push_back({T_IGT, key, {}, (double)v}); }
void Options::int_ge(const std::string &key, int v) { tests_.push_back({T_IGE, key, {}, (double)v}); }
void Options::one_of(const std::string &key, const std::vector<std::string> &values) {
    tests_.push_back({T_SET, key, values, 0.0});
}

Option *Options::find(const std::string &key) {
    for (auto &o : opts_)
        if (o.key == key) return &o;
    return nullptr;
}
const Option *Options::find(const std::string &key) const {
    for (auto &o : opts_)
        if (o.key == key) return &o;
    return nullptr;
}
const Option &Options::at(const std::string &key) const {
    const Option *o = find(key);
    if (!o) throw std::invalid_argument("No parameter '" + key + "' has been specified!");
    if (!o->valid) throw std::invalid_argument("Parameter " + key + " not set");
    return *o;
}

void Options::set_value(const std::string &key, const std::string &value) {
    Option *o = find(key);
    if (!o) throw std::invalid_argument("Unexpected argument: '" + key + "'");
    if (o->kind == Kind::Flag)
        throw std::invalid_argument("Option " + key + " appears to have an argument '" + value +
                                    "', but is not expecting one.");
    char *end = nullptr;
    switch (o->kind) {
    case Kind::Int:
        o->i = (int)std::strtol(value.c_str(), &end, 10);
        if (end == value.c_str()) throw std::invalid_argument("Option " + key + ": '" + value + "' is not an integer");
        break;
    case Kind::Float:
        o->f = std::strtod(value.c_str(), &end);
        if (end == value.c_str()) throw std::invalid_argument("Option " + key + ": '" + value + "' is not a number");
        break;
    default:
        o->s = value;
    }
    o->found = true;
    o->valid = true;
}

void Options::set_flag(const std::string &key) {
    Option *o = find(key);
    if (!o) throw std::invalid_argument("Unexpected argument: '" + key + "'");
    if (o->kind != Kind::Flag)
        throw std::invalid_argument("Option " + key + " expects an argument, but none appears to be provided.");
    o->found = true;
}

bool Options::is_long(const std::string &a) const { return a.size() >= 3 && a.compare(0, 2, "--") == 0; }
bool Options::is_short(const std::string &a) const {
    return a.size() >= 2 && a[0] == '-' && a[1] != '-' && !std::isdigit((unsigned char)a[1]);
}

void Options::parse_command_line(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (is_long(arg)) {
            const std::string body = arg.substr(arg.find_first_not_of('-'));
            const size_t d = body.find_first_of(delims_);
            if (d != std::string::npos) {
                set_value(body.substr(0, d), body.substr(d + 1));
                continue;
            }
            const Option *o = find(body);
            if (!o) throw std::invalid_argument("Unknown option '" + arg + "'.");
            if (o->kind == Kind::Flag) {
                set_flag(body);
            } else {
                if (i + 1 >= argc)
                    throw std::invalid_argument("Option '" + body + "' expects a value, but none was provided.");
                const std::string v = argv[i + 1];
                if (is_long(v) || is_short(v))
                    throw std::invalid_argument("Value '" + v + "' for option '" + body +
                                                "' appears to be another argument.");
                set_value(body, v);
                ++i;
            }
        } else if (is_short(arg)) {
            for (char c : arg.substr(1)) {
                const std::string k(1, c);
                if (!find(k)) throw std::invalid_argument("Unexpected argument '-" + k + "'.");
                set_flag(k);
            }
        }
        // anything else is left unparsed, as the reference does
    }
}

int Options::parse_file(const std::string &path) {
    std::ifstream in(path);
    if (!in.good()) return 0;
    int n = 0;
    std::string line;
    while (std::getline(in, line)) {
        line = trim(line);
        // a line is skipped when it is blank or starts with a comment character
        if (line.empty() || line.find_first_of(comments_) == 0) continue;
        std::vector<std::string> tok;
        size_t p = 0;
        while (p < line.size()) {
            const size_t b = line.find_first_not_of(delims_, p);
            if (b == std::string::npos) break;
            const size_t e = line.find_first_of(delims_, b);
            tok.push_back(line.substr(b, e == std::string::npos ? std::string::npos : e - b));
            if (e == std::string::npos) break;
            p = e;
        }
        if (tok.size() > 2) throw std::invalid_argument("Ambiguous line in file " + path + ": '" + line + "'.");
        if (tok.size() == 1) {
            set_flag(trim(tok[0]));
        } else if (tok.size() == 2) {
            set_value(trim(tok[0]), trim(tok[1]));
        }
        ++n;
    }
    return n;
}

bool Options::validate() {
    failed_.clear();
    for (const Test &t : tests_) {
        const Option *o = find(t.key);
        bool ok = true;
        std::string msg;
        switch (t.type) {
        case T_REQ:
            ok = o && o->found;
            msg = t.key + " is not present.";
            break;
        case T_REQ_IF: {
            const Option *other = find(t.names[0]);
            if (other && other->found) ok = o && o->found;
            msg = "One of " + set_string(t.names) + " is set, but " + t.key + " is not set.";
            break;
        }
        case T_RADIO: {
            int matched = 0;
            for (auto &k : t.names) {
                const Option *b = find(k);
                if (b && b->found) ++matched;
            }
            ok = matched == 1;
            msg = t.key + ": " + std::to_string(matched) + " of " + set_string(t.names) +
                  " are present; must be one and only one.";
            break;
        }
        case T_GT:
        case T_GE:
        case T_LT:
        case T_LE: {
            if (!o || !o->valid) break;
            const double x = o->kind == Kind::Int ? (double)o->i : o->f;
            const char *rel = "";
            if (t.type == T_GT) { ok = x > t.v; rel = "must be greater than"; }
            if (t.type == T_GE) { ok = x >= t.v; rel = "must be greater than or equal to"; }
            if (t.type == T_LT) { ok = x < t.v; rel = "must be less than"; }
            if (t.type == T_LE) { ok = x <= t.v; rel = "must be less than or equal to"; }
            msg = t.key + " (" + num(x) + ") " + rel + " " + num(t.v) + ".";
            break;
        }
        case T_IGT:
        case T_IGE: {
            if (!o || !o->valid) break;
            const int x = o->kind == Kind::Int ? o->i : (int)o->f;
            ok = t.type == T_IGT ? x > (int)t.v : x >= (int)t.v;
            msg = t.key + " (" + std::to_string(x) + ") must be greater than " +
                  std::string(t.type == T_IGE ? "or equal to " : "") + std::to_string((int)t.v) + ".";
            break;
        }
        case T_SET: {
            if (!o || !o->valid) break;
            ok = false;
            for (auto &v : t.names) ok = ok || (o->s == v);
            msg = t.key + ": \"" + o->s + "\" is not in " + set_string(t.names) + ".";
            break;
        }
        }
        if (!ok) failed_.push_back(msg);
    }
    return failed_.empty();
}

void Options::print_failed(std::ostream &os) const {
    for (auto &m : failed_) os << m << std::endl;
}

bool Options::has(const std::string &key) const { return find(key) != nullptr; }

bool Options::was_found(const std::string &key) const {
    const Option *o = find(key);
    if (!o) throw std::invalid_argument("No parameter '" + key + "' has been specified!");
    return o->found;
}
bool Options::get_bool(const std::string &key) const {
    const Option &o = at(key);
    switch (o.kind) {
    case Kind::Flag: return o.found;
    case Kind::Int: return o.i != 0;
    case Kind::Float: return o.f != 0.0;
    default: return !o.s.empty();
    }
}
int Options::get_int(const std::string &key) const {
    const Option &o = at(key);
    return o.kind == Kind::Int ? o.i : (o.kind == Kind::Float ? (int)o.f : (o.found ? 1 : 0));
}
double Options::get_real(const std::string &key) const {
    const Option &o = at(key);
    return o.kind == Kind::Float ? o.f : (o.kind == Kind::Int ? (double)o.i : (o.found ? 1.0 : 0.0));
}
std::string Options::get_string(const std::string &key) const {
    const Option &o = at(key);
    if (o.kind == Kind::Flag) return o.found ? "true" : "false";
    if (o.kind == Kind::String) return o.s;
    std::ostringstream oss;
    if (o.kind == Kind::Int) oss << o.i; else oss << o.f;
    return oss.str();
}

}  // namespace ncpacli
