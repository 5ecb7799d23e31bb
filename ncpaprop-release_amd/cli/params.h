// Command-line / parameter-file options with the semantics of ncpaprop's NCPA::ParameterSet
// (src/common/parameterset.cpp): --key value, --key=value (any of the delimiters), -abc short
// flags, strict mode (unknown options are an error), a parameter file of "key<delim>value"
// lines read between two command-line passes, and validation tests whose failure messages
// are printed before the program exits.
#pragma once
#include <ostream>
#include <string>
#include <vector>

namespace ncpacli {

enum class Kind { Flag, Int, Float, String };

struct Option {
    std::string key;
    Kind kind;
    bool found = false;      // given on the command line / in the file (flags: set)
    bool valid = false;      // holds a value (a default or a parsed one)
    double f = 0.0;
    int i = 0;
    std::string s;
};

class Options {
public:
    explicit Options(std::string comments = "#%", std::string delims = "=: ");

    void flag(const std::string &key, bool default_set = false);
    void integer(const std::string &key, int def);
    void integer(const std::string &key);
    void real(const std::string &key, double def);
    void real(const std::string &key);
    void string(const std::string &key, const std::string &def);
    void string(const std::string &key);

    // validation tests (parameterset.cpp:1230-2035)
    void required(const std::string &key);
    void required_if(const std::string &key, const std::string &other);
    void radio(const std::string &name, const std::vector<std::string> &keys);
    void real_gt(const std::string &key, double v);
    void real_ge(const std::string &key, double v);
    void real_lt(const std::string &key, double v);
    void real_le(const std::string &key, double v);
    void int_gt(const std::string &key, int v);
    void int_ge(const std::string &key, int v);
    void one_of(const std::string &key, const std::vector<std::string> &values);

    // parsing; throw std::invalid_argument on malformed or unexpected options
    void parse_command_line(int argc, char **argv);
    int parse_file(const std::string &path);   // returns lines parsed (0 if the file is absent)
    bool validate();
    void print_failed(std::ostream &os) const;

    bool has(const std::string &key) const;   // the option is configured
    bool was_found(const std::string &key) const;
    bool get_bool(const std::string &key) const;
    int get_int(const std::string &key) const;
    double get_real(const std::string &key) const;
    std::string get_string(const std::string &key) const;

private:
    struct Test {
        int type;
        std::string key;
        std::vector<std::string> names;
        double v = 0.0;
    };
    Option *find(const std::string &key);
    const Option *find(const std::string &key) const;
    const Option &at(const std::string &key) const;
    void add(const std::string &key, Kind kind, bool valid);
    void set_value(const std::string &key, const std::string &value);
    void set_flag(const std::string &key);
    bool is_long(const std::string &a) const;
    bool is_short(const std::string &a) const;

    std::string comments_, delims_;
    std::vector<Option> opts_;
    std::vector<Test> tests_;
    std::vector<std::string> failed_;
};

}  // namespace ncpacli
