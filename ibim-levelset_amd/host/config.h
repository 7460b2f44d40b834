// config.h -- `key = value` options file (Config.h:15-22, Config.cpp:15-49).
#pragma once

#include <iostream>
#include <map>
#include <string>

class Config {
public:
    std::map<std::string, std::string> options;
    // one `key = value` per line, `#` starts a comment line; a malformed line
    // is reported and skipped (Config.cpp:15-41)
    void parse(std::istream &in);
    void print() const;
    // typed access with a default for a missing key (the reference's atoi/atof
    // on a missing key silently gives 0)
    double getd(const std::string &k, double dflt) const;
    int geti(const std::string &k, int dflt) const;
    std::string gets(const std::string &k, const std::string &dflt) const;
    bool has(const std::string &k) const { return options.count(k) != 0; }
};
