#include "config.h"

#include <cstdlib>
#include <iomanip>
#include <sstream>

void Config::parse(std::istream &in) {
    for (std::string line; std::getline(in, line);) {
        std::istringstream iss(line);
        std::string id, eq, val;
        bool error = false;
        if (!(iss >> id)) {
            error = !line.empty() && line.find_first_not_of(" \t\r") != std::string::npos;
            if (!error) continue;
        } else if (id[0] == '#') {
            continue;
        } else if (!(iss >> eq >> val) || eq != "=") {
            error = true;
        } else {
            // nothing may follow the value (Config.cpp:26); tested by trying to
            // read one more token, which also works when `val` ended the line
            std::string extra;
            if (iss >> extra) error = true;
        }
        if (error) {
            std::cout << "Incomplete configuration! Did you forget the space?" << std::endl;
            continue;
        }
        options[id] = val;
    }
}

void Config::print() const {
    std::cout << std::setw(15) << "=============== OPTIONS ===============" << std::endl;
    for (const auto &kv : options) std::cout << std::setw(15) << kv.first << " = " << kv.second << std::endl;
    std::cout << std::setw(15) << "=======================================" << std::endl;
}

double Config::getd(const std::string &k, double d) const {
    auto it = options.find(k);
    return it == options.end() ? d : std::atof(it->second.c_str());
}
int Config::geti(const std::string &k, int d) const {
    auto it = options.find(k);
    return it == options.end() ? d : std::atoi(it->second.c_str());
}
std::string Config::gets(const std::string &k, const std::string &d) const {
    auto it = options.find(k);
    return it == options.end() ? d : it->second;
}
