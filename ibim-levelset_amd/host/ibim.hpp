// ibim.hpp -- C++ ownership helpers over the C ABI (include/ibim.h).
#pragma once

#include <stdexcept>
#include <string>

#include "../../include/ibim.h"

namespace ibim {

struct Error : std::runtime_error {
    explicit Error(const std::string &m) : std::runtime_error(m) {}
};

inline int check(int rc, const char *what) {
    if (rc < 0) throw Error(std::string(what) + ": " + ibim_last_error());
    return rc;
}

}  // namespace ibim
