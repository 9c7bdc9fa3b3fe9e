// Shared pieces of the jacobi2D / jacobi3D / cg command-line drivers.
//
// Flag parsing keeps the reference's semantics (Stencil/jacobi2D/include/common.h:28-39,
// CG/include/common.h:21-30): a linear scan of argv where the FIRST occurrence
// of a flag wins and its value is parsed with an istringstream.
#pragma once

#include <algorithm>
#include <array>
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cfm.h"

namespace cli {

template <typename T>
T get_argval(char **begin, char **end, const std::string &arg, const T default_val) {
    T argval = default_val;
    char **itr = std::find(begin, end, arg);
    if (itr != end && ++itr != end) {
        std::istringstream inbuf(*itr);
        inbuf >> argval;
    }
    return argval;
}

inline bool get_arg(char **begin, char **end, const std::string &arg) {
    return std::find(begin, end, arg) != end;
}

template <size_t N>
void print_versions(const std::array<const char *, N> &names, size_t selected) {
    std::cout << "Versions (select with -v):" << "\n";
    for (size_t i = 0; i < N; ++i) std::cout << i << ":\t" << names[i] << "\n";
    std::cout << std::endl;
    std::cout << "Running " << names[selected] << "\n" << std::endl;
}

inline int fail(const char *what, int st) {
    const char *msg = cfm_last_error();
    std::fprintf(stderr, "ERROR: %s failed: %s (%d)%s%s\n", what, cfm_status_str(st), st, msg && *msg ? ": " : "",
                 msg ? msg : "");
    return 1;
}

}  // namespace cli
