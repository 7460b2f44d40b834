// Prints the FP16 M2L pair-result offsets of ibim_fp24.h for the pair rows,
// row strides and value indices given on the command line (triples), one per
// line, so that a test can compare them with exact integer arithmetic.
#include <cstdio>
#include <cstdlib>

#include "../../ibim-levelset_amd/csrc/ibim_fp24.h"

int main(int argc, char **argv) {
    for (int i = 1; i + 2 < argc; i += 3) {
        const long long pair = std::atoll(argv[i]);
        const int ys = std::atoi(argv[i + 1]), v = std::atoi(argv[i + 2]);
        std::printf("%zu %zu\n", ibim::fp24_offset(pair, ys, v), ibim::fp24_buffer_dwords(pair + 1, ys));
    }
    return 0;
}
