// Loads a TOML case with dacti::load_case_file, grows the initial tree with
// dacti::build_tree and writes both as raw bytes:
//   dacti_case | int64 n | parent | child_base | level | flags | bc | center | size | p
// Errors print "ConfigError: <message>" and exit 3.
// usage: case_tree_dump <case.toml> <euler|navier_stokes|multi> <dim> <limiter> <out> [grid.bin n0 n1 n2]
// (grid.bin: the float64 primitive grid of a "grid" initial condition)
#include <cstdio>
#include <cstring>
#include <string>

#include "dacti/config/case.hpp"

template <typename T>
static void put(FILE *f, const std::vector<T> &v) {
    std::fwrite(v.data(), sizeof(T), v.size(), f);
}

int main(int argc, char **argv) {
    if (argc != 6 && argc != 10) return 2;
    const std::string m = argv[2];
    const int model = m == "euler" ? DACTI_MODEL_EULER : m == "navier_stokes" ? DACTI_MODEL_NAVIER_STOKES : DACTI_MODEL_MULTI;
    try {
        dacti::Case cs = dacti::load_case_file(argv[1], model, std::atoi(argv[3]), dacti::limiter_id(argv[4]));
        if (argc == 10) {
            size_t total = (size_t) cs.num_eqns();
            for (int d = 0; d < 3; ++d) {
                cs.ic_grid_shape[d] = std::atoi(argv[7 + d]);
                if (d < cs.dim) total *= (size_t) cs.ic_grid_shape[d];
            }
            cs.ic_grid.resize(total);
            FILE *g = std::fopen(argv[6], "rb");
            if (!g || std::fread(cs.ic_grid.data(), 8, total, g) != total) return 4;
            std::fclose(g);
        }
        dacti_case c;
        std::memset(&c, 0, sizeof c);
        c = dacti::to_dacti_case(cs);
        const dacti::Tree t = dacti::build_tree(cs);
        FILE *f = std::fopen(argv[5], "wb");
        std::fwrite(&c, sizeof c, 1, f);
        std::fwrite(&t.n, 8, 1, f);
        put(f, t.parent), put(f, t.child_base), put(f, t.level), put(f, t.flags), put(f, t.bc);
        put(f, t.center), put(f, t.size), put(f, t.p);
        std::fclose(f);
    } catch (const std::exception &e) {
        std::printf("ConfigError: %s\n", e.what());
        return 3;
    }
    return 0;
}
