// ExactCount.cpp -- the device counterpart of the naive_implementation
// programs.  Reads their text edge list ("<vertices> <edges>", then one
// "<src> <dst>" pair per edge: naive_implementation/FiveChain.cpp:33-42) and
// prints "count: <exact count>" as they do, for the program named:
//   TriangleCounting  naive_implementation/TriangleCounting.cpp:44-72
//   FourClique        naive_implementation/FourClique.cpp:44-78
//   ThreeChain        naive_implementation/ThreeChain.cpp:44-51
//   FourChain         naive_implementation/FourChain.cpp:46-66
//   FiveChain         naive_implementation/FiveChain.cpp:45-72
//   FiveStar          naive_implementation/FiveStar.cpp:46-60
//   FiveHouse         naive_implementation/FiveHouse.cpp:25-130
//   FourCycle         4-cycles (no reference program)
// Counts are non-induced, on the simple graph underlying the edge list.
// usage: ExactCount <graph.txt> <program>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "AppCommon.hpp"

int main(int argc, char **argv)
{
    static const struct { const char *name; int what; } kinds[] = {
        {"TriangleCounting", 0}, {"FourClique", 1}, {"ThreeChain", 2}, {"FourChain", 3},
        {"FiveStar", 4},         {"FiveHouse", 5},  {"FiveChain", 6},  {"FourCycle", 7}};
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <graph.txt> <program>\n", argv[0]);
        return 2;
    }
    int what = -1;
    for (const auto &k : kinds)
        if (std::strcmp(argv[2], k.name) == 0) what = k.what;
    if (what < 0) {
        std::fprintf(stderr, "unknown program %s\n", argv[2]);
        return 2;
    }
    std::ifstream fin(argv[1]);
    unsigned long long n = 0, m = 0;
    if (!(fin >> n >> m) || n == 0 || n > 0xffffffffull) {
        std::fprintf(stderr, "cannot read the header of %s\n", argv[1]);
        return 2;
    }
    std::vector<uint32_t> src(m), dst(m);
    for (unsigned long long i = 0; i < m; ++i) {
        unsigned long long a = 0, b = 0;
        if (!(fin >> a >> b) || a >= n || b >= n) {
            std::fprintf(stderr, "bad edge %llu in %s\n", i, argv[1]);
            return 2;
        }
        src[i] = (uint32_t)a;
        dst[i] = (uint32_t)b;
    }
    APP_MAIN_GUARD({
        Graph<Empty> graph;
        graph.create(src, dst, nullptr, (VertexId)n, 1);
        std::printf("count: %llu\n", (unsigned long long)graph.exact_count(what));
    })
    return 0;
}
