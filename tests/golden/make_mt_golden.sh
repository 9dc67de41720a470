#!/bin/sh
# Writes the libstdc++ golden draws of std::mt19937 that the reference's optimisers
# consume (compiled here with g++; no reference source involved):
#  mt19937_uniform_2pi.txt: std::uniform_real_distribution<>(0, 2*M_PI), seed 1
#    (Decomposition_Base.cpp:32-42, 426-436: the RANDOM initial guess);
#  mt19937_agents_draws.txt: interleaved std::uniform_int_distribution<>(0, 83),
#    std::uniform_real_distribution<>(0.0, 1.0) and (-2*M_PI, 2*M_PI), seed 7
#    (AGENTS.cpp:280-296, 696-703; Optimization_Interface.cpp:576-602);
#  mt19937_decomposition_draws.txt: the decomposition driver's generator, seeds 1 and
#    42: the object's id, std::uniform_int_distribution<>(0, INT_MAX), drawn by the
#    constructor (Optimization_Interface.cpp:179-180), then 48 draws of
#    std::uniform_real_distribution<>(0, 2*M_PI) (the RANDOM initial guess,
#    Decomposition_Base.cpp:426-429, then the restart jitter, BFGS.cpp:130-133).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
TMP=$(mktemp -d)
cat > "$TMP/gen.cpp" <<'CPP'
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    const unsigned seed = 1;
    const int n = 64;
    std::mt19937 gen(seed);
    std::uniform_real_distribution<> dist(0.0, 2 * M_PI);
    std::printf("%u %d\n", seed, n);
    for (int i = 0; i < n; ++i) std::printf("%.17g\n", dist(gen));
    return 0;
}
CPP
cat > "$TMP/gen2.cpp" <<'CPP'
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    const unsigned seed = 7;
    const int n = 200;
    std::mt19937 gen(seed);
    std::uniform_int_distribution<> di(0, 83);
    std::uniform_real_distribution<> dp(0.0, 1.0);
    std::uniform_real_distribution<> dr(-2 * M_PI, 2 * M_PI);
    std::printf("%u %d\n", seed, n);
    for (int i = 0; i < n; ++i) {
        const int a = di(gen);
        const double b = dp(gen);
        const double c = dr(gen);
        std::printf("%d %.17g %.17g\n", a, b, c);
    }
    return 0;
}
CPP
cat > "$TMP/gen3.cpp" <<'CPP'
#include <climits>
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    const unsigned seeds[2] = {1, 42};
    const int n = 48;
    for (unsigned seed : seeds) {
        std::mt19937 gen(seed);
        std::uniform_int_distribution<> di(0, INT_MAX);
        std::uniform_real_distribution<> dr(0.0, 2 * M_PI);
        const int id = di(gen);
        std::printf("%u %d %d\n", seed, id, n);
        for (int i = 0; i < n; ++i) std::printf("%.17g\n", dr(gen));
    }
    return 0;
}
CPP
g++ -O2 -o "$TMP/gen" "$TMP/gen.cpp"
g++ -O2 -o "$TMP/gen3" "$TMP/gen3.cpp"
g++ -O2 -o "$TMP/gen2" "$TMP/gen2.cpp"
"$TMP/gen" > "$HERE/mt19937_uniform_2pi.txt"
"$TMP/gen2" > "$HERE/mt19937_agents_draws.txt"
"$TMP/gen3" > "$HERE/mt19937_decomposition_draws.txt"
rm -rf "$TMP"
