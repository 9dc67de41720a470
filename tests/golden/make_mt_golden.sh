#!/bin/sh
# Writes mt19937_uniform_2pi.txt: draws of std::uniform_real_distribution<>(0, 2*M_PI)
# from std::mt19937(seed) as libstdc++ produces them (the RNG the reference's
# decomposition driver consumes: Decomposition_Base.cpp:32-42, 426-436).
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
g++ -O2 -o "$TMP/gen" "$TMP/gen.cpp"
"$TMP/gen" > "$HERE/mt19937_uniform_2pi.txt"
rm -rf "$TMP"
