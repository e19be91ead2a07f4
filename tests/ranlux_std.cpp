// Independent RANLUX check: the C++ standard library's subtract-with-carry
// engine (lags 24/10, base 2^24) under a discard-block adaptor keeping 24 of
// every 223 values is the generator of F. James's RANLUX at luxury level 3.
// Reads the 24 seeds oldest-first and the carry, prints N values.
//   usage: ranlux_std N < state
#include <cstdio>
#include <iostream>
#include <random>
#include <sstream>

int main(int argc, char** argv)
{
    const int n = argc > 1 ? std::atoi(argv[1]) : 100;
    if (argc > 2 && std::string(argv[2]) == "kat") {
        // [rand.predef]: the 10000th value of a default-constructed engine
        std::ranlux24_base b;
        b.discard(9999);
        std::ranlux24 d;
        d.discard(9999);
        std::printf("%lu %lu\n", (unsigned long)b(), (unsigned long)d());
        return 0;
    }
    std::string state;
    std::getline(std::cin, state);
    std::istringstream base_state(state);
    std::ranlux24_base base;
    base_state >> base;
    std::discard_block_engine<std::ranlux24_base, 223, 24> eng(base);
    for (int i = 0; i < n; ++i) std::printf("%lu\n", (unsigned long)eng());
    return 0;
}
