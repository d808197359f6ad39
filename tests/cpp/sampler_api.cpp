// sampler_api.cpp — the 2D samplers of the C++ host API (include/petrichor/Petrichor.h)
// as host code uses them: MicroJitteredSampler(seed).Initialize(length) and
// RandomSampler2D(seedX, seedY), each asked for N points.
//
//   sampler_api OUT SEED LENGTH N    (LENGTH 0: RandomSampler2D(SEED, SEED + 1))
// OUT: N (x, y) float pairs
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "petrichor/Petrichor.h"

int main(int argc, char** argv)
{
    if (argc != 5) {
        std::fprintf(stderr, "usage: sampler_api OUT SEED LENGTH N\n");
        return 2;
    }
    const unsigned seed = (unsigned)std::strtoul(argv[2], nullptr, 10);
    const int length = std::atoi(argv[3]), n = std::atoi(argv[4]);
    Petrichor::Core::MicroJitteredSampler mj(seed);
    Petrichor::Core::RandomSampler2D rnd(seed, seed + 1);
    Petrichor::Core::ISampler2D* s = &rnd;
    if (length > 0) {
        mj.Initialize(length);
        s = &mj;
    }
    std::vector<float> out;
    for (int i = 0; i < n; ++i) {
        const auto [x, y] = s->Next();
        out.push_back(x);
        out.push_back(y);
    }
    FILE* f = std::fopen(argv[1], "wb");
    if (!f) return 2;
    std::fwrite(out.data(), 4, out.size(), f);
    std::fclose(f);
    return 0;
}
