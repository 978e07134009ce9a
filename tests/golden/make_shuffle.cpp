// Writes the VPL permutation of VPLProcessor.cpp:58-62 for several VPL
// counts: std::shuffle(iota(n), std::mt19937(1979982)), as raw int32 files
// shuffle_<n>.i32. Before writing, checks the engine against the known answer
// the C++ standard fixes ([rand.predef]: the 10000th invocation of a
// default-constructed std::mt19937 produces 4123659995).
//
//   g++ -O2 -std=c++17 -o /tmp/make_shuffle tests/golden/make_shuffle.cpp
//   /tmp/make_shuffle tests/golden
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <string>
#include <vector>

int main(int argc, char** argv) {
  std::mt19937 kat;
  kat.discard(9999);
  if (kat() != 4123659995u) {
    std::fprintf(stderr, "mt19937 known-answer test failed\n");
    return 1;
  }
  const std::string dir = argc > 1 ? argv[1] : ".";
  for (int n : {256, 1024, 2048, 4096, 16384}) {
    std::vector<int> v(n);
    std::iota(v.begin(), v.end(), 0);
    std::mt19937 g(1979982);
    std::shuffle(v.begin(), v.end(), g);
    const std::string path = dir + "/shuffle_" + std::to_string(n) + ".i32";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return 2;
    std::fwrite(v.data(), sizeof(int), v.size(), f);
    std::fclose(f);
  }
  return 0;
}
