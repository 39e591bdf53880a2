// Writes draws of std::mt19937 and std::uniform_int_distribution<int> as
// implemented by this toolchain's libstdc++ (the library the reference's
// random_phase relies on, src/engine.cpp:10-30), as JSON on stdout.
#include <cstdio>
#include <random>

int main() {
  std::printf("{\n");
  {
    std::mt19937 g;  // default seed 5489
    std::printf("  \"mt19937_default_first\": [");
    for (int k = 0; k < 8; k++) std::printf("%s%lu", k ? ", " : "", (unsigned long)g());
    std::printf("],\n");
  }
  const int seeds[] = {1, 42, 2024};
  const int ranges[][2] = {{0, 1}, {0, 3}, {0, 5}, {0, 6}, {3, 1000}, {-7, 7}};
  std::printf("  \"uniform_int\": [\n");
  bool first = true;
  for (int s : seeds)
    for (auto& r : ranges) {
      std::mt19937 g(s);
      std::uniform_int_distribution<int> d(r[0], r[1]);
      std::printf("%s    {\"seed\": %d, \"a\": %d, \"b\": %d, \"draws\": [", first ? "" : ",\n", s, r[0], r[1]);
      first = false;
      for (int k = 0; k < 64; k++) std::printf("%s%d", k ? ", " : "", d(g));
      std::printf("], \"next_raw\": %lu}", (unsigned long)g());
    }
  std::printf("\n  ]\n}\n");
  return 0;
}
