// Writes colour schedules straight from libstdc++'s generators, as the
// reference's host loops produce them: std::mt19937(seed) % 2 for domino
// (DominoTiler.cpp:51,70) and std::default_random_engine(seed) with
// std::uniform_int_distribution<int>(0,2) for lozenge / rect-triangle
// (LozengeTiler.cpp:86-91), and the bibone move schedule.  Output: one JSON
// object on stdout.
#include <cstdio>
#include <random>

int main() {
  const long seeds[] = {1, 2, 7, 345903, 2147483646L, 4000000001L};
  std::printf("{\"mt19937\": {");
  for (int k = 0; k < 6; k++) {
    std::mt19937 g((uint32_t)seeds[k]);
    std::printf("%s\"%ld\": [", k ? ", " : "", seeds[k]);
    for (int i = 0; i < 64; i++) std::printf("%s%u", i ? ", " : "", (unsigned)g());
    std::printf("]");
  }
  std::printf("}, \"uniform012\": {");
  for (int k = 0; k < 6; k++) {
    std::default_random_engine g;
    std::uniform_int_distribution<int> d(0, 2);
    g.seed(seeds[k]);
    std::printf("%s\"%ld\": [", k ? ", " : "", seeds[k]);
    for (int i = 0; i < 4096; i++) std::printf("%s%d", i ? ", " : "", d(g));
    std::printf("]");
  }
  // Bibone (TriangleDimerTiler.cpp:272-298, 447-448): per step move type,
  // r1, r2 and, for butterfly steps, p1, p2 from one engine; -1 = not drawn.
  std::printf("}, \"bibone\": {");
  for (int k = 0; k < 6; k++) {
    std::default_random_engine g;
    std::uniform_int_distribution<int> d1(0, 2), d2(0, 1);
    g.seed(seeds[k]);
    std::printf("%s\"%ld\": [", k ? ", " : "", seeds[k]);
    for (int i = 0; i < 512; i++) {
      const int type = d1(g), r1 = d1(g), r2 = d2(g);
      int p1 = -1, p2 = -1;
      if (type == 2) {
        p1 = d1(g);
        p2 = d2(g);
      }
      std::printf("%s%d, %d, %d, %d, %d", i ? ", " : "", type, r1, r2, p1, p2);
    }
    std::printf("]");
  }
  std::printf("}}\n");
  return 0;
}
