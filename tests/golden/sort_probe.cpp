// Prints the order std::sort leaves N elements in when every key compares
// equal (the reference sorts elements by partition id, RootStructure.cpp:797-798,
// and on one partition all ids are equal).  Used by the tests as an
// independent libstdc++ probe of the element-order permutation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

struct E {
  int pid;
  int id;
};

int main(int argc, char** argv) {
  int n = argc > 1 ? std::atoi(argv[1]) : 40;
  std::vector<E> es(n);
  std::vector<E*> p(n);
  for (int i = 0; i < n; ++i) {
    es[i] = {0, i};
    p[i] = &es[i];
  }
  std::sort(p.begin(), p.end(), [](E* a, E* b) { return a->pid < b->pid; });
  for (int i = 0; i < n; ++i) std::printf("%d%c", p[i]->id, i + 1 == n ? '\n' : ' ');
  return 0;
}
