// Reverse Cuthill-McKee (see reorder.hpp).
#include "reorder.hpp"

#include <algorithm>
#include <stdexcept>

namespace s355 {

namespace {

// Breadth-first levels from `root` over the unvisited rows of one component;
// returns the last level and the number of levels (marks are reset).
struct Bfs {
  const int* rowptr;
  const int* col;
  std::vector<int> mark, queue;
  int stamp = 0;

  int sweep(int root, std::vector<int>& last) {
    ++stamp;
    queue.clear();
    queue.push_back(root);
    mark[root] = stamp;
    size_t head = 0;
    int levels = 0;
    while (head < queue.size()) {
      const size_t end = queue.size();
      last.assign(queue.begin() + static_cast<long>(head), queue.begin() + static_cast<long>(end));
      for (; head < end; ++head) {
        const int r = queue[head];
        for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) {
          const int c = col[p];
          if (mark[c] != stamp && mark[c] >= 0) {
            mark[c] = stamp;
            queue.push_back(c);
          }
        }
      }
      ++levels;
    }
    return levels;
  }
};

}  // namespace

std::vector<int> rcm_order(int n, const int* rowptr, const int* col) {
  if (n < 0) throw std::invalid_argument("rcm_order: negative size");
  std::vector<int> deg(n);
  for (int r = 0; r < n; ++r) deg[r] = rowptr[r + 1] - rowptr[r];
  // rows by ascending degree, ties by id: the starts of the components
  std::vector<int> by_deg(n);
  for (int r = 0; r < n; ++r) by_deg[r] = r;
  std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  Bfs bfs{rowptr, col, std::vector<int>(n, 0), {}, 0};
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> placed(n, 0);
  std::vector<int> last, nbr;
  size_t next = 0;
  while (order.size() < static_cast<size_t>(n)) {
    while (placed[by_deg[next]]) ++next;
    int root = by_deg[next];
    // pseudo-peripheral root: move to the lowest-degree row of the last
    // level while the level count grows
    int levels = bfs.sweep(root, last);
    for (int k = 0; k < 8; ++k) {
      int cand = last[0];
      for (int r : last)
        if (deg[r] < deg[cand] || (deg[r] == deg[cand] && r < cand)) cand = r;
      std::vector<int> last2;
      const int l2 = bfs.sweep(cand, last2);
      if (l2 <= levels) break;
      root = cand;
      levels = l2;
      last.swap(last2);
    }
    // Cuthill-McKee from the root, neighbours by ascending degree
    const size_t head0 = order.size();
    order.push_back(root);
    placed[root] = 1;
    for (size_t h = head0; h < order.size(); ++h) {
      const int r = order[h];
      nbr.clear();
      for (int p = rowptr[r]; p < rowptr[r + 1]; ++p)
        if (!placed[col[p]]) {
          placed[col[p]] = 1;
          nbr.push_back(col[p]);
        }
      std::sort(nbr.begin(), nbr.end(), [&](int a, int b) { return deg[a] < deg[b] || (deg[a] == deg[b] && a < b); });
      order.insert(order.end(), nbr.begin(), nbr.end());
    }
    // rows of this component are never swept again
    for (size_t h = head0; h < order.size(); ++h) bfs.mark[order[h]] = -1;
  }
  std::reverse(order.begin(), order.end());
  return order;
}

}  // namespace s355
