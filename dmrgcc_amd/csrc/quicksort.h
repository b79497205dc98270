// The reference's QuickSort (src/util/general.ih:6-121), restated: first
// element as pivot, hole-filling partition, not stable.  Its tie order fixes
// the quantum-number table order (real_tensor_space.cc:716-736) and the
// truncation tie breaks (real_tensor_lattice.cc:538), so the same partition
// sequence is kept.  flag 0 = descending, 1 = ascending.  An explicit stack
// replaces the recursion (sub-ranges are independent, so the result is the
// same) to stay safe on long presorted inputs.
#pragma once

#include <utility>
#include <vector>

namespace dmrg {

template <typename T>
void quicksort(std::vector<T>& a, std::vector<int>* idx, int flag) {
  if (a.size() < 2) return;
  std::vector<std::pair<int, int>> stack;
  stack.emplace_back(0, (int)a.size() - 1);
  while (!stack.empty()) {
    const int left = stack.back().first, right = stack.back().second;
    stack.pop_back();
    if (left >= right) continue;
    int i = left, j = right;
    const T pivot = a[i];
    const int pidx = idx ? (*idx)[i] : 0;
    if (flag == 0) {
      while (i < j) {
        while (i < j && a[j] <= pivot) --j;
        if (i < j) { a[i] = a[j]; if (idx) (*idx)[i] = (*idx)[j]; ++i; }
        while (i < j && a[i] >= pivot) ++i;
        if (i < j) { a[j] = a[i]; if (idx) (*idx)[j] = (*idx)[i]; --j; }
      }
    } else {
      while (i < j) {
        while (i < j && a[j] >= pivot) --j;
        if (i < j) { a[i] = a[j]; if (idx) (*idx)[i] = (*idx)[j]; ++i; }
        while (i < j && a[i] <= pivot) ++i;
        if (i < j) { a[j] = a[i]; if (idx) (*idx)[j] = (*idx)[i]; --j; }
      }
    }
    a[i] = pivot;
    if (idx) (*idx)[i] = pidx;
    stack.emplace_back(i + 1, right);
    stack.emplace_back(left, i - 1);
  }
}

}  // namespace dmrg
