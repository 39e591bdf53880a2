// SpMatElems / SpMatBsr: the host block-sparse format of the engine API.
// Behaviour follows src/spmat.cpp of the reference: build() sorts COO blocks
// by (row, column) and sums duplicate blocks; blocks are column-major;
// symmetrize() replaces each block pair (i,j), (j,i) by its Hermitian mean.
#include <algorithm>
#include <cassert>
#include <cstring>
#include <iostream>
#include <numeric>
#include <stdexcept>
#include <string>

#include "fastkpm.h"

namespace fkpm {

template <typename T>
SpMatElems<T>::SpMatElems(int n_rows, int n_cols, int b_len)
    : n_rows(n_rows), n_cols(n_cols), b_len(b_len) {}

template <typename T>
int SpMatElems<T>::n_blocks() const {
  return (int)row_idx.size();
}

template <typename T>
void SpMatElems<T>::clear() {
  row_idx.clear();
  col_idx.clear();
  val.clear();
}

// src/spmat.cpp:28-33
template <typename T>
void SpMatElems<T>::add(int i, int j, T const* v) {
  assert(i >= 0 && i < n_rows && j >= 0 && j < n_cols);
  row_idx.push_back(i);
  col_idx.push_back(j);
  val.insert(val.end(), v, v + (size_t)b_len * b_len);
}

template <typename T>
SpMatBsr<T>::SpMatBsr() {}

template <typename T>
SpMatBsr<T>::SpMatBsr(SpMatElems<T> const& elems) {
  build(elems);
}

// src/spmat.cpp:46-129.  A counting sort by row followed by a stable sort of
// each row by column gives the (row, column) order; equal (i, j) blocks are
// then adjacent and summed in input order.
template <typename T>
void SpMatBsr<T>::build(SpMatElems<T> const& elems) {
  n_rows = elems.n_rows;
  n_cols = elems.n_cols;
  b_len = elems.b_len;
  const size_t bb = (size_t)b_len * b_len;
  const int nin = elems.n_blocks();
  Vec<int> start(n_rows + 1, 0);
  for (int p = 0; p < nin; p++) start[elems.row_idx[p] + 1]++;
  std::partial_sum(start.begin(), start.end(), start.begin());
  Vec<int> order(nin);
  {
    Vec<int> fill(start.begin(), start.end() - 1);
    for (int p = 0; p < nin; p++) order[fill[elems.row_idx[p]]++] = p;
  }
  for (int i = 0; i < n_rows; i++)
    std::stable_sort(order.begin() + start[i], order.begin() + start[i + 1],
                     [&](int a, int b) { return elems.col_idx[a] < elems.col_idx[b]; });
  row_idx.clear();
  col_idx.clear();
  val.clear();
  row_idx.reserve(nin);
  col_idx.reserve(nin);
  val.reserve(nin * bb);
  row_ptr.assign(n_rows + 1, 0);
  for (int p : order) {
    const int i = elems.row_idx[p], j = elems.col_idx[p];
    const T* src = &elems.val[(size_t)p * bb];
    const int cnt = (int)row_idx.size();
    if (cnt > 0 && row_idx[cnt - 1] == i && col_idx[cnt - 1] == j) {
      T* dst = &val[(size_t)(cnt - 1) * bb];
      for (size_t k = 0; k < bb; k++) dst[k] += src[k];
    } else {
      row_idx.push_back(i);
      col_idx.push_back(j);
      val.insert(val.end(), src, src + bb);
      row_ptr[i + 1]++;
    }
  }
  std::partial_sum(row_ptr.begin(), row_ptr.end(), row_ptr.begin());
}

template <typename T>
int SpMatBsr<T>::n_blocks() const {
  return (int)row_idx.size();
}

template <typename T>
void SpMatBsr<T>::clear() {
  row_idx.clear();
  col_idx.clear();
  row_ptr.clear();
  val.clear();
}

// src/spmat.cpp:145-153 (the reference aborts; this throws)
template <typename T>
int SpMatBsr<T>::find_index(int i, int j) const {
  for (int p = row_ptr[i]; p < row_ptr[i + 1]; p++)
    if (col_idx[p] == j) return p;
  throw std::out_of_range("Could not find block index (" + std::to_string(i) + "," +
                          std::to_string(j) + ") in BSR matrix");
}

template <typename T>
T* SpMatBsr<T>::operator()(int i, int j) {
  return &val[(size_t)b_len * b_len * find_index(i, j)];
}

template <typename T>
T const* SpMatBsr<T>::operator()(int i, int j) const {
  return &val[(size_t)b_len * b_len * find_index(i, j)];
}

template <typename T>
void SpMatBsr<T>::zeros() {
  std::fill(val.begin(), val.end(), T(0));
}

// src/spmat.cpp:171-189
template <typename T>
void SpMatBsr<T>::symmetrize() {
  const int b = b_len;
  for (int k = 0; k < n_blocks(); k++) {
    const int i = row_idx[k], j = col_idx[k];
    if (i < j) continue;
    T* v1b = &val[(size_t)b * b * k];
    T* v2b = (*this)(j, i);
    for (int bj = 0; bj < b; bj++)
      for (int bi = 0; bi < b; bi++) {
        T& v1 = v1b[bj * b + bi];
        T& v2 = v2b[bi * b + bj];
        const T s = (v1 + conj(v2)) / T(2);
        v1 = s;
        v2 = conj(s);
      }
  }
}

template <typename T>
void SpMatBsr<T>::scale(T alpha) {
  for (auto& v : val) v *= alpha;
}

template <typename T>
DenseMat<T> SpMatBsr<T>::to_dense() const {
  const int b = b_len;
  DenseMat<T> out((size_t)b * n_rows, (size_t)b * n_cols);
  for (int k = 0; k < n_blocks(); k++)
    for (int bj = 0; bj < b; bj++)
      for (int bi = 0; bi < b; bi++)
        out((size_t)b * row_idx[k] + bi, (size_t)b * col_idx[k] + bj) +=
            val[(size_t)b * b * k + bj * b + bi];
  return out;
}

template class SpMatElems<float>;
template class SpMatElems<double>;
template class SpMatElems<cx_float>;
template class SpMatElems<cx_double>;
template class SpMatBsr<float>;
template class SpMatBsr<double>;
template class SpMatBsr<cx_float>;
template class SpMatBsr<cx_double>;

}  // namespace fkpm
