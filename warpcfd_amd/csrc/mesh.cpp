// Host-side face matching for the mesh preprocessing (no GPU work).
//
// Every cell lists its faces as node ids in a fixed local order.  Two slots
// (cell, local face) hold the same face when their node sets are equal.  The
// unique faces are numbered in the lexicographic order of their ascending
// node ids, and each face keeps its one (boundary) or two (internal) slots in
// ascending slot order, i.e. the lower cell id first.
//
// Method: each slot's node ids are sorted (at most four, insertion sort);
// the slots are then ordered by a stable counting sort on the smallest node
// id, and each run of equal smallest id (a handful of slots: the faces that
// meet at one node) is finished by an insertion sort on the remaining ids
// and the slot index.  O(slots + nodes) work, no allocation proportional to
// anything but those two counts.
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"

namespace {

constexpr int kMaxFaceNodes = 4;

struct Key {
  int32_t v[kMaxFaceNodes];
};

inline bool key_less(const Key& a, const Key& b, int n) {
  for (int i = 0; i < n; ++i)
    if (a.v[i] != b.v[i]) return a.v[i] < b.v[i];
  return false;
}

inline bool key_equal(const Key& a, const Key& b, int n) {
  for (int i = 0; i < n; ++i)
    if (a.v[i] != b.v[i]) return false;
  return true;
}

}  // namespace

extern "C" int64_t wcfd_match_faces(const int32_t* face_nodes, int64_t n_cells,
                                    int32_t faces_per_cell, int32_t nodes_per_face,
                                    int32_t* face_of_slot, int32_t* unique_nodes,
                                    int32_t* face_slots, int64_t capacity) {
  if (!face_nodes || !face_of_slot || !unique_nodes || !face_slots || n_cells < 0 ||
      faces_per_cell <= 0 || nodes_per_face < 1 || nodes_per_face > kMaxFaceNodes ||
      capacity < 0)
    return -WCFD_EINVAL;
  const int n = nodes_per_face;
  const int64_t S = n_cells * faces_per_cell;
  if (S > INT32_MAX) return -WCFD_EINVAL;
  std::vector<Key> key((size_t)S);
  int32_t max_node = -1;
  for (int64_t s = 0; s < S; ++s) {
    Key k{};
    for (int i = 0; i < n; ++i) {
      int32_t x = face_nodes[s * n + i];
      if (x < 0) return -WCFD_EINVAL;
      int j = i;
      for (; j > 0 && k.v[j - 1] > x; --j) k.v[j] = k.v[j - 1];
      k.v[j] = x;
    }
    if (k.v[n - 1] > max_node) max_node = k.v[n - 1];
    key[(size_t)s] = k;
  }
  // stable counting sort of the slots on the smallest node id
  std::vector<int32_t> start((size_t)max_node + 2, 0);
  for (int64_t s = 0; s < S; ++s) ++start[(size_t)key[(size_t)s].v[0] + 1];
  for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
  std::vector<int32_t> order((size_t)S);
  {
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int64_t s = 0; s < S; ++s) order[(size_t)fill[(size_t)key[(size_t)s].v[0]]++] = (int32_t)s;
  }
  // each run of one smallest id: insertion sort on (key, slot)
  for (size_t r = 0; r + 1 < start.size(); ++r) {
    const int32_t lo = start[r], hi = start[r + 1];
    for (int32_t i = lo + 1; i < hi; ++i) {
      const int32_t s = order[(size_t)i];
      int32_t j = i;
      while (j > lo) {
        const int32_t t = order[(size_t)j - 1];
        const Key& a = key[(size_t)s];
        const Key& b = key[(size_t)t];
        if (!(key_less(a, b, n) || (key_equal(a, b, n) && s < t))) break;
        order[(size_t)j] = t;
        --j;
      }
      order[(size_t)j] = s;
    }
  }
  // number the faces and record their slots
  int64_t K = 0;
  for (int64_t i = 0; i < S;) {
    const int32_t s0 = order[(size_t)i];
    const Key& k = key[(size_t)s0];
    int64_t j = i + 1;
    while (j < S && key_equal(key[(size_t)order[(size_t)j]], k, n)) ++j;
    if (j - i > 2) return -WCFD_EMESH;
    if (K >= capacity) return -WCFD_ECAPACITY;
    std::memcpy(unique_nodes + K * n, k.v, sizeof(int32_t) * (size_t)n);
    face_slots[2 * K] = s0;
    face_slots[2 * K + 1] = j - i == 2 ? order[(size_t)i + 1] : -1;
    for (int64_t q = i; q < j; ++q) face_of_slot[order[(size_t)q]] = (int32_t)K;
    ++K;
    i = j;
  }
  return K;
}
