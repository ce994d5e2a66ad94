// SPR (subtree prune and regraft) neighbourhood of a rooted binary tree, host
// side.  The reference's generators (generate_spr, lib/treeNeighborhoods.ml:
// 27-75) never compiled; its Tree.break / Tree.join (lib/tree.ml:299-494)
// define the operation on unrooted trees with handles.  Here, on the rooted
// layout of phy_topo_create:
//
//   prune s (any non-root node): its parent p is removed and its sibling x
//     takes p's place under g = parent(p), with t_x += t_p; if p was the root,
//     x becomes the root;
//   regraft at v (any node of the pruned tree except its root): p is reused
//     as a new node on the edge above v, children (v, s), t_p = t_v / 2 and
//     t_v = t_v / 2; t_s is unchanged.
//
// A move is the pair (s, v) with v outside the subtree of s and v != x (x
// would rebuild the same tree).  Different moves can give the same tree;
// moves are enumerated, not unique trees.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/phylo_hip.h"

namespace {

struct Tree {
  int32_t N;
  std::vector<int32_t> ch, parent;
  int32_t root;
};

bool load(int32_t N, const int32_t* children, int32_t root, Tree* t) {
  if (N < 3 || !children || root < N || root > 2 * N - 2) return false;
  t->N = N;
  t->ch.assign(children, children + 2 * (N - 1));
  t->parent.assign(2 * N - 1, -1);
  for (int32_t k = 0; k < N - 1; ++k)
    for (int c = 0; c < 2; ++c) {
      const int32_t x = t->ch[2 * k + c];
      if (x < 0 || x > 2 * N - 2 || t->parent[x] >= 0) return false;
      t->parent[x] = N + k;
    }
  t->root = root;
  return t->parent[root] < 0;
}

// mark the subtree of s
void mark_subtree(const Tree& t, int32_t s, std::vector<char>* in) {
  in->assign(2 * t.N - 1, 0);
  std::vector<int32_t> st = {s};
  while (!st.empty()) {
    const int32_t v = st.back();
    st.pop_back();
    (*in)[v] = 1;
    if (v >= t.N) {
      st.push_back(t.ch[2 * (v - t.N)]);
      st.push_back(t.ch[2 * (v - t.N) + 1]);
    }
  }
}

}  // namespace

extern "C" {

int phy_spr_moves(int32_t N, const int32_t* children, int32_t root, int64_t* n_moves,
                  int32_t* moves, int64_t capacity) {
  Tree t;
  if (!n_moves || !load(N, children, root, &t)) return PHY_EINVAL;
  int64_t n = 0;
  std::vector<char> in;
  for (int32_t s = 0; s < 2 * N - 1; ++s) {
    if (s == root) continue;
    const int32_t p = t.parent[s];
    const int32_t x = (t.ch[2 * (p - N)] == s) ? t.ch[2 * (p - N) + 1] : t.ch[2 * (p - N)];
    const int32_t new_root = (p == root) ? x : root;
    mark_subtree(t, s, &in);
    for (int32_t v = 0; v < 2 * N - 1; ++v) {
      if (in[v] || v == p || v == x || v == new_root) continue;
      if (moves && n < capacity) {
        moves[2 * n] = s;
        moves[2 * n + 1] = v;
      }
      ++n;
    }
  }
  *n_moves = n;
  return PHY_OK;
}

int phy_spr_apply(int32_t N, const int32_t* children, int32_t root, const double* blen, int32_t s,
                  int32_t v, int32_t* children_out, int32_t* root_out, double* blen_out) {
  Tree t;
  if (!children_out || !root_out || !blen || !blen_out || !load(N, children, root, &t))
    return PHY_EINVAL;
  const int32_t M = 2 * N - 1;
  if (s < 0 || s >= M || s == root || v < 0 || v >= M) return PHY_EINVAL;
  const int32_t p = t.parent[s];
  const int32_t x = (t.ch[2 * (p - N)] == s) ? t.ch[2 * (p - N) + 1] : t.ch[2 * (p - N)];
  const int32_t g = t.parent[p];
  std::vector<char> in;
  mark_subtree(t, s, &in);
  const int32_t new_root = (p == root) ? x : root;
  if (in[v] || v == p || v == x || v == new_root) return PHY_EINVAL;
  std::vector<int32_t> ch = t.ch, par = t.parent;
  std::vector<double> bl(blen, blen + M);
  // prune
  int32_t r = root;
  if (g >= 0) {
    int32_t* slot = (ch[2 * (g - N)] == p) ? &ch[2 * (g - N)] : &ch[2 * (g - N) + 1];
    *slot = x;
    par[x] = g;
    bl[x] += bl[p];
  } else {
    r = x;
    par[x] = -1;
    bl[x] = 0.0;
  }
  // regraft on the edge above v
  const int32_t u = par[v];
  int32_t* slot = (ch[2 * (u - N)] == v) ? &ch[2 * (u - N)] : &ch[2 * (u - N) + 1];
  *slot = p;
  par[p] = u;
  ch[2 * (p - N)] = v;
  ch[2 * (p - N) + 1] = s;
  par[v] = p;
  par[s] = p;
  const double half = 0.5 * bl[v];
  bl[v] = half;
  bl[p] = half;
  bl[r] = 0.0;
  std::copy(ch.begin(), ch.end(), children_out);
  std::copy(bl.begin(), bl.end(), blen_out);
  *root_out = r;
  return PHY_OK;
}

}  // extern "C"
