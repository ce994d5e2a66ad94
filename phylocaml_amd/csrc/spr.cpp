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
#include <new>
#include <vector>

#include "../../include/phylo_hip.h"

namespace {

struct Tree {
  int32_t N;
  std::vector<int32_t> ch, parent;
  int32_t root;
};

// A rooted binary tree of N leaves: every node but the root has exactly one
// parent, and every node reaches the root (no node is its own child and no
// cycle is detached from the root).
bool load(int32_t N, const int32_t* children, int32_t root, Tree* t) {
  if (N < 3 || !children || root < N || root > 2 * N - 2) return false;
  const int32_t M = 2 * N - 1;
  t->N = N;
  t->ch.assign(children, children + 2 * (N - 1));
  t->parent.assign(M, -1);
  for (int32_t k = 0; k < N - 1; ++k)
    for (int c = 0; c < 2; ++c) {
      const int32_t x = t->ch[2 * k + c];
      if (x < 0 || x >= M || x == N + k || t->parent[x] >= 0) return false;
      t->parent[x] = N + k;
    }
  t->root = root;
  if (t->parent[root] >= 0) return false;
  // walk up from every node, marking nodes known to reach the root: each
  // node is walked over at most once after it is marked, so O(M) overall
  std::vector<char> ok(M, 0);
  ok[root] = 1;
  std::vector<int32_t> path;
  for (int32_t v = 0; v < M; ++v) {
    path.clear();
    int32_t y = v;
    while (!ok[y]) {
      path.push_back(y);
      if ((int32_t)path.size() > M) return false;  // a cycle
      y = t->parent[y];
      if (y < 0) return false;  // a second root
    }
    for (int32_t z : path) ok[z] = 1;
  }
  return true;
}

// C entry points return a status and never let an exception out: an
// allocation failure is PHY_ENOMEM, anything else PHY_EINVAL.
template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return PHY_ENOMEM;
  } catch (...) {
    return PHY_EINVAL;
  }
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

// Reroot the subtree of s (internal) on the edge above w (a descendant of s
// other than s): a root at the middle of that edge (id s reused), the path
// s -> w reversed with its branch lengths, and the former root s suppressed
// (its other child c' hangs from the first path node with the two edge
// lengths added).  Lengths in the subtree are conserved; t_s is unchanged.
void reroot_subtree(int32_t N, std::vector<int32_t>& ch, std::vector<int32_t>& par,
                    std::vector<double>& bl, int32_t s, int32_t w) {
  std::vector<int32_t> path;  // w, ..., s
  for (int32_t y = w; y != s; y = par[y]) path.push_back(y);
  path.push_back(s);
  std::reverse(path.begin(), path.end());  // u0 = s, ..., uk = w
  const int k = (int)path.size() - 1;
  auto slot = [&](int32_t u, int32_t c) -> int32_t& {
    return ch[2 * (u - N)] == c ? ch[2 * (u - N)] : ch[2 * (u - N) + 1];
  };
  const int32_t u1 = path[1];
  const int32_t cs = (ch[2 * (s - N)] == u1) ? ch[2 * (s - N) + 1] : ch[2 * (s - N)];
  const std::vector<double> old = bl;
  const double half = 0.5 * old[w];
  if (k == 1) {
    ch[2 * (s - N)] = w;
    ch[2 * (s - N) + 1] = cs;
    bl[w] = half;
    bl[cs] = half + old[cs];
    return;
  }
  for (int i = k - 1; i >= 1; --i) {
    const int32_t u = path[i];
    const int32_t nc = (i >= 2) ? path[i - 1] : cs;
    slot(u, path[i + 1]) = nc;
    par[nc] = u;
    bl[nc] = (i >= 2) ? old[u] : old[u1] + old[cs];
  }
  ch[2 * (s - N)] = w;
  ch[2 * (s - N) + 1] = path[k - 1];
  par[w] = s;
  par[path[k - 1]] = s;
  bl[w] = half;
  bl[path[k - 1]] = half;
}

int32_t sibling(const Tree& t, int32_t s) {
  const int32_t p = t.parent[s];
  return (t.ch[2 * (p - t.N)] == s) ? t.ch[2 * (p - t.N) + 1] : t.ch[2 * (p - t.N)];
}

// Edge distances in the unrooted tree left after pruning s: edges are keyed
// by their lower node in the rooted layout, the two edges at the pruned
// tree's root being one edge; the start is the edge the prune point lies on.
// dist[y] = number of edges between the start and the edge above y (-1 for
// nodes of the subtree of s, for p and for the pruned tree's root).
void prune_distances(const Tree& t, int32_t s, std::vector<int32_t>* dist) {
  const int32_t N = t.N, M = 2 * N - 1;
  const int32_t p = t.parent[s], x = sibling(t, s), g = t.parent[p];
  if (g < 0 && x < N) {  // the pruned tree is a single leaf: no edge left
    dist->assign(M, -1);
    return;
  }
  std::vector<int32_t> par = t.parent, ch = t.ch;
  int32_t r = t.root;
  if (g >= 0) {
    (ch[2 * (g - N)] == p ? ch[2 * (g - N)] : ch[2 * (g - N) + 1]) = x;
    par[x] = g;
  } else {
    r = x;
    par[x] = -1;
  }
  std::vector<char> in;
  mark_subtree(t, s, &in);
  in[p] = 1;
  auto key = [&](int32_t y) -> int32_t {  // unrooted edge above y
    if (par[y] == r) return std::min(ch[2 * (r - N)], ch[2 * (r - N) + 1]);
    return y;
  };
  const int32_t start = (g >= 0) ? key(x) : std::min(ch[2 * (x - N)], ch[2 * (x - N) + 1]);
  dist->assign(M, -1);
  std::vector<int32_t> ed(M, -1), q;
  ed[start] = 0;
  q.push_back(start);
  for (size_t h = 0; h < q.size(); ++h) {
    const int32_t e = q[h];
    // the edge e joins lower node e and upper node par[e] (or, for the merged
    // root edge, the two children of r)
    int32_t ends[2] = {e, par[e] == r ? (ch[2 * (r - N)] == e ? ch[2 * (r - N) + 1] : ch[2 * (r - N)])
                                      : par[e]};
    for (int32_t y : ends) {
      if (y < 0) continue;
      // edges meeting at node y: the edge above y and y's child edges
      int32_t cand[3] = {y != r ? key(y) : -1, y >= N ? ch[2 * (y - N)] : -1,
                         y >= N ? ch[2 * (y - N) + 1] : -1};
      for (int32_t c : cand) {
        if (c < 0) continue;
        const int32_t kc = key(c);
        if (ed[kc] < 0) {
          ed[kc] = ed[e] + 1;
          q.push_back(kc);
        }
      }
    }
  }
  for (int32_t y = 0; y < M; ++y)
    if (!in[y] && y != r) (*dist)[y] = ed[key(y)];
}

int phy_spr_apply_impl(int32_t N, const int32_t* children, int32_t root, const double* blen,
                       int32_t s, int32_t v, int32_t* children_out, int32_t* root_out,
                       double* blen_out);

int phy_tbr_apply_impl(int32_t N, const int32_t* children, int32_t root, const double* blen, int32_t s,
                       int32_t w, int32_t v, int32_t* children_out, int32_t* root_out,
                       double* blen_out) {
  Tree t;
  if (!children_out || !root_out || !blen || !blen_out || !load(N, children, root, &t))
    return PHY_EINVAL;
  const int32_t M = 2 * N - 1;
  if (s < 0 || s >= M || s == root || w < 0 || w >= M) return PHY_EINVAL;
  if (w == s) return phy_spr_apply_impl(N, children, root, blen, s, v, children_out, root_out, blen_out);
  std::vector<char> in;
  mark_subtree(t, s, &in);
  if (!in[w] || s < N) return PHY_EINVAL;
  std::vector<int32_t> ch = t.ch, par = t.parent;
  std::vector<double> bl(blen, blen + M);
  reroot_subtree(N, ch, par, bl, s, w);
  if (v == -1) {  // keep the attachment: the rerooted subtree stays in place
    std::copy(ch.begin(), ch.end(), children_out);
    std::copy(bl.begin(), bl.end(), blen_out);
    *root_out = root;
    return PHY_OK;
  }
  return phy_spr_apply_impl(N, ch.data(), root, bl.data(), s, v, children_out, root_out, blen_out);
}

int phy_tbr_moves_impl(int32_t N, const int32_t* children, int32_t root, int32_t max_join,
                       int32_t max_reroot, int64_t* n_moves, int32_t* moves, int64_t capacity) {
  Tree t;
  if (!n_moves || !load(N, children, root, &t)) return PHY_EINVAL;
  const int32_t M = 2 * N - 1;
  std::vector<int32_t> depth(M, 0), post;
  {
    std::vector<int32_t> st = {root};
    while (!st.empty()) {
      const int32_t y = st.back();
      st.pop_back();
      if (y >= N)
        for (int c = 0; c < 2; ++c) {
          const int32_t z = t.ch[2 * (y - N) + c];
          depth[z] = depth[y] + 1;
          st.push_back(z);
        }
    }
  }
  int64_t n = 0;
  std::vector<char> in;
  std::vector<int32_t> dist;
  for (int32_t s = 0; s < M; ++s) {
    if (s == root) continue;
    mark_subtree(t, s, &in);
    prune_distances(t, s, &dist);
    // reroot edges: s itself, or an edge strictly below the children of s
    std::vector<int32_t> ws = {s};
    if (s >= N && max_reroot != 0)
      for (int32_t y = 0; y < M; ++y) {
        const int d = depth[y] - depth[s] - 1;  // edges between s's child edges and y's
        if (in[y] && d >= 1 && (max_reroot < 0 || d <= max_reroot)) ws.push_back(y);
      }
    auto put = [&](int32_t w, int32_t v) {
      if (moves && n < capacity) {
        moves[3 * n] = s;
        moves[3 * n + 1] = w;
        moves[3 * n + 2] = v;
      }
      ++n;
    };
    // v = -1: the subtree keeps its attachment (join distance 0), so only a
    // reroot changes the tree
    for (size_t i = 1; i < ws.size(); ++i) put(ws[i], -1);
    for (int32_t v = 0; v < M; ++v) {
      const int32_t d = dist[v];
      if (d < 1 || (max_join >= 0 && d > max_join)) continue;
      for (int32_t w : ws) put(w, v);
    }
  }
  *n_moves = n;
  return PHY_OK;
}

int phy_nni_moves_impl(int32_t N, const int32_t* children, int32_t root, int64_t* n_moves,
                       int32_t* moves, int64_t capacity) {
  Tree t;
  if (!n_moves || !load(N, children, root, &t)) return PHY_EINVAL;
  int64_t n = 0;
  auto put = [&](int32_t s, int32_t v) {
    if (moves && n < capacity) {
      moves[2 * n] = s;
      moves[2 * n + 1] = v;
    }
    ++n;
  };
  const int32_t ra = t.ch[2 * (root - N)], rb = t.ch[2 * (root - N) + 1];
  for (int32_t w = N; w < 2 * N - 1; ++w) {
    if (w == root) continue;
    const int32_t u = t.parent[w];
    const int32_t c1 = t.ch[2 * (w - N)], c2 = t.ch[2 * (w - N) + 1];
    if (u != root) {
      // internal edge (u, w): the two swaps of w's children with w's sibling
      const int32_t y = sibling(t, w);
      put(c1, y);
      put(c2, y);
    } else if (w == std::min(ra, rb) && ra >= N && rb >= N) {
      // the root's two edges are one internal edge between w and y
      const int32_t y = (w == ra) ? rb : ra;
      put(c1, t.ch[2 * (y - N)]);
      put(c1, t.ch[2 * (y - N) + 1]);
    }
  }
  *n_moves = n;
  return PHY_OK;
}

int phy_spr_moves_impl(int32_t N, const int32_t* children, int32_t root, int64_t* n_moves,
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

int phy_spr_apply_impl(int32_t N, const int32_t* children, int32_t root, const double* blen, int32_t s,
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

}  // namespace

extern "C" {

int phy_tbr_apply(int32_t N, const int32_t* children, int32_t root, const double* blen, int32_t s,
                  int32_t w, int32_t v, int32_t* children_out, int32_t* root_out,
                  double* blen_out) {
  return guarded([&] {
    return phy_tbr_apply_impl(N, children, root, blen, s, w, v, children_out, root_out, blen_out);
  });
}

int phy_tbr_moves(int32_t N, const int32_t* children, int32_t root, int32_t max_join,
                  int32_t max_reroot, int64_t* n_moves, int32_t* moves, int64_t capacity) {
  return guarded([&] {
    return phy_tbr_moves_impl(N, children, root, max_join, max_reroot, n_moves, moves, capacity);
  });
}

int phy_nni_moves(int32_t N, const int32_t* children, int32_t root, int64_t* n_moves,
                  int32_t* moves, int64_t capacity) {
  return guarded([&] { return phy_nni_moves_impl(N, children, root, n_moves, moves, capacity); });
}

int phy_spr_moves(int32_t N, const int32_t* children, int32_t root, int64_t* n_moves,
                  int32_t* moves, int64_t capacity) {
  return guarded([&] { return phy_spr_moves_impl(N, children, root, n_moves, moves, capacity); });
}

int phy_spr_apply(int32_t N, const int32_t* children, int32_t root, const double* blen, int32_t s,
                  int32_t v, int32_t* children_out, int32_t* root_out, double* blen_out) {
  return guarded([&] {
    return phy_spr_apply_impl(N, children, root, blen, s, v, children_out, root_out, blen_out);
  });
}

}  // extern "C"
