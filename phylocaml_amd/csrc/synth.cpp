// Seeded synthetic workloads (host only).  See include/phylo_synth.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/phylo_hip.h"
#include "../../include/phylo_synth.h"
#include "model_host.hpp"
#include "rng.hpp"

namespace {

unsigned host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  return std::min(n, 16u);
}

template <class F>
void parallel_for(int64_t n, F f) {
  const unsigned nt = host_threads();
  if (n < 4096 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min<int64_t>(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

uint64_t hash_bytes(const uint8_t* p, size_t n) {
  uint64_t h = 0x243F6A8885A308D3ull;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t v;
    std::memcpy(&v, p + i, 8);
    h = phy::mix64(h ^ v);
  }
  uint64_t tail = 0;
  for (size_t k = 0; i < n; ++i, ++k) tail |= (uint64_t)p[i] << (8 * k);
  return phy::mix64(h ^ tail ^ (n * 0x9E3779B97F4A7C15ull));
}

}  // namespace

extern "C" {

int phy_synth_yule(int32_t N, uint64_t seed, double mean_bl, int32_t* children, int32_t* root,
                   double* blen) {
  if (N < 2 || !children || !root || !blen) return 1;
  // temporary nodes: 0 = root; a split turns a tip into an internal node with
  // two new tips.
  phy::Rng rng(seed, 1);
  std::vector<int32_t> tips = {1, 2};
  std::vector<int32_t> kid(2 * (size_t)(2 * N - 1), -1);
  kid[0] = 1;
  kid[1] = 2;
  int32_t next = 3;
  while ((int32_t)tips.size() < N) {
    const uint32_t k = rng.below((uint32_t)tips.size());
    const int32_t t = tips[k];
    kid[2 * (size_t)t] = next;
    kid[2 * (size_t)t + 1] = next + 1;
    tips[k] = next;
    tips.push_back(next + 1);
    next += 2;
  }
  // relabel: tips -> 0..N-1 in creation order, internals -> N.. in creation order
  std::vector<int32_t> lab(2 * N - 1, -1);
  int32_t nl = 0, ni = N;
  for (int32_t v = 0; v < 2 * N - 1; ++v) lab[v] = (kid[2 * (size_t)v] < 0) ? nl++ : ni++;
  for (int32_t v = 0; v < 2 * N - 1; ++v) {
    if (kid[2 * (size_t)v] < 0) continue;
    const int32_t k = lab[v] - N;
    children[2 * k] = lab[kid[2 * (size_t)v]];
    children[2 * k + 1] = lab[kid[2 * (size_t)v + 1]];
  }
  *root = lab[0];
  phy::Rng brng(seed, 2);
  for (int32_t v = 0; v < 2 * N - 1; ++v) blen[v] = brng.exponential(mean_bl);
  blen[*root] = 0.0;
  return 0;
}

int phy_synth_fitch_leaves(int32_t N, int64_t S, uint64_t seed, uint64_t* out, int64_t stride) {
  if (N < 1 || S < 0 || !out) return 1;
  const int64_t W = phy_words(S);
  if (stride < W) return 1;
  static const uint8_t two[6] = {3, 5, 6, 9, 10, 12};
  parallel_for((int64_t)N * W, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      const int64_t leaf = i / W, w = i % W;
      phy::Rng r(seed, (uint64_t)leaf * 0x100000000ull + (uint64_t)w + 7);
      uint64_t word = 0;
      for (int k = 0; k < 16; ++k) {
        const int64_t s = w * 16 + k;
        if (s >= S) break;
        const uint64_t x = r.next();
        const uint32_t u = (uint32_t)(x >> 32) % 10000u;  // 0.01% resolution
        uint64_t code;
        if (u < 9500u) code = 1ull << ((x >> 8) & 3);
        else if (u < 9900u) code = two[(uint32_t)(x & 0xffff) % 6u];
        else code = 0xF;
        word |= code << (4 * k);
      }
      out[leaf * stride + w] = word;
    }
  });
  for (int64_t leaf = 0; leaf < N; ++leaf)
    for (int64_t w = W; w < stride; ++w) out[leaf * stride + w] = 0;
  return 0;
}

int phy_synth_dirichlet(int32_t n, double conc, uint64_t seed, double* x) {
  if (n < 1 || !(conc > 0) || !x) return 1;
  phy::Rng r(seed, 3);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += (x[i] = r.gamma(conc));
  for (int i = 0; i < n; ++i) x[i] /= s;
  return 0;
}

int phy_synth_gtr_params(int32_t n, double pi_conc, uint64_t seed, double* pi, double* co) {
  if (n < 2 || !pi || !co) return 1;
  const int m = n * (n - 1) / 2;
  std::vector<double> ex(m);
  phy_synth_dirichlet(m, 1.0, phy::mix64(seed ^ 0x5151), ex.data());
  for (int k = 0; k < m - 1; ++k) co[k] = ex[k] / ex[m - 1];
  phy_synth_dirichlet(n, pi_conc, phy::mix64(seed ^ 0xA1A1), pi);
  return 0;
}

int phy_synth_simulate(int32_t N, const int32_t* children, int32_t root, const double* blen,
                       int32_t K, const double* pi, const double* Q, int32_t R,
                       const double* rates, int64_t S, uint64_t seed, int32_t unique,
                       uint8_t* states) {
  if (N < 2 || K < 2 || K > 64 || R < 1 || S < 0 || !states) return 1;
  const int32_t M = 2 * N - 1;
  std::vector<double> U((size_t)K * K), lam(K), Ui((size_t)K * K);
  std::string err;
  if (!phy::eigen_reversible(K, Q, U.data(), lam.data(), Ui.data(), &err)) return 1;
  // cumulative transition rows per (node, rate)
  std::vector<double> cum((size_t)M * R * K * K);
  std::vector<double> P((size_t)K * K);
  for (int32_t v = 0; v < M; ++v)
    for (int r = 0; r < R; ++r) {
      phy::compose(K, U.data(), lam.data(), Ui.data(), rates[r] * blen[v], P.data());
      double* c = &cum[(((size_t)v * R + r) * K) * K];
      for (int i = 0; i < K; ++i) {
        double s = 0.0;
        for (int j = 0; j < K; ++j) {
          s += std::max(0.0, P[(size_t)i * K + j]);
          c[(size_t)i * K + j] = s;
        }
        for (int j = 0; j < K; ++j) c[(size_t)i * K + j] /= s;
      }
    }
  std::vector<double> picum(K);
  {
    double s = 0.0;
    for (int i = 0; i < K; ++i) picum[i] = (s += pi[i]);
    for (int i = 0; i < K; ++i) picum[i] /= s;
  }
  // pre-order node list
  std::vector<int32_t> order;
  order.reserve(M);
  std::vector<int32_t> st = {root};
  while (!st.empty()) {
    const int32_t v = st.back();
    st.pop_back();
    order.push_back(v);
    if (v >= N) {
      st.push_back(children[2 * (v - N)]);
      st.push_back(children[2 * (v - N) + 1]);
    }
  }
  auto draw = [&](phy::Rng& r, const double* c) {
    const double u = r.uniform();
    int j = 0;
    while (j < K - 1 && u >= c[j]) ++j;
    return j;
  };
  // simulate one site with stream = draw index; write leaf column into col (N bytes)
  auto sim_site = [&](int64_t idx, uint8_t* col, std::vector<uint8_t>& nodest) {
    phy::Rng r(seed, (uint64_t)idx + 11);
    const int rate = (int)r.below((uint32_t)R);
    nodest[root] = (uint8_t)draw(r, picum.data());
    // children drawn from parents, in pre-order
    for (size_t q = 0; q < order.size(); ++q) {
      const int32_t v = order[q];
      if (v < N) continue;
      for (int c = 0; c < 2; ++c) {
        const int32_t ch = children[2 * (v - N) + c];
        const double* row = &cum[(((size_t)ch * R + rate) * K + nodest[v]) * K];
        nodest[ch] = (uint8_t)draw(r, row);
      }
    }
    for (int32_t l = 0; l < N; ++l) col[l] = nodest[l];
  };
  if (!unique) {
    parallel_for(S, [&](int64_t lo, int64_t hi) {
      std::vector<uint8_t> nodest(M), col(N);
      for (int64_t s = lo; s < hi; ++s) {
        sim_site(s, col.data(), nodest);
        for (int32_t l = 0; l < N; ++l) states[(size_t)l * S + s] = col[l];
      }
    });
    return 0;
  }
  // unique: draw in chunks (parallel), filter sequentially in draw order
  std::unordered_map<uint64_t, int64_t> seen;
  seen.reserve((size_t)S * 2);
  const int64_t chunk = 1 << 16;
  std::vector<uint8_t> buf((size_t)chunk * N);
  int64_t got = 0, drawn = 0;
  while (got < S) {
    if (drawn > 64 * S + 1024) return 1;
    parallel_for(chunk, [&](int64_t lo, int64_t hi) {
      std::vector<uint8_t> nodest(M);
      for (int64_t s = lo; s < hi; ++s) sim_site(drawn + s, &buf[(size_t)s * N], nodest);
    });
    for (int64_t s = 0; s < chunk && got < S; ++s) {
      const uint8_t* col = &buf[(size_t)s * N];
      const uint64_t h = hash_bytes(col, N);
      auto it = seen.find(h);
      bool dup = false;
      if (it != seen.end()) {
        dup = true;
        for (int32_t l = 0; l < N; ++l)
          if (states[(size_t)l * S + it->second] != col[l]) {
            dup = false;
            break;
          }
      }
      if (dup) continue;
      if (it == seen.end()) seen.emplace(h, got);
      for (int32_t l = 0; l < N; ++l) states[(size_t)l * S + got] = col[l];
      ++got;
    }
    drawn += chunk;
  }
  return 0;
}

int phy_synth_compress(int32_t N, int64_t S, const uint8_t* states, uint8_t* patterns,
                       uint32_t* weights, int64_t* n_patterns) {
  if (N < 1 || S < 0 || !states || !patterns || !weights || !n_patterns) return 1;
  std::unordered_multimap<uint64_t, int64_t> seen;
  std::vector<uint8_t> col(N);
  int64_t np = 0;
  for (int64_t s = 0; s < S; ++s) {
    for (int32_t l = 0; l < N; ++l) col[l] = states[(size_t)l * S + s];
    const uint64_t h = hash_bytes(col.data(), N);
    auto range = seen.equal_range(h);
    int64_t found = -1;
    for (auto it = range.first; it != range.second && found < 0; ++it) {
      bool eq = true;
      for (int32_t l = 0; l < N; ++l)
        if (patterns[(size_t)l * S + it->second] != col[l]) {
          eq = false;
          break;
        }
      if (eq) found = it->second;
    }
    if (found >= 0) {
      weights[found]++;
      continue;
    }
    for (int32_t l = 0; l < N; ++l) patterns[(size_t)l * S + np] = col[l];
    weights[np] = 1;
    seen.emplace(h, np);
    ++np;
  }
  *n_patterns = np;
  return 0;
}

int phy_synth_poisson_weights(int64_t n, double lambda, uint64_t seed, uint32_t* w) {
  if (n < 0 || !w) return 1;
  parallel_for(n, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      phy::Rng r(seed, (uint64_t)i + 13);
      w[i] = r.poisson(lambda) + 1u;
    }
  });
  return 0;
}

int phy_pack_dna(int32_t n_rows, int64_t S, const uint8_t* st, int64_t pitch, uint64_t* out,
                 int64_t stride) {
  const int64_t W = phy_words(S);
  if (n_rows < 0 || S < 0 || !st || !out || stride < W || pitch < S) return 1;
  parallel_for((int64_t)n_rows * stride, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      const int64_t row = i / stride, w = i % stride;
      uint64_t word = 0;
      for (int k = 0; k < 16; ++k) {
        const int64_t s = w * 16 + k;
        if (s >= S) break;
        const uint8_t x = st[row * pitch + s];
        const uint64_t code = (x < 4) ? (1ull << x) : 0xFull;
        word |= code << (4 * k);
      }
      out[i] = word;
    }
  });
  return 0;
}

}  // extern "C"
