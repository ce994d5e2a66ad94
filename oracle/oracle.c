/*
 * oracle.c -- CPU restatement of the reference algorithms for the Fitch and
 * Felsenstein hot paths.  TEST INFRASTRUCTURE ONLY (checker for tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg).  Each function
 * cites the reference file:line it follows; see oracle.h for what pins it.
 *
 * Written for clarity, in plain C with OpenMP over independent sites.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================== */
/* Fitch                                                                    */
/* ======================================================================== */

/* NonAdditive_c.median_2 (lib/nonAdditive_c.ml:19-35): test_one x_i y_i =
 * let m = x_i land y_i in if m = 0 then (x_i lor y_i, 1) else (m, 0);
 * cost accumulates x.weights.(i) * c. */
double orc_fitch_median2(const uint8_t* x, const uint8_t* y, uint8_t* out, int64_t n,
                         const double* w) {
  double cost = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t m = x[i] & y[i];
    int c = 0;
    uint8_t r = m;
    if (m == 0) {
      r = x[i] | y[i];
      c = 1;
    }
    if (out) out[i] = r;
    cost += (w ? w[i] : 1.0) * (double)c;
  }
  return cost;
}

/* bv_distance (lib/bitvector/bv.c:46-55). */
int64_t orc_fitch_distance(const uint8_t* a, const uint8_t* b, int64_t n) {
  int64_t r = 0;
  for (int64_t i = 0; i < n; ++i)
    if ((a[i] & b[i]) == 0) r++;
  return r;
}

double orc_fitch_median3(const uint8_t* a, const uint8_t* b, const uint8_t* c, uint8_t* out,
                         int64_t n, const double* w) {
  double cost = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    int best = 0;
    uint8_t set = 0;
    for (int s = 0; s < 8; ++s) {
      const int cnt = ((a[i] >> s) & 1) + ((b[i] >> s) & 1) + ((c[i] >> s) & 1);
      if (cnt > best) {
        best = cnt;
        set = (uint8_t)(1u << s);
      } else if (cnt == best && cnt > 0) {
        set |= (uint8_t)(1u << s);
      }
    }
    out[i] = set;
    if (best > 0) cost += (w ? w[i] : 1.0) * (double)(3 - best);
  }
  return cost;
}

/* Fitch (1971) final-state rules (I-V); no reference implementation exists
 * (Node.Make1D.final_states is TODO, lib/node.ml:258-270). */
static inline uint8_t final_rule(uint8_t pv, uint8_t pa, uint8_t pb, uint8_t fp) {
  if ((fp & (uint8_t)~pv) == 0) return fp;         /* F_p subset of P_v      */
  if ((pa & pb) != 0) return pv | (fp & (pa | pb)); /* P_v from intersection  */
  return pv | fp;                                   /* P_v from union         */
}

void orc_fitch_adjust3(const uint8_t* pv, const uint8_t* pa, const uint8_t* pb,
                       const uint8_t* fp, uint8_t* out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = final_rule(pv[i], pa[i], pb[i], fp[i]);
}

/* post-order node list (lib/tree.ml:171-187 visits children before parents) */
static int32_t* post_order(int32_t N, const int32_t* ch, int32_t root) {
  const int32_t M = 2 * N - 1;
  int32_t* out = (int32_t*)malloc(sizeof(int32_t) * M);
  int32_t* st = (int32_t*)malloc(sizeof(int32_t) * 2 * M);
  int32_t* stage = (int32_t*)calloc(M, sizeof(int32_t));
  int32_t sp = 0, k = 0;
  st[sp++] = root;
  while (sp) {
    const int32_t v = st[sp - 1];
    if (v < N || stage[v] == 2) {
      out[k++] = v;
      --sp;
      continue;
    }
    st[sp++] = ch[2 * (v - N) + stage[v]];
    stage[v]++;
  }
  free(st);
  free(stage);
  return out;
}

uint64_t orc_fitch_downpass(int32_t N, const int32_t* ch, int32_t root, uint8_t* sets, int64_t S,
                            const uint32_t* w, uint32_t* site_cost, uint64_t* node_cost,
                            int32_t n_threads) {
  const int32_t M = 2 * N - 1;
  int32_t* post = post_order(N, ch, root);
  const int64_t B = 4096; /* site block kept in cache across all nodes */
  const int64_t nb = (S + B - 1) / B;
  uint64_t total = 0;
  if (node_cost) memset(node_cost, 0, sizeof(uint64_t) * (N - 1));
#ifdef _OPENMP
  if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1) reduction(+ : total)
#endif
  for (int64_t blk = 0; blk < nb; ++blk) {
    const int64_t s0 = blk * B, s1 = (s0 + B < S) ? s0 + B : S;
    uint32_t cnt[4096];
    memset(cnt, 0, sizeof(cnt));
    for (int32_t q = 0; q < M; ++q) {
      const int32_t v = post[q];
      if (v < N) continue;
      const uint8_t* a = sets + (int64_t)ch[2 * (v - N)] * S;
      const uint8_t* b = sets + (int64_t)ch[2 * (v - N) + 1] * S;
      uint8_t* o = sets + (int64_t)v * S;
      uint64_t nc = 0;
      for (int64_t s = s0; s < s1; ++s) {
        const uint8_t m = a[s] & b[s];
        if (m == 0) {
          o[s] = a[s] | b[s];
          cnt[s - s0]++;
          nc += w ? w[s] : 1u;
        } else {
          o[s] = m;
        }
      }
      if (node_cost && nc) {
#ifdef _OPENMP
#pragma omp atomic
#endif
        node_cost[v - N] += nc;
      }
    }
    for (int64_t s = s0; s < s1; ++s) {
      if (site_cost) site_cost[s] = cnt[s - s0];
      total += (uint64_t)cnt[s - s0] * (w ? w[s] : 1u);
    }
  }
  free(post);
  return total;
}

void orc_fitch_uppass(int32_t N, const int32_t* ch, int32_t root, const uint8_t* sets,
                      uint8_t* fin, int64_t S, int32_t n_threads) {
  const int32_t M = 2 * N - 1;
  int32_t* post = post_order(N, ch, root);
  /* reverse post-order visits parents before children */
  memcpy(fin + (int64_t)root * S, sets + (int64_t)root * S, (size_t)S);
  (void)n_threads;
  for (int32_t q = M - 1; q >= 0; --q) {
    const int32_t v = post[q];
    if (v < N) continue;
    const int32_t a = ch[2 * (v - N)], b = ch[2 * (v - N) + 1];
    const uint8_t* fv = fin + (int64_t)v * S;
    const uint8_t* pa = sets + (int64_t)a * S;
    const uint8_t* pb = sets + (int64_t)b * S;
    for (int c = 0; c < 2; ++c) {
      const int32_t x = c ? b : a;
      const uint8_t* px = sets + (int64_t)x * S;
      uint8_t* fx = fin + (int64_t)x * S;
      if (x < N) {
        for (int64_t s = 0; s < S; ++s) {
          const uint8_t i = fv[s] & px[s];
          fx[s] = i ? i : px[s];
        }
      } else {
        const uint8_t* pxa = sets + (int64_t)ch[2 * (x - N)] * S;
        const uint8_t* pxb = sets + (int64_t)ch[2 * (x - N) + 1] * S;
        for (int64_t s = 0; s < S; ++s) fx[s] = final_rule(px[s], pxa[s], pxb[s], fv[s]);
      }
    }
    (void)pa;
    (void)pb;
  }
  free(post);
}

/* Sankoff (1975) over an integer cost matrix, in the cost convention of the
 * median search of lib/costMatrix.ml:52-65 (a child in state j under a parent
 * in state k costs tcm[j][k]):  C_leaf[k] = 0 for k in the leaf set (an empty
 * set means every state), else infinite;  C_v[k] = sum over children of
 * min_j (C_c[j] + tcm[j][k]);  site cost = min_k C_root[k].  int64 throughout.
 * Returns sum_s w_s site_cost_s. */
uint64_t orc_sankoff(int32_t N, const int32_t* ch, int32_t root, int32_t K, const int32_t* tcm,
                     const uint8_t* leaves, int64_t S, const uint32_t* w, uint32_t* site_cost,
                     int32_t n_threads) {
  const int32_t M = 2 * N - 1;
  const int64_t INF = INT64_MAX / 4;
  int32_t* post = post_order(N, ch, root);
  uint64_t total = 0;
#ifdef _OPENMP
  if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(static) reduction(+ : total)
#endif
  for (int64_t s = 0; s < S; ++s) {
    int64_t* C = (int64_t*)malloc(sizeof(int64_t) * (size_t)M * K);
    for (int32_t q = 0; q < M; ++q) {
      const int32_t v = post[q];
      int64_t* cv = C + (int64_t)v * K;
      if (v < N) {
        uint8_t code = leaves[(int64_t)v * S + s] & (uint8_t)((1u << K) - 1u);
        if (code == 0) code = (uint8_t)((1u << K) - 1u);
        for (int32_t k = 0; k < K; ++k) cv[k] = ((code >> k) & 1) ? 0 : INF;
        continue;
      }
      for (int32_t k = 0; k < K; ++k) cv[k] = 0;
      for (int c = 0; c < 2; ++c) {
        const int64_t* cc = C + (int64_t)ch[2 * (v - N) + c] * K;
        for (int32_t k = 0; k < K; ++k) {
          int64_t best = INF;
          for (int32_t j = 0; j < K; ++j) {
            if (cc[j] >= INF) continue;
            const int64_t x = cc[j] + tcm[j * K + k];
            if (x < best) best = x;
          }
          cv[k] += best;
        }
      }
    }
    const int64_t* cr = C + (int64_t)root * K;
    int64_t best = cr[0];
    for (int32_t k = 1; k < K; ++k)
      if (cr[k] < best) best = cr[k];
    if (site_cost) site_cost[s] = (uint32_t)best;
    total += (uint64_t)best * (w ? w[s] : 1u);
    free(C);
  }
  free(post);
  return total;
}

int64_t orc_fitch_brute(int32_t N, const int32_t* ch, int32_t root, const uint8_t* leaves,
                        int32_t K, uint8_t* mpr) {
  const int32_t nint = N - 1;
  int64_t total = 1;
  for (int32_t i = 0; i < nint; ++i) total *= K;
  int64_t best = INT64_MAX;
  int32_t st[16];
  uint8_t acc[16];
  memset(acc, 0, sizeof(acc));
  (void)root;
  for (int64_t code = 0; code < total; ++code) {
    int64_t c = code;
    for (int32_t i = 0; i < nint; ++i) {
      st[i] = (int32_t)(c % K);
      c /= K;
    }
    int64_t cost = 0;
    for (int32_t k = 0; k < nint; ++k) {
      const int32_t su = st[k];
      for (int q = 0; q < 2; ++q) {
        const int32_t x = ch[2 * k + q];
        if (x < N) cost += ((leaves[x] >> su) & 1) ? 0 : 1;
        else cost += (st[x - N] != su);
      }
    }
    if (cost < best) {
      best = cost;
      memset(acc, 0, sizeof(acc));
    }
    if (cost == best)
      for (int32_t i = 0; i < nint; ++i) acc[i] |= (uint8_t)(1u << st[i]);
  }
  if (mpr)
    for (int32_t i = 0; i < nint; ++i) mpr[N + i] = acc[i];
  return best;
}

/* ======================================================================== */
/* Model                                                                    */
/* ======================================================================== */

/* MlModel.m_gtr_independent (lib/mlModel.ml:390-421) followed by
 * m_meanrate (lib/mlModel.ml:204-213). */
void orc_gtr_q(int32_t n, const double* pi, const double* co, double* Q) {
  const int32_t m = n * (n - 1) / 2;
  double* c = (double*)malloc(sizeof(double) * m);
  for (int32_t k = 0; k < m; ++k) c[k] = (k == m - 1) ? 1.0 : co[k];
  memset(Q, 0, sizeof(double) * n * n);
  int32_t k = 0;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = i + 1; j < n; ++j) {
      Q[i * n + j] = c[k] * pi[j];
      Q[j * n + i] = c[k] * pi[i];
      ++k;
    }
  for (int32_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t j = 0; j < n; ++j)
      if (i != j) d += Q[i * n + j];
    Q[i * n + i] = -d;
  }
  double mr = 0.0;
  for (int32_t i = 0; i < n; ++i) mr += -Q[i * n + i] * pi[i];
  for (int32_t i = 0; i < n * n; ++i) Q[i] /= mr;
  free(c);
}

static void matmul(int32_t n, const double* A, const double* B, double* C) {
  for (int32_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < n; ++j) {
      double s = 0.0;
      for (int32_t k = 0; k < n; ++k) s += A[i * n + k] * B[k * n + j];
      C[i * n + j] = s;
    }
}

/* compose_gtr rules (lib/mlmodel.c:325-342) around exp(Qt) computed by
 * scaling and squaring of a degree-24 Taylor polynomial. */
void orc_compose(int32_t n, const double* Q, double t, double* P) {
  if (t == -1.0) {
    memcpy(P, Q, sizeof(double) * n * n);
    return;
  }
  if (!(t >= 1e-10)) {
    for (int32_t i = 0; i < n * n; ++i) P[i] = (i % (n + 1) == 0) ? 1.0 : 0.0;
    return;
  }
  double* A = (double*)malloc(sizeof(double) * n * n);
  double* T = (double*)malloc(sizeof(double) * n * n);
  double* X = (double*)malloc(sizeof(double) * n * n);
  double norm = 0.0;
  for (int32_t j = 0; j < n; ++j) {
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s += fabs(Q[i * n + j] * t);
    if (s > norm) norm = s;
  }
  int32_t sq = 0;
  while (norm > 0.25) {
    norm *= 0.5;
    sq++;
  }
  const double f = ldexp(t, -sq);
  for (int32_t i = 0; i < n * n; ++i) A[i] = Q[i] * f;
  /* P = I + A + A^2/2! + ... (Horner) */
  for (int32_t i = 0; i < n * n; ++i) P[i] = (i % (n + 1) == 0) ? 1.0 : 0.0;
  for (int32_t k = 24; k >= 1; --k) {
    matmul(n, A, P, T);
    for (int32_t i = 0; i < n * n; ++i) P[i] = ((i % (n + 1) == 0) ? 1.0 : 0.0) + T[i] / k;
  }
  for (int32_t s = 0; s < sq; ++s) {
    matmul(n, P, P, X);
    memcpy(P, X, sizeof(double) * n * n);
  }
  free(A);
  free(T);
  free(X);
}

/* ======================================================================== */
/* Felsenstein pruning                                                      */
/* ======================================================================== */

typedef struct {
  double s, c;
} neu_t; /* Neumaier compensated sum */
static inline void neu_add(neu_t* a, double x) {
  const double t = a->s + x;
  if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
  else a->c += (x - t) + a->s;
  a->s = t;
}

#define ORC_SCALE_BELOW 0x1p-256

/* P, P' = r Q P, P'' = r^2 Q^2 P for every (node, rate): [v][r][K][K] */
static void branch_mats(int32_t N, const double* blen, int32_t K, const double* Q, int32_t R,
                        const double* rates, double* P, double* P1, double* P2) {
  const int32_t M = 2 * N - 1;
  double* QP = (double*)malloc(sizeof(double) * K * K);
  double* QQP = (double*)malloc(sizeof(double) * K * K);
  for (int32_t v = 0; v < M; ++v)
    for (int32_t r = 0; r < R; ++r) {
      double* p = P + ((int64_t)v * R + r) * K * K;
      orc_compose(K, Q, rates[r] * blen[v], p);
      if (P1) {
        matmul(K, Q, p, QP);
        matmul(K, Q, QP, QQP);
        for (int32_t i = 0; i < K * K; ++i) {
          P1[((int64_t)v * R + r) * K * K + i] = rates[r] * QP[i];
          P2[((int64_t)v * R + r) * K * K + i] = rates[r] * rates[r] * QQP[i];
        }
      }
    }
  free(QP);
  free(QQP);
}

static inline void rescale(double* x, int32_t n, int32_t* e) {
  double mx = 0.0;
  for (int32_t i = 0; i < n; ++i)
    if (x[i] > mx) mx = x[i];
  if (mx > 0.0 && mx < ORC_SCALE_BELOW) {
    int ex;
    frexp(mx, &ex);
    for (int32_t i = 0; i < n; ++i) x[i] = ldexp(x[i], -ex);
    *e += ex;
  }
}

/* y = P x for a K x K row-major P */
static inline void matvec(int32_t K, const double* P, const double* x, double* y) {
  for (int32_t i = 0; i < K; ++i) {
    double s = 0.0;
    for (int32_t j = 0; j < K; ++j) s += P[i * K + j] * x[j];
    y[i] = s;
  }
}

double orc_lnl(int32_t N, const int32_t* ch, int32_t root, const double* blen, int32_t K,
               const double* pi, const double* Q, int32_t R, const double* rates,
               const double* probs, const uint32_t* tips, int64_t P, const uint32_t* weights,
               double* site_lnl, double* d1, double* d2, double* d1abs, double* d2abs,
               int32_t n_threads) {
  const int32_t M = 2 * N - 1;
  const int32_t RK = R * K;
  const int want_d = d1 != NULL && d2 != NULL;
  const int want_abs = want_d && d1abs != NULL && d2abs != NULL;
  int32_t* post = post_order(N, ch, root);
  int32_t* parent = (int32_t*)malloc(sizeof(int32_t) * M);
  for (int32_t v = 0; v < M; ++v) parent[v] = -1;
  for (int32_t k = 0; k < N - 1; ++k) {
    parent[ch[2 * k]] = N + k;
    parent[ch[2 * k + 1]] = N + k;
  }
  double* Pm = (double*)malloc(sizeof(double) * M * R * K * K);
  double* P1 = want_d ? (double*)malloc(sizeof(double) * M * R * K * K) : NULL;
  double* P2 = want_d ? (double*)malloc(sizeof(double) * M * R * K * K) : NULL;
  branch_mats(N, blen, K, Q, R, rates, Pm, P1, P2);
  neu_t total = {0.0, 0.0};
  double* slnl = (double*)malloc(sizeof(double) * P);
#ifdef _OPENMP
  if (n_threads < 1) n_threads = 1;
#else
  n_threads = 1;
#endif
  /* per-thread compensated derivative sums, merged in thread order: d1, d2
   * and (want_abs) the sums of their terms' magnitudes */
  neu_t* tdd = want_d ? (neu_t*)calloc((size_t)n_threads * 4 * M, sizeof(neu_t)) : NULL;
#ifdef _OPENMP
#pragma omp parallel num_threads(n_threads)
#endif
  {
#ifdef _OPENMP
    const int tno = omp_get_thread_num();
#else
    const int tno = 0;
#endif
    neu_t* my1 = want_d ? tdd + (size_t)tno * 4 * M : NULL;
    neu_t* my2 = want_d ? my1 + M : NULL;
    neu_t* my3 = want_d ? my1 + 2 * M : NULL;
    neu_t* my4 = want_d ? my1 + 3 * M : NULL;
    double* L = (double*)malloc(sizeof(double) * M * RK);
    int32_t* e = (int32_t*)malloc(sizeof(int32_t) * M);
    double* A = want_d ? (double*)malloc(sizeof(double) * M * RK) : NULL;
    double x[64], y[64];
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t p = 0; p < P; ++p) {
      for (int32_t q = 0; q < M; ++q) {
        const int32_t v = post[q];
        double* Lv = L + (int64_t)v * RK;
        if (v < N) {
          const uint32_t code = tips[(int64_t)v * P + p];
          for (int32_t r = 0; r < R; ++r)
            for (int32_t i = 0; i < K; ++i) Lv[r * K + i] = (double)((code >> i) & 1u);
          e[v] = 0;
          continue;
        }
        const int32_t a = ch[2 * (v - N)], b = ch[2 * (v - N) + 1];
        for (int32_t r = 0; r < R; ++r) {
          matvec(K, Pm + ((int64_t)a * R + r) * K * K, L + (int64_t)a * RK + r * K, x);
          matvec(K, Pm + ((int64_t)b * R + r) * K * K, L + (int64_t)b * RK + r * K, y);
          for (int32_t i = 0; i < K; ++i) Lv[r * K + i] = x[i] * y[i];
        }
        e[v] = e[a] + e[b];
        rescale(Lv, RK, &e[v]);
      }
      const double* Lr = L + (int64_t)root * RK;
      double lk = 0.0;
      for (int32_t r = 0; r < R; ++r) {
        double s = 0.0;
        for (int32_t i = 0; i < K; ++i) s += pi[i] * Lr[r * K + i];
        lk += probs[r] * s;
      }
      slnl[p] = log(lk) + (double)e[root] * M_LN2;
      if (!want_d) continue;
      /* up-partials in pre-order (reverse post-order) */
      for (int32_t q = M - 1; q >= 0; --q) {
        const int32_t c = post[q];
        if (c == root) continue;
        const int32_t pa = parent[c];
        const int32_t s = (ch[2 * (pa - N)] == c) ? ch[2 * (pa - N) + 1] : ch[2 * (pa - N)];
        double* Ac = A + (int64_t)c * RK;
        int32_t dummy = 0;
        for (int32_t r = 0; r < R; ++r) {
          matvec(K, Pm + ((int64_t)s * R + r) * K * K, L + (int64_t)s * RK + r * K, y);
          if (pa == root) {
            for (int32_t i = 0; i < K; ++i) Ac[r * K + i] = pi[i] * y[i];
          } else {
            const double* Ap = A + (int64_t)pa * RK + r * K;
            const double* Pp = Pm + ((int64_t)pa * R + r) * K * K;
            for (int32_t i = 0; i < K; ++i) {
              double m = 0.0;
              for (int32_t k = 0; k < K; ++k) m += Ap[k] * Pp[k * K + i];
              Ac[r * K + i] = m * y[i];
            }
          }
        }
        rescale(Ac, RK, &dummy);
        double l0 = 0.0, l1 = 0.0, l2 = 0.0;
        for (int32_t r = 0; r < R; ++r) {
          const double* Dc = L + (int64_t)c * RK + r * K;
          const double* Ar = Ac + r * K;
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
          matvec(K, Pm + ((int64_t)c * R + r) * K * K, Dc, x);
          for (int32_t i = 0; i < K; ++i) s0 += Ar[i] * x[i];
          matvec(K, P1 + ((int64_t)c * R + r) * K * K, Dc, x);
          for (int32_t i = 0; i < K; ++i) s1 += Ar[i] * x[i];
          matvec(K, P2 + ((int64_t)c * R + r) * K * K, Dc, x);
          for (int32_t i = 0; i < K; ++i) s2 += Ar[i] * x[i];
          l0 += probs[r] * s0;
          l1 += probs[r] * s1;
          l2 += probs[r] * s2;
        }
        const double g = l1 / l0;
        const double wp = weights ? (double)weights[p] : 1.0;
        const double h = l2 / l0 - g * g;
        neu_add(&my1[c], wp * g);
        neu_add(&my2[c], wp * h);
        neu_add(&my3[c], wp * fabs(g));
        neu_add(&my4[c], wp * (fabs(l2 / l0) + g * g));
      }
    }
    free(L);
    free(e);
    free(A);
  }
  for (int64_t p = 0; p < P; ++p) {
    const double wp = weights ? (double)weights[p] : 1.0;
    neu_add(&total, wp * slnl[p]);
    if (site_lnl) site_lnl[p] = slnl[p];
  }
  if (want_d) {
    const int32_t ra = ch[2 * (root - N)], rb = ch[2 * (root - N) + 1];
    for (int32_t c = 0; c < M; ++c) {
      neu_t a[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
      for (int t = 0; t < n_threads; ++t)
        for (int q = 0; q < 4; ++q) {
          neu_add(&a[q], tdd[(size_t)t * 4 * M + q * M + c].s);
          neu_add(&a[q], tdd[(size_t)t * 4 * M + q * M + c].c);
        }
      d1[c] = a[0].s + a[0].c;
      d2[c] = a[1].s + a[1].c;
      if (want_abs) {
        d1abs[c] = a[2].s + a[2].c;
        d2abs[c] = a[3].s + a[3].c;
      }
    }
    d1[root] = d2[root] = 0.0;
    d1[rb] = d1[ra]; /* one unrooted edge across the root */
    d2[rb] = d2[ra];
    if (want_abs) {
      d1abs[root] = d2abs[root] = 0.0;
      d1abs[rb] = d1abs[ra];
      d2abs[rb] = d2abs[ra];
    }
  }
  free(post);
  free(parent);
  free(Pm);
  free(P1);
  free(P2);
  free(tdd);
  free(slnl);
  return total.s + total.c;
}

double orc_lnl_brute(int32_t N, const int32_t* ch, int32_t root, const double* blen, int32_t K,
                     const double* pi, const double* Q, int32_t R, const double* rates,
                     const double* probs, const uint32_t* tips, int64_t P,
                     const uint32_t* weights) {
  const int32_t M = 2 * N - 1, nint = N - 1;
  double* Pm = (double*)malloc(sizeof(double) * M * R * K * K);
  branch_mats(N, blen, K, Q, R, rates, Pm, NULL, NULL);
  int64_t total = 1;
  for (int32_t i = 0; i < nint; ++i) total *= K;
  int32_t st[64];
  double lnl = 0.0;
  for (int64_t p = 0; p < P; ++p) {
    double lk = 0.0;
    for (int32_t r = 0; r < R; ++r) {
      double sr = 0.0;
      for (int64_t code = 0; code < total; ++code) {
        int64_t c = code;
        for (int32_t i = 0; i < nint; ++i) {
          st[i] = (int32_t)(c % K);
          c /= K;
        }
        double pr = pi[st[root - N]];
        for (int32_t k = 0; k < nint; ++k) {
          const int32_t su = st[k];
          for (int q = 0; q < 2; ++q) {
            const int32_t x = ch[2 * k + q];
            const double* Px = Pm + ((int64_t)x * R + r) * K * K + su * K;
            if (x < N) {
              const uint32_t code_t = tips[(int64_t)x * P + p];
              double s = 0.0;
              for (int32_t j = 0; j < K; ++j)
                if ((code_t >> j) & 1u) s += Px[j];
              pr *= s;
            } else {
              pr *= Px[st[x - N]];
            }
          }
        }
        sr += pr;
      }
      lk += probs[r] * sr;
    }
    lnl += (weights ? (double)weights[p] : 1.0) * log(lk);
  }
  free(Pm);
  return lnl;
}
