// Host-side graph tools (include/lvgraph.h): generators, batch tidy/apply,
// Matrix Market and temporal readers.  C++17 + OpenMP; results depend on
// the seed only (counter-based random streams), never on the thread count.
#include "lvgraph.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

struct lvg_graph {
  uint64_t S = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> keys;
  std::vector<float> w;
  std::vector<uint8_t> ex;  // empty = every id exists
  uint64_t order = 0;
};

struct lvg_batch {
  std::vector<uint32_t> u[2], v[2];
  std::vector<float> w[2];
};

namespace {

inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// k-th draw of stream s under seed: uniform in [0, 1)
inline double draw(uint64_t seed, uint64_t s, uint64_t k) {
  uint64_t x = mix64(mix64(seed ^ mix64(s + 0x632BE59BD9B4E019ull)) + k);
  return (double)(x >> 11) * 0x1.0p-53;
}

// CSR from directed triples; rows sorted by key, duplicates resolved by the
// last occurrence (index order), as LazyBitset's union keeps the last write.
lvg_graph* build_csr(uint64_t S, uint64_t n, const uint32_t* U, const uint32_t* V, const float* W,
                     const uint8_t* ex) {
  auto* g = new lvg_graph();
  g->S = S;
  std::vector<std::atomic<uint64_t>> cnt(S + 1);
  for (auto& c : cnt) c.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) cnt[U[i]].fetch_add(1, std::memory_order_relaxed);
  std::vector<uint64_t> start(S + 1, 0);
  for (uint64_t u = 0; u < S; ++u) start[u + 1] = start[u] + cnt[u].load(std::memory_order_relaxed);
  for (uint64_t u = 0; u < S; ++u) cnt[u].store(start[u], std::memory_order_relaxed);
  // (v << 32 | index) packed so a row sort orders by key then by input order
  bool wide = n >= (1ull << 32);
  std::vector<uint64_t> slot(n);
  std::vector<uint64_t> idx(wide ? n : 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    uint64_t p = cnt[U[i]].fetch_add(1, std::memory_order_relaxed);
    if (wide) {
      slot[p] = (uint64_t)V[i];
      idx[p] = (uint64_t)i;
    } else {
      slot[p] = ((uint64_t)V[i] << 32) | (uint64_t)i;
    }
  }
  std::vector<uint64_t> deg(S, 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)S; ++u) {
    uint64_t a = start[u], b = start[u + 1];
    if (wide) {
      std::vector<std::pair<uint64_t, uint64_t>> row;
      for (uint64_t k = a; k < b; ++k) row.push_back({slot[k], idx[k]});
      std::sort(row.begin(), row.end());
      uint64_t o = a;
      for (size_t k = 0; k < row.size(); ++k) {
        if (k + 1 < row.size() && row[k + 1].first == row[k].first) continue;  // keep the last
        slot[o] = row[k].first;
        idx[o] = row[k].second;
        ++o;
      }
      deg[u] = o - a;
    } else {
      std::sort(slot.begin() + a, slot.begin() + b);
      uint64_t o = a;
      for (uint64_t k = a; k < b; ++k) {
        if (k + 1 < b && (slot[k + 1] >> 32) == (slot[k] >> 32)) continue;  // keep the last
        slot[o++] = slot[k];
      }
      deg[u] = o - a;
    }
  }
  g->off.assign(S + 1, 0);
  for (uint64_t u = 0; u < S; ++u) g->off[u + 1] = g->off[u] + deg[u];
  g->keys.resize(g->off[S]);
  g->w.resize(g->off[S]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < (int64_t)S; ++u) {
    uint64_t a = start[u], o = g->off[u];
    for (uint64_t k = 0; k < deg[u]; ++k) {
      uint64_t s = slot[a + k];
      uint64_t i = wide ? idx[a + k] : (s & 0xFFFFFFFFull);
      g->keys[o + k] = wide ? (uint32_t)s : (uint32_t)(s >> 32);
      g->w[o + k] = W ? W[i] : 1.0f;
    }
  }
  if (ex) {
    g->ex.assign(ex, ex + S);
    uint64_t c = 0;
    for (uint64_t u = 0; u < S; ++u) c += g->ex[u] != 0;
    g->order = c;
  } else {
    g->order = S;
  }
  return g;
}

bool has_edge(const lvg_graph* g, uint32_t u, uint32_t v) {
  if (u >= g->S) return false;
  auto b = g->keys.begin() + g->off[u], e = g->keys.begin() + g->off[u + 1];
  return std::binary_search(b, e, v);
}

lvg_batch* tidy(const lvg_graph* g, std::vector<uint32_t> du, std::vector<uint32_t> dv, std::vector<float> dw,
                std::vector<uint32_t> iu, std::vector<uint32_t> iv, std::vector<float> iw) {
  auto* b = new lvg_batch();
  auto run = [&](int which, std::vector<uint32_t>& U, std::vector<uint32_t>& V, std::vector<float>& W, bool exist) {
    size_t n = U.size();
    std::vector<uint8_t> keep(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) keep[i] = has_edge(g, U[i], V[i]) == exist;
    std::vector<std::pair<uint64_t, float>> e;
    e.reserve(n);
    for (size_t i = 0; i < n; ++i)
      if (keep[i]) e.push_back({((uint64_t)U[i] << 32) | V[i], W[i]});
    std::stable_sort(e.begin(), e.end(), [](const auto& a, const auto& c) { return a.first < c.first; });
    auto it = std::unique(e.begin(), e.end(), [](const auto& a, const auto& c) { return a.first == c.first; });
    e.erase(it, e.end());
    b->u[which].resize(e.size());
    b->v[which].resize(e.size());
    b->w[which].resize(e.size());
    for (size_t i = 0; i < e.size(); ++i) {
      b->u[which][i] = (uint32_t)(e[i].first >> 32);
      b->v[which][i] = (uint32_t)e[i].first;
      b->w[which][i] = e[i].second;
    }
  };
  run(0, du, dv, dw, true);
  run(1, iu, iv, iw, false);
  return b;
}

uint64_t first_id(const lvg_graph* g) { return (!g->ex.empty() && g->S && !g->ex[0]) ? 1 : 0; }

}  // namespace

extern "C" {

lvg_graph* lvg_from_edges(uint64_t S, const uint32_t* u, const uint32_t* v, const float* w, uint64_t n,
                          const uint8_t* ex) {
  return build_csr(S, n, u, v, w, ex);
}

lvg_graph* lvg_rmat(int scale, int edge_factor, double a, double b, double c, uint64_t seed) {
  uint64_t S = 1ull << scale;
  uint64_t m = (uint64_t)edge_factor << scale;
  std::vector<uint32_t> U(2 * m), V(2 * m);
  double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m; ++e) {
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      double r = draw(seed, (uint64_t)e, (uint64_t)l);
      uint32_t ub = r >= ab, vb = (r >= a && r < ab) || r >= abc;
      u |= ub << l;
      v |= vb << l;
    }
    U[2 * e] = u;
    V[2 * e] = v;
    U[2 * e + 1] = v;
    V[2 * e + 1] = u;
  }
  return build_csr(S, 2 * m, U.data(), V.data(), nullptr, nullptr);
}

lvg_graph* lvg_planted_partition(uint64_t n, uint32_t blocks, double exponent, double avg_degree, double p_inter,
                                 uint64_t seed, int one_indexed) {
  if (blocks == 0 || n < 2ull * blocks) return nullptr;
  std::vector<uint64_t> size(blocks), start(blocks);
  double z = 0;
  for (uint32_t k = 0; k < blocks; ++k) z += 1.0 / std::pow(k + 1.0, exponent - 1.0);
  uint64_t tot = 0;
  for (uint32_t k = 0; k < blocks; ++k) {
    size[k] = std::max<uint64_t>(2, (uint64_t)std::floor(1.0 / std::pow(k + 1.0, exponent - 1.0) / z * n));
    tot += size[k];
  }
  if (tot > n) {  // trim the largest block so the sizes sum to n
    uint64_t over = tot - n;
    size[0] = size[0] > over + 2 ? size[0] - over : 2;
  } else {
    size[0] += n - tot;
  }
  start[0] = 0;
  for (uint32_t k = 1; k < blocks; ++k) start[k] = start[k - 1] + size[k - 1];
  uint64_t nn = start[blocks - 1] + size[blocks - 1];
  std::vector<uint32_t> block(nn);
  for (uint32_t k = 0; k < blocks; ++k)
    for (uint64_t i = 0; i < size[k]; ++i) block[start[k] + i] = k;
  uint64_t m = (uint64_t)(nn * avg_degree / 2);
  uint64_t base = one_indexed ? 1 : 0;
  std::vector<uint32_t> U(2 * m), V(2 * m);
  std::vector<uint8_t> ok(m);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)m; ++e) {
    uint64_t u = (uint64_t)(draw(seed, (uint64_t)e, 0) * nn);
    uint64_t v;
    if (draw(seed, (uint64_t)e, 1) < p_inter) v = (uint64_t)(draw(seed, (uint64_t)e, 2) * nn);
    else {
      uint32_t k = block[u];
      v = start[k] + (uint64_t)(draw(seed, (uint64_t)e, 2) * size[k]);
    }
    ok[e] = u != v;
    U[2 * e] = (uint32_t)(u + base);
    V[2 * e] = (uint32_t)(v + base);
    U[2 * e + 1] = (uint32_t)(v + base);
    V[2 * e + 1] = (uint32_t)(u + base);
  }
  // drop self-loops
  uint64_t k = 0;
  for (uint64_t e = 0; e < m; ++e) {
    if (!ok[e]) continue;
    U[2 * k] = U[2 * e];
    V[2 * k] = V[2 * e];
    U[2 * k + 1] = U[2 * e + 1];
    V[2 * k + 1] = V[2 * e + 1];
    ++k;
  }
  uint64_t S = nn + base;
  std::vector<uint8_t> ex;
  if (one_indexed) {
    ex.assign(S, 1);
    ex[0] = 0;
  }
  return build_csr(S, 2 * k, U.data(), V.data(), nullptr, one_indexed ? ex.data() : nullptr);
}

lvg_graph* lvg_read_mtx(const char* path) {
  std::ifstream f(path);
  if (!f) return nullptr;
  std::string line;
  if (!std::getline(f, line)) return nullptr;
  bool symmetric = line.find("symmetric") != std::string::npos;
  bool pattern = line.find("pattern") != std::string::npos;
  while (std::getline(f, line))
    if (!line.empty() && line[0] != '%') break;
  uint64_t rows = 0, cols = 0, nnz = 0;
  std::istringstream hs(line);
  if (!(hs >> rows >> cols >> nnz)) return nullptr;
  uint64_t S = std::max(rows, cols) + 1;
  std::vector<uint32_t> U, V;
  std::vector<float> W;
  U.reserve(symmetric ? 2 * nnz : nnz);
  for (uint64_t k = 0; k < nnz && std::getline(f, line); ++k) {
    std::istringstream ls(line);
    uint64_t i, j;
    double w = 1;
    if (!(ls >> i >> j)) break;
    if (!pattern) ls >> w;
    if (i >= S || j >= S) return nullptr;
    U.push_back((uint32_t)i); V.push_back((uint32_t)j); W.push_back((float)w);
    if (symmetric) { U.push_back((uint32_t)j); V.push_back((uint32_t)i); W.push_back((float)w); }
  }
  std::vector<uint8_t> ex(S, 1);
  ex[0] = 0;
  return build_csr(S, U.size(), U.data(), V.data(), W.data(), ex.data());
}

int lvg_write_mtx(const lvg_graph* g, const char* path) {
  FILE* f = std::fopen(path, "w");
  if (!f) return -1;
  bool one = first_id(g) == 1;
  uint64_t shift = one ? 0 : 1, rows = one ? g->S - 1 : g->S;
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%llu %llu %llu\n", (unsigned long long)rows,
               (unsigned long long)rows, (unsigned long long)g->off[g->S]);
  for (uint64_t u = 0; u < g->S; ++u)
    for (uint64_t i = g->off[u]; i < g->off[u + 1]; ++i)
      std::fprintf(f, "%llu %llu %.9g\n", (unsigned long long)(u + shift), (unsigned long long)(g->keys[i] + shift),
                   (double)g->w[i]);
  std::fclose(f);
  return 0;
}

lvg_graph* lvg_read_temporal(const char* path, uint64_t rows, uint64_t lines, int symmetric) {
  std::ifstream f(path);
  if (!f) return nullptr;
  uint64_t S = rows + 1;
  std::vector<uint32_t> U, V;
  std::string line;
  for (uint64_t k = 0; k < lines && std::getline(f, line); ++k) {
    std::istringstream ls(line);
    uint64_t u, v;
    if (!(ls >> u >> v)) break;
    if (u >= S || v >= S) S = std::max(u, v) + 1;
    U.push_back((uint32_t)u); V.push_back((uint32_t)v);
    if (symmetric) { U.push_back((uint32_t)v); V.push_back((uint32_t)u); }
  }
  std::vector<uint8_t> ex(S, 0);
  for (uint64_t u = 1; u <= rows && u < S; ++u) ex[u] = 1;
  for (size_t i = 0; i < U.size(); ++i) ex[U[i]] = ex[V[i]] = 1;
  return build_csr(S, U.size(), U.data(), V.data(), nullptr, ex.data());
}

void lvg_free(lvg_graph* g) { delete g; }
uint64_t lvg_span(const lvg_graph* g) { return g->S; }
uint64_t lvg_edges(const lvg_graph* g) { return g->off[g->S]; }
uint64_t lvg_order(const lvg_graph* g) { return g->order; }
const uint64_t* lvg_offsets(const lvg_graph* g) { return g->off.data(); }
const uint32_t* lvg_keys(const lvg_graph* g) { return g->keys.data(); }
const float* lvg_weights(const lvg_graph* g) { return g->w.data(); }
const uint8_t* lvg_exists(const lvg_graph* g) { return g->ex.empty() ? nullptr : g->ex.data(); }

lvg_batch* lvg_random_batch(const lvg_graph* g, uint64_t n_del, uint64_t n_ins, float w, uint64_t seed) {
  uint64_t lo = first_id(g), n = g->S - lo;
  std::vector<uint32_t> du, dv, iu(2 * n_ins), iv(2 * n_ins);
  std::vector<float> dw, iwv(2 * n_ins, w);
  for (uint64_t k = 0; k < n_del; ++k) {
    for (uint64_t r = 0; r < 5; ++r) {  // retry(..., 5), inc/batch.hxx:111
      uint64_t u = lo + (uint64_t)(draw(seed, 2 * k, 2 * r) * n);
      uint64_t d = g->off[u + 1] - g->off[u];
      if (d == 0) continue;
      uint64_t j = g->off[u] + (uint64_t)(draw(seed, 2 * k, 2 * r + 1) * d);
      uint32_t v = g->keys[j];
      du.push_back((uint32_t)u); dv.push_back(v); dw.push_back(g->w[j]);
      du.push_back(v); dv.push_back((uint32_t)u); dw.push_back(g->w[j]);
      break;
    }
  }
  for (uint64_t k = 0; k < n_ins; ++k) {
    uint32_t u = (uint32_t)(lo + (uint64_t)(draw(seed, 2 * k + 1, 0) * n));
    uint32_t v = (uint32_t)(lo + (uint64_t)(draw(seed, 2 * k + 1, 1) * n));
    iu[2 * k] = u; iv[2 * k] = v;
    iu[2 * k + 1] = v; iv[2 * k + 1] = u;
  }
  return tidy(g, du, dv, dw, iu, iv, iwv);
}

lvg_batch* lvg_tidy_batch(const lvg_graph* g, const uint32_t* du, const uint32_t* dv, const float* dw, uint64_t nd,
                          const uint32_t* iu, const uint32_t* iv, const float* iw, uint64_t ni) {
  return tidy(g, std::vector<uint32_t>(du, du + nd), std::vector<uint32_t>(dv, dv + nd),
              std::vector<float>(dw, dw + nd), std::vector<uint32_t>(iu, iu + ni), std::vector<uint32_t>(iv, iv + ni),
              std::vector<float>(iw, iw + ni));
}

void lvg_batch_free(lvg_batch* b) { delete b; }
uint64_t lvg_batch_deletions(const lvg_batch* b) { return b->u[0].size(); }
uint64_t lvg_batch_insertions(const lvg_batch* b) { return b->u[1].size(); }
const uint32_t* lvg_batch_u(const lvg_batch* b, int which) { return b->u[which & 1].data(); }
const uint32_t* lvg_batch_v(const lvg_batch* b, int which) { return b->v[which & 1].data(); }
const float* lvg_batch_w(const lvg_batch* b, int which) { return b->w[which & 1].data(); }

lvg_graph* lvg_apply_batch(const lvg_graph* g, const lvg_batch* b) {
  auto* y = new lvg_graph();
  uint64_t S = g->S;
  for (int k = 0; k < 2; ++k)
    for (size_t i = 0; i < b->u[k].size(); ++i) S = std::max<uint64_t>(S, std::max(b->u[k][i], b->v[k][i]) + 1);
  y->S = S;
  const auto& DU = b->u[0];
  const auto& IU = b->u[1];
  auto range = [&](const std::vector<uint32_t>& U, uint64_t u) {
    auto lo = std::lower_bound(U.begin(), U.end(), (uint32_t)u) - U.begin();
    auto hi = std::upper_bound(U.begin(), U.end(), (uint32_t)u) - U.begin();
    return std::make_pair((uint64_t)lo, (uint64_t)hi);
  };
  std::vector<uint64_t> deg(S);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < (int64_t)S; ++u) {
    uint64_t d0 = (uint64_t)u < g->S ? g->off[u + 1] - g->off[u] : 0;
    auto dr = range(DU, u), ir = range(IU, u);
    deg[u] = d0 - (dr.second - dr.first) + (ir.second - ir.first);
  }
  y->off.assign(S + 1, 0);
  for (uint64_t u = 0; u < S; ++u) y->off[u + 1] = y->off[u] + deg[u];
  y->keys.resize(y->off[S]);
  y->w.resize(y->off[S]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t u = 0; u < (int64_t)S; ++u) {
    uint64_t a = (uint64_t)u < g->S ? g->off[u] : 0, e = (uint64_t)u < g->S ? g->off[u + 1] : 0;
    auto dr = range(DU, u), ir = range(IU, u);
    uint64_t o = y->off[u], di = dr.first, ii = ir.first;
    for (uint64_t k = a; k < e || ii < ir.second;) {
      bool take_old = k < e && (ii >= ir.second || g->keys[k] < b->v[1][ii]);
      if (take_old) {
        while (di < dr.second && b->v[0][di] < g->keys[k]) ++di;
        bool gone = di < dr.second && b->v[0][di] == g->keys[k];
        if (!gone) { y->keys[o] = g->keys[k]; y->w[o] = g->w[k]; ++o; }
        ++k;
      } else {
        y->keys[o] = b->v[1][ii]; y->w[o] = b->w[1][ii]; ++o; ++ii;
      }
    }
  }
  if (!g->ex.empty() || S != g->S) {
    y->ex.assign(S, 0);
    for (uint64_t u = 0; u < S; ++u) y->ex[u] = u < g->S ? (g->ex.empty() ? 1 : g->ex[u]) : 0;
    for (size_t i = 0; i < IU.size(); ++i) y->ex[IU[i]] = y->ex[b->v[1][i]] = 1;
    uint64_t c = 0;
    for (uint64_t u = 0; u < S; ++u) c += y->ex[u] != 0;
    y->order = c;
  } else {
    y->order = S;
  }
  return y;
}

}  // extern "C"
