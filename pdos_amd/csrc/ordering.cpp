// ordering.cpp -- the direct plugin's fill-reducing ordering of the KKT
// matrix: approximate minimum degree on the quotient graph, after the
// published method of Amestoy, Davis and Duff ("An approximate minimum degree
// ordering algorithm", SIAM J. Matrix Anal. Appl. 17(4):886-905, 1996).
//
// The graph is kept as a quotient graph: an eliminated node becomes an
// element, the clique it created kept implicitly as the list of its
// variables.  Each step takes a principal variable of least approximate
// degree (ties to the lowest index), forms its element from the elements it
// touches (absorbing them) and its remaining variable neighbours, and updates
// the variables of the new element:
//   * w(e) = |L_e \ L_p| for every element e next to them (weighted sizes);
//   * the degree bound of the paper,
//       d_i <= min(n - k - |i|, d_i' + |L_p \ i|,
//                  |A_i \ i| + |L_p \ i| + sum_{e in E_i \ p} w(e));
//   * an element with w(e) = 0 lies inside L_p and is absorbed;
//   * variables with identical variable and element lists (found by a hash
//     and checked exactly) are merged into one supervariable and ordered
//     together.
// Nodes whose degree exceeds max(16, 6 sqrt(n)) are taken out of the graph
// and ordered last (the dense rows of a portfolio's budget and factor
// constraints).  Their adjacency is still counted exactly as bit masks (up to
// 4,096 such nodes): a variable's degree adds the dense nodes it reaches
// directly or through its elements, so a variable tied to many dense rows is
// not taken before its sparse neighbours.  Any ordering yields the same
// solution up to roundoff.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "host.hpp"

namespace pdos_amd {

// Nodes of degree above kDenseFactor sqrt(N) count as dense.  The published
// method's default is 10; 6 also takes out the hub rows of factor models
// (each tied to a few hundred assets), whose element lists otherwise make
// every elimination next to them rescan a few hundred elements: for the
// n_a = 500 portfolio (N = 1,523) the ordering takes 0.85 ms instead of 6 ms
// with the same fill and level counts.
constexpr double kDenseFactor = 6.0;

bool amd_order(int N, const std::vector<int>& Kp, const std::vector<int>& Ki, std::vector<int>& perm,
               std::string& err) {
  perm.clear();
  if (N <= 0) return true;
  // symmetric adjacency without the diagonal, deduplicated (each list sized
  // once: the lists are filled from counted degrees)
  std::vector<std::vector<int>> A(N);
  {
    std::vector<int> cnt(N, 0);
    for (int j = 0; j < N; ++j)
      for (int p = Kp[j]; p < Kp[j + 1]; ++p) {
        const int i = Ki[p];
        if (i < 0 || i >= N) { err = "ordering: index out of range"; return false; }
        if (i == j) continue;
        ++cnt[i];
        ++cnt[j];
      }
    for (int i = 0; i < N; ++i) A[i].reserve(cnt[i]);
  }
  for (int j = 0; j < N; ++j)
    for (int p = Kp[j]; p < Kp[j + 1]; ++p) {
      const int i = Ki[p];
      if (i == j) continue;
      A[i].push_back(j);
      A[j].push_back(i);
    }
  for (auto& a : A) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  enum : char { VAR = 0, ELEM = 1, GONE = 2 };
  std::vector<char> st(N, VAR);
  // dense nodes: out of the graph, ordered last (by degree, then index)
  const long dense_deg = std::max<long>(16, (long)(kDenseFactor * std::sqrt((double)N)));
  std::vector<int> dense;
  for (int i = 0; i < N; ++i)
    if ((long)A[i].size() > dense_deg) dense.push_back(i);
  if ((long)dense.size() == N) dense.clear();  // nothing sparse left: order everything normally
  std::vector<long> full_deg(N);
  for (int i = 0; i < N; ++i) full_deg[i] = (long)A[i].size();
  std::vector<char> is_dense(N, 0);
  for (int i : dense) is_dense[i] = 1;
  if (!dense.empty())
    for (int i = 0; i < N; ++i) {
      if (is_dense[i]) { A[i].clear(); continue; }
      auto& a = A[i];
      a.erase(std::remove_if(a.begin(), a.end(), [&](int j) { return is_dense[j] != 0; }), a.end());
    }
  for (int i : dense) st[i] = GONE;
  // dense adjacency as bit masks: W words per node (variables: own edges and
  // those inherited from absorbed elements; elements: their clique's)
  const int W = dense.empty() || dense.size() > 4096 ? 0 : (int)((dense.size() + 63) / 64);
  std::vector<uint64_t> msk((size_t)N * W, 0);
  if (W) {
    std::vector<int> didx(N, -1);
    for (size_t t = 0; t < dense.size(); ++t) didx[dense[t]] = (int)t;
    for (int j = 0; j < N; ++j)
      for (int p = Kp[j]; p < Kp[j + 1]; ++p) {
        const int i = Ki[p];
        if (i == j || is_dense[i] == is_dense[j]) continue;
        const int v = is_dense[i] ? j : i, dn = is_dense[i] ? didx[i] : didx[j];
        msk[(size_t)v * W + dn / 64] |= 1ull << (dn % 64);
      }
  }
  std::vector<uint64_t> acc(W);
  auto mask_or = [&](uint64_t* dst, const uint64_t* src) { for (int t = 0; t < W; ++t) dst[t] |= src[t]; };
  auto popc = [&](const uint64_t* a) { long c = 0; for (int t = 0; t < W; ++t) c += __builtin_popcountll(a[t]); return c; };

  std::vector<std::vector<int>> E(N), L(N);
  std::vector<long> nv(N, 1), deg(N, 0), sdeg(N, 0), esize(N, 0), w(N, 0), newdeg(N, 0), newsdeg(N, 0);
  std::vector<int> mark(N, 0), wmark(N, 0), chain_next(N, -1), chain_tail(N);
  std::vector<unsigned long> hsh(N, 0);
  for (int i = 0; i < N; ++i) chain_tail[i] = i;
  long nsparse = 0;
  std::set<std::pair<long, int>> Q;
  for (int i = 0; i < N; ++i)
    if (st[i] == VAR) {
      sdeg[i] = (long)A[i].size();
      deg[i] = sdeg[i] + (W ? popc(&msk[(size_t)i * W]) : 0);
      Q.insert({deg[i], i});
      ++nsparse;
    }
  int stamp = 0, wround = 0;
  long k = 0;  // weight eliminated so far
  perm.reserve(N);
  std::vector<int> Lp, cand;
  while (!Q.empty()) {
    const int p = Q.begin()->second;
    Q.erase(Q.begin());
    // the new element: variables of the elements p touches (absorbed) and
    // p's remaining variable neighbours
    ++stamp;
    mark[p] = stamp;
    Lp.clear();
    for (int e : E[p]) {
      if (st[e] != ELEM) continue;
      if (W) mask_or(&msk[(size_t)p * W], &msk[(size_t)e * W]);
      for (int i : L[e])
        if (st[i] == VAR && mark[i] != stamp) { mark[i] = stamp; Lp.push_back(i); }
      st[e] = GONE;
      std::vector<int>().swap(L[e]);
    }
    for (int i : A[p])
      if (st[i] == VAR && mark[i] != stamp) { mark[i] = stamp; Lp.push_back(i); }
    std::vector<int>().swap(A[p]);
    std::vector<int>().swap(E[p]);
    for (int v = p; v >= 0; v = chain_next[v]) perm.push_back(v);
    k += nv[p];
    st[p] = ELEM;
    long lpw = 0;
    for (int i : Lp) lpw += nv[i];
    esize[p] = lpw;
    L[p] = Lp;
    // w(e) = |L_e \ L_p| for the other elements next to L_p
    ++wround;
    for (int i : Lp)
      for (int e : E[i]) {
        if (st[e] != ELEM || e == p) continue;
        if (wmark[e] != wround) { wmark[e] = wround; w[e] = esize[e]; }
        w[e] -= nv[i];
      }
    const long rem = nsparse - k;
    for (int i : Lp) {
      auto& Ei = E[i];
      size_t o = 0;
      long ext = 0;
      unsigned long h = (unsigned long)p;
      for (int e : Ei) {
        if (st[e] != ELEM || e == p) continue;
        if (wmark[e] == wround && w[e] <= 0) {  // L_e inside L_p: absorbed into p
          if (W)  // its members keep the dense nodes its clique reached
            for (int v : L[e])
              if (st[v] == VAR) mask_or(&msk[(size_t)v * W], &msk[(size_t)e * W]);
          st[e] = GONE;
          std::vector<int>().swap(L[e]);
          continue;
        }
        Ei[o++] = e;
        ext += wmark[e] == wround ? w[e] : esize[e];
        h += (unsigned long)e;
      }
      Ei.resize(o);
      Ei.push_back(p);
      auto& Ai = A[i];
      o = 0;
      long aw = 0;
      for (int j : Ai) {
        if (st[j] != VAR || mark[j] == stamp || j == i) continue;
        Ai[o++] = j;
        aw += nv[j];
        h += (unsigned long)j;
      }
      Ai.resize(o);
      long d = aw + (lpw - nv[i]) + ext;
      d = std::min(d, sdeg[i] + (lpw - nv[i]));
      d = std::min(d, rem - nv[i]);
      newsdeg[i] = std::max<long>(d, 0);
      long dd = 0;
      if (W) {
        std::copy(&msk[(size_t)i * W], &msk[(size_t)i * W] + W, acc.begin());
        for (int e : Ei) mask_or(acc.data(), &msk[(size_t)e * W]);
        dd = popc(acc.data());
        for (int t = 0; t < W; ++t) h += acc[t] * 0x9e3779b97f4a7c15ull;
      }
      newdeg[i] = newsdeg[i] + dd;
      hsh[i] = h;
    }
    // supervariables: same variable and element lists
    cand.assign(Lp.begin(), Lp.end());
    std::sort(cand.begin(), cand.end(), [&](int a, int b) { return hsh[a] != hsh[b] ? hsh[a] < hsh[b] : a < b; });
    for (size_t s = 0; s < cand.size();) {
      size_t t = s;
      while (t < cand.size() && hsh[cand[t]] == hsh[cand[s]]) ++t;
      if (t - s > 1) {
        for (size_t a = s; a < t; ++a) {
          const int i = cand[a];
          if (st[i] != VAR) continue;
          std::sort(A[i].begin(), A[i].end());
          std::sort(E[i].begin(), E[i].end());
          for (size_t b = a + 1; b < t; ++b) {
            const int j = cand[b];
            if (st[j] != VAR) continue;
            std::sort(A[j].begin(), A[j].end());
            std::sort(E[j].begin(), E[j].end());
            if (A[i] != A[j] || E[i] != E[j]) continue;
            if (W && !std::equal(&msk[(size_t)i * W], &msk[(size_t)i * W] + W, &msk[(size_t)j * W])) continue;
            // j joins i: ordered right after i's members
            nv[i] += nv[j];
            newdeg[i] -= nv[j];
            newsdeg[i] -= nv[j];
            nv[j] = 0;
            st[j] = GONE;
            chain_next[chain_tail[i]] = j;
            chain_tail[i] = chain_tail[j];
            Q.erase({deg[j], j});
            std::vector<int>().swap(A[j]);
            std::vector<int>().swap(E[j]);
          }
        }
      }
      s = t;
    }
    for (int i : Lp) {
      if (st[i] != VAR) continue;
      Q.erase({deg[i], i});
      deg[i] = std::max<long>(newdeg[i], 0);
      sdeg[i] = std::max<long>(newsdeg[i], 0);
      Q.insert({deg[i], i});
    }
  }
  std::stable_sort(dense.begin(), dense.end(), [&](int a, int b) { return full_deg[a] < full_deg[b]; });
  for (int i : dense) perm.push_back(i);
  if ((int)perm.size() != N) { err = "ordering: internal error (" + std::to_string(perm.size()) + " of " +
                                     std::to_string(N) + " nodes ordered)"; return false; }
  return true;
}

// The ordering is a function of the pattern alone, so a process keeps the
// orderings of the last kOrderCacheEntries patterns (up to kOrderCacheNnz
// entries each) and returns a stored one when the same pattern comes again
// (parametric problems: a pdos() call per instance of a family; a batch
// member whose matrix group is new but whose pattern is not).  The pattern is
// compared in full on a hash match.
namespace {
constexpr size_t kOrderCacheEntries = 16;
constexpr size_t kOrderCacheNnz = 1 << 22;
struct OrderEntry {
  uint64_t hash;
  int N;
  std::vector<int> Kp, Ki, perm;
};
std::mutex g_order_mu;
std::list<OrderEntry> g_order_cache;  // most recently used first

uint64_t pattern_hash(int N, const std::vector<int>& Kp, const std::vector<int>& Ki) {
  uint64_t h = 1469598103934665603ull ^ (uint64_t)N;
  auto mix = [&](const std::vector<int>& v) {
    for (int x : v) h = (h ^ (uint32_t)x) * 1099511628211ull;
  };
  mix(Kp);
  mix(Ki);
  return h;
}
}  // namespace

bool amd_order_cached(int N, const std::vector<int>& Kp, const std::vector<int>& Ki, std::vector<int>& perm,
                      std::string& err) {
  const bool cacheable = Ki.size() <= kOrderCacheNnz;
  uint64_t h = 0;
  if (cacheable) {
    h = pattern_hash(N, Kp, Ki);
    std::lock_guard<std::mutex> lk(g_order_mu);
    for (auto it = g_order_cache.begin(); it != g_order_cache.end(); ++it)
      if (it->hash == h && it->N == N && it->Kp == Kp && it->Ki == Ki) {
        perm = it->perm;
        g_order_cache.splice(g_order_cache.begin(), g_order_cache, it);
        return true;
      }
  }
  if (!amd_order(N, Kp, Ki, perm, err)) return false;
  if (cacheable) {
    std::lock_guard<std::mutex> lk(g_order_mu);
    g_order_cache.push_front(OrderEntry{h, N, Kp, Ki, perm});
    if (g_order_cache.size() > kOrderCacheEntries) g_order_cache.pop_back();
  }
  return true;
}

}  // namespace pdos_amd
