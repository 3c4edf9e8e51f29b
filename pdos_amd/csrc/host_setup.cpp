// host_setup.cpp -- problem setup (scaling, lambda, transpose) and the
// direct plugin's host factorization (ordering, LDL', level sets).
//
// Setup follows common.h:35-218 step for step in binary64, so the scaled data
// the GPU iterates on are bit-identical to the reference's.  Unlike the
// reference it never writes to the caller's arrays.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace pdos_amd {

namespace {
constexpr double kSqrtRatio = 1e3;  // common.h:16
constexpr double kRangeMax = 1e6;   // common.h:17
constexpr double kRangeMin = 1e-6;  // common.h:18

double norm2(const double* v, long n) {
  double a = 0.0;
  for (long i = 0; i < n; ++i) a += v[i] * v[i];
  return std::sqrt(a);
}

double clamp_rsqrt(double v) {
  v = std::fabs(v);
  v = v > kRangeMin ? v : kRangeMin;
  v = v > kRangeMax ? kRangeMax : v;
  return 1.0 / std::sqrt(v);
}

inline double vmax(double a, double b) { return a > b ? a : b; }

// Ruiz-type scaling of ruiz_equil_common.h:130-269 into S.D, S.E (before the
// x1e3), on a working copy of |A| and b in which every entry of a column
// inside a second-order cone takes the column's largest |value| in that cone
// at the cone's first row, and b's cone entries the cone's largest |b|
// (ruiz_equil_common.h:17-81; as there, the scan of a column's entries in a
// cone is bounded by nnz, not by the column, so the next column's leading
// entries in the same cone are rewritten too).  Three sweeps: row maxima pi
// (the last row: c), cone rows share their first row's pi, D = sqrt(D / pi);
// column maxima delta (the last column: b), E = sqrt(E / delta); a maximum
// <= 1e-6 resets the scale to 1.
void ruiz_scale(const Data* d, const Cone* k, Setup& S) {
  const long n = d->n, m = d->m, nz = d->Ap[n], top = k->f + k->l;
  std::vector<double> ax(d->Ax, d->Ax + nz), bc(d->b, d->b + m);
  std::vector<long> ai(d->Ai, d->Ai + nz);
  for (long col = 0; col < n && k->qsize > 0; ++col) {
    long r0 = top, p = d->Ap[col], cone = 0;
    while (p < d->Ap[col + 1] && d->Ai[p] < r0) ++p;
    while (p < d->Ap[col + 1] && cone < k->qsize) {  // rows out of order: stop at the last cone
      const long r1 = r0 + k->q[cone];
      double mx = 0.0;
      for (long t = p; t < nz && d->Ai[t] >= r0 && d->Ai[t] < r1; ++t) mx = vmax(mx, std::fabs(ax[t]));
      for (; p < nz && d->Ai[p] >= r0 && d->Ai[p] < r1; ++p) {
        ax[p] = mx;
        ai[p] = r0;
      }
      r0 = r1;
      ++cone;
    }
  }
  for (int c = 0; c < S.nsoc(); ++c) {
    double mx = 0.0;
    for (int i = S.soc_ptr[c]; i < S.soc_ptr[c + 1]; ++i) mx = vmax(mx, std::fabs(bc[i]));
    for (int i = S.soc_ptr[c]; i < S.soc_ptr[c + 1]; ++i) bc[i] = mx;
  }
  std::vector<double> pi(m, 0.0), delta(n, 0.0);
  std::fill(S.D.begin(), S.D.end(), 1.0);
  std::fill(S.E.begin(), S.E.end(), 1.0);
  double dl = 1.0, el = 1.0;
  for (int sweep = 0; sweep < 3; ++sweep) {
    double pil = 0.0, dlt = 0.0;
    for (long j = 0; j < n; ++j) {
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) pi[ai[p]] = vmax(pi[ai[p]], std::fabs(ax[p]) * S.E[j]);
      pil = vmax(pil, std::fabs(d->c[j]) * S.E[j]);
    }
    for (long i = 0; i < m; ++i) pi[i] = vmax(pi[i], std::fabs(bc[i]) * el);
    for (int c = 0; c < S.nsoc(); ++c)
      for (int i = S.soc_ptr[c]; i < S.soc_ptr[c + 1]; ++i) pi[i] = pi[S.soc_ptr[c]];
    for (long i = 0; i < m; ++i) {
      S.D[i] = std::fabs(pi[i]) > 1e-6 ? std::sqrt(S.D[i] / pi[i]) : 1.0;
      pi[i] = 0.0;
    }
    dl = std::fabs(pil) > 1e-6 ? std::sqrt(dl / pil) : 1.0;
    for (long j = 0; j < n; ++j) {
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) delta[j] = vmax(delta[j], std::fabs(ax[p]) * S.D[ai[p]]);
      delta[j] = vmax(delta[j], std::fabs(d->c[j]) * dl);
    }
    for (long i = 0; i < m; ++i) dlt = vmax(dlt, std::fabs(bc[i]) * S.D[i]);
    for (long j = 0; j < n; ++j) {
      S.E[j] = std::fabs(delta[j]) > 1e-6 ? std::sqrt(S.E[j] / delta[j]) : 1.0;
      delta[j] = 0.0;
    }
    el = std::fabs(dlt) > 1e-6 ? std::sqrt(el / dlt) : 1.0;
  }
}
}  // namespace

// Checks the caller's data and copies it into S unscaled (soc offsets, the CSC
// arrays, b, c and the warm start, zeros where absent).
static bool load_input(const Data* d, const Cone* k, Setup& S, std::string& err) {
  if (!d || !k) { err = "NULL data or cone"; return false; }
  if (!d->p) { err = "NULL params"; return false; }
  if (d->n < 1 || d->m < 1) { err = "n and m must be positive"; return false; }
  if (!d->Ap || !d->b || !d->c) { err = "NULL Ap, b or c"; return false; }
  const long n = d->n, m = d->m, nz = d->Ap[n];
  if (nz < 0 || nz > 0x7fffffffL || n > 0x3fffffffL || m > 0x3fffffffL) {
    err = "problem too large for 32-bit device indices"; return false;
  }
  if (nz > 0 && (!d->Ai || !d->Ax)) { err = "NULL Ai or Ax"; return false; }
  long tot = k->f + k->l;
  if (k->f < 0 || k->l < 0 || k->qsize < 0 || (k->qsize > 0 && !k->q)) { err = "bad cone"; return false; }
  for (long i = 0; i < k->qsize; ++i) {
    if (k->q[i] < 1) { err = "SOC size < 1"; return false; }
    tot += k->q[i];
  }
  if (tot != m) { err = "cone sizes do not sum to m"; return false; }
  for (long j = 0; j < n; ++j)
    if (d->Ap[j + 1] < d->Ap[j]) { err = "Ap not monotone"; return false; }
  for (long p = 0; p < nz; ++p)
    if (d->Ai[p] < 0 || d->Ai[p] >= m) { err = "row index out of range"; return false; }

  S.n = (int)n; S.m = (int)m; S.f = (int)k->f; S.l = (int)k->l;
  S.soc_ptr.assign(k->qsize + 1, 0);
  S.soc_ptr[0] = S.f + S.l;
  for (long i = 0; i < k->qsize; ++i) S.soc_ptr[i + 1] = S.soc_ptr[i] + (int)k->q[i];

  S.cp.resize(n + 1);
  S.ci.resize(nz);
  for (long j = 0; j <= n; ++j) S.cp[j] = (int)d->Ap[j];
  for (long p = 0; p < nz; ++p) S.ci[p] = (int)d->Ai[p];
  S.cx.assign(d->Ax, d->Ax + nz);
  S.b.assign(d->b, d->b + m);
  S.c.assign(d->c, d->c + n);
  S.x.assign(n, 0.0);
  S.s.assign(m, 0.0);
  S.y.assign(m, 0.0);
  if (d->x) std::memcpy(S.x.data(), d->x, sizeof(double) * n);
  if (d->s) std::memcpy(S.s.data(), d->s, sizeof(double) * m);
  if (d->y) std::memcpy(S.y.data(), d->y, sizeof(double) * m);
  return true;
}

bool build_setup(const Data* d, const Cone* k, Setup& S, std::string& err, int equil) {
  if (!load_input(d, k, S, err)) return false;
  const long n = d->n, m = d->m, nz = d->Ap[n];
  S.D.assign(m, 0.0);
  S.E.assign(n, 0.0);

  // stop-test scales use the caller's data (pdos.c:76-77)
  S.pscale = 1.0 + norm2(d->b, m);
  S.dscale = 1.0 + norm2(d->c, n);

  if (equil != 0 && equil != 1) { err = "unknown scaling (equil)"; return false; }
  if (d->p->NORMALIZE && equil == 1) {
    ruiz_scale(d, k, S);
    for (long j = 0; j < n; ++j)
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) S.cx[p] = S.D[d->Ai[p]] * d->Ax[p] * S.E[j];
    for (long j = 0; j < n; ++j) S.c[j] = d->c[j] * S.E[j];
    for (long i = 0; i < m; ++i) S.b[i] = d->b[i] * S.D[i];
    for (long i = 0; i < m; ++i) S.y[i] /= S.D[i];
    for (long j = 0; j < n; ++j) S.x[j] /= S.E[j];
    for (long i = 0; i < m; ++i) S.s[i] *= S.D[i];
  } else if (d->p->NORMALIZE) {
    for (long j = 0; j < n; ++j)
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) S.D[d->Ai[p]] += d->Ax[p] * d->Ax[p];
    for (int c = 0; c < S.nsoc(); ++c) {
      const int r0 = S.soc_ptr[c], q = S.soc_ptr[c + 1] - r0;
      double acc = 0.0;
      for (int i = 0; i < q; ++i) acc += S.D[r0 + i];
      for (int i = 0; i < q; ++i) S.D[r0 + i] = acc / (long)q;
    }
    for (long i = 0; i < m; ++i) S.D[i] = clamp_rsqrt(S.D[i]);
    for (long j = 0; j < n; ++j)
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) S.cx[p] *= S.D[d->Ai[p]];
    for (long j = 0; j < n; ++j)
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) S.E[j] += S.cx[p] * S.cx[p];
    for (long j = 0; j < n; ++j) S.E[j] = clamp_rsqrt(S.E[j]);
    for (long j = 0; j < n; ++j)
      for (long p = d->Ap[j]; p < d->Ap[j + 1]; ++p) S.cx[p] *= S.E[j];
    for (long j = 0; j < n; ++j) S.c[j] = d->c[j] * S.E[j];
    for (long i = 0; i < m; ++i) S.b[i] = d->b[i] * S.D[i];
    for (long i = 0; i < m; ++i) S.y[i] /= S.D[i];
    for (long j = 0; j < n; ++j) S.x[j] /= S.E[j];
    for (long i = 0; i < m; ++i) S.s[i] *= S.D[i];
  } else {
    std::fill(S.D.begin(), S.D.end(), 1.0);
    std::fill(S.E.begin(), S.E.end(), 1.0);
  }
  S.E0 = S.E;
  const double leps = equil == 1 ? 1e-6 : 1e-15;  // ruiz_equil_common.h:288, common.h:182
  S.lambda = (leps + norm2(S.b.data(), m)) / (leps + norm2(S.c.data(), n));
  for (long j = 0; j < n; ++j) {
    S.E[j] *= kSqrtRatio;
    S.c[j] *= kSqrtRatio;
    S.x[j] /= kSqrtRatio;
  }
  for (long p = 0; p < nz; ++p) S.cx[p] *= kSqrtRatio;

  // rows of A: stable counting sort of the CSC entries by row
  S.rp.assign(m + 1, 0);
  S.ri.resize(nz);
  S.rx.resize(nz);
  std::vector<int> next(m + 1, 0);
  for (long p = 0; p < nz; ++p) next[S.ci[p] + 1]++;
  for (long i = 0; i < m; ++i) next[i + 1] += next[i];
  std::copy(next.begin(), next.end(), S.rp.begin());
  for (long j = 0; j < n; ++j)
    for (int p = S.cp[j]; p < S.cp[j + 1]; ++p) {
      const int q = next[S.ci[p]]++;
      S.ri[q] = (int)j;
      S.rx[q] = S.cx[p];
    }
  return true;
}

bool build_setup_raw(const Data* d, const Cone* k, Setup& S, std::string& err) {
  if (!load_input(d, k, S, err)) return false;
  const long n = d->n, m = d->m, nz = d->Ap[n];
  S.dev = true;
  S.normalize = d->p->NORMALIZE ? 1 : 0;
  // the row pattern of A by the stable counting sort of build_setup, with the
  // CSC entry each row entry comes from
  S.rp.assign(m + 1, 0);
  S.ri.resize(nz);
  S.csr_src.resize(nz);
  std::vector<int> next(m + 1, 0);
  for (long p = 0; p < nz; ++p) next[S.ci[p] + 1]++;
  for (long i = 0; i < m; ++i) next[i + 1] += next[i];
  std::copy(next.begin(), next.end(), S.rp.begin());
  for (long j = 0; j < n; ++j)
    for (int p = S.cp[j]; p < S.cp[j + 1]; ++p) {
      const int q = next[S.ci[p]]++;
      S.ri[q] = (int)j;
      S.csr_src[q] = p;
    }
  return true;
}

// ---------------------------------------------------------------------------
// Direct plugin: KKT assembly, ordering, LDL', level sets.

namespace {

// Upper triangle of K = [[-I_n, A'], [A, I_m]] in CSC with the entries of each
// column in the order direct/private.c:77-120 produces them: column n+r holds
// A(r, j) for increasing j, then the +1 diagonal.
void form_kkt(const Setup& S, std::vector<int>& Kp, std::vector<int>& Ki, std::vector<double>& Kx) {
  const int n = S.n, m = S.m, N = n + m;
  Kp.assign(N + 1, 0);
  for (int j = 0; j < n; ++j) Kp[j + 1] = 1;
  for (int r = 0; r < m; ++r) Kp[n + r + 1] = (S.rp[r + 1] - S.rp[r]) + 1;
  for (int j = 0; j < N; ++j) Kp[j + 1] += Kp[j];
  Ki.resize(Kp[N]);
  Kx.resize(Kp[N]);
  for (int j = 0; j < n; ++j) { Ki[Kp[j]] = j; Kx[Kp[j]] = -1.0; }
  for (int r = 0; r < m; ++r) {
    int q = Kp[n + r];
    for (int p = S.rp[r]; p < S.rp[r + 1]; ++p) { Ki[q] = S.ri[p]; Kx[q] = S.rx[p]; ++q; }
    Ki[q] = n + r;
    Kx[q] = 1.0;
  }
}

}  // namespace

bool kkt_order(const Setup& S, std::vector<int>& perm, std::string& err) {
  std::vector<int> Kp, Ki;
  std::vector<double> Kx;
  form_kkt(S, Kp, Ki, Kx);
  return amd_order_cached(S.n + S.m, Kp, Ki, perm, err);
}

bool build_factor(const Setup& S, const std::vector<int>& perm_in, Factor& F, std::string& err) {
  const int N = S.n + S.m;
  F.N = N;
  std::vector<int> Kp, Ki;
  std::vector<double> Kx;
  form_kkt(S, Kp, Ki, Kx);
  if (!perm_in.empty()) {
    if ((int)perm_in.size() != N) { err = "permutation has wrong length"; return false; }
    F.P = perm_in;
  } else if (!amd_order_cached(N, Kp, Ki, F.P, err)) {
    return false;
  }
  std::vector<int> pinv(N, -1);
  for (int k = 0; k < N; ++k) {
    if (F.P[k] < 0 || F.P[k] >= N || pinv[F.P[k]] >= 0) { err = "invalid permutation"; return false; }
    pinv[F.P[k]] = k;
  }
  // M = lower triangle of P K P' by columns: entry (max(i2, j2), min(i2, j2))
  // of each stored K(i, j), i <= j, in column min(i2, j2).
  std::vector<int> Mp(N + 1, 0), Mi(Kp[N]);
  std::vector<double> Mx(Kp[N]);
  for (int j = 0; j < N; ++j)
    for (int p = Kp[j]; p < Kp[j + 1]; ++p) {
      if (Ki[p] > j) continue;
      Mp[std::min(pinv[Ki[p]], pinv[j]) + 1]++;
    }
  for (int j = 0; j < N; ++j) Mp[j + 1] += Mp[j];
  {
    std::vector<int> nxt(Mp.begin(), Mp.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int p = Kp[j]; p < Kp[j + 1]; ++p) {
        if (Ki[p] > j) continue;
        const int i2 = pinv[Ki[p]], j2 = pinv[j];
        const int q = nxt[std::min(i2, j2)]++;
        Mi[q] = std::max(i2, j2);
        Mx[q] = Kx[p];
      }
  }
  // Elimination tree (Liu's algorithm with path-compressed ancestors): row r
  // of M links each earlier column c it touches to r via c's current root.
  std::vector<int> parent(N, -1), anc(N, -1);
  std::vector<int> rp(N + 1, 0), rc(Mp[N]);  // M by rows: columns c <= r of row r
  {
    for (int p = 0; p < Mp[N]; ++p) rp[Mi[p] + 1]++;
    for (int r = 0; r < N; ++r) rp[r + 1] += rp[r];
    std::vector<int> nxt(rp.begin(), rp.end() - 1);
    for (int c = 0; c < N; ++c)
      for (int p = Mp[c]; p < Mp[c + 1]; ++p) rc[nxt[Mi[p]]++] = c;
    for (int r = 0; r < N; ++r)
      for (int q = rp[r]; q < rp[r + 1]; ++q) {
        int c = rc[q];
        if (c >= r) continue;
        while (c != -1 && c != r) {
          const int up = anc[c];
          anc[c] = r;
          if (up == -1) parent[c] = r;
          c = up;
        }
      }
  }
  // Column structures of L from row subtrees: row i of L has a nonzero in
  // every column on the tree paths from the columns k < i of row i of M up
  // to i (each path stops at a column already visited for this row).  One
  // pass counts, the second appends i to those columns; rows are taken in
  // increasing order, so every column's rows come out sorted.  The work is
  // nnz(M) + nnz(L).
  F.Lp.assign(N + 1, 0);
  {
    std::vector<int> mark(N, -1);
    auto walk = [&](int i, auto&& visit) {
      mark[i] = i;
      for (int q = rp[i]; q < rp[i + 1]; ++q)
        for (int k = rc[q]; k >= 0 && k < i && mark[k] != i; k = parent[k]) {
          mark[k] = i;
          visit(k);
        }
    };
    for (int i = 0; i < N; ++i) walk(i, [&](int k) { F.Lp[k + 1]++; });
    for (int j = 0; j < N; ++j) F.Lp[j + 1] += F.Lp[j];
    F.Li.resize(F.Lp[N]);
    std::vector<int> pos(F.Lp.begin(), F.Lp.end() - 1);
    std::fill(mark.begin(), mark.end(), -1);
    for (int i = 0; i < N; ++i) walk(i, [&](int k) { F.Li[pos[k]++] = i; });
  }
  const long lnz_total = F.Lp[N];
  F.Lx.assign(lnz_total, 0.0);
  F.Ld.assign(N, 0.0);
  // Left-looking numeric factorization: column j gathers M(:, j) and
  // subtracts L(:, k) d_k L(j, k) for every earlier column k with L(j, k) != 0;
  // the columns waiting for row r are kept in a linked list per row, each at
  // the position of its next row.
  std::vector<double> wv(N, 0.0);
  std::vector<int> head(N, -1), link(N, -1), at(N, 0);
  for (int j = 0; j < N; ++j) {
    for (int p = Mp[j]; p < Mp[j + 1]; ++p) wv[Mi[p]] += Mx[p];
    for (int k = head[j]; k != -1;) {
      const int knext = link[k];
      const int q = at[k];
      const double ljk = F.Lx[q], f = ljk * F.Ld[k];
      wv[j] -= ljk * f;
      const int qe = F.Lp[k + 1];
      for (int t = q + 1; t < qe; ++t) wv[F.Li[t]] -= F.Lx[t] * f;
      at[k] = q + 1;
      if (q + 1 < qe) {
        const int r = F.Li[q + 1];
        link[k] = head[r];
        head[r] = k;
      }
      k = knext;
    }
    head[j] = -1;
    const double dj = wv[j];
    wv[j] = 0.0;
    if (dj == 0.0 || !std::isfinite(dj)) { err = "zero pivot in LDL'"; return false; }
    F.Ld[j] = dj;
    for (int t = F.Lp[j]; t < F.Lp[j + 1]; ++t) {
      F.Lx[t] = wv[F.Li[t]] / dj;
      wv[F.Li[t]] = 0.0;
    }
    if (F.Lp[j] < F.Lp[j + 1]) {
      at[j] = F.Lp[j];
      const int r = F.Li[F.Lp[j]];
      link[j] = head[r];
      head[r] = j;
    }
  }
  // Rows of L (CSR) for the forward sweep, columns ascending within a row.
  const int lnz_tot = F.Lp[N];
  F.Rp.assign(N + 1, 0);
  F.Ri.resize(lnz_tot);
  F.Rx.resize(lnz_tot);
  for (int p = 0; p < lnz_tot; ++p) F.Rp[F.Li[p] + 1]++;
  for (int i = 0; i < N; ++i) F.Rp[i + 1] += F.Rp[i];
  {
    std::vector<int> nxt(F.Rp.begin(), F.Rp.end() - 1);
    for (int j = 0; j < N; ++j)
      for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) {
        const int q = nxt[F.Li[p]]++;
        F.Ri[q] = j;
        F.Rx[q] = F.Lx[p];
      }
  }
  // Level sets: forward level of row i = 1 + max level of the columns it
  // reads; backward level of column j = 1 + max level of the rows it reads.
  std::vector<int> lev(N, 0);
  int nlev = 0;
  for (int j = 0; j < N; ++j) {
    for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) lev[F.Li[p]] = std::max(lev[F.Li[p]], lev[j] + 1);
    nlev = std::max(nlev, lev[j] + 1);
  }
  auto bucket = [&](const std::vector<int>& L, int nl, std::vector<int>& ptr, std::vector<int>& rows) {
    ptr.assign(nl + 1, 0);
    for (int i = 0; i < N; ++i) ptr[L[i] + 1]++;
    for (int t = 0; t < nl; ++t) ptr[t + 1] += ptr[t];
    rows.resize(N);
    std::vector<int> nxt(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < N; ++i) rows[nxt[L[i]]++] = i;
  };
  bucket(lev, nlev, F.fl_ptr, F.fl_rows);
  std::fill(lev.begin(), lev.end(), 0);
  nlev = 0;
  for (int j = N - 1; j >= 0; --j) {
    int lv = 0;
    for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) lv = std::max(lv, lev[F.Li[p]] + 1);
    lev[j] = lv;
    nlev = std::max(nlev, lv + 1);
  }
  bucket(lev, nlev, F.bl_ptr, F.bl_rows);
  return true;
}

}  // namespace pdos_amd
