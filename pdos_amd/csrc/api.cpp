// api.cpp -- the C ABI: pdos(), freeData(), freeSol() (pdos.h:85-87), the
// demo timers (util.h:37-47) and the extensions of pdos_ext.h.
//
// Flow of one call: host setup (common.h semantics, binary64) -> optional host
// LDL' factorization -> one packed upload into a per-thread device arena ->
// one engine launch on a per-thread stream -> one download -> host unscale
// (pdos.c:251-294) into a malloc'd Sol.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pdos.h"
#include "../../include/pdos_ext.h"
#include "device.hpp"
#include "engine.hpp"
#include "host.hpp"

#ifndef PDOS_DEFAULT_PLUGIN
#define PDOS_DEFAULT_PLUGIN PDOS_INDIRECT
#endif

namespace pdos_amd {

thread_local std::string g_err;
thread_local PdosInfo g_info;
thread_local bool g_info_valid = false;

Ctx& ctx() {
  thread_local Ctx c;
  return c;
}

bool Ctx::ensure(int device, std::string& err) {
  if (dev == device && stream) return true;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) { err = std::string("hipSetDevice: ") + hipGetErrorString(e); return false; }
  if (stream) (void)hipStreamDestroy(stream);
  e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  if (e != hipSuccess) { err = std::string("hipStreamCreate: ") + hipGetErrorString(e); return false; }
  if (!ev0 && hipEventCreate(&ev0) != hipSuccess) { err = "hipEventCreate failed"; return false; }
  if (!ev1 && hipEventCreate(&ev1) != hipSuccess) { err = "hipEventCreate failed"; return false; }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    max_lds = (int)prop.sharedMemPerBlock;
    num_cu = prop.multiProcessorCount;
  }
  if (dbuf) { (void)hipFree(dbuf); dbuf = nullptr; dcap = 0; }
  dev = device;
  return true;
}

bool Ctx::reserve(size_t dbytes, size_t hbytes, std::string& err) {
  if (dbytes > dcap) {
    if (dbuf) (void)hipFree(dbuf);
    dbuf = nullptr;
    size_t want = std::max(dbytes, dcap + dcap / 2);
    hipError_t e = hipMalloc(&dbuf, want);
    if (e != hipSuccess) {
      dcap = 0;
      err = std::string("hipMalloc: ") + hipGetErrorString(e);
      return false;
    }
    dcap = want;
  }
  if (hbytes > hcap) {
    if (hbuf) (void)hipHostFree(hbuf);
    hbuf = nullptr;
    size_t want = std::max(hbytes, hcap + hcap / 2);
    hipError_t e = hipHostMalloc(&hbuf, want, hipHostMallocDefault);
    if (e != hipSuccess) {
      hcap = 0;
      err = std::string("hipHostMalloc: ") + hipGetErrorString(e);
      return false;
    }
    hcap = want;
  }
  return true;
}

void Packer::grow(size_t need) {
  size_t cap = std::max<size_t>(cap_ ? cap_ : 4096, 4096);
  while (cap < need) cap *= 2;
  std::unique_ptr<unsigned char[]> m(new unsigned char[cap]);
  if (len_) std::memcpy(m.get(), mem_.get(), len_);
  mem_ = std::move(m);
  cap_ = cap;
}

size_t Packer::slot(size_t bytes) {
  const size_t off = (len_ + 255) & ~size_t(255);
  const size_t end = off + std::max<size_t>(bytes, 1);
  if (end > cap_) grow(end);
  if (off > len_) std::memset(mem_.get() + len_, 0, off - len_);  // alignment padding
  len_ = end;
  return off;
}

size_t Packer::put(const void* src, size_t bytes) {
  const size_t off = slot(bytes);
  if (bytes) std::memcpy(mem_.get() + off, src, bytes);
  return off;
}

size_t Packer::zeros(size_t bytes) {
  const size_t off = slot(bytes);
  std::memset(mem_.get() + off, 0, std::max<size_t>(bytes, 1));
  return off;
}

// Row blocks {r0, r1, p0, p1} (flattened) of a CSR matrix for the workgroup
// engine's SpMV: runs of whole rows of one class, at most kWgChunk entries
// and kWgRowsPerBlock rows per block.  Rows shorter than kWgLongRow entries
// form blocks {r0, r1, ..}; longer rows form blocks {r0, -r1, ..} (a wave
// per row); a row longer than kWgChunk is a block of its own.
std::vector<int> wg_row_blocks(const std::vector<int>& ptr) {
  std::vector<int> out;
  const int nrows = (int)ptr.size() - 1;
  auto len = [&](int r) { return ptr[r + 1] - ptr[r]; };
  int r = 0;
  while (r < nrows) {
    if (len(r) > kWgChunk) {
      out.insert(out.end(), {r, r + 1, ptr[r], ptr[r + 1]});
      ++r;
      continue;
    }
    const bool lng = len(r) >= kWgLongRow;
    int e = r;
    while (e < nrows && e - r < kWgRowsPerBlock && len(e) <= kWgChunk && (len(e) >= kWgLongRow) == lng &&
           ptr[e + 1] - ptr[r] <= kWgChunk)
      ++e;
    out.insert(out.end(), {r, lng ? -e : e, ptr[r], ptr[e]});
    r = e;
  }
  return out;
}

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <typename V>
size_t put_vec(Packer& P, const std::vector<V>& v) {
  return P.put(v.data(), v.size() * sizeof(V));
}

static size_t put_vals(Packer& P, const std::vector<double>& v, bool fp64) {
  if (fp64) return put_vec(P, v);
  std::vector<float> f(v.begin(), v.end());
  return put_vec(P, f);
}

WgBatch::~WgBatch() {
  if (own_mem && dbuf) (void)hipFree(dbuf);
  if (own_mem && ev0) (void)hipEventDestroy(ev0);
  if (own_mem && ev1) (void)hipEventDestroy(ev1);
  for (int c = 0; c < kClasses - 1; ++c) {
    if (xev[c]) (void)hipEventDestroy(xev[c]);
    if (xs[c]) (void)hipStreamDestroy(xs[c]);
  }
}

// Hash of a byte range, 8 bytes per step (multiply-xorshift mixing); only
// grouping uses it, and a hash match is confirmed by comparing the data.
static uint64_t fnv(uint64_t h, const void* p, size_t bytes) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  size_t i = 0;
  for (; i + 8 <= bytes; i += 8) {
    uint64_t w;
    std::memcpy(&w, c + i, 8);
    h = (h ^ w) * 0x9e3779b97f4a7c15ull;
    h ^= h >> 29;
  }
  for (; i < bytes; ++i) h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

template <typename V>
static uint64_t fnv_vec(uint64_t h, const std::vector<V>& v) {
  return fnv(h, v.data(), v.size() * sizeof(V));
}

// Runs body(i) for i in [0, count) on up to `host_threads()` threads, each
// taking a contiguous share; returns the lowest index whose body failed
// (with its error) or -1.
static int host_threads() {
  const char* e = std::getenv("PDOS_HOST_THREADS");
  int t = e ? std::atoi(e) : (int)std::thread::hardware_concurrency();
  return t < 1 ? 1 : (t > 64 ? 64 : t);
}

template <typename Body>
static int parallel_for(int count, Body body, std::string& err) {
  const int nt = std::min(host_threads(), std::max(1, count / 4));
  std::vector<int> bad(nt, -1);
  std::vector<std::string> msg(nt);
  auto run = [&](int t) {
    const int i0 = (int)((long)count * t / nt), i1 = (int)((long)count * (t + 1) / nt);
    for (int i = i0; i < i1; ++i)
      if (!body(i, msg[t])) { bad[t] = i; return; }
  };
  if (nt == 1) {
    run(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(run, t);
    run(0);
    for (auto& x : th) x.join();
  }
  for (int t = 0; t < nt; ++t)
    if (bad[t] >= 0) { err = msg[t]; return bad[t]; }
  return -1;
}

// Same scaled matrix: with host setup the scaled values and scalings match;
// with device setup (unscaled values) the values and NORMALIZE match, which
// fixes the same D, E and scaled values.
static bool same_matrix(const Setup& a, const Setup& b) {
  return a.n == b.n && a.m == b.m && a.f == b.f && a.l == b.l && a.soc_ptr == b.soc_ptr &&
         a.dev == b.dev && a.normalize == b.normalize && a.cp == b.cp && a.ci == b.ci && a.cx == b.cx &&
         a.D == b.D && a.E == b.E;
}

bool WgBatch::group_and_factor(std::string& err) {
  const int count = (int)S.size();
  grp.assign(count, -1);
  rep.clear();
  std::unordered_multimap<uint64_t, int> keys;  // hash -> group
  std::vector<uint64_t> hs(count);
  {
    std::string herr;
    parallel_for(count, [&](int i, std::string&) {
      const Setup& s = S[i];
      uint64_t h = 1469598103934665603ull;
      const int dims[6] = {s.n, s.m, s.f, s.l, s.dev ? 1 : 0, s.normalize};
      h = fnv(h, dims, sizeof(dims));
      h = fnv_vec(h, s.soc_ptr);
      h = fnv_vec(h, s.cp);
      h = fnv_vec(h, s.ci);
      h = fnv_vec(h, s.cx);
      h = fnv_vec(h, s.D);
      h = fnv_vec(h, s.E);
      hs[i] = h;
      return true;
    }, herr);
  }
  for (int i = 0; i < count; ++i) {
    const Setup& s = S[i];
    const uint64_t h = hs[i];
    auto range = keys.equal_range(h);
    for (auto it = range.first; it != range.second; ++it)
      if (same_matrix(S[rep[it->second]], s)) { grp[i] = it->second; break; }
    if (grp[i] < 0) {
      grp[i] = (int)rep.size();
      keys.emplace(h, grp[i]);
      rep.push_back(i);
    }
  }
  for (int i = 0; i < count; ++i) {
    if (rep[grp[i]] == i) continue;
    Setup& s = S[i];  // drop the duplicate matrix copies (the device shares the group's)
    std::vector<int>().swap(s.cp); std::vector<int>().swap(s.ci); std::vector<double>().swap(s.cx);
    std::vector<int>().swap(s.rp); std::vector<int>().swap(s.ri); std::vector<double>().swap(s.rx);
    std::vector<int>().swap(s.csr_src);
  }
  if (direct) {  // one factorization per distinct matrix, on the host threads
    F.assign(rep.size(), Factor());
    if (parallel_for((int)rep.size(), [&](int g, std::string& e) { return build_factor(S[rep[g]], {}, F[g], e); },
                     err) >= 0)
      return false;
  }
  return true;
}

// The LDS image of a group for the resident variant (k_wg_res): scalings,
// both CSR forms of A (and, direct plugin, A's column indices as pivot
// positions), the rows' wave batches, cone offsets and, direct plugin, the
// permutation, its inverse and both sweeps of L in level order.  Fills the
// layout, including the per-instance part after the image.  epl: entries per
// lane that the rows' wave batches aim for.
// With s.dev (setup on the device) the value slots and D, E are left zero:
// the device writes them, walking the rows' positions {row, e0, e1}.
static void res_image(const Setup& s, const Factor* f, bool fp64, int epl, std::vector<unsigned char>& img,
                      ResLayout& L) {
  img.clear();
  auto put = [&](const void* src, size_t bytes) {
    const size_t off = (img.size() + 15) & ~size_t(15);
    img.resize(off + bytes);
    if (bytes) std::memcpy(img.data() + off, src, bytes);
    return (int)off;
  };
  auto putv = [&](const std::vector<double>& v) {
    if (fp64) return put(v.data(), v.size() * 8);
    std::vector<float> w(v.begin(), v.end());
    return put(w.data(), w.size() * 4);
  };
  // items (rows) in class order and their wave batches: a row takes the
  // fewest lanes (a power of two, at most 64) that leave each lane at most
  // `epl` of its entries
  auto lanes = [epl](int len) {
    int lg = 0;
    while (lg < 6 && len > epl * (1 << lg)) ++lg;
    return lg;
  };
  auto batches = [&](const std::vector<int>& items, const std::vector<int>& len, std::vector<int>& ord,
                     std::vector<int4>& bat) {
    std::vector<int> o(items);
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return lanes(len[a]) < lanes(len[b]); });
    size_t k = 0;
    while (k < o.size()) {
      const int lg = lanes(len[o[k]]);
      size_t e = k;
      while (e < o.size() && (int)(e - k) < (64 >> lg) && lanes(len[o[e]]) == lg) ++e;
      bat.push_back(int4{(int)(ord.size() + k), (int)(e - k), lg, 0});
      k = e;
    }
    ord.insert(ord.end(), o.begin(), o.end());
  };
  // a CSR matrix with its rows in batch order: row ids, pointers, indices
  // (and, with `map`, the indices mapped), values
  auto positions = [&](const std::vector<int>& ord, const std::vector<int>& qp) {
    std::vector<int4> pos(ord.size());
    for (size_t q = 0; q < ord.size(); ++q) pos[q] = int4{ord[q], qp[q], qp[q + 1], 0};
    return put(pos.data(), pos.size() * sizeof(int4));
  };
  // zero_vals (device setup): the value slots stay zero
  auto spmv_batches = [&](const std::vector<int>& ptr, const std::vector<int>& ix, const std::vector<double>& vx,
                          const std::vector<int>* map, int& o_pos, int& o_idx, int& o_val, int* o_map,
                          int& o_bat, int& nbat, bool zero_vals) {
    const int nr = (int)ptr.size() - 1;
    std::vector<int> items(nr), len(nr), ord;
    std::vector<int4> bat;
    for (int r = 0; r < nr; ++r) { items[r] = r; len[r] = ptr[r + 1] - ptr[r]; }
    batches(items, len, ord, bat);
    std::vector<int> qp(nr + 1, 0), qi, qm;
    std::vector<double> qv;
    for (int q = 0; q < nr; ++q) {
      const int r = ord[q];
      qp[q + 1] = qp[q] + len[r];
      qi.insert(qi.end(), ix.begin() + ptr[r], ix.begin() + ptr[r + 1]);
      if (zero_vals) qv.resize(qv.size() + len[r], 0.0);
      else qv.insert(qv.end(), vx.begin() + ptr[r], vx.begin() + ptr[r + 1]);
      if (map)
        for (int e = ptr[r]; e < ptr[r + 1]; ++e) qm.push_back((*map)[ix[e]]);
    }
    o_pos = positions(ord, qp);
    qi.resize(qi.size() + 256, 0);  // padding read by the kernel's unconditional loads
    qv.resize(qv.size() + 256, 0.0);
    if (map) qm.resize(qm.size() + 256, 0);
    o_idx = put(qi.data(), qi.size() * 4);
    o_val = putv(qv);
    if (map) *o_map = put(qm.data(), qm.size() * 4);
    o_bat = put(bat.data(), bat.size() * sizeof(int4));
    nbat = (int)bat.size();
  };
  std::memset(&L, 0, sizeof(L));
  if (s.dev) {  // formed on the device
    const std::vector<double> zm(s.m, 0.0), zn(s.n, 0.0);
    L.D = put(zm.data(), zm.size() * 8);
    L.E = put(zn.data(), zn.size() * 8);
  } else {
    L.D = put(s.D.data(), s.D.size() * 8);
    L.E = put(s.E.data(), s.E.size() * 8);
  }

  std::vector<int> pinv;
  if (f) {
    pinv.resize(f->N);
    for (int k = 0; k < f->N; ++k) pinv[f->P[k]] = k;
  }
  spmv_batches(s.rp, s.ri, s.rx, f ? &pinv : nullptr, L.rpos, L.ridx, L.rval, &L.ridp, L.rbat, L.nrbat, s.dev);
  spmv_batches(s.cp, s.ci, s.cx, nullptr, L.cpos, L.cidx, L.cval, nullptr, L.cbat, L.ncbat, s.dev);
  L.soc = put(s.soc_ptr.data(), s.soc_ptr.size() * 4);
  {
    std::vector<int> big;
    for (int c = 0; c + 1 < (int)s.soc_ptr.size(); ++c)
      if (s.soc_ptr[c + 1] - s.soc_ptr[c] > kWgBigCone) { big.push_back(s.soc_ptr[c]); big.push_back(s.soc_ptr[c + 1]); }
    L.nbig = (int)big.size() / 2;
    if (big.empty()) big.push_back(0);
    L.big = put(big.data(), big.size() * 4);
  }
  L.mbat = std::max(L.nrbat, L.ncbat);
  if (f) {
    const int N = f->N;
    L.perm = put(f->P.data(), N * 4);
    L.pinv = put(pinv.data(), N * 4);
    // The dense trailing chain T: the forward levels from the last one with
    // more than one row on each hold a single row (a portfolio's budget and
    // factor rows); T's rows only depend on earlier rows and on each other,
    // and T's columns only have entries in T.  At most 64 rows.  Its solve,
    // z_T <- (L_T D_T L_T')^-1 z_T, is one dense product with that inverse,
    // formed here once per factorization (forward and back substitution on
    // the unit vectors, binary64): one wave applies it with every load in
    // flight at once, instead of 2 t dependent substitution steps (each a
    // cross-lane read, ~100 cycles) or 2 t barrier-separated levels.
    const int nfl = f->nfl();
    int lam0 = nfl;
    while (lam0 > 0 && f->fl_ptr[lam0] - f->fl_ptr[lam0 - 1] == 1 && nfl - (lam0 - 1) <= 64) --lam0;
    const int t = nfl - lam0 >= 3 ? nfl - lam0 : 0;
    if (!t) lam0 = nfl;
    std::vector<int> tpos(N, -1), trow(t);
    for (int k = 0; k < t; ++k) {
      trow[k] = f->fl_rows[f->fl_ptr[lam0 + k]];
      tpos[trow[k]] = k;
    }
    std::vector<double> tl((size_t)t * t, 0.0), tdq(t), tm((size_t)t * t), col(t);
    for (int k = 0; k < t; ++k) {
      tdq[k] = f->Ld[trow[k]];
      for (int p = f->Rp[trow[k]]; p < f->Rp[trow[k] + 1]; ++p)
        if (tpos[f->Ri[p]] >= 0) tl[(size_t)k * t + tpos[f->Ri[p]]] = f->Rx[p];
    }
    for (int c = 0; c < t; ++c) {  // column c of the inverse: L D L' col = e_c
      for (int k = 0; k < t; ++k) {
        double v = k == c ? 1.0 : 0.0;
        for (int j = 0; j < k; ++j) v -= tl[(size_t)k * t + j] * col[j];
        col[k] = v;
      }
      for (int k = 0; k < t; ++k) col[k] /= tdq[k];
      for (int k = t - 1; k >= 0; --k) {
        double v = col[k];
        for (int i = k + 1; i < t; ++i) v -= tl[(size_t)i * t + k] * col[i];
        col[k] = v;
      }
      for (int k = 0; k < t; ++k) tm[(size_t)k * t + c] = col[k];
    }
    L.t = t;
    L.trow = put(trow.data(), trow.size() * 4);
    L.tm = put(tm.data(), tm.size() * 8);
    // backward levels of the columns outside T, with T's values final
    std::vector<int> blev(N, -1);
    int nbl = 0;
    for (int q = N - 1; q >= 0; --q) {
      const int j = q;  // pivot order: columns after j are solved first
      if (tpos[j] >= 0) continue;
      int lv = 0;
      for (int p = f->Lp[j]; p < f->Lp[j + 1]; ++p)
        if (tpos[f->Li[p]] < 0) lv = std::max(lv, blev[f->Li[p]] + 1);
      blev[j] = lv;
      nbl = std::max(nbl, lv + 1);
    }
    for (int back = 0; back < 2; ++back) {
      const std::vector<int>& P = back ? f->Lp : f->Rp;
      const std::vector<int>& I = back ? f->Li : f->Ri;
      const std::vector<double>& V = back ? f->Lx : f->Rx;
      // levels: forward, the levels before T, then T's rows with their
      // entries outside T (one more level); backward, the columns outside T
      std::vector<std::vector<int>> levels;
      if (!back) {
        for (int lv = 0; lv < lam0; ++lv)
          levels.emplace_back(f->fl_rows.begin() + f->fl_ptr[lv], f->fl_rows.begin() + f->fl_ptr[lv + 1]);
        if (t) levels.push_back(trow);
      } else {
        levels.resize(nbl);
        for (int j = 0; j < N; ++j)
          if (blev[j] >= 0) levels[blev[j]].push_back(j);
      }
      std::vector<int> len(N, 0), order, lb(1, 0);
      std::vector<int4> bat;
      auto kept = [&](int r, int p) { return back || tpos[r] < 0 || tpos[I[p]] < 0; };
      for (int i = 0; i < N; ++i)
        for (int p = P[i]; p < P[i + 1]; ++p) len[i] += kept(i, p) ? 1 : 0;
      // forward: a row without entries (all of level 0) leaves z unchanged,
      // so it is left out, and a level without such rows with it
      std::vector<int> work;
      for (const auto& items : levels) {
        work.clear();
        for (int r : items)
          if (back || len[r] > 0) work.push_back(r);
        if (work.empty()) continue;
        batches(work, len, order, bat);
        L.mbat = std::max(L.mbat, (int)bat.size() - lb.back());
        lb.push_back((int)bat.size());
      }
      const int npos = (int)order.size();
      std::vector<int> ptr(npos + 1, 0), idx;
      std::vector<double> val, dq;
      for (int q = 0; q < npos; ++q) {
        const int r = order[q];
        ptr[q + 1] = ptr[q] + len[r];
        for (int p = P[r]; p < P[r + 1]; ++p)
          if (kept(r, p)) { idx.push_back(I[p]); val.push_back(V[p]); }
        if (back) dq.push_back(f->Ld[r]);
      }
      const int o_pos = positions(order, ptr);
      idx.resize(idx.size() + 256, 0);  // padding read by the kernel's unconditional loads
      val.resize(val.size() + 256, 0.0);
      const int o_idx = put(idx.data(), idx.size() * 4), o_val = put(val.data(), val.size() * 8);
      const int o_bat = put(bat.data(), bat.size() * sizeof(int4)), o_lb = put(lb.data(), lb.size() * 4);
      if (back) {
        L.bpos = o_pos; L.bidx = o_idx; L.bval = o_val; L.bbat = o_bat; L.blb = o_lb;
        L.bdq = put(dq.data(), std::max<size_t>(dq.size(), 1) * 8);
        L.nbl = (int)lb.size() - 1;
      } else {
        L.fpos = o_pos; L.fidx = o_idx; L.fval = o_val; L.fbat = o_bat; L.flb = o_lb;
        L.nfl = (int)lb.size() - 1;
      }
    }
  }
  img.resize((img.size() + 15) & ~size_t(15));
  L.img = (int)img.size();
  L.b = L.img;
  L.c = L.b + (int)(((long)s.m * 8 + 15) & ~15L);
  L.vec = L.c + (int)(((long)s.n * 8 + 15) & ~15L);
  L.total = L.vec + (int)res_vec_bytes(s.n, s.m, s.n + s.m, f != nullptr, fp64 ? 8 : 4);
}

bool WgBatch::prepare(int device_, bool fp64_, bool direct_, bool borrow, std::string& err) {
  device = device_; fp64 = fp64_; direct = direct_; own_mem = !borrow;
  Ctx& C = ctx();
  if (!C.ensure(device, err)) return false;
  stream = C.stream;
  const double t0 = now_ms();
  const int count = (int)S.size();
  const int tsz = fp64 ? 8 : 4;
  Packer P;
  struct Off { size_t soc, rp, ri, rx, cp, ci, cx, rb, cb, b, c, D, E, x0, s0, y0;
               int nrb, ncb;
               size_t perm, lrp, lri, lrx, lcp, lci, lcx, ld, flp, flr, blp, blr;
               size_t img; int res; ResLayout lay; bool stream;
               size_t dax, dcp, dci, drp, dsrc, dE0; };
  std::vector<Off> pk(count);
  xo.resize(count); so.resize(count); yo.resize(count);
  ho.assign(count, 0);
  if (grp.size() != (size_t)count) {  // no grouping requested: every instance is its own group
    grp.resize(count);
    rep.resize(count);
    for (int i = 0; i < count; ++i) grp[i] = rep[i] = i;
  }
  variant_env = wg_variant_env();
  persist_env = -1;
  if (const char* pe = std::getenv("PDOS_WG_PERSISTENT")) persist_env = std::atoi(pe) ? 1 : 0;
  threads_env = 0;
  if (const char* te = std::getenv("PDOS_WG_THREADS")) {
    const int t = std::atoi(te);
    threads_env = t == 256 || t == 512 || t == 1024 ? t : 0;
  }
  // the kernels' own LDS (the work-queue slot) comes first
  const int lds_cap = C.max_lds - 64;
  // members too large for one workgroup's LDS go to the whole-GPU engine
  is_big.assign(count, 0);
  big.clear();
  for (int i = 0; i < count; ++i)
    if (wg_lds_bytes(S[i].n, S[i].m, S[i].n + S[i].m, direct, tsz) > lds_cap) {
      if (S[i].dev) { err = "internal: an instance past the LDS was set up for the device"; return false; }
      is_big[i] = 1;
      big.push_back(i);
    }
  // LDS images of the groups (on the host threads), unless the streaming
  // variant is forced
  // per group: a member forced to the streaming variant (it needs the
  // streaming copies of the matrices); every member forced to it (no image)
  const int ngrp = (int)rep.size();
  std::vector<char> grp_stream(ngrp, variant_env == 0 ? 1 : 0), grp_all_stream(ngrp, 1);
  for (int i = 0; i < count; ++i) {
    const int vf = variant_env >= 0 ? variant_env : (var_force.empty() ? -1 : var_force[i]);
    if (vf == 0 || pack_all) grp_stream[grp[i]] = 1;
    if (vf != 0) grp_all_stream[grp[i]] = 0;
  }
  std::vector<std::vector<unsigned char>> imgs(count);
  std::vector<ResLayout> lays(count);
  {
    std::string ierr;
    parallel_for(count, [&](int i, std::string&) {
      if (!is_big[i] && rep[grp[i]] == i && !grp_all_stream[grp[i]])
        res_image(S[i], direct ? &F[grp[i]] : nullptr, fp64, epl.empty() ? 4 : epl[i], imgs[i], lays[i]);
      return true;
    }, ierr);
  }
  vlds.assign(count, VarLds{0, 0, 0, false});
  {  // the images and per-instance vectors, so the staging buffer grows at most once more
    size_t est = 0;
    for (int i = 0; i < count; ++i) est += imgs[i].size() + (size_t)(S[i].n + 2 * (size_t)S[i].m) * 64 + 4096;
    P.reserve(est);
  }
  std::vector<int> img_jobs;  // images copied into their slots by the host threads, below
  // the host arrays (alive until the end of prepare) likewise: a slot now,
  // the copy on the host threads with the images'
  struct CopyJob { size_t off; const void* src; size_t bytes; };
  std::vector<CopyJob> jobs;
  auto defer = [&](const auto& v) {
    const size_t bytes = v.size() * sizeof(v[0]);
    const size_t off = P.slot(bytes);
    if (bytes) jobs.push_back(CopyJob{off, v.data(), bytes});
    return off;
  };
  std::vector<int> dev_groups, dev_insts;  // representatives and instances set up on the device
  for (int i = 0; i < count; ++i) {
    const Setup& s = S[i];
    Off& k = pk[i];
    if (is_big[i]) continue;
    if (rep[grp[i]] != i && !is_big[rep[grp[i]]]) {  // shares the representative's matrix, scalings and factor
      const Off& r = pk[rep[grp[i]]];
      k = r;
    } else {
      // the image, kept when at least the variant with the image in device
      // memory fits (the variant is chosen per record afterwards)
      k.res = 0;
      k.lay = lays[i];
      if (!imgs[i].empty() && k.lay.total - k.lay.img <= lds_cap) {
        k.res = 1;
        k.img = P.slot(imgs[i].size());
        img_jobs.push_back(i);
      }
      // the streaming variant's copies of the matrices (and factor): only
      // where a record may use it (forced, or no image fits; with an image
      // the automatic choice never streams)
      k.stream = k.res == 0 || grp_stream[grp[i]];
      k.soc = defer(s.soc_ptr);
      // device setup: D, E and the values are written by k_setup_group
      // (batch_setup.hip) from the caller's values, uploaded below; D and E
      // get their slots after this loop, all in one range (one download)
      if (!s.dev) {
        k.D = defer(s.D);
        k.E = defer(s.E);
      }
      if (k.stream) {
        k.rp = defer(s.rp);
        k.ri = defer(s.ri);
        k.rx = s.dev ? P.zeros((size_t)s.nnz() * tsz) : put_vals(P, s.rx, fp64);
        k.cp = defer(s.cp);
        k.ci = defer(s.ci);
        k.cx = s.dev ? P.zeros((size_t)s.nnz() * tsz) : put_vals(P, s.cx, fp64);
        const std::vector<int> rb = wg_row_blocks(s.rp), cb = wg_row_blocks(s.cp);
        k.rb = put_vec(P, rb);
        k.cb = put_vec(P, cb);
        k.nrb = (int)rb.size() / 4;
        k.ncb = (int)cb.size() / 4;
      }
      if (s.dev) {
        k.dax = defer(s.cx);
        k.dcp = defer(s.cp);
        k.dci = k.res ? 0 : defer(s.ci);  // with an image its column row indices serve
        k.drp = defer(s.rp);
        k.dsrc = defer(s.csr_src);
        k.dE0 = P.zeros(sizeof(double) * s.n);
        dev_groups.push_back(i);
      }
    }
    vlds[i].x = wg_lds_bytes(s.n, s.m, s.n + s.m, direct, tsz);
    vlds[i].s = k.stream;
    if (k.res) {
      vlds[i].y = k.lay.total <= lds_cap ? k.lay.total : 0;
      vlds[i].z = k.lay.total - k.lay.img;
    }
    if (s.dev) dev_insts.push_back(i);
    k.b = defer(s.b);
    k.c = defer(s.c);
    k.x0 = defer(s.x);
    k.s0 = defer(s.s);
    k.y0 = defer(s.y);
    xo[i] = P.zeros(sizeof(double) * s.n);
    so[i] = P.zeros(sizeof(double) * s.m);
    yo[i] = P.zeros(sizeof(double) * s.m);
    if (count == 1) {  // the history and the result right after: fetch takes them in one copy
      if (want_hist && par[i].VERBOSE) ho[i] = P.zeros(sizeof(double) * (direct ? 5 : 25) * verbose_hist_rows(par[i].MAX_ITERS));
      res_off = P.zeros(sizeof(DevResult));
    }
    if (direct && rep[grp[i]] == i && k.stream) {
      const Factor& f = F[grp[i]];
      k.perm = defer(f.P);
      k.lrp = defer(f.Rp);
      k.lri = defer(f.Ri);
      k.lrx = defer(f.Rx);
      k.lcp = defer(f.Lp);
      k.lci = defer(f.Li);
      k.lcx = defer(f.Lx);
      k.ld = defer(f.Ld);
      k.flp = defer(f.fl_ptr);
      k.flr = defer(f.fl_rows);
      k.blp = defer(f.bl_ptr);
      k.blr = defer(f.bl_rows);
    }
  }
  // device setup: every group's D and E in one range, downloaded in one copy
  const size_t scales_lo = (P.size() + 255) & ~size_t(255);
  for (int i : dev_groups) {
    pk[i].D = P.zeros(sizeof(double) * S[i].m);
    pk[i].E = P.zeros(sizeof(double) * S[i].n);
  }
  const size_t scales_hi = P.size();
  for (int i : dev_insts)
    if (rep[grp[i]] != i) {
      pk[i].D = pk[rep[grp[i]]].D;
      pk[i].E = pk[rep[grp[i]]].E;
    }
  if (count != 1 || is_big[0]) {
    for (int i = 0; i < count; ++i)  // per row: 5 residuals, then (indirect) 10 x {CG steps, |q|}
      if (want_hist && par[i].VERBOSE && !is_big[i]) ho[i] = P.zeros(sizeof(double) * (direct ? 5 : 25) * verbose_hist_rows(par[i].MAX_ITERS));
    res_off = P.zeros(sizeof(DevResult) * count);
  }
  queue_off = P.zeros(sizeof(int) * kClasses);
  inst_off = P.zeros(sizeof(WgInst) * count);
  cinst_off = P.zeros(sizeof(WgInst) * count);
  const size_t sg_off = P.zeros(sizeof(SetupGroup) * dev_groups.size());
  const size_t si_off = P.zeros(sizeof(SetupInst) * dev_insts.size());
  const size_t sout_off = P.zeros(sizeof(double) * 3 * dev_insts.size());
  bytes = P.size();
  {
    // largest first, dealt round-robin into one contiguous run per thread
    // (parallel_for gives each thread a contiguous share), so the threads
    // copy about the same bytes
    for (int i : img_jobs) jobs.push_back(CopyJob{pk[i].img, imgs[i].data(), imgs[i].size()});
    std::stable_sort(jobs.begin(), jobs.end(), [](const CopyJob& a, const CopyJob& b) { return a.bytes > b.bytes; });
    const int nj = (int)jobs.size();
    const int nt = std::min(host_threads(), std::max(1, nj / 4));
    std::vector<CopyJob> dealt;
    dealt.reserve(nj);
    for (int t = 0; t < nt; ++t)
      for (int q = t; q < nj; q += nt) dealt.push_back(jobs[q]);
    std::string cerr;
    parallel_for(nj, [&](int q, std::string&) {
      const CopyJob& j = dealt[q];
      std::memcpy(P.data() + j.off, j.src, j.bytes);
      return true;
    }, cerr);
    for (int i : img_jobs) std::vector<unsigned char>().swap(imgs[i]);
  }
  if (borrow) {
    if (!C.reserve(bytes, bytes, err)) return false;
    dbuf = C.dbuf;
  } else {
    hipError_t e = hipMalloc(&dbuf, bytes);
    if (e != hipSuccess) { dbuf = nullptr; err = std::string("hipMalloc: ") + hipGetErrorString(e); return false; }
  }
  unsigned char* base = static_cast<unsigned char*>(dbuf);
  insts.assign(count, WgInst());
  for (int i = 0; i < count; ++i) {
    if (is_big[i]) continue;
    const Setup& s = S[i];
    const Off& k = pk[i];
    const Setup& sr = S[rep[grp[i]]];  // holds the matrix arrays
    WgInst& w = insts[i];
    std::memset(&w, 0, sizeof(w));
    w.n = s.n; w.m = s.m; w.f = s.f; w.l = s.l; w.nsoc = sr.nsoc(); w.N = s.n + s.m;
    w.nrb = k.nrb;
    w.ncb = k.ncb;
    w.max_iters = (int)std::min<long>(std::max<long>(par[i].MAX_ITERS, 0), 0x7fffffff);
    w.cg_max_its = (int)std::min<long>(std::max<long>(par[i].CG_MAX_ITS, 0), 0x7fffffff);
    w.eps = par[i].EPS_ABS;
    w.alpha = par[i].ALPHA;
    w.cg_tol_sq = par[i].CG_TOL * par[i].CG_TOL;
    w.lambda = s.lambda;
    w.pscale = s.pscale;
    w.dscale = s.dscale;
    w.soc_ptr = reinterpret_cast<const int*>(base + k.soc);
    if (k.stream) {  // else null: only the image variants run this record
      w.rp = reinterpret_cast<const int*>(base + k.rp);
      w.ri = reinterpret_cast<const int*>(base + k.ri);
      w.rx = base + k.rx;
      w.cp = reinterpret_cast<const int*>(base + k.cp);
      w.ci = reinterpret_cast<const int*>(base + k.ci);
      w.cx = base + k.cx;
      w.rblk = reinterpret_cast<const int*>(base + k.rb);
      w.cblk = reinterpret_cast<const int*>(base + k.cb);
    }
    w.b = reinterpret_cast<const double*>(base + k.b);
    w.c = reinterpret_cast<const double*>(base + k.c);
    w.D = reinterpret_cast<const double*>(base + k.D);
    w.E = reinterpret_cast<const double*>(base + k.E);
    w.x0 = reinterpret_cast<const double*>(base + k.x0);
    w.s0 = reinterpret_cast<const double*>(base + k.s0);
    w.y0 = reinterpret_cast<const double*>(base + k.y0);
    w.x = reinterpret_cast<double*>(base + xo[i]);
    w.s = reinterpret_cast<double*>(base + so[i]);
    w.y = reinterpret_cast<double*>(base + yo[i]);
    if (direct && k.stream) {
      const Factor& f = F[grp[i]];
      w.perm = reinterpret_cast<const int*>(base + k.perm);
      w.lrp = reinterpret_cast<const int*>(base + k.lrp);
      w.lri = reinterpret_cast<const int*>(base + k.lri);
      w.lrx = reinterpret_cast<const double*>(base + k.lrx);
      w.lcp = reinterpret_cast<const int*>(base + k.lcp);
      w.lci = reinterpret_cast<const int*>(base + k.lci);
      w.lcx = reinterpret_cast<const double*>(base + k.lcx);
      w.ld = reinterpret_cast<const double*>(base + k.ld);
      w.fl_ptr = reinterpret_cast<const int*>(base + k.flp);
      w.fl_rows = reinterpret_cast<const int*>(base + k.flr);
      w.bl_ptr = reinterpret_cast<const int*>(base + k.blp);
      w.bl_rows = reinterpret_cast<const int*>(base + k.blr);
      w.nfl = f.nfl();
      w.nbl = f.nbl();
    }
    w.res = reinterpret_cast<DevResult*>(base + res_off) + i;
    w.hist_rows = want_hist && par[i].VERBOSE ? verbose_hist_rows(par[i].MAX_ITERS) : 0;
    w.hist = w.hist_rows ? reinterpret_cast<double*>(base + ho[i]) : nullptr;
    w.cg_hist = w.hist_rows && !direct ? w.hist + (size_t)5 * w.hist_rows : nullptr;
    if (k.res) {
      w.img = base + k.img;
      w.lay = k.lay;
    }
  }
  set_shapes();
  std::memcpy(P.data() + inst_off, insts.data(), sizeof(WgInst) * count);
  for (size_t k = 0; k < order.size(); ++k)
    std::memcpy(P.data() + cinst_off + sizeof(WgInst) * k, &insts[order[k]], sizeof(WgInst));
  // device setup records: per representative its arrays and where D, E and
  // the values go (the image, and the streaming copies where packed); per
  // instance its vectors, its group's D and E0 and its two records
  if (!dev_groups.empty()) {
    std::vector<SetupGroup> G(dev_groups.size());
    for (size_t q = 0; q < dev_groups.size(); ++q) {
      const int i = dev_groups[q];
      const Setup& s = S[i];
      const Off& k = pk[i];
      SetupGroup& g = G[q];
      std::memset(&g, 0, sizeof(g));
      g.n = s.n; g.m = s.m; g.nnz = (int)s.nnz(); g.normalize = s.normalize; g.nsoc = s.nsoc();
      g.cp = reinterpret_cast<const int*>(base + k.dcp);
      g.ci = k.res ? nullptr : reinterpret_cast<const int*>(base + k.dci);
      g.ax = reinterpret_cast<double*>(base + k.dax);
      g.rp = reinterpret_cast<const int*>(base + k.drp);
      g.csr_src = reinterpret_cast<const int*>(base + k.dsrc);
      g.soc = reinterpret_cast<const int*>(base + k.soc);
      g.E0 = reinterpret_cast<double*>(base + k.dE0);
      if (k.res) {
        g.D = reinterpret_cast<double*>(base + k.img + k.lay.D);
        g.E = reinterpret_cast<double*>(base + k.img + k.lay.E);
        g.D2 = reinterpret_cast<double*>(base + k.D);
        g.E2 = reinterpret_cast<double*>(base + k.E);
        g.rval = base + k.img + k.lay.rval;
        g.cval = base + k.img + k.lay.cval;
        g.rpos = reinterpret_cast<const int4*>(base + k.img + k.lay.rpos);
        g.cpos = reinterpret_cast<const int4*>(base + k.img + k.lay.cpos);
        g.cidx = reinterpret_cast<const int*>(base + k.img + k.lay.cidx);
      } else {
        g.D = reinterpret_cast<double*>(base + k.D);
        g.E = reinterpret_cast<double*>(base + k.E);
      }
      if (k.stream) {
        g.srx = base + k.rx;
        g.scx = base + k.cx;
      }
    }
    std::memcpy(P.data() + sg_off, G.data(), sizeof(SetupGroup) * G.size());
  }
  if (!dev_insts.empty()) {
    std::vector<int> cls_pos(count, 0);
    for (size_t q = 0; q < order.size(); ++q) cls_pos[order[q]] = (int)q;
    std::vector<SetupInst> I(dev_insts.size());
    for (size_t q = 0; q < dev_insts.size(); ++q) {
      const int i = dev_insts[q];
      const Off& k = pk[i];
      const Off& kr = pk[rep[grp[i]]];
      SetupInst& t = I[q];
      std::memset(&t, 0, sizeof(t));
      t.n = S[i].n; t.m = S[i].m; t.rec = i; t.cls = cls_pos[i];
      t.b = reinterpret_cast<double*>(base + k.b);
      t.c = reinterpret_cast<double*>(base + k.c);
      t.x = reinterpret_cast<double*>(base + k.x0);
      t.s = reinterpret_cast<double*>(base + k.s0);
      t.y = reinterpret_cast<double*>(base + k.y0);
      t.D = reinterpret_cast<const double*>(base + kr.D);
      t.E0 = reinterpret_cast<const double*>(base + kr.dE0);
      t.out = reinterpret_cast<double*>(base + sout_off) + 3 * q;
    }
    std::memcpy(P.data() + si_off, I.data(), sizeof(SetupInst) * I.size());
  }
  hipError_t e;
  if (borrow) {
    // from the context's pinned buffer, which no other work of this thread
    // touches before this call's fetch has synchronized the stream: the
    // launch follows the copy on the same stream without a host wait
    std::memcpy(C.hbuf, P.data(), bytes);
    e = hipMemcpyAsync(dbuf, C.hbuf, bytes, hipMemcpyHostToDevice, stream);
    ev0 = C.ev0;
    ev1 = C.ev1;
  } else {
    e = hipMemcpyAsync(dbuf, P.data(), bytes, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess && !ev0) e = hipEventCreate(&ev0);
    if (e == hipSuccess && !ev1) e = hipEventCreate(&ev1);
  }
  if (e != hipSuccess) { err = std::string("upload: ") + hipGetErrorString(e); return false; }
  if (!dev_groups.empty() || !dev_insts.empty()) {
    // the setup kernels, then D, E (with the x1e3) per group and lambda and
    // the stop-test scales per instance back to the host records (the
    // unscaling at fetch and later uploads of the records use them)
    e = batch_setup_launch(reinterpret_cast<const SetupGroup*>(base + sg_off), (int)dev_groups.size(),
                           reinterpret_cast<const SetupInst*>(base + si_off), (int)dev_insts.size(),
                           reinterpret_cast<WgInst*>(base + inst_off), reinterpret_cast<WgInst*>(base + cinst_off),
                           fp64, stream);
    std::vector<double> outs(3 * dev_insts.size());
    std::vector<unsigned char> sc(scales_hi - scales_lo);
    if (e == hipSuccess && !sc.empty())
      e = hipMemcpyAsync(sc.data(), base + scales_lo, sc.size(), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess && !outs.empty())
      e = hipMemcpyAsync(outs.data(), base + sout_off, sizeof(double) * outs.size(), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) { err = std::string("device setup: ") + hipGetErrorString(e); return false; }
    for (int i : dev_groups) {
      Setup& s = S[i];
      const double* d = reinterpret_cast<const double*>(sc.data() + (pk[i].D - scales_lo));
      const double* ee = reinterpret_cast<const double*>(sc.data() + (pk[i].E - scales_lo));
      s.D.assign(d, d + s.m);
      s.E.assign(ee, ee + s.n);
    }
    for (size_t q = 0; q < dev_insts.size(); ++q) {
      const int i = dev_insts[q];
      Setup& s = S[i];
      s.lambda = outs[3 * q];
      s.pscale = outs[3 * q + 1];
      s.dscale = outs[3 * q + 2];
      insts[i].lambda = s.lambda;
      insts[i].pscale = s.pscale;
      insts[i].dscale = s.dscale;
      if (rep[grp[i]] != i) {
        s.D = S[rep[grp[i]]].D;
        s.E = S[rep[grp[i]]].E;
      }
    }
  }
  gsol.clear();
  for (int i : big) {  // own device memory: the batch may hold the thread's arena
    Setup full;
    const Setup* sp = &S[i];
    if (rep[grp[i]] != i) {  // the group's matrix arrays live with its representative
      full = S[i];
      const Setup& r = S[rep[grp[i]]];
      full.cp = r.cp; full.ci = r.ci; full.cx = r.cx; full.rp = r.rp; full.ri = r.ri; full.rx = r.rx;
      sp = &full;
    }
    std::unique_ptr<GridSolver> g = grid_prepare(*sp, par[i], fp64, direct ? &F[grp[i]] : nullptr, device, false, err);
    if (!g) { err = "instance " + std::to_string(i) + " (whole-GPU engine): " + err; return false; }
    gsol.push_back(std::move(g));
  }
  upload_ms = now_ms() - t0;
  return true;
}

// -1 automatic; PDOS_WG_VARIANT = stream | resident | image forces 0, 1, 2
// (where the variant fits), as does PDOS_WG_RESIDENT=0 the streaming one.
int wg_variant_env() {
  int v = -1;
  if (const char* e = std::getenv("PDOS_WG_VARIANT")) {
    if (!std::strcmp(e, "stream")) v = 0;
    else if (!std::strcmp(e, "resident")) v = 1;
    else if (!std::strcmp(e, "image")) v = 2;
  }
  if (const char* rs = std::getenv("PDOS_WG_RESIDENT"))
    if (!std::strcmp(rs, "0")) v = 0;
  return v;
}

// Kernel variant and launch mode of every record, then the launch classes.
// Variant: forced (environment, else the tuner's choice) where it fits (a
// resident choice that does not fit takes the image in device memory);
// automatic: the resident variant when the image and iterates fit in LDS
// without fewer workgroups per CU than the streaming variant (or the launch
// has at most one workgroup per CU anyway), else the image in device memory
// where its iterates fit, else the streaming variant.  Classes group records
// by (workgroups per CU that the LDS and the waves allow, threads per
// workgroup, launch mode: PDOS_WG_PERSISTENT = 1 / 0 forces the work-queue
// launch on or off), biggest
// first, and within a class the instances with the most nonzeros first.
void WgBatch::set_shapes() {
  const int count = (int)insts.size();
  const long cu_lds = 160 * 1024;
  auto wpc = [&](long b) { return (int)std::max<long>(1, std::min<long>(kOccMax, cu_lds / b)); };
  const int nwg_total = count - (int)big.size();
  for (int i = 0; i < count; ++i) {
    if (is_big[i]) continue;
    WgInst& w = insts[i];
    const VarLds& L = vlds[i];
    const int vf = variant_env >= 0 ? variant_env : (var_force.empty() ? -1 : var_force[i]);
    int v = 0;
    if (vf == 1) v = L.y ? 1 : (L.z ? 2 : 0);
    else if (vf == 2) v = L.z ? 2 : 0;
    else if (vf < 0) v = L.y && (nwg_total <= ctx().num_cu || wpc(L.y) >= wpc(L.x)) ? 1 : (L.z ? 2 : 0);
    else v = L.s ? 0 : (L.y ? 1 : 2);  // forced streaming where its copies were packed
    // (v == 0 only with L.s: without an image the copies are always packed)
    w.resident = v;
    w.lds_bytes = v == 1 ? L.y : (v == 2 ? L.z : L.x);
  }
  // threads per workgroup: 256 for the streaming variant; the image
  // variants 256, 512 or 1,024 (PDOS_WG_THREADS, else the tuner's choice,
  // else: 512 when there are no more instances than CUs -- latency decides,
  // and 8 waves deal a level's rows and batches over more SIMDs: the n=101
  // portfolio 0.94 -> 0.85 ms per call -- or 1,024 when some phase has more
  // than 16 wave batches (two per wave at 512), and otherwise enough threads that
  // the workgroups the LDS allows per CU hold 16 waves: 256 at 4 per CU, 512
  // at 2 or 3, 1,024 at 1)
  auto threads = [&](int i) {
    if (insts[i].resident == 0) return 256;
    if (threads_env) return threads_env;
    if (!wgsize.empty()) return (int)wgsize[i];
    if (nwg_total <= ctx().num_cu) return insts[i].lay.mbat > 16 ? 1024 : 512;
    const int w = wpc(insts[i].lds_bytes);
    return w >= 4 ? 256 : (w >= 2 ? 512 : 1024);
  };
  auto tcode = [&](int i) { const int t = threads(i); return t == 1024 ? 2 : (t == 512 ? 1 : 0); };
  // workgroups per CU: what the LDS allows, and at most 16 waves per CU
  // (four per SIMD at the kernels' 128 VGPRs)
  auto occ = [&](int i) { return std::min(wpc(insts[i].lds_bytes), 16 / (threads(i) / 64)); };
  auto key = [&](int i) {
    const int p = persist_env >= 0 ? persist_env : (persist.empty() ? 0 : persist[i]);
    return ((occ(i) - 1) * kSizes + tcode(i)) * 2 + p;
  };
  order.clear();
  for (int i = 0; i < count; ++i)
    if (!is_big[i]) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    const int ka = key(a), kb = key(b);
    if (ka != kb) return ka < kb;
    return S[rep[grp[a]]].nnz() > S[rep[grp[b]]].nnz();
  });
  ncls = 0;
  const int nwg = (int)order.size();
  for (int k = 0; k < nwg; ++k) {
    if (k == 0 || key(order[k]) != key(order[k - 1])) {
      cls_start[ncls] = k;
      cls_lds[ncls] = 0;
      cls_persist[ncls] = key(order[k]) & 1;
      cls_occ[ncls] = occ(order[k]);
      cls_threads[ncls] = threads(order[k]);
      ++ncls;
    }
    cls_lds[ncls - 1] = std::max(cls_lds[ncls - 1], insts[order[k]].lds_bytes);
  }
  cls_start[ncls] = nwg;
  // classes to streams, longest processing time first: a class's work is
  // estimated as its instances' nonzeros and iterates (A, A' and, direct
  // plugin, L), divided by the workgroups that run at once (a class that fits
  // more workgroups per CU finishes the same work sooner)
  std::vector<double> work(ncls, 0.0);
  for (int c = 0; c < ncls; ++c) {
    for (int k = cls_start[c]; k < cls_start[c + 1]; ++k) {
      const int i = order[k];
      const Setup& r = S[rep[grp[i]]];
      double w = 2.0 * (double)r.nnz() + 3.0 * (S[i].n + S[i].m);
      if (direct) w += 2.0 * (double)F[grp[i]].Lp.back();
      work[c] += w;
    }
    work[c] /= cls_occ[c];
  }
  for (int c = 0; c < ncls; ++c) cls_seq[c] = c;
  std::stable_sort(cls_seq, cls_seq + ncls, [&](int a, int b) { return work[a] > work[b]; });
  double load[kStreams] = {0};
  for (int q = 0; q < ncls; ++q) {
    const int c = cls_seq[q];
    int best = 0;
    for (int t = 1; t < kStreams; ++t)
      if (load[t] < load[best]) best = t;
    cls_stream[c] = best;
    load[best] += work[c];
  }
}

// Launch-shape tuning.  The candidates per size bucket (half octaves of
// n + m) are the kernel variant (streaming, resident, image in device
// memory), the launch mode (a workgroup per instance, or persistent
// workgroups that take instances from a work queue) and, for the image
// variants, the entries per lane that the wave batches aim for (4, 8) and
// the threads per workgroup (256, 512, 1,024).
// Each is timed (HIP events) on a sample of the bucket's instances, every
// one running a fixed kTuneIters iterations (EPS 0, so no early stop): when
// the whole batch holds more instances than there are CUs, enough records to
// fill the GPU at the highest occupancy (replicas of up to kTuneDistinct
// instances), else the bucket itself.  Decisions are cached per process by (device, precision,
// plugin, load, bucket).
namespace {
constexpr int kTuneIters = 50;
constexpr int kTuneDistinct = 64;
constexpr double kPersistMargin = 0.02;
struct TuneChoice { signed char v, p, e; short t; };
std::mutex g_tune_mu;
std::map<std::vector<int>, std::pair<TuneChoice, std::string>> g_tune_cache;
const char* variant_name(int v) { return v == 1 ? "resident" : (v == 2 ? "image" : "stream"); }
}  // namespace

bool WgBatch::tune(int dev, bool fp64_, std::string& err) {
  tuning.clear();
  if (wg_variant_env() >= 0) return true;  // the environment forces the variant
  Ctx& C = ctx();
  if (!C.ensure(dev, err)) return false;
  const int count = (int)S.size();
  const int tsz = fp64_ ? 8 : 4;
  const int lds_cap = C.max_lds - 64;
  var_force.assign(count, -1);
  persist.assign(count, 0);
  epl.assign(count, 4);
  wgsize.assign(count, 256);
  std::map<int, std::vector<int>> buckets;
  int nwg = 0;
  for (int i = 0; i < count; ++i) {
    const int N = S[i].n + S[i].m;
    if (wg_lds_bytes(S[i].n, S[i].m, N, direct, tsz) > lds_cap) continue;  // whole-GPU engine
    ++nwg;
    buckets[(int)std::floor(2.0 * std::log2((double)std::max(N, 2)))].push_back(i);
  }
  const int load = nwg > C.num_cu ? 1 : 0;
  std::string js = "[";
  for (const auto& bk : buckets) {
    const std::vector<int>& mem = bk.second;
    const std::vector<int> key = {dev, fp64_ ? 1 : 0, direct ? 1 : 0, load, bk.first};
    TuneChoice best{-1, 0, 4, 256};
    std::string cands;
    bool cached = false;
    {
      std::lock_guard<std::mutex> g(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) { best = it->second.first; cands = it->second.second; cached = true; }
    }
    int ns = (int)mem.size();
    if (!cached) {
      // a loaded GPU: as many records as the most workgroups per CU any
      // variant allows times the CUs, so every candidate runs at its full
      // occupancy; drawn from at most kTuneDistinct instances of the bucket
      // (replicas share their instance's packed matrices)
      const int nd = load ? std::min<int>((int)mem.size(), kTuneDistinct) : (int)mem.size();
      if (load) ns = kOccMax * C.num_cu;
      // the sample as a batch of its own, sharing matrix groups like this one
      auto make = [&](int e, std::unique_ptr<WgBatch>& T) {
        T.reset(new WgBatch);
        T->direct = direct;
        T->pack_all = true;  // every candidate variant runs on the same records
        T->S.resize(ns);
        T->par.resize(ns);
        T->grp.resize(ns);
        T->epl.assign(ns, (signed char)e);
        std::map<int, int> gmap;
        for (int k = 0; k < ns; ++k) {
          const int i = mem[(size_t)(k % nd) * mem.size() / nd];
          auto g = gmap.find(grp[i]);
          if (g == gmap.end()) {
            g = gmap.emplace(grp[i], (int)T->rep.size()).first;
            T->rep.push_back(k);
            T->S[k] = S[i];
            const Setup& r = S[rep[grp[i]]];
            if (rep[grp[i]] != i) {
              T->S[k].cp = r.cp; T->S[k].ci = r.ci; T->S[k].cx = r.cx;
              T->S[k].rp = r.rp; T->S[k].ri = r.ri; T->S[k].rx = r.rx; T->S[k].csr_src = r.csr_src;
            }
            if (direct) T->F.push_back(F[grp[i]]);
          } else {  // the matrix arrays live with the group's first sample
            const Setup& a = S[i];
            Setup& t = T->S[k];
            t.n = a.n; t.m = a.m; t.f = a.f; t.l = a.l;
            t.soc_ptr = a.soc_ptr; t.b = a.b; t.c = a.c; t.D = a.D; t.E = a.E; t.E0 = a.E0;
            t.lambda = a.lambda; t.pscale = a.pscale; t.dscale = a.dscale;
            t.x = a.x; t.s = a.s; t.y = a.y;
            t.dev = a.dev; t.normalize = a.normalize;
          }
          T->grp[k] = g->second;
          Params q = par[i];
          q.MAX_ITERS = kTuneIters;
          q.EPS_ABS = 0.0;
          q.VERBOSE = 0;
          T->par[k] = q;
        }
        return T->prepare(dev, fp64_, direct, false, err);
      };
      // one untimed run per sample batch (first launch costs), then one timed
      // run per candidate
      auto time = [&](WgBatch& T, int v, int p, double& ms, int th = 256) {
        T.var_force.assign(ns, (signed char)v);
        T.persist.assign(ns, (signed char)p);
        T.wgsize.assign(ns, (short)th);
        T.set_shapes();
        if (!T.upload_records(err)) return false;
        if (!T.warm) {
          if (!T.run(err)) return false;
          T.warm = true;
        }
        if (!T.run(err)) return false;
        ms = T.kernel_ms;
        return true;
      };
      auto has = [&](WgBatch& T, int v) {
        for (const VarLds& L : T.vlds)
          if (v == 1 ? L.y != 0 : L.z != 0) return true;
        return false;
      };
      char buf[96];
      double best_ms = 1e30;
      // the work-queue launch and the wider workgroups must win by
      // kPersistMargin each: every launch mode and workgroup size is a launch
      // class of its own, and classes beyond the device's hardware queues run
      // one after another
      auto eff = [](double ms, int p, int th) { return ms * (1.0 + kPersistMargin * (p + (th != 256 ? 1 : 0))); };
      auto note = [&](int v, int p, int e, int th, double ms) {
        std::snprintf(buf, sizeof buf, "%s[\"%s\",%d,%d,%d,%.4f]", cands.empty() ? "" : ",", variant_name(v), p, e,
                      th, ms);
        cands += buf;
      };
      std::unique_ptr<WgBatch> T, T2;
      if (!make(4, T)) { err = "tuning: " + err; return false; }
      // stage 1: every variant and launch mode; stage 2: from the best
      // image variant, 8 entries per lane and then 512 / 1,024 threads (the
      // image variants only have those knobs, and the streaming variant
      // winning stage 1 does not mean an image variant cannot win stage 2)
      TuneChoice img{-1, 0, 4, 256};
      double img_ms = 1e30;
      auto consider = [&](const TuneChoice& c, double ms) {
        note(c.v, c.p, c.e, c.t, ms);
        const double x = eff(ms, c.p, c.t);
        if (x < best_ms) { best_ms = x; best = c; }
        if (c.v != 0 && x < img_ms) { img_ms = x; img = c; }
      };
      for (int v = 0; v < kVariants; ++v) {
        if (v && !has(*T, v)) continue;
        for (int p = 0; p < 2; ++p) {
          double ms;
          if (!time(*T, v, p, ms)) { err = "tuning: " + err; return false; }
          consider(TuneChoice{(signed char)v, (signed char)p, 4, 256}, ms);
        }
      }
      if (img.v > 0) {
        double ms;
        if (!make(8, T2) || !time(*T2, img.v, img.p, ms)) { err = "tuning: " + err; return false; }
        consider(TuneChoice{img.v, img.p, 8, 256}, ms);
        const TuneChoice base = img;
        WgBatch& Tb = base.e == 8 ? *T2 : *T;
        for (int th : {512, 1024}) {
          if (!time(Tb, base.v, base.p, ms, th)) { err = "tuning: " + err; return false; }
          consider(TuneChoice{base.v, base.p, base.e, (short)th}, ms);
        }
      }
      std::lock_guard<std::mutex> g(g_tune_mu);
      g_tune_cache[key] = {best, cands};
    }
    for (int i : mem) {
      var_force[i] = best.v;
      persist[i] = best.p;
      epl[i] = best.e;
      wgsize[i] = best.t;
    }
    char head[256];
    std::snprintf(head, sizeof head,
                  "%s{\"n_plus_m\":[%d,%d],\"count\":%d,\"sample\":%d,\"cached\":%s,\"variant\":\"%s\","
                  "\"persistent\":%d,\"entries_per_lane\":%d,\"threads\":%d,\"candidates\":[",
                  js.size() > 1 ? "," : "", (int)std::ceil(std::exp2(bk.first / 2.0)),
                  (int)std::ceil(std::exp2((bk.first + 1) / 2.0)) - 1, (int)mem.size(), cached ? 0 : ns,
                  cached ? "true" : "false", variant_name(best.v), best.p, best.e, best.t);
    js += head + cands + "]}";
  }
  tuning = js + "]";
  return true;
}

// Device records in instance order and in class order, from `insts`.
bool WgBatch::upload_records(std::string& err) {
  unsigned char* base = static_cast<unsigned char*>(dbuf);
  const size_t count = insts.size();
  std::vector<WgInst> byc(order.size());
  for (size_t k = 0; k < order.size(); ++k) byc[k] = insts[order[k]];
  hipError_t e = hipMemcpyAsync(base + inst_off, insts.data(), sizeof(WgInst) * count, hipMemcpyHostToDevice,
                                stream);
  if (e == hipSuccess && !byc.empty())
    e = hipMemcpyAsync(base + cinst_off, byc.data(), sizeof(WgInst) * byc.size(), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) { err = std::string("records: ") + hipGetErrorString(e); return false; }
  return true;
}

bool WgBatch::set_warm(const float* dW, long ld, int check, std::string& err) {
  if (dW && !big.empty()) {
    err = "set_warm: device warm starts are not supported for members on the whole-GPU engine (instance " +
          std::to_string(big[0]) + ")";
    return false;
  }
  for (size_t i = 0; i < insts.size(); ++i) {
    if (!is_big.empty() && is_big[i]) continue;
    insts[i].wsrc = dW ? dW + (size_t)i * ld : nullptr;
    insts[i].warm_check = dW ? check : 0;
  }
  if (!upload_records(err)) {
    err = "set_warm: " + err;
    return false;
  }
  return true;
}

// One launch per class, on the class's stream (the extra streams forked from
// and joined back into `stream`; ev0 .. ev1 covers all of them).
bool WgBatch::run(std::string& err) {
  unsigned char* base = static_cast<unsigned char*>(dbuf);
  const WgInst* recs = reinterpret_cast<const WgInst*>(base + cinst_off);
  const double t0 = now_ms();
  hipError_t e = hipSuccess;
  // the extra streams and their join events (created on first use)
  bool used[kStreams] = {false};
  for (int c = 0; c < ncls; ++c) used[cls_stream[c]] = true;
  for (int t = 1; t < kStreams && e == hipSuccess; ++t) {
    if (!used[t]) continue;
    if (!xs[t - 1]) e = hipStreamCreateWithFlags(&xs[t - 1], hipStreamNonBlocking);
    if (e == hipSuccess && !xev[t - 1]) e = hipEventCreateWithFlags(&xev[t - 1], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventRecord(ev0, stream);
  for (int t = 1; t < kStreams && e == hipSuccess; ++t)
    if (used[t]) e = hipStreamWaitEvent(xs[t - 1], ev0, 0);
  for (int q = 0; q < ncls && e == hipSuccess; ++q) {
    const int c = cls_seq[q];
    hipStream_t sc = cls_stream[c] == 0 ? stream : xs[cls_stream[c] - 1];
    // a CU per workgroup (no more workgroups in the whole batch than CUs),
    // image variants only
    bool solo = !cls_persist[c] && (int)order.size() <= ctx().num_cu;
    for (int k = cls_start[c]; solo && k < cls_start[c + 1]; ++k) solo = insts[order[k]].resident != 0;
    e = wg_launch(recs + cls_start[c], cls_start[c + 1] - cls_start[c], cls_lds[c], cls_threads[c], fp64, direct,
                  cls_persist[c] ? reinterpret_cast<int*>(base + queue_off) + c : nullptr,
                  cls_occ[c] * ctx().num_cu, solo, sc);
  }
  for (int t = 1; t < kStreams && e == hipSuccess; ++t) {
    if (!used[t]) continue;
    e = hipEventRecord(xev[t - 1], xs[t - 1]);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, xev[t - 1], 0);
  }
  if (e == hipSuccess) e = hipEventRecord(ev1, stream);
  if (e == hipSuccess) e = hipEventSynchronize(ev1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
  if (e != hipSuccess) { err = std::string("wg engine: ") + hipGetErrorString(e); return false; }
  kernel_ms = ms;
  for (size_t b = 0; b < gsol.size(); ++b) {
    if (!gsol[b]->run(err)) { err = "instance " + std::to_string(big[b]) + ": " + err; return false; }
    kernel_ms += gsol[b]->kernel_ms;
  }
  wall_ms = now_ms() - t0;
  return true;
}

bool WgBatch::fetch(std::vector<DevResult>& res, std::string& err) {
  unsigned char* base = static_cast<unsigned char*>(dbuf);
  const int count = (int)S.size();
  res.resize(count);
  if (!own_mem && count == 1 && !is_big[0]) {
    // one copy into the context's pinned buffer: x, s, y, the history and
    // the result lie in that order in [xo, res_off + result)
    Ctx& C = ctx();
    const size_t lo = xo[0], hi = res_off + sizeof(DevResult);
    hipError_t e = hi - lo <= C.hcap ? hipMemcpyAsync(C.hbuf, base + lo, hi - lo, hipMemcpyDeviceToHost, stream)
                                     : hipErrorInvalidValue;
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) { err = std::string("download: ") + hipGetErrorString(e); return false; }
    const unsigned char* h = static_cast<const unsigned char*>(C.hbuf) - lo;
    Setup& s = S[0];
    std::memcpy(s.x.data(), h + xo[0], sizeof(double) * s.n);
    std::memcpy(s.s.data(), h + so[0], sizeof(double) * s.m);
    std::memcpy(s.y.data(), h + yo[0], sizeof(double) * s.m);
    std::memcpy(res.data(), h + res_off, sizeof(DevResult));
    hist.assign(1, {});
    if (insts[0].hist_rows) {
      hist[0].resize((size_t)(direct ? 5 : 25) * insts[0].hist_rows);
      std::memcpy(hist[0].data(), h + ho[0], sizeof(double) * hist[0].size());
    }
    return true;
  }
  hipError_t e = hipMemcpyAsync(res.data(), base + res_off, sizeof(DevResult) * count,
                                hipMemcpyDeviceToHost, stream);
  for (int i = 0; i < count && e == hipSuccess; ++i) {
    if (is_big[i]) continue;
    Setup& s = S[i];
    e = hipMemcpyAsync(s.x.data(), base + xo[i], sizeof(double) * s.n, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.s.data(), base + so[i], sizeof(double) * s.m, hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.y.data(), base + yo[i], sizeof(double) * s.m, hipMemcpyDeviceToHost, stream);
  }
  hist.assign(count, {});
  for (int i = 0; i < count && e == hipSuccess; ++i)
    if (insts[i].hist_rows) {
      hist[i].assign((size_t)(direct ? 5 : 25) * insts[i].hist_rows, 0.0);
      e = hipMemcpyAsync(hist[i].data(), base + ho[i], sizeof(double) * hist[i].size(), hipMemcpyDeviceToHost,
                         stream);
    }
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) { err = std::string("download: ") + hipGetErrorString(e); return false; }
  for (size_t b = 0; b < gsol.size(); ++b)
    if (!gsol[b]->fetch(S[big[b]], res[big[b]], nullptr, err)) return false;
  return true;
}

bool solve_wg(std::vector<Setup>& S, std::vector<Factor>& F, const std::vector<const Params*>& par,
              bool fp64, bool direct, int device, std::vector<DevResult>& res,
              double& upload_ms, double& solve_ms, std::string& err, std::vector<double>* hist0) {
  WgBatch B;
  B.S.swap(S);
  B.F.swap(F);
  for (const Params* p : par) B.par.push_back(*p);
  B.want_hist = hist0 != nullptr && B.S.size() == 1;
  bool ok = B.prepare(device, fp64, direct, true, err) && B.run(err) && B.fetch(res, err);
  B.S.swap(S);
  B.F.swap(F);
  upload_ms = B.upload_ms;
  solve_ms = B.wall_ms;
  if (ok && hist0 && !B.hist.empty()) hist0->swap(B.hist[0]);
  return ok;
}

// x_out = E x, y_out = D y, s_out = s / D (pdos.c:251-294).
Sol* make_sol(const Setup& s, int status) {
  Sol* sol = static_cast<Sol*>(std::malloc(sizeof(Sol)));
  sol->n = s.n;
  sol->m = s.m;
  sol->x = static_cast<double*>(std::malloc(sizeof(double) * s.n));
  sol->y = static_cast<double*>(std::malloc(sizeof(double) * s.m));
  sol->s = static_cast<double*>(std::malloc(sizeof(double) * s.m));
  for (int j = 0; j < s.n; ++j) sol->x[j] = s.E[j] * s.x[j];
  for (int i = 0; i < s.m; ++i) sol->y[i] = s.D[i] * s.y[i];
  for (int i = 0; i < s.m; ++i) sol->s[i] = s.s[i] / s.D[i];
  std::memset(sol->status, 0, sizeof(sol->status));
  std::strcpy(sol->status, status == 0 ? "Solved" : "Indeterminate");
  return sol;
}

void fill_info(PdosInfo* info, const DevResult& r, const Setup& s, int engine, double setup_ms,
               double solve_ms, double lnz) {
  if (!info) return;
  info->iters = r.iters;
  info->cg_iters = r.cg_iters;
  info->status = r.status;
  info->engine = engine;
  info->fallback = r.fallback;
  info->p_res = r.p_res;
  info->d_res = r.d_res;
  info->p_obj = r.p_obj;
  info->d_obj = r.d_obj;
  info->eta = r.eta;
  info->lambda = s.lambda;
  info->setup_ms = setup_ms;
  info->solve_ms = solve_ms;
  info->total_ms = setup_ms + solve_ms;
  info->lnz = lnz;
}

int env_precision() {
  const char* p = std::getenv("PDOS_PRECISION");
  if (p && (!std::strcmp(p, "mixed") || !std::strcmp(p, "fp32"))) return PDOS_MIXED;
  return PDOS_FP64;
}

int env_equil() {
  const char* p = std::getenv("PDOS_EQUIL");
  return p && !std::strcmp(p, "ruiz") ? PDOS_EQUIL_RUIZ : PDOS_EQUIL_REFERENCE;
}

// The VERBOSE log in the reference's layout (pdos.c:79-92, 107, 124-127,
// 306-335; direct/private.c:157): banner, factorization time (direct),
// lambda, the column header and its rule, a row for every check i with
// i % 10 == 0 that did not stop the loop (from the engine's history), the
// final row (updates done, last check's residuals) and the solve time.
void print_verbose(const Params* p, const DevResult& r, double lam, double secs, const std::vector<double>& hist,
                   double factor_secs) {
  if (!p->VERBOSE) return;
  static const char* const kHead[6] = {" iter", " ||Ax+s-b||", " ||A'y+c||", "    c'x    ", "   -b'y    ",
                                       "    eta   "};
  auto row = [&](long i, const double* v) {
    std::printf("%*li | ", (int)std::strlen(kHead[0]), i);
    for (int k = 0; k < 4; ++k) std::printf("%*.3e   ", (int)std::strlen(kHead[k + 1]), v[k]);
    std::printf("%*.3e   \n", (int)std::strlen(kHead[5]), v[4]);
  };
  std::printf("\nPDOS - A Primal-Dual Operator Splitting for Cone Programming.\n");
  std::printf("       MI355X (gfx950) build: HIP kernels, FP64 or mixed precision.\n\n");
  if (factor_secs >= 0) std::printf("Factorization time: %4.8fs\n", factor_secs);
  std::printf("lambda: %5.3e\n\n", lam);
  size_t len = 0;
  for (int k = 0; k < 5; ++k) {
    std::printf(k == 0 ? "%s | " : "%s   ", kHead[k]);
    len += std::strlen(kHead[k]) + 3;
  }
  std::printf("%s\n", kHead[5]);
  len += std::strlen(kHead[5]);
  std::printf("%s\n", std::string(len + 3, '=').c_str());
  // history: rows x 5 residuals, then (workgroup engine, indirect plugin)
  // rows x 10 updates' {CG steps, final |q|}, printed after each update as
  // indirect/private.c:100 does
  const size_t rows = (size_t)verbose_hist_rows(p->MAX_ITERS);
  const bool cg = hist.size() == 25 * rows;
  for (long i = 0; i < r.iters && (size_t)i < 10 * rows && hist.size() >= 5 * rows; ++i) {
    if (i % 10 == 0) row(i, &hist[(size_t)(i / 10) * 5]);
    if (cg) {
      const double* c = &hist[5 * rows + (size_t)i * 2];
      std::printf("terminating cg residual = %4f, took %li itns\n", c[1], (long)c[0]);
    }
  }
  const double last[5] = {r.p_res, r.d_res, r.p_obj, r.d_obj, r.eta};
  row(r.iters, last);
  std::printf("Total solve time is %4.8fs\n", secs);
}

}  // namespace pdos_amd

using namespace pdos_amd;

extern "C" {

void pdos_default_options(PdosOptions* opt) {
  std::memset(opt, 0, sizeof(*opt));
  opt->plugin = PDOS_DEFAULT_PLUGIN;
  opt->precision = PDOS_FP64;
  opt->engine = PDOS_ENGINE_AUTO;
}

const char* pdos_last_error(void) { return g_err.c_str(); }

const char* pdos_version(void) { return "pdos-mi355x 0.1 (gfx950)"; }

int pdos_last_info(PdosInfo* info) {
  if (!g_info_valid || !info) return -1;
  *info = g_info;
  return 0;
}

Sol* pdos_ex(const Data* d, const Cone* k, const PdosOptions* opt, PdosInfo* info) {
  g_err.clear();
  g_info_valid = false;
  if (d == nullptr || k == nullptr) return nullptr;  // pdos.c:61-63
  PdosOptions o;
  if (opt) o = *opt;
  else pdos_default_options(&o);
  double t0 = now_ms();
  std::vector<Setup> S(1);
  std::vector<Factor> F;
  std::string err;
  if (!ctx().ensure(o.device, err)) {
    g_err = err;
    return nullptr;
  }
  if (!build_setup(d, k, S[0], err, o.equil)) {
    g_err = err;
    return nullptr;
  }
  const bool direct = o.plugin == PDOS_DIRECT;
  double lnz = 0, factor_secs = -1.0;
  if (direct) {
    F.resize(1);
    const double tf = now_ms();
    const bool fok = build_factor(S[0], {}, F[0], err);
    factor_secs = (now_ms() - tf) * 1e-3;
    if (!fok) {
      g_err = err;
      return nullptr;
    }
    lnz = (double)F[0].Lp.back();
  }
  double t1 = now_ms();
  std::vector<const Params*> par{d->p};
  std::vector<DevResult> res;
  std::vector<double> hist;
  double up = 0, sv = 0;
  int engine = PDOS_ENGINE_WG;
  bool ok;
  if (o.engine == PDOS_ENGINE_GRID ||
      (o.engine == PDOS_ENGINE_AUTO && grid_preferred(S[0], o.precision == PDOS_FP64, direct))) {
    engine = PDOS_ENGINE_GRID;
    ok = solve_grid(S[0], *d->p, o.precision == PDOS_FP64, direct ? &F[0] : nullptr, o.device, res, up, sv, err,
                    &hist);
  } else {
    ok = solve_wg(S, F, par, o.precision == PDOS_FP64, direct, o.device, res, up, sv, err, &hist);
  }
  if (!ok) {
    g_err = err;
    if (d->p->VERBOSE) std::fprintf(stderr, "pdos: %s\n", err.c_str());
    return nullptr;
  }
  Sol* sol = make_sol(S[0], res[0].status);
  PdosInfo inf;
  fill_info(&inf, res[0], S[0], engine, (t1 - t0) + up, sv, lnz);
  g_info = inf;
  g_info_valid = true;
  if (info) *info = inf;
  print_verbose(d->p, res[0], S[0].lambda, sv * 1e-3, hist, factor_secs);
  return sol;
}

Sol* pdos(const Data* d, const Cone* k) {
  if (d == nullptr || k == nullptr) return nullptr;
  long tot = k->f + k->l;
  for (long i = 0; i < k->qsize; ++i) tot += k->q[i];
  if (d->m != tot) {  // the reference asserts (pdos.c:65-71)
    std::fprintf(stderr, "pdos: cone sizes (%ld) do not match m (%ld)\n", tot, (long)d->m);
    std::abort();
  }
  PdosOptions o;
  pdos_default_options(&o);
  o.precision = env_precision();
  o.equil = env_equil();
  return pdos_ex(d, k, &o, nullptr);
}

}  // extern "C"

struct PdosBatchHandle {
  pdos_amd::WgBatch B;
  double setup_ms = 0;
  // host wall time of prepare's parts: setup (scaling, transpose), grouping
  // and factorization, launch-shape tuning, packing and upload
  double part_ms[4] = {0, 0, 0, 0};
};

static PdosBatchHandle* batch_prepare(const Data* const* ds, const Cone* const* ks, int count,
                                      const PdosOptions* opt, bool borrow) {
  g_err.clear();
  if (count <= 0 || !ds || !ks) {
    g_err = "bad batch arguments";
    return nullptr;
  }
  PdosOptions o;
  if (opt) o = *opt;
  else pdos_default_options(&o);
  const bool direct = o.plugin == PDOS_DIRECT;
  const double t0 = now_ms();
  PdosBatchHandle* h = new PdosBatchHandle;
  h->B.S.resize(count);
  h->B.par.resize(count);
  h->B.direct = direct;
  std::string err;
  // the setup on the device (batch_setup.hip) for the indirect plugin's
  // instances that the workgroup engine runs, with the reference scaling;
  // the direct plugin factors the scaled matrix on the host, and members
  // past one workgroup's LDS run on the whole-GPU engine from a host setup.
  // PDOS_HOST_SETUP=1 keeps every setup on the host.
  bool dev_setup = !direct && o.equil == PDOS_EQUIL_REFERENCE;
  if (const char* hs = std::getenv("PDOS_HOST_SETUP")) dev_setup = dev_setup && std::atoi(hs) == 0;
  int lds_cap = 0;
  if (dev_setup) {
    Ctx& C = ctx();
    if (!C.ensure(o.device, err)) { g_err = err; delete h; return nullptr; }
    lds_cap = C.max_lds - 64;
  }
  const int tsz = o.precision == PDOS_FP64 ? 8 : 4;
  double tp = now_ms();
  const int bad = parallel_for(count, [&](int i, std::string& e) {
    if (!ds[i] || !ks[i]) return false;
    const bool raw = dev_setup && ds[i]->n > 0 && ds[i]->m > 0 &&
                     wg_lds_bytes((int)ds[i]->n, (int)ds[i]->m, (int)(ds[i]->n + ds[i]->m), false, tsz) <= lds_cap;
    if (!(raw ? build_setup_raw(ds[i], ks[i], h->B.S[i], e) : build_setup(ds[i], ks[i], h->B.S[i], e, o.equil)))
      return false;
    h->B.par[i] = *ds[i]->p;
    return true;
  }, err);
  auto lap = [&](int k) {
    const double t = now_ms();
    h->part_ms[k] = t - tp;
    tp = t;
  };
  lap(0);
  if (bad >= 0) {
    g_err = "instance " + std::to_string(bad) + ": " + (err.empty() ? "NULL input" : err);
    delete h;
    return nullptr;
  }
  if (!h->B.group_and_factor(err)) {
    g_err = err;
    delete h;
    return nullptr;
  }
  lap(1);
  if (o.tune && !h->B.tune(o.device, o.precision == PDOS_FP64, err)) {
    g_err = err;
    delete h;
    return nullptr;
  }
  lap(2);
  if (!h->B.prepare(o.device, o.precision == PDOS_FP64, direct, borrow, err)) {
    g_err = err;
    delete h;
    return nullptr;
  }
  lap(3);
  h->setup_ms = now_ms() - t0;
  return h;
}

static int batch_fetch(PdosBatchHandle* h, Sol** sols, PdosInfo* infos) {
  std::vector<DevResult> res;
  std::string err;
  if (!h->B.fetch(res, err)) {
    g_err = err;
    return -1;
  }
  for (size_t i = 0; i < res.size(); ++i) {
    sols[i] = make_sol(h->B.S[i], res[i].status);
    if (infos)
      fill_info(&infos[i], res[i], h->B.S[i], h->B.is_big[i] ? PDOS_ENGINE_GRID : PDOS_ENGINE_WG, h->setup_ms,
                h->B.wall_ms,
                h->B.direct ? (double)h->B.F[h->B.grp[i]].Lp.back() : 0.0);
  }
  return 0;
}

extern "C" {

PdosBatchHandle* pdos_batch_prepare(const Data* const* ds, const Cone* const* ks, int count,
                                    const PdosOptions* opt) {
  return batch_prepare(ds, ks, count, opt, false);
}

int pdos_batch_run(PdosBatchHandle* h, double* kernel_ms) {
  if (!h) return -1;
  std::string err;
  if (!h->B.run(err)) {
    g_err = err;
    return -1;
  }
  if (kernel_ms) *kernel_ms = h->B.kernel_ms;
  return 0;
}

int pdos_batch_fetch(PdosBatchHandle* h, Sol** sols, PdosInfo* infos) {
  if (!h || !sols) return -1;
  return batch_fetch(h, sols, infos);
}

void pdos_batch_free(PdosBatchHandle* h) { delete h; }

size_t pdos_batch_device_bytes(const PdosBatchHandle* h) { return h ? h->B.bytes : 0; }

const char* pdos_batch_tuning(const PdosBatchHandle* h) { return h ? h->B.tuning.c_str() : ""; }

int pdos_batch_prepare_times(const PdosBatchHandle* h, double* ms) {
  if (!h || !ms) return -1;
  for (int k = 0; k < 4; ++k) ms[k] = h->part_ms[k];
  return 0;
}

int pdos_batch_scales(const PdosBatchHandle* h, int i, double* D, double* E, double* lambda) {
  if (!h || i < 0 || i >= (int)h->B.S.size()) return -1;
  const Setup& s = h->B.S[i];
  if (D) std::copy(s.D.begin(), s.D.end(), D);
  if (E) std::copy(s.E.begin(), s.E.end(), E);
  if (lambda) *lambda = s.lambda;
  return s.dev ? 1 : 0;
}

int pdos_batch_set_warm_dev(PdosBatchHandle* h, const float* dW, long ld, int check) {
  if (!h) return -1;
  std::string err;
  if (!h->B.set_warm(dW, ld, check, err)) {
    g_err = err;
    return -1;
  }
  return 0;
}

void* pdos_device_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
    g_err = "hipMalloc failed";
    return nullptr;
  }
  return p;
}

void pdos_device_free(void* p) {
  if (p) (void)hipFree(p);
}

int pdos_device_upload(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

int pdos_device_download(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}

int pdos_device_copy(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return 0;
  if (!dst || !src) return -1;
  if (pdos_amd::device_copy(dst, src, bytes, nullptr) != hipSuccess) return -1;
  return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}

int pdos_batch(const Data* const* ds, const Cone* const* ks, int count, const PdosOptions* opt,
               Sol** sols, PdosInfo* infos) {
  g_err.clear();
  if (count < 0 || (count > 0 && (!ds || !ks || !sols))) {
    g_err = "bad batch arguments";
    return -1;
  }
  if (count == 0) return 0;
  for (int i = 0; i < count; ++i) sols[i] = nullptr;
  PdosBatchHandle* h = batch_prepare(ds, ks, count, opt, true);
  if (!h) return -1;
  std::string err;
  int rc = h->B.run(err) ? batch_fetch(h, sols, infos) : (g_err = err, -1);
  delete h;
  return rc;
}

int pdos_setup_host(const Data* d, const Cone* k, int equil, double* D, double* E, double* lambda) {
  g_err.clear();
  Setup S;
  std::string err;
  if (!build_setup(d, k, S, err, equil)) {
    g_err = err;
    return -1;
  }
  if (D) std::copy(S.D.begin(), S.D.end(), D);
  if (E) std::copy(S.E.begin(), S.E.end(), E);
  if (lambda) *lambda = S.lambda;
  return 0;
}

int pdos_kkt_order(const Data* d, const Cone* k, long* perm) {
  g_err.clear();
  Setup S;
  std::vector<int> P;
  std::string err;
  if (!build_setup(d, k, S, err) || !kkt_order(S, P, err)) {
    g_err = err;
    return -1;
  }
  for (size_t i = 0; i < P.size(); ++i) perm[i] = P[i];
  return 0;
}

int pdos_kkt_solve_host(const Data* d, const Cone* k, const double* rhs, double* z, long* stats) {
  g_err.clear();
  Setup S;
  Factor F;
  std::string err;
  if (!build_setup(d, k, S, err) || !build_factor(S, {}, F, err)) {
    g_err = err;
    return -1;
  }
  const int N = F.N;
  std::vector<double> bp(N);
  for (int j = 0; j < N; ++j) bp[j] = rhs[F.P[j]];
  for (int lv = 0; lv < F.nfl(); ++lv)
    for (int r = F.fl_ptr[lv]; r < F.fl_ptr[lv + 1]; ++r) {
      const int i = F.fl_rows[r];
      double v = bp[i];
      for (int p = F.Rp[i]; p < F.Rp[i + 1]; ++p) v -= F.Rx[p] * bp[F.Ri[p]];
      bp[i] = v;
    }
  for (int lv = 0; lv < F.nbl(); ++lv)
    for (int r = F.bl_ptr[lv]; r < F.bl_ptr[lv + 1]; ++r) {
      const int j = F.bl_rows[r];
      double v = bp[j] / F.Ld[j];
      for (int p = F.Lp[j]; p < F.Lp[j + 1]; ++p) v -= F.Lx[p] * bp[F.Li[p]];
      bp[j] = v;
    }
  for (int j = 0; j < N; ++j) z[F.P[j]] = bp[j];
  if (stats) {
    stats[0] = N;
    stats[1] = F.Lp[N];
    stats[2] = F.nfl();
    stats[3] = F.nbl();
  }
  return 0;
}

void freeData(Data** d, Cone** k) {
  if (d && *d) {
    Data* p = *d;
    std::free(p->b);
    std::free(p->c);
    std::free(p->Ax);
    std::free(p->Ai);
    std::free(p->Ap);
    std::free(p->p);
    std::free(p->x);
    std::free(p->y);
    std::free(p->s);
    std::free(p);
  }
  if (k && *k) {
    std::free((*k)->q);
    std::free(*k);
  }
  if (d) *d = nullptr;
  if (k) *k = nullptr;
}

void freeSol(Sol** sol) {
  if (sol && *sol) {
    std::free((*sol)->x);
    std::free((*sol)->y);
    std::free((*sol)->s);
    std::free(*sol);
  }
  if (sol) *sol = nullptr;
}

void tic(timer* t) { clock_gettime(CLOCK_MONOTONIC, &t->tic); }

double tocq(timer* t) {
  clock_gettime(CLOCK_MONOTONIC, &t->toc);
  return double(t->toc.tv_sec - t->tic.tv_sec) + 1e-9 * double(t->toc.tv_nsec - t->tic.tv_nsec);
}

double toc(timer* t) {
  const double v = tocq(t);
  std::printf("time: %8.4f seconds.\n", v);
  return v;
}

}  // extern "C"
