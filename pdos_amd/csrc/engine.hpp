// engine.hpp -- per-thread device context and the engine entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/pdos.h"
#include "device.hpp"
#include "host.hpp"

namespace pdos_amd {

// Per-thread device state: one non-blocking stream and grow-only device and
// pinned-host arenas, so a call allocates nothing once warm.
struct Ctx {
  int dev = -1;
  hipStream_t stream = nullptr;
  void* dbuf = nullptr;
  size_t dcap = 0;
  void* hbuf = nullptr;
  size_t hcap = 0;
  int max_lds = 65536;
  int num_cu = 256;
  bool ensure(int device, std::string& err);
  bool reserve(size_t dbytes, size_t hbytes, std::string& err);
};
Ctx& ctx();

// Host staging buffer: arrays at 256-byte aligned offsets.
struct Packer {
  std::vector<unsigned char> buf;
  size_t put(const void* src, size_t bytes);
  size_t zeros(size_t bytes);
};

double now_ms();

hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
hipError_t wg_launch(const WgInst* d_insts, int count, int lds_bytes, bool fp64, bool direct,
                     hipStream_t st);

// A prepared batch for the workgroup engine: inputs resident on the device,
// re-runnable from the same start (the kernel reads x0/s0/y0, writes x/s/y).
struct WgBatch {
  std::vector<Setup> S;
  std::vector<int> grp;       // instance -> matrix group (instances with identical scaled A)
  std::vector<int> rep;       // group -> representative instance (keeps the matrix arrays)
  std::vector<Factor> F;      // per group (direct plugin)
  std::vector<Params> par;
  int device = 0;
  bool fp64 = true, direct = false, own_mem = true;
  // VERBOSE residual histories are kept only for a single pdos()/pdos_ex()
  // instance (the only path that prints them)
  bool want_hist = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void* dbuf = nullptr;
  size_t bytes = 0, res_off = 0, inst_off = 0;
  std::vector<size_t> xo, so, yo;
  std::vector<size_t> ho;           // VERBOSE residual history offsets (rows in insts[i].hist_rows)
  std::vector<std::vector<double>> hist;  // fetched histories (VERBOSE instances)
  std::vector<WgInst> insts;  // host copy of the device records
  int lds = 0;
  // Launch classes: instances grouped by how many of their workgroups fit
  // in a CU's LDS, each class launched with its own LDS size on its own
  // stream; the device holds a second copy of the records in class order.
  static constexpr int kClasses = 4;
  std::vector<int> order;                 // class order -> instance
  int ncls = 0, cls_start[kClasses + 1] = {0}, cls_lds[kClasses] = {0};
  size_t cinst_off = 0;
  hipStream_t xs[kClasses - 1] = {nullptr, nullptr, nullptr};
  hipEvent_t xev[kClasses - 1] = {nullptr, nullptr, nullptr};
  bool upload_records(std::string& err);
  double upload_ms = 0, kernel_ms = 0, wall_ms = 0;
  ~WgBatch();
  // Groups instances by their scaled matrix (and scalings, cones) and drops
  // the duplicate host copies; with `direct`, factors once per group.
  bool group_and_factor(std::string& err);
  bool prepare(int device, bool fp64, bool direct, bool borrow_ctx_arena, std::string& err);
  bool run(std::string& err);
  bool fetch(std::vector<DevResult>& res, std::string& err);
  bool set_warm(const float* dW, long ld, int check, std::string& err);
};

bool solve_wg(std::vector<Setup>& S, std::vector<Factor>& F, const std::vector<const Params*>& par,
              bool fp64, bool direct, int device, std::vector<DevResult>& res, double& upload_ms,
              double& solve_ms, std::string& err, std::vector<double>* hist0 = nullptr);

// Whole-GPU engine for one large instance (indirect plugin).
bool grid_preferred(const Setup& S, bool fp64);
// hist (optional): with par.VERBOSE, the residuals of checks 0, 10, 20, ...
// (5 per row, verbose_hist_rows(MAX_ITERS) rows)
bool solve_grid(Setup& S, const Params& par, bool fp64, int device, std::vector<DevResult>& res,
                double& upload_ms, double& solve_ms, std::string& err, std::vector<double>* hist = nullptr);

}  // namespace pdos_amd
