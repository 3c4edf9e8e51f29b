// engine.hpp -- per-thread device context and the engine entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
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
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // timing events of calls that borrow the context
  bool ensure(int device, std::string& err);
  bool reserve(size_t dbytes, size_t hbytes, std::string& err);
};
Ctx& ctx();

// Host staging buffer: arrays at 256-byte aligned offsets.  The storage
// grows by doubling and is not zero-filled; slot() reserves bytes that the
// caller writes later (several threads may fill separate slots at once, once
// no more space is being added).
class Packer {
 public:
  size_t put(const void* src, size_t bytes);
  size_t zeros(size_t bytes);
  size_t slot(size_t bytes);
  void reserve(size_t bytes) { if (bytes > cap_) grow(bytes); }
  unsigned char* data() { return mem_.get(); }
  size_t size() const { return len_; }

 private:
  std::unique_ptr<unsigned char[]> mem_;
  size_t cap_ = 0, len_ = 0;
  void grow(size_t need);
};

double now_ms();

hipError_t device_copy(void* dst, const void* src, size_t bytes, hipStream_t st);

// Dense shared-A products on split bf16 MFMA (dense_gemm.hip), binary32
// feature-major right-hand sides (rows x ldb, ldb a multiple of 64):
//   dense_at_pair  U, V planes (splits x n x ldb, stride `plane`) of A'Y, A'W
//                  from A' split copies ATh / ATl (n64 x m64)
//   dense_a        Z = A X from A split copies ARh / ARl (m64 x n64)
//   dense_at_splits  the K split dense_at_pair uses for a width
int dense_at_splits(int n, int m64, int ldb);
// rows of the split copies of A' (n) and A (m) padded to the products' tile
// height (the kernels read whole tiles)
int dense_row_pad(int rows);
hipError_t dense_at_pair(hipStream_t st, int m, int n, int ldb, int m64, const __bf16* ATh, const __bf16* ATl,
                         const float* Y, const float* W, float* Up, float* Vp, size_t plane, int splits);
hipError_t dense_a(hipStream_t st, int m, int n, int ldb, int n64, const __bf16* ARh, const __bf16* ARl,
                   const float* X, float* Z);
// Setup of a batch's instances on the device (batch_setup.hip): groups, then
// instances, on stream st; rec / rec_cls are the instance records in instance
// and class order (lambda and the stop-test scales are written into both).
hipError_t batch_setup_launch(const SetupGroup* d_groups, int ngroups, const SetupInst* d_insts, int ninsts,
                              WgInst* rec, WgInst* rec_cls, bool fp64, hipStream_t st);
// solo: every workgroup of the launch has a CU of its own (the kernel may use
// more registers per lane; 512-thread launches only).
hipError_t wg_launch(const WgInst* d_insts, int count, int lds_bytes, int threads, bool fp64, bool direct, int* queue,
                     int grid, bool solo, hipStream_t st);

// Whole-GPU engine for one large instance (both plugins), prepared once and
// re-runnable from the prepared start.
class GridSolver {
 public:
  virtual ~GridSolver() = default;
  virtual bool run(std::string& err) = 0;
  // the scaled x, s, y into S, the result and (VERBOSE) the residual history
  virtual bool fetch(Setup& S, DevResult& r, std::vector<double>* hist, std::string& err) = 0;
  double upload_ms = 0, solve_ms = 0, kernel_ms = 0;
};

// F: the direct plugin's factorization (nullptr: indirect plugin).  borrow:
// the calling thread's device arena and stream (one pdos() call) instead of
// memory and a stream of its own (a batch member).
std::unique_ptr<GridSolver> grid_prepare(const Setup& S, const Params& par, bool fp64, const Factor* F, int device,
                                         bool borrow, std::string& err);

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
  // Members whose iterates exceed a workgroup's LDS run on the whole-GPU
  // engine, one after another, after the workgroup launches.
  std::vector<int> big;
  std::vector<std::unique_ptr<GridSolver>> gsol;
  std::vector<char> is_big;
  // per instance: LDS bytes of the streaming (x), resident (y) and
  // image-in-device-memory (z) variants, 0 where the image variant is
  // unavailable; s: the streaming variant's matrix copies were packed
  struct VarLds { int x, y, z; bool s; };
  std::vector<VarLds> vlds;
  int variant_env = -1, persist_env = -1, threads_env = 0;
  bool pack_all = false;  // pack the streaming copies even where an image fits (the tuner's samples)
  // Launch classes: instances grouped by how many of their workgroups fit
  // in a CU's LDS, each class launched with its own LDS size on its own
  // stream; the device holds a second copy of the records in class order.
  // A class is (workgroups per CU, launch mode: a workgroup per instance or
  // persistent workgroups draining a work queue); each record names its own
  // kernel variant (streaming, resident, image in device memory).
  static constexpr int kOccMax = 4;
  static constexpr int kVariants = 3;
  static constexpr int kSizes = 3;  // threads per workgroup: 256, 512, 1,024 (image variants)
  static constexpr int kClasses = 2 * kSizes * kOccMax;
  std::vector<int> order;                 // class order -> instance
  int ncls = 0, cls_start[kClasses + 1] = {0}, cls_lds[kClasses] = {0};
  int cls_persist[kClasses] = {0}, cls_occ[kClasses] = {0}, cls_threads[kClasses] = {0};
  // Classes run concurrently on at most kStreams streams (the device's
  // hardware queues: more streams would share queues and serialize behind
  // each other anyway); cls_seq lists the classes in launch order,
  // cls_stream gives each its stream (0 = `stream`, s > 0 = xs[s - 1]).
  static constexpr int kStreams = 4;
  int cls_seq[kClasses] = {0}, cls_stream[kClasses] = {0};
  size_t cinst_off = 0, queue_off = 0;
  // tuner choices per instance (empty: automatic): kernel variant (-1
  // automatic), launch mode (1 persistent), entries per lane of the wave
  // batches
  std::vector<signed char> var_force, persist, epl;
  std::vector<short> wgsize;  // threads per workgroup of an image-variant record (empty: 256)
  std::string tuning;  // the tuner's decisions (JSON), empty if it did not run
  bool warm = false;   // run at least once (the tuner's untimed first run)
  hipStream_t xs[kClasses - 1] = {};
  hipEvent_t xev[kClasses - 1] = {};
  bool upload_records(std::string& err);
  void set_shapes();
  // Chooses var_force / persist / epl per size bucket by timing candidate
  // launches on samples of the batch (before prepare).
  bool tune(int device, bool fp64, std::string& err);
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

int wg_variant_env();

bool solve_wg(std::vector<Setup>& S, std::vector<Factor>& F, const std::vector<const Params*>& par,
              bool fp64, bool direct, int device, std::vector<DevResult>& res, double& upload_ms,
              double& solve_ms, std::string& err, std::vector<double>* hist0 = nullptr);

// Whether one instance goes to the whole-GPU engine (its iterates exceed a
// workgroup's LDS, or it is large enough to fill the GPU).
bool grid_preferred(const Setup& S, bool fp64, bool direct);
// One solve on the whole-GPU engine; hist (optional): with par.VERBOSE, the
// residuals of checks 0, 10, 20, ... (5 per row, verbose_hist_rows(MAX_ITERS)
// rows)
bool solve_grid(Setup& S, const Params& par, bool fp64, const Factor* F, int device, std::vector<DevResult>& res,
                double& upload_ms, double& solve_ms, std::string& err, std::vector<double>* hist = nullptr);

}  // namespace pdos_amd
