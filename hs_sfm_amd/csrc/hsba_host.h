// hsba_host.h — host-side problem layout and the device engine.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/hsba.h"
#include "hsba_kernels.h"

namespace hsba {

struct HostWindow {
  int64_t P = 0, I = 0, C = 0, N = 0;
  int64_t x_pt = 0, x_img = 0, x_cam = 0, x_size = 0, y_size = 0;
  std::vector<int64_t> cons_xidx;  // window-local x index of each constraint row
  bool unobserved = false;         // a free image / camera with no key
};

// Host copy of everything the kernels index; built once per problem / batch.
struct Layout {
  int IMG = 0, D = 0, E = 0;
  uint32_t comp = 0, free_bits = 0, fix = 0;
  int fidx[kFull] = {0};
  bool pts_free = true, imgs_free = true, cams_free = true, has_cov = false;
  int nwin = 0;
  std::vector<HostWindow> hw;
  std::vector<WinInfo> wi;
  std::vector<int32_t> key_img, key_pt, key_orig;
  std::vector<int32_t> pt_ptr, pt_win;
  std::vector<int32_t> img_ptr, img_keys, img_win, img_cam, img_row;
  // image-key chunks of at most kImgChunk keys: chunk -> image, begin / end in
  // img_keys; img_ch_ptr[i] .. img_ch_ptr[i+1] are image i's chunks in order
  std::vector<int32_t> ich_img, ich_begin, ich_end, img_ch_ptr;
  std::vector<int32_t> key_ipos, ikey_pt;  // PCG: key -> image-major row, row -> point
  std::vector<int32_t> cam_ptr, cam_imgs, cam_win, cam_row;
  std::vector<int64_t> pair_ptr;
  std::vector<uint64_t> inc;
  std::vector<int32_t> pair_win, pair_img;
  std::vector<int64_t> blk_ptr;
  std::vector<Contrib> contrib;
  std::vector<int32_t> row_blk, row_sub;
  std::vector<ConstraintRow> cons;
  std::vector<int32_t> chunk_kind, chunk_begin, chunk_end, chunk_win;
  std::vector<double> cam_base;    // [C][10] held / identity values
  std::vector<double> fix_pts, fix_imgs, fix_cam_free;
  std::vector<int> small_list, large_list;
  int64_t P = 0, I = 0, C = 0, N = 0, npairs = 0;
  int64_t S_total = 0, rows_total = 0, dinv_total = 0, lower_total = 0;
  int64_t incidences = 0;
  bool pcg = false;       // implicit Schur: diagonal pairs only, no dense reduced matrix
};

int build_layout(const hsba_problem_desc *descs, int64_t count, Layout &L, std::string *why,
                 bool pcg = false);
int64_t layout_device_bytes(const Layout &L, int precision);
// Upper bound of the device bytes an engine allocates for one problem, from
// its description alone (one O(keys) pass over the observations).
int64_t estimate_device_bytes(const hsba_problem_desc &desc, int precision, bool pcg, bool has_cov);
// x and y lengths of one problem in the reference layouts (before validation).
int problem_sizes(const hsba_problem_desc &desc, int64_t *x_size, int64_t *y_size);
// Index structure of window 0 (hsba_problem_structure).
int export_structure(const Layout &L, hsba_structure_sizes *sizes, int32_t *key_order, int32_t *point_ptr,
                     int32_t *point_image, int32_t *point_camera, int32_t *image_camera, int32_t *pairs,
                     int64_t *pair_ptr, int32_t *incidences);
// The checks build_layout makes on one problem alone (sizes, id ranges,
// duplicate keys, constraint records), without building anything: lets the
// serving loop refuse a malformed job at submit instead of failing its batch.
int validate_desc(const hsba_problem_desc &desc, std::string *why);

// Runs f() and maps an escaping C++ exception to an error code, so that no
// exception crosses an extern "C" entry point.
template <class F>
int guarded(F &&f) noexcept {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    return HSBA_ERR_OOM;
  } catch (...) {
    return HSBA_ERR_INVALID;
  }
}

// Autotuner table (hsba_tune.cpp).  Callers hold tune_mutex() around a
// lookup, the timing of the candidates and the store, so one class is
// measured once even when several engines are created concurrently.
std::mutex &tune_mutex();
constexpr int kPairVariants = 3;  // pair kernels the autotuner chooses between
bool tune_lookup_locked(int precision, int E, int bucket, int *variant);
void tune_store_locked(int precision, int E, int bucket, int variant, const double ms[kPairVariants],
                       int64_t pairs);

// First-fit sub-allocator over one device allocation, with coalescing of
// freed neighbours (256-byte granules).  Thread-safe.  The serving loop
// allocates its HBM budget once and every job's buffers come from it, so no
// hipMalloc / hipFree (which synchronise the device) runs while jobs execute.
class DeviceArena {
 public:
  ~DeviceArena();
  int init(int64_t bytes);
  void *alloc(int64_t bytes);  // nullptr when no free block is large enough
  void release(void *p);
  int64_t capacity() const { return cap_; }
  int64_t in_use();

 private:
  std::mutex mu_;
  char *base_ = nullptr;
  int64_t cap_ = 0, used_ = 0;
  std::map<int64_t, int64_t> free_;  // offset -> bytes
  std::map<int64_t, int64_t> live_;  // offset -> bytes
};

struct EngineBase {
  virtual ~EngineBase() {}
  Layout L;
  DeviceArena *arena = nullptr;  // buffers come from here when set
  int precision = HSBA_PRECISION_MIXED;
  int solver = HSBA_SOLVER_DENSE;
  double pcg_rtol = 1e-6;
  int pcg_max_iterations = 500;
  int loss = HSBA_LOSS_TRIVIAL;
  double loss_scale = 1.0;
  int glog2 = 3;
  // K1 lanes per point from each point's own track length (2 / 4 / 8); false when
  // the caller fixed point_group_lanes (then glog2 for every point)
  bool lin_by_track = true;
  int pair_kernel = 0;  // HSBA_PAIRS_*: 0 tuned, else the forced pair kernel
  virtual int init(hipStream_t stream) = 0;
  virtual int upload_x(const double *const *xs) = 0;
  virtual int upload_obs(const double *const *ys, const hsba_ycov_inverse *ycovs) = 0;
  virtual int download_x(double *const *xs) = 0;
  // Full LM solve; returns the worst per-window status (or a device error).
  virtual int lm(const hsba_lm_options *lm, hsba_lm_stats *stats) = 0;
  // The same solve for callers that read each window's status from stats:
  // returns only device / allocation errors, never a window's own failure.
  virtual int lm_run(const hsba_lm_options *lm, hsba_lm_stats *stats) = 0;
  virtual int lm_begin(const hsba_lm_options *lm) = 0;
  virtual int lm_iterate(int iterations, hsba_stage_times *times) = 0;
  virtual int lm_end(hsba_lm_stats *stats) = 0;
  virtual int build_normal(double *V, double *Ui, double *Uc, double *g) = 0;
  virtual int download_normal(double *V, double *Ui, double *Uc, double *g) = 0;
  virtual int solve(double mu, double *delta, double *S, double *rhs) = 0;
  // scratch: null = device buffers are allocated for the call; else the
  // fixed scratch region (no allocation; sizes bounded by kScratchMaxLd /
  // kScratchMaxPoints, HSBA_ERR_TOO_LARGE beyond them)
  virtual int covariance(double *Vp, double *Ui, double *Uc, double *reduced, double *scratch) = 0;
  // dense pseudo-inverse of N at the last build (window 0), [x_size]^2 in the
  // x layout (hsba_x_covariance)
  virtual int x_covariance(double *out, int method, hsba_covariance_info *info) = 0;
  // coupling blocks W_pi / W_pc / W_ic of the last build (window 0)
  virtual int coupling(double *pi, double *pc, double *ic) = 0;
  virtual int evaluate(double *y) = 0;
  virtual int jacobian(double *jp, double *ji, double *jc) = 0;
  virtual int scalars(double *maxdiag, double *ginf) = 0;
  virtual int cost(double *value) = 0;
  virtual hipStream_t stream() const = 0;
};

EngineBase *make_engine(int precision);

// Whether the automatic solver choice takes PCG for this problem (reduced
// system beyond 4096 unknowns).
bool auto_pcg(const hsba_problem_desc &desc);
// Builds the layout and an initialised engine for `count` problems on
// `stream` (nullptr: the engine creates its own), with buffers from `arena`
// when given.
int create_engine(const hsba_problem_desc *descs, int64_t count, const hsba_options *options,
                  hipStream_t stream, DeviceArena *arena, std::unique_ptr<EngineBase> &out);

}  // namespace hsba
