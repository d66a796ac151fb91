// hsba_engine.cpp — device-resident state of a problem / batch and the LM
// driver.  One LM iteration is a fixed sequence of launches whose work is
// gated by per-window flags on the device (WinState), so the sequence is
// captured once into a hipGraph and replayed max_iterations times with no
// host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hsba_host.h"

namespace hsba {

namespace {

#define HCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      std::fprintf(stderr, "hsba: %s failed: %s\n", #expr, hipGetErrorString(e_));    \
      return e_ == hipErrorOutOfMemory ? HSBA_ERR_OOM : HSBA_ERR_DEVICE;               \
    }                                                                                  \
  } while (0)

// Device buffer of n elements, from the arena when the engine has one (the
// serving loop's pre-allocated budget), else from hipMalloc.
template <typename X>
int dalloc(X *&p, int64_t n, std::vector<void *> &owned, DeviceArena *arena) {
  p = nullptr;
  if (n <= 0) n = 1;
  if (arena) {
    p = (X *)arena->alloc((int64_t)sizeof(X) * n);
    if (!p) return HSBA_ERR_OOM;
  } else {
    HCHK(hipMalloc((void **)&p, sizeof(X) * (size_t)n));
  }
  owned.push_back((void *)p);
  return HSBA_OK;
}

// Device buffers of one call (hipMalloc), freed when the call returns.
struct CallBuffers {
  std::vector<void *> ptrs;
  ~CallBuffers() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename X>
  int get(X *&p, int64_t n) {
    p = nullptr;
    if (hipMalloc((void **)&p, sizeof(X) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return HSBA_ERR_OOM;
    ptrs.push_back((void *)p);
    return HSBA_OK;
  }
};

// Singular values at or below this are dropped from the pseudo-inverse
// (camera_shared_analytical_x_covariance_calculator.hpp:78).
constexpr double kPinvThreshold = 1e-6;
constexpr int kJacobiMaxSweeps = 64;

template <typename X>
int upload(X *dst, const std::vector<X> &src, hipStream_t st) {
  if (src.empty()) return HSBA_OK;
  HCHK(hipMemcpyAsync(dst, src.data(), sizeof(X) * src.size(), hipMemcpyHostToDevice, st));
  return HSBA_OK;
}

}  // namespace

template <typename T>
struct Engine final : EngineBase {
  KernelSet<T> *ks = nullptr;
  Dev<T> d;
  std::vector<void *> owned;
  hipStream_t st = nullptr;
  bool own_stream = false;
  hipGraphExec_t gexec = nullptr;  // `graph_iters` LM iterations in one graph
  hipGraph_t graph = nullptr;
  int graph_iters = 0;
  static constexpr int kMaxGraphIterations = 64;
  int *small_list_dev = nullptr;
  CholWin *large_wins_dev = nullptr;
  // K1 point classes by track length: <= 6 keys 2 lanes, <= 12 keys 4 lanes,
  // longer 8 lanes (the lanes of a point then follow from its own track, so a
  // window alone and in a batch sums each point in the same order)
  static constexpr int kLinClasses = 3;
  int32_t *lin_list = nullptr;
  int lin_cnt[kLinClasses] = {0, 0, 0};
  std::vector<int32_t> lin_host;
  std::vector<CholWin> large_wins_host;  // source of the upload, kept while it is in flight
  CholPlan plan;
  std::vector<WinState> ws_host;
  bool built = false;  // a linearization is resident

  ~Engine() override {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    drop_stage_graphs();
    for (int i = 0; i <= kStages; ++i)
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    for (int i = 0; i < 3; ++i) {
      if (pg_exec[i]) (void)hipGraphExecDestroy(pg_exec[i]);
      if (pg_graph[i]) (void)hipGraphDestroy(pg_graph[i]);
    }
    if (st) (void)hipStreamSynchronize(st);
    for (int i = 0; i < kPairVariants; ++i) {
      if (pair_st[i]) (void)hipStreamSynchronize(pair_st[i]), (void)hipStreamDestroy(pair_st[i]);
      if (pair_ev[i]) (void)hipEventDestroy(pair_ev[i]);
    }
    if (pcg_host) (void)hipHostFree(pcg_host);
    for (void *p : owned) {
      if (arena) arena->release(p);
      else (void)hipFree(p);
    }
    if (own_stream && st) (void)hipStreamDestroy(st);
    delete ks;
  }

  hipStream_t stream() const override { return st; }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    drop_stage_graphs();
    drop_pcg_graphs();
  }
  void drop_pcg_graphs() {
    for (int i = 0; i < 3; ++i) {
      if (pg_exec[i]) (void)hipGraphExecDestroy(pg_exec[i]);
      if (pg_graph[i]) (void)hipGraphDestroy(pg_graph[i]);
      pg_exec[i] = nullptr;
      pg_graph[i] = nullptr;
    }
  }
  void drop_stage_graphs() {
    for (int i = 0; i < kStages; ++i) {
      if (sgexec[i]) (void)hipGraphExecDestroy(sgexec[i]);
      if (sgraph[i]) (void)hipGraphDestroy(sgraph[i]);
      sgexec[i] = nullptr;
      sgraph[i] = nullptr;
    }
  }

  int init(hipStream_t stream_in) override {
    ks = make_kernel_set<T>(L.IMG, L.D);
    if (!ks) return HSBA_ERR_UNSUPPORTED;
    ks->pcg_params(pcg_rtol, pcg_max_iterations);
    if (stream_in) st = stream_in;
    else { HCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)); own_stream = true; }
    std::memset(&d, 0, sizeof(d));
    const int E = L.E, D = L.D;
    d.P = (int32_t)L.P; d.I = (int32_t)L.I; d.C = (int32_t)L.C; d.N = (int32_t)L.N;
    d.nwin = L.nwin; d.npairs = (int32_t)L.npairs; d.ncons = (int32_t)L.cons.size();
    d.nchunks = (int32_t)L.chunk_kind.size();
    d.lower_max = 0;
    for (const WinInfo &q : L.wi) d.lower_max = std::max<int64_t>(d.lower_max, (int64_t)q.n * (q.n + 1) / 2);
    d.pts_free = L.pts_free; d.imgs_free = L.imgs_free; d.cams_free = L.cams_free; d.has_cov = 0;
    d.pcg_rows = L.pcg ? (int32_t)L.rows_total : 0;
    d.loss = loss;
    d.loss_scale = loss_scale;
    for (int j = 0; j < kFull; ++j) d.fidx[j] = L.fidx[j];
    int rc = 0;
#define A(ptr, n) if ((rc = dalloc(ptr, (n), owned, arena))) return rc
    int32_t *key_img, *key_pt, *key_orig, *pt_ptr, *pt_win, *img_ptr, *img_keys, *img_win, *img_cam, *img_row;
    int32_t *cam_ptr, *cam_imgs, *cam_win, *cam_row, *pair_win, *pair_img, *row_blk, *row_sub;
    int32_t *ck, *cbg, *cen, *cw;
    int64_t *pair_ptr, *blk_ptr;
    uint64_t *inc;
    Contrib *contrib;
    ConstraintRow *cons;
    WinInfo *wi;
    A(key_img, L.N); A(key_pt, L.N); A(key_orig, L.N);
    A(pt_ptr, L.P + 1); A(pt_win, L.P);
    A(img_ptr, L.I + 1); A(img_keys, L.N); A(img_win, L.I); A(img_cam, L.I); A(img_row, L.I);
    A(cam_ptr, L.C + 1); A(cam_imgs, L.I); A(cam_win, L.C); A(cam_row, L.C);
    A(pair_ptr, L.npairs + 1); A(inc, L.incidences); A(pair_win, L.npairs); A(pair_img, L.npairs);
    A(blk_ptr, (int64_t)L.blk_ptr.size()); A(contrib, (int64_t)L.contrib.size());
    A(row_blk, (int64_t)L.row_blk.size()); A(row_sub, (int64_t)L.row_sub.size());
    A(cons, (int64_t)L.cons.size());
    A(ck, d.nchunks); A(cbg, d.nchunks); A(cen, d.nchunks); A(cw, d.nchunks);
    A(wi, L.nwin);
    A(d.ws, L.nwin);
    // observations and the whitening factors of their inverse covariances
    // (allocated up front, so that no later upload allocates)
    double *key_uv; A(key_uv, 2 * L.N);
    double *key_cov; A(key_cov, 3 * L.N);
    A(d.pts, 3 * L.P); A(d.pts_new, 3 * L.P); A(d.dpt, 3 * L.P);
    A(d.V, 6 * L.P); A(d.gp, 3 * L.P); A(d.Linv, 6 * L.P); A(d.q, 3 * L.P); A(d.tpt, 3 * L.P);
    A(d.imgs, 6 * L.I); A(d.imgs_new, 6 * L.I); A(d.imgq, L.I); A(d.imgq_new, L.I);
    A(d.Dimg, (int64_t)E * E * L.I); A(d.gr, (int64_t)E * L.I); A(d.rp, (int64_t)E * L.I);
    A(d.cam_free, (int64_t)std::max(D, 1) * L.C); A(d.cam_free_new, (int64_t)std::max(D, 1) * L.C);
    A(d.cam_full, kFull * L.C); A(d.cam_full_new, kFull * L.C);
    A(d.Uc, (int64_t)D * D * L.C); A(d.gc, (int64_t)D * L.C);
    A(d.W, 3LL * E * L.N); A(d.Z, (int64_t)z_stride(E, sizeof(T)) * L.N);
    d.Zi = nullptr; d.key_ipos = nullptr; d.ikey_pt = nullptr;
    if (L.pcg && L.pts_free && E > 0) {
      int32_t *kip, *ikp;
      A(d.Zi, (int64_t)z_stride(E, sizeof(T)) * L.N); A(kip, L.N); A(ikp, L.N);
      if ((rc = upload(kip, L.key_ipos, st)) || (rc = upload(ikp, L.ikey_pt, st))) return rc;
      d.key_ipos = kip; d.ikey_pt = ikp;
    }
    A(d.pair_out, (int64_t)E * E * L.npairs);
    A(d.S, L.S_total); A(d.rhs, L.rows_total); A(d.dred, L.rows_total); A(d.Dinv, L.dinv_total);
    A(d.partial, 4LL * d.nchunks);
    A(d.pv, L.rows_total); A(d.pr, L.rows_total); A(d.pz, L.rows_total); A(d.pp, L.rows_total);
    A(d.pAp, L.rows_total); A(d.Minv, (int64_t)E * E * L.I);
    A(d.tpd, 3 * L.P); A(d.pcg, L.nwin); A(d.pdot, 3 * L.I);
    d.nich = (int32_t)L.ich_img.size();
    {
      int32_t *ich_img, *ich_begin, *ich_end, *img_ch_ptr;
      A(ich_img, L.ich_img.size()); A(ich_begin, L.ich_begin.size()); A(ich_end, L.ich_end.size());
      A(img_ch_ptr, L.img_ch_ptr.size());
      A(d.ipart, (int64_t)(E * (E + 1) / 2 + E) * std::max<int64_t>(1, L.ich_img.size()));
      if ((rc = upload(ich_img, L.ich_img, st)) || (rc = upload(ich_begin, L.ich_begin, st)) ||
          (rc = upload(ich_end, L.ich_end, st)) || (rc = upload(img_ch_ptr, L.img_ch_ptr, st)))
        return rc;
      d.ich_img = ich_img; d.ich_begin = ich_begin; d.ich_end = ich_end; d.img_ch_ptr = img_ch_ptr;
    }
    A(small_list_dev, (int64_t)L.small_list.size());
    lin_host.clear();
    for (int c = 0; c < kLinClasses; ++c) lin_cnt[c] = 0;
    if (lin_by_track) {
      for (int c = 0; c < kLinClasses; ++c)
        for (int64_t p = 0; p < L.P; ++p) {
          const int len = L.pt_ptr[p + 1] - L.pt_ptr[p];
          if ((len <= 6 ? 0 : len <= 12 ? 1 : 2) == c) {
            lin_host.push_back((int32_t)p);
            ++lin_cnt[c];
          }
        }
      A(lin_list, (int64_t)lin_host.size());
    }
    A(large_wins_dev, (int64_t)L.large_list.size());
    large_wins_host.clear();
    for (int w : L.large_list) {
      const WinInfo &q = L.wi[w];
      large_wins_host.push_back(CholWin{q.s_off, q.dinv_off, q.ld, q.nb, q.n, q.r_off, w, 0});
    }
#undef A
    if ((rc = upload(key_img, L.key_img, st)) || (rc = upload(key_pt, L.key_pt, st)) ||
        (rc = upload(key_orig, L.key_orig, st)) || (rc = upload(pt_ptr, L.pt_ptr, st)) ||
        (rc = upload(pt_win, L.pt_win, st)) || (rc = upload(img_ptr, L.img_ptr, st)) ||
        (rc = upload(img_keys, L.img_keys, st)) || (rc = upload(img_win, L.img_win, st)) ||
        (rc = upload(img_cam, L.img_cam, st)) || (rc = upload(img_row, L.img_row, st)) ||
        (rc = upload(cam_ptr, L.cam_ptr, st)) || (rc = upload(cam_imgs, L.cam_imgs, st)) ||
        (rc = upload(cam_win, L.cam_win, st)) || (rc = upload(cam_row, L.cam_row, st)) ||
        (rc = upload(pair_ptr, L.pair_ptr, st)) || (rc = upload(inc, L.inc, st)) ||
        (rc = upload(pair_win, L.pair_win, st)) || (rc = upload(pair_img, L.pair_img, st)) ||
        (rc = upload(blk_ptr, L.blk_ptr, st)) || (rc = upload(contrib, L.contrib, st)) ||
        (rc = upload(row_blk, L.row_blk, st)) || (rc = upload(row_sub, L.row_sub, st)) ||
        (rc = upload(cons, L.cons, st)) || (rc = upload(ck, L.chunk_kind, st)) ||
        (rc = upload(cbg, L.chunk_begin, st)) || (rc = upload(cen, L.chunk_end, st)) ||
        (rc = upload(cw, L.chunk_win, st)) || (rc = upload(wi, L.wi, st)) ||
        (rc = upload(small_list_dev, L.small_list, st)) || (rc = upload(large_wins_dev, large_wins_host, st)) ||
        (lin_by_track && (rc = upload(lin_list, lin_host, st))))
      return rc;
    d.key_img = key_img; d.key_pt = key_pt; d.key_orig = key_orig; d.key_uv = key_uv; d.key_cov = key_cov;
    d.pt_ptr = pt_ptr; d.pt_win = pt_win;
    d.img_ptr = img_ptr; d.img_keys = img_keys; d.img_win = img_win; d.img_cam = img_cam; d.img_row = img_row;
    d.cam_ptr = cam_ptr; d.cam_imgs = cam_imgs; d.cam_win = cam_win; d.cam_row = cam_row;
    d.pair_ptr = pair_ptr; d.inc = inc; d.pair_win = pair_win; d.pair_img = pair_img;
    d.wi = wi; d.blk_ptr = blk_ptr; d.contrib = contrib; d.row_blk = row_blk; d.row_sub = row_sub;
    d.cons = cons; d.chunk_kind = ck; d.chunk_begin = cbg; d.chunk_end = cen; d.chunk_win = cw;
    d.win_of_chunk = cw;
    // fixed blocks and camera base vectors
    if (!L.pts_free && (rc = upload(d.pts, L.fix_pts, st))) return rc;
    if (!L.pts_free && (rc = upload(d.pts_new, L.fix_pts, st))) return rc;
    if (!L.imgs_free && (rc = upload(d.imgs, L.fix_imgs, st))) return rc;
    if (!L.imgs_free) ks->image_rotations(d, st);
    if (!L.imgs_free && (rc = upload(d.imgs_new, L.fix_imgs, st))) return rc;
    if (!L.cams_free) {
      std::vector<double> full = L.cam_base;
      const int dd = __builtin_popcount(L.free_bits);
      for (int64_t c = 0; c < L.C; ++c)
        for (int q = 0; q < dd; ++q) full[c * kFull + L.fidx[q]] = L.fix_cam_free[c * dd + q];
      if ((rc = upload(d.cam_full, full, st)) || (rc = upload(d.cam_full_new, full, st))) return rc;
    }
    plan.nwin = L.nwin; plan.wi = wi; plan.ws = d.ws; plan.S = d.S; plan.rhs = d.rhs; plan.dred = d.dred;
    plan.Dinv = d.Dinv; plan.wi_host = L.wi.data(); plan.large_wins = large_wins_dev;
    plan.Lf = nullptr; plan.bwork = nullptr; plan.ywork = nullptr;
    {
      int64_t large_s = 0;
      for (int w : L.large_list) large_s = std::max<int64_t>(large_s, L.wi[w].s_off + (int64_t)L.wi[w].ld * L.wi[w].ld);
      if ((rc = dalloc(plan.Lf, large_s, owned, arena)) || (rc = dalloc(plan.bwork, L.rows_total, owned, arena)) ||
          (rc = dalloc(plan.ywork, L.rows_total, owned, arena)))
        return rc;
    }
    HCHK(hipMemsetAsync(d.S, 0, sizeof(double) * (size_t)std::max<int64_t>(1, L.S_total), st));
    if (!L.pcg) chol_init_padding(plan, st);  // no dense reduced matrix in PCG mode
    HCHK(hipGetLastError());
    ws_host.assign(L.nwin, WinState());
    HCHK(hipStreamSynchronize(st));
    return setup_pairs();
  }

  // ---- K4 kernel choice (autotuner) -------------------------------------
  // Pairs run by each pair kernel (index 0 VALU, 1 matrix cores); a list that
  // covers every pair is passed as null (pairs 0 .. npairs).
  int32_t *plist_dev[kPairVariants] = {nullptr, nullptr, nullptr};
  int plist_n[kPairVariants] = {0, 0, 0};
  bool plist_all[kPairVariants] = {false, false, false};

  // Pair kernels of the variants in use.  Each writes its own pairs' blocks,
  // so with more than one variant (a batch whose window classes took
  // different kernels) they run as concurrent branches: the others on helper
  // streams forked from and joined back to st by events (branches of the
  // captured graph), the one with the most pairs on st.
  hipStream_t pair_st[kPairVariants] = {};
  hipEvent_t pair_ev[kPairVariants] = {};
  bool pair_fork = false;
  void enqueue_pairs() {
    int order[kPairVariants], na = 0;
    for (int v = 0; v < kPairVariants; ++v)
      if (plist_n[v]) order[na++] = v;
    std::sort(order, order + na, [&](int a, int b) { return plist_n[a] > plist_n[b]; });
    if (na <= 1 || !pair_fork) {
      for (int i = 0; i < na; ++i) ks->pairs(d, st, order[i], plist_all[order[i]] ? nullptr : plist_dev[order[i]],
                                              plist_n[order[i]]);
      return;
    }
    (void)hipEventRecord(pair_ev[0], st);
    for (int i = 1; i < na; ++i) {
      const int v = order[i];
      (void)hipStreamWaitEvent(pair_st[i], pair_ev[0], 0);
      ks->pairs(d, pair_st[i], v, plist_all[v] ? nullptr : plist_dev[v], plist_n[v]);
      (void)hipEventRecord(pair_ev[i], pair_st[i]);
    }
    ks->pairs(d, st, order[0], plist_all[order[0]] ? nullptr : plist_dev[order[0]], plist_n[order[0]]);
    for (int i = 1; i < na; ++i) (void)hipStreamWaitEvent(st, pair_ev[i], 0);
  }

  // Average device time (ms) of `reps` launches of pair kernel `variant` over
  // the device list `dl` of `n` pairs, after one untimed launch.
  int time_pairs(int variant, const int32_t *dl, int n, int reps, double *ms) {
    hipEvent_t e0, e1;
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    ks->pairs(d, st, variant, dl, n);
    HCHK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) ks->pairs(d, st, variant, dl, n);
    HCHK(hipEventRecord(e1, st));
    HCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / reps;
    return HSBA_OK;
  }

  // Chooses the pair kernel of every window.  A window's class is (precision,
  // E, floor(log2(mean incidences per pair))); the first engine that meets a
  // class with two candidate kernels times both on that class's pairs (zeroed
  // Z rows, every window in a trial) and the process-wide table keeps the
  // fastest one.  The two matrix-core kernels are candidates for E <= 15 in
  // mixed precision (the FP64 path keeps the register-tile kernel for E <= 6).
  int setup_pairs() {
    if (!L.pts_free || L.E == 0 || L.npairs == 0) return HSBA_OK;
    std::vector<int64_t> inc_w(L.nwin, 0), np_w(L.nwin, 0);
    for (int64_t q = 0; q < L.npairs; ++q) {
      const int w = L.pair_win[q];
      np_w[w] += 1;
      inc_w[w] += L.pair_ptr[q + 1] - L.pair_ptr[q];
    }
    std::map<int, std::vector<int>> classes;  // bucket -> windows
    for (int w = 0; w < L.nwin; ++w) {
      if (!np_w[w]) continue;
      const double mean = (double)inc_w[w] / (double)np_w[w];
      int b = 0;
      while (b < 40 && std::ldexp(1.0, b + 1) <= mean) ++b;
      classes[b].push_back(w);
    }
    const bool mfma = ks->has_pair_variant(1) && precision == HSBA_PRECISION_MIXED;
    std::vector<int> wvar(L.nwin, 0);
    {
      std::lock_guard<std::mutex> g(tune_mutex());
      for (const auto &kv : classes) {
        int var = 0;
        if (pair_kernel == HSBA_PAIRS_VALU || !mfma) {
          var = 0;
        } else if (pair_kernel == HSBA_PAIRS_MFMA) {
          var = 1;
        } else if (pair_kernel == HSBA_PAIRS_MFMA_WAVE) {
          var = 2;
        } else if (!tune_lookup_locked(precision, L.E, kv.first, &var)) {
          std::vector<char> in(L.nwin, 0);
          for (int w : kv.second) in[w] = 1;
          std::vector<int32_t> cl;
          for (int64_t q = 0; q < L.npairs; ++q)
            if (in[L.pair_win[q]]) cl.push_back((int32_t)q);
          int32_t *dl = nullptr;
          int rc = dalloc(dl, (int64_t)cl.size(), owned, arena);
          if (rc) return rc;
          if ((rc = upload(dl, cl, st))) return rc;
          if ((rc = set_state(nullptr, 0.0, 1, 0, 1))) return rc;
          HCHK(hipMemsetAsync(d.Z, 0, sizeof(T) * (size_t)z_stride(L.E, sizeof(T)) * (size_t)std::max<int64_t>(1, L.N), st));
          HCHK(hipMemsetAsync(d.q, 0, sizeof(T) * 3 * (size_t)std::max<int64_t>(1, L.P), st));
          double ms[kPairVariants] = {0.0, 0.0, 0.0};
          for (int v = 0; v < kPairVariants; ++v)
            if ((rc = time_pairs(v, dl, (int)cl.size(), 3, &ms[v]))) return rc;
          var = 0;
          for (int v = 1; v < kPairVariants; ++v)
            if (ms[v] < ms[var]) var = v;
          tune_store_locked(precision, L.E, kv.first, var, ms, (int64_t)cl.size());
        }
        for (int w : kv.second) wvar[w] = var;
      }
    }
    std::vector<int32_t> lists[kPairVariants];
    for (int64_t q = 0; q < L.npairs; ++q) lists[wvar[L.pair_win[q]]].push_back((int32_t)q);
    for (int v = 0; v < kPairVariants; ++v) {
      plist_n[v] = (int)lists[v].size();
      plist_all[v] = plist_n[v] == L.npairs;
      if (plist_n[v] && !plist_all[v]) {
        int rc = dalloc(plist_dev[v], plist_n[v], owned, arena);
        if (rc) return rc;
        if ((rc = upload(plist_dev[v], lists[v], st))) return rc;
      }
    }
    int na = 0;
    for (int v = 0; v < kPairVariants; ++v) na += plist_n[v] > 0;
    if (na > 1 && !pair_fork) {
      for (int i = 0; i < kPairVariants; ++i) {
        HCHK(hipEventCreateWithFlags(&pair_ev[i], hipEventDisableTiming));
        if (i > 0) HCHK(hipStreamCreateWithFlags(&pair_st[i], hipStreamNonBlocking));
      }
      pair_fork = true;
    }
    HCHK(hipStreamSynchronize(st));
    return HSBA_OK;
  }

  // ---- host <-> device parameter transfer in the reference x layout
  int upload_x(const double *const *xs) override {
    const int D = L.D;
    std::vector<double> pts, imgs, camf, full;
    if (L.pts_free) pts.resize(3 * L.P);
    if (L.imgs_free) imgs.resize(6 * L.I);
    if (L.cams_free) { camf.resize((int64_t)std::max(D, 1) * L.C); full = L.cam_base; }
    for (int w = 0; w < L.nwin; ++w) {
      const HostWindow &h = L.hw[w];
      const WinInfo &wi = L.wi[w];
      const double *x = xs[w];
      if (!x && h.x_size) return HSBA_ERR_INVALID;
      if (L.pts_free) std::copy(x + h.x_pt, x + h.x_pt + 3 * h.P, pts.begin() + 3 * (int64_t)wi.pt_off);
      if (L.imgs_free) std::copy(x + h.x_img, x + h.x_img + 6 * h.I, imgs.begin() + 6 * (int64_t)wi.img_off);
      if (L.cams_free)
        for (int64_t c = 0; c < h.C; ++c)
          for (int q = 0; q < D; ++q) {
            const double v = x[h.x_cam + D * c + q];
            camf[(wi.cam_off + c) * D + q] = v;
            full[(wi.cam_off + c) * kFull + L.fidx[q]] = v;
          }
    }
    int rc = 0;
    if (L.pts_free && ((rc = upload(d.pts, pts, st)) || (rc = upload(d.pts_new, pts, st)))) return rc;
    if (L.imgs_free && ((rc = upload(d.imgs, imgs, st)) || (rc = upload(d.imgs_new, imgs, st)))) return rc;
    if (L.imgs_free) ks->image_rotations(d, st);
    if (L.cams_free && ((rc = upload(d.cam_free, camf, st)) || (rc = upload(d.cam_free_new, camf, st)) ||
                        (rc = upload(d.cam_full, full, st)) || (rc = upload(d.cam_full_new, full, st))))
      return rc;
    built = false;
    return HSBA_OK;
  }

  int download_x(double *const *xs) override {
    const int D = L.D;
    std::vector<double> pts(L.pts_free ? 3 * L.P : 0), imgs(L.imgs_free ? 6 * L.I : 0),
        camf(L.cams_free ? (int64_t)D * L.C : 0);
    if (!pts.empty()) HCHK(hipMemcpyAsync(pts.data(), d.pts, 8 * pts.size(), hipMemcpyDeviceToHost, st));
    if (!imgs.empty()) HCHK(hipMemcpyAsync(imgs.data(), d.imgs, 8 * imgs.size(), hipMemcpyDeviceToHost, st));
    if (!camf.empty()) HCHK(hipMemcpyAsync(camf.data(), d.cam_free, 8 * camf.size(), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    for (int w = 0; w < L.nwin; ++w) {
      const HostWindow &h = L.hw[w];
      const WinInfo &wi = L.wi[w];
      double *x = xs[w];
      if (!x) continue;
      if (L.pts_free) std::copy(pts.begin() + 3 * (int64_t)wi.pt_off, pts.begin() + 3 * ((int64_t)wi.pt_off + h.P), x + h.x_pt);
      if (L.imgs_free) std::copy(imgs.begin() + 6 * (int64_t)wi.img_off, imgs.begin() + 6 * ((int64_t)wi.img_off + h.I), x + h.x_img);
      if (L.cams_free) std::copy(camf.begin() + (int64_t)D * wi.cam_off, camf.begin() + (int64_t)D * (wi.cam_off + h.C), x + h.x_cam);
    }
    return HSBA_OK;
  }

  int upload_obs(const double *const *ys, const hsba_ycov_inverse *ycovs) override {
    std::vector<double> uv(2 * L.N);
    bool any_cov = false;
    for (int w = 0; w < L.nwin; ++w)
      if (ycovs && ycovs[w].key_blocks) any_cov = true;
    std::vector<double> cov(any_cov ? 3 * L.N : 0);
    for (int w = 0; w < L.nwin; ++w) {
      const WinInfo &wi = L.wi[w];
      const HostWindow &h = L.hw[w];
      const double *y = ys[w];
      if (!y && h.N) return HSBA_ERR_INVALID;
      const double *kb = (ycovs && ycovs[w].key_blocks) ? ycovs[w].key_blocks : nullptr;
      for (int64_t k = wi.key_off; k < wi.key_off + h.N; ++k) {
        const int64_t o = L.key_orig[k] - wi.key_off;
        uv[2 * k] = y[2 * o];
        uv[2 * k + 1] = y[2 * o + 1];
        if (any_cov) {
          double l00 = 1.0, l10 = 0.0, l11 = 1.0;
          if (kb) {
            // Lambda = L L^T; whitening uses L^T = [[l00, l10], [0, l11]]
            const double a = kb[4 * o], b = 0.5 * (kb[4 * o + 1] + kb[4 * o + 2]), c = kb[4 * o + 3];
            if (!(a > 0.0)) return HSBA_ERR_INVALID;
            l00 = std::sqrt(a);
            l10 = b / l00;
            const double t = c - l10 * l10;
            if (!(t > 0.0)) return HSBA_ERR_INVALID;
            l11 = std::sqrt(t);
          }
          cov[3 * k] = l00; cov[3 * k + 1] = l10; cov[3 * k + 2] = l11;
        }
      }
      const double *pb = ycovs ? ycovs[w].point_constraint_blocks : nullptr;
      for (int q = wi.con_begin; q < wi.con_end; ++q) {
        const int local = q - wi.con_begin;
        ConstraintRow &cr = L.cons[q];
        cr.y = y[2 * h.N + local];
        cr.weight = (ycovs && ycovs[w].constraint_weights) ? ycovs[w].constraint_weights[local] : 1.0;
        cr.block = 0;
        cr.brow[0] = cr.brow[1] = cr.brow[2] = 0.0;
        if (pb && cr.kind == 0) {
          // point constraints come first in y order; each must hold all three
          // coordinates, so rows local, local+1, local+2 form one record
          if (cr.rec_mask != 7) return HSBA_ERR_INVALID;
          const int r0 = local - cr.param;
          const double *B = pb + 3 * r0;  // record r0 / 3, row-major 3x3
          for (int j = 0; j < 3; ++j) {
            if (B[3 * cr.param + j] != B[3 * j + cr.param]) return HSBA_ERR_INVALID;  // symmetric
            cr.brow[j] = B[3 * cr.param + j];
          }
          cr.block = 1;
        }
      }
    }
    int rc = upload((double *)d.key_uv, uv, st);
    if (rc) return rc;
    if (any_cov && (rc = upload((double *)d.key_cov, cov, st))) return rc;
    if (d.has_cov != (any_cov ? 1 : 0) && gexec) {
      drop_graph();
    }
    d.has_cov = any_cov ? 1 : 0;
    if (!L.cons.empty() && (rc = upload((ConstraintRow *)d.cons, L.cons, st))) return rc;
    built = false;
    return HSBA_OK;
  }

  // ---- window state helpers
  int set_state(const hsba_lm_options *lm, double mu, int active, int need_lin, int trial, bool upload = true) {
    for (int w = 0; w < L.nwin; ++w) {
      WinState &s = ws_host[w];
      std::memset(&s, 0, sizeof(s));
      s.mu = mu; s.nu = 2.0;
      s.tau = lm ? lm->tau : 0.0; s.eps1 = lm ? lm->eps1 : 0.0; s.eps2 = lm ? lm->eps2 : 0.0;
      s.max_iter = lm ? lm->max_iterations : 0;
      s.active = active; s.need_lin = need_lin; s.trial = trial;
    }
    if (upload) HCHK(hipMemcpyAsync(d.ws, ws_host.data(), sizeof(WinState) * L.nwin, hipMemcpyHostToDevice, st));
    return HSBA_OK;
  }
  int get_state() {
    HCHK(hipMemcpyAsync(ws_host.data(), d.ws, sizeof(WinState) * L.nwin, hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    return HSBA_OK;
  }

  void enqueue_lin_kernel() {
    LinSeg seg{};
    auto whole_waves = [](int64_t threads) { return (int32_t)((threads + 63) / 64 * 64); };
    if (!lin_by_track) {
      // one segment: every point in order, 2^glog2 lanes
      seg.g[0] = seg.g[1] = seg.g[2] = glog2;
      seg.cnt[2] = (int32_t)L.P;
      seg.t0[3] = whole_waves(L.P << glog2);
      ks->linearize(d, st, seg, nullptr);
      return;
    }
    int32_t base = 0;
    for (int c = 0; c < kLinClasses; ++c) {
      seg.g[c] = c + 1;
      seg.base[c] = base;
      seg.cnt[c] = lin_cnt[c];
      seg.t0[c + 1] = seg.t0[c] + whole_waves((int64_t)lin_cnt[c] << (c + 1));
      base += lin_cnt[c];
    }
    ks->linearize(d, st, seg, lin_list);
  }
  void enqueue_linearize() {
    enqueue_lin_kernel();
    ks->image_reduce(d, st);
    ks->camera_reduce(d, st);
    ks->constraints(d, st);
  }
  void enqueue_trial() {
    ks->point_factor(d, st);
    ks->key_z(d, st);
    enqueue_pairs();
    ks->assemble(d, st, L.lower_total, (int)L.rows_total);
    chol_factor_solve_impl(plan, small_list_dev, (int)L.small_list.size(), L.large_list.data(),
                           (int)L.large_list.size(), st);
    ks->backsub(d, st, glog2);
  }

  // ---- PCG form of a trial: setup graph, chunks of PCG iterations until
  // every window has converged (host checks flags between chunks), rest.
  static constexpr int kPcgChunk = 8;
  hipGraphExec_t pg_exec[3] = {};
  hipGraph_t pg_graph[3] = {};
  PcgState *pcg_host = nullptr;  // pinned: the convergence flags read back between chunks

  void enqueue_pcg_part(int part) {
    if (part == 0) {
      ks->lm_control(d, st, 1);
      ks->point_factor(d, st);
      ks->key_z(d, st);
      enqueue_pairs();
      ks->assemble(d, st, 0, (int)L.rows_total);
      ks->pcg_setup(d, st, glog2);
    } else if (part == 1) {
      for (int i = 0; i < kPcgChunk; ++i) ks->pcg_iteration(d, st, glog2);
    } else {
      ks->pcg_finish(d, st);
      ks->backsub(d, st, glog2);
      ks->step(d, st);
      ks->lm_control(d, st, 2);
      ks->cost(d, st, 1);
      ks->lm_control(d, st, 3);
      ks->commit(d, st);
      enqueue_linearize();
      ks->grad_stats(d, st);
      ks->lm_control(d, st, 4);
    }
  }

  int capture_pcg() {
    if (pg_exec[0]) return HSBA_OK;
    for (int part = 0; part < 3; ++part) {
      HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue_pcg_part(part);
      hipError_t e1 = hipStreamEndCapture(st, &pg_graph[part]);
      if (e1 != hipSuccess) return HSBA_ERR_DEVICE;
      HCHK(hipGraphInstantiate(&pg_exec[part], pg_graph[part], nullptr, nullptr, 0));
    }
    if (!pcg_host && hipHostMalloc((void **)&pcg_host, sizeof(PcgState) * std::max(L.nwin, 1)) != hipSuccess) {
      pcg_host = nullptr;
      return HSBA_ERR_OOM;
    }
    return HSBA_OK;
  }

  // Runs the PCG chunks of one trial; returns after every window converged
  // or the iteration budget is spent.
  int pcg_solve_loop() {
    const int chunks = (pcg_max_iterations + kPcgChunk - 1) / kPcgChunk;
    for (int c = 0; c < chunks; ++c) {
      HCHK(hipGraphLaunch(pg_exec[1], st));
      HCHK(hipMemcpyAsync(pcg_host, d.pcg, sizeof(PcgState) * L.nwin, hipMemcpyDeviceToHost, st));
      HCHK(hipStreamSynchronize(st));
      bool all = true;
      for (int w = 0; w < L.nwin; ++w) all = all && pcg_host[w].done;
      if (all) break;
    }
    return HSBA_OK;
  }

  int pcg_iterate(int iterations, hsba_stage_times *times) {
    int rc = capture_pcg();
    if (rc) return rc;
    hipEvent_t e[4] = {};
    if (times)
      for (int i = 0; i < 4; ++i) HCHK(hipEventCreate(&e[i]));
    double acc[3] = {0, 0, 0}, tot = 0;
    for (int it = 0; it < iterations; ++it) {
      if (times) HCHK(hipEventRecord(e[0], st));
      HCHK(hipGraphLaunch(pg_exec[0], st));
      if (times) HCHK(hipEventRecord(e[1], st));
      if ((rc = pcg_solve_loop())) return rc;
      if (times) HCHK(hipEventRecord(e[2], st));
      HCHK(hipGraphLaunch(pg_exec[2], st));
      if (times) {
        HCHK(hipEventRecord(e[3], st));
        HCHK(hipEventSynchronize(e[3]));
        float ms;
        for (int i = 0; i < 3; ++i) { HCHK(hipEventElapsedTime(&ms, e[i], e[i + 1])); acc[i] += ms; }
        HCHK(hipEventElapsedTime(&ms, e[0], e[3]));
        tot += ms;
      }
    }
    HCHK(hipStreamSynchronize(st));
    if (times) {
      for (int i = 0; i < 4; ++i) (void)hipEventDestroy(e[i]);
      std::memset(times, 0, sizeof(*times));
      const double inv = 1.0 / iterations;
      times->point_factor_ms = acc[0] * inv;  // factor, Z, diagonal pairs, rhs, preconditioner
      times->cholesky_ms = acc[1] * inv;      // PCG iterations
      times->step_cost_ms = acc[2] * inv;     // back-substitution, step, cost, relinearization
      times->iteration_ms = tot * inv;
      times->iterations = iterations;
    }
    return HSBA_OK;
  }
  void enqueue_iteration() {
    ks->lm_control(d, st, 1);
    enqueue_trial();
    ks->step(d, st);
    ks->lm_control(d, st, 2);
    ks->cost(d, st, 1);
    ks->lm_control(d, st, 3);
    ks->commit(d, st);
    enqueue_linearize();
    ks->grad_stats(d, st);
    ks->lm_control(d, st, 4);
  }

  int lm(const hsba_lm_options *lm, hsba_lm_stats *stats) override {
    int rc = lm_begin(lm);
    if (rc == HSBA_OK && lm->max_iterations > 0) rc = lm_iterate(lm->max_iterations, nullptr);
    const int rc2 = lm_end(stats);
    return rc ? rc : rc2;
  }

  int lm_run(const hsba_lm_options *lm, hsba_lm_stats *stats) override {
    int rc = lm_begin(lm);
    if (rc == HSBA_OK && lm->max_iterations > 0) rc = lm_iterate(lm->max_iterations, nullptr);
    if (rc) return rc;
    rc = lm_end(stats);
    return (rc == HSBA_ERR_DEVICE || rc == HSBA_ERR_OOM) ? rc : HSBA_OK;
  }

  int lm_begin(const hsba_lm_options *lm) override {
    int rc = set_state(lm, 0.0, 1, 1, 0, /*upload=*/false);
    if (rc) return rc;
    // A free image or camera without any key leaves its diagonal block out of
    // the reference's reduced system, so SimplicialLDLT fails for that
    // problem: that window stops at once with the reference's -1, and the
    // other windows of the batch run as if alone.
    for (int w = 0; w < L.nwin; ++w)
      if (L.hw[w].unobserved) {
        WinState &s = ws_host[w];
        s.active = 0;
        s.need_lin = 0;
        s.status = HSBA_ERR_GENERIC;
        s.stop_reason = HSBA_STOP_FAILED;
      }
    HCHK(hipMemcpyAsync(d.ws, ws_host.data(), sizeof(WinState) * L.nwin, hipMemcpyHostToDevice, st));
    // initial cost, linearization, mu0 and the gradient test
    ks->cost(d, st, 0);
    enqueue_linearize();
    ks->grad_stats(d, st);
    ks->lm_control(d, st, 0);
    HCHK(hipGetLastError());
    HCHK(hipStreamSynchronize(st));
    return HSBA_OK;
  }

  // Timed mode: each stage of an iteration is its own captured graph and the
  // HIP events are recorded on the stream between stage launches (timing
  // queries on events recorded inside a captured graph are not supported).
  static constexpr int kStages = 8;
  hipEvent_t ev[kStages + 1] = {};
  hipGraphExec_t sgexec[kStages] = {};
  hipGraph_t sgraph[kStages] = {};

  void enqueue_stage(int i) {
    switch (i) {
      case 0:  // K3 point factor + Z
        ks->lm_control(d, st, 1);
        ks->point_factor(d, st);
        ks->key_z(d, st);
        break;
      case 1: enqueue_pairs(); break;  // K4
      case 2: ks->assemble(d, st, L.lower_total, (int)L.rows_total); break;
      case 3:  // K5
        chol_factor_solve_impl(plan, small_list_dev, (int)L.small_list.size(), L.large_list.data(),
                               (int)L.large_list.size(), st);
        break;
      case 4: ks->backsub(d, st, glog2); break;  // K6
      case 5:  // K7 + K8
        ks->step(d, st);
        ks->lm_control(d, st, 2);
        ks->cost(d, st, 1);
        ks->lm_control(d, st, 3);
        ks->commit(d, st);
        break;
      case 6: enqueue_lin_kernel(); break;  // K1 + K2
      default:
        ks->image_reduce(d, st);
        ks->camera_reduce(d, st);
        ks->constraints(d, st);
        ks->grad_stats(d, st);
        ks->lm_control(d, st, 4);
        break;
    }
  }

  int lm_iterate(int iterations, hsba_stage_times *times) override {
    if (iterations <= 0) return HSBA_OK;
    if (L.pcg) return pcg_iterate(iterations, times);
    if (!times) {
      // The whole run of `iterations` LM iterations is one captured graph and
      // one launch (up to kMaxGraphIterations; longer runs replay a graph of
      // that many iterations).  The graph is kept for the next call with the
      // same count.
      const int n = std::min(iterations, kMaxGraphIterations);
      if (!gexec || graph_iters != n) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < n; ++it) enqueue_iteration();
        hipError_t e1 = hipStreamEndCapture(st, &graph);
        if (e1 != hipSuccess) {
          std::fprintf(stderr, "hsba: capture failed: %s\n", hipGetErrorString(e1));
          return HSBA_ERR_DEVICE;
        }
        HCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        graph_iters = n;
      }
      int left = iterations;
      while (left >= n) {
        HCHK(hipGraphLaunch(gexec, st));
        left -= n;
      }
      if (left > 0) {
        // remainder: eager launches (same kernels, same order)
        for (int it = 0; it < left; ++it) enqueue_iteration();
        HCHK(hipGetLastError());
      }
      HCHK(hipStreamSynchronize(st));
      return HSBA_OK;
    }
    if (!sgexec[0]) {
      for (int i = 0; i <= kStages; ++i) HCHK(hipEventCreate(&ev[i]));
      for (int i = 0; i < kStages; ++i) {
        HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_stage(i);
        hipError_t e1 = hipStreamEndCapture(st, &sgraph[i]);
        if (e1 != hipSuccess) {
          std::fprintf(stderr, "hsba: capture failed: %s\n", hipGetErrorString(e1));
          return HSBA_ERR_DEVICE;
        }
        HCHK(hipGraphInstantiate(&sgexec[i], sgraph[i], nullptr, nullptr, 0));
      }
    }
    double acc[kStages + 1] = {0};
    for (int it = 0; it < iterations; ++it) {
      HCHK(hipEventRecord(ev[0], st));
      for (int i = 0; i < kStages; ++i) {
        HCHK(hipGraphLaunch(sgexec[i], st));
        HCHK(hipEventRecord(ev[i + 1], st));
      }
      HCHK(hipEventSynchronize(ev[kStages]));
      for (int i = 0; i < kStages; ++i) {
        float ms = 0.f;
        HCHK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        acc[i] += ms;
      }
      float tot = 0.f;
      HCHK(hipEventElapsedTime(&tot, ev[0], ev[kStages]));
      acc[kStages] += tot;
    }
    const double inv = 1.0 / iterations;
    std::memset(times, 0, sizeof(*times));
    times->point_factor_ms = acc[0] * inv;
    times->pairs_ms = acc[1] * inv;
    times->assemble_ms = acc[2] * inv;
    times->cholesky_ms = acc[3] * inv;
    times->backsub_ms = acc[4] * inv;
    times->step_cost_ms = acc[5] * inv;
    times->linearize_ms = acc[6] * inv;
    times->image_reduce_ms = acc[7] * inv;
    times->iteration_ms = acc[kStages] * inv;
    times->iterations = iterations;
    return HSBA_OK;
  }

  int lm_end(hsba_lm_stats *stats) override {
    int rc = get_state();
    if (rc) return rc;
    built = true;
    int worst = HSBA_OK;
    for (int w = 0; w < L.nwin; ++w) {
      const WinState &s = ws_host[w];
      hsba_lm_stats &o = stats[w];
      std::memset(&o, 0, sizeof(o));
      o.status = s.status;
      o.stop_reason = s.status ? HSBA_STOP_FAILED : s.stop_reason;
      o.iterations = s.iter;
      o.accepted = s.accepted;
      o.initial_cost = s.cost0;
      o.final_cost = s.cost;
      o.final_mu = s.mu;
      o.pcg_iterations = s.pcg_iters;
      if (s.status) worst = s.status;
    }
    return worst;
  }

  int build_normal(double *V, double *Ui, double *Uc, double *g) override {
    int rc = set_state(nullptr, 0.0, 1, 1, 0);
    if (rc) return rc;
    enqueue_linearize();
    HCHK(hipGetLastError());
    built = true;
    return download_normal(V, Ui, Uc, g);
  }

  int download_normal(double *V, double *Ui, double *Uc, double *g) override {
    if (!built) return HSBA_ERR_GENERIC;
    const int E = L.E, D = L.D, IMG = L.IMG;
    std::vector<T> hV(6 * L.P), hg(3 * L.P), hD((int64_t)E * E * L.I), hgr((int64_t)E * L.I),
        hU((int64_t)D * D * L.C), hgc((int64_t)D * L.C);
    if (L.pts_free && L.P) {
      HCHK(hipMemcpyAsync(hV.data(), d.V, sizeof(T) * hV.size(), hipMemcpyDeviceToHost, st));
      HCHK(hipMemcpyAsync(hg.data(), d.gp, sizeof(T) * hg.size(), hipMemcpyDeviceToHost, st));
    }
    if (E && L.I) {
      HCHK(hipMemcpyAsync(hD.data(), d.Dimg, sizeof(T) * hD.size(), hipMemcpyDeviceToHost, st));
      HCHK(hipMemcpyAsync(hgr.data(), d.gr, sizeof(T) * hgr.size(), hipMemcpyDeviceToHost, st));
    }
    if (D && L.C) {
      HCHK(hipMemcpyAsync(hU.data(), d.Uc, sizeof(T) * hU.size(), hipMemcpyDeviceToHost, st));
      HCHK(hipMemcpyAsync(hgc.data(), d.gc, sizeof(T) * hgc.size(), hipMemcpyDeviceToHost, st));
    }
    HCHK(hipStreamSynchronize(st));
    const HostWindow &h = L.hw[0];
    static const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    if (V && L.pts_free)
      for (int64_t p = 0; p < L.P; ++p)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) V[9 * p + 3 * a + b] = (double)hV[6 * p + sym[a][b]];
    if (Ui && IMG)
      for (int64_t i = 0; i < L.I; ++i)
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 6; ++b) Ui[36 * i + 6 * a + b] = (double)hD[(int64_t)i * E * E + a * E + b];
    if (Uc && D)
      for (int64_t j = 0; j < (int64_t)D * D * L.C; ++j) Uc[j] = (double)hU[j];
    if (g) {
      if (L.pts_free) for (int64_t j = 0; j < 3 * L.P; ++j) g[h.x_pt + j] = (double)hg[j];
      if (IMG)
        for (int64_t i = 0; i < L.I; ++i)
          for (int a = 0; a < 6; ++a) g[h.x_img + 6 * i + a] = (double)hgr[(int64_t)i * E + a];
      if (D)
        for (int64_t j = 0; j < (int64_t)D * L.C; ++j) g[h.x_cam + j] = (double)hgc[j];
    }
    return HSBA_OK;
  }

  // Coupling blocks of the last build in the reference's block order:
  // W_pi per key, W_pc per distinct (point, camera), W_ic per distinct
  // (image, camera).  Device W_k = J_p^T Lambda [J_i | J_c] (3 x E per
  // point-major key) and D_i (whose image-camera block is W_ic) are read back.
  int coupling(double *pi, double *pc, double *ic) override {
    if (!built) return HSBA_ERR_GENERIC;
    const int E = L.E, IMG = L.IMG, D = L.D;
    hsba_structure_sizes sz;
    int rc = export_structure(L, &sz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    std::vector<int32_t> pcs(2 * sz.num_point_camera), ics(2 * sz.num_image_camera);
    if ((rc = export_structure(L, nullptr, nullptr, nullptr, nullptr, pcs.data(), ics.data(), nullptr, nullptr,
                               nullptr)))
      return rc;
    std::vector<T> hW, hD;
    if (L.pts_free && E > 0 && L.N > 0 && (pi || pc)) {
      hW.resize((size_t)3 * E * L.N);
      HCHK(hipMemcpyAsync(hW.data(), d.W, sizeof(T) * hW.size(), hipMemcpyDeviceToHost, st));
    }
    if (ic && sz.num_image_camera > 0) {
      hD.resize((size_t)E * E * L.I);
      HCHK(hipMemcpyAsync(hD.data(), d.Dimg, sizeof(T) * hD.size(), hipMemcpyDeviceToHost, st));
    }
    HCHK(hipStreamSynchronize(st));
    if (pi && sz.num_point_image > 0)
      for (int64_t s = 0; s < L.N; ++s) {
        const int64_t k = L.key_orig[s];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 6; ++b) pi[18 * k + 6 * a + b] = (double)hW[(size_t)(3 * s + a) * E + b];
      }
    if (pc && sz.num_point_camera > 0) {
      for (int64_t q = 0; q < sz.num_point_camera; ++q) {
        const int32_t p = pcs[2 * q], c = pcs[2 * q + 1];
        double *o = pc + (size_t)q * 3 * D;
        for (int j = 0; j < 3 * D; ++j) o[j] = 0.0;
        for (int32_t s = L.pt_ptr[p]; s < L.pt_ptr[p + 1]; ++s) {
          if (L.img_cam[L.key_img[s]] != c) continue;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < D; ++b) o[a * D + b] += (double)hW[(size_t)(3 * s + a) * E + IMG + b];
        }
      }
    }
    if (ic && sz.num_image_camera > 0)
      for (int64_t q = 0; q < sz.num_image_camera; ++q) {
        const int32_t im = ics[2 * q];
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < D; ++b) ic[(q * 6 + a) * D + b] = (double)hD[(size_t)im * E * E + a * E + IMG + b];
      }
    return HSBA_OK;
  }

  // One damped Newton step on the resident normal equations (window 0):
  // CameraSharedNormalEquationSolver::operator().  S / rhs, when given,
  // receive the reduced system instead of (or besides) the step.
  // Diagonal blocks of N^-1 at the last build (no damping) and the dense
  // inverse of the reduced system: the Schur complement S = U - W^T V^-1 W is
  // factored, S^-1 formed from the factor on the device, and each point's
  // block from S^-1 and its Z rows.
  int covariance(double *Vp, double *Ui, double *Uc, double *red, double *scratch) override {
    if (!built) return HSBA_ERR_GENERIC;
    if (L.pcg || L.nwin != 1) return HSBA_ERR_UNSUPPORTED;
    if (L.hw[0].unobserved) return HSBA_ERR_NOT_SPD;
    const WinInfo &wi = L.wi[0];
    if (scratch && (wi.ld > kScratchMaxLd || L.P > kScratchMaxPoints)) return HSBA_ERR_TOO_LARGE;
    int rc = set_state(nullptr, 0.0, 1, 0, 1);
    if (rc) return rc;
    ks->point_factor(d, st);
    ks->key_z(d, st);
    enqueue_pairs();
    ks->assemble(d, st, L.lower_total, (int)L.rows_total);
    chol_factor_solve_impl(plan, small_list_dev, (int)L.small_list.size(), L.large_list.data(),
                           (int)L.large_list.size(), st);
    HCHK(hipGetLastError());
    if ((rc = get_state())) return rc;
    if (ws_host[0].chol_fail) return HSBA_ERR_NOT_SPD;
    const int64_t ld = wi.ld, n = wi.n;
    // L^-T (Y), S^-1 and the point blocks: from the fixed scratch region, or
    // allocated for this call.  Every element that is read is written first
    // (Y's upper block triangle column by column, both triangles of S^-1,
    // every point block), so the scratch needs no clearing between calls.
    double *Y = nullptr, *Sinv = nullptr, *pc = nullptr;
    const bool own = scratch == nullptr;
    auto release = [&]() {
      if (!own) return;
      if (Y) (void)hipFree(Y);
      if (Sinv) (void)hipFree(Sinv);
      if (pc) (void)hipFree(pc);
    };
    std::vector<double> hS;
    if (n > 0) {
      if (own) {
        const size_t bytes = sizeof(double) * (size_t)(wi.s_off + ld * ld);
        if (hipMalloc((void **)&Y, bytes) != hipSuccess || hipMalloc((void **)&Sinv, bytes) != hipSuccess) {
          release();
          return HSBA_ERR_OOM;
        }
      } else {
        // one window: wi.s_off == 0
        Y = scratch;
        Sinv = scratch + kScratchMaxLd * kScratchMaxLd;
      }
      const bool large = !L.large_list.empty();
      chol_inverse_impl(plan, 0, large, Y, Sinv, st);
    }
    if (L.pts_free && L.P && Vp) {
      if (!own) {
        pc = scratch + 2 * kScratchMaxLd * kScratchMaxLd;
      } else if (hipMalloc((void **)&pc, sizeof(double) * 9 * L.P) != hipSuccess) {
        release();
        return HSBA_ERR_OOM;
      }
      ks->point_cov(d, st, Sinv, ld, wi.r_off, pc);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
      release();
      return HSBA_ERR_DEVICE;
    }
    if (pc && hipMemcpy(Vp, pc, sizeof(double) * 9 * L.P, hipMemcpyDeviceToHost) != hipSuccess) {
      release();
      return HSBA_ERR_DEVICE;
    }
    if (n > 0 && (Ui || Uc || red)) {
      hS.resize((size_t)ld * ld);
      if (hipMemcpy(hS.data(), Sinv + wi.s_off, sizeof(double) * hS.size(), hipMemcpyDeviceToHost) != hipSuccess) {
        release();
        return HSBA_ERR_DEVICE;
      }
    }
    release();
    auto at = [&](int64_t i, int64_t j) { return hS[(size_t)(i - wi.r_off) * ld + (j - wi.r_off)]; };
    if (Ui && L.IMG)
      for (int64_t i = 0; i < L.I; ++i)
        for (int a = 0; a < 6; ++a)
          for (int b = 0; b < 6; ++b) Ui[36 * i + 6 * a + b] = at(L.img_row[i] + a, L.img_row[i] + b);
    if (Uc && L.D)
      for (int64_t c = 0; c < L.C; ++c)
        for (int a = 0; a < L.D; ++a)
          for (int b = 0; b < L.D; ++b)
            Uc[(c * L.D + a) * L.D + b] = at(L.cam_row[c] + a, L.cam_row[c] + b);
    if (red)
      for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j) red[i * n + j] = hS[(size_t)i * ld + j];
    return HSBA_OK;
  }

  // Point factors, Z, pairs, the reduced system and its factorization of
  // window 0 at damping mu (no LM state); the state is read back, so
  // ws_host[0].chol_fail tells whether N + mu I was positive definite.
  int factor_at(double mu) {
    int rc = set_state(nullptr, mu, 1, 0, 1);
    if (rc) return rc;
    ks->point_factor(d, st);
    ks->key_z(d, st);
    enqueue_pairs();
    ks->assemble(d, st, L.lower_total, (int)L.rows_total);
    chol_factor_solve_impl(plan, small_list_dev, (int)L.small_list.size(), L.large_list.data(),
                           (int)L.large_list.size(), st);
    HCHK(hipGetLastError());
    return get_state();
  }

  // Full dense covariance (window 0) in the x layout: the pseudo-inverse of N
  // with singular values at or below kPinvThreshold dropped.  Schur route:
  // C_rr = S^-1, C_pr = -G S^-1 and C_pp = V^-1 + G S^-1 G^T with the dense
  // point rows G = V^-1 W = L^-T Z, as matrix-core GEMMs.  Dense route: N
  // assembled densely (V blocks, W rows and their transpose, U = the reduced
  // matrix without pair terms or damping) and pseudo-inverted through its
  // eigen-decomposition by the cyclic Jacobi method.
  int x_covariance(double *out, int method, hsba_covariance_info *info) override {
    if (!built) return HSBA_ERR_GENERIC;
    if (L.pcg || L.nwin != 1) return HSBA_ERR_UNSUPPORTED;
    if (L.hw[0].unobserved) return HSBA_ERR_NOT_SPD;
    if (method < HSBA_COV_AUTO || method > HSBA_COV_DENSE) return HSBA_ERR_INVALID;
    const WinInfo &wi = L.wi[0];
    const int64_t n = wi.n, ld = wi.ld;
    const int64_t P3 = L.pts_free ? 3 * L.P : 0, xs = P3 + n;
    hsba_covariance_info inf;
    std::memset(&inf, 0, sizeof(inf));
    if (xs == 0) {
      if (info) *info = inf;
      return HSBA_OK;
    }
    int rc;
    bool schur = method == HSBA_COV_SCHUR;
    if (method == HSBA_COV_AUTO) {
      // every singular value of N exceeds the cut-off exactly when
      // N - cut-off I is positive definite
      if ((rc = factor_at(-kPinvThreshold))) return rc;
      schur = !ws_host[0].chol_fail;
    }
    if (!schur && xs > HSBA_COV_DENSE_MAX) return HSBA_ERR_TOO_LARGE;
    CallBuffers buf;
    double *dout = nullptr;
    if ((rc = buf.get(dout, xs * xs))) return rc;
    if (schur) {
      if ((rc = factor_at(0.0))) return rc;
      if (ws_host[0].chol_fail) return HSBA_ERR_NOT_SPD;
      double *Y = nullptr, *Sinv = nullptr;
      if (n > 0) {
        if ((rc = buf.get(Y, ld * ld)) || (rc = buf.get(Sinv, ld * ld))) return rc;
        chol_inverse_impl(plan, 0, !L.large_list.empty(), Y, Sinv, st);
      }
      if (P3 > 0 && n > 0) {
        const int64_t P3p = (P3 + 63) / 64 * 64;
        double *G = nullptr, *Tn = nullptr;
        if ((rc = buf.get(G, P3p * ld)) || (rc = buf.get(Tn, P3p * ld))) return rc;
        HCHK(hipMemsetAsync(G, 0, sizeof(double) * (size_t)(P3p * ld), st));
        ks->point_rows(d, st, 1, G, ld, wi.r_off);
        dense_gemm(false, false, P3p, ld, ld, -1.0, G, ld, Sinv, ld, Tn, ld, false, st);  // Tn = -G S^-1 = C_pr
        dense_gemm(false, true, P3, P3, ld, -1.0, Tn, ld, G, ld, dout, xs, false, st);    // G S^-1 G^T
        dense_copy(Tn, ld, dout + P3, xs, P3, n, 0, st);
        dense_copy(Tn, ld, dout + P3 * xs, xs, n, P3, 1, st);
      } else if (P3 > 0) {
        HCHK(hipMemsetAsync(dout, 0, sizeof(double) * (size_t)(xs * xs), st));
      }
      if (P3 > 0) ks->point_diag(d, st, 1, dout, xs);
      if (n > 0) dense_copy(Sinv, ld, dout + P3 * xs + P3, xs, n, n, 0, st);
      inf.method = HSBA_COV_SCHUR;
      inf.rank = (int32_t)xs;
    } else {
      const int64_t xp = (xs + 63) / 64 * 64;
      double *A = nullptr, *Vm = nullptr, *work = nullptr;
      int *rot = nullptr;
      if ((rc = buf.get(A, xp * xp)) || (rc = buf.get(Vm, xp * xp)) || (rc = buf.get(work, 2 * xs + 2)) ||
          (rc = buf.get(rot, kJacobiMaxSweeps)))
        return rc;
      HCHK(hipMemsetAsync(A, 0, sizeof(double) * (size_t)(xp * xp), st));
      if (P3 > 0) ks->point_diag(d, st, 0, A, xp);
      if (P3 > 0 && n > 0) {
        ks->point_rows(d, st, 0, A + P3, xp, wi.r_off);
        dense_copy(A + P3, xp, A + P3 * xp, xp, n, P3, 1, st);
      }
      if (n > 0) {
        if (L.npairs > 0) HCHK(hipMemsetAsync(d.pair_out, 0, sizeof(T) * (size_t)(L.E * L.E * L.npairs), st));
        if ((rc = set_state(nullptr, 0.0, 1, 0, 1))) return rc;
        ks->assemble(d, st, L.lower_total, (int)L.rows_total);
        dense_copy(d.S + wi.s_off, ld, A + P3 * xp + P3, xp, n, n, 2, st);
      }
      HCHK(hipGetLastError());
      int sweeps = 0;
      std::vector<double> lam;
      if ((rc = dense_pinv(A, Vm, xs, xp, kPinvThreshold, dout, xs, kJacobiMaxSweeps, rot, work, lam, &sweeps, st)))
        return rc;
      inf.method = HSBA_COV_DENSE;
      inf.sweeps = sweeps;
      for (double v : lam)
        if (std::fabs(v) > kPinvThreshold) {
          ++inf.rank;
          inf.smallest_kept = inf.rank == 1 ? std::fabs(v) : std::min(inf.smallest_kept, std::fabs(v));
        }
    }
    HCHK(hipGetLastError());
    HCHK(hipMemcpyAsync(out, dout, sizeof(double) * (size_t)(xs * xs), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    if (info) *info = inf;
    return HSBA_OK;
  }

  int solve(double mu, double *delta, double *S, double *rhs) override {
    if (!built) return HSBA_ERR_GENERIC;
    if (L.hw[0].unobserved) return HSBA_ERR_GENERIC;
    int rc = set_state(nullptr, mu, 1, 0, 1);
    if (rc) return rc;
    const WinInfo &wi = L.wi[0];
    if (L.pcg) {
      if (S || rhs) return HSBA_ERR_UNSUPPORTED;  // no dense reduced matrix in PCG mode
      if ((rc = capture_pcg())) return rc;
      ks->point_factor(d, st);
      ks->key_z(d, st);
      enqueue_pairs();
      ks->assemble(d, st, 0, (int)L.rows_total);
      ks->pcg_setup(d, st, glog2);
      if ((rc = pcg_solve_loop())) return rc;
      ks->backsub(d, st, glog2);
      HCHK(hipGetLastError());
      if ((rc = get_state())) return rc;
      if (ws_host[0].chol_fail) return HSBA_ERR_GENERIC;
      const HostWindow &h = L.hw[0];
      std::vector<double> dp(L.pts_free ? 3 * L.P : 0), dr(std::max<int64_t>(1, L.rows_total));
      if (!dp.empty()) HCHK(hipMemcpy(dp.data(), d.dpt, 8 * dp.size(), hipMemcpyDeviceToHost));
      HCHK(hipMemcpy(dr.data(), d.dred, 8 * dr.size(), hipMemcpyDeviceToHost));
      if (L.pts_free) std::copy(dp.begin(), dp.end(), delta + h.x_pt);
      const int64_t r0 = L.pts_free ? 3 * L.P : 0;
      for (int j = 0; j < wi.n; ++j) delta[r0 + j] = dr[wi.r_off + j];
      return HSBA_OK;
    }
    ks->point_factor(d, st);
    ks->key_z(d, st);
    enqueue_pairs();
    ks->assemble(d, st, L.lower_total, (int)L.rows_total);
    if (S || rhs) {
      HCHK(hipStreamSynchronize(st));
      std::vector<double> hS((size_t)wi.ld * wi.ld), hr(wi.ld);
      if (wi.n) {
        HCHK(hipMemcpy(hS.data(), d.S + wi.s_off, 8 * hS.size(), hipMemcpyDeviceToHost));
        HCHK(hipMemcpy(hr.data(), d.rhs + wi.r_off, 8 * (size_t)wi.n, hipMemcpyDeviceToHost));
      }
      for (int i = 0; i < wi.n; ++i)
        for (int j = 0; j <= i; ++j) {
          if (S) { S[(int64_t)i * wi.n + j] = hS[(int64_t)i * wi.ld + j]; S[(int64_t)j * wi.n + i] = hS[(int64_t)i * wi.ld + j]; }
        }
      if (rhs) for (int i = 0; i < wi.n; ++i) rhs[i] = hr[i];
      if (!delta) return HSBA_OK;
    }
    chol_factor_solve_impl(plan, small_list_dev, (int)L.small_list.size(), L.large_list.data(),
                           (int)L.large_list.size(), st);
    ks->backsub(d, st, glog2);
    HCHK(hipGetLastError());
    if ((rc = get_state())) return rc;
    if (ws_host[0].chol_fail) return HSBA_ERR_GENERIC;
    const HostWindow &h = L.hw[0];
    std::vector<double> dp(L.pts_free ? 3 * L.P : 0), dr(std::max<int64_t>(1, L.rows_total));
    if (!dp.empty()) HCHK(hipMemcpy(dp.data(), d.dpt, 8 * dp.size(), hipMemcpyDeviceToHost));
    HCHK(hipMemcpy(dr.data(), d.dred, 8 * dr.size(), hipMemcpyDeviceToHost));
    if (L.pts_free) std::copy(dp.begin(), dp.end(), delta + h.x_pt);
    const int64_t r0 = L.pts_free ? 3 * L.P : 0;
    for (int j = 0; j < wi.n; ++j) delta[r0 + j] = dr[wi.r_off + j];
    return HSBA_OK;
  }

  int evaluate(double *y) override {
    double *dy;
    HCHK(hipMalloc((void **)&dy, 16 * (size_t)std::max<int64_t>(1, L.N)));
    ks->export_keys(d, st, dy, nullptr, nullptr, nullptr, 0);
    HCHK(hipGetLastError());
    HCHK(hipMemcpyAsync(y, dy, 16 * (size_t)L.N, hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    HCHK(hipFree(dy));
    return HSBA_OK;
  }

  int jacobian(double *jp, double *ji, double *jc) override {
    const int D = std::max(L.D, 1);
    double *dp = nullptr, *di = nullptr, *dc = nullptr;
    const size_t n = (size_t)std::max<int64_t>(1, L.N);
    HCHK(hipMalloc((void **)&dp, 8 * 6 * n));
    HCHK(hipMalloc((void **)&di, 8 * 12 * n));
    HCHK(hipMalloc((void **)&dc, 8 * 2 * D * n));
    ks->export_keys(d, st, nullptr, dp, di, dc, 1);
    HCHK(hipGetLastError());
    if (jp) HCHK(hipMemcpyAsync(jp, dp, 8 * 6 * (size_t)L.N, hipMemcpyDeviceToHost, st));
    if (ji) HCHK(hipMemcpyAsync(ji, di, 8 * 12 * (size_t)L.N, hipMemcpyDeviceToHost, st));
    if (jc && L.D) HCHK(hipMemcpyAsync(jc, dc, 8 * 2 * (size_t)L.D * L.N, hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    HCHK(hipFree(dp));
    HCHK(hipFree(di));
    HCHK(hipFree(dc));
    return HSBA_OK;
  }

  int scalars(double *maxdiag, double *ginf) override {
    if (!built) return HSBA_ERR_GENERIC;
    int rc = set_state(nullptr, 0.0, 1, 1, 0);
    if (rc) return rc;
    ks->grad_stats(d, st);
    std::vector<double> part(4 * L.chunk_kind.size());
    HCHK(hipMemcpyAsync(part.data(), d.partial, 8 * part.size(), hipMemcpyDeviceToHost, st));
    HCHK(hipStreamSynchronize(st));
    double gm = 0.0, dm = -1e308;
    for (size_t c = 0; c < L.chunk_kind.size(); ++c)
      if (L.chunk_kind[c] < 3) { gm = std::max(gm, part[4 * c]); dm = std::max(dm, part[4 * c + 1]); }
    if (maxdiag) *maxdiag = dm;
    if (ginf) *ginf = gm;
    return HSBA_OK;
  }

  int cost(double *value) override {
    int rc = set_state(nullptr, 0.0, 1, 0, 0);
    if (rc) return rc;
    ks->cost(d, st, 0);
    HCHK(hipGetLastError());
    std::vector<double> part(4 * L.chunk_kind.size());
    HCHK(hipMemcpyAsync(part.data(), d.partial, 8 * part.size(), hipMemcpyDeviceToHost, st));
    if ((rc = get_state())) return rc;
    double F = 0.0;
    for (size_t c = 0; c < L.chunk_kind.size(); ++c)
      if (L.chunk_kind[c] == 3 && L.chunk_win[c] == 0) F += part[4 * c + 3];
    *value = F + ws_host[0].cost_new;
    return HSBA_OK;
  }
};

EngineBase *make_engine(int precision) {
  if (precision == HSBA_PRECISION_FP64) return new Engine<double>();
  return new Engine<float>();
}

}  // namespace hsba
