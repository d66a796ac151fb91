// hsba_layout.cpp — host-side structure of a problem or batch of windows.
//
// Mirrors the sparsity pass of CameraSharedNormalEquationBuilder::
// ComputeNormalMatrix (include/hs_sfm/bundle_adjustment/
// camera_shared_normal_equation_builder.hpp:137-205): instead of Eigen sparse
// maps of (point, image) / (point, camera) / (image, camera) blocks it builds
//   * keys in point-major order (a point's keys contiguous, sorted by image),
//   * point CSR, image CSR (image-major key lists), camera -> image lists,
//   * the camera-pair incidence lists of the Schur complement
//     (pairs (ia <= ib) of images that share a point, one incidence per point),
//   * the reduced-system layout [images 6I | cameras dC] of
//     camera_shared_normal_equation_solver.hpp:131-135 with per-block
//     contribution lists,
//   * chunk tables for the deterministic per-window reductions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>

#include "hsba_host.h"

namespace hsba {

namespace {

uint32_t enabled_bits(uint32_t comp) {
  uint32_t b = 0;
  if (comp & HSBA_COMPUTE_RADIAL_DISTORTION) b |= 0x7u;
  if (comp & HSBA_COMPUTE_DECENTERING_DISTORTION) b |= 0x18u;
  if (comp & HSBA_COMPUTE_INTRINSIC_PARAMS) b |= 0x3E0u;
  return b;
}

double identity_value(int j) { return (j == IF || j == IPR) ? 1.0 : 0.0; }

int fail(std::string *why, const std::string &msg, int code = HSBA_ERR_INVALID) {
  if (why) *why = msg;
  return code;
}

// Entries per reduction chunk (one workgroup of 256 threads each): small
// enough that the cost and step reductions of a 1M-key window fill the GPU.
constexpr int kChunk = 1024;

// Host threads for the O(keys x track) passes: one per 64k keys, at most 16.
int layout_threads(int64_t keys) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  return (int)std::max<int64_t>(1, std::min<int64_t>({16, hw, keys / 65536}));
}

template <class F>
void run_threads(int nt, F &&f) {
  if (nt <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(f, t);
  f(0);
  for (auto &x : th) x.join();
}

void add_chunks(Layout &L, int kind, int64_t begin, int64_t end, int w) {
  for (int64_t b = begin; b < end; b += kChunk) {
    L.chunk_kind.push_back(kind);
    L.chunk_begin.push_back((int32_t)b);
    L.chunk_end.push_back((int32_t)std::min<int64_t>(end, b + kChunk));
    L.chunk_win.push_back(w);
  }
}

}  // namespace

int build_layout(const hsba_problem_desc *descs, int64_t count, Layout &L, std::string *why, bool pcg) {
  if (!descs || count <= 0) return fail(why, "no problems");
  L = Layout();
  L.pcg = pcg;
  L.nwin = (int)count;
  const hsba_problem_desc &d0 = descs[0];
  L.comp = d0.intrinsic_computations_mask & 7u;
  L.fix = d0.fix_mask & 7u;
  if (L.fix == 7u) return fail(why, "every block fixed");
  const uint32_t en = enabled_bits(L.comp);
  L.free_bits = d0.intrinsic_free_mask ? d0.intrinsic_free_mask : en;
  if (L.free_bits & ~en) return fail(why, "free mask outside the enabled groups");
  int d = 0;
  for (int j = 0; j < kFull; ++j)
    if (L.free_bits & (1u << j)) L.fidx[d++] = j;
  for (int j = d; j < kFull; ++j) L.fidx[j] = 0;
  L.pts_free = !(L.fix & HSBA_FIX_POINTS);
  L.imgs_free = !(L.fix & HSBA_FIX_IMAGES);
  L.cams_free = !(L.fix & HSBA_FIX_CAMERAS);
  L.IMG = L.imgs_free ? 6 : 0;
  L.D = L.cams_free ? d : 0;
  L.E = L.IMG + L.D;
  const int D = L.D, IMG = L.IMG;

  // ---- totals and per-window validation
  L.hw.resize(count);
  L.wi.resize(count);
  for (int64_t w = 0; w < count; ++w) {
    const hsba_problem_desc &ds = descs[w];
    if ((ds.intrinsic_computations_mask & 7u) != L.comp || (ds.fix_mask & 7u) != L.fix ||
        (ds.intrinsic_free_mask ? ds.intrinsic_free_mask : enabled_bits(ds.intrinsic_computations_mask)) !=
            L.free_bits)
      return fail(why, "windows of one batch must share masks", HSBA_ERR_UNSUPPORTED);
    if (ds.num_points < 0 || ds.num_images < 0 || ds.num_cameras < 0 || ds.num_keys < 0)
      return fail(why, "negative size");
    if (ds.num_keys > 0 && !ds.feature_maps) return fail(why, "feature_maps missing");
    if (ds.num_images > 0 && !ds.image_camera_map) return fail(why, "image_camera_map missing");
    HostWindow &h = L.hw[w];
    h.P = ds.num_points; h.I = ds.num_images; h.C = ds.num_cameras; h.N = ds.num_keys;
    h.x_pt = 0;
    h.x_img = L.pts_free ? 3 * h.P : 0;
    h.x_cam = h.x_img + (L.imgs_free ? 6 * h.I : 0);
    h.x_size = h.x_cam + (L.cams_free ? (int64_t)d * h.C : 0);
    if (!L.pts_free && !ds.fix_points && h.P) return fail(why, "fix_points missing");
    if (!L.imgs_free && !ds.fix_images && h.I) return fail(why, "fix_images missing");
    if (!L.cams_free && !ds.fix_cameras && h.C && d) return fail(why, "fix_cameras missing");
    for (int64_t i = 0; i < h.I; ++i)
      if (ds.image_camera_map[i] < 0 || ds.image_camera_map[i] >= h.C) return fail(why, "image_camera_map out of range");
    for (int64_t k = 0; k < h.N; ++k) {
      const int32_t im = ds.feature_maps[2 * k], pt = ds.feature_maps[2 * k + 1];
      if (im < 0 || im >= h.I || pt < 0 || pt >= h.P) return fail(why, "feature_maps out of range");
    }
    L.P += h.P; L.I += h.I; L.C += h.C; L.N += h.N;
  }
  if (L.P >= (1LL << 31) || L.N >= (1LL << 31)) return fail(why, "too large", HSBA_ERR_UNSUPPORTED);

  L.key_img.resize(L.N); L.key_pt.resize(L.N); L.key_orig.resize(L.N);
  L.pt_ptr.assign(L.P + 1, 0); L.pt_win.resize(L.P);
  L.img_ptr.assign(L.I + 1, 0); L.img_keys.resize(L.N); L.img_win.resize(L.I); L.img_cam.resize(L.I);
  L.img_row.assign(L.I, -1);
  L.cam_ptr.assign(L.C + 1, 0); L.cam_imgs.resize(L.I); L.cam_win.resize(L.C); L.cam_row.assign(L.C, -1);
  L.cam_base.assign(L.C * kFull, 0.0);
  if (!L.pts_free) L.fix_pts.resize(3 * L.P);
  if (!L.imgs_free) L.fix_imgs.resize(6 * L.I);
  if (!L.cams_free) L.fix_cam_free.resize((int64_t)d * L.C);
  L.blk_ptr.assign(1, 0);

  int64_t pt_off = 0, img_off = 0, cam_off = 0, key_off = 0;
  int64_t r_off = 0, s_off = 0, dinv_off = 0, lo_off = 0, blk_off = 0;
  std::vector<int64_t> pcount;
  for (int64_t w = 0; w < count; ++w) {
    const hsba_problem_desc &ds = descs[w];
    HostWindow &h = L.hw[w];
    WinInfo &wi = L.wi[w];
    std::memset(&wi, 0, sizeof(wi));
    wi.key_off = key_off; wi.pt_off = (int32_t)pt_off; wi.img_off = (int32_t)img_off; wi.cam_off = (int32_t)cam_off;
    wi.num_pts = (int32_t)h.P; wi.num_imgs = (int32_t)h.I; wi.num_cams = (int32_t)h.C; wi.num_keys = (int32_t)h.N;
    // ---- parameters held fixed and camera base vectors
    for (int64_t c = 0; c < h.C; ++c)
      for (int j = 0; j < kFull; ++j) {
        double v = identity_value(j);
        if ((en >> j) & 1u) v = ds.intrinsic_defaults ? ds.intrinsic_defaults[c * kFull + j] : identity_value(j);
        L.cam_base[(cam_off + c) * kFull + j] = v;
      }
    if (!L.pts_free) std::copy(ds.fix_points, ds.fix_points + 3 * h.P, L.fix_pts.begin() + 3 * pt_off);
    if (!L.imgs_free) std::copy(ds.fix_images, ds.fix_images + 6 * h.I, L.fix_imgs.begin() + 6 * img_off);
    if (!L.cams_free && d)
      std::copy(ds.fix_cameras, ds.fix_cameras + (int64_t)d * h.C, L.fix_cam_free.begin() + (int64_t)d * cam_off);
    // ---- point-major key order: counting sort by point, images ascending inside
    pcount.assign(h.P + 1, 0);
    for (int64_t k = 0; k < h.N; ++k) pcount[ds.feature_maps[2 * k + 1] + 1]++;
    for (int64_t p = 0; p < h.P; ++p) pcount[p + 1] += pcount[p];
    {
      std::vector<int64_t> fill(pcount.begin(), pcount.end() - 1);
      for (int64_t k = 0; k < h.N; ++k) {
        const int64_t slot = key_off + fill[ds.feature_maps[2 * k + 1]]++;
        L.key_orig[slot] = (int32_t)(key_off + k);
        L.key_img[slot] = (int32_t)(img_off + ds.feature_maps[2 * k]);
        L.key_pt[slot] = (int32_t)(pt_off + ds.feature_maps[2 * k + 1]);
      }
    }
    int max_track = 0;
    {
      const int nt = layout_threads(h.N);
      std::vector<int> tmax(nt, 0), tdup(nt, 0);
      run_threads(nt, [&](int t) {
        for (int64_t p = h.P * t / nt; p < h.P * (t + 1) / nt; ++p) {
          const int64_t b = key_off + pcount[p], e = key_off + pcount[p + 1];
          tmax[t] = std::max<int>(tmax[t], (int)(e - b));
          // insertion sort of the point's keys by image (stable in key order)
          for (int64_t a = b + 1; a < e; ++a) {
            int32_t im = L.key_img[a], ko = L.key_orig[a];
            int64_t c = a - 1;
            while (c >= b && L.key_img[c] > im) {
              L.key_img[c + 1] = L.key_img[c]; L.key_orig[c + 1] = L.key_orig[c]; --c;
            }
            L.key_img[c + 1] = im; L.key_orig[c + 1] = ko;
          }
          for (int64_t a = b + 1; a < e; ++a)
            if (L.key_img[a] == L.key_img[a - 1]) tdup[t] = 1;
          L.pt_ptr[pt_off + p + 1] = (int32_t)e;
          L.pt_win[pt_off + p] = (int32_t)w;
        }
      });
      for (int t = 0; t < nt; ++t) {
        if (tdup[t]) return fail(why, "duplicate (image, point) key");
        max_track = std::max(max_track, tmax[t]);
      }
    }
    if (h.P == 0) L.pt_ptr[pt_off] = (int32_t)key_off;
    // ---- images and cameras
    for (int64_t i = 0; i < h.I; ++i) {
      L.img_win[img_off + i] = (int32_t)w;
      L.img_cam[img_off + i] = (int32_t)(cam_off + ds.image_camera_map[i]);
    }
    for (int64_t c = 0; c < h.C; ++c) L.cam_win[cam_off + c] = (int32_t)w;
    // ---- reduced layout
    wi.cb = (int32_t)(IMG ? 6 * h.I : 0);
    wi.n = (int32_t)(wi.cb + (D ? D * h.C : 0));
    wi.ld = wi.n > 0 ? ((wi.n + 63) / 64) * 64 : 0;
    wi.nb = wi.ld / 64;
    wi.r_off = (int32_t)r_off;
    wi.s_off = s_off;
    wi.dinv_off = dinv_off;
    wi.lo_off = lo_off;
    wi.blk_off = blk_off;
    wi.small = wi.nb <= 4 ? 1 : 0;
    if (pcg) {
      wi.s_off = 0;
      wi.dinv_off = 0;
      wi.lo_off = 0;
      wi.blk_off = 0;
    } else if (wi.n > 0) {
      (wi.small ? L.small_list : L.large_list).push_back((int)w);
    }
    for (int64_t i = 0; i < h.I && IMG; ++i) L.img_row[img_off + i] = (int32_t)(r_off + 6 * i);
    for (int64_t c = 0; c < h.C && D; ++c) L.cam_row[cam_off + c] = (int32_t)(r_off + wi.cb + D * c);
    const int64_t nblk = pcg ? 0 : (IMG ? h.I : 0) + (D ? h.C : 0);
    if (pcg && D) {
      std::vector<int> per_cam(h.C, 0);
      for (int64_t i = 0; i < h.I; ++i) per_cam[ds.image_camera_map[i]]++;
      for (int64_t c = 0; c < h.C; ++c)
        if (per_cam[c] != 1) return fail(why, "PCG needs one image per camera", HSBA_ERR_UNSUPPORTED);
    }
    auto bimg = [&](int64_t i) { return i; };
    auto bcam = [&](int64_t c) { return (IMG ? h.I : 0) + c; };
    L.row_blk.resize(r_off + wi.ld, 0);
    L.row_sub.resize(r_off + wi.ld, 0);
    for (int64_t R = 0; R < wi.n; ++R) {
      if (R < wi.cb) { L.row_blk[r_off + R] = (int32_t)(R / 6); L.row_sub[r_off + R] = (int32_t)(R % 6); }
      else {
        L.row_blk[r_off + R] = (int32_t)((IMG ? h.I : 0) + (R - wi.cb) / D);
        L.row_sub[r_off + R] = (int32_t)((R - wi.cb) % D);
      }
    }
    // observed images / cameras ("exist" rule of the reference Schur loops)
    std::vector<int64_t> img_n(h.I, 0), cam_n(h.C, 0);
    for (int64_t k = 0; k < h.N; ++k) {
      img_n[ds.feature_maps[2 * k]]++;
      cam_n[ds.image_camera_map[ds.feature_maps[2 * k]]]++;
    }
    if (L.pts_free) {
      for (int64_t i = 0; i < h.I && IMG; ++i) if (!img_n[i]) h.unobserved = true;
      for (int64_t c = 0; c < h.C && D; ++c) if (!cam_n[c]) h.unobserved = true;
    }
    // ---- camera pairs (dense pair table of the window)
    std::vector<std::vector<Contrib>> blocks((size_t)(nblk * (nblk + 1) / 2));
    auto add = [&](int64_t bx, int64_t by, Contrib c) {
      blocks[(size_t)(bx * (bx + 1) / 2 + by)].push_back(c);
    };
    for (int64_t i = 0; i < h.I && IMG && !pcg; ++i) add(bimg(i), bimg(i), Contrib{0, (int32_t)(img_off + i), 0, 0, 0});
    for (int64_t c = 0; c < h.C && D && !pcg; ++c) add(bcam(c), bcam(c), Contrib{1, (int32_t)(cam_off + c), 0, 0, 0});
    if (IMG && D && !pcg)
      for (int64_t i = 0; i < h.I; ++i)
        add(bcam(ds.image_camera_map[i]), bimg(i), Contrib{2, (int32_t)(img_off + i), 0, 0, 0});
    if (L.pts_free && L.E > 0 && h.I > 0) {
      // Incidences are counted and filled by threads over point ranges; a
      // pair's list is the concatenation of the threads' parts in thread
      // order, i.e. in point order, the same as a serial pass.
      const int64_t II = h.I * h.I;
      const int nt = (pcg || II > (int64_t(1) << 20)) ? 1 : layout_threads(h.N);
      std::vector<int64_t> pbeg(nt + 1);
      for (int t = 0; t <= nt; ++t) pbeg[t] = h.P * t / nt;
      std::vector<std::vector<int32_t>> tcnt(nt);
      run_threads(nt, [&](int t) {
        tcnt[t].assign((size_t)II, 0);
        int32_t *cn = tcnt[t].data();
        for (int64_t p = pbeg[t]; p < pbeg[t + 1]; ++p) {
          const int64_t b = key_off + pcount[p], e = key_off + pcount[p + 1];
          for (int64_t u = b; u < e; ++u) {
            const int64_t row = (L.key_img[u] - img_off) * h.I - img_off;
            for (int64_t v = u; v < (pcg ? u + 1 : e); ++v) cn[row + L.key_img[v]]++;
          }
        }
      });
      std::vector<int32_t> pid((size_t)II, -1);
      std::vector<int64_t> cnt((size_t)II, 0);
      for (int t = 0; t < nt; ++t)
        for (int64_t q = 0; q < II; ++q) cnt[q] += tcnt[t][q];
      const int64_t first_pair = L.npairs;
      std::vector<int64_t> cursor;
      for (int64_t ia = 0; ia < h.I; ++ia)
        for (int64_t ib = ia; ib < h.I; ++ib) {
          const int64_t c = cnt[(size_t)(ia * h.I + ib)];
          if (!c) continue;
          pid[(size_t)(ia * h.I + ib)] = (int32_t)L.npairs++;
          L.pair_win.push_back((int32_t)w);
          L.pair_img.push_back(ia == ib ? (int32_t)(img_off + ia) : -1);
          cursor.push_back(L.incidences);
          L.pair_ptr.push_back(L.incidences);
          L.incidences += c;
          // contributions of this pair block
          struct Side { int64_t blk; int16_t off; };
          Side xs[2], ys[2];
          int nx = 0, ny = 0;
          if (IMG) { xs[nx++] = {bimg(ia), 0}; ys[ny++] = {bimg(ib), 0}; }
          if (D) {
            xs[nx++] = {bcam(ds.image_camera_map[ia]), (int16_t)IMG};
            ys[ny++] = {bcam(ds.image_camera_map[ib]), (int16_t)IMG};
          }
          const int32_t pg = (int32_t)(L.npairs - 1);
          for (int a = 0; a < nx && !pcg; ++a)
            for (int b2 = 0; b2 < ny; ++b2) {
              const Side X = xs[a], Y = ys[b2];
              if (X.blk >= Y.blk) add(X.blk, Y.blk, Contrib{3, pg, X.off, Y.off, 0});
              if (ia != ib && Y.blk >= X.blk) add(Y.blk, X.blk, Contrib{3, pg, Y.off, X.off, 1});
            }
        }
      L.inc.resize(L.incidences);
      const int64_t npw = L.npairs - first_pair;
      std::vector<std::vector<int64_t>> tcur(nt, std::vector<int64_t>((size_t)npw));
      for (int64_t ia = 0, qq = 0; ia < h.I; ++ia)
        for (int64_t ib = ia; ib < h.I; ++ib) {
          const int64_t q = ia * h.I + ib;
          if (pid[(size_t)q] < 0) continue;
          int64_t off = cursor[qq];
          for (int t = 0; t < nt; ++t) {
            tcur[t][qq] = off;
            off += tcnt[t][q];
          }
          ++qq;
        }
      run_threads(nt, [&](int t) {
        int64_t *cur = tcur[t].data();
        for (int64_t p = pbeg[t]; p < pbeg[t + 1]; ++p) {
          const int64_t b = key_off + pcount[p], e = key_off + pcount[p + 1];
          for (int64_t u = b; u < e; ++u) {
            const int64_t row = (L.key_img[u] - img_off) * h.I - img_off;
            for (int64_t v = u; v < (pcg ? u + 1 : e); ++v) {
              const int32_t q = pid[(size_t)(row + L.key_img[v])];
              L.inc[cur[q - first_pair]++] = inc_code(u, v);
            }
          }
        }
      });
    }
    // flatten contribution lists
    for (auto &bl : blocks) {
      for (const Contrib &c : bl) L.contrib.push_back(c);
      L.blk_ptr.push_back((int64_t)L.contrib.size());
    }
    // ---- constraints in y order (camera_shared_vector_function.hpp:485-629)
    auto push = [&](int kind, int64_t ent_local, int param, int64_t xidx) {
      ConstraintRow cr;
      cr.kind = kind; cr.param = param; cr.win = (int32_t)w; cr.weight = 1.0; cr.y = 0.0;
      cr.brow[0] = cr.brow[1] = cr.brow[2] = 0.0; cr.block = 0; cr.rec_mask = 0;
      cr.entity = (int32_t)(ent_local + (kind == 0 ? pt_off : (kind == 1 ? img_off : cam_off)));
      L.cons.push_back(cr);
      h.cons_xidx.push_back(xidx);
    };
    wi.con_begin = (int32_t)L.cons.size();
    if (L.pts_free)
      for (int64_t q = 0; q < ds.num_point_constraints; ++q) {
        const hsba_constraint &c = ds.point_constraints[q];
        if (c.id < 0 || c.id >= h.P) return fail(why, "point constraint id");
        for (int b = 0; b < 3; ++b)
          if (c.mask & (1u << b)) {
            push(0, c.id, b, h.x_pt + 3 * c.id + b);
            L.cons.back().rec_mask = (int32_t)(c.mask & 7u);
          }
      }
    if (L.imgs_free)
      for (int64_t q = 0; q < ds.num_image_constraints; ++q) {
        const hsba_constraint &c = ds.image_constraints[q];
        if (c.id < 0 || c.id >= h.I) return fail(why, "image constraint id");
        if (c.mask & 1u) for (int b = 0; b < 3; ++b) push(1, c.id, b, h.x_img + 6 * c.id + b);
        for (int b = 1; b < 4; ++b) if (c.mask & (1u << b)) push(1, c.id, 2 + b, h.x_img + 6 * c.id + 2 + b);
      }
    if (L.cams_free)
      for (int64_t q = 0; q < ds.num_camera_constraints; ++q) {
        const hsba_constraint &c = ds.camera_constraints[q];
        if (c.id < 0 || c.id >= h.C) return fail(why, "camera constraint id");
        if ((c.mask & 0x3FFu) & ~en) return fail(why, "camera constraint on a disabled group");
        for (int j = 0; j < kFull; ++j) {
          if (!(c.mask & (1u << j))) continue;
          int local = -1;
          for (int q2 = 0; q2 < d; ++q2) if (L.fidx[q2] == j) local = q2;
          if (local < 0) return fail(why, "camera constraint on a held parameter");
          push(2, c.id, local, h.x_cam + (int64_t)d * c.id + local);
        }
      }
    wi.con_end = (int32_t)L.cons.size();
    h.y_size = 2 * h.N + (int64_t)h.cons_xidx.size();
    // ---- reduction chunks (sorted by window)
    if (L.pts_free) add_chunks(L, 0, 3 * pt_off, 3 * (pt_off + h.P), (int)w);
    if (L.imgs_free) add_chunks(L, 1, 6 * img_off, 6 * (img_off + h.I), (int)w);
    if (L.cams_free && D) add_chunks(L, 2, (int64_t)D * cam_off, (int64_t)D * (cam_off + h.C), (int)w);
    if (h.N > 0) add_chunks(L, 3, key_off, key_off + h.N, (int)w);
    else { L.chunk_kind.push_back(3); L.chunk_begin.push_back((int32_t)key_off); L.chunk_end.push_back((int32_t)key_off); L.chunk_win.push_back((int)w); }
    // ---- advance offsets
    pt_off += h.P; img_off += h.I; cam_off += h.C; key_off += h.N;
    r_off += wi.ld;
    if (!pcg) {
      s_off += (int64_t)wi.ld * wi.ld;
      dinv_off += (int64_t)wi.nb * 64 * 64;
      lo_off += (int64_t)wi.n * (wi.n + 1) / 2;
    }
    blk_off += nblk * (nblk + 1) / 2;
  }
  L.pt_ptr[0] = 0;
  // image CSR over point-major keys (each image's keys in point order)
  for (int64_t k = 0; k < L.N; ++k) L.img_ptr[L.key_img[k] + 1]++;
  for (int64_t i = 0; i < L.I; ++i) L.img_ptr[i + 1] += L.img_ptr[i];
  {
    std::vector<int32_t> fill(L.img_ptr.begin(), L.img_ptr.end() - 1);
    for (int64_t k = 0; k < L.N; ++k) L.img_keys[fill[L.key_img[k]]++] = (int32_t)k;
  }
  L.img_ch_ptr.assign(L.I + 1, 0);
  L.ich_img.clear(); L.ich_begin.clear(); L.ich_end.clear();
  for (int64_t i = 0; i < L.I; ++i) {
    for (int32_t b = L.img_ptr[i]; b < L.img_ptr[i + 1]; b += kImgChunk) {
      L.ich_img.push_back((int32_t)i);
      L.ich_begin.push_back(b);
      L.ich_end.push_back(std::min<int32_t>(b + kImgChunk, L.img_ptr[i + 1]));
    }
    L.img_ch_ptr[i + 1] = (int32_t)L.ich_img.size();
  }
  if (pcg) {
    L.key_ipos.resize(L.N);
    L.ikey_pt.resize(L.N);
    for (int64_t t = 0; t < L.N; ++t) {
      L.key_ipos[L.img_keys[t]] = (int32_t)t;
      L.ikey_pt[t] = L.key_pt[L.img_keys[t]];
    }
  }
  for (int64_t i = 0; i < L.I; ++i) L.cam_ptr[L.img_cam[i] + 1]++;
  for (int64_t c = 0; c < L.C; ++c) L.cam_ptr[c + 1] += L.cam_ptr[c];
  {
    std::vector<int32_t> fill(L.cam_ptr.begin(), L.cam_ptr.end() - 1);
    for (int64_t i = 0; i < L.I; ++i) L.cam_imgs[fill[L.img_cam[i]]++] = (int32_t)i;
  }
  L.pair_ptr.push_back(L.incidences);
  if (pcg && L.pts_free && L.E > 0) {
    // diagonal pairs in image order: pair id == image id
    if (L.npairs != L.I) return fail(why, "PCG needs every image observed");
    for (int64_t i = 0; i < L.I; ++i)
      if (L.pair_img[i] != (int32_t)i) return fail(why, "PCG pair order");
  }
  L.S_total = s_off; L.rows_total = r_off; L.dinv_total = dinv_off; L.lower_total = lo_off;
  if (L.row_blk.empty()) { L.row_blk.push_back(0); L.row_sub.push_back(0); }
  return HSBA_OK;
}

int64_t layout_device_bytes(const Layout &L, int precision) {
  const int64_t t = precision == HSBA_PRECISION_FP64 ? 8 : 4;
  const int64_t E = L.E, D = L.D;
  int64_t b = 0;
  b += L.N * (4 * 3 + 16 + 24);                                   // key arrays, whitening factors
  b += L.N * (3 * E * t + z_stride((int)E, (int)t) * t);           // W, Z
  if (L.pcg) b += L.N * (z_stride((int)E, (int)t) * t + 8);        // image-major Z copy
  b += L.P * (8 + 3 * 8 * 3 + t * 18);                              // points
  b += L.I * (4 * 5 + 6 * 8 * 2 + t * (E * E + 2 * E));             // images
  b += L.C * (4 * 3 + 8 * (2 * D + 20) + t * (D * D + D));           // cameras
  b += L.npairs * (8 + 8 + t * E * E) + L.incidences * 8;           // pairs
  b += (L.S_total + L.dinv_total) * 8 + L.rows_total * (8 * 8 + 8); // reduced systems
  b += (int64_t)L.contrib.size() * sizeof(Contrib) + (int64_t)L.blk_ptr.size() * 8;
  b += (int64_t)L.chunk_kind.size() * (16 + 32) + (int64_t)L.nwin * (sizeof(WinInfo) + sizeof(WinState));
  b += (int64_t)L.ich_img.size() * (12 + t * (E * (E + 1) / 2 + E)) + (L.I + 1) * 4;  // image-key chunks
  return b;
}

int export_structure(const Layout &L, hsba_structure_sizes *sz, int32_t *key_order, int32_t *point_ptr,
                     int32_t *point_image, int32_t *point_camera, int32_t *image_camera, int32_t *pairs,
                     int64_t *pair_ptr, int32_t *incidences) {
  // window 0 of the layout (a single problem: every offset is zero)
  const HostWindow &h = L.hw[0];
  const int64_t N = h.N, P = h.P;
  std::vector<int32_t> slot_of(N);
  for (int64_t s = 0; s < N; ++s) slot_of[L.key_orig[s]] = (int32_t)s;
  const bool has_pi = L.pts_free && L.imgs_free;
  const bool has_pc = L.pts_free && L.cams_free && L.D > 0;
  const bool has_ic = L.imgs_free && L.cams_free && L.D > 0;
  // first appearance in key order of each distinct (point, camera) and
  // (image, camera): the reference's std::set insertion numbering
  std::vector<int32_t> pc, ic;
  if (has_pc) {
    std::unordered_map<uint64_t, int64_t> seen;
    seen.reserve((size_t)N * 2);
    for (int64_t k = 0; k < N; ++k) {
      const int32_t s = slot_of[k];
      const int32_t p = L.key_pt[s], c = L.img_cam[L.key_img[s]];
      if (seen.emplace(((uint64_t)(uint32_t)p << 32) | (uint32_t)c, (int64_t)pc.size() / 2).second) {
        pc.push_back(p);
        pc.push_back(c);
      }
    }
  }
  if (has_ic) {
    std::vector<char> seen(h.I, 0);
    for (int64_t k = 0; k < N; ++k) {
      const int32_t im = L.key_img[slot_of[k]];
      if (!seen[im]) {
        seen[im] = 1;
        ic.push_back(im);
        ic.push_back(L.img_cam[im]);
      }
    }
  }
  if (sz) {
    sz->num_keys = N;
    sz->num_points = P;
    sz->num_point_image = has_pi ? N : 0;
    sz->num_point_camera = (int64_t)pc.size() / 2;
    sz->num_image_camera = (int64_t)ic.size() / 2;
    sz->num_pairs = L.npairs;
    sz->num_incidences = L.incidences;
  }
  if (key_order) std::copy(L.key_orig.begin(), L.key_orig.begin() + N, key_order);
  if (point_ptr) std::copy(L.pt_ptr.begin(), L.pt_ptr.begin() + P + 1, point_ptr);
  if (point_image && has_pi)
    for (int64_t k = 0; k < N; ++k) {
      point_image[2 * k] = L.key_pt[slot_of[k]];
      point_image[2 * k + 1] = L.key_img[slot_of[k]];
    }
  if (point_camera) std::copy(pc.begin(), pc.end(), point_camera);
  if (image_camera) std::copy(ic.begin(), ic.end(), image_camera);
  for (int64_t q = 0; q < L.npairs; ++q) {
    const int64_t b = L.pair_ptr[q], e = L.pair_ptr[q + 1];
    if (pairs) {
      // both images of a pair from its first incidence (every pair has one)
      pairs[2 * q] = L.key_img[inc_a(L.inc[b])];
      pairs[2 * q + 1] = L.key_img[inc_b(L.inc[b])];
    }
    if (pair_ptr) pair_ptr[q] = b;
    if (incidences)
      for (int64_t j = b; j < e; ++j) {
        incidences[2 * j] = L.key_orig[inc_a(L.inc[j])];
        incidences[2 * j + 1] = L.key_orig[inc_b(L.inc[j])];
      }
  }
  if (pair_ptr) pair_ptr[L.npairs] = L.incidences;
  return HSBA_OK;
}

int validate_desc(const hsba_problem_desc &ds, std::string *why) {
  if (ds.num_points < 0 || ds.num_images < 0 || ds.num_cameras < 0 || ds.num_keys < 0)
    return fail(why, "negative size");
  if (ds.num_points >= (1LL << 31) || ds.num_keys >= (1LL << 31)) return fail(why, "too large", HSBA_ERR_UNSUPPORTED);
  if (ds.num_keys > 0 && !ds.feature_maps) return fail(why, "feature_maps missing");
  if (ds.num_images > 0 && !ds.image_camera_map) return fail(why, "image_camera_map missing");
  const uint32_t fix = ds.fix_mask & 7u;
  if (fix == 7u) return fail(why, "every block fixed");
  const uint32_t en = enabled_bits(ds.intrinsic_computations_mask & 7u);
  const uint32_t fb = ds.intrinsic_free_mask ? ds.intrinsic_free_mask : en;
  if (fb & ~en) return fail(why, "free mask outside the enabled groups");
  const int d = __builtin_popcount(fb);
  if ((fix & HSBA_FIX_POINTS) && !ds.fix_points && ds.num_points) return fail(why, "fix_points missing");
  if ((fix & HSBA_FIX_IMAGES) && !ds.fix_images && ds.num_images) return fail(why, "fix_images missing");
  if ((fix & HSBA_FIX_CAMERAS) && !ds.fix_cameras && ds.num_cameras && d) return fail(why, "fix_cameras missing");
  for (int64_t i = 0; i < ds.num_images; ++i)
    if (ds.image_camera_map[i] < 0 || ds.image_camera_map[i] >= ds.num_cameras)
      return fail(why, "image_camera_map out of range");
  // ranges, then duplicate (image, point) keys: keys bucketed by point, each
  // bucket checked by sorting its images
  std::vector<int64_t> ptr(ds.num_points + 1, 0);
  for (int64_t k = 0; k < ds.num_keys; ++k) {
    const int32_t im = ds.feature_maps[2 * k], pt = ds.feature_maps[2 * k + 1];
    if (im < 0 || im >= ds.num_images || pt < 0 || pt >= ds.num_points) return fail(why, "feature_maps out of range");
    ptr[pt + 1]++;
  }
  for (int64_t p = 0; p < ds.num_points; ++p) ptr[p + 1] += ptr[p];
  {
    std::vector<int32_t> imgs(ds.num_keys);
    std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < ds.num_keys; ++k) imgs[fill[ds.feature_maps[2 * k + 1]]++] = ds.feature_maps[2 * k];
    for (int64_t p = 0; p < ds.num_points; ++p) {
      std::sort(imgs.begin() + ptr[p], imgs.begin() + ptr[p + 1]);
      for (int64_t a = ptr[p] + 1; a < ptr[p + 1]; ++a)
        if (imgs[a] == imgs[a - 1]) return fail(why, "duplicate (image, point) key");
    }
  }
  if (!(fix & HSBA_FIX_POINTS) && ds.num_point_constraints > 0) {
    if (!ds.point_constraints) return fail(why, "point_constraints missing");
    for (int64_t q = 0; q < ds.num_point_constraints; ++q)
      if (ds.point_constraints[q].id < 0 || ds.point_constraints[q].id >= ds.num_points)
        return fail(why, "point constraint id");
  }
  if (!(fix & HSBA_FIX_IMAGES) && ds.num_image_constraints > 0) {
    if (!ds.image_constraints) return fail(why, "image_constraints missing");
    for (int64_t q = 0; q < ds.num_image_constraints; ++q)
      if (ds.image_constraints[q].id < 0 || ds.image_constraints[q].id >= ds.num_images)
        return fail(why, "image constraint id");
  }
  if (!(fix & HSBA_FIX_CAMERAS) && ds.num_camera_constraints > 0) {
    if (!ds.camera_constraints) return fail(why, "camera_constraints missing");
    for (int64_t q = 0; q < ds.num_camera_constraints; ++q) {
      const hsba_constraint &c = ds.camera_constraints[q];
      if (c.id < 0 || c.id >= ds.num_cameras) return fail(why, "camera constraint id");
      if ((c.mask & 0x3FFu) & ~en) return fail(why, "camera constraint on a disabled group");
      if ((c.mask & 0x3FFu) & ~fb) return fail(why, "camera constraint on a held parameter");
    }
  }
  return HSBA_OK;
}

int problem_sizes(const hsba_problem_desc &ds, int64_t *x_size, int64_t *y_size) {
  if (ds.num_points < 0 || ds.num_images < 0 || ds.num_cameras < 0 || ds.num_keys < 0) return HSBA_ERR_INVALID;
  const uint32_t fix = ds.fix_mask & 7u;
  const uint32_t en = enabled_bits(ds.intrinsic_computations_mask & 7u);
  const uint32_t fb = ds.intrinsic_free_mask ? ds.intrinsic_free_mask : en;
  const bool pts = !(fix & HSBA_FIX_POINTS), imgs = !(fix & HSBA_FIX_IMAGES), cams = !(fix & HSBA_FIX_CAMERAS);
  *x_size = (pts ? 3 * ds.num_points : 0) + (imgs ? 6 * ds.num_images : 0) +
            (cams ? (int64_t)__builtin_popcount(fb) * ds.num_cameras : 0);
  int64_t rows = 0;
  if (pts && ds.point_constraints)
    for (int64_t q = 0; q < ds.num_point_constraints; ++q) rows += __builtin_popcount(ds.point_constraints[q].mask & 7u);
  if (imgs && ds.image_constraints)
    for (int64_t q = 0; q < ds.num_image_constraints; ++q) {
      const uint32_t m = ds.image_constraints[q].mask;
      rows += ((m & 1u) ? 3 : 0) + __builtin_popcount(m & 0xEu);
    }
  if (cams && ds.camera_constraints)
    for (int64_t q = 0; q < ds.num_camera_constraints; ++q) rows += __builtin_popcount(ds.camera_constraints[q].mask & 0x3FFu);
  *y_size = 2 * ds.num_keys + rows;
  return HSBA_OK;
}

int64_t estimate_device_bytes(const hsba_problem_desc &ds, int precision, bool pcg, bool has_cov) {
  const int64_t t = precision == HSBA_PRECISION_FP64 ? 8 : 4;
  const uint32_t fix = ds.fix_mask & 7u;
  const uint32_t en = enabled_bits(ds.intrinsic_computations_mask & 7u);
  const uint32_t fb = ds.intrinsic_free_mask ? (ds.intrinsic_free_mask & en) : en;
  const int64_t IMG = (fix & HSBA_FIX_IMAGES) ? 0 : 6;
  const int64_t D = (fix & HSBA_FIX_CAMERAS) ? 0 : __builtin_popcount(fb);
  const int64_t E = IMG + D;
  const int64_t N = std::max<int64_t>(0, ds.num_keys), P = std::max<int64_t>(0, ds.num_points);
  const int64_t I = std::max<int64_t>(0, ds.num_images), C = std::max<int64_t>(0, ds.num_cameras);
  // camera-pair incidences: one per point and pair of images seeing it
  int64_t inc = 0;
  if (pcg) {
    inc = N;
  } else if (ds.feature_maps) {
    std::vector<int32_t> k(P, 0);
    for (int64_t q = 0; q < N; ++q) {
      const int32_t p = ds.feature_maps[2 * q + 1];
      if (p >= 0 && p < P) ++k[p];
    }
    for (int64_t p = 0; p < P; ++p) inc += (int64_t)k[p] * (k[p] + 1) / 2;
  }
  const int64_t npairs = pcg ? I : std::min<int64_t>(I * (I + 1) / 2, inc);
  const int64_t n = IMG * I + D * C, ld = (n + 63) / 64 * 64, nb = ld / 64;
  const int64_t S = pcg ? 0 : ld * ld, dinv = pcg ? 0 : nb * 64 * 64;
  const int64_t nblk = pcg ? 0 : I + C;
  const int64_t chunks = (N + kChunk - 1) / kChunk + (3 * P + kChunk - 1) / kChunk + (6 * I + kChunk - 1) / kChunk +
                         (D * C + kChunk - 1) / kChunk + 4;
  int64_t b = 0;
  b += N * (4 * 3 + 16 + 24);
  b += N * (3 * E * t + z_stride((int)E, (int)t) * t) + N * 4;     // W, Z, image key lists
  if (pcg) b += N * (z_stride((int)E, (int)t) * t + 8);             // image-major Z copy
  b += P * (8 + 3 * 8 * 3 + t * 18 + 24);
  b += I * (4 * 5 + 6 * 8 * 2 + t * (2 * E * E + 2 * E) + 8);
  b += C * (4 * 4 + 8 * (2 * D + 20) + t * (D * D + D));
  b += npairs * (8 + 8 + t * E * E) + inc * 8;
  b += (2 * S + dinv) * 8 + ld * (8 * 10 + 8);                       // S, factor, reduced vectors
  b += (4 * npairs + 2 * nblk) * (int64_t)sizeof(Contrib) + (nblk * (nblk + 1) / 2 + 1) * 8;
  b += chunks * (16 + 32) + 4096;
  b += (N / kImgChunk + I) * (12 + t * (E * (E + 1) / 2 + E)) + (I + 1) * 4;  // image-key chunks
  // allocation granules (about 70 buffers) and slack
  return b + b / 8 + 70 * 256 + (1 << 20);
}

}  // namespace hsba
