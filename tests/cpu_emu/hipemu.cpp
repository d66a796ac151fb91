// Runtime of the CPU HIP emulation (see include/hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <semaphore>

namespace hipemu {

thread_local dim3 threadIdx_, blockIdx_;
dim3 blockDim_, gridDim_;
std::vector<Job> *capture = nullptr;

namespace {

struct Pool {
  std::vector<std::thread> workers;
  std::vector<std::unique_ptr<std::binary_semaphore>> go;
  std::counting_semaphore<1024> done{0};
  const std::function<void()> *body = nullptr;
  unsigned nthreads = 0;
  bool stop = false;
  dim3 cur_block;
  std::unique_ptr<std::barrier<>> block_bar;
  std::vector<std::unique_ptr<std::barrier<>>> wave_bar;
  std::vector<unsigned char> xbuf;  // 8 bytes per thread

  Pool() {
    const unsigned n = 1024;
    xbuf.resize(8 * n);
    for (unsigned t = 0; t < n; ++t) go.push_back(std::make_unique<std::binary_semaphore>(0));
    for (unsigned t = 0; t < n; ++t) workers.emplace_back([this, t] { loop(t); });
  }
  ~Pool() {
    stop = true;
    for (auto &g : go) g->release();
    for (auto &w : workers) w.join();
  }
  void loop(unsigned t) {
    for (;;) {
      go[t]->acquire();
      if (stop) return;
      threadIdx_ = dim3(t, 0, 0);
      blockIdx_ = cur_block;
      (*body)();
      block_bar->arrive_and_drop();
      wave_bar[t / 64]->arrive_and_drop();
      done.release();
    }
  }
  void run_block(dim3 b, unsigned n, const std::function<void()> &f) {
    body = &f;
    nthreads = n;
    cur_block = b;
    block_bar = std::make_unique<std::barrier<>>(n);
    wave_bar.clear();
    for (unsigned w = 0; w < (n + 63) / 64; ++w)
      wave_bar.push_back(std::make_unique<std::barrier<>>(std::min<unsigned>(64, n - 64 * w)));
    for (unsigned t = 0; t < n; ++t) go[t]->release();
    for (unsigned t = 0; t < n; ++t) done.acquire();
  }
};

Pool &pool() {
  static Pool p;
  return p;
}

}  // namespace

int lane_id() { return (int)threadIdx_.x; }

void syncthreads() { pool().block_bar->arrive_and_wait(); }

void wave_exchange(const void *in, void *out, int bytes, int src_lane) {
  Pool &p = pool();
  const unsigned t = threadIdx_.x;
  std::memcpy(&p.xbuf[8 * t], in, bytes);
  p.wave_bar[t / 64]->arrive_and_wait();
  std::memcpy(out, &p.xbuf[8 * (unsigned)src_lane], bytes);
  p.wave_bar[t / 64]->arrive_and_wait();
}

void wave_gather(const void *in, void *all, int bytes) {
  Pool &p = pool();
  const unsigned t = threadIdx_.x, base = t & ~63u;
  std::memcpy(&p.xbuf[8 * t], in, bytes);
  p.wave_bar[t / 64]->arrive_and_wait();
  for (unsigned q = 0; q < 64; ++q) std::memcpy((char *)all + q * bytes, &p.xbuf[8 * (base + q)], bytes);
  p.wave_bar[t / 64]->arrive_and_wait();
}

void wave_sync() { pool().wave_bar[threadIdx_.x / 64]->arrive_and_wait(); }

// Kernels from several host threads (the serving loop's workers) run one at
// a time: the pool and the launch dimensions are shared.
std::mutex launch_mu;

void run_kernel(dim3 grid, dim3 block, const std::function<void()> &body) {
  std::lock_guard<std::mutex> g(launch_mu);
  blockDim_ = block;
  gridDim_ = grid;
  for (unsigned by = 0; by < grid.y; ++by)
    for (unsigned bx = 0; bx < grid.x; ++bx) pool().run_block(dim3(bx, by, 0), block.x, body);
}

}  // namespace hipemu
