// Block executor of the CPU build of the library (shim/hip/hip_runtime.h):
// a launch runs the grid's blocks in order, each on blockDim real threads.
// Test infrastructure only.
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace hipcpu {

thread_local Idx thread_idx, block_idx;
dim3 block_dim, grid_dim;

namespace {

struct Block {
  std::unique_ptr<std::barrier<>> bar;                 // all threads of the block
  std::vector<std::unique_ptr<std::barrier<>>> waves;  // the threads of each 64-lane wave
  std::vector<unsigned long long> slots;               // shuffle exchange, one per thread
};

Block *g_block = nullptr;
thread_local int tl_tid = 0;
std::mutex g_launch;

}  // namespace

void block_sync() { g_block->bar->arrive_and_wait(); }

unsigned long long wave_exchange(unsigned long long v, int mask) {
  const int w = tl_tid / 64, lane = tl_tid % 64;
  g_block->slots[tl_tid] = v;
  g_block->waves[w]->arrive_and_wait();
  const int src = w * 64 + (lane ^ mask);
  const unsigned long long r = src < (int)g_block->slots.size() ? g_block->slots[src] : v;
  g_block->waves[w]->arrive_and_wait();
  return r;
}

void launch(const std::function<void()> &body, dim3 grid, dim3 block) {
  std::lock_guard<std::mutex> lock(g_launch);
  const int nt = (int)(block.x * block.y * block.z);
  const long long nb = (long long)grid.x * grid.y * grid.z;
  if (nt <= 0 || nb <= 0) return;
  block_dim = block;
  grid_dim = grid;
  Block B;
  B.slots.assign(nt, 0);
  g_block = &B;
  std::barrier<> pool(nt);
  auto worker = [&](int tid) {
    tl_tid = tid;
    thread_idx.x = tid % block.x;
    thread_idx.y = (tid / block.x) % block.y;
    thread_idx.z = tid / (block.x * block.y);
    for (long long b = 0; b < nb; ++b) {
      if (tid == 0) {
        B.bar = std::make_unique<std::barrier<>>(nt);
        B.waves.clear();
        for (int w = 0; w * 64 < nt; ++w)
          B.waves.push_back(std::make_unique<std::barrier<>>(std::min(64, nt - w * 64)));
      }
      pool.arrive_and_wait();
      block_idx.x = (unsigned)(b % grid.x);
      block_idx.y = (unsigned)((b / grid.x) % grid.y);
      block_idx.z = (unsigned)(b / ((long long)grid.x * grid.y));
      body();
      // a thread that has finished no longer takes part in the block's barriers
      B.waves[tid / 64]->arrive_and_drop();
      B.bar->arrive_and_drop();
      pool.arrive_and_wait();
    }
  };
  std::vector<std::thread> threads;
  threads.reserve(nt);
  for (int t = 0; t < nt; ++t) threads.emplace_back(worker, t);
  for (auto &t : threads) t.join();
  g_block = nullptr;
}

}  // namespace hipcpu
