// Lockstep wavefront execution for the CPU emulation header (hip/hip_runtime.h).
#include "hip/hip_runtime.h"
#include <barrier>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
thread_local uint3e blockIdx, threadIdx, blockDim, gridDim;
thread_local int emu_lane = 0;

namespace {
struct Pool {
  std::vector<std::thread> th;
  std::barrier<> start{65}, end{65};
  const std::function<void()>* job = nullptr;
  dim3 grid, block;
  unsigned bx = 0, by = 0, bz = 0, wave = 0;
  std::barrier<>* wave_bar = nullptr;
  float slots[64];
  bool quit = false;
  Pool() {
    for (int l = 0; l < 64; ++l) th.emplace_back([this, l] { worker(l); });
  }
  ~Pool() {
    quit = true;
    start.arrive_and_wait();
    for (auto& t : th) t.join();
  }
  void worker(int l) {
    emu_lane = l;
    for (;;) {
      start.arrive_and_wait();
      if (quit) return;
      const unsigned nthreads = block.x * block.y * block.z;
      const unsigned t = wave * 64 + l;
      if (t < nthreads) {
        gridDim = {grid.x, grid.y, grid.z};
        blockDim = {block.x, block.y, block.z};
        blockIdx = {bx, by, bz};
        threadIdx = {t % block.x, (t / block.x) % block.y, t / (block.x * block.y)};
        (*job)();
      }
      wave_bar->arrive_and_drop();
      end.arrive_and_wait();
    }
  }
};
Pool& pool() { static Pool p; return p; }
}  // namespace

float emu_exchange(float v, int src) {
  Pool& p = pool();
  p.slots[emu_lane] = v;
  p.wave_bar->arrive_and_wait();
  const float r = p.slots[src];
  p.wave_bar->arrive_and_wait();
  return r;
}

void emu_run_grid(dim3 g, dim3 b, const std::function<void()>& job) {
  Pool& p = pool();
  p.job = &job;
  p.grid = g;
  p.block = b;
  const unsigned waves = (b.x * b.y * b.z + 63) / 64;
  for (unsigned z = 0; z < g.z; ++z)
    for (unsigned y = 0; y < g.y; ++y)
      for (unsigned x = 0; x < g.x; ++x)
        for (unsigned w = 0; w < waves; ++w) {
          p.bx = x; p.by = y; p.bz = z; p.wave = w;
          std::barrier<> bar(64);
          p.wave_bar = &bar;
          p.start.arrive_and_wait();
          p.end.arrive_and_wait();
        }
}

namespace {
std::mutex launch_mutex;
std::map<std::string, long>& launch_counts() { static std::map<std::string, long> m; return m; }
}  // namespace

void emu_note_launch(const char* name) {
  std::lock_guard<std::mutex> lock(launch_mutex);
  ++launch_counts()[name];
}

// launches so far of the kernels whose launch expression contains `part`
extern "C" long emu_kernel_launches(const char* part) {
  std::lock_guard<std::mutex> lock(launch_mutex);
  long n = 0;
  for (const auto& kv : launch_counts())
    if (std::strstr(kv.first.c_str(), part)) n += kv.second;
  return n;
}
