// Host-thread model of the queue/counter protocol of the persistent batch
// kernel (fd3d_tsn_amd/csrc/batch.hip, k_batch_persistent), which the
// library refuses to launch: on MI355X its first steps did not finish.
//
// Threads stand in for resident workgroups.  The protocol is the kernel's:
//   * work items are numbered stage-major (all members' items of stage 0,
//     then of stage 1, ...; a member with no items in a stage has no
//     segment there) and taken from a shared head counter;
//   * an item of stage s of member p first waits until every item of the
//     stage before s in that member's stage order (dep) has been counted
//     done, polling the per-(member, stage) done counter with relaxed
//     loads, then issues an acquire fence; the wait is bounded in
//     wall-clock time, and a wait that runs out sets a sticky give-up flag
//     that ends every other wait at once;
//   * after its work an item releases its stores and increments its own
//     (member, stage) counter;
//   * one launch per step; the host resets the head and the counters
//     between launches.
// The model checks, for every item, that each item of its member's earlier
// stages had finished when it started (the dependency the counters
// enforce), that no wait ran out, and that every launch drains.
//
// With --stale the polls of a thread see only the increments made by
// threads of its own group of eight (a per-group copy of each counter and
// of the flag, as a polling workgroup would see if its XCD's L2 kept a
// stale copy of a coarse-grained line): then waits on items done in other
// groups can only run out.
//
// Usage: persistent_model [--stale] members stages workers steps seed wait_ms
// Prints one line: "drained=<0|1> timeouts=<n> violations=<n> items=<n> max_launch_ms=<t>".
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

struct Seg { int start, member, stage; };

struct Model {
  int P = 0, K = 0;
  std::vector<int> n;      // items of (member, stage): n[p * K + s]
  std::vector<int> dep;    // stage before s in member p's order, or -1
  std::vector<Seg> seg;
  int total = 0;
  // shared state of one launch
  std::atomic<unsigned> head{0}, flag{0};
  std::vector<std::atomic<unsigned>> cnt;   // done items of (member, stage)
  std::vector<std::atomic<unsigned>> fin;   // per item: 1 once finished
  // --stale: each group of eight threads sees its own copies
  bool stale = false;
  int groups = 1;
  std::vector<std::atomic<unsigned>> gcnt, gflag;
  std::atomic<long> timeouts{0}, violations{0};
  long long wait_ns = 0;
  std::atomic<int> arrived{0};   // start barrier: all workers begin together, as a launch's workgroups do

  int find_seg(int item) const {
    int lo = 0, hi = (int)seg.size() - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (seg[mid].start <= item) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
  unsigned load_cnt(int g, int idx) const {
    return stale ? gcnt[(size_t)g * cnt.size() + idx].load(std::memory_order_relaxed)
                 : cnt[idx].load(std::memory_order_relaxed);
  }
  unsigned load_flag(int g) const {
    return stale ? gflag[g].load(std::memory_order_relaxed) : flag.load(std::memory_order_relaxed);
  }
  void worker(int g, int nworkers) {
    using clk = std::chrono::steady_clock;
    arrived.fetch_add(1);
    while (arrived.load() < nworkers) std::this_thread::yield();
    for (;;) {
      const int item = (int)head.fetch_add(1, std::memory_order_relaxed);
      if (item >= total) return;
      const Seg S = seg[find_seg(item)];
      const int p = S.member, s = S.stage, d = dep[p * K + s];
      if (d >= 0) {
        const unsigned need = (unsigned)n[p * K + d];
        const auto t0 = clk::now();
        while (load_cnt(g, p * K + d) < need) {
          std::this_thread::yield();
          if (load_flag(g) != 0u) break;
          if (std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - t0).count() > wait_ns) {
            flag.fetch_or(1u, std::memory_order_relaxed);
            if (stale) gflag[g].fetch_or(1u, std::memory_order_relaxed);
            timeouts.fetch_add(1);
            break;
          }
        }
        std::atomic_thread_fence(std::memory_order_acquire);
      }
      // the work: check that every item of this member's earlier stages has finished
      for (int q = 0; q < (int)seg.size(); ++q) {
        const Seg &E = seg[q];
        if (E.member != p || E.stage >= s) continue;
        for (int it = E.start; it < E.start + n[p * K + E.stage]; ++it)
          if (fin[it].load(std::memory_order_relaxed) == 0u) { violations.fetch_add(1); break; }
      }
      // an item's work takes 2-20 microseconds, so the items of a stage
      // spread over many workers
      const auto tw = clk::now();
      const long long spin_ns = 2000 + (long long)((unsigned)item * 2654435761u % 18000u);
      while (std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - tw).count() < spin_ns) {
      }
      fin[item].store(1u, std::memory_order_relaxed);
      cnt[p * K + s].fetch_add(1u, std::memory_order_release);
      if (stale) gcnt[(size_t)g * cnt.size() + p * K + s].fetch_add(1u, std::memory_order_release);
    }
  }
};

}  // namespace

int main(int argc, char **argv) {
  int a = 1;
  Model m;
  if (a < argc && std::strcmp(argv[a], "--stale") == 0) { m.stale = true; ++a; }
  if (argc - a != 6) {
    std::fprintf(stderr, "usage: %s [--stale] members stages workers steps seed wait_ms\n", argv[0]);
    return 2;
  }
  m.P = std::atoi(argv[a]); m.K = std::atoi(argv[a + 1]);
  const int W = std::atoi(argv[a + 2]), steps = std::atoi(argv[a + 3]);
  const unsigned seed = (unsigned)std::atol(argv[a + 4]);
  m.wait_ns = (long long)(std::atof(argv[a + 5]) * 1e6);
  // item counts: stage sizes of uneven members, some stages absent
  std::mt19937 rng(seed);
  m.n.assign((size_t)m.P * m.K, 0);
  m.dep.assign((size_t)m.P * m.K, -1);
  for (int p = 0; p < m.P; ++p) {
    int prev = -1;
    for (int s = 0; s < m.K; ++s) {
      const int r = (int)(rng() % 10);
      m.n[p * m.K + s] = r < 2 ? 0 : 1 + (int)(rng() % (r < 5 ? 4 : 60));
      if (m.n[p * m.K + s] > 0) { m.dep[p * m.K + s] = prev; prev = s; }
    }
  }
  for (int s = 0; s < m.K; ++s)
    for (int p = 0; p < m.P; ++p)
      if (m.n[p * m.K + s] > 0) { m.seg.push_back({m.total, p, s}); m.total += m.n[p * m.K + s]; }
  m.cnt = std::vector<std::atomic<unsigned>>((size_t)m.P * m.K);
  m.fin = std::vector<std::atomic<unsigned>>((size_t)m.total);
  m.groups = m.stale ? (W + 7) / 8 : 1;
  m.gcnt = std::vector<std::atomic<unsigned>>((size_t)m.groups * m.P * m.K);
  m.gflag = std::vector<std::atomic<unsigned>>((size_t)m.groups);
  double max_ms = 0.0;
  bool drained = true;
  for (int st = 0; st < steps; ++st) {
    // per-step reset (the kernel the graph runs before the persistent launch)
    m.head.store(0);
    m.arrived.store(0);
    for (auto &c : m.cnt) c.store(0);
    for (auto &c : m.gcnt) c.store(0);
    for (auto &f : m.fin) f.store(0);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int w = 0; w < W; ++w) th.emplace_back([&m, w, W] { m.worker(w / 8, W); });
    for (auto &t : th) t.join();
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms > max_ms) max_ms = ms;
    if ((int)m.head.load() < m.total) drained = false;
  }
  std::printf("drained=%d timeouts=%ld violations=%ld items=%d max_launch_ms=%.3f\n", drained ? 1 : 0,
              m.timeouts.load(), m.violations.load(), m.total, max_ms);
  return 0;
}
