// Launch plan of the batched LU (host code only; no HIP types, so the plan
// can be compiled and checked on its own).
//
// The LU runs in segments: runs of the largest-first point order that share
// a work-group shape (threads, panel rows per thread).  Each segment runs its
// 64-column block steps K0 = 0, 64, ... in turn; at step K0
//   * the panel kernel takes the first npanel points of the segment, those of
//     padded order above K0 (the order is largest first, so they are a prefix),
//   * the trailing kernel takes ntask (point, c0) tasks from tasks[2 toff ..]:
//     one per strip [c0, c0 + gw) right of the block, for every such point
//     with columns right of the block.
// The strips of one matrix go to one of eight queues (the least loaded,
// matrices largest first) and the queues are interleaved, task b from queue
// b % 8, so that with work-groups dealt to the eight XCDs in turn all strips
// of a matrix share one XCD's L2 (they all read the same L21 and X).  Queues
// shorter than the longest are padded with point -1.
#pragma once
#include <algorithm>
#include <vector>

struct LuStep { int K0, npanel, ntask, toff; };
struct LuSeg { int nt, R, start, cnt; std::vector<LuStep> steps; };

// Split segments until there are `want` of them (one per LU stream), each
// time the one with the most LU work (sum of npad^3) at the point where its
// work halves; a segment of one point is not split.  The parts keep the
// segment's work-group shape and stay runs of the largest-first order.
inline void lu_split(std::vector<LuSeg> &segs, const std::vector<int> &npad, const std::vector<int> &horder,
                     int want) {
  auto work = [&](const LuSeg &g, int from, int to) {
    double w = 0;
    for (int i = from; i < to; ++i) w += (double)npad[horder[g.start + i]] * npad[horder[g.start + i]] *
                                         npad[horder[g.start + i]];
    return w;
  };
  while ((int)segs.size() < want) {
    int best = -1;
    double bw = 0;
    for (int k = 0; k < (int)segs.size(); ++k) {
      const double w = work(segs[k], 0, segs[k].cnt);
      if (segs[k].cnt >= 2 && w > bw) { bw = w; best = k; }
    }
    if (best < 0) break;
    LuSeg a = segs[best];
    double acc = 0;
    int cut = 1;
    for (int i = 0; i < a.cnt - 1; ++i) {
      acc += work(a, i, i + 1);
      cut = i + 1;
      if (acc >= 0.5 * bw) break;
    }
    LuSeg b = a;
    a.cnt = cut;
    b.start = a.start + cut;
    b.cnt -= cut;
    segs[best] = a;
    segs.insert(segs.begin() + best + 1, b);
  }
}

// npad[p]: padded order of point p; horder: points largest first; gw: strip
// width of the trailing kernel (64 FP32, 32 FP64).
inline void lu_plan(std::vector<LuSeg> &segs, const std::vector<int> &npad, const std::vector<int> &horder, int gw,
                    std::vector<int> &tasks) {
  constexpr int KB = 64, NQ = 8;
  tasks.clear();
  for (LuSeg &g : segs) {
    g.steps.clear();
    if (g.cnt <= 0) continue;
    const int top = npad[horder[g.start]];
    for (int K0 = 0; K0 < top; K0 += KB) {
      LuStep st{K0, 0, 0, (int)(tasks.size() / 2)};
      std::vector<int> q[NQ];
      long long load[NQ] = {0};
      for (int i = 0; i < g.cnt; ++i) {
        const int pi = horder[g.start + i], n = npad[pi];
        if (n <= K0) break;
        ++st.npanel;
        const int kend = K0 + std::min(KB, n - K0);
        if (kend >= n) continue;
        int best = 0;
        for (int x = 1; x < NQ; ++x)
          if (load[x] < load[best]) best = x;
        for (int c0 = kend; c0 < n; c0 += gw) {
          q[best].push_back(pi);
          q[best].push_back(c0);
          ++load[best];
        }
      }
      size_t len = 0;
      for (int x = 0; x < NQ; ++x) len = std::max(len, q[x].size() / 2);
      for (size_t i = 0; i < len; ++i)
        for (int x = 0; x < NQ; ++x) {
          const bool have = i < q[x].size() / 2;
          tasks.push_back(have ? q[x][2 * i] : -1);
          tasks.push_back(have ? q[x][2 * i + 1] : 0);
        }
      st.ntask = (int)(NQ * len);
      g.steps.push_back(st);
    }
  }
}
