// Merge-split conditioning of an aggregation (aggregatorType 3): the size
// bounds of AggMIS's fine aggregation.
//
// Reference: MergeSplitConditionerGPU::Condition and the routines it cycles,
// src/core/aggmis/cuda/AggMIS_MergeSplitConditioner.cu:1562-1715 (driver),
// 1985-2420 (MarkMerges, MarkSplits, MarkMergeSplits, MakeSplits, MakeMerges),
// kernels :121-417 and :1380-1545; called by NewMIS,
// src/core/cuda/ComputePermutationMethods.cu:352-431, as
// Condition((min + max) / 2, respectUpper = true, 0.1, 0.1, 10).
//
// Where the reference lets GPU threads race (which of two mutual desires is
// seen first, which of the farthest nodes becomes a split root, which
// border nodes move while balancing), the choices here are fixed: rounds
// evaluate every aggregate against the previous round's state, roots are
// the lowest local index among the farthest nodes, a growing node takes the
// side of its last assigned neighbour in adjacency order, balancing moves
// border nodes in ascending local index.  The result is a deterministic
// function of the graph and the input aggregation.  Nodes a split's search
// cannot reach keep their aggregate (the reference loops forever there).
//
// The aggregation of a level is at most a few million nodes and its graph a
// few tens of millions of edges; the conditioner runs on the host between
// the device MIS / allocation and the device prolongator build.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <queue>
#include <thread>
#include <vector>

#include "fem_internal.hpp"

namespace fem {

namespace {

class MergeSplit {
 public:
  MergeSplit(int64_t n, const int32_t* rp, const int32_t* ci, std::vector<int32_t>& agg, int64_t nagg, int lo,
             int hi)
      : n_(n), rp_(rp), ci_(ci), agg_(agg), nagg_(nagg), lo_(lo), hi_(hi) {
    rebuild();
  }

  int64_t condition() {
    const int desired = (lo_ + hi_) / 2;
    const float tolerance = 0.1f, min_improve = 0.1f;
    const int max_cycles = 10;
    cycle_merges(false);
    cycle_splits(false);
    cycle_merge_splits(min_improve, desired);
    float ratio = (float)outsized_ / (float)nagg_;
    bool high = false;
    for (int cycle = 0; ratio > tolerance && cycle < max_cycles; ++cycle) {
      if (high) cycle_merges(true);
      else cycle_splits(true);
      cycle_merge_splits(min_improve, desired);
      if (!high) ratio = (float)outsized_ / (float)nagg_;  // respectUpper: read after the low cycles
      high = !high;
    }
    cycle_splits(true);
    cycle_merges(false);
    compact();
    return nagg_;
  }

 private:
  // sizes and the aggregate adjacency (sorted, distinct, no self), in time
  // linear in the graph: members grouped by aggregate, each aggregate's
  // neighbours collected once (stamp) and sorted (a few tens each)
  void rebuild() {
    const size_t na = (size_t)nagg_;
    size_.assign(na, 0);
    for (int64_t i = 0; i < n_; ++i) ++size_[(size_t)agg_[(size_t)i]];
    std::vector<int64_t> mptr(na + 1, 0);
    for (size_t a = 0; a < na; ++a) mptr[a + 1] = mptr[a] + size_[a];
    std::vector<int32_t> members((size_t)n_);
    {
      std::vector<int64_t> at(mptr.begin(), mptr.end() - 1);
      for (int64_t i = 0; i < n_; ++i) members[(size_t)at[(size_t)agg_[(size_t)i]]++] = (int32_t)i;
    }
    // contiguous ranges of aggregates per worker thread; each list depends on
    // its aggregate only, so the result does not depend on the split
    const int nt = workers(n_);
    std::vector<std::vector<int32_t>> part((size_t)nt);
    std::vector<int32_t> cnt(na, 0);
    auto run = [&](int w) {
      const size_t a0 = na * (size_t)w / (size_t)nt, a1 = na * (size_t)(w + 1) / (size_t)nt;
      std::vector<int32_t> stamp(na, -1);
      std::vector<int32_t>& out = part[(size_t)w];
      for (size_t a = a0; a < a1; ++a) {
        const size_t first = out.size();
        for (int64_t t = mptr[a]; t < mptr[a + 1]; ++t) {
          const int32_t i = members[(size_t)t];
          for (int32_t q = rp_[i]; q < rp_[i + 1]; ++q) {
            const int32_t b = agg_[(size_t)ci_[q]];
            if (b != (int32_t)a && stamp[(size_t)b] != (int32_t)a) {
              stamp[(size_t)b] = (int32_t)a;
              out.push_back(b);
            }
          }
        }
        std::sort(out.begin() + (std::ptrdiff_t)first, out.end());
        cnt[a] = (int32_t)(out.size() - first);
      }
    };
    parallel(nt, run);
    gptr_.assign(na + 1, 0);
    for (size_t a = 0; a < na; ++a) gptr_[a + 1] = gptr_[a] + cnt[a];
    gadj_.resize((size_t)gptr_[na]);
    size_t at = 0;
    for (const auto& p : part) {
      std::copy(p.begin(), p.end(), gadj_.begin() + (std::ptrdiff_t)at);
      at += p.size();
    }
  }

  // worker threads for the linear passes: one per 64K nodes, at most the
  // hardware threads and 16
  static int workers(int64_t n) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)hw, 16, n / 65536 + 1}));
  }

  // fn(0 .. nt-1) on worker threads; an exception in a worker (allocation
  // failure) is rethrown here after every worker has joined, and workers
  // that cannot be started run on the calling thread
  template <typename Fn>
  static void parallel(int nt, Fn&& fn) {
    std::vector<std::exception_ptr> err((size_t)nt);
    auto guarded = [&](int w) {
      try {
        fn(w);
      } catch (...) {
        err[(size_t)w] = std::current_exception();
      }
    };
    std::vector<std::thread> pool;
    try {
      for (int w = 1; w < nt; ++w) pool.emplace_back(guarded, w);
    } catch (...) {
    }
    for (int w = 1 + (int)pool.size(); w < nt; ++w) guarded(w);
    guarded(0);
    for (auto& t : pool) t.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
  }

  bool out_of_bounds(int s) const { return s < lo_ || s > hi_; }

  int count_outsized() const {
    int k = 0;
    for (int s : size_) k += out_of_bounds(s) ? 1 : 0;
    return k;
  }

  // FindDesirableMerges (kind 0: unforced, 1: forced) and
  // FindDesirableMergeSplits (kind 2)
  float want(int kind, int desired, int32_t a, int32_t b) const {
    const int sa = size_[(size_t)a], sb = size_[(size_t)b], tot = sa + sb;
    if (kind == 2) {
      const bool legal = (out_of_bounds(sa) || out_of_bounds(sb)) && tot > 2 * lo_ && tot < 2 * hi_;
      return legal ? (float)(1.0 / std::abs(desired - tot)) : 0.0f;
    }
    float d = (sa < lo_ ? (float)(lo_ - sa) : 0.0f) + (sb < lo_ ? (float)(lo_ - sb) : 0.0f);
    if (tot > hi_) d = (float)((double)d * (kind == 1 ? 1.0 / (double)(tot - hi_) : 0.0));
    return d;
  }

  // MarkMerges / MarkMergeSplits: partner[a] = the lower aggregate that a
  // (the higher of a mutual pair) merges into, else -1; returns the count
  int64_t pair_up(int kind, int desired, std::vector<int32_t>& partner) const {
    const size_t na = (size_t)nagg_;
    partner.assign(na, -1);
    std::vector<char> busy(na, 0);  // merging, or without a possible merge
    std::vector<int32_t> pick(na, -1);
    for (;;) {
      for (size_t a = 0; a < na; ++a) {
        pick[a] = -1;
        if (busy[a]) continue;
        float best = 0.0f;
        for (int32_t q = gptr_[a]; q < gptr_[a + 1]; ++q) {
          const int32_t b = gadj_[(size_t)q];
          if (busy[(size_t)b]) continue;
          const float d = want(kind, desired, (int32_t)a, b);
          if (d > best) {
            best = d;
            pick[a] = b;
          }
        }
      }
      bool waiting = false, paired = false;
      std::vector<size_t> newly;
      for (size_t a = 0; a < na; ++a) {
        if (busy[a]) continue;
        const int32_t b = pick[a];
        if (b < 0) {
          busy[a] = 1;
        } else if (pick[(size_t)b] == (int32_t)a) {
          if ((size_t)b > a) partner[(size_t)b] = (int32_t)a;
          newly.push_back(a);
          paired = true;
        } else {
          waiting = true;
        }
      }
      for (size_t a : newly) busy[a] = 1;
      if (!waiting || !paired) break;
    }
    int64_t k = 0;
    for (int32_t p : partner) k += p >= 0 ? 1 : 0;
    return k;
  }

  // MakeMerges: compacted ids; a merging aggregate takes its partner's id
  void merge(const std::vector<int32_t>& partner) {
    std::vector<int32_t> shift((size_t)nagg_);
    int32_t run = 0;
    for (int64_t a = 0; a < nagg_; ++a) {
      if (partner[(size_t)a] >= 0) ++run;
      shift[(size_t)a] = run;
    }
    for (int64_t i = 0; i < n_; ++i) {
      int32_t& g = agg_[(size_t)i];
      const int32_t t = partner[(size_t)g] >= 0 ? partner[(size_t)g] : g;
      g = t - shift[(size_t)t];
    }
    nagg_ -= run;
    rebuild();
  }

  bool flag_splits(bool force, std::vector<char>& flag) const {
    flag.assign((size_t)nagg_, 0);
    bool any = false;
    for (int64_t a = 0; a < nagg_; ++a) {
      const int s = size_[(size_t)a];
      flag[(size_t)a] = (s > hi_ && (force || s > 2 * lo_)) ? 1 : 0;
      any = any || flag[(size_t)a];
    }
    return any;
  }

  // breadth-first distances from root over the local adjacency; returns the
  // lowest local index among the farthest reached nodes
  static int farthest(const std::vector<int>& lp, const std::vector<int>& lj, int root, std::vector<int>& dist) {
    const int m = (int)lp.size() - 1;
    dist.assign((size_t)m, -1);
    dist[(size_t)root] = 0;
    int far = root;
    std::vector<int> front{root}, next;
    for (int r = 1; !front.empty(); ++r) {
      next.clear();
      for (int t = 0; t < m; ++t) {
        if (dist[(size_t)t] != -1) continue;
        for (int q = lp[(size_t)t]; q < lp[(size_t)t + 1]; ++q)
          if (dist[(size_t)lj[(size_t)q]] == r - 1) {
            next.push_back(t);
            break;
          }
      }
      for (int t : next) dist[(size_t)t] = r;
      if (!next.empty()) far = next.front();  // ascending t: the lowest index at the largest distance
      front.swap(next);
    }
    return far;
  }

  // MakeSplits: two-root split of every flagged aggregate of at most 256
  // nodes; the ones of up to 64 nodes take the first new ids
  void split(const std::vector<char>& flag) {
    std::vector<int32_t> order;
    for (int pass = 0; pass < 2; ++pass)
      for (int64_t a = 0; a < nagg_; ++a) {
        const int s = size_[(size_t)a];
        if (flag[(size_t)a] && (pass == 0 ? s <= 64 : (s > 64 && s <= 256))) order.push_back((int32_t)a);
      }
    if (order.empty()) return;
    // members of every aggregate, ascending node id
    std::vector<int64_t> mptr((size_t)nagg_ + 1, 0);
    for (int64_t i = 0; i < n_; ++i) ++mptr[(size_t)agg_[(size_t)i] + 1];
    for (int64_t a = 0; a < nagg_; ++a) mptr[(size_t)a + 1] += mptr[(size_t)a];
    std::vector<int32_t> members((size_t)n_);
    {
      std::vector<int64_t> at(mptr.begin(), mptr.end() - 1);
      for (int64_t i = 0; i < n_; ++i) members[(size_t)at[(size_t)agg_[(size_t)i]]++] = (int32_t)i;
    }
    std::vector<int32_t> result(agg_);
    const int64_t base = nagg_;
    // each flagged aggregate splits on its own (reads the old aggregation,
    // writes its members only): contiguous ranges of them per worker thread
    const int nt = workers(n_);
    auto run = [&](int w) {
    const size_t k0 = order.size() * (size_t)w / (size_t)nt, k1 = order.size() * (size_t)(w + 1) / (size_t)nt;
    std::vector<int> lp, lj, dist, side, nd, ns;
    for (size_t k = k0; k < k1; ++k) {
      const int32_t cur = order[k];
      const int32_t fresh = (int32_t)(base + (int64_t)k);
      const int32_t* ids = members.data() + mptr[(size_t)cur];
      const int m = (int)(mptr[(size_t)cur + 1] - mptr[(size_t)cur]);
      lp.assign((size_t)m + 1, 0);
      lj.clear();
      for (int t = 0; t < m; ++t) {
        for (int32_t q = rp_[ids[t]]; q < rp_[ids[t] + 1]; ++q) {
          const int32_t* hit = std::lower_bound(ids, ids + m, ci_[q]);
          if (hit != ids + m && *hit == ci_[q]) lj.push_back((int)(hit - ids));
        }
        lp[(size_t)t + 1] = (int)lj.size();
      }
      const int ra = farthest(lp, lj, 0, dist);
      const int rb = farthest(lp, lj, ra, dist);
      side.assign((size_t)m, -1);
      dist.assign((size_t)m, -1);
      side[(size_t)ra] = cur;
      dist[(size_t)ra] = 0;
      int na = 1, nb = 0;
      if (rb != ra) {
        side[(size_t)rb] = fresh;
        dist[(size_t)rb] = 0;
        nb = 1;
      }
      // grow both sides one layer per round
      for (bool grew = true; grew;) {
        grew = false;
        nd = dist;
        ns = side;
        for (int t = 0; t < m; ++t) {
          if (side[(size_t)t] != -1) continue;
          for (int q = lp[(size_t)t]; q < lp[(size_t)t + 1]; ++q) {
            const int u = lj[(size_t)q];
            if (side[(size_t)u] != -1) {
              ns[(size_t)t] = side[(size_t)u];
              nd[(size_t)t] = dist[(size_t)u] + 1;
            }
          }
          if (ns[(size_t)t] != -1) {
            grew = true;
            (ns[(size_t)t] == fresh ? nb : na) += 1;
          }
        }
        dist.swap(nd);
        side.swap(ns);
      }
      // balance: border nodes of the larger side move, deepest first excluded
      const bool to_cur = na < nb;
      const int quota = std::abs(na - nb) / 2;
      for (int moved = 0; moved < quota;) {
        bool any = false;
        nd = dist;
        ns = side;
        for (int t = 0; t < m && moved < quota; ++t) {
          const int st = side[(size_t)t];
          if (st == -1) continue;
          bool movable = to_cur ? st == fresh : st == cur;
          bool border = false;
          int through = 0x7fffffff;
          for (int q = lp[(size_t)t]; q < lp[(size_t)t + 1]; ++q) {
            const int u = lj[(size_t)q];
            const int su = side[(size_t)u];
            if (su == -1) continue;
            if (su == st && dist[(size_t)u] > dist[(size_t)t]) movable = false;
            if (su != st) {
              border = true;
              through = std::min(through, dist[(size_t)u] + 1);
            }
          }
          if (movable && border) {
            ns[(size_t)t] = to_cur ? cur : fresh;
            nd[(size_t)t] = through;
            ++moved;
            any = true;
          }
        }
        dist.swap(nd);
        side.swap(ns);
        if (!any) break;
      }
      for (int t = 0; t < m; ++t)
        if (side[(size_t)t] != -1) result[(size_t)ids[t]] = side[(size_t)t];
    }
    };
    parallel(nt, run);
    agg_.swap(result);
    nagg_ = base + (int64_t)order.size();
    rebuild();
  }

  // ids without nodes (a split whose two roots coincide) are removed, the
  // others keep their order
  void compact() {
    std::vector<int32_t> id((size_t)nagg_, -1);
    int32_t next = 0;
    for (int64_t a = 0; a < nagg_; ++a)
      if (size_[(size_t)a] > 0) id[(size_t)a] = next++;
    if (next == nagg_) return;
    for (int32_t& g : agg_) g = id[(size_t)g];
    nagg_ = next;
    rebuild();
  }

  void cycle_merges(bool force) {
    std::vector<int32_t> partner;
    while (pair_up(force ? 1 : 0, 0, partner) > 0) merge(partner);
  }

  void cycle_splits(bool force) {
    std::vector<char> flag;
    for (int k = 0; k < 10 && flag_splits(force, flag); ++k) split(flag);
  }

  bool merge_split_once(int desired) {
    std::vector<int32_t> partner;
    if (pair_up(2, desired, partner) == 0) return false;
    merge(partner);
    std::vector<char> flag;
    if (flag_splits(false, flag)) split(flag);
    return true;
  }

  void cycle_merge_splits(float min_improve, int desired) {
    bool done = merge_split_once(desired);
    int after = count_outsized();
    float gain = (float)(outsized_ - after) / (float)outsized_;
    outsized_ = after;
    while (gain > min_improve && done) {
      done = merge_split_once(desired);
      after = count_outsized();
      gain = (float)(outsized_ - after) / (float)outsized_;
      outsized_ = after;
    }
  }

  int64_t n_;
  const int32_t* rp_;
  const int32_t* ci_;
  std::vector<int32_t>& agg_;
  int64_t nagg_;
  int lo_, hi_;
  int outsized_ = 0;
  std::vector<int> size_;
  std::vector<int32_t> gptr_, gadj_;
};

}  // namespace

int64_t condition_aggregates(int64_t n, const int32_t* rp, const int32_t* ci, std::vector<int32_t>& agg,
                             int64_t nagg, int min_size, int max_size) {
  if (n == 0 || nagg == 0) return nagg;
  MergeSplit ms(n, rp, ci, agg, nagg, min_size, max_size);
  return ms.condition();
}

// ---- aggregatorType 4 and 5: flood-fill MIS and aggregation to the nearest root (host) ----

int64_t flood_fill_starter(int64_t n, unsigned long long seed) {
  // the reference starts at rand() % n after srand(time(NULL)); here a
  // splitmix64 step of the seed, so the result is reproducible
  uint64_t z = seed + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return n > 0 ? (int64_t)(z % (uint64_t)n) : 0;
}

// Distance-k maximal independent set by flood fill (AggMIS::MIS::FloodFillMIS,
// src/core/aggmis/cuda/AggMIS_MIS_CPU.cu:25-140).  A root is taken from the
// fringe (nodes just beyond distance k of the roots so far, most often
// reached first), its k-neighbourhood is marked out breadth-first, and nodes
// at distance k + 1 join the fringe; when the fringe runs dry the lowest
// undecided node starts the next component.  The reference leaves ties in
// the fringe's visit counts to its heap layout; here they go in push order.
std::vector<int32_t> flood_fill_mis(int64_t n, const int32_t* rp, const int32_t* ci, int k, int64_t starter) {
  std::vector<int32_t> mis((size_t)n, -1), visits((size_t)n, 0), dist((size_t)n, 1000);
  if (n == 0) return mis;
  struct Entry {
    int32_t node, visits;
    int64_t seq;
  };
  struct Lower {  // the top is the most visits, then the earliest push
    bool operator()(const Entry& a, const Entry& b) const {
      return a.visits != b.visits ? a.visits < b.visits : a.seq > b.seq;
    }
  };
  std::priority_queue<Entry, std::vector<Entry>, Lower> fringe;
  std::queue<int32_t> frontier;
  int64_t seq = 0, cursor = 0, start = starter;
  while (true) {
    if (mis[(size_t)start] == -1) fringe.push(Entry{(int32_t)start, 1, seq++});
    while (true) {
      int32_t root = -1;
      while (!fringe.empty()) {
        const Entry c = fringe.top();
        fringe.pop();
        if (dist[(size_t)c.node] > k) {
          root = c.node;
          break;
        }
      }
      if (root < 0) break;
      mis[(size_t)root] = 1;
      dist[(size_t)root] = 0;
      for (int32_t q = rp[root]; q < rp[root + 1]; ++q) {
        const int32_t v = ci[q];
        if (dist[(size_t)v] > 1) {
          dist[(size_t)v] = 1;
          frontier.push(v);
        }
      }
      while (!frontier.empty()) {
        const int32_t u = frontier.front();
        frontier.pop();
        const int32_t d = dist[(size_t)u];
        mis[(size_t)u] = 0;
        for (int32_t q = rp[u]; q < rp[u + 1]; ++q) {
          const int32_t v = ci[q];
          if (d < k && dist[(size_t)v] > d + 1) {
            dist[(size_t)v] = d + 1;
            frontier.push(v);
          } else if (d == k && dist[(size_t)v] >= d + 1) {
            dist[(size_t)v] = d + 1;
            fringe.push(Entry{v, ++visits[(size_t)v], seq++});
          }
        }
      }
    }
    // decided nodes stay decided, so the lowest undecided one is found by a
    // cursor that only moves forward
    while (cursor < n && mis[(size_t)cursor] != -1) ++cursor;
    if (cursor == n) break;
    start = cursor;
  }
  return mis;
}

// Every node to an aggregate of the nearest root (AggMIS::Aggregation::
// AggregateToNearest, src/core/aggmis/cuda/AggMIS_Aggregation_CPU.cu:4-92):
// roots are numbered in index order and queue their neighbours; a queued
// node still unassigned joins the assigned neighbour of least root distance
// (the first in adjacency order among equals) and queues its unassigned
// neighbours.  Returns the number of aggregates.
int64_t aggregate_to_nearest(int64_t n, const int32_t* rp, const int32_t* ci, const std::vector<int32_t>& roots,
                             std::vector<int32_t>& agg) {
  agg.assign((size_t)n, -1);
  std::vector<int32_t> rd((size_t)n, -1);
  std::queue<int32_t> todo;
  int32_t next = 0;
  for (int64_t i = 0; i < n; ++i)
    if (roots[(size_t)i] == 1) {
      rd[(size_t)i] = 0;
      agg[(size_t)i] = next++;
      for (int32_t q = rp[i]; q < rp[i + 1]; ++q) todo.push(ci[q]);
    }
  while (!todo.empty()) {
    const int32_t u = todo.front();
    todo.pop();
    if (agg[(size_t)u] != -1 && rd[(size_t)u] != -1) continue;
    int32_t best_agg = -1, best_d = -1;
    for (int32_t q = rp[u]; q < rp[u + 1]; ++q) {
      const int32_t v = ci[q];
      const int32_t d = rd[(size_t)v];
      if (d >= 0) {
        if (best_d == -1 || d < best_d) {
          best_d = d;
          best_agg = agg[(size_t)v];
        }
      } else {
        todo.push(v);
      }
    }
    agg[(size_t)u] = best_agg;
    rd[(size_t)u] = best_d + 1;
  }
  return next;
}

}  // namespace fem
