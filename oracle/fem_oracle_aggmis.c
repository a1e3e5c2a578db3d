/* fem_oracle_aggmis.c -- CPU restatement of AggMIS's merge-split conditioner
 * (aggregatorType 3, "NewMIS" fine aggregation).  Test infrastructure only:
 * never linked into the product.
 *
 * Reference: src/core/aggmis/cuda/AggMIS_MergeSplitConditioner.cu
 *   Condition                 :1625-1715
 *   CycleMerges / CycleSplits :1580-1599, CycleMergeSplits :1600-1624
 *   MarkMerges :1985-2056, MarkSplits :2057-2099, MarkMergeSplits :2100-2162
 *   MakeSplits :2163-2332 (kernel MakeSplits :166-417), MakeMerges :2333-2409
 *   kernels MarkSplits :1380-1393, FindDesirableMerges :1394-1452,
 *   FindDesirableMergeSplits :1453-1506, MarkMerges :1507-1545, MakeMerges :121-141
 * called from NewMIS, src/core/cuda/ComputePermutationMethods.cu:352-431 with
 * Condition((min + max) / 2, true, 0.1, 0.1, 10).
 *
 * The reference resolves several choices through races between GPU threads;
 * this restatement fixes each of them (the library makes the same choices):
 *   - merges: desires are evaluated for all active aggregates against the
 *     previous round's flags, then every mutual pair is marked (rounds until
 *     no desire is left unanswered);
 *   - split roots: the farthest node of a breadth-first search, the lowest
 *     local index among equally far ones;
 *   - split growth: a node joins the side of its last assigned neighbour in
 *     adjacency order (the reference's loop order), nodes the search does not
 *     reach keep the old aggregate (the reference loops forever there);
 *   - split balancing: movable border nodes are taken in ascending local
 *     index order.
 * Aggregates larger than 256 nodes are not split (the reference's split
 * filters cover sizes 0-64 and 65-256 only). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int64_t n;
  const int32_t* ap;
  const int32_t* aj;
  int32_t* agg;
  int64_t nagg;
  int minsz, maxsz;
  int32_t* sizes;  /* nagg */
  int32_t* gp;     /* induced graph of the aggregates, ascending neighbours */
  int32_t* gj;
  int outsized;
} cond_t;

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* sizes and the induced graph (GraphHelpers::GetInducedGraph, getPartSizes) */
static void refresh(cond_t* c) {
  free(c->sizes);
  free(c->gp);
  free(c->gj);
  c->sizes = (int32_t*)calloc((size_t)c->nagg + 1, sizeof(int32_t));
  for (int64_t i = 0; i < c->n; ++i) c->sizes[c->agg[i]]++;
  /* neighbour lists: collect (a, b) pairs then sort and unique per a */
  int64_t m = 0;
  for (int64_t i = 0; i < c->n; ++i)
    for (int32_t q = c->ap[i]; q < c->ap[i + 1]; ++q)
      if (c->agg[c->aj[q]] != c->agg[i]) ++m;
  int32_t* cnt = (int32_t*)calloc((size_t)c->nagg + 1, sizeof(int32_t));
  for (int64_t i = 0; i < c->n; ++i)
    for (int32_t q = c->ap[i]; q < c->ap[i + 1]; ++q)
      if (c->agg[c->aj[q]] != c->agg[i]) cnt[c->agg[i] + 1]++;
  for (int64_t a = 0; a < c->nagg; ++a) cnt[a + 1] += cnt[a];
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m + 1));
  int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c->nagg + 1));
  memcpy(pos, cnt, sizeof(int32_t) * (size_t)(c->nagg + 1));
  for (int64_t i = 0; i < c->n; ++i)
    for (int32_t q = c->ap[i]; q < c->ap[i + 1]; ++q)
      if (c->agg[c->aj[q]] != c->agg[i]) tmp[pos[c->agg[i]]++] = c->agg[c->aj[q]];
  c->gp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c->nagg + 1));
  c->gj = (int32_t*)malloc(sizeof(int32_t) * (size_t)(m + 1));
  int64_t o = 0;
  c->gp[0] = 0;
  for (int64_t a = 0; a < c->nagg; ++a) {
    int32_t s = cnt[a], e = cnt[a + 1];
    qsort(tmp + s, (size_t)(e - s), sizeof(int32_t), cmp_i32);
    for (int32_t q = s; q < e; ++q)
      if (q == s || tmp[q] != tmp[q - 1]) c->gj[o++] = tmp[q];
    c->gp[a + 1] = (int32_t)o;
  }
  free(tmp);
  free(pos);
  free(cnt);
}

static int count_outsized(const cond_t* c) {
  int k = 0;
  for (int64_t a = 0; a < c->nagg; ++a) k += c->sizes[a] < c->minsz || c->sizes[a] > c->maxsz;
  return k;
}

/* desirability of merging aggregate a with neighbour b:
 * mode 0 FindDesirableMerges(force=0), 1 (force=1), 2 FindDesirableMergeSplits */
static float desirability(const cond_t* c, int mode, int desired, int32_t a, int32_t b) {
  int cur = c->sizes[a], nb = c->sizes[b], tot = cur + nb;
  if (mode == 2) {
    int cout = cur < c->minsz || cur > c->maxsz;
    int nout = nb < c->minsz || nb > c->maxsz;
    int legal = (nout || cout) && tot > c->minsz * 2 && tot < c->maxsz * 2;
    return legal ? (float)(1.0 / abs(desired - tot)) : 0.0f;
  }
  float d = 0.0f;
  d += cur < c->minsz ? (float)(c->minsz - cur) : 0.0f;
  d += nb < c->minsz ? (float)(c->minsz - nb) : 0.0f;
  if (tot > c->maxsz) d = (float)((double)d * (mode == 1 ? 1.0 / (tot - c->maxsz) : 0.0));
  return d;
}

/* MarkMerges / MarkMergeSplits: merge_to[a] = lower partner of a merging
 * higher aggregate a, else -1; returns the number of merges marked */
static int64_t mark_merges(const cond_t* c, int mode, int desired, int32_t* merge_to) {
  int64_t na = c->nagg;
  int32_t* want = (int32_t*)malloc(sizeof(int32_t) * (size_t)na);
  char* merging = (char*)calloc((size_t)na, 1);
  for (int64_t a = 0; a < na; ++a) merge_to[a] = -1;
  int64_t marked = 0;
  for (;;) {
    for (int64_t a = 0; a < na; ++a) {
      want[a] = -1;
      if (merging[a]) continue;
      float best = 0.0f;
      for (int32_t q = c->gp[a]; q < c->gp[a + 1]; ++q) {
        int32_t b = c->gj[q];
        if (merging[b]) continue;
        float d = desirability(c, mode, desired, (int32_t)a, b);
        if (d > best) {
          best = d;
          want[a] = b;
        }
      }
    }
    int unanswered = 0, any = 0;
    for (int64_t a = 0; a < na; ++a) {
      if (merging[a]) continue;
      int32_t b = want[a];
      if (b < 0) {
        merging[a] = 1; /* no possible merge: inactive */
        continue;
      }
      if (want[b] == (int32_t)a) {
        if (b > a) merge_to[b] = (int32_t)a;
        else merge_to[a] = b;
        any = 1;
      } else {
        unanswered = 1;
      }
    }
    for (int64_t a = 0; a < na; ++a)
      if (!merging[a] && want[a] >= 0 && want[want[a]] == (int32_t)a) merging[a] = 1;
    if (!unanswered) break;
    if (!any) { /* (cannot happen with symmetric desirability) */
      break;
    }
  }
  for (int64_t a = 0; a < na; ++a) marked += merge_to[a] >= 0;
  free(want);
  free(merging);
  return marked;
}

/* MakeMerges: ids compacted, merged aggregates take their partner's id */
static void make_merges(cond_t* c, const int32_t* merge_to) {
  int64_t na = c->nagg;
  int32_t* off = (int32_t*)malloc(sizeof(int32_t) * (size_t)na);
  int32_t run = 0;
  for (int64_t a = 0; a < na; ++a) {
    run += merge_to[a] >= 0;
    off[a] = run; /* inclusive scan */
  }
  for (int64_t i = 0; i < c->n; ++i) {
    int32_t a = c->agg[i], t = merge_to[a];
    c->agg[i] = t < 0 ? a - off[a] : t - off[t];
  }
  c->nagg = na - run;
  free(off);
  refresh(c);
}

static int64_t mark_splits(const cond_t* c, int force, char* split) {
  int64_t k = 0;
  for (int64_t a = 0; a < c->nagg; ++a) {
    int s = c->sizes[a];
    split[a] = s > c->maxsz && (force || s > c->minsz * 2);
    k += split[a];
  }
  return k;
}

/* BFS over local adjacency from root; returns the lowest local index among
 * the farthest reached nodes (dist[] = -1 for unreached) */
static int bfs_far(int m, const int32_t* lp, const int32_t* lj, int root, int* dist) {
  for (int t = 0; t < m; ++t) dist[t] = -1;
  dist[root] = 0;
  int far = root, fd = 0;
  for (int r = 1;; ++r) {
    int changed = 0;
    for (int t = 0; t < m; ++t) {
      if (dist[t] != -1) continue;
      for (int q = lp[t]; q < lp[t + 1]; ++q)
        if (dist[lj[q]] == r - 1) {
          dist[t] = -2; /* marked this round */
          break;
        }
    }
    for (int t = 0; t < m; ++t)
      if (dist[t] == -2) {
        dist[t] = r;
        changed = 1;
        if (r > fd) {
          fd = r;
          far = t;
        }
      }
    if (!changed) break;
  }
  return far;
}

/* MakeSplits: aggregates flagged, sizes 0-64 first, then 65-256 */
static void make_splits(cond_t* c, const char* split) {
  int64_t na = c->nagg;
  int64_t nsplit = 0;
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(na + 1));
  for (int pass = 0; pass < 2; ++pass)
    for (int64_t a = 0; a < na; ++a) {
      int s = c->sizes[a];
      if (split[a] && (pass == 0 ? s <= 64 : (s >= 65 && s <= 256))) order[nsplit++] = (int32_t)a;
    }
  if (nsplit == 0) {
    free(order);
    return;
  }
  /* nodes of each aggregate in ascending order */
  int32_t* mp = (int32_t*)calloc((size_t)na + 1, sizeof(int32_t));
  for (int64_t i = 0; i < c->n; ++i) mp[c->agg[i] + 1]++;
  for (int64_t a = 0; a < na; ++a) mp[a + 1] += mp[a];
  int32_t* mem = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c->n + 1));
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(na + 1));
  memcpy(fill, mp, sizeof(int32_t) * (size_t)na);
  for (int64_t i = 0; i < c->n; ++i) mem[fill[c->agg[i]]++] = (int32_t)i;
  int32_t* newagg = (int32_t*)malloc(sizeof(int32_t) * (size_t)c->n);
  memcpy(newagg, c->agg, sizeof(int32_t) * (size_t)c->n);
  for (int64_t k = 0; k < nsplit; ++k) {
    int32_t cur = order[k], nid = (int32_t)(na + k);
    int m = mp[cur + 1] - mp[cur];
    const int32_t* ids = mem + mp[cur];
    /* local adjacency (neighbours inside the aggregate, adjacency order) */
    int* lp = (int*)malloc(sizeof(int) * (size_t)(m + 1));
    int cap = 0;
    for (int t = 0; t < m; ++t) cap += c->ap[ids[t] + 1] - c->ap[ids[t]];
    int* lj = (int*)malloc(sizeof(int) * (size_t)(cap + 1));
    int o = 0;
    for (int t = 0; t < m; ++t) {
      lp[t] = o;
      for (int32_t q = c->ap[ids[t]]; q < c->ap[ids[t] + 1]; ++q) {
        int32_t v = c->aj[q];
        /* binary search v in ids */
        int lo = 0, hi = m;
        while (lo < hi) {
          int mid = (lo + hi) / 2;
          if (ids[mid] < v) lo = mid + 1; else hi = mid;
        }
        if (lo < m && ids[lo] == v) lj[o++] = lo;
      }
    }
    lp[m] = o;
    int* dist = (int*)malloc(sizeof(int) * (size_t)m);
    int* side = (int*)malloc(sizeof(int) * (size_t)m);
    int* nd = (int*)malloc(sizeof(int) * (size_t)m);
    int* ns = (int*)malloc(sizeof(int) * (size_t)m);
    int ra = bfs_far(m, lp, lj, 0, dist);
    int rb = bfs_far(m, lp, lj, ra, dist);
    for (int t = 0; t < m; ++t) {
      side[t] = -1;
      dist[t] = -1;
    }
    side[ra] = cur;
    dist[ra] = 0;
    side[rb] = nid;
    dist[rb] = 0;
    int acount = 1, bcount = (rb != ra) ? 1 : 0;
    if (rb == ra) side[ra] = cur;
    for (;;) {
      int changed = 0;
      for (int t = 0; t < m; ++t) {
        nd[t] = dist[t];
        ns[t] = side[t];
        if (side[t] != -1) continue;
        for (int q = lp[t]; q < lp[t + 1]; ++q)
          if (side[lj[q]] != -1) { /* the last assigned neighbour wins */
            nd[t] = dist[lj[q]] + 1;
            ns[t] = side[lj[q]];
          }
        if (ns[t] != -1) {
          changed = 1;
          if (ns[t] == nid) ++bcount; else ++acount;
        }
      }
      memcpy(dist, nd, sizeof(int) * (size_t)m);
      memcpy(side, ns, sizeof(int) * (size_t)m);
      if (!changed) break;
    }
    /* balance: move border nodes from the larger side */
    int diff = acount > bcount ? acount - bcount : bcount - acount;
    int to_a = acount < bcount;
    int tomove = diff / 2, moved = 0;
    while (moved < tomove) {
      int any = 0;
      for (int t = 0; t < m; ++t) {
        nd[t] = dist[t];
        ns[t] = side[t];
      }
      for (int t = 0; t < m && moved < tomove; ++t) {
        if (side[t] == -1) continue;
        int can = to_a ? side[t] == nid : side[t] == cur;
        int border = 0, newd = 0x7fffffff;
        for (int q = lp[t]; q < lp[t + 1]; ++q) {
          int u = lj[q];
          if (side[u] == -1) continue;
          if (side[u] == side[t] && dist[u] > dist[t]) can = 0;
          if (side[u] != side[t]) {
            if (dist[u] + 1 < newd) newd = dist[u] + 1;
            border = 1;
          }
        }
        if (border && can) {
          ns[t] = to_a ? cur : nid;
          nd[t] = newd;
          ++moved;
          any = 1;
        }
      }
      memcpy(dist, nd, sizeof(int) * (size_t)m);
      memcpy(side, ns, sizeof(int) * (size_t)m);
      if (!any) break;
    }
    for (int t = 0; t < m; ++t)
      if (side[t] != -1) newagg[ids[t]] = side[t];
    free(lp);
    free(lj);
    free(dist);
    free(side);
    free(nd);
    free(ns);
  }
  memcpy(c->agg, newagg, sizeof(int32_t) * (size_t)c->n);
  c->nagg = na + nsplit;
  free(newagg);
  free(mem);
  free(fill);
  free(mp);
  free(order);
  refresh(c);
}

static void cycle_merges(cond_t* c, int force) {
  int32_t* mt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c->nagg + 1));
  for (;;) {
    mt = (int32_t*)realloc(mt, sizeof(int32_t) * (size_t)(c->nagg + 1));
    if (mark_merges(c, force ? 1 : 0, 0, mt) == 0) break;
    make_merges(c, mt);
  }
  free(mt);
}

static void cycle_splits(cond_t* c, int force) {
  for (int count = 0; count < 10; ++count) {
    char* sp = (char*)malloc((size_t)c->nagg + 1);
    int64_t k = mark_splits(c, force, sp);
    if (k == 0) {
      free(sp);
      break;
    }
    make_splits(c, sp);
    free(sp);
  }
}

static int make_merge_splits(cond_t* c, int desired) {
  int32_t* mt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c->nagg + 1));
  int64_t k = mark_merges(c, 2, desired, mt);
  if (k == 0) {
    free(mt);
    return 0;
  }
  make_merges(c, mt);
  free(mt);
  char* sp = (char*)malloc((size_t)c->nagg + 1);
  if (mark_splits(c, 0, sp)) make_splits(c, sp);
  free(sp);
  return 1;
}

static void cycle_merge_splits(cond_t* c, float min_improve, int desired) {
  int done = make_merge_splits(c, desired);
  int after = count_outsized(c);
  float improvement = (float)(c->outsized - after) / (float)c->outsized;
  c->outsized = after;
  while (improvement > min_improve && done) {
    done = make_merge_splits(c, desired);
    after = count_outsized(c);
    improvement = (float)(c->outsized - after) / (float)c->outsized;
    c->outsized = after;
  }
}

/* Condition(desired, respectUpper = true, tolerance, minImprove, maxCycles)
 * on aggregation agg (ids 0..nagg-1) of graph (ap, aj); agg is rewritten,
 * the new number of aggregates is returned. */
int64_t orc_condition(int64_t n, const int32_t* ap, const int32_t* aj, int32_t* agg, int64_t nagg, int min_size,
                      int max_size) {
  cond_t c;
  memset(&c, 0, sizeof(c));
  c.n = n;
  c.ap = ap;
  c.aj = aj;
  c.agg = agg;
  c.nagg = nagg;
  c.minsz = min_size;
  c.maxsz = max_size;
  refresh(&c);
  const int desired = (min_size + max_size) / 2;
  const float tolerance = 0.1f, min_improve = 0.1f;
  const int max_cycles = 10;
  cycle_merges(&c, 0);
  cycle_splits(&c, 0);
  cycle_merge_splits(&c, min_improve, desired);
  float ratio = (float)c.outsized / (float)c.nagg;
  int counter = 0, high = 0;
  while (ratio > tolerance && counter++ < max_cycles) {
    if (high) cycle_merges(&c, 1);
    else cycle_splits(&c, 1);
    cycle_merge_splits(&c, min_improve, desired);
    if (!high) ratio = (float)c.outsized / (float)c.nagg; /* respectUpper: updated after low cycles */
    high = !high;
  }
  cycle_splits(&c, 1);
  cycle_merges(&c, 0);
  /* drop ids without nodes (a split whose two roots coincide), keeping order */
  {
    int32_t* id = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c.nagg + 1));
    int32_t next = 0;
    for (int64_t a = 0; a < c.nagg; ++a) id[a] = c.sizes[a] > 0 ? next++ : -1;
    if (next != c.nagg) {
      for (int64_t i = 0; i < n; ++i) agg[i] = id[agg[i]];
      c.nagg = next;
    }
    free(id);
  }
  int64_t out = c.nagg;
  free(c.sizes);
  free(c.gp);
  free(c.gj);
  return out;
}

/* ---- aggregatorType 4 / 5: flood-fill MIS and aggregation to the nearest root ----
 *
 * Reference: AggMIS::MIS::FloodFillMIS, src/core/aggmis/cuda/AggMIS_MIS_CPU.cu:25-140, and
 * AggMIS::Aggregation::AggregateToNearest, src/core/aggmis/cuda/AggMIS_Aggregation_CPU.cu:4-92,
 * called by NewMIS_CPU (type 4, then the merge-split conditioner) and LightMIS_CPU (type 5),
 * src/core/cuda/ComputePermutationMethods.cu:566-860.  The reference starts at rand() % n after
 * srand(time(NULL)); here the start node is an argument.  Ties in the fringe's visit counts,
 * which the reference leaves to std::priority_queue's heap layout, go in push order. */

typedef struct {
  int32_t node, visits;
  int64_t seq;
} orc_fringe_entry;

/* a before b in the fringe: more visits, then earlier push */
static int orc_fringe_first(const orc_fringe_entry* a, const orc_fringe_entry* b) {
  return a->visits != b->visits ? a->visits > b->visits : a->seq < b->seq;
}

typedef struct {
  orc_fringe_entry* e;
  int64_t n, cap;
} orc_heap;

static void orc_heap_push(orc_heap* h, orc_fringe_entry x) {
  if (h->n == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->e = (orc_fringe_entry*)realloc(h->e, (size_t)h->cap * sizeof(orc_fringe_entry));
  }
  int64_t i = h->n++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (!orc_fringe_first(&x, &h->e[p])) break;
    h->e[i] = h->e[p];
    i = p;
  }
  h->e[i] = x;
}

static orc_fringe_entry orc_heap_pop(orc_heap* h) {
  orc_fringe_entry top = h->e[0], last = h->e[--h->n];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= h->n) break;
    if (c + 1 < h->n && orc_fringe_first(&h->e[c + 1], &h->e[c])) ++c;
    if (!orc_fringe_first(&h->e[c], &last)) break;
    h->e[i] = h->e[c];
    i = c;
  }
  if (h->n > 0) h->e[i] = last;
  return top;
}

typedef struct {
  int32_t* v;
  int64_t head, tail, cap;
} orc_fifo;

static void orc_fifo_push(orc_fifo* q, int32_t x) {
  if (q->tail == q->cap) {
    if (q->head > 0) { /* slide the live part to the front */
      memmove(q->v, q->v + q->head, (size_t)(q->tail - q->head) * sizeof(int32_t));
      q->tail -= q->head;
      q->head = 0;
    }
    if (q->tail == q->cap) {
      q->cap = q->cap ? 2 * q->cap : 1024;
      q->v = (int32_t*)realloc(q->v, (size_t)q->cap * sizeof(int32_t));
    }
  }
  q->v[q->tail++] = x;
}

/* mis[i] = 1 for a root, 0 otherwise; every node within distance k of a root, no two roots
 * within distance k of each other. */
void orc_flood_fill_mis(int64_t n, const int32_t* ap, const int32_t* aj, int k, int64_t starter, int32_t* mis) {
  int32_t* visits = (int32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
  int32_t* dist = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i) {
    mis[i] = -1;
    dist[i] = 1000;
  }
  orc_heap fringe = {0};
  orc_fifo frontier = {0};
  int64_t seq = 0, start = starter;
  while (n > 0) {
    if (mis[start] == -1) {
      orc_fringe_entry s0 = {(int32_t)start, 1, seq++};
      orc_heap_push(&fringe, s0);
    }
    for (;;) {
      int32_t root = -1;
      while (fringe.n > 0) {
        orc_fringe_entry c = orc_heap_pop(&fringe);
        if (dist[c.node] > k) {
          root = c.node;
          break;
        }
      }
      if (root < 0) break;
      mis[root] = 1;
      dist[root] = 0;
      for (int32_t q = ap[root]; q < ap[root + 1]; ++q)
        if (dist[aj[q]] > 1) {
          dist[aj[q]] = 1;
          orc_fifo_push(&frontier, aj[q]);
        }
      while (frontier.head < frontier.tail) {
        const int32_t u = frontier.v[frontier.head++];
        const int32_t d = dist[u];
        mis[u] = 0;
        if (d < k) {
          for (int32_t q = ap[u]; q < ap[u + 1]; ++q)
            if (dist[aj[q]] > d + 1) {
              dist[aj[q]] = d + 1;
              orc_fifo_push(&frontier, aj[q]);
            }
        }
        if (d == k) {
          for (int32_t q = ap[u]; q < ap[u + 1]; ++q)
            if (dist[aj[q]] >= d + 1) {
              dist[aj[q]] = d + 1;
              orc_fringe_entry f = {aj[q], ++visits[aj[q]], seq++};
              orc_heap_push(&fringe, f);
            }
        }
      }
    }
    int64_t next = -1;
    for (int64_t i = 0; i < n; ++i)
      if (mis[i] == -1) {
        next = i;
        break;
      }
    if (next < 0) break;
    start = next;
  }
  free(fringe.e);
  free(frontier.v);
  free(visits);
  free(dist);
}

/* agg[i] = aggregate of node i (roots numbered in index order); returns the aggregate count. */
int64_t orc_aggregate_to_nearest(int64_t n, const int32_t* ap, const int32_t* aj, const int32_t* roots,
                                 int32_t* agg) {
  int32_t* rd = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  orc_fifo todo = {0};
  int32_t next = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (roots[i] == 1) {
      rd[i] = 0;
      agg[i] = next++;
      for (int32_t q = ap[i]; q < ap[i + 1]; ++q) orc_fifo_push(&todo, aj[q]);
    } else {
      rd[i] = -1;
      agg[i] = -1;
    }
  }
  while (todo.head < todo.tail) {
    const int32_t u = todo.v[todo.head++];
    if (agg[u] != -1 && rd[u] != -1) continue;
    int32_t best = -1, bd = -1;
    for (int32_t q = ap[u]; q < ap[u + 1]; ++q) {
      const int32_t v = aj[q], d = rd[v];
      if (d >= 0) {
        if (bd == -1) {
          bd = d;
          best = agg[v];
        } else if (d < bd) {
          best = agg[v];
          bd = d;
        }
      } else {
        orc_fifo_push(&todo, v);
      }
    }
    agg[u] = best;
    rd[u] = bd + 1;
  }
  free(todo.v);
  free(rd);
  return next;
}
