// rtree.hip -- rt_rcount's R-tree: host build and device count.
//
// Build (host).  Points go in one at a time, in order, as insertDataRTree
// puts them (src/objects/RTree/RTree.c:35-49, rt_rcount/main.c:711-715), by
// Guttman's insertion with the quadratic split as src/objects/RTree/
// Build.c implements it; every FP32 operation that decides where a point
// goes is the reference's, in its order:
//   area      ((1 * dx) * dy) * dz (Mbr.c:50-70; the SSE branch multiplies
//             in the same order)
//   union     low = a < b ? a : b, high = a > b ? a : b per axis, the minps/
//             maxps of enlargeMbr (Mbr.c:136-144); a node's own box is
//             getMbrRTreeNode's scalar min/max (Node.c:74-116)
//   descent   the child of least |area(union) - area|, ties to the least
//             area, further ties to the first (ChooseSubTree, Build.c:
//             271-320, whose glibc qsort is a stable merge sort); when a
//             NaN makes those comparisons inconsistent, that merge order
//             itself
//   split     seeds: the first pair of greatest area(union) - area(a) -
//             area(b) (PickSeeds, :42-75); then the entry of greatest
//             |d1 - d2| goes to the group it enlarges less, on a tie to the
//             one with fewer entries, the second on equal sizes (PickNext /
//             DistributeEntry, :77-140), until a group holds MAX_ENT -
//             MIN_ENT + 1 entries and the other takes the rest (:147-189)
//   adjust    the parent's box of each node on the path, a split sibling
//             added or split in turn, a new root on a root split (:191-241);
//             adding an entry refreshes the parent's box of the node from its
//             entries before the new one counts (Node.c:54-72), as the
//             reference does.
//
// Count (device).  multiRangeCountWithPriodicBoundRTree (src/objects/RTree/
// MultiRangeCount.c:13-74): the query box c -+ r and, for each axis whose
// box leaves [0, 1000], one more query per non-empty subset of those axes
// with the centre moved by 1000; MultiRangeCountRecursive (:86-114): an
// inner node descends into the children whose box overlaps the query box
// (cmpnlt both ways, Mbr.c:167-182); a leaf tests every entry's distance to
// the (moved) centre with no box test, counts it in the first shell k with
// dist < r2[k], and stops at the first entry at distance exactly 0, which
// is not counted nor is any entry after it in that leaf.
//
// One query per 16 lanes: a node's (at most 10) entries are tested side by
// side, the children to visit are pushed on the query's stack in LDS, a
// leaf's first zero distance is found by a ballot.  Counts go to per-query
// LDS bins and are stored once per halo after its last image.  The queries
// come in Morton order, and each XCD takes one contiguous range of them, so
// queries that share nodes run close together on one L2.
#include "rtree.h"

#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>

#include "pairs.h"

#pragma clang fp contract(off)

namespace sjcs {

// ---- host build ----

namespace {

inline float box_area(const RtBox& b) {
  float a = 1.0f;
  a *= b.hi[0] - b.lo[0];
  a *= b.hi[1] - b.lo[1];
  a *= b.hi[2] - b.lo[2];
  return a;
}

inline RtBox box_union(RtBox a, const RtBox& e) {
  for (int d = 0; d < 3; ++d) {
    a.lo[d] = a.lo[d] < e.lo[d] ? a.lo[d] : e.lo[d];
    a.hi[d] = a.hi[d] > e.hi[d] ? a.hi[d] : e.hi[d];
  }
  return a;
}

// glibc's qsort order for these small arrays: halves sorted recursively, the
// merge takes the left element unless it compares greater than the right.
void merge_order(int* idx, int n, const float* key, int* tmp) {
  if (n <= 1) return;
  const int h = n / 2;
  merge_order(idx, h, key, tmp);
  merge_order(idx + h, n - h, key, tmp);
  int i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = key[idx[i]] > key[idx[j]] ? idx[j++] : idx[i++];
  while (i < h) tmp[k++] = idx[i++];
  while (j < n) tmp[k++] = idx[j++];
  memcpy(idx, tmp, sizeof(int) * (size_t)n);
}

// ChooseSubTree's pick among n children by its two sorts, NaNs included.
int pick_by_sorts(const float* cost, const float* area, int n) {
  int idx[kRtMax], tmp[kRtMax];
  for (int i = 0; i < n; ++i) idx[i] = i;
  merge_order(idx, n, cost, tmp);
  int m = 1;
  while (m < n && !(cost[idx[m]] != cost[idx[m - 1]])) ++m;
  if (m > 1) merge_order(idx, m, area, tmp);
  return idx[0];
}

}  // namespace

RtTree::RtTree() { root_ = new_node(true); }

uint32_t RtTree::new_node(bool leaf) {
  nodes_.emplace_back();
  Node& nd = nodes_.back();
  memset(&nd, 0, sizeof(Node));
  nd.parent = -1;
  nd.leaf = leaf;
  return (uint32_t)(nodes_.size() - 1);
}

int RtTree::slot(uint32_t parent, uint32_t u) const {
  const Node& p = nodes_[parent];
  int i = 0;
  while (i < p.n && p.ref[i] != u) ++i;
  return i;
}

RtBox RtTree::node_box(uint32_t u) const {
  const Node& nd = nodes_[u];
  Box b = nd.box[0];
  for (int k = 1; k < nd.n; ++k)
    for (int d = 0; d < 3; ++d) {
      if (b.lo[d] > nd.box[k].lo[d]) b.lo[d] = nd.box[k].lo[d];
      if (b.hi[d] < nd.box[k].hi[d]) b.hi[d] = nd.box[k].hi[d];
    }
  return b;
}

void RtTree::add(uint32_t u, const Entry& e) {
  if (nodes_[u].n >= kRtMax) {
    bad_ = true;
    why_ = "R-tree node overflow";
    return;
  }
  nodes_[u].box[nodes_[u].n] = e.box;
  const int32_t p = nodes_[u].parent;
  if (p >= 0) {
    const int i = slot((uint32_t)p, u);
    if (i >= kRtMax) {
      bad_ = true;
      why_ = "R-tree node missing from its parent";
      return;
    }
    nodes_[p].box[i] = node_box(u);  // from the entries before this one
  }
  Node& nd = nodes_[u];
  nd.ref[nd.n] = e.ref;
  nd.n++;
  if (!nd.leaf) nodes_[e.ref].parent = (int32_t)u;
}

uint32_t RtTree::choose(const Box& e) const {
  uint32_t u = root_;
  for (int level = level_; level > 0; --level) {
    const Node& nd = nodes_[u];
    float cost[kRtMax], area[kRtMax];
    bool nan = false;
    for (int i = 0; i < nd.n; ++i) {
      area[i] = box_area(nd.box[i]);
      cost[i] = fabsf(box_area(box_union(nd.box[i], e)) - area[i]);
      nan |= cost[i] != cost[i] || area[i] != area[i];
    }
    int best = 0;
    if (!nan) {
      for (int i = 1; i < nd.n; ++i)
        if (cost[i] < cost[best] || (cost[i] == cost[best] && area[i] < area[best])) best = i;
    } else {
      best = pick_by_sorts(cost, area, nd.n);
    }
    u = nd.ref[best];
  }
  return u;
}

int RtTree::pick_next(const Entry* e, int remain, uint32_t a, uint32_t b, int* to) const {
  const Box ba = node_box(a), bb = node_box(b);
  const float aa = box_area(ba), ab = box_area(bb);
  float best = -FLT_MAX;
  int pick = -1;
  for (int i = 0; i < remain; ++i) {
    const float d1 = box_area(box_union(ba, e[i].box)) - aa;
    const float d2 = box_area(box_union(bb, e[i].box)) - ab;
    const float diff = fabsf(d1 - d2);
    if (diff > best) {
      best = diff;
      pick = i;
      *to = d1 < d2 ? 0 : d2 < d1 ? 1 : (nodes_[a].n < nodes_[b].n ? 0 : 1);
    }
  }
  return pick;
}

uint32_t RtTree::split(uint32_t u, const Entry& in) {
  Entry e[kRtMax + 1];
  for (int i = 0; i < kRtMax; ++i) e[i] = Entry{nodes_[u].box[i], nodes_[u].ref[i]};
  e[kRtMax] = in;
  nodes_[u].n = 0;
  const uint32_t v = new_node(nodes_[u].leaf);

  float ea[kRtMax + 1];
  for (int i = 0; i <= kRtMax; ++i) ea[i] = box_area(e[i].box);
  float best = -FLT_MAX;
  int s0 = 0, s1 = 1;
  for (int i = 0; i < kRtMax; ++i)
    for (int j = i + 1; j <= kRtMax; ++j) {
      const float d = box_area(box_union(e[i].box, e[j].box)) - ea[i] - ea[j];
      if (d > best) {
        best = d;
        s0 = i;
        s1 = j;
      }
    }
  add(u, e[s0]);
  add(v, e[s1]);
  int remain = kRtMax + 1;
  std::copy(e + s0 + 1, e + remain, e + s0);
  --remain;
  std::copy(e + s1, e + remain, e + s1 - 1);
  --remain;

  uint32_t rest = u;
  bool take_rest = false;
  for (; remain > 0; --remain) {
    if (nodes_[u].n == kRtMax - kRtMin + 1) {
      rest = v;
      take_rest = true;
      break;
    }
    if (nodes_[v].n == kRtMax - kRtMin + 1) {
      rest = u;
      take_rest = true;
      break;
    }
    int to = 0;
    const int k = pick_next(e, remain, u, v, &to);
    if (k < 0) {
      bad_ = true;
      why_ = "R-tree split undefined for these coordinates (NaN areas)";
      return v;
    }
    add(to == 0 ? u : v, e[k]);
    std::copy(e + k + 1, e + remain, e + k);
  }
  if (take_rest)
    for (int i = 0; i < remain; ++i) add(rest, e[i]);
  return v;
}

void RtTree::adjust(uint32_t u, int64_t* sib) {
  while (nodes_[u].parent >= 0 && !bad_) {
    const uint32_t p = (uint32_t)nodes_[u].parent;
    const int i = slot(p, u);
    if (i >= kRtMax) {
      bad_ = true;
      why_ = "R-tree node missing from its parent";
      return;
    }
    nodes_[p].box[i] = node_box(u);
    if (*sib >= 0) {
      const Entry s{node_box((uint32_t)*sib), (uint32_t)*sib};
      if (nodes_[p].n < kRtMax) {
        add(p, s);
        *sib = -1;
      } else {
        *sib = split(p, s);
      }
    }
    u = p;
  }
}

void RtTree::insert_one(const float* pt, uint32_t id) {
  Entry e;
  for (int d = 0; d < 3; ++d) e.box.lo[d] = e.box.hi[d] = pt[d];
  e.ref = id;
  const uint32_t t = choose(e.box);
  int64_t sib = -1;
  if (nodes_[t].n < kRtMax)
    add(t, e);
  else
    sib = split(t, e);
  if (!bad_) adjust(t, &sib);
  if (sib >= 0 && !bad_) {
    const uint32_t old = root_;
    const uint32_t r = new_node(false);
    add(r, Entry{node_box(old), old});
    add(r, Entry{node_box((uint32_t)sib), (uint32_t)sib});
    root_ = r;
    ++level_;
  }
}

bool RtTree::insert(const float* xyz, uint64_t n, std::string* err) {
  if (bad_) {
    *err = why_;
    return false;
  }
  if (n > kRtMaxPoints - npoints_) {
    *err = "an R-tree holds at most 2^29-1 points";
    return false;
  }
  for (uint64_t i = 0; i < n; ++i) {
    insert_one(xyz + 3 * i, (uint32_t)npoints_);
    if (bad_) {
      *err = why_;
      return false;
    }
    ++npoints_;
  }
  return true;
}

uint64_t RtTree::leaves() const {
  uint64_t k = 0;
  for (const Node& nd : nodes_) k += nd.leaf;
  return k;
}

void RtTree::leaf_ids(uint64_t* ids, uint32_t* len) const {
  std::vector<uint32_t> stack{root_};
  uint64_t ni = 0, nl = 0;
  while (!stack.empty()) {
    const Node& nd = nodes_[stack.back()];
    stack.pop_back();
    if (!nd.leaf) {
      for (int k = nd.n - 1; k >= 0; --k) stack.push_back(nd.ref[k]);
      continue;
    }
    for (int k = 0; k < nd.n; ++k) ids[ni++] = nd.ref[k];
    len[nl++] = (uint32_t)nd.n;
  }
}

void RtTree::flatten(RtFlat* f) const {
  std::vector<uint32_t> order;
  order.reserve(nodes_.size());
  order.push_back(root_);
  std::vector<uint32_t> num(nodes_.size());
  for (size_t i = 0; i < order.size(); ++i) {
    num[order[i]] = (uint32_t)i;
    const Node& nd = nodes_[order[i]];
    if (!nd.leaf)
      for (int k = 0; k < nd.n; ++k) order.push_back(nd.ref[k]);
  }
  uint32_t inner = 0;
  while (inner < order.size() && !nodes_[order[inner]].leaf) ++inner;
  const size_t m = order.size();
  f->inner = inner;
  f->root = (uint32_t)nodes_[root_].n << kRtRefBits;
  for (int d = 0; d < 3; ++d) {
    f->lo[d].assign(m * kRtLanes, 0.0f);
    f->hi[d].assign((size_t)std::max<uint32_t>(inner, 1) * kRtLanes, 0.0f);
  }
  f->child.assign((size_t)std::max<uint32_t>(inner, 1) * kRtLanes, 0);
  for (size_t i = 0; i < m; ++i) {
    const Node& nd = nodes_[order[i]];
    for (int k = 0; k < nd.n; ++k) {
      const size_t at = i * kRtLanes + k;
      for (int d = 0; d < 3; ++d) f->lo[d][at] = nd.box[k].lo[d];
      if (!nd.leaf) {
        for (int d = 0; d < 3; ++d) f->hi[d][at] = nd.box[k].hi[d];
        f->child[at] = num[nd.ref[k]] | (uint32_t)nodes_[nd.ref[k]].n << kRtRefBits;
      }
    }
  }
}

// ---- device count ----

namespace {

// first k with d < r2[k], or -1 (MultiRangeCount.c:107-113); r2 in LDS
template <bool kSorted>
__device__ __forceinline__ int rt_shell(float d, const float* r2, int nrads) {
  if constexpr (kSorted) {
    if (!(d < r2[nrads - 1])) return -1;
    int lo = 0, hi = nrads - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (d < r2[mid])
        hi = mid;
      else
        lo = mid + 1;
    }
    return lo;
  } else {
    for (int k = 0; k < nrads; ++k)
      if (d < r2[k]) return k;
    return -1;
  }
}

// Workgroups are dealt to the 8 XCDs in turn; this gives each XCD one
// contiguous range of the (Morton-ordered) queries, so the nodes that
// neighbouring queries share are fetched into one L2.
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, x = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
  const uint32_t q = nb / kXcds, r = nb % kXcds;
  return x * q + min(x, r) + i;
}

template <bool kSorted>
__global__ __launch_bounds__(kRtThreads) void rt_count_kernel(RtView t, uint32_t root,
                                                              uint32_t inner,
                                                              const float* __restrict__ hx,
                                                              const float* __restrict__ hy,
                                                              const float* __restrict__ hz,
                                                              uint64_t n,
                                                              const float* __restrict__ r2g,
                                                              int nrads, float rbox,
                                                              uint32_t* __restrict__ out) {
  extern __shared__ uint32_t rt_lds[];
  const int q = threadIdx.x / kRtLanes;           // query slot of the workgroup
  const int k = threadIdx.x % kRtLanes;           // entry lane
  const int base = threadIdx.x % 64 - k;          // the query's first lane in its wave
  uint32_t* stack = rt_lds + q * kRtStack;
  uint32_t* bins = rt_lds + kRtQueries * kRtStack + q * nrads;
  float* r2 = reinterpret_cast<float*>(rt_lds + kRtQueries * (kRtStack + nrads));
  for (int i = threadIdx.x; i < nrads; i += kRtThreads) r2[i] = r2g[i];
  for (int i = threadIdx.x; i < kRtQueries * nrads; i += kRtThreads)
    rt_lds[kRtQueries * kRtStack + i] = 0;
  __syncthreads();

  const uint64_t h = (uint64_t)xcd_block() * kRtQueries + q;
  const bool valid = h < n;
  float c[3] = {0.0f, 0.0f, 0.0f}, shift[3] = {0.0f, 0.0f, 0.0f};
  int axes = 0;  // bit d: axis d's query box leaves [0, 1000]
  if (valid) {
    c[0] = hx[h];
    c[1] = hy[h];
    c[2] = hz[h];
    for (int d = 0; d < 3; ++d) {
      if (c[d] - rbox < 0.0f) {
        axes |= 1 << d;
        shift[d] = 1000.0f;
      } else if (c[d] + rbox > 1000.0f) {
        axes |= 1 << d;
        shift[d] = -1000.0f;
      }
    }
  }
  // images: every subset of `axes`, enumerated by sub = (sub - axes) & axes
  int sub = 0;
  bool more = valid;
  float qc[3], qlo[3], qhi[3];
  int sp = 0;
  auto start = [&]() {
    for (int d = 0; d < 3; ++d) {
      qc[d] = (sub >> d & 1) ? c[d] + shift[d] : c[d];
      qlo[d] = qc[d] - rbox;
      qhi[d] = qc[d] + rbox;
    }
    if (k == 0) stack[0] = root;
    sp = 1;
  };
  if (more) start();

  while (true) {
    if (more && sp == 0) {
      sub = (sub - axes) & axes;
      if (sub == 0)
        more = false;
      else
        start();
    }
    wave_sync();  // the stack's last writes (a push, a new image's root) before its reads
    const bool busy = sp > 0;
    if (__ballot(busy) == 0) break;
    if (busy) {
      const uint32_t ref = stack[sp - 1];
      --sp;
      const uint32_t u = ref & kRtNodeMask;
      const bool live = k < (int)(ref >> kRtRefBits);
      const size_t at = (size_t)u * kRtLanes + k;
      if (u >= inner) {
        float d = 1.0f;
        if (live) d = dpps_distance(t.lox[at], t.loy[at], t.loz[at], qc[0], qc[1], qc[2]);
        const uint32_t zero = (uint32_t)(__ballot(live && d == 0.0f) >> base) & 0xffffu;
        const int first = zero ? __builtin_ctz(zero) : kRtLanes;
        if (live && k < first) {
          const int s = rt_shell<kSorted>(d, r2, nrads);
          if (s >= 0) atomicAdd(&bins[s], 1u);
        }
      } else {
        bool hit = false;
        if (live)
          hit = !(t.hix[at] < qlo[0]) && !(qhi[0] < t.lox[at]) && !(t.hiy[at] < qlo[1]) &&
                !(qhi[1] < t.loy[at]) && !(t.hiz[at] < qlo[2]) && !(qhi[2] < t.loz[at]);
        const uint32_t hits = (uint32_t)(__ballot(hit) >> base) & 0xffffu;
        if (hit) stack[sp + __builtin_popcount(hits & ((1u << k) - 1u))] = t.child[at];
        sp += __builtin_popcount(hits);
      }
    }
  }
  wave_sync();  // every bin increment of this wave before the bins are read
  if (valid)
    for (int s = k; s < nrads; s += kRtLanes) out[h * (uint64_t)nrads + s] = bins[s];
}

__global__ __launch_bounds__(256) void rt_add_counts_kernel(const uint32_t* __restrict__ c32,
                                                             const uint32_t* __restrict__ inv,
                                                             uint64_t* __restrict__ dst,
                                                             uint64_t total, uint32_t K) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * 256) {
    const uint64_t r = i / K, k = i - r * K;
    dst[i] += c32[(uint64_t)inv[r] * K + k];
  }
}

}  // namespace

hipError_t rt_add_counts(const uint32_t* c32, const uint32_t* inv, uint64_t* dst, uint64_t n,
                         int K, hipStream_t s) {
  const uint64_t total = n * (uint64_t)K;
  if (total == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(rt_add_counts_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, c32, inv, dst,
                     total, (uint32_t)K);
  return hipGetLastError();
}

hipError_t rt_count(const RtView& t, uint32_t root, uint32_t inner, const float* hx,
                    const float* hy, const float* hz, uint64_t n, const float* r2, int nrads,
                    bool sorted, float rbox, uint32_t* counts, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (nrads < 1 || nrads > kRtMaxShells) return hipErrorInvalidValue;
  const uint64_t blocks = (n + kRtQueries - 1) / kRtQueries;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t lds = sizeof(uint32_t) * ((size_t)kRtQueries * (kRtStack + nrads) + nrads);
  if (sorted)
    hipLaunchKernelGGL(rt_count_kernel<true>, dim3((uint32_t)blocks), dim3(kRtThreads), lds, s, t,
                       root, inner, hx, hy, hz, n, r2, nrads, rbox, counts);
  else
    hipLaunchKernelGGL(rt_count_kernel<false>, dim3((uint32_t)blocks), dim3(kRtThreads), lds, s,
                       t, root, inner, hx, hy, hz, n, r2, nrads, rbox, counts);
  return hipGetLastError();
}

}  // namespace sjcs
