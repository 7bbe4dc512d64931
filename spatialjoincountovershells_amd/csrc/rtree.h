// rtree.h -- rt_rcount's index: the Guttman R-tree with quadratic split,
// built on the host by inserting the points one at a time in order (the
// order fixes which points share a leaf, and rt_rcount's count depends on
// that), then flattened into node arrays that the count kernel walks on the
// device, one query per 16 lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace sjcs {

constexpr int kRtMax = 10;      // entries per node (include/objects/RTree/Node.h:7)
constexpr int kRtMin = 4;       // least entries of a split half (Node.h:8)
constexpr int kRtLanes = 16;    // lanes per query: one node's entries side by side
constexpr int kRtStack = 192;   // nodes pending per query, in LDS
constexpr int kRtThreads = 256;
constexpr int kRtQueries = kRtThreads / kRtLanes;  // queries per workgroup
constexpr int kRtRefBits = 28;  // a node reference: node number | entries << 28
constexpr uint32_t kRtNodeMask = (1u << kRtRefBits) - 1;
// a halo's count over its (at most 8) images stays below 2^32 in LDS
constexpr uint64_t kRtMaxPoints = (1ull << 29) - 1;
constexpr int kRtMaxShells = 1024;
constexpr uint32_t kXcds = 8;  // MI355X: XCDs that workgroups are dealt to in turn

// The tree on the device.  Node u's entry k is at u * kRtLanes + k: the low
// corner of its box (a leaf entry's box is its point), and for an inner node
// the high corner and a reference to the child: its node number | its entry
// count << kRtRefBits, so a walk knows how many entries to read without
// another load.  Inner nodes are numbered first: u >= inner is a leaf.
struct RtView {
  const float* lox;
  const float* loy;
  const float* loz;
  const float* hix;  // inner nodes only
  const float* hiy;
  const float* hiz;
  const uint32_t* child;
};

// An entry's box: low and high corner (include/objects/RTree/Mbr.h:13-16
// without its padding lane, which stays 0 in every box the reference's tree
// holds and so never changes an area or an overlap test).
struct RtBox {
  float lo[3], hi[3];
};

// Host copy of that layout, breadth first from the root (every leaf is at
// depth `height`, so the inner nodes come first).
struct RtFlat {
  std::vector<float> lo[3], hi[3];
  std::vector<uint32_t> child;
  uint32_t inner = 0;
  uint32_t root = 0;  // the root's reference (node 0)
};

class RtTree {
 public:
  RtTree();
  // Inserts n points (packed xyz) with ids continuing from those already
  // held.  False with *err set when the reference's split is undefined for
  // the coordinates (every enlargement difference NaN), or past the limits.
  bool insert(const float* xyz, uint64_t n, std::string* err);
  // false after a failed insert: the tree may then hold part of a split
  bool ok() const { return !bad_; }
  uint64_t points() const { return npoints_; }
  int height() const { return level_; }
  uint64_t nodes() const { return nodes_.size(); }
  uint64_t leaves() const;
  // ids[points()] grouped by leaf, leaves depth first, entries in node
  // order; len[leaves()] the leaf sizes.
  void leaf_ids(uint64_t* ids, uint32_t* len) const;
  void flatten(RtFlat* f) const;

 private:
  using Box = RtBox;
  struct Entry {
    Box box;
    uint32_t ref;  // node number, or point id in a leaf
  };
  struct Node {
    Box box[kRtMax];
    uint32_t ref[kRtMax];
    int32_t parent;  // -1: the root
    int32_t n;
    bool leaf;
  };

  uint32_t new_node(bool leaf);
  int slot(uint32_t parent, uint32_t u) const;
  Box node_box(uint32_t u) const;
  void add(uint32_t u, const Entry& e);
  uint32_t choose(const Box& e) const;
  uint32_t split(uint32_t u, const Entry& in);
  int pick_next(const Entry* e, int remain, uint32_t a, uint32_t b, int* to) const;
  void adjust(uint32_t u, int64_t* split);
  void insert_one(const float* p, uint32_t id);

  std::vector<Node> nodes_;
  uint32_t root_ = 0;
  int level_ = 0;
  uint64_t npoints_ = 0;
  bool bad_ = false;
  std::string why_;
};

// counts[h * nrads + k] = rt_rcount's shell-k count of centre h (stored, not
// added), for n centres (SoA x, y, z on the device; in Morton order for
// cache reuse) against the tree `t` whose root has reference `root` and
// whose first `inner` nodes are inner nodes.  r2: nrads squared
// thresholds on the device; sorted: they are non-decreasing (a binary search
// finds the first match); rbox: the query box half-edge,
// (float)sqrt(r2[nrads-1]).
hipError_t rt_count(const RtView& t, uint32_t root, uint32_t inner, const float* hx,
                    const float* hy,
                    const float* hz, uint64_t n, const float* r2, int nrads, bool sorted,
                    float rbox, uint32_t* counts, hipStream_t s);

// dst[r * K + k] += c32[inv[r] * K + k] for r < n: counts from the sorted
// query order back to input rows, added.
hipError_t rt_add_counts(const uint32_t* c32, const uint32_t* inv, uint64_t* dst, uint64_t n,
                         int K, hipStream_t s);

}  // namespace sjcs
