/*
 * sjcs_oracle_rt.c -- CPU restatement of rt_rcount's index and count: the
 * Guttman R-tree with quadratic split, built by inserting the particles one
 * at a time in file order, and its periodic multi-shell count.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP R-tree count (and
 * the library's host-side R-tree build) are compared against; the product
 * never links it.
 *
 * rt_rcount counts differently from astr_rcount, and the difference depends
 * on how particles are grouped into leaves, so the grouping is replayed
 * step for step:
 *   insert       src/objects/RTree/RTree.c:35-49: an entry whose box is the
 *                point (low = upp = p, padding lane 0), id = file position.
 *   InsertRTree  src/objects/RTree/Build.c:243-269 (I1-I4).
 *   choose leaf  Build.c:271-320: per level, children sorted by
 *                |area(enlarged) - area| with glibc's qsort -- a stable top-
 *                down merge sort (msort) for arrays this small -- then the
 *                leading run of equal costs re-sorted by area; first wins.
 *   split        Build.c:147-189 (QS1-QS3): seeds by PickSeeds (:42-75,
 *                first pair of largest area(union) - area(a) - area(b)),
 *                then PickNext/DistributeEntry (:77-140) until one group has
 *                MAX_ENT - MIN_ENT + 1 = 7 entries; ties in enlargement go
 *                to the group with fewer entries, the second on equal sizes.
 *   adjust       Build.c:191-222 (AT2-AT4) and SplitRoot :224-241.
 *   node ops     src/objects/RTree/Node.c: addChildEntryToRTreeNode (:54-72)
 *                refreshes the parent's box of the node *before* counting
 *                the new entry (a stale box that AdjustTree later fixes),
 *                getMbrRTreeNode (:74-116, the scalar branch),
 *                findEntryIndexFromParentRTreeNode (:43-52).
 *   box ops      src/objects/RTree/Mbr.c (SSE branches, the ones it
 *                defines): area = ((1 * dx) * dy) * dz in FP32 (:50-70),
 *                enlarge by minps/maxps (:136-144), overlap by cmpnlt both
 *                ways (:167-182), query box c -+ r (:207-220), out-of-bounds
 *                flags (:284-313).
 *   count        src/objects/RTree/MultiRangeCount.c:13-74 (the periodic
 *                images, the same set as astr_rcount's) and :86-114: inner
 *                nodes descend into children whose box overlaps the query
 *                box; a leaf tests every entry's distance with no box test,
 *                and an entry at distance exactly 0 ends that leaf's scan
 *                (its own and later entries are not counted).
 * Distances, shells, images and the box radius are sjcs_oracle.c's.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "sjcs_oracle.h"

#define RT_MAX 10 /* include/objects/RTree/Node.h:7 */
#define RT_MIN 4  /* include/objects/RTree/Node.h:8 */

/* Mbr (include/objects/RTree/Mbr.h:13-16): three coordinates and a padding
 * lane that stays 0 in every box the tree holds. */
typedef struct {
    float lo[4], hi[4];
} rt_box;

typedef struct {
    int leaf, nch;
    int64_t parent; /* -1: root */
    rt_box box[RT_MAX];
    int64_t child[RT_MAX]; /* node index, or particle id in a leaf */
} rt_node;

typedef struct {
    rt_box box;
    int64_t child;
} rt_entry;

struct oracle_rtree {
    rt_node *node;
    int64_t nnodes, cap, root;
    int level;
    uint64_t npoints;
    int failed;
};

static int64_t rt_new_node(oracle_rtree *t, int leaf)
{
    if (t->nnodes == t->cap) {
        int64_t cap = t->cap ? 2 * t->cap : 64;
        rt_node *p = realloc(t->node, sizeof(rt_node) * (size_t)cap);
        if (!p) {
            t->failed = 1;
            return -1;
        }
        t->node = p;
        t->cap = cap;
    }
    memset(&t->node[t->nnodes], 0, sizeof(rt_node));
    t->node[t->nnodes].leaf = leaf;
    t->node[t->nnodes].parent = -1;
    return t->nnodes++;
}

static float rt_area(const rt_box *b)
{
    float a = 1.0f;
    a *= b->hi[0] - b->lo[0];
    a *= b->hi[1] - b->lo[1];
    a *= b->hi[2] - b->lo[2];
    return a;
}

/* minps / maxps: the first operand when it compares true, else the second */
static void rt_enlarge(rt_box *b, const rt_box *e)
{
    int d;
    for (d = 0; d < 4; d++) {
        b->lo[d] = b->lo[d] < e->lo[d] ? b->lo[d] : e->lo[d];
        b->hi[d] = b->hi[d] > e->hi[d] ? b->hi[d] : e->hi[d];
    }
}

static int rt_overlap(const rt_box *a, const rt_box *q)
{
    int d;
    for (d = 0; d < 4; d++)
        if (a->hi[d] < q->lo[d] || q->hi[d] < a->lo[d])
            return 0;
    return 1;
}

static rt_box rt_node_box(const rt_node *nd)
{
    rt_box b = nd->box[0];
    int k, d;
    for (k = 1; k < nd->nch; k++)
        for (d = 0; d < 3; d++) {
            if (b.lo[d] > nd->box[k].lo[d])
                b.lo[d] = nd->box[k].lo[d];
            if (b.hi[d] < nd->box[k].hi[d])
                b.hi[d] = nd->box[k].hi[d];
        }
    return b;
}

static int rt_slot(const oracle_rtree *t, int64_t parent, int64_t node)
{
    const rt_node *p = &t->node[parent];
    int i;
    for (i = 0; i < p->nch; i++)
        if (p->child[i] == node)
            break;
    return i;
}

static void rt_add(oracle_rtree *t, int64_t node, const rt_entry *e)
{
    rt_node *nd = &t->node[node];
    if (nd->nch >= RT_MAX) {
        t->failed = 1;
        return;
    }
    nd->box[nd->nch] = e->box;
    if (nd->parent >= 0) {
        int i = rt_slot(t, nd->parent, node);
        if (i >= RT_MAX) {
            t->failed = 1;
            return;
        }
        t->node[nd->parent].box[i] = rt_node_box(nd);
    }
    nd->child[nd->nch] = e->child;
    nd->nch++;
    if (!nd->leaf)
        t->node[e->child].parent = node;
}

typedef struct {
    int64_t node;
    float cost;
} rt_cost;

/* glibc msort's order: halves sorted recursively, the merge takes the left
 * element unless it compares greater than the right one. */
static void rt_msort(rt_cost *a, int n, rt_cost *tmp)
{
    int n1 = n / 2, i = 0, j = n1, k = 0;
    if (n <= 1)
        return;
    rt_msort(a, n1, tmp);
    rt_msort(a + n1, n - n1, tmp);
    while (i < n1 && j < n)
        tmp[k++] = (a[i].cost > a[j].cost) ? a[j++] : a[i++];
    while (i < n1)
        tmp[k++] = a[i++];
    while (j < n)
        tmp[k++] = a[j++];
    memcpy(a, tmp, sizeof(rt_cost) * (size_t)n);
}

static int64_t rt_choose(oracle_rtree *t, const rt_box *e)
{
    int64_t node = t->root;
    int level;
    for (level = t->level; level > 0; level--) {
        const rt_node *nd = &t->node[node];
        rt_cost c[RT_MAX], tmp[RT_MAX];
        int i, n;
        for (i = 0; i < nd->nch; i++) {
            rt_box x = nd->box[i];
            rt_enlarge(&x, e);
            c[i].node = nd->child[i];
            c[i].cost = fabsf(rt_area(&x) - rt_area(&nd->box[i]));
        }
        rt_msort(c, nd->nch, tmp);
        for (n = 1; n < nd->nch; n++)
            if (c[n].cost != c[n - 1].cost)
                break;
        if (n > 1) {
            for (i = 0; i < n; i++)
                c[i].cost = rt_area(&nd->box[rt_slot(t, node, c[i].node)]);
            rt_msort(c, n, tmp);
        }
        node = c[0].node;
    }
    return node;
}

static void rt_remove(rt_entry *e, int at, int count)
{
    memmove(&e[at], &e[at + 1], sizeof(rt_entry) * (size_t)(count - at - 1));
}

/* Next entry to place and where: >= 0 (its index) with *to = 1 for the
 * first group, 2 for the second; -1 when no difference compared (NaN). */
static int rt_pick_next(oracle_rtree *t, const rt_entry *e, int remain, int64_t g1, int64_t g2,
                        int *to)
{
    rt_box b1 = rt_node_box(&t->node[g1]), b2 = rt_node_box(&t->node[g2]);
    float best = -FLT_MAX;
    int i, pick = -1;
    for (i = 0; i < remain; i++) {
        rt_box x1 = b1, x2 = b2;
        float d1, d2, diff;
        rt_enlarge(&x1, &e[i].box);
        d1 = rt_area(&x1) - rt_area(&b1);
        rt_enlarge(&x2, &e[i].box);
        d2 = rt_area(&x2) - rt_area(&b2);
        diff = fabsf(d1 - d2);
        if (diff > best) {
            best = diff;
            pick = i;
            if (d1 < d2)
                *to = 1;
            else if (d2 < d1)
                *to = 2;
            else
                *to = t->node[g1].nch < t->node[g2].nch ? 1 : 2;
        }
    }
    return pick;
}

/* Splits `node` (full) plus entry `in`; the new sibling is *split. */
static void rt_split(oracle_rtree *t, int64_t node, const rt_entry *in, int64_t *split)
{
    rt_entry e[RT_MAX + 1];
    int remain = RT_MAX + 1, i, j, s0 = 0, s1 = 1;
    float best = -FLT_MAX;
    int64_t rest = -1;

    for (i = 0; i < RT_MAX; i++) {
        e[i].box = t->node[node].box[i];
        e[i].child = t->node[node].child[i];
    }
    e[RT_MAX] = *in;
    t->node[node].nch = 0;
    *split = rt_new_node(t, t->node[node].leaf);
    if (*split < 0)
        return;
    /* seeds: the pair that wastes the most area together */
    for (i = 0; i < RT_MAX; i++)
        for (j = i + 1; j < RT_MAX + 1; j++) {
            rt_box u = e[i].box;
            float d;
            rt_enlarge(&u, &e[j].box);
            d = rt_area(&u) - rt_area(&e[i].box) - rt_area(&e[j].box);
            if (d > best) {
                best = d;
                s0 = i;
                s1 = j;
            }
        }
    rt_add(t, node, &e[s0]);
    rt_add(t, *split, &e[s1]);
    rt_remove(e, s0, RT_MAX + 1);
    rt_remove(e, s1 - 1, RT_MAX);
    remain -= 2;

    for (; remain > 0; remain--) {
        int to = 0, k;
        if (t->node[node].nch == RT_MAX - RT_MIN + 1) {
            rest = *split;
            break;
        }
        if (t->node[*split].nch == RT_MAX - RT_MIN + 1) {
            rest = node;
            break;
        }
        k = rt_pick_next(t, e, remain, node, *split, &to);
        if (k < 0) {
            t->failed = 1;
            return;
        }
        rt_add(t, to == 1 ? node : *split, &e[k]);
        rt_remove(e, k, remain);
    }
    for (i = 0; i < remain; i++)
        rt_add(t, rest, &e[i]);
}

static void rt_adjust(oracle_rtree *t, int64_t node, int64_t *split)
{
    while (t->node[node].parent >= 0 && !t->failed) {
        int64_t parent = t->node[node].parent;
        int i = rt_slot(t, parent, node);
        t->node[parent].box[i] = rt_node_box(&t->node[node]);
        if (*split >= 0) {
            rt_entry s;
            s.box = rt_node_box(&t->node[*split]);
            s.child = *split;
            if (t->node[parent].nch < RT_MAX) {
                rt_add(t, parent, &s);
                *split = -1;
            } else {
                rt_split(t, parent, &s, split);
            }
        }
        node = parent;
    }
}

static void rt_insert(oracle_rtree *t, const rt_entry *e)
{
    int64_t target = rt_choose(t, &e->box), split = -1;
    if (t->node[target].nch < RT_MAX)
        rt_add(t, target, e);
    else
        rt_split(t, target, e, &split);
    rt_adjust(t, target, &split);
    if (split >= 0 && !t->failed) {
        int64_t old = t->root, root = rt_new_node(t, 0);
        rt_entry a, b;
        if (root < 0)
            return;
        a.box = rt_node_box(&t->node[old]);
        a.child = old;
        b.box = rt_node_box(&t->node[split]);
        b.child = split;
        rt_add(t, root, &a);
        rt_add(t, root, &b);
        t->root = root;
        t->level++;
    }
}

oracle_rtree *oracle_rt_build(const float *xyz, uint64_t n)
{
    oracle_rtree *t = calloc(1, sizeof(*t));
    uint64_t i;
    if (!t)
        return NULL;
    t->root = rt_new_node(t, 1);
    for (i = 0; i < n && !t->failed; i++) {
        rt_entry e;
        memset(&e, 0, sizeof(e));
        memcpy(e.box.lo, xyz + 3 * i, 3 * sizeof(float));
        memcpy(e.box.hi, xyz + 3 * i, 3 * sizeof(float));
        e.child = (int64_t)i;
        rt_insert(t, &e);
    }
    t->npoints = n;
    if (t->failed) {
        oracle_rt_free(t);
        return NULL;
    }
    return t;
}

void oracle_rt_free(oracle_rtree *t)
{
    if (!t)
        return;
    free(t->node);
    free(t);
}

int oracle_rt_height(const oracle_rtree *t)
{
    return t->level;
}

static void rt_leaves(const oracle_rtree *t, int64_t node, uint64_t *ids, uint32_t *len,
                      uint64_t *nid, uint64_t *nleaf)
{
    const rt_node *nd = &t->node[node];
    int k;
    if (!nd->leaf) {
        for (k = 0; k < nd->nch; k++)
            rt_leaves(t, nd->child[k], ids, len, nid, nleaf);
        return;
    }
    for (k = 0; k < nd->nch; k++) {
        if (ids)
            ids[*nid] = (uint64_t)nd->child[k];
        (*nid)++;
    }
    if (len)
        len[*nleaf] = (uint32_t)nd->nch;
    (*nleaf)++;
}

uint64_t oracle_rt_leaves(const oracle_rtree *t, uint64_t *ids, uint32_t *leaf_len)
{
    uint64_t nid = 0, nleaf = 0;
    rt_leaves(t, t->root, ids, leaf_len, &nid, &nleaf);
    return nleaf;
}

static void rt_query(const oracle_rtree *t, int64_t node, const rt_box *q, const float c[3],
                     const float *r2, int nrads, uint64_t *row)
{
    const rt_node *nd = &t->node[node];
    int k;
    if (!nd->leaf) {
        for (k = 0; k < nd->nch; k++)
            if (rt_overlap(&nd->box[k], q))
                rt_query(t, nd->child[k], q, c, r2, nrads, row);
        return;
    }
    for (k = 0; k < nd->nch; k++) {
        float dist = oracle_distance(nd->box[k].lo, c);
        int s;
        if (dist == 0)
            return;
        s = oracle_shell_of(dist, r2, nrads);
        if (s >= 0)
            row[s]++;
    }
}

void oracle_rt_count(const oracle_rtree *t, const void *halo_points, size_t stride_bytes,
                     uint64_t nh, const float *r2, int nrads, uint64_t *counts)
{
    const float rbox = oracle_box_radius(r2, nrads);
    uint64_t h;
    for (h = 0; h < nh; h++) {
        const float *c = (const float *)((const char *)halo_points + h * stride_bytes);
        float cen[8][3];
        int m = oracle_images(c, rbox, cen), i, d;
        for (i = 0; i < m; i++) {
            rt_box q;
            memset(&q, 0, sizeof(q));
            for (d = 0; d < 3; d++) {
                q.lo[d] = cen[i][d] - rbox;
                q.hi[d] = cen[i][d] + rbox;
            }
            rt_query(t, t->root, &q, cen[i], r2, nrads, counts + h * (uint64_t)nrads);
        }
    }
}
