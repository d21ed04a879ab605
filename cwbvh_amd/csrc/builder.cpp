// SAH BVH builder and GPU layout converters (host).
//
// Replaces the reference's Embree-based builder (src/EmbreeBVHBuilder.cpp:303-478,
// which needs Embree3 + TBB) and its converters (src/GPUBVHConverter.cpp:10-403).
//
// One binned-SAH binary tree with single-reference leaves is built per mesh,
// on a work queue of subtrees; with spatial splits (split_alpha > 0) a
// reference can be the part of a triangle on one side of split planes, the
// role of the reference's SplitPrimitiveTriangle (EmbreeBVHBuilder.cpp:110-247).
// Two products are derived from it, each written on all threads in the byte
// order of the serial depth-first pass that defines it:
//   * BVH2: leaves merged bottom-up wherever the SAH leaf cost is lower
//     (bounded by bvh2_max_leaf), written in the reference BVH2 layout
//     (GPUBVHConverter.cpp:10-121): 4 float4 per node, child ints >= 0 are
//     float4 node addresses, < 0 are ~(Woop float4 index); leaves are Woop
//     triangles closed by a 0x80000000 terminator float4.
//   * CWBVH: the binary tree collapsed to 8-wide nodes with <= 3 triangles per
//     leaf by the dynamic program of Ylitie, Karras, Laine (HPG 2017, Sec. 3.1),
//     written in the reference compressed layout (GPUBVHConverter.cpp:123-378):
//       n0 = (p.xyz, e.x e.y e.z imask bytes), n1 = (child base node index,
//       triangle base index, meta[0..3], meta[4..7]), n2..n4 = 8-bit child
//       bounds qlo.x[8] qlo.y[8] qlo.z[8] qhi.x[8] qhi.y[8] qhi.z[8].
//     Inner-child meta = 0b001_11sss (slot s), leaf meta = unary(count)<<5 |
//     triangle offset.  Children are assigned to octant slots greedily by
//     centroid direction.  Optionally each node is padded to 8 float4 (128 B).
//
// Deviations from the reference converter, all keeping the layout:
//   * per-axis exponent is the smallest e with extent <= 255*2^e, clamped to
//     [-126, 127] (reference: ceilf(log2f(extent/255)) gives -inf -> UB for a
//     flat axis and can round to a qhi of 256, stored as 0);
//   * quantization runs in double so floor/ceil are conservative;
//   * a root that is a leaf, and an empty mesh, get a valid inner root;
//   * no fixed-size DFS stacks;
//   * Woop row0.x of -0.0 is stored as +0.0 so it cannot read as a terminator.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <thread>

#include "common.h"

using namespace cw;

namespace {

struct BNode {
    Box box;
    int32_t left = -1, right = -1;  // internal: children
    int32_t first = 0, count = 0;   // leaf: range in refs
    int32_t prims = 0;              // triangles in subtree
};

struct Tree {
    std::vector<BNode> nodes;
    std::vector<int32_t> refs;
    int32_t root = 0;
    bool split_refs = false;  // built with spatial splits: a leaf can name a triangle twice
};

// A triangle reference with its box and centroid, so the recursive passes
// scan and partition one contiguous array instead of gathering from two.
// With spatial splits a triangle can have several references, each with the
// box of the part of the triangle inside its subtree.
struct PrimRef {
    Box box;
    V3 cen;
    int32_t id;
};

bool box_empty(const Box& b) { return !(b.lo.x <= b.hi.x && b.lo.y <= b.hi.y && b.lo.z <= b.hi.z); }

Box box_and(const Box& a, const Box& b) {
    Box r;
    r.lo = V3{std::max(a.lo.x, b.lo.x), std::max(a.lo.y, b.lo.y), std::max(a.lo.z, b.lo.z)};
    r.hi = V3{std::min(a.hi.x, b.hi.x), std::min(a.hi.y, b.hi.y), std::min(a.hi.z, b.hi.z)};
    return r;
}

void set_comp(V3& v, int a, float x) { (a == 0 ? v.x : (a == 1 ? v.y : v.z)) = x; }

// Neighbouring floats (finite inputs; +-0 step to the smallest denormal).
inline float float_up(float f) {
    const uint32_t u = f2u(f);
    if ((u & 0x7FFFFFFFu) == 0) return u2f(1u);
    if (f == INFINITY) return f;
    return u2f((u & 0x80000000u) ? u - 1 : u + 1);
}
inline float float_down(float f) { return -float_up(-f); }

// Grows `b` by a point computed in double, rounded outwards to float and
// widened by one more ulp, so the box holds the exact point.  The split axis
// gets `axis_value` exactly.
void grow_outward(Box& b, const double p[3], int axis, float axis_value) {
    V3 lo{0.0f, 0.0f, 0.0f}, hi{0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 3; ++k) {
        const float f = (float)p[k];
        set_comp(lo, k, float_down((double)f > p[k] ? float_down(f) : f));
        set_comp(hi, k, float_up((double)f < p[k] ? float_up(f) : f));
    }
    set_comp(lo, axis, axis_value);
    set_comp(hi, axis, axis_value);
    b.grow(lo);
    b.grow(hi);
}

// Boxes of the parts of triangle (v0, v1, v2) on either side of the plane
// x[axis] = pos (Stich et al. 2009, "Spatial splits in bounding volume
// hierarchies"): the vertices on each side plus the points where edges
// cross the plane.  Both boxes are conservative.
void split_triangle(const V3 v[3], int axis, float pos, Box& L, Box& R) {
    L = Box();
    R = Box();
    for (int i = 0; i < 3; ++i) {
        const V3 p = v[i], q = v[(i + 1) % 3];
        const float pa = comp(p, axis), qa = comp(q, axis);
        if (pa <= pos) L.grow(p);
        if (pa >= pos) R.grow(p);
        if ((pa < pos && qa > pos) || (pa > pos && qa < pos)) {
            const double t = ((double)pos - pa) / ((double)qa - pa);
            const double x[3] = {p.x + ((double)q.x - p.x) * t, p.y + ((double)q.y - p.y) * t,
                                 p.z + ((double)q.z - p.z) * t};
            grow_outward(L, x, axis, pos);
            grow_outward(R, x, axis, pos);
        }
    }
}

// The same split for estimates only (spatial binning): float arithmetic, no
// outward rounding.
void split_triangle_estimate(const V3 v[3], int axis, float pos, Box& L, Box& R) {
    L = Box();
    R = Box();
    for (int i = 0; i < 3; ++i) {
        const V3 p = v[i], q = v[(i + 1) % 3];
        const float pa = comp(p, axis), qa = comp(q, axis);
        if (pa <= pos) L.grow(p);
        if (pa >= pos) R.grow(p);
        if ((pa < pos && qa > pos) || (pa > pos && qa < pos)) {
            V3 x = p + (q - p) * ((pos - pa) / (qa - pa));
            set_comp(x, axis, pos);
            L.grow(x);
            R.grow(x);
        }
    }
}

struct Builder {
    const cwbvh_mesh* mesh;
    std::vector<BNode>& nodes;
    // Node ranges [b, e) own refs[b, c): the tail [e, c) is room for the
    // references spatial splits add below the node.
    std::vector<PrimRef>& refs;
    std::atomic<int32_t> next_node{0};
    int max_threads;
    int bins;
    int max_depth;
    // spatial splits are searched where the children of the best object
    // split overlap by more than this fraction of the root's area (<= 0: never)
    float split_alpha;
    double root_area = 0.0;

    // Subtrees at least this large are handed to the worker pool; smaller
    // ones are finished by the thread that split them.
    static constexpr int32_t kTaskMin = 4096;
    // nodes this large bin their references on all threads
    static constexpr int32_t kParallelBinMin = 1 << 18;
    struct Task {
        int32_t idx, b, e, c, depth;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Task> tasks;
    int64_t outstanding = 0;

    Builder(const cwbvh_mesh* m, std::vector<BNode>& n, std::vector<PrimRef>& r, int threads, int nbins, int maxd,
            float alpha)
        : mesh(m), nodes(n), refs(r), max_threads(threads), bins(nbins), max_depth(maxd), split_alpha(alpha) {}

    int32_t alloc2() { return next_node.fetch_add(2); }

    struct Bin {
        Box box;
        int32_t n = 0;
    };
    struct SBin {
        Box box;
        int32_t enter = 0, exit = 0;
    };

    // Centroid bins of one axis over [lo, hi]: nb bins of equal width, in
    // double so that tiny or huge extents neither overflow nor produce NaN.
    // Binning and partitioning use the same mapping.
    struct AxisBins {
        double lo = 0.0, scale = 0.0;
        int nb = 0;
        bool ok = false;
        AxisBins(float l, float h, int n) : nb(n) {
            const double ext = (double)h - (double)l;
            ok = ext > 0.0 && std::isfinite(ext);
            lo = l;
            scale = ok ? (double)n / ext : 0.0;
        }
        int bin(float x) const {
            const double f = ((double)x - lo) * scale;
            if (!(f > 0.0)) return 0;
            return f >= nb - 1 ? nb - 1 : (int)f;
        }
    };

    // Planes of the spatial bins of one axis: plane[k], k = 0..nb, with
    // plane[0] and plane[nb] the node's bounds.  A reference with extent
    // [lo, hi] on the axis enters at the last bin whose plane is <= lo and
    // leaves at the last bin whose plane is < hi (not before it enters), so
    // that the plane k splits it iff lo < plane[k] < hi.
    struct SpatialAxis {
        float plane[257];
        double lo = 0.0, scale = 0.0;
        int nb = 0;
        bool ok = false;
        SpatialAxis(float l, float h, int n) : nb(n) {
            const double ext = (double)h - (double)l;
            ok = ext > 0.0 && std::isfinite(ext);
            lo = l;
            scale = ok ? (double)n / ext : 0.0;
            for (int k = 0; k <= n; ++k) plane[k] = (float)((double)l + ext * k / n);
            plane[n] = h;
        }
        int guess(float x) const {
            const double f = ((double)x - lo) * scale;
            if (!(f > 0.0)) return 0;
            return f >= nb - 1 ? nb - 1 : (int)f;
        }
        int enter_bin(float x) const {
            int k = guess(x);
            while (k > 0 && plane[k] > x) --k;
            while (k < nb - 1 && plane[k + 1] <= x) ++k;
            return k;
        }
        int exit_bin(float x, int enter) const {
            int k = guess(x);
            while (k > 0 && !(plane[k] < x)) --k;
            while (k < nb - 1 && plane[k + 1] < x) ++k;
            return std::max(k, enter);
        }
    };

    void triangle(int32_t id, V3 v[3]) const { mesh->tri(id, v[0], v[1], v[2]); }

    // Runs bin_range(lo, hi, out) over [b, e) into `bins_out`, on all threads
    // for large ranges (partial bins merged in thread order).
    // Helper threads of the parallel binning, over all tasks at once: a task
    // takes them only while the total stays within max_threads, so nested
    // parallelism never multiplies the thread count.
    std::atomic<int> bin_helpers{0};

    template <class B, class F, class M>
    void bin_all(int32_t b, int32_t e, size_t count, std::vector<B>& bins_out, F&& bin_range, M&& merge) {
        const int32_t n = e - b;
        int t = 1;
        if (n >= kParallelBinMin && max_threads > 1) {
            t = std::min(max_threads, 8);
            if (bin_helpers.fetch_add(t) + t > max_threads) {
                bin_helpers.fetch_sub(t);
                t = 1;
            }
        }
        if (t > 1) {
            std::vector<std::vector<B>> part(t, std::vector<B>(count));
            std::vector<std::thread> pool;
            int32_t chunk = (n + t - 1) / t;
            for (int i = 0; i < t; ++i) {
                int32_t lo = b + i * chunk, hi = std::min(e, lo + chunk);
                if (lo >= hi) break;
                pool.emplace_back([&, lo, hi, i] { bin_range(lo, hi, part[i]); });
            }
            for (auto& th : pool) th.join();
            bin_helpers.fetch_sub(t);
            for (auto& pt : part)
                for (size_t k = 0; k < count; ++k) merge(bins_out[k], pt[k]);
        } else {
            bin_range(b, e, bins_out);
        }
    }

    struct Choice {
        int kind = 0;  // 0: none, 1: object split, 2: spatial split
        int axis = -1, bin = -1;
        float pos = 0.0f;  // spatial split: the plane
        double cost = INFINITY;
        Box left, right;   // object split: the children's boxes
    };

    // Binned SAH over the centroids of [b, e).
    Choice object_split(int32_t b, int32_t e, int nb, const AxisBins ax[3]) {
        Choice best;
        const int32_t n = e - b;
        thread_local std::vector<Bin> bin;
        bin.assign(3 * (size_t)nb, Bin());
        bin_all(
            b, e, 3 * (size_t)nb, bin,
            [&](int32_t lo, int32_t hi, std::vector<Bin>& out) {
                for (int32_t i = lo; i < hi; ++i) {
                    const PrimRef& r = refs[i];
                    for (int a = 0; a < 3; ++a) {
                        if (!ax[a].ok) continue;
                        Bin& bb = out[a * nb + ax[a].bin(comp(r.cen, a))];
                        bb.box.grow(r.box);
                        bb.n++;
                    }
                }
            },
            [](Bin& into, const Bin& from) {
                into.box.grow(from.box);
                into.n += from.n;
            });
        double right_cost[257];
        Box right_box[257];
        for (int a = 0; a < 3; ++a) {
            if (!ax[a].ok) continue;
            Box acc;
            int32_t cnt = 0;
            for (int k = nb - 1; k >= 1; --k) {
                acc.grow(bin[a * nb + k].box);
                cnt += bin[a * nb + k].n;
                right_cost[k] = cnt ? acc.area() * cnt : 0.0;
                right_box[k] = acc;
            }
            acc = Box();
            cnt = 0;
            for (int k = 0; k < nb - 1; ++k) {
                acc.grow(bin[a * nb + k].box);
                cnt += bin[a * nb + k].n;
                if (cnt == 0 || cnt == n) continue;
                const double c = acc.area() * cnt + right_cost[k + 1];
                if (c < best.cost) {
                    best.cost = c;
                    best.kind = 1;
                    best.axis = a;
                    best.bin = k;
                    best.left = acc;
                    best.right = right_box[k + 1];
                }
            }
        }
        return best;
    }

    // Binned spatial-split SAH over the node box `nb_box`: references are
    // chopped at the bin planes they straddle, each bin grows by the box of
    // the triangle's part inside it.  Only planes whose split fits the free
    // room [e, c) are candidates.
    Choice spatial_split(int32_t b, int32_t e, int32_t c, const Box& nb_box, int nb) {
        Choice best;
        const SpatialAxis S[3] = {SpatialAxis(nb_box.lo.x, nb_box.hi.x, nb), SpatialAxis(nb_box.lo.y, nb_box.hi.y, nb),
                                  SpatialAxis(nb_box.lo.z, nb_box.hi.z, nb)};
        if (!S[0].ok && !S[1].ok && !S[2].ok) return best;
        thread_local std::vector<SBin> sbin;
        sbin.assign(3 * (size_t)nb, SBin());
        bin_all(
            b, e, 3 * (size_t)nb, sbin,
            [&](int32_t lo, int32_t hi, std::vector<SBin>& out) {
                V3 v[3];
                for (int32_t i = lo; i < hi; ++i) {
                    const PrimRef& r = refs[i];
                    bool fetched = false;
                    for (int a = 0; a < 3; ++a) {
                        if (!S[a].ok) continue;
                        SBin* ob = &out[(size_t)a * nb];
                        const int k0 = S[a].enter_bin(comp(r.box.lo, a));
                        const int k1 = S[a].exit_bin(comp(r.box.hi, a), k0);
                        ob[k0].enter++;
                        ob[k1].exit++;
                        if (k0 == k1) {
                            ob[k0].box.grow(r.box);
                            continue;
                        }
                        if (!fetched) {
                            triangle(r.id, v);
                            fetched = true;
                        }
                        Box rest = r.box;
                        for (int k = k0; k < k1 && !box_empty(rest); ++k) {
                            Box L, R;
                            split_triangle_estimate(v, a, S[a].plane[k + 1], L, R);
                            const Box part = box_and(L, rest);
                            if (!box_empty(part)) ob[k].box.grow(part);
                            rest = box_and(R, rest);
                        }
                        if (!box_empty(rest)) ob[k1].box.grow(rest);
                    }
                }
            },
            [](SBin& into, const SBin& from) {
                into.box.grow(from.box);
                into.enter += from.enter;
                into.exit += from.exit;
            });
        for (int a = 0; a < 3; ++a) {
            if (!S[a].ok) continue;
            const SBin* sb = &sbin[(size_t)a * nb];
            double right_cost[257];
            int32_t right_n[257];
            Box acc;
            int32_t cnt = 0;
            for (int k = nb - 1; k >= 1; --k) {
                acc.grow(sb[k].box);
                cnt += sb[k].exit;
                right_cost[k] = cnt ? acc.area() * cnt : 0.0;
                right_n[k] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (int k = 0; k < nb - 1; ++k) {
                acc.grow(sb[k].box);
                cnt += sb[k].enter;
                if (cnt == 0 || right_n[k + 1] == 0) continue;
                if ((int64_t)cnt + right_n[k + 1] > (int64_t)c - b) continue;
                const double cost = acc.area() * cnt + right_cost[k + 1];
                if (cost < best.cost) {
                    best.cost = cost;
                    best.kind = 2;
                    best.axis = a;
                    best.bin = k + 1;
                    best.pos = S[a].plane[k + 1];
                }
            }
        }
        return best;
    }

    // Moves the right part [m, e) of an in-place partition so that the left
    // child gets its share of the free room [e, c), proportional to its size.
    // Returns the first index of the right child.
    int32_t place_right(int32_t b, int32_t m, int32_t e, int32_t c) {
        const int64_t room = (int64_t)c - e;
        if (room <= 0) return m;
        const int32_t extra = (int32_t)(room * (m - b) / (e - b));
        if (extra > 0) std::move_backward(refs.begin() + m, refs.begin() + e, refs.begin() + e + extra);
        return m + extra;
    }

    // Spatial partition at plane `pos` of `axis`: references entirely on one
    // side go there, straddling ones are split into both.  Returns false
    // (refs untouched) when a side would be empty or the room is short.
    bool spatial_partition(int32_t b, int32_t e, int32_t c, int axis, float pos, int32_t& m_out,
                           int32_t& rb_out, int32_t& re_out) {
        thread_local std::vector<PrimRef> tl, tr;
        tl.clear();
        tr.clear();
        V3 v[3];
        for (int32_t i = b; i < e; ++i) {
            const PrimRef& r = refs[i];
            const float lo = comp(r.box.lo, axis), hi = comp(r.box.hi, axis);
            if (lo >= pos) {
                tr.push_back(r);
            } else if (hi <= pos) {
                tl.push_back(r);
            } else {
                triangle(r.id, v);
                Box L, R;
                split_triangle(v, axis, pos, L, R);
                L = box_and(L, r.box);
                R = box_and(R, r.box);
                set_comp(L.hi, axis, std::min(comp(L.hi, axis), pos));
                set_comp(R.lo, axis, std::max(comp(R.lo, axis), pos));
                if (!box_empty(L)) tl.push_back(PrimRef{L, L.center(), r.id});
                if (!box_empty(R)) tr.push_back(PrimRef{R, R.center(), r.id});
            }
        }
        const int64_t nl = (int64_t)tl.size(), nr = (int64_t)tr.size();
        if (nl == 0 || nr == 0 || nl + nr > (int64_t)c - b) return false;
        const int64_t room = (int64_t)c - b - nl - nr;
        const int32_t extra = (int32_t)(room * nl / (nl + nr));
        std::copy(tl.begin(), tl.end(), refs.begin() + b);
        const int32_t rb = b + (int32_t)nl + extra;
        std::copy(tr.begin(), tr.end(), refs.begin() + rb);
        m_out = b + (int32_t)nl;
        rb_out = rb;
        re_out = rb + (int32_t)nr;
        return true;
    }

    void enqueue(const Task& t) {
        {
            std::lock_guard<std::mutex> g(mu);
            tasks.push_back(t);
            ++outstanding;
        }
        cv.notify_one();
    }

    void worker() {
        for (;;) {
            Task t;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !tasks.empty() || outstanding == 0; });
                if (tasks.empty()) return;
                t = tasks.back();
                tasks.pop_back();
            }
            build(t.idx, t.b, t.e, t.c, t.depth);
            bool last;
            {
                std::lock_guard<std::mutex> g(mu);
                last = --outstanding == 0;
            }
            if (last) cv.notify_all();
        }
    }

    // Build the whole tree over refs[0, n), room up to refs.size(), with the
    // worker pool.
    void run(int32_t n) {
        next_node = 1;
        const int32_t c = (int32_t)refs.size();
        if (max_threads <= 1 || n < 2 * kTaskMin) {
            build(0, 0, n, c, 1);
            return;
        }
        enqueue(Task{0, 0, n, c, 1});
        std::vector<std::thread> pool;
        for (int i = 0; i < max_threads; ++i) pool.emplace_back([this] { worker(); });
        for (auto& th : pool) th.join();
    }

    void build(int32_t idx, int32_t b, int32_t e, int32_t c, int depth) {
        BNode& nd = nodes[idx];
        Box box, cb;
        for (int32_t i = b; i < e; ++i) {
            box.grow(refs[i].box);
            cb.grow(refs[i].cen);
        }
        nd.box = box;
        nd.prims = e - b;
        if (e - b == 1) {
            nd.first = b;
            nd.count = 1;
            return;
        }
        const int32_t n = e - b;
        int log2n = 0;
        while ((1 << log2n) < n) ++log2n;
        const bool force_median = depth + log2n >= max_depth;
        const int nb = std::min(bins, std::max(4, n));
        const AxisBins ax[3] = {AxisBins(cb.lo.x, cb.hi.x, nb), AxisBins(cb.lo.y, cb.hi.y, nb),
                                AxisBins(cb.lo.z, cb.hi.z, nb)};
        int32_t m = -1, rb = -1, re = e;
        if (!force_median) {
            Choice obj = object_split(b, e, nb, ax);
            // overlap of the object split's children, relative to the root
            bool try_spatial = split_alpha > 0.0f && c > e && root_area > 0.0;
            if (try_spatial && obj.kind == 1) {
                const Box ov = box_and(obj.left, obj.right);
                try_spatial = !box_empty(ov) && ov.area() > split_alpha * root_area;
            }
            if (try_spatial) {
                const Choice sp = spatial_split(b, e, c, box, nb);
                if (sp.kind == 2 && sp.cost < obj.cost &&
                    spatial_partition(b, e, c, sp.axis, sp.pos, m, rb, re)) {
                    obj.kind = 0;
                }
            }
            if (m < 0 && obj.kind == 1) {
                const AxisBins& A = ax[obj.axis];
                const int axis = obj.axis, bin = obj.bin;
                auto mid = std::partition(refs.begin() + b, refs.begin() + e, [&](const PrimRef& r) {
                    return A.bin(comp(r.cen, axis)) <= bin;
                });
                const int32_t mm = (int32_t)(mid - refs.begin());
                if (mm != b && mm != e) {
                    m = mm;
                    rb = place_right(b, m, e, c);
                    re = rb + (e - m);
                }
            }
        }
        if (m < 0) {
            // median split along the widest centroid axis
            V3 d = cb.hi - cb.lo;
            int a = (d.x >= d.y && d.x >= d.z) ? 0 : (d.y >= d.z ? 1 : 2);
            m = b + (e - b) / 2;
            std::nth_element(refs.begin() + b, refs.begin() + m, refs.begin() + e,
                             [&](const PrimRef& p, const PrimRef& q) {
                                 const float cp = comp(p.cen, a), cq = comp(q.cen, a);
                                 return cp < cq || (cp == cq && p.id < q.id);
                             });
            rb = place_right(b, m, e, c);
            re = rb + (e - m);
        }
        int32_t ch = alloc2();
        nd.left = ch;
        nd.right = ch + 1;
        // children: left [b, m) with room up to rb, right [rb, re) with room up
        // to c.  The larger child goes to the pool when it is worth a task;
        // this thread continues with the other (the tree does not depend on
        // which thread builds what, only node numbering does).
        const bool left_big = (m - b) >= (re - rb);
        const Task L{ch, b, m, rb, depth + 1}, R{ch + 1, rb, re, c, depth + 1};
        const Task& big = left_big ? L : R;
        const Task& small = left_big ? R : L;
        if (max_threads > 1 && big.e - big.b >= kTaskMin) enqueue(big);
        else build(big.idx, big.b, big.e, big.c, big.depth);
        build(small.idx, small.b, small.e, small.c, small.depth);
    }
};

bool finite_vertex(const V3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

void build_tree(const cwbvh_mesh* mesh, const cwbvh_build_params& P, Tree& T) {
    const int64_t mesh_nt = mesh->nt();
    int threads = num_threads_or_default(P.num_threads);
    // A triangle with a NaN or infinite coordinate stays out of the tree:
    // woopify gives it rows that never report a hit, and its box would make
    // every ancestor's box (and the quantized child boxes of a CWBVH node)
    // infinite, hiding the triangles beside it.
    std::vector<int32_t> valid;
    valid.reserve((size_t)mesh_nt);
    for (int64_t t = 0; t < mesh_nt; ++t) {
        V3 a, bb, c;
        mesh->tri(t, a, bb, c);
        if (finite_vertex(a) && finite_vertex(bb) && finite_vertex(c)) valid.push_back((int32_t)t);
    }
    const int64_t nt = (int64_t)valid.size();
    // room for the references spatial splits add
    const int64_t cap = P.split_alpha > 0.0f ? nt + (int64_t)((double)nt * P.split_budget) : nt;
    T.split_refs = cap > nt;
    std::vector<PrimRef> refs((size_t)cap);
    parallel_for(nt, threads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const int32_t t = valid[(size_t)i];
            V3 a, bb, c;
            mesh->tri(t, a, bb, c);
            PrimRef& r = refs[(size_t)i];
            r.box = Box();
            r.box.grow(a);
            r.box.grow(bb);
            r.box.grow(c);
            r.cen = r.box.center();
            r.id = t;
        }
    });
    T.nodes.assign(std::max<int64_t>(1, 2 * cap - 1), BNode());
    T.root = 0;
    if (nt == 0) {
        T.nodes.resize(1);
        T.nodes[0].count = 0;
        T.nodes[0].prims = 0;
        return;
    }
    Builder B(mesh, T.nodes, refs, threads, P.sah_bins, P.max_depth, P.split_alpha);
    Box root;
    for (int64_t t = 0; t < nt; ++t) root.grow(refs[(size_t)t].box);
    B.root_area = root.area();
    if (!(B.root_area > 0.0) || !std::isfinite(B.root_area)) B.split_alpha = 0.0f;
    B.run((int32_t)nt);
    T.nodes.resize(B.next_node.load());
    // leaves name ranges of refs; entries outside every leaf's range are
    // never read
    T.refs.resize((size_t)cap);
    parallel_for(cap, threads, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) T.refs[(size_t)i] = refs[(size_t)i].id;
    });
}

// ------------------------------------------------------------------ Woop

void woopify(V3 v0, V3 v1, V3 v2, cwbvh_float4 out[3]) {
    // Rows of A^-1 for A = [v0-v2 | v1-v2 | n], n = (v0-v2)x(v1-v2):
    //   r0 = (e1 x n)/|n|^2, r1 = (n x e0)/|n|^2, r2 = n/|n|^2.
    // Local coordinates (u, v, w) of x are (r0, r1, r2) . (x - v2).
    double e0[3] = {(double)v0.x - v2.x, (double)v0.y - v2.y, (double)v0.z - v2.z};
    double e1[3] = {(double)v1.x - v2.x, (double)v1.y - v2.y, (double)v1.z - v2.z};
    double n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
    double det = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (!(det > 0.0) || !std::isfinite(det)) {
        // degenerate: Dz = 0 for every ray, so t is never finite inside (tmin, tmax)
        out[0] = cwbvh_float4{0.f, 0.f, 0.f, 1.f};
        out[1] = cwbvh_float4{0.f, 0.f, 0.f, 0.f};
        out[2] = cwbvh_float4{0.f, 0.f, 0.f, 0.f};
        return;
    }
    double r0[3] = {(e1[1] * n[2] - e1[2] * n[1]) / det, (e1[2] * n[0] - e1[0] * n[2]) / det,
                    (e1[0] * n[1] - e1[1] * n[0]) / det};
    double r1[3] = {(n[1] * e0[2] - n[2] * e0[1]) / det, (n[2] * e0[0] - n[0] * e0[2]) / det,
                    (n[0] * e0[1] - n[1] * e0[0]) / det};
    double r2[3] = {n[0] / det, n[1] / det, n[2] / det};
    double p[3] = {v2.x, v2.y, v2.z};
    auto d3 = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    out[0] = cwbvh_float4{(float)r2[0], (float)r2[1], (float)r2[2], (float)d3(r2, p)};
    out[1] = cwbvh_float4{(float)r0[0], (float)r0[1], (float)r0[2], (float)-d3(r0, p)};
    out[2] = cwbvh_float4{(float)r1[0], (float)r1[1], (float)r1[2], (float)-d3(r1, p)};
    if (f2u(out[0].x) == 0x80000000u) out[0].x = 0.0f;
}

// Fork-join over the subtrees of a tree: run() starts a task on a new thread
// while fewer than `limit` threads work, and runs it on the caller otherwise.
// The layout tasks write disjoint ranges that are fixed before they start, so
// the bytes do not depend on which thread ran what.
class Forker {
public:
    explicit Forker(int limit) : limit_(limit) {}
    template <class F>
    void run(bool big, std::vector<std::thread>& group, F fn) {
        if (big && limit_ > 1 && extra_.fetch_add(1) < limit_ - 1) {
            group.emplace_back([this, fn]() {
                fn();
                extra_.fetch_sub(1);
            });
            return;
        }
        if (big && limit_ > 1) extra_.fetch_sub(1);
        fn();
    }
    static void join(std::vector<std::thread>& group) {
        for (auto& t : group) t.join();
        group.clear();
    }
    // subtrees with at least this many references are worth a thread
    static constexpr int32_t kMinPrims = 1 << 14;

private:
    std::atomic<int> extra_{0};  // threads started beyond the caller's
    int limit_;
};

// Woop rows and prim_idx entries [prim, -1, -1] of one triangle at row `row`.
void write_triangle(const cwbvh_mesh* mesh, int32_t prim, int64_t row, cwbvh_layout& L) {
    V3 a, b, c;
    mesh->tri(prim, a, b, c);
    woopify(a, b, c, &L.woop[(size_t)row]);
    L.prim_idx[(size_t)row] = prim;
    L.prim_idx[(size_t)row + 1] = -1;
    L.prim_idx[(size_t)row + 2] = -1;
}

// Triangle ids under tree node n, left to right; `st` is scratch.
void gather_prims(const Tree& T, int32_t n, std::vector<int32_t>& st, std::vector<int32_t>& out) {
    st.clear();
    st.push_back(n);
    while (!st.empty()) {
        const int32_t x = st.back();
        st.pop_back();
        const BNode& b = T.nodes[x];
        if (b.left < 0) {
            for (int32_t k = 0; k < b.count; ++k) out.push_back(T.refs[b.first + k]);
        } else {
            st.push_back(b.right);
            st.push_back(b.left);
        }
    }
}

// Drops repeated triangle ids (a triangle split into several references
// can land twice in one leaf), keeping the first of each.
void unique_in_order(std::vector<int32_t>& v) {
    size_t n = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        bool seen = false;
        for (size_t j = 0; j < n && !seen; ++j) seen = v[j] == v[i];
        if (!seen) v[n++] = v[i];
    }
    v.resize(n);
}

// ------------------------------------------------------------------ BVH2

// Child boxes of a BVH2 node at float4 `addr`: (lo.x, hi.x, lo.y, hi.y) per
// child in rows 0/1, the z extents of both in row 2, child links in row 3.
void bvh2_put_box(cwbvh_layout& L, int64_t addr, int slot, const Box& b) {
    cwbvh_float4* n = &L.nodes[(size_t)addr];
    if (slot == 0) {
        n[0] = cwbvh_float4{b.lo.x, b.hi.x, b.lo.y, b.hi.y};
        n[2].x = b.lo.z;
        n[2].y = b.hi.z;
    } else {
        n[1] = cwbvh_float4{b.lo.x, b.hi.x, b.lo.y, b.hi.y};
        n[2].z = b.lo.z;
        n[2].w = b.hi.z;
    }
}

void bvh2_put_child(cwbvh_layout& L, int64_t addr, int slot, int32_t value) {
    if (slot == 0) L.nodes[(size_t)addr + 3].x = i2f(value);
    else L.nodes[(size_t)addr + 3].y = i2f(value);
}

// The BVH2 layout in the order of a pre-order DFS (left child first): inner
// nodes take consecutive 64 B slots, each leaf its triangles' Woop rows plus a
// terminator row where the DFS reaches it.  The bottom-up pass counts the
// inner nodes and Woop rows under every node, which fixes every subtree's
// place, so subtrees are written in parallel.
struct Bvh2Layout {
    const cwbvh_mesh* mesh;
    const Tree& T;
    cwbvh_layout& L;
    std::vector<uint8_t> leaf;
    std::vector<int32_t> sub_nodes, sub_rows;
    bool dup = false;  // spatial splits: a leaf can name a triangle twice
    Forker fork;
    std::atomic<int64_t> leaves{0}, tris{0};

    Bvh2Layout(const cwbvh_mesh* m, const Tree& t, cwbvh_layout& l, int threads)
        : mesh(m), T(t), L(l), fork(threads) {}

    void leaf_prims(int32_t n, std::vector<int32_t>& out) const {
        thread_local std::vector<int32_t> scratch;
        out.clear();
        gather_prims(T, n, scratch, out);
        if (dup) unique_in_order(out);
    }

    // rows of leaf n at `row`: its triangles, then the terminator
    void write_leaf(int32_t n, int64_t row) {
        thread_local std::vector<int32_t> prims;
        leaf_prims(n, prims);
        for (size_t k = 0; k < prims.size(); ++k) write_triangle(mesh, prims[k], row + 3 * (int64_t)k, L);
        const int64_t end = row + 3 * (int64_t)prims.size();
        L.woop[(size_t)end] = cwbvh_float4{u2f(0x80000000u), 0.f, 0.f, 0.f};
        L.prim_idx[(size_t)end] = -1;
        leaves.fetch_add(1, std::memory_order_relaxed);
        tris.fetch_add((int64_t)prims.size(), std::memory_order_relaxed);
    }

    // inner node n at float4 `addr`, its subtree's rows from `row`; returns
    // the depth of the deepest inner node below
    int write_node(int32_t n, int64_t addr, int64_t row, int depth) {
        const BNode& nd = T.nodes[n];
        const int32_t kid[2] = {nd.left, nd.right};
        int deepest[2] = {depth, depth};
        std::vector<std::thread> group;
        int64_t a = addr + 4, r = row;
        for (int slot = 0; slot < 2; ++slot) {
            const int32_t c = kid[slot];
            bvh2_put_box(L, addr, slot, T.nodes[c].box);
            if (leaf[c]) {
                bvh2_put_child(L, addr, slot, ~(int32_t)r);
                write_leaf(c, r);
            } else {
                bvh2_put_child(L, addr, slot, (int32_t)a);
                int* out = &deepest[slot];
                fork.run(T.nodes[c].prims >= Forker::kMinPrims, group,
                         [this, c, a, r, depth, out] { *out = write_node(c, a, r, depth + 1); });
            }
            a += 4 * (int64_t)sub_nodes[c];
            r += sub_rows[c];
        }
        Forker::join(group);
        return std::max(deepest[0], deepest[1]);
    }
};

void layout_bvh2(const cwbvh_mesh* mesh, const Tree& T, const cwbvh_build_params& P, cwbvh_layout& L) {
    const auto& N = T.nodes;
    const int32_t nn = (int32_t)N.size();
    Bvh2Layout B(mesh, T, L, num_threads_or_default(P.num_threads));
    B.dup = T.split_refs;
    // bottom-up SAH: make a leaf wherever it is cheaper (children indices > parent)
    std::vector<double> cost(nn);
    B.leaf.assign(nn, 0);
    B.sub_nodes.assign(nn, 0);
    B.sub_rows.assign(nn, 0);
    std::vector<int32_t> prims;
    for (int32_t i = nn - 1; i >= 0; --i) {
        const BNode& nd = N[i];
        double A = nd.box.area();
        double lc = A * nd.prims * P.bvh2_tri_cost;
        bool is_leaf = nd.left < 0;
        if (is_leaf) {
            cost[i] = lc;
        } else {
            double ic = A * P.bvh2_node_cost + cost[nd.left] + cost[nd.right];
            is_leaf = nd.prims <= P.bvh2_max_leaf && lc <= ic;
            cost[i] = is_leaf ? lc : ic;
        }
        B.leaf[i] = is_leaf ? 1 : 0;
        if (is_leaf) {
            int32_t cnt = nd.prims;
            if (B.dup) {
                B.leaf_prims(i, prims);
                cnt = (int32_t)prims.size();
            }
            B.sub_rows[i] = 3 * cnt + 1;
        } else {
            B.sub_nodes[i] = 1 + B.sub_nodes[nd.left] + B.sub_nodes[nd.right];
            B.sub_rows[i] = B.sub_rows[nd.left] + B.sub_rows[nd.right];
        }
    }
    double root_area = N[T.root].box.area();
    L.sah_cost = root_area > 0 ? cost[T.root] / root_area : 0.0;

    L.kind = CWBVH_KIND_BVH2;
    L.node_stride = 4;
    const BNode& root = N[T.root];
    if (root.prims == 0 || B.leaf[T.root]) {
        // wrap: inner root whose first child is the leaf (or an empty leaf) and
        // whose second child is an empty leaf that can never be hit
        Box inverted;
        inverted.lo = V3{FLT_MAX, FLT_MAX, FLT_MAX};
        inverted.hi = V3{-FLT_MAX, -FLT_MAX, -FLT_MAX};
        const int64_t rows = root.prims ? B.sub_rows[T.root] : 1;
        L.nodes.assign(4, cwbvh_float4{0.f, 0.f, 0.f, 0.f});
        L.woop.assign((size_t)rows + 1, cwbvh_float4{0.f, 0.f, 0.f, 0.f});
        L.prim_idx.assign((size_t)rows + 1, -1);
        bvh2_put_box(L, 0, 0, root.prims ? root.box : inverted);
        bvh2_put_child(L, 0, 0, ~0);
        if (root.prims) {
            B.write_leaf(T.root, 0);
        } else {
            L.woop[0] = cwbvh_float4{u2f(0x80000000u), 0.f, 0.f, 0.f};
            B.leaves.fetch_add(1);
        }
        bvh2_put_box(L, 0, 1, inverted);
        bvh2_put_child(L, 0, 1, ~(int32_t)rows);
        L.woop[(size_t)rows] = cwbvh_float4{u2f(0x80000000u), 0.f, 0.f, 0.f};
        B.leaves.fetch_add(1);
        L.num_nodes = 1;
        L.num_leaves = B.leaves.load();
        L.num_tri_refs = B.tris.load();
        L.max_depth = 1;
        L.stack_need = 2;
        return;
    }
    L.nodes.assign(4 * (size_t)B.sub_nodes[T.root], cwbvh_float4{0.f, 0.f, 0.f, 0.f});
    L.woop.assign((size_t)B.sub_rows[T.root], cwbvh_float4{0.f, 0.f, 0.f, 0.f});
    L.prim_idx.assign((size_t)B.sub_rows[T.root], -1);
    const int maxd = B.write_node(T.root, 0, 0, 1);
    L.num_nodes = B.sub_nodes[T.root];
    L.num_leaves = B.leaves.load();
    L.num_tri_refs = B.tris.load();
    L.max_depth = maxd;
    L.stack_need = maxd + 1;
}

// ----------------------------------------------------------------- CWBVH

struct Collapse {
    const Tree& T;
    const cwbvh_build_params& P;
    std::vector<float> C;        // C[n*8 + i], i = 1..7 (index 0 unused)
    std::vector<uint8_t> dec;    // dec[n*8 + i]: i=1: 1 = leaf, 0 = inner; i>=2: split k (0 = use i-1)
    std::vector<uint8_t> split8;

    Collapse(const Tree& t, const cwbvh_build_params& p) : T(t), P(p) {}

    double inv_root = 0.0;

    // The DP of every node, children before parents: a post-order walk
    // whose large subtrees run on other threads (each node's entries depend
    // on its children's only, so the table is the same for any thread count).
    void run(int threads) {
        const auto& N = T.nodes;
        int32_t nn = (int32_t)N.size();
        C.assign((size_t)nn * 8, 0.f);
        dec.assign((size_t)nn * 8, 0);
        split8.assign(nn, 0);
        // areas relative to the root keep the float cost table finite for any
        // coordinate range
        const double root_area = nn ? N[T.root].box.area() : 0.0;
        inv_root = root_area > 0.0 && std::isfinite(root_area) ? 1.0 / root_area : 0.0;
        if (nn == 0) return;
        Forker fork(threads);
        solve(T.root, fork);
    }

    void solve(int32_t n, Forker& fork) {
        const BNode& nd = T.nodes[n];
        if (nd.left >= 0) {
            std::vector<std::thread> group;
            const int32_t l = nd.left;
            fork.run(T.nodes[l].prims >= Forker::kMinPrims, group, [this, l, &fork] { solve(l, fork); });
            solve(nd.right, fork);
            Forker::join(group);
        }
        solve_node(n);
    }

    void solve_node(int32_t n) {
        const auto& N = T.nodes;
        const BNode& nd = N[n];
        const float A = inv_root > 0.0 ? (float)(nd.box.area() * inv_root) : 0.f;
        float* c = &C[(size_t)n * 8];
        uint8_t* d = &dec[(size_t)n * 8];
        if (nd.left < 0) {
            float lc = A * nd.prims * P.cwbvh_tri_cost;
            for (int i = 1; i < 8; ++i) c[i] = lc;
            d[1] = 1;
            return;
        }
        const float* cl = &C[(size_t)nd.left * 8];
        const float* cr = &C[(size_t)nd.right * 8];
        float dist[9];
        uint8_t dk[9];
        for (int j = 2; j <= 8; ++j) {
            float best = INFINITY;
            uint8_t bk = 1;
            for (int k = 1; k < j; ++k) {
                float v = cl[k] + cr[j - k];
                if (v < best) {
                    best = v;
                    bk = (uint8_t)k;
                }
            }
            dist[j] = best;
            dk[j] = bk;
        }
        float inner = A * P.cwbvh_node_cost + dist[8];
        float leafc = nd.prims <= 3 ? A * nd.prims * P.cwbvh_tri_cost : INFINITY;
        // a leaf slot holds at most 3 triangles (unary count in the meta
        // byte), whatever the costs say
        if (nd.prims <= 3 && leafc <= inner) {
            c[1] = leafc;
            d[1] = 1;
        } else {
            c[1] = inner;
            d[1] = 0;
        }
        split8[n] = dk[8];
        for (int i = 2; i < 8; ++i) {
            if (dist[i] < c[i - 1]) {
                c[i] = dist[i];
                d[i] = dk[i];
            } else {
                c[i] = c[i - 1];
                d[i] = 0;
            }
        }
    }

    // children of the 8-wide node made from binary node n, left to right
    void collect(int32_t n, int i, int32_t out[8], int& cnt) const {
        const BNode& nd = T.nodes[n];
        while (i > 1 && dec[(size_t)n * 8 + i] == 0) --i;
        if (i == 1 || nd.left < 0) {
            out[cnt++] = n;
            return;
        }
        int k = dec[(size_t)n * 8 + i];
        collect(nd.left, k, out, cnt);
        collect(nd.right, i - k, out, cnt);
    }
    int children_of(int32_t n, int32_t out[8]) const {
        const BNode& nd = T.nodes[n];
        int cnt = 0;
        if (nd.left < 0) {
            out[cnt++] = n;
            return cnt;
        }
        int k = split8[n];
        collect(nd.left, k, out, cnt);
        collect(nd.right, 8 - k, out, cnt);
        return cnt;
    }
    bool is_leaf(int32_t n) const { return T.nodes[n].left < 0 || dec[(size_t)n * 8 + 1] == 1; }
};

int exponent_for(double extent) {
    if (!(extent > 0.0)) return -126;
    int e = (int)std::ceil(std::log2(extent / 255.0));
    if (e < -126) e = -126;
    if (e > 127) e = 127;
    while (e < 127 && std::ldexp(255.0, e) < extent) ++e;
    while (e > -126 && std::ldexp(255.0, e - 1) >= extent) --e;
    return e;
}

void assign_slots(const Box& nb, const Box* cbox, int nc, int slot_of[8]) {
    V3 nc_c = nb.center();
    struct Cand { float cost; int s, c; };
    // at most 8 slots x 8 children, sorted by cost with ties kept in
    // (slot, child) order (insertion sort is stable)
    Cand cand[64];
    int ncand = 0;
    for (int s = 0; s < 8; ++s) {
        V3 ds{(s & 4) ? -1.f : 1.f, (s & 2) ? -1.f : 1.f, (s & 1) ? -1.f : 1.f};
        for (int c = 0; c < nc; ++c) {
            const Cand x{dot(cbox[c].center() - nc_c, ds), s, c};
            int j = ncand++;
            while (j > 0 && x.cost < cand[j - 1].cost) {
                cand[j] = cand[j - 1];
                --j;
            }
            cand[j] = x;
        }
    }
    bool used[8] = {};
    for (int c = 0; c < nc; ++c) slot_of[c] = -1;
    int assigned = 0;
    for (int i = 0; i < ncand && assigned < nc; ++i) {
        const Cand& x = cand[i];
        if (used[x.s] || slot_of[x.c] >= 0) continue;
        used[x.s] = true;
        slot_of[x.c] = x.s;
        ++assigned;
    }
}

// floor and ceil of a child's offset in quantization steps, clamped to the
// byte range (NaN gives 0)
inline int quant_floor(double x) {
    if (!(x > 0.0)) return 0;
    return x >= 255.0 ? 255 : (int)x;
}
inline int quant_ceil(double x) {
    if (!(x > 0.0)) return 0;
    if (x >= 255.0) return 255;
    const int t = (int)x;
    return t + ((double)t < x ? 1 : 0);
}

// The CWBVH layout in the order of a DFS that takes a node's inner children
// in reverse slot order (the order a stack pushed in slot order pops them):
// a node's inner children occupy consecutive node slots, allocated when the
// node is reached, and its leaf triangles consecutive Woop rows, appended
// when it is reached.  A first pass counts, under every 8-wide node, the
// nodes its subtree allocates and the triangles it appends; those counts fix
// where every subtree lands, so the second pass writes subtrees in parallel.
struct CwbvhLayout {
    const cwbvh_mesh* mesh;
    const Tree& T;
    const Collapse& CL;
    const cwbvh_build_params& P;
    cwbvh_layout& L;
    const int S;
    double root_area = 0.0;
    bool dup = false;  // spatial splits: a leaf can name a triangle twice
    // per binary node that roots an 8-wide node: nodes and triangles its
    // subtree allocates (its inner children and everything below them)
    std::vector<int32_t> sub_nodes, sub_tris;
    std::vector<double> sah_part;  // per output node, summed in node order
    std::atomic<int64_t> leaves{0};
    Forker fork;

    CwbvhLayout(const cwbvh_mesh* m, const Tree& t, const Collapse& cl, const cwbvh_build_params& p,
                cwbvh_layout& l, int threads)
        : mesh(m), T(t), CL(cl), P(p), L(l), S(p.cwbvh_node_stride), fork(threads) {}

    void leaf_prims(int32_t n, std::vector<int32_t>& out) const {
        thread_local std::vector<int32_t> scratch;
        out.clear();
        gather_prims(T, n, scratch, out);
        if (dup) unique_in_order(out);
    }
    int32_t leaf_count(int32_t n) const {
        if (!dup) return T.nodes[n].prims;
        thread_local std::vector<int32_t> prims;
        leaf_prims(n, prims);
        return (int32_t)prims.size();
    }

    void count(int32_t b) {
        int32_t kids[8];
        const int nc = CL.children_of(b, kids);
        int32_t n_inner = 0, tris = 0;
        std::vector<std::thread> group;
        for (int c = 0; c < nc; ++c) {
            const int32_t k = kids[c];
            if (CL.is_leaf(k)) {
                tris += leaf_count(k);
            } else {
                ++n_inner;
                fork.run(T.nodes[k].prims >= Forker::kMinPrims, group, [this, k] { count(k); });
            }
        }
        Forker::join(group);
        int32_t nodes = n_inner;
        for (int c = 0; c < nc; ++c) {
            if (CL.is_leaf(kids[c])) continue;
            nodes += sub_nodes[kids[c]];
            tris += sub_tris[kids[c]];
        }
        sub_nodes[b] = nodes;
        sub_tris[b] = tris;
    }

    // 8-wide node made from binary node b at node slot `index`; its inner
    // children take node slots from `alloc`, its triangles Woop triangle
    // slots from `wo`.  Returns the depth of the deepest node below.
    int write(int32_t b, int32_t index, int32_t alloc, int32_t wo, int depth) {
        int32_t kids[8];
        const int nc = CL.children_of(b, kids);
        Box cbox[8];
        Box nb;
        for (int c = 0; c < nc; ++c) {
            cbox[c] = T.nodes[kids[c]].box;
            nb.grow(cbox[c]);
        }
        double sah = root_area > 0 ? P.cwbvh_node_cost * nb.area() / root_area : 0.0;
        int slot_of[8];
        assign_slots(nb, cbox, nc, slot_of);
        int32_t child_at[8];
        for (int s = 0; s < 8; ++s) child_at[s] = -1;
        for (int c = 0; c < nc; ++c) child_at[slot_of[c]] = c;

        int ex[3];
        for (int a = 0; a < 3; ++a) ex[a] = exponent_for((double)comp(nb.hi, a) - (double)comp(nb.lo, a));
        uint8_t imask = 0;
        int32_t inner_count = 0, tri_count = 0;
        int32_t child_index[8];
        cwbvh_float4* node = &L.nodes[(size_t)index * S];
        uint8_t* n0b = reinterpret_cast<uint8_t*>(&node[0]);
        uint8_t* meta = reinterpret_cast<uint8_t*>(&node[1]) + 8;
        uint8_t* q = reinterpret_cast<uint8_t*>(&node[2]);
        node[0].x = nb.lo.x;
        node[0].y = nb.lo.y;
        node[0].z = nb.lo.z;
        double scale[3];
        for (int a = 0; a < 3; ++a) scale[a] = std::ldexp(1.0, -ex[a]);
        thread_local std::vector<int32_t> prims;
        for (int s = 0; s < 8; ++s) {
            child_index[s] = -1;
            const int c = child_at[s];
            if (c < 0) continue;
            const Box& bx = cbox[c];
            for (int a = 0; a < 3; ++a) {
                q[a * 8 + s] = (uint8_t)quant_floor(((double)comp(bx.lo, a) - (double)comp(nb.lo, a)) * scale[a]);
                q[24 + a * 8 + s] = (uint8_t)quant_ceil(((double)comp(bx.hi, a) - (double)comp(nb.lo, a)) * scale[a]);
            }
            const int32_t bn = kids[c];
            if (!CL.is_leaf(bn)) {
                imask |= (uint8_t)(1u << s);
                meta[s] = (uint8_t)((1u << 5) | (24u + (unsigned)s));
                child_index[s] = alloc + inner_count++;
            } else {
                leaf_prims(bn, prims);
                const int cnt = (int)prims.size();
                const unsigned unary = cnt == 1 ? 1u : (cnt == 2 ? 3u : 7u);
                meta[s] = (uint8_t)((unary << 5) | (unsigned)tri_count);
                for (int k = 0; k < cnt; ++k)
                    write_triangle(mesh, prims[(size_t)k], 3 * ((int64_t)wo + tri_count + k), L);
                tri_count += cnt;
                leaves.fetch_add(1, std::memory_order_relaxed);
                if (root_area > 0) sah += P.cwbvh_tri_cost * cnt * bx.area() / root_area;
            }
        }
        n0b[12] = (uint8_t)(int8_t)ex[0];
        n0b[13] = (uint8_t)(int8_t)ex[1];
        n0b[14] = (uint8_t)(int8_t)ex[2];
        n0b[15] = imask;
        node[1].x = i2f(inner_count ? alloc : 0);
        node[1].y = i2f(wo);
        sah_part[(size_t)index] = sah;

        // inner children, highest slot first, each with the allocations of
        // the subtrees taken before it
        int deepest[8];
        std::vector<std::thread> group;
        int32_t a_next = alloc + inner_count, w_next = wo + tri_count;
        for (int s = 7; s >= 0; --s) {
            deepest[s] = depth;
            if (child_index[s] < 0) continue;
            const int32_t bn = kids[child_at[s]];
            const int32_t ci = child_index[s], ca = a_next, cw = w_next;
            int* out = &deepest[s];
            fork.run(T.nodes[bn].prims >= Forker::kMinPrims, group,
                     [this, bn, ci, ca, cw, depth, out] { *out = write(bn, ci, ca, cw, depth + 1); });
            a_next += sub_nodes[bn];
            w_next += sub_tris[bn];
        }
        Forker::join(group);
        int d = depth;
        for (int s = 0; s < 8; ++s) d = std::max(d, deepest[s]);
        return d;
    }
};

void layout_cwbvh(const cwbvh_mesh* mesh, const Tree& T, const cwbvh_build_params& P, cwbvh_layout& L) {
    const int threads = num_threads_or_default(P.num_threads);
    const int S = P.cwbvh_node_stride;
    L.kind = CWBVH_KIND_CWBVH;
    L.node_stride = S;
    if (T.nodes.empty() || T.nodes[T.root].prims == 0) {
        L.nodes.assign((size_t)S, cwbvh_float4{0.f, 0.f, 0.f, 0.f});
        L.num_nodes = 1;
        L.max_depth = 1;
        L.stack_need = 2;
        L.sah_cost = 0.0;
        return;
    }
    Collapse CL(T, P);
    CL.run(threads);
    CwbvhLayout W(mesh, T, CL, P, L, threads);
    W.root_area = T.nodes[T.root].box.area();
    W.dup = T.split_refs;
    W.sub_nodes.assign(T.nodes.size(), 0);
    W.sub_tris.assign(T.nodes.size(), 0);
    W.count(T.root);
    const int64_t num_nodes = 1 + (int64_t)W.sub_nodes[T.root];
    const int64_t num_tris = W.sub_tris[T.root];
    L.nodes.assign((size_t)(num_nodes * S), cwbvh_float4{0.f, 0.f, 0.f, 0.f});
    L.woop.assign((size_t)(3 * num_tris), cwbvh_float4{0.f, 0.f, 0.f, 0.f});
    L.prim_idx.assign((size_t)(3 * num_tris), -1);
    W.sah_part.assign((size_t)num_nodes, 0.0);
    const int maxd = W.write(T.root, 0, 1, 0, 1);
    double sah = 0.0;
    for (double v : W.sah_part) sah += v;
    L.num_nodes = num_nodes;
    L.num_leaves = W.leaves.load();
    L.num_tri_refs = num_tris;
    L.max_depth = maxd;
    L.stack_need = maxd + 1;
    L.sah_cost = sah;
}

bool valid_params(const cwbvh_build_params& P) {
    return P.sah_bins >= 4 && P.sah_bins <= 256 && P.bvh2_max_leaf >= 1 && P.bvh2_max_leaf <= 32 &&
           (P.cwbvh_node_stride == 5 || P.cwbvh_node_stride == 8) && P.max_depth >= 8 &&
           P.max_depth <= CWBVH_BVH2_STACK_CAPACITY - 2 && P.bvh2_node_cost >= 0 && P.bvh2_tri_cost > 0 &&
           P.cwbvh_node_cost >= 0 && P.cwbvh_tri_cost > 0 && P.split_alpha >= 0.0f && P.split_alpha <= 1.0f &&
           P.split_budget >= 0.0f && P.split_budget <= 4.0f;
}

}  // namespace

namespace cw {
void woopify_host(const float* v0, const float* v1, const float* v2, cwbvh_float4* out3) {
    woopify(V3{v0[0], v0[1], v0[2]}, V3{v1[0], v1[1], v1[2]}, V3{v2[0], v2[1], v2[2]}, out3);
}
}  // namespace cw

extern "C" {

int32_t cwbvh_build_params_default(cwbvh_build_params* p) {
    CW_CHECK_ARG(p != nullptr, "params is NULL");
    p->num_threads = 0;
    p->sah_bins = 32;
    p->bvh2_node_cost = 1.0f;
    p->bvh2_tri_cost = 1.0f;
    p->bvh2_max_leaf = 8;
    p->cwbvh_node_cost = 1.0f;
    // a triangle test at 0.6 node visits: what one costs in the interleaved
    // traversal loop once its lower lane utilization is counted (0.6..1.5 is
    // a flat optimum on AO rays; 0.3 leaves 4% on the table)
    p->cwbvh_tri_cost = 0.6f;
    p->cwbvh_node_stride = 5;
    p->max_depth = 60;
    p->split_alpha = 0.0f;
    p->split_budget = 0.3f;
    return CWBVH_OK;
}

int32_t cwbvh_build_layouts(const cwbvh_mesh* mesh, const cwbvh_build_params* params, cwbvh_layout** bvh2_out,
                            cwbvh_layout** cwbvh_out) {
    CW_CHECK_ARG(mesh != nullptr, "mesh is NULL");
    CW_CHECK_ARG(bvh2_out != nullptr || cwbvh_out != nullptr, "no output requested");
    cwbvh_build_params P;
    if (params) P = *params;
    else cwbvh_build_params_default(&P);
    CW_CHECK_ARG(valid_params(P), "invalid build parameters");
    CW_CHECK_ARG(mesh->nt() < (int64_t)0x7FFFFFFF / 4, "too many triangles for 32-bit layouts");
    Tree T;
    build_tree(mesh, P, T);
    std::unique_ptr<cwbvh_layout> bvh2(bvh2_out ? new cwbvh_layout() : nullptr);
    std::unique_ptr<cwbvh_layout> cwb(cwbvh_out ? new cwbvh_layout() : nullptr);
    // the two layouts only read the tree: with both requested and more than
    // one thread, the BVH2 one runs beside the CWBVH one
    if (bvh2 && cwb && num_threads_or_default(P.num_threads) > 1) {
        std::thread side([&] { layout_bvh2(mesh, T, P, *bvh2); });
        layout_cwbvh(mesh, T, P, *cwb);
        side.join();
    } else {
        if (bvh2) layout_bvh2(mesh, T, P, *bvh2);
        if (cwb) layout_cwbvh(mesh, T, P, *cwb);
    }
    if (bvh2_out) *bvh2_out = bvh2.release();
    if (cwbvh_out) *cwbvh_out = cwb.release();
    return CWBVH_OK;
}

namespace {

// Structural checks of caller-supplied layouts: every index the kernels derive
// from a reachable node stays inside the arrays, so a malformed array is
// refused on the host instead of being read out of bounds on the device.
// Each node may be reached once (a tree), which also bounds the walk.
bool validate_cwbvh(cwbvh_layout& L, int32_t& maxd, std::string& why) {
    const int64_t S = L.node_stride, nn = L.num_nodes, ntri = (int64_t)L.woop.size() / 3;
    std::vector<uint8_t> seen((size_t)nn, 0);
    std::vector<std::pair<int64_t, int32_t>> st{{0, 1}};
    while (!st.empty()) {
        auto [n, d] = st.back();
        st.pop_back();
        if (n < 0 || n >= nn) { why = "child index out of range"; return false; }
        if (seen[(size_t)n]) { why = "node " + std::to_string(n) + " reached twice"; return false; }
        if (d > CWBVH_STACK_CAPACITY * 8) { why = "tree too deep"; return false; }
        seen[(size_t)n] = 1;
        maxd = std::max(maxd, d);
        const cwbvh_float4* nd = &L.nodes[(size_t)(n * S)];
        const uint8_t imask = reinterpret_cast<const uint8_t*>(&nd[0])[15];
        int32_t base, tbase;
        std::memcpy(&base, &nd[1].x, 4);
        std::memcpy(&tbase, &nd[1].y, 4);
        const uint8_t* meta = reinterpret_cast<const uint8_t*>(&nd[1]) + 8;
        int k = 0;
        for (int s = 0; s < 8; ++s) {
            const unsigned m = meta[s];
            const bool inner_meta = (m & 0x18u) == 0x18u;
            if (imask & (1u << s)) {
                if (m != (0x20u | 0x18u | (unsigned)s)) { why = "inner slot meta does not match imask"; return false; }
                st.push_back({(int64_t)base + k++, d + 1});
            } else if (m != 0) {
                const unsigned unary = m >> 5, off = m & 31u;
                const int cnt = unary == 1 ? 1 : unary == 3 ? 2 : unary == 7 ? 3 : 0;
                if (inner_meta || cnt == 0 || off + (unsigned)cnt > 24u) { why = "bad leaf slot meta"; return false; }
                const int64_t first = (int64_t)tbase + off;
                if (tbase < 0 || first + cnt > ntri) { why = "leaf triangles out of range"; return false; }
                L.num_leaves++;
            }
        }
    }
    return true;
}

bool validate_bvh2(cwbvh_layout& L, int32_t& maxd, std::string& why) {
    const int64_t nf4 = (int64_t)L.nodes.size(), wf4 = (int64_t)L.woop.size();
    std::vector<uint8_t> seen((size_t)(nf4 / 4 + 1), 0);
    std::vector<std::pair<int64_t, int32_t>> st{{0, 1}};
    while (!st.empty()) {
        auto [a, d] = st.back();
        st.pop_back();
        if (a < 0 || a % 4 != 0 || a + 3 >= nf4) { why = "node address out of range"; return false; }
        if (seen[(size_t)(a / 4)]) { why = "node reached twice"; return false; }
        if (d > CWBVH_BVH2_STACK_CAPACITY * 8) { why = "tree too deep"; return false; }
        seen[(size_t)(a / 4)] = 1;
        maxd = std::max(maxd, d);
        int32_t c[2];
        std::memcpy(&c[0], &L.nodes[(size_t)a + 3].x, 4);
        std::memcpy(&c[1], &L.nodes[(size_t)a + 3].y, 4);
        for (int j = 0; j < 2; ++j) {
            if (c[j] >= 0) {
                st.push_back({c[j], d + 1});
                continue;
            }
            // a leaf: Woop rows of 3 float4 closed by a 0x80000000 terminator
            int64_t w = ~(int64_t)c[j];
            for (;;) {
                if (w >= wf4) { why = "leaf runs past the Woop array"; return false; }
                uint32_t x;
                std::memcpy(&x, &L.woop[(size_t)w].x, 4);
                if (x == 0x80000000u) break;
                if (w + 2 >= wf4) { why = "leaf runs past the Woop array"; return false; }
                w += 3;
            }
            L.num_leaves++;
        }
    }
    return true;
}

}  // namespace

int32_t cwbvh_layout_from_arrays(int32_t kind, int32_t node_stride, const cwbvh_float4* nodes, int64_t nodes_f4,
                                 const cwbvh_float4* woop, int64_t woop_f4, const int32_t* prim_idx,
                                 int64_t prim_idx_len, cwbvh_layout** out) {
    CW_CHECK_ARG(out != nullptr, "out is NULL");
    CW_CHECK_ARG(kind == CWBVH_KIND_BVH2 || kind == CWBVH_KIND_CWBVH, "bad kind");
    CW_CHECK_ARG(kind == CWBVH_KIND_BVH2 ? node_stride == 4 : (node_stride == 5 || node_stride == 8),
                 "bad node stride");
    CW_CHECK_ARG(nodes_f4 >= node_stride && nodes_f4 % node_stride == 0 && nodes, "bad node array");
    CW_CHECK_ARG(woop_f4 >= 0 && (woop_f4 == 0 || woop), "bad woop array");
    CW_CHECK_ARG(prim_idx_len >= 0 && (prim_idx_len == 0 || prim_idx), "bad prim index array");
    CW_CHECK_ARG(prim_idx_len == 0 || prim_idx_len >= woop_f4, "prim index array shorter than the Woop array");
    auto* L = new cwbvh_layout();
    L->kind = kind;
    L->node_stride = node_stride;
    L->nodes.assign(nodes, nodes + nodes_f4);
    L->woop.assign(woop, woop + woop_f4);
    L->prim_idx.assign(prim_idx, prim_idx + prim_idx_len);
    L->num_nodes = nodes_f4 / node_stride;
    std::string why;
    int32_t maxd = 0;
    const bool ok = kind == CWBVH_KIND_BVH2 ? validate_bvh2(*L, maxd, why) : validate_cwbvh(*L, maxd, why);
    if (!ok) {
        delete L;
        set_error("malformed " + std::string(kind == CWBVH_KIND_BVH2 ? "BVH2" : "CWBVH") + " layout: " + why);
        return CWBVH_ERR_INVALID_ARGUMENT;
    }
    L->stack_need = maxd + 1;
    L->max_depth = maxd;
    for (int32_t v : L->prim_idx)
        if (v >= 0) L->num_tri_refs++;
    *out = L;
    return CWBVH_OK;
}

int32_t cwbvh_layout_get_view(const cwbvh_layout* L, cwbvh_layout_view* v) {
    CW_CHECK_ARG(L != nullptr && v != nullptr, "NULL argument");
    v->kind = L->kind;
    v->node_stride = L->node_stride;
    v->nodes = L->nodes.data();
    v->nodes_f4 = (int64_t)L->nodes.size();
    v->woop = L->woop.data();
    v->woop_f4 = (int64_t)L->woop.size();
    v->prim_idx = L->prim_idx.data();
    v->prim_idx_len = (int64_t)L->prim_idx.size();
    v->max_depth = L->max_depth;
    v->stack_need = L->stack_need;
    v->num_nodes = L->num_nodes;
    v->num_leaves = L->num_leaves;
    v->num_tri_refs = L->num_tri_refs;
    v->sah_cost = L->sah_cost;
    return CWBVH_OK;
}

int32_t cwbvh_layout_restride(const cwbvh_layout* in, int32_t node_stride, cwbvh_layout** out) {
    CW_CHECK_ARG(in != nullptr && out != nullptr, "NULL argument");
    CW_CHECK_ARG(in->kind == CWBVH_KIND_CWBVH, "restride applies to CWBVH layouts");
    CW_CHECK_ARG(node_stride == 5 || node_stride == 8, "node stride must be 5 or 8");
    auto* L = new cwbvh_layout(*in);
    L->node_stride = node_stride;
    L->nodes.assign((size_t)in->num_nodes * node_stride, cwbvh_float4{0.f, 0.f, 0.f, 0.f});
    for (int64_t n = 0; n < in->num_nodes; ++n)
        for (int k = 0; k < 5; ++k) L->nodes[n * node_stride + k] = in->nodes[n * in->node_stride + k];
    *out = L;
    return CWBVH_OK;
}

void cwbvh_layout_free(cwbvh_layout* layout) { delete layout; }

int32_t cwbvh_woopify(const float* v0, const float* v1, const float* v2, cwbvh_float4* out3) {
    CW_CHECK_ARG(v0 && v1 && v2 && out3, "NULL argument");
    cw::woopify_host(v0, v1, v2, out3);
    return CWBVH_OK;
}

}  // extern "C"
