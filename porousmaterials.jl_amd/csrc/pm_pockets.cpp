// pm_pockets.cpp -- pocket blocking of an accessibility grid, host code.
//
// compute_accessibility_grid (src/grid.jl:689-774) thresholds an energy grid
// and, with block_inaccessible_pockets, keeps only the points of segments
// through which a molecule can leave the unit cell:
//
//   1. flood fill (src/grid.jl:383-456): points are scanned with i outermost
//      and k innermost; an unassigned point with energy > tol is marked -1,
//      any other unassigned point seeds the next segment, whose FIFO fill
//      takes each unassigned 26-neighbour (inside the grid, no periodic wrap)
//      with energy < tol and marks the other unassigned neighbours -1;
//   2. connections (src/grid.jl:485-530): each pair of labels facing each
//      other across the c, a and b faces of the cell (points k = n3 and 1,
//      i = n1 and 1, j = n2 and 1) gives a directed connection (src, dst,
//      +unit direction) and its reverse (dst, src, -direction), once;
//   3. graph (src/grid.jl:533-566): vertices are the segments and one vertex
//      per connection, edges src -> connection -> dst;
//   4. classification (src/grid.jl:569-629): a segment is accessible if it
//      lies on a simple cycle of that graph whose connection directions sum
//      to a nonzero translation.  (The reference's follow-up pass over
//      segment -> segment edges finds none: every edge ends or starts at a
//      connection vertex.)  Other segments are blocked (:632-646) and the
//      accessible points are those left in accessible segments (:750-759).
//
// Simple cycles are enumerated with Johnson's algorithm (D. B. Johnson,
// "Finding all the elementary circuits of a directed graph", SIAM J. Comput.
// 4 (1975) 77-84), iteratively.  Every cycle passes through a segment vertex
// and segment vertices are numbered first, so only segment vertices start a
// search.
#include "pm_pockets.h"

#include <set>
#include <tuple>
#include <vector>

namespace pmhost {
namespace {

struct Connection {
    int src, dst;  // 0-based segment vertices
    int d[3];      // unit-cell translation crossed going src -> dst
};

int flood_fill(const int64_t n0, const int64_t n1, const int64_t n2, const double* data, double tol,
               std::vector<int32_t>& seg) {
    const int64_t npts = n0 * n1 * n2;
    seg.assign((size_t)npts, 0);
    std::vector<int64_t> queue;
    int32_t n_seg = 0;
    for (int64_t i = 0; i < n0; ++i)
        for (int64_t j = 0; j < n1; ++j)
            for (int64_t k = 0; k < n2; ++k) {
                const int64_t p = i + n0 * (j + n1 * k);
                if (seg[p] != 0) continue;
                if (data[p] > tol) {
                    seg[p] = -1;
                    continue;
                }
                ++n_seg;
                queue.clear();
                queue.push_back(p);
                for (size_t head = 0; head < queue.size(); ++head) {
                    const int64_t q = queue[head];
                    seg[q] = n_seg;
                    const int64_t qi = q % n0, qj = (q / n0) % n1, qk = q / (n0 * n1);
                    for (int di = -1; di <= 1; ++di)
                        for (int dj = -1; dj <= 1; ++dj)
                            for (int dk = -1; dk <= 1; ++dk) {
                                if (di == 0 && dj == 0 && dk == 0) continue;
                                const int64_t a = qi + di, b = qj + dj, c = qk + dk;
                                if (a < 0 || b < 0 || c < 0 || a >= n0 || b >= n1 || c >= n2) continue;
                                const int64_t r = a + n0 * (b + n1 * c);
                                if (seg[r] != 0) continue;
                                if (data[r] < tol) {
                                    seg[r] = n_seg;
                                    queue.push_back(r);
                                } else {
                                    seg[r] = -1;
                                }
                            }
                }
            }
    return n_seg;
}

struct ConnectionList {
    std::vector<Connection> list;
    std::set<std::tuple<int, int, int, int, int>> seen;

    void note(int32_t s1, int32_t s2, int d0, int d1, int d2) {
        if (s1 == -1 || s2 == -1) return;
        if (!seen.insert(std::make_tuple(s1, s2, d0, d1, d2)).second) return;
        seen.insert(std::make_tuple(s2, s1, -d0, -d1, -d2));
        list.push_back({s1 - 1, s2 - 1, {d0, d1, d2}});
        list.push_back({s2 - 1, s1 - 1, {-d0, -d1, -d2}});
    }
};

// Johnson's circuit search from every segment vertex; marks the segments of
// each cycle with a nonzero net translation.  Returns false past max_cycles.
bool classify(int n_seg, const std::vector<Connection>& conns, int64_t max_cycles,
              std::vector<char>& accessible, int64_t& n_cycles) {
    const int nv = n_seg + (int)conns.size();
    std::vector<std::vector<int>> adj((size_t)nv);
    for (size_t c = 0; c < conns.size(); ++c) {
        const int cv = n_seg + (int)c;
        adj[(size_t)conns[c].src].push_back(cv);
        adj[(size_t)cv].push_back(conns[c].dst);
    }
    accessible.assign((size_t)n_seg, 0);
    n_cycles = 0;
    std::vector<char> blocked((size_t)nv, 0);
    std::vector<std::vector<int>> B((size_t)nv);
    struct Frame {
        int v;
        size_t next;
        bool found;
    };
    std::vector<Frame> frames;
    std::vector<int> work;
    auto unblock = [&](int u) {
        blocked[(size_t)u] = 0;
        work.assign(1, u);
        while (!work.empty()) {
            const int x = work.back();
            work.pop_back();
            for (int w : B[(size_t)x])
                if (blocked[(size_t)w]) {
                    blocked[(size_t)w] = 0;
                    work.push_back(w);
                }
            B[(size_t)x].clear();
        }
    };
    auto dir = [&](int v, int axis) { return v >= n_seg ? conns[(size_t)(v - n_seg)].d[axis] : 0; };
    for (int s = 0; s < n_seg; ++s) {
        for (int v = s; v < nv; ++v) {
            blocked[(size_t)v] = 0;
            B[(size_t)v].clear();
        }
        long long t[3] = {0, 0, 0};  // net translation of the path
        frames.assign(1, Frame{s, 0, false});
        blocked[(size_t)s] = 1;
        while (!frames.empty()) {
            Frame& f = frames.back();
            const int v = f.v;
            if (f.next < adj[(size_t)v].size()) {
                const int w = adj[(size_t)v][f.next++];
                if (w < s) continue;  // the search from s stays on vertices >= s
                if (w == s) {
                    f.found = true;
                    if (++n_cycles > max_cycles) return false;
                    if (t[0] != 0 || t[1] != 0 || t[2] != 0)
                        for (const Frame& g : frames)
                            if (g.v < n_seg) accessible[(size_t)g.v] = 1;
                } else if (!blocked[(size_t)w]) {
                    blocked[(size_t)w] = 1;
                    for (int a = 0; a < 3; ++a) t[a] += dir(w, a);
                    frames.push_back(Frame{w, 0, false});
                }
                continue;
            }
            const bool found = f.found;
            if (found) {
                unblock(v);
            } else {
                for (int w : adj[(size_t)v]) {
                    if (w < s) continue;
                    std::vector<int>& bw = B[(size_t)w];
                    bool in = false;
                    for (int x : bw) in |= x == v;
                    if (!in) bw.push_back(v);
                }
            }
            for (int a = 0; a < 3; ++a) t[a] -= dir(v, a);
            frames.pop_back();
            if (found && !frames.empty()) frames.back().found = true;
        }
    }
    return true;
}

}  // namespace

std::string block_pockets(const int32_t n[3], const double* data, double tol, bool block,
                          int64_t max_cycles, uint8_t* out_acc, int32_t* out_seg, int64_t* counts) {
    const int64_t n0 = n[0], n1 = n[1], n2 = n[2];
    const int64_t npts = n0 * n1 * n2;
    if (counts) counts[0] = counts[1] = counts[2] = 0;
    if (!block) {
        for (int64_t p = 0; p < npts; ++p) out_acc[p] = data[p] < tol ? 1 : 0;
        return "";
    }
    std::vector<int32_t> seg;
    const int n_seg = flood_fill(n0, n1, n2, data, tol, seg);
    if (out_seg)
        for (int64_t p = 0; p < npts; ++p) out_seg[p] = seg[(size_t)p];
    auto at = [&](int64_t i, int64_t j, int64_t k) { return seg[(size_t)(i + n0 * (j + n1 * k))]; };
    ConnectionList cl;
    for (int64_t i = 0; i < n0; ++i)
        for (int64_t j = 0; j < n1; ++j) cl.note(at(i, j, n2 - 1), at(i, j, 0), 0, 0, 1);
    for (int64_t j = 0; j < n1; ++j)
        for (int64_t k = 0; k < n2; ++k) cl.note(at(n0 - 1, j, k), at(0, j, k), 1, 0, 0);
    for (int64_t i = 0; i < n0; ++i)
        for (int64_t k = 0; k < n2; ++k) cl.note(at(i, n1 - 1, k), at(i, 0, k), 0, 1, 0);
    std::vector<char> acc;
    int64_t n_cycles = 0;
    if (!classify(n_seg, cl.list, max_cycles, acc, n_cycles))
        return "pm_block_pockets: more than max_cycles simple cycles in the segment graph";
    int64_t n_blocked = 0;
    for (int s = 0; s < n_seg; ++s) n_blocked += acc[(size_t)s] ? 0 : 1;
    for (int64_t p = 0; p < npts; ++p) {
        const int32_t l = seg[(size_t)p];
        out_acc[p] = (l > 0 && acc[(size_t)(l - 1)]) ? 1 : 0;
    }
    if (counts) {
        counts[0] = n_seg;
        counts[1] = n_blocked;
        counts[2] = n_cycles;
    }
    return "";
}

}  // namespace pmhost
