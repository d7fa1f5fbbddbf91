// HiG-Tree mesh, simulation domains and point-stencil builder (host side).
//
// This is the product's host surface for HiG-Tree (sd_get_stencil /
// sfd_get_stencil / stn_* / dp_interpolate_from_stencil, higtree/src/domain.h
// :317-387, pdomain.h:292-315).  It is also the generator of the device
// operators: every point evaluation the step makes is resolved here once, at
// setup, into an affine row  value = -rhs + sum_k w_k * field[id_k],  whose
// rhs is kept as linear coefficients of boundary values so time-dependent
// boundary data only re-folds constants.
//
// Semantics follow the reference exactly (candidate order, pruning, glibc
// msort tie order, absolute-coordinate monomials, EPS tolerances) because the
// selected WLS samples decide the weights to O(1e-3).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "hfh_progress.hpp"

namespace hfh {

constexpr double EPSDELTA = 1.0e-8;  // higtree/src/utils.h:16
constexpr double EPSMACH = 1.0e-12;  // higtree/src/utils.h:25
inline bool flt_gt(double x, double y) { return x > y + EPSMACH; }
inline bool flt_lt(double x, double y) { return x < y - EPSMACH; }
inline bool flt_le(double x, double y) { return x < y + EPSMACH; }
inline bool flt_eq(double x, double y) { return std::fabs(x - y) < EPSMACH; }
inline bool pos_le(double x, double y) { return x < y + EPSDELTA; }
inline bool pos_ge(double x, double y) { return x > y - EPSDELTA; }
inline double rmax(double a, double b) { return a > b ? a : b; }


enum BcType { DIRICHLET = 0, NEUMANN = 1 };

template <int D>
struct Vec {
    double v[D];
    double &operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};

// One tree node.  Children of a refined node are stored contiguously in the
// arena, ordered by hig_toposition (last dimension fastest).
template <int D>
struct Node {
    Vec<D> lo, hi;
    int nc[D];
    int parent = -1;
    int pos = -1;
    int child0 = -1;  // arena index of first child, -1 for a leaf
    int cid = 0;
    int fid[2 * D];
};

template <int D>
struct Facet {
    int cell;  // arena index
    int dim, dir;
};

struct IdGen {
    int next = 1;
    int get() { return next++; }
};

template <int D>
inline int toposition(const int *nc, const int *pp) {
    int r = 0;
    for (int i = 0; i < D; i++) r = r * nc[i] + pp[i];
    return r;
}

template <int D>
inline void tobase(int pos, const int *nc, int *pp) {
    for (int i = D - 1; i >= 0; i--) {
        pp[i] = pos % nc[i];
        pos /= nc[i];
    }
}

template <int D>
inline double distance(const Vec<D> &a, const Vec<D> &b) {
    double s = 0;
    for (int i = 0; i < D; i++) {
        const double d = a[i] - b[i];
        s += d * d;
    }
    return std::sqrt(s);
}

template <int D>
inline bool point_equal(const Vec<D> &a, const Vec<D> &b) {
    for (int i = 0; i < D; i++)
        if (std::fabs(a[i] - b[i]) >= EPSMACH) return false;
    return true;
}

template <int D>
inline bool rect_contains(const Vec<D> &lo, const Vec<D> &hi, const Vec<D> &x) {
    for (int i = 0; i < D; i++) {
        if (!pos_le(x[i], hi[i])) return false;
        if (!pos_ge(x[i], lo[i])) return false;
    }
    return true;
}

// A HiG-Tree: arena of nodes, root at index 0.
template <int D>
class Tree {
  public:
    std::vector<Node<D>> n;

    int nchildren(int c) const {
        int k = 1;
        for (int i = 0; i < D; i++) k *= n[c].nc[i];
        return n[c].child0 < 0 ? 0 : k;
    }
    const Node<D> &root() const { return n[0]; }

    Vec<D> center(int c) const {  // higtree.c:61-66
        Vec<D> p;
        for (int i = 0; i < D; i++) {
            p[i] = n[c].lo[i] + n[c].hi[i];
            p[i] = p[i] / 2.0;
        }
        return p;
    }
    Vec<D> delta(int c) const {
        Vec<D> p;
        for (int i = 0; i < D; i++) p[i] = n[c].hi[i] - n[c].lo[i];
        return p;
    }
    Vec<D> facet_center(const Facet<D> &f) const {  // higtree.c:676-687
        Vec<D> p = center(f.cell);
        p[f.dim] = f.dir == 0 ? n[f.cell].lo[f.dim] : n[f.cell].hi[f.dim];
        return p;
    }

    int make_node(IdGen &g, const Vec<D> &lo, const Vec<D> &hi) {
        Node<D> c;
        c.lo = lo;
        c.hi = hi;
        for (int i = 0; i < D; i++) c.nc[i] = 0;
        c.cid = g.get();
        for (int i = 0; i < 2 * D; i++) c.fid[i] = g.get();
        n.push_back(c);
        return (int)n.size() - 1;
    }

    // higtree.c:257-284 (same floating-point operation order)
    void refine(IdGen &g, int c, const int *nc) {
        Vec<D> dl;
        int k = 1;
        for (int i = 0; i < D; i++) {
            n[c].nc[i] = nc[i];
            dl[i] = n[c].hi[i] - n[c].lo[i];
            dl[i] = dl[i] / nc[i];
            k *= nc[i];
        }
        int first = (int)n.size();
        n[c].child0 = first;
        for (int j = 0; j < k; j++) {
            int pp[D];
            tobase<D>(j, nc, pp);
            Vec<D> lo, hi;
            for (int i = 0; i < D; i++) {
                lo[i] = dl[i] * pp[i];
                lo[i] = lo[i] + n[c].lo[i];
                hi[i] = lo[i] + dl[i];
            }
            int ch = make_node(g, lo, hi);
            n[ch].parent = c;
            n[ch].pos = j;
        }
    }

    // higtree.c:371-415
    void coords_of_point(int c, const Vec<D> &x, int *pp) const {
        const Node<D> &cell = n[c];
        for (int i = 0; i < D; i++) pp[i] = 0;
        for (int dim = 0; dim < D; dim++) {
            int first = 0, last = cell.nc[dim] - 1;
            while (last >= first) {
                pp[dim] = first;
                if (last == first) break;
                int pf = cell.child0 + toposition<D>(cell.nc, pp);
                pp[dim] = last;
                int pl = cell.child0 + toposition<D>(cell.nc, pp);
                double dl = (n[pl].hi[dim] - n[pf].lo[dim]) / (last - first + 1);
                int pos = (int)((x[dim] - n[pf].lo[dim]) / dl) + first;
                if (pos > last) pos = last;
                if (pos < first) pos = first;
                pp[dim] = pos;
                int pq = cell.child0 + toposition<D>(cell.nc, pp);
                if (flt_le(x[dim], n[pq].lo[dim])) last = pos - 1;
                else if (flt_gt(x[dim], n[pq].hi[dim])) first = pos + 1;
                else break;
            }
        }
        for (int dim = 0; dim < D; dim++) {
            if (flt_lt(x[dim], cell.lo[dim])) pp[dim] = -1;
            if (flt_gt(x[dim], cell.hi[dim])) pp[dim] = cell.nc[dim];
        }
    }

    // higtree.c:417-433 (descends from node `top`)
    int cell_with_point(const Vec<D> &x, int top = 0) const {
        int c = top;
        for (int i = 0; i < D; i++)
            if (flt_gt(n[c].lo[i], x[i]) || flt_lt(n[c].hi[i], x[i])) return -1;
        while (n[c].child0 >= 0) {
            int pp[D];
            coords_of_point(c, x, pp);
            c = n[c].child0 + toposition<D>(n[c].nc, pp);
        }
        return c;
    }

    // higtree.c:445-455
    int cell_with_point_from(int c, const Vec<D> &x) const {
        while (c >= 0) {
            bool out = false;
            for (int i = 0; i < D; i++)
                if (flt_lt(x[i], n[c].lo[i]) || flt_gt(x[i], n[c].hi[i])) out = true;
            if (!out) return cell_with_point(x, c);
            c = n[c].parent;
        }
        return -1;
    }

    int narrowest_dim() const {  // higtree.c:173-184
        double minw = 0;
        int dim = -1;
        for (int i = 0; i < D; i++) {
            double w = n[0].hi[i] - n[0].lo[i];
            if (dim == -1 || w < minw) {
                minw = w;
                dim = i;
            }
        }
        return dim;
    }

    // higtree.c:710-750
    bool facet_with_point(int dim, const Vec<D> &x, Facet<D> &f) const {
        int c = cell_with_point(x);
        if (c < 0) return false;
        int dir = -1;
        if (flt_eq(x[dim], n[c].lo[dim])) dir = 0;
        else if (flt_eq(x[dim], n[c].hi[dim])) dir = 1;
        if (dir == -1) return false;
        if (n[c].fid[dim * 2 + dir] > 0) {
            f = {c, dim, dir};
            return true;
        }
        Vec<D> x2 = x;
        if (dir == 0) x2[dim] -= EPSDELTA;
        else x2[dim] += EPSDELTA;
        int c2 = cell_with_point_from(c, x2);
        if (c2 < 0) return false;
        f = {c2, dim, 1 - dir};
        return true;
    }

    bool intersects(int c, const Vec<D> &lo, const Vec<D> &hi) const {  // higtree.c:479-489
        for (int i = 0; i < D; i++)
            if (flt_gt(lo[i], n[c].hi[i]) || flt_lt(hi[i], n[c].lo[i])) return false;
        return true;
    }

    bool facet_intersects(const Facet<D> &f, const Vec<D> &lo, const Vec<D> &hi) const {
        Vec<D> fl = n[f.cell].lo, fh = n[f.cell].hi;
        if (f.dir == 1) fl[f.dim] = n[f.cell].hi[f.dim];
        else fh[f.dim] = n[f.cell].lo[f.dim];
        for (int i = 0; i < D; i++)
            if (!((fl[i] < hi[i]) && (lo[i] < fh[i]))) return false;
        return true;
    }

    template <class F>
    void leaves(F &&fn, int c = 0) const {  // depth-first, children in order
        if (n[c].child0 < 0) {
            fn(c);
            return;
        }
        int k = nchildren(c);
        for (int j = 0; j < k; j++) leaves(fn, n[c].child0 + j);
    }

    // higtree-iterator.c:151-249: per level an odometer between the corner
    // children (last dim fastest); the low corner child is always visited.
    template <class F>
    void bbox_cells(const Vec<D> &lo, const Vec<D> &hi, F &&fn) const {
        if (intersects(0, lo, hi)) bbox_rec(0, lo, hi, fn);
    }

    template <class F>
    void bbox_facets(const int *interest, const Vec<D> &lo, const Vec<D> &hi, F &&fn) const {
        bbox_cells(lo, hi, [&](int c) {
            for (int dim = 0; dim < D; dim++)
                for (int dir = 0; dir < 2; dir++) {
                    if (!interest[dim] || n[c].fid[dim * 2 + dir] <= 0) continue;
                    Facet<D> f{c, dim, dir};
                    if (facet_intersects(f, lo, hi)) fn(f);
                }
        });
    }

    // bbox_cells / bbox_facets in the same order, skipping every child box
    // whose squared distance from x exceeds far2() (read before each child,
    // never increasing).  Children of a node form a regular grid, so a box
    // too far in the leading dimensions ends its run along the last one, as
    // does a box too far beyond x in the last dimension.
    template <class F, class G>
    void bbox_cells_near(const Vec<D> &lo, const Vec<D> &hi, const Vec<D> &x, G &&far2, F &&fn) const {
        if (intersects(0, lo, hi)) near_rec(0, lo, hi, x, far2, fn);
    }

    template <class F, class G>
    void bbox_facets_near(const int *interest, const Vec<D> &lo, const Vec<D> &hi, const Vec<D> &x, G &&far2,
                          F &&fn) const {
        bbox_cells_near(lo, hi, x, far2, [&](int c) {
            for (int dim = 0; dim < D; dim++)
                for (int dir = 0; dir < 2; dir++) {
                    if (!interest[dim] || n[c].fid[dim * 2 + dir] <= 0) continue;
                    Facet<D> f{c, dim, dir};
                    if (facet_intersects(f, lo, hi)) fn(f);
                }
        });
    }

  private:
    template <class F, class G>
    void near_rec(int c, const Vec<D> &lo, const Vec<D> &hi, const Vec<D> &x, G &far2, F &fn) const {
        if (n[c].child0 < 0) {
            fn(c);
            return;
        }
        int pl[D], ph[D], ps[D];
        coords_of_point(c, lo, pl);
        coords_of_point(c, hi, ph);
        for (int i = 0; i < D; i++) {
            if (pl[i] < 0) pl[i] = 0;
            if (ph[i] >= n[c].nc[i]) ph[i] = n[c].nc[i] - 1;
            ps[i] = pl[i];
        }
        for (;;) {
            const int ch = n[c].child0 + toposition<D>(n[c].nc, ps);
            const Node<D> &q = n[ch];
            const double lim = far2();
            double a = 0;
            for (int i = 0; i < D - 1; i++) {
                const double g = rmax(rmax(q.lo[i] - x[i], x[i] - q.hi[i]), 0.0);
                a += g * g;
            }
            const double g = rmax(rmax(q.lo[D - 1] - x[D - 1], x[D - 1] - q.hi[D - 1]), 0.0);
            if (a > lim) ps[D - 1] = ph[D - 1];
            else if (a + g * g > lim) {
                if (q.lo[D - 1] > x[D - 1]) ps[D - 1] = ph[D - 1];
            } else {
                near_rec(ch, lo, hi, x, far2, fn);
            }
            int i = D - 1;
            while (i >= 0 && ps[i] >= ph[i]) {
                ps[i] = pl[i];
                i--;
            }
            if (i < 0) break;
            ps[i]++;
        }
    }

    template <class F>
    void bbox_rec(int c, const Vec<D> &lo, const Vec<D> &hi, F &fn) const {
        if (n[c].child0 < 0) {
            fn(c);
            return;
        }
        int pl[D], ph[D], ps[D];
        coords_of_point(c, lo, pl);
        coords_of_point(c, hi, ph);
        for (int i = 0; i < D; i++) {
            if (pl[i] < 0) pl[i] = 0;
            if (ph[i] >= n[c].nc[i]) ph[i] = n[c].nc[i] - 1;
            ps[i] = pl[i];
        }
        bbox_rec(n[c].child0 + toposition<D>(n[c].nc, ps), lo, hi, fn);
        for (;;) {
            int i = D - 1;
            while (i >= 0 && ps[i] >= ph[i]) {
                ps[i] = pl[i];
                i--;
            }
            if (i < 0) break;
            ps[i]++;
            bbox_rec(n[c].child0 + toposition<D>(n[c].nc, ps), lo, hi, fn);
        }
    }
};

// higtree-io.c:269-293 (parse) and :97-153 (build).  Returns false on a
// malformed text.
template <int D>
bool read_amr(IdGen &g, const std::string &text, Tree<D> &t) {
    const char *p = text.c_str();
    char *end;
    auto nd = [&](double &v) {
        v = std::strtod(p, &end);
        if (end == p) return false;
        p = end;
        return true;
    };
    auto ni = [&](int &v) {
        v = (int)std::strtol(p, &end, 10);
        if (end == p) return false;
        p = end;
        return true;
    };
    Vec<D> l, h;
    for (int d = 0; d < D; d++)
        if (!nd(l[d]) || !nd(h[d])) return false;
    int nl;
    if (!ni(nl) || nl < 1) return false;
    std::vector<Vec<D>> ldelta(nl);
    std::vector<std::vector<int>> init(nl), size(nl);
    for (int lv = 0; lv < nl; lv++) {
        for (int d = 0; d < D; d++)
            if (!nd(ldelta[lv][d])) return false;
        int np;
        if (!ni(np) || np < 0) return false;
        init[lv].resize(np * D);
        size[lv].resize(np * D);
        for (int q = 0; q < np; q++) {
            for (int d = 0; d < D; d++)
                if (!ni(init[lv][q * D + d])) return false;
            for (int d = 0; d < D; d++)
                if (!ni(size[lv][q * D + d])) return false;
        }
    }
    if (size[0].size() < (size_t)D) return false;
    for (int d = 0; d < D; d++)
        if (size[0][d] < 1) return false;
    t.n.clear();
    int r = t.make_node(g, l, h);
    t.n[r].parent = -1;
    t.refine(g, r, size[0].data());
    int two[D];
    for (int d = 0; d < D; d++) two[d] = 2;
    for (int lv = 1; lv < nl; lv++) {
        int np = (int)size[lv].size() / D;
        for (int q = 0; q < np; q++) {
            const int *pi = &init[lv][q * D];
            const int *psz = &size[lv][q * D];
            Vec<D> ip;
            for (int d = 0; d < D; d++) {
                ip[d] = ldelta[lv][d] * pi[d];
                ip[d] = ip[d] + l[d];
            }
            int k[D] = {0};
            for (;;) {
                Vec<D> pt;
                for (int d = 0; d < D; d++) {
                    double off = ldelta[lv][d] * k[d];
                    pt[d] = ip[d] + off;
                }
                int c = t.cell_with_point(pt);
                if (c >= 0) {
                    Vec<D> cd = t.delta(c);
                    bool enough = true;
                    for (int d = 0; d < D; d++)
                        if (flt_lt(ldelta[lv][d], cd[d])) {
                            enough = false;
                            break;
                        }
                    if (!enough) t.refine(g, c, two);
                }
                int i = 0;
                while (i < D && k[i] + 2 >= psz[i] - 1) {
                    k[i] = 0;
                    i++;
                }
                if (i < D) k[i] += 2;
                else break;
            }
        }
    }
    return true;
}

// Dense id -> local-index map (ids are small positive integers).
struct IdMap {
    std::vector<int> lid;  // lid+1, 0 undefined
    int count = 0;
    void init(int maxid) {
        lid.assign(maxid + 1, 0);
        count = 0;
    }
    int lookup(int id) const {
        if (id <= 0 || id >= (int)lid.size()) return -1;
        return lid[id] - 1;
    }
    void assign(int id) {
        if (id <= 0 || id >= (int)lid.size()) return;
        if (lid[id] == 0) lid[id] = ++count;
    }
};

// sim_boundary (higtree/src/domain.h:29-38).  `base` is this boundary's first
// slot in the solver-wide boundary-value vector.
template <int D>
struct Boundary {
    const Tree<D> *tree;
    int type, userid;
    IdMap map;
    std::vector<int> cells;  // local id -> node
    int base = 0;
    int pdir = 0;
    void build(int maxid) {
        map.init(maxid);
        tree->leaves([&](int c) { map.assign(tree->n[c].cid); });
        cells.assign(map.count, -1);
        tree->leaves([&](int c) { cells[map.lookup(tree->n[c].cid)] = c; });
        pdir = tree->narrowest_dim();
    }
};

// Index over an append-only list of ids that accumulates repeated ids
// (stencils, row combinations): up to 16 entries the list is scanned, past
// that an open-addressed table holds position + 1 per slot.  Positions are
// those of the list, so first-insertion order and the order of the sums
// are those of a plain scan.
class IdIndex {
  public:
    void clear() { slot_.clear(); }
    // position of id among the n entries (id of entry i: key(i)), or -1
    template <class K>
    int find(int id, size_t n, const K &key) const {
        if (slot_.empty()) {
            for (size_t i = 0; i < n; i++)
                if (key(i) == id) return (int)i;
            return -1;
        }
        const size_t m = slot_.size() - 1;
        for (size_t h = hash(id) & m; slot_[h]; h = (h + 1) & m)
            if (key((size_t)slot_[h] - 1) == id) return slot_[h] - 1;
        return -1;
    }
    // id was appended as entry n - 1
    template <class K>
    void appended(int id, size_t n, const K &key) {
        if (slot_.empty() ? n > 16 : 2 * n > slot_.size()) {
            size_t cap = 64;
            while (cap < 4 * n) cap *= 2;
            slot_.assign(cap, 0);
            for (size_t i = 0; i < n; i++) put(key(i), (int)i + 1);
        } else if (!slot_.empty()) {
            put(id, (int)n);
        }
    }

  private:
    std::vector<int> slot_;
    static size_t hash(int id) { return (size_t)((uint32_t)id * 0x9E3779B1u) >> 7; }
    void put(int id, int v) {
        const size_t m = slot_.size() - 1;
        size_t h = hash(id) & m;
        while (slot_[h]) h = (h + 1) & m;
        slot_[h] = v;
    }
};

// Affine point stencil with boundary values kept symbolic:
//   value = -rhs0 - sum_b bc[b].second * bcval[bc[b].first] + sum_k w_k f[id_k]
struct Stencil {
    std::vector<int> ids;
    std::vector<double> w;
    std::vector<std::pair<int, double>> bc;  // (global bc slot, coefficient of rhs)
    double rhs0 = 0.0;
    void reset() {
        ids.clear();
        w.clear();
        bc.clear();
        rhs0 = 0.0;
        index_.clear();
    }
    void add(int id, double v) {  // domain.c:1994-2010
        auto key = [this](size_t i) { return ids[i]; };
        const int i = index_.find(id, ids.size(), key);
        if (i >= 0) {
            w[i] += v;
            return;
        }
        ids.push_back(id);
        w.push_back(v);
        index_.appended(id, ids.size(), key);
    }
    void add_rhs_bc(int slot, double coef) { bc.emplace_back(slot, coef); }
    double rhs(const std::vector<double> &bcval) const {
        double r = rhs0;
        for (auto &b : bc) r += b.second * bcval[b.first];
        return r;
    }
    // room for `more` new ids: one allocation for a WLS sample set instead
    // of a doubling series (growth stays geometric across calls)
    void reserve_more(size_t more) {
        const size_t need = ids.size() + more;
        if (ids.capacity() < need) ids.reserve(std::max(need, 2 * ids.capacity()));
        if (w.capacity() < need) w.reserve(std::max(need, 2 * w.capacity()));
    }

  private:
    IdIndex index_;
};

struct WItem {
    double x[3];
    double dist;
    int bc;  // boundary index in the domain's Dirichlet list, -1 otherwise
    int id;
    int pos;  // place in the order candidates were accepted
    unsigned char type;
};

// Moving least squares (higtree/src/wls.c:15-124, 236-292, 495-532).
class Wls {
  public:
    Wls(int dim, int ord, int maxpts) : dim_(dim), ord_(ord), maxpts_(maxpts) {
        npoly_ = num_min_points(dim, ord);
        A_.assign((size_t)(maxpts + 1) * (npoly_ + 1), 0.0);
        Q_.assign((size_t)(maxpts + 1) * (maxpts + 1), 0.0);
        Dv_.assign(maxpts + 1, 0.0);
        b_.assign(maxpts + 1, 0.0);
        wv_.assign(maxpts + 1, 0.0);
    }
    static int num_min_points(int dim, int ord) {
        int n = 1;
        for (int i = 1; i <= ord; i++) {
            int nn = dim + i - 1, c = 1, j;
            for (j = nn; j > nn - i; j--) c *= j;
            for (j = 1; j <= i; j++) c /= j;
            n += c;
        }
        return n;
    }
    // weights of `npts` samples for the value at x
    void weights(int npts, const WItem *items, const double *x, double *out) {
        for (int i = 0; i < npts; i++) row(npts, &A(i, 0), items[i].x);
        for (int i = 0; i < npts; i++) {
            double m = 0.0;
            for (int j = 0; j < dim_; j++) m += (x[j] - items[i].x[j]) * (x[j] - items[i].x[j]);
            Dv_[i] = std::sqrt(1.0 / (m + EPSMACH));
            for (int j = 0; j < npoly_; j++) A(i, j) *= Dv_[i];
        }
        qr(npts);
        row(npts, b_.data(), x);
        solve(npts);
        for (int i = 0; i < npts; i++) out[i] = Dv_[i] * wv_[i];
    }
    int maxpts() const { return maxpts_; }

  private:
    int dim_, ord_, maxpts_, npoly_;
    std::vector<double> A_, Q_, Dv_, b_, wv_;
    double &A(int i, int j) { return A_[(size_t)i * (npoly_ + 1) + j]; }
    double &Q(int i, int j) { return Q_[(size_t)i * (maxpts_ + 1) + j]; }
    void row(int m, double *a, const double *x) {
        int k = 0;
        if (ord_ >= 0) {
            a[k++] = 1;
            if (k > m) return;
        }
        if (ord_ >= 1)
            for (int i1 = 0; i1 < dim_; i1++) {
                a[k++] = x[i1];
                if (k > m) return;
            }
        if (ord_ >= 2)
            for (int i1 = 0; i1 < dim_; i1++)
                for (int i2 = i1; i2 < dim_; i2++) {
                    a[k++] = x[i1] * x[i2];
                    if (k > m) return;
                }
        if (ord_ >= 3)
            for (int i1 = 0; i1 < dim_; i1++)
                for (int i2 = i1; i2 < dim_; i2++)
                    for (int i3 = i2; i3 < dim_; i3++) {
                        a[k++] = x[i1] * x[i2] * x[i3];
                        if (k > m) return;
                    }
    }
    // Rows i and k of Q have no entries beyond column k when qr() rotates
    // them (rotations run i-major, k ascending), so the Q update stops there;
    // the skipped entries are zero in both rows and stay zero.
    void rotate(int m, int n, int i, int k, double s1, double s2) {
        double s, t;
        if (std::fabs(s2) + std::fabs(s1) > 0.0) {
            if (std::fabs(s2) >= std::fabs(s1)) s = std::sqrt(1.0 + (s1 / s2) * (s1 / s2)) * std::fabs(s2);
            else s = std::sqrt(1.0 + (s2 / s1) * (s2 / s1)) * std::fabs(s1);
            s1 = s1 / s;
            s2 = s2 / s;
            for (int j = 0; j < n; j++) {
                s = s1 * A(i, j) + s2 * A(k, j);
                t = -s2 * A(i, j) + s1 * A(k, j);
                A(i, j) = s;
                A(k, j) = t;
            }
            const int qe = std::min(m, std::max(i, k) + 1);
            for (int j = 0; j < qe; j++) {
                s = s1 * Q(i, j) + s2 * Q(k, j);
                t = -s2 * Q(i, j) + s1 * Q(k, j);
                Q(i, j) = s;
                Q(k, j) = t;
            }
        }
    }
    void qr(int m) {
        for (int i = 0; i < m; i++) {
            for (int j = 0; j < m; j++) Q(i, j) = 0.0;
            Q(i, i) = 1.0;
        }
        for (int i = 0; i < npoly_; i++)
            for (int j = i + 1; j < m; j++) rotate(m, npoly_, i, j, A(i, i), A(j, i));
    }
    void solve(int m) {
        int i, j;
        double *x = wv_.data(), *y = b_.data();
        for (i = 0; i < npoly_; i++) {
            if (std::fabs(A(i, i)) > EPSMACH) {
                for (j = 0, x[i] = y[i]; j < i; j++) x[i] -= A(j, i) * x[j];
                x[i] = x[i] / A(i, i);
            } else {
                x[i] = 0.0;
            }
        }
        for (i = 0; i < npoly_; i++) y[i] = x[i];
        for (i = npoly_; i < m; i++) y[i] = 0.0;
        for (i = 0; i < m; i++)
            for (j = 0, x[i] = 0.0; j < m; j++) x[i] += Q(j, i) * y[j];
    }
};

}  // namespace hfh
