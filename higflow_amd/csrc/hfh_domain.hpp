// Simulation domains (cell and facet), unknown-facet rule and the
// point-stencil resolver of higtree/src/domain.c:489-1623 and
// pdomain.c:1146-1221, with boundary values kept symbolic.
#pragma once
#include <array>
#include <stdexcept>
#include <atomic>
#include <limits>
#include <map>
#include <thread>

#include "hfh_mesh.hpp"

namespace hfh {

struct KeyHash {
    size_t operator()(const std::array<int64_t, 3> &k) const {
        uint64_t h = 1469598103934665603ull;
        for (int d = 0; d < 3; d++) {
            h ^= (uint64_t)k[d];
            h *= 1099511628211ull;
        }
        return (size_t)(h ^ (h >> 29));
    }
};

// higtree/src/point-mapper.c:3-8
template <int D>
inline std::array<int64_t, 3> point_key(const Vec<D> &x) {
    std::array<int64_t, 3> k{0, 0, 0};
    for (int d = 0; d < D; d++) k[d] = (int64_t)(x[d] / EPSMACH + 0.5);
    return k;
}

struct CachedWls {
    int numpts = 0, numbcpts = 0;
    std::vector<int> ids;     // numpts ids then numbcpts boundary-local ids
    std::vector<double> w;
    std::vector<int> bcidx;   // Dirichlet-list index of each boundary sample
    // request of a deferred computation (first call's point and search size)
    bool pending = false;
    double x[3] = {0, 0, 0};
    double delta = 0;
    bool use_dir = false;
};

// Per-thread WLS work space.  `items` holds the kept candidates (a heap
// while a search runs), `count` the candidates accepted so far.
struct Scratch {
    std::vector<WItem> items;
    double max_dist = 0;
    int count = 0;
    std::unique_ptr<Wls> wls, bcwls;
};

// Threads for the setup's WLS computations: HFH_SETUP_THREADS, else the
// hardware threads (at most 16).
inline int setup_threads() {
    if (const char *e = std::getenv("HFH_SETUP_THREADS")) {
        int n = std::atoi(e);
        if (n >= 1) return n;
    }
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h));
}

// A point request of a parallel first set-up pass (Domain::begin_logging).
template <int D>
struct PointReq {
    Vec<D> x;
    double delta;
    bool use_dir;
};

template <int D>
class Domain {
  public:
    using Req = PointReq<D>;

  public:
    // trees are shared with the other domains of the solver
    std::vector<const Tree<D> *> trees;
    std::vector<Boundary<D> *> dir, neu;
    int fdim = -1;  // facet dimension, -1 for a cell domain
    int interest[D] = {0};
    IdMap map;  // cell id (cell domain) or facet id (facet domain) -> local id
    std::vector<std::pair<int, int>> cells;  // local id -> (tree, node)
    std::vector<std::pair<int, Facet<D>>> facets;  // local id -> (tree, facet)
    std::vector<std::vector<unsigned char>> sfbi_l, sfbi_h;
    bool use_cache = true;

    Domain(int order, int facet_dim) : fdim(facet_dim) {
        order_ = order;
        maxpts_ = calc_num_points(D, order);
        bc_maxpts_ = calc_num_points(D - 1, 1);
        if (fdim >= 0) interest[fdim] = 1;
    }

    int size() const { return fdim < 0 ? (int)cells.size() : (int)facets.size(); }
    int maxpts() const { return maxpts_; }

    static int calc_num_points(int dim, int order) {  // domain.c:102-118
        if (dim == 2) {
            if (order == 3) return 60;
            if (order == 2) return 45;
        } else if (dim == 3) {
            if (order == 3) return 150;
            if (order == 2) return 120;
        }
        return 2 * dim * Wls::num_min_points(dim, order);
    }

    Vec<D> center_of(int lid) const {
        if (fdim < 0) return trees[cells[lid].first]->center(cells[lid].second);
        return trees[facets[lid].first]->facet_center(facets[lid].second);
    }
    // lowpoint and highpoint of cell lid (cell domains)
    void bounds_of(int lid, double *lo, double *hi) const {
        const auto &nd = trees[cells[lid].first]->n[cells[lid].second];
        for (int i = 0; i < D; i++) {
            lo[i] = nd.lo[i];
            hi[i] = nd.hi[i];
        }
    }
    Vec<D> delta_of(int lid) const {
        if (fdim < 0) return trees[cells[lid].first]->delta(cells[lid].second);
        return trees[facets[lid].first]->delta(facets[lid].second.cell);
    }

    // pdomain.c:154-175
    void build_cell_map(int maxid) {
        map.init(maxid);
        for (size_t t = 0; t < trees.size(); t++)
            trees[t]->leaves([&](int c) { map.assign(trees[t]->n[c].cid); });
        cells.assign(map.count, {0, 0});
        for (size_t t = 0; t < trees.size(); t++)
            trees[t]->leaves([&](int c) { cells[map.lookup(trees[t]->n[c].cid)] = {(int)t, c}; });
    }

    // domain.c:1440-1521 (mutates the shared trees' facet ids)
    static void adjust_facet_ids(std::vector<Tree<D> *> &ts, int dim) {
        for (size_t i = 0; i < ts.size(); i++) {
            Tree<D> &t = *ts[i];
            std::vector<int> leaves;
            t.leaves([&](int c) { leaves.push_back(c); });
            for (int c : leaves) {
                Vec<D> ofp = t.center(c);
                for (int dir = 0; dir < 2; dir++) {
                    ofp[dim] = dir == 0 ? t.n[c].lo[dim] - EPSDELTA : t.n[c].hi[dim] + EPSDELTA;
                    bool different = false, smaller = false;
                    const Tree<D> *ot = &t;
                    int ofc = t.cell_with_point(ofp);
                    for (size_t j = 0; j < ts.size() && ofc < 0; j++) {
                        if (j != i) {
                            ofc = ts[j]->cell_with_point(ofp);
                            ot = ts[j];
                        }
                    }
                    if (ofc >= 0) {
                        for (int k = 0; k < D; k++) {
                            if (k == dim) continue;
                            if (!flt_eq(t.n[c].lo[k], ot->n[ofc].lo[k]) ||
                                !flt_eq(t.n[c].hi[k], ot->n[ofc].hi[k])) {
                                different = true;
                                if (flt_gt(t.n[c].hi[k] - t.n[c].lo[k], ot->n[ofc].hi[k] - ot->n[ofc].lo[k]))
                                    smaller = true;
                            }
                        }
                    }
                    if (different) {
                        if (smaller) t.n[c].fid[dim * 2 + dir] = 0;
                    } else if (ofc >= 0) {
                        if (dir == 0) t.n[c].fid[dim * 2 + dir] = -ot->n[ofc].fid[dim * 2 + 1];
                    }
                }
            }
        }
    }

    // pdomain.c:1146-1221: computed before any boundary exists
    void compute_sfbi() {
        sfbi_l.assign(trees.size(), {});
        sfbi_h.assign(trees.size(), {});
        for (size_t t = 0; t < trees.size(); t++) {
            const Tree<D> &tr = *trees[t];
            int nb = tr.nchildren(0);
            sfbi_l[t].assign(nb, 0);
            sfbi_h[t].assign(nb, 0);
            for (int b = 0; b < nb; b++) {
                int pp[D];
                tobase<D>(b, tr.n[0].nc, pp);
                Vec<D> pl = block_facet_center(tr, b, 0), ph = block_facet_center(tr, b, 1);
                sfbi_l[t][b] = in_neumann(pl);
                bool flow = false;
                for (size_t t2 = 0; t2 < trees.size(); t2++)
                    if (t2 != t && trees[t2]->cell_with_point(ph) >= 0) flow = true;
                sfbi_h[t][b] = pp[fdim] < tr.n[0].nc[fdim] - 1 || in_neumann(ph) || flow;
            }
        }
    }

    // domain.c:1834-1894: per block, facets inside the block box
    template <class F>
    void domain_facets(F &&fn) const {
        for (size_t t = 0; t < trees.size(); t++) {
            const Tree<D> &tr = *trees[t];
            int nb = tr.nchildren(0);
            for (int b = 0; b < nb; b++) {
                const Node<D> &c = tr.n[tr.n[0].child0 + b];
                Vec<D> l, h;
                for (int i = 0; i < D; i++) {
                    l[i] = c.lo[i] + EPSDELTA;
                    h[i] = c.hi[i] - EPSDELTA;
                }
                if (sfbi_l[t][b]) l[fdim] -= 2.0 * EPSDELTA;
                if (sfbi_h[t][b]) h[fdim] += 2.0 * EPSDELTA;
                tr.bbox_facets(interest, l, h, [&](const Facet<D> &f) { fn((int)t, f); });
            }
        }
    }

    void build_facet_map(int maxid) {  // pdomain.c:780-797
        map.init(maxid);
        domain_facets([&](int t, const Facet<D> &f) {
            map.assign(trees[t]->n[f.cell].fid[f.dim * 2 + f.dir]);
        });
        facets.assign(map.count, {0, Facet<D>{0, 0, 0}});
        domain_facets([&](int t, const Facet<D> &f) {
            facets[map.lookup(trees[t]->n[f.cell].fid[f.dim * 2 + f.dir])] = {t, f};
        });
    }

    int cell_with_point(const Vec<D> &x, int *tree = nullptr) const {
        for (size_t t = 0; t < trees.size(); t++) {
            int c = trees[t]->cell_with_point(x);
            if (c >= 0) {
                if (tree) *tree = (int)t;
                return c;
            }
        }
        return -1;
    }

    bool facet_with_point(const Vec<D> &x) const {  // domain.c:1927-1938
        Facet<D> f;
        for (auto *t : trees)
            if (t->facet_with_point(fdim, x, f)) return true;
        return false;
    }

    // sd_get_stencil (domain.c:1429-1433) / sfd_get_stencil (:1604-1623)
    void get_stencil(const Vec<D> &center, const Vec<D> &x, double alpha, Stencil &stn) {
        double delta = distance<D>(center, x);
        if (fdim >= 0 && flt_eq(delta, 0.0)) {
            for (auto *t : trees) {
                int c = t->cell_with_point(x);
                if (c >= 0) {
                    delta = t->delta(c)[0];
                    break;
                }
            }
        }
        resolve(delta, x, alpha, stn, true, true);
    }

  private:
    int order_, maxpts_, bc_maxpts_;
    struct Entry {
        std::shared_ptr<CachedWls> with_bc, without_bc;
        bool undecided = false;  // without_bc follows with_bc once its samples are known
    };
    std::unordered_map<std::array<int64_t, 3>, Entry, KeyHash> cache_;
    Scratch sc_;
    // Concurrent reads (the parallel row-recording pass): the point cache is
    // read-only, each thread fits boundary values in scratch of its own,
    // and a request the cache cannot answer is only flagged.
    bool concurrent_ = false;
    bool logging_ = false;
    static std::vector<Req> &req_log(const Domain *d) {
        thread_local std::map<const Domain *, std::vector<Req>> logs;
        return logs[d];
    }
    std::atomic<bool> missed_{false};
    Scratch &scratch() {
        if (!concurrent_) return sc_;
        thread_local std::map<std::pair<int, int>, Scratch> pool;  // per (order, maxpts)
        return pool[{order_, maxpts_}];
    }
    bool defer_ = false;
    std::vector<std::shared_ptr<CachedWls>> pending_;

    Wls &wls(Scratch &sc) const {
        if (!sc.wls) sc.wls.reset(new Wls(D, order_, maxpts_));
        return *sc.wls;
    }
    Wls &bcwls(Scratch &sc) const {
        if (!sc.bcwls) sc.bcwls.reset(new Wls(D - 1, 1, bc_maxpts_));
        return *sc.bcwls;
    }

    Vec<D> block_facet_center(const Tree<D> &tr, int b, int dir) const {
        int c = tr.n[0].child0 + b;
        Vec<D> p = tr.center(c);
        p[fdim] = dir == 0 ? tr.n[c].lo[fdim] : tr.n[c].hi[fdim];
        return p;
    }
    bool in_neumann(const Vec<D> &p) const {
        for (auto *b : neu)
            if (b->tree->cell_with_point(p) >= 0) return true;
        return false;
    }

    // Candidate lists (domain.c:373-412).  The reference appends a candidate
    // unless the list already holds maxpts and the candidate is no nearer
    // than the farthest one taken while the list filled, sorts the list with
    // qsort on distance (glibc's merge sort: of two equal distances the later
    // candidate comes first) and keeps the first maxpts.  Only those are ever
    // read, so a heap of the maxpts first candidates under that same order
    // (distance, then later first) is kept instead of the whole list: same
    // points, same order.
    static bool before(const WItem &a, const WItem &b) {
        return a.dist < b.dist || (a.dist == b.dist && a.pos > b.pos);
    }
    static void begin_list(Scratch &sc) {
        sc.items.clear();
        sc.max_dist = 0.0;
        sc.count = 0;
    }
    // the reference's append test, without its side effect
    static bool rejected(const Scratch &sc, int maxpts, double dist) {
        return dist >= sc.max_dist && sc.count >= maxpts;
    }
    static void accept(Scratch &sc, int maxpts, WItem &it) {
        if (it.dist >= sc.max_dist) sc.max_dist = it.dist;
        it.pos = sc.count++;
        if ((int)sc.items.size() < maxpts) {
            sc.items.push_back(it);
            std::push_heap(sc.items.begin(), sc.items.end(), before);
        } else if (it.dist <= sc.items.front().dist) {  // a later equal distance ranks first
            std::pop_heap(sc.items.begin(), sc.items.end(), before);
            sc.items.back() = it;
            std::push_heap(sc.items.begin(), sc.items.end(), before);
        }
    }
    static void end_list(Scratch &sc) { std::sort(sc.items.begin(), sc.items.end(), before); }
    // Once maxpts candidates are kept, one farther than the farthest kept
    // cannot enter: squared bound for the tree walk, with a relative margin
    // far above the rounding of the distances (boxes at the bound are walked).
    static double far2(const Scratch &sc, int maxpts) {
        if (sc.count < maxpts) return std::numeric_limits<double>::infinity();
        const double r = sc.items.front().dist;
        return r * r * (1.0 + 1e-9);
    }

    static void search_tree_cells(Scratch &sc, const Tree<D> &t, const IdMap &m, const Vec<D> &x,
                                  const Vec<D> &lo, const Vec<D> &hi, int maxpts, int bcidx, unsigned char type) {
        t.bbox_cells_near(lo, hi, x, [&]() { return far2(sc, maxpts); }, [&](int c) {
            Vec<D> cc = t.center(c);
            double dist = distance<D>(x, cc);
            if (rejected(sc, maxpts, dist)) return;
            int id = m.lookup(t.n[c].cid);
            if (id < 0) return;
            WItem it;
            it.dist = dist;
            it.id = id;
            it.bc = bcidx;
            it.type = type;
            for (int d = 0; d < D; d++) it.x[d] = cc[d];
            accept(sc, maxpts, it);
        });
    }

    void search_box(Scratch &sc, const Vec<D> &x, const Vec<D> &lo, const Vec<D> &hi) const {
        if (fdim < 0) {
            for (auto *t : trees) search_tree_cells(sc, *t, map, x, lo, hi, maxpts_, -1, 0);
            return;
        }
        for (auto *t : trees) {
            t->bbox_facets_near(interest, lo, hi, x, [&]() { return far2(sc, maxpts_); }, [&](const Facet<D> &f) {
                Vec<D> fc = t->facet_center(f);
                double dist = distance<D>(x, fc);
                if (rejected(sc, maxpts_, dist)) return;
                int id = map.lookup(t->n[f.cell].fid[f.dim * 2 + f.dir]);
                if (id < 0) return;
                WItem it;
                it.dist = dist;
                it.id = id;
                it.bc = -1;
                it.type = 0;
                for (int d = 0; d < D; d++) it.x[d] = fc[d];
                accept(sc, maxpts_, it);
            });
        }
    }

    bool find_in_center(const Vec<D> &x, double alpha, bool &in_domain, Stencil &stn) {
        in_domain = false;
        for (auto *t : trees) {
            if (!rect_contains<D>(t->n[0].lo, t->n[0].hi, x)) continue;
            in_domain = true;
            if (fdim < 0) {
                int c = t->cell_with_point(x);
                if (c >= 0 && point_equal<D>(t->center(c), x)) {
                    stn.add(map.lookup(t->n[c].cid), alpha);
                    return true;
                }
                break;
            } else {
                Facet<D> f;
                if (t->facet_with_point(fdim, x, f) && point_equal<D>(t->facet_center(f), x)) {
                    int id = map.lookup(t->n[f.cell].fid[f.dim * 2 + f.dir]);
                    if (id >= 0) {
                        stn.add(id, alpha);
                        return true;
                    }
                }
            }
        }
        return false;
    }

    bool dirichlet_center(const Vec<D> &x, double alpha, Stencil &stn) {  // domain.c:1298-1319
        for (auto *b : dir) {
            int c = b->tree->cell_with_point(x);
            if (c >= 0 && point_equal<D>(b->tree->center(c), x)) {
                stn.add_rhs_bc(b->base + b->map.lookup(b->tree->n[c].cid), -alpha);
                return true;
            }
        }
        return false;
    }

    // boundary value at px as (slot, weight) pairs: exact at a boundary-cell
    // centre, else the DIM-1 linear WLS fit sampled around boxc
    std::vector<std::pair<int, double>> boundary_value(Boundary<D> *b, int bcell, const Vec<D> &px,
                                                        const Vec<D> &boxc, double bd) {
        std::vector<std::pair<int, double>> out;
        // first pass of a deferred build: its rows are discarded and no
        // request depends on boundary terms, so the fit is skipped
        if (defer_) return out;
        if (point_equal<D>(b->tree->center(bcell), px)) {
            out.emplace_back(b->base + b->map.lookup(b->tree->n[bcell].cid), 1.0);
            return out;
        }
        Scratch &sc = scratch();
        begin_list(sc);
        Vec<D> lo, hi;
        for (int i = 0; i < D; i++) {
            lo[i] = boxc[i] - 3.0 * bd;
            hi[i] = boxc[i] + 3.0 * bd;
        }
        search_tree_cells(sc, *b->tree, b->map, px, lo, hi, bc_maxpts_, -1, 2);
        end_list(sc);
        for (auto &it : sc.items)
            for (int d = b->pdir + 1; d < D; d++) it.x[d - 1] = it.x[d];
        std::vector<double> w(sc.items.size() + 1);
        bcwls(sc).weights((int)sc.items.size(), sc.items.data(), px.v, w.data());
        for (size_t i = 0; i < sc.items.size(); i++) out.emplace_back(b->base + sc.items[i].id, w[i]);
        return out;
    }

    struct Proj {
        Boundary<D> *b = nullptr;
        int cell = -1;
        Vec<D> px;
        double dist = 1e300;
        int pd = -1;
    };

    Proj project(const std::vector<Boundary<D> *> &list, const Vec<D> &x) const {
        Proj best;
        best.dist = 1.7976931348623157e308;
        for (auto *b : list) {
            int pd = b->pdir;
            Vec<D> tc = b->tree->center(0);
            Vec<D> px = x;
            px[pd] = tc[pd];
            int c = b->tree->cell_with_point(px);
            if (c < 0) continue;
            double dist = x[pd] - px[pd];
            if (std::fabs(dist) < std::fabs(best.dist)) {
                best.b = b;
                best.cell = c;
                best.px = px;
                best.dist = dist;
                best.pd = pd;
            }
        }
        return best;
    }

    double cell_delta_at(const Vec<D> &at, int pd) const {
        for (auto *t : trees) {
            int c = t->cell_with_point(at);
            if (c >= 0) return t->delta(c)[pd];
        }
        // a Dirichlet boundary whose projection of a ghost point finds no
        // domain cell: malformed boundary trees.  Set-up reports it through
        // the entry point's error (hfh_last_error) instead of ending the
        // caller's process.
        throw std::runtime_error("hfh: boundary projection outside the domain (boundary tree does not border a "
                                 "domain cell)");
    }

    double side_delta(Boundary<D> *b, int bcell, int pd) const {
        Vec<D> bdl = b->tree->delta(bcell);
        double bd = 0.0;
        for (int k = 0; k < D; k++)
            if (k != pd) bd = rmax(bd, bdl[k]);
        return bd;
    }

    bool neumann_outside(const Vec<D> &x, double alpha, Stencil &stn) {  // domain.c:692-870
        Proj best = project(neu, x);
        if (!best.b) return false;
        double delta = std::fabs(best.dist);
        for (auto *b : dir) {
            const Node<D> &r = b->tree->n[0];
            double acc = 0;
            for (int k = 0; k < D; k++) {
                const double dd = rmax(rmax(r.lo[k] - x[k], 0.0), x[k] - r.hi[k]);
                acc += dd * dd;
            }
            if (std::sqrt(acc) < delta) return false;
        }
        double mh = std::copysign(cell_delta_at(best.px, best.pd), best.dist) * -0.5;
        if (fdim >= 0 && fdim == best.pd) mh *= 2;
        Vec<D> inside = best.px;
        inside[best.pd] += mh;
        double bd = side_delta(best.b, best.cell, best.pd);
        resolve(bd, inside, alpha, stn, true, false);
        for (auto &sw : boundary_value(best.b, best.cell, best.px, x, bd))
            stn.add_rhs_bc(sw.first, alpha * sw.second * (mh + best.dist));
        return true;
    }

    bool dirichlet_outside(const Vec<D> &x, double alpha, Stencil &stn) {  // domain.c:877-1029
        Proj best = project(dir, x);
        if (!best.b) return false;
        double mh = std::copysign(cell_delta_at(best.px, best.pd), best.dist) * -0.5;
        if (fdim >= 0 && fdim == best.pd) mh *= 2;
        Vec<D> inside = best.px;
        inside[best.pd] += mh;
        double ratio = (mh - best.dist) / (mh);
        double bd = side_delta(best.b, best.cell, best.pd);
        resolve(bd, inside, (1.0 - ratio) * alpha, stn, false, false);
        for (auto &sw : boundary_value(best.b, best.cell, best.px, x, bd))
            stn.add_rhs_bc(sw.first, alpha * sw.second * (-ratio));
        return true;
    }

    bool neumann_on_boundary(const Vec<D> &x, double alpha, Stencil &stn) {  // domain.c:1037-1192
        Boundary<D> *cb = nullptr;
        int cc = -1;
        for (auto *b : neu) {
            int c = b->tree->cell_with_point(x);
            if (c >= 0) {
                cb = b;
                cc = c;
            }
        }
        if (!cb) return false;
        int pd = cb->pdir;
        int out_sign = 1;
        Vec<D> xl = x, xr = x;
        xl[pd] -= EPSDELTA + EPSMACH;
        xr[pd] += EPSDELTA + EPSMACH;
        for (auto *t : trees) {
            if (rect_contains<D>(t->n[0].lo, t->n[0].hi, xl)) {
                out_sign = 1;
                break;
            }
            if (rect_contains<D>(t->n[0].lo, t->n[0].hi, xr)) {
                out_sign = -1;
                break;
            }
        }
        double mh = std::copysign(cell_delta_at(x, pd), (double)out_sign) * -0.5;
        if (fdim >= 0 && fdim == pd) mh *= 2;
        Vec<D> inside = x;
        inside[pd] += mh;
        double bd = side_delta(cb, cc, pd);
        resolve(bd, inside, alpha, stn, false, false);
        for (auto &sw : boundary_value(cb, cc, x, x, bd)) stn.add_rhs_bc(sw.first, alpha * sw.second * mh);
        return true;
    }

    bool dirichlet_on_boundary(const Vec<D> &x, double alpha, Stencil &stn) {  // domain.c:1200-1294
        Boundary<D> *cb = nullptr;
        int cc = -1;
        for (auto *b : dir) {
            int c = b->tree->cell_with_point(x);
            if (c >= 0) {
                cb = b;
                cc = c;
            }
        }
        if (!cb) return false;
        double bd = side_delta(cb, cc, cb->pdir);
        for (auto &sw : boundary_value(cb, cc, x, x, bd)) stn.add_rhs_bc(sw.first, alpha * sw.second * (-1.0));
        return true;
    }

    // domain.c:527-599: samples in the box x +- 10 delta (unknowns, then the
    // Dirichlet boundary centres), pruned, sorted, clipped, WLS weights; the
    // cached form drops |w| < EPSMACH (domain.c:619-648).  Reads only the
    // domain geometry, so deferred requests run on worker threads.
    void compute_wls(CachedWls &cw, Scratch &sc) const {
        Vec<D> x;
        for (int k = 0; k < D; k++) x[k] = cw.x[k];
        const double delta = cw.delta;
        begin_list(sc);
        Vec<D> lo, hi;
        for (int k = 0; k < D; k++) {
            lo[k] = x[k] - 10.0 * delta;
            hi[k] = x[k] + 10.0 * delta;
        }
        search_box(sc, x, lo, hi);
        if (cw.use_dir)
            for (size_t i = 0; i < dir.size(); i++)
                search_tree_cells(sc, *dir[i]->tree, dir[i]->map, x, lo, hi, maxpts_, (int)i, 1);
        end_list(sc);
        std::vector<double> w(sc.items.size() + 1);
        wls(sc).weights((int)sc.items.size(), sc.items.data(), x.v, w.data());
        cw.ids.clear();
        cw.w.clear();
        cw.bcidx.clear();
        cw.numpts = cw.numbcpts = 0;
        for (size_t i = 0; i < sc.items.size(); i++)
            if (!flt_eq(w[i], 0.0) && sc.items[i].type == 0) {
                cw.ids.push_back(sc.items[i].id);
                cw.w.push_back(w[i]);
                cw.numpts++;
            }
        for (size_t i = 0; i < sc.items.size(); i++)
            if (!flt_eq(w[i], 0.0) && sc.items[i].type == 1) {
                cw.ids.push_back(sc.items[i].id);
                cw.w.push_back(w[i]);
                cw.bcidx.push_back(sc.items[i].bc);
                cw.numbcpts++;
            }
        cw.pending = false;
    }

    void interpolate(const Vec<D> &x, double delta, double alpha, Stencil &stn, bool use_dir) {
        if (logging_) {  // parallel first pass: the request is replayed in order later
            req_log(this).push_back(Req{x, delta, use_dir});
            return;
        }
        auto key = point_key<D>(x);
        Entry *e = nullptr;
        if (use_cache) {
            auto it = cache_.find(key);
            if (it != cache_.end()) e = &it->second;
        }
        std::shared_ptr<CachedWls> cw;
        if (e) {
            if (!use_dir && e->undecided) {
                if (concurrent_) {
                    missed_ = true;
                    return;
                }
                if (e->with_bc->pending) compute_wls(*e->with_bc, sc_);
                if (e->with_bc->numbcpts == 0) e->without_bc = e->with_bc;
                e->undecided = false;
            }
            cw = use_dir ? e->with_bc : e->without_bc;
        }
        if (!cw && concurrent_) {
            missed_ = true;
            return;
        }
        if (!cw) {
            if (flt_eq(delta, 0.0)) {
                bool got = false;
                for (auto *t : trees) {
                    int c = t->cell_with_point(x);
                    if (c >= 0) {
                        Vec<D> dl = t->delta(c);
                        delta = dl[0];
                        for (int k = 1; k < D; k++) delta *= dl[k];
                        delta = std::pow(delta, 1.0 / D);
                        got = true;
                        break;
                    }
                }
                if (!got || flt_eq(delta, 0.0)) {
                    const Tree<D> &t0 = *trees[0];
                    int c = 0;
                    while (t0.n[c].child0 >= 0) c = t0.n[c].child0;
                    Vec<D> dl = t0.delta(c);
                    delta = dl[0];
                    for (int k = 1; k < D; k++) delta *= dl[k];
                    delta = std::pow(delta, 1.0 / D);
                }
            }
            cw = std::make_shared<CachedWls>();
            for (int k = 0; k < D; k++) cw->x[k] = x[k];
            cw->delta = delta;
            cw->use_dir = use_dir;
            if (defer_ && use_cache) {
                cw->pending = true;
                pending_.push_back(cw);
            } else {
                compute_wls(*cw, sc_);
            }
            if (use_cache) {
                Entry &n = cache_[key];
                if (e) {
                    if (use_dir) n.with_bc = cw;
                    else n.without_bc = cw;
                } else if (use_dir) {
                    n.with_bc = cw;
                    if (cw->pending) n.undecided = true;
                    else n.without_bc = cw->numbcpts == 0 ? cw : nullptr;
                } else {
                    n.without_bc = cw;
                }
            }
        }
        if (cw->pending) return;  // first pass of a deferred build: contents follow
        stn.reserve_more(cw->numpts);
        for (int i = 0; i < cw->numpts; i++) stn.add(cw->ids[i], alpha * cw->w[i]);
        for (int i = 0; i < cw->numbcpts; i++) {
            Boundary<D> *b = dir[cw->bcidx[i]];
            stn.add_rhs_bc(b->base + cw->ids[cw->numpts + i], -alpha * cw->w[cw->numpts + i]);
        }
    }

  public:
    // Deferred build: while set, WLS requests are only recorded (first call
    // wins, as with immediate computation); resolve_pending() computes them on
    // `nthreads` threads and settles the with/without-boundary sharing.
    void begin_deferred() { defer_ = true; }
    // Parallel first pass: each thread only logs its point requests (with
    // the delta and boundary choice of the call); replaying the logs in the
    // serial walk's order gives the cache the serial first pass builds.
    void begin_logging() { logging_ = true; }
    void end_logging() { logging_ = false; }
    // this thread's log for this domain, moved out
    std::vector<Req> take_log() { return std::move(req_log(this)); }
    void replay(const std::vector<Req> &log) {
        Stencil s;
        for (const Req &q : log) {
            s.reset();
            interpolate(q.x, q.delta, 1.0, s, q.use_dir);
        }
    }
    void begin_concurrent() {
        concurrent_ = true;
        missed_ = false;
    }
    // false when a request missed the cache during the concurrent pass
    bool end_concurrent() {
        concurrent_ = false;
        return !missed_;
    }
    void resolve_pending(int nthreads) {
        std::vector<CachedWls *> work;
        for (auto &p : pending_)
            if (p->pending) work.push_back(p.get());
        std::atomic<size_t> next{0};
        auto run = [&]() {
            Scratch sc;
            for (size_t i; (i = next.fetch_add(1)) < work.size();) {
                compute_wls(*work[i], sc);
                if ((i & 4095) == 0) progress("least-squares fits", (long)i, (long)work.size());
            }
        };
        int nt = (int)std::min<size_t>((size_t)std::max(1, nthreads), work.size());
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(run);
        run();
        for (auto &t : th) t.join();
        for (auto &kv : cache_) {
            Entry &e = kv.second;
            if (e.undecided) {
                if (e.with_bc->numbcpts == 0) e.without_bc = e.with_bc;
                e.undecided = false;
            }
        }
        pending_.clear();
        defer_ = false;
    }
    size_t cache_size() const { return cache_.size(); }

  public:
    // domain.c:1321-1379
    void resolve(double delta, const Vec<D> &x, double alpha, Stencil &stn, bool use_dir, bool use_neu) {
        bool in_domain;
        if (find_in_center(x, alpha, in_domain, stn)) return;
        if (in_domain) {
            if (dirichlet_center(x, alpha, stn)) return;
            if (use_neu && neumann_on_boundary(x, alpha, stn)) return;
            if (use_dir && dirichlet_on_boundary(x, alpha, stn)) return;
        } else {
            if (use_neu && neumann_outside(x, alpha, stn)) return;
            if (use_dir && dirichlet_outside(x, alpha, stn)) return;
        }
        interpolate(x, delta, alpha, stn, use_dir);
    }
};

}  // namespace hfh
