// Simulation domains (cell and facet), unknown-facet rule and the
// point-stencil resolver of higtree/src/domain.c:489-1623 and
// pdomain.c:1146-1221, with boundary values kept symbolic.
#pragma once
#include <array>

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
};

template <int D>
class Domain {
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
    std::unique_ptr<Wls> wls_, bcwls_;
    std::unordered_map<std::array<int64_t, 3>, std::pair<std::shared_ptr<CachedWls>, std::shared_ptr<CachedWls>>,
                       KeyHash>
        cache_;  // (with_bc, without_bc)
    std::vector<WItem> items_;
    double max_dist_ = 0;

    Wls &wls() {
        if (!wls_) wls_.reset(new Wls(D, order_, maxpts_));
        return *wls_;
    }
    Wls &bcwls() {
        if (!bcwls_) bcwls_.reset(new Wls(D - 1, 1, bc_maxpts_));
        return *bcwls_;
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

    WItem *new_item(int maxpts, double dist) {  // domain.c:388-412
        if (dist >= max_dist_) {
            if ((int)items_.size() >= maxpts) return nullptr;
            max_dist_ = dist;
        }
        items_.emplace_back();
        items_.back().dist = dist;
        return &items_.back();
    }

    void search_tree_cells(const Tree<D> &t, const IdMap &m, const Vec<D> &x, const Vec<D> &lo,
                           const Vec<D> &hi, int maxpts, int bcidx, unsigned char type) {
        t.bbox_cells(lo, hi, [&](int c) {
            int id = m.lookup(t.n[c].cid);
            if (id < 0) return;
            Vec<D> cc = t.center(c);
            double dist = distance<D>(x, cc);
            WItem *it = new_item(maxpts, dist);
            if (it) {
                it->id = id;
                it->bc = bcidx;
                it->type = type;
                for (int d = 0; d < D; d++) it->x[d] = cc[d];
            }
        });
    }

    void search_box(const Vec<D> &x, const Vec<D> &lo, const Vec<D> &hi) {
        if (fdim < 0) {
            for (auto *t : trees) search_tree_cells(*t, map, x, lo, hi, maxpts_, -1, 0);
            return;
        }
        for (auto *t : trees) {
            t->bbox_facets(interest, lo, hi, [&](const Facet<D> &f) {
                int id = map.lookup(t->n[f.cell].fid[f.dim * 2 + f.dir]);
                if (id < 0) return;
                Vec<D> fc = t->facet_center(f);
                double dist = distance<D>(x, fc);
                WItem *it = new_item(maxpts_, dist);
                if (it) {
                    it->id = id;
                    it->bc = -1;
                    it->type = 0;
                    for (int d = 0; d < D; d++) it->x[d] = fc[d];
                }
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
        if (point_equal<D>(b->tree->center(bcell), px)) {
            out.emplace_back(b->base + b->map.lookup(b->tree->n[bcell].cid), 1.0);
            return out;
        }
        items_.clear();
        max_dist_ = 0.0;
        Vec<D> lo, hi;
        for (int i = 0; i < D; i++) {
            lo[i] = boxc[i] - 3.0 * bd;
            hi[i] = boxc[i] + 3.0 * bd;
        }
        search_tree_cells(*b->tree, b->map, px, lo, hi, bc_maxpts_, -1, 2);
        sort_items();
        if ((int)items_.size() > bc_maxpts_) items_.resize(bc_maxpts_);
        for (auto &it : items_)
            for (int d = b->pdir + 1; d < D; d++) it.x[d - 1] = it.x[d];
        std::vector<double> w(items_.size() + 1);
        bcwls().weights((int)items_.size(), items_.data(), px.v, w.data());
        for (size_t i = 0; i < items_.size(); i++) out.emplace_back(b->base + items_[i].id, w[i]);
        return out;
    }

    void sort_items() {
        if (items_.size() <= 1) return;
        std::vector<WItem> tmp(items_.size());
        merge_sort_items(items_.data(), (int)items_.size(), tmp.data());
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
        std::fprintf(stderr, "hfh: boundary projection outside the domain\n");
        std::abort();
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

    void interpolate(const Vec<D> &x, double delta, double alpha, Stencil &stn, bool use_dir) {
        auto key = point_key<D>(x);
        auto found = use_cache ? cache_.find(key) : cache_.end();
        std::shared_ptr<CachedWls> cw;
        if (found != cache_.end()) cw = use_dir ? found->second.first : found->second.second;
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
            items_.clear();
            max_dist_ = 0.0;
            Vec<D> lo, hi;
            for (int k = 0; k < D; k++) {
                lo[k] = x[k] - 10.0 * delta;
                hi[k] = x[k] + 10.0 * delta;
            }
            search_box(x, lo, hi);
            if (use_dir)
                for (size_t i = 0; i < dir.size(); i++)
                    search_tree_cells(*dir[i]->tree, dir[i]->map, x, lo, hi, maxpts_, (int)i, 1);
            sort_items();
            if ((int)items_.size() > maxpts_) items_.resize(maxpts_);
            std::vector<double> w(items_.size() + 1);
            wls().weights((int)items_.size(), items_.data(), x.v, w.data());
            // cached form drops |w| < EPSMACH (domain.c:619-648); rows are
            // built once and reused every step, so the cached form is used
            cw = std::make_shared<CachedWls>();
            for (size_t i = 0; i < items_.size(); i++)
                if (!flt_eq(w[i], 0.0) && items_[i].type == 0) {
                    cw->ids.push_back(items_[i].id);
                    cw->w.push_back(w[i]);
                    cw->numpts++;
                }
            for (size_t i = 0; i < items_.size(); i++)
                if (!flt_eq(w[i], 0.0) && items_[i].type == 1) {
                    cw->ids.push_back(items_[i].id);
                    cw->w.push_back(w[i]);
                    cw->bcidx.push_back(items_[i].bc);
                    cw->numbcpts++;
                }
            if (use_cache) {
                auto &e = cache_[key];
                if (found != cache_.end()) {
                    if (use_dir) e.first = cw;
                    else e.second = cw;
                } else if (use_dir) {
                    e.first = cw;
                    e.second = cw->numbcpts == 0 ? cw : nullptr;
                } else {
                    e.first = nullptr;
                    e.second = cw;
                }
            }
        }
        for (int i = 0; i < cw->numpts; i++) stn.add(cw->ids[i], alpha * cw->w[i]);
        for (int i = 0; i < cw->numbcpts; i++) {
            Boundary<D> *b = dir[cw->bcidx[i]];
            stn.add_rhs_bc(b->base + cw->ids[cw->numpts + i], -alpha * cw->w[cw->numpts + i]);
        }
    }

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
