/*
 * CPU oracle (test infrastructure): HiG-Tree restatement.
 * Tree cells, point location, bounding-box and facet iterators and the
 * `.amr` reader, following higtree/src/higtree.c, higtree-iterator.c and
 * higtree-io.c.  See hfo_internal.h for the oracle's role and pinning status.
 */
#include "hfo_internal.h"

/* higtree/src/higtree.c:23-30 -- last dimension varies fastest */
int hfo_toposition(const int nc[DIM], const int pp[DIM])
{
    int res = 0;
    for (int i = 0; i < DIM; i++) {
        res *= nc[i];
        res += pp[i];
    }
    return res;
}

/* higtree/src/higtree.c:32-37 */
void hfo_tobase(int pos, const int nc[DIM], int pp[DIM])
{
    for (int i = DIM - 1; i >= 0; i--) {
        pp[i] = pos % nc[i];
        pos /= nc[i];
    }
}

static int next_id(idgen *g) { return g->next++; }

static hcell *new_cell(idgen *g, const Point lo, const Point hi)
{
    hcell *c = calloc(1, sizeof *c);
    for (int i = 0; i < DIM; i++) {
        c->lo[i] = lo[i];
        c->hi[i] = hi[i];
        c->nc[i] = 0;
    }
    c->cid = next_id(g);
    for (int i = 0; i < 2 * DIM; i++) c->fid[i] = next_id(g);
    return c;
}

/* higtree/src/higtree.c:250-255 */
hcell *hfo_create_root(idgen *g, const Point lo, const Point hi)
{
    hcell *c = new_cell(g, lo, hi);
    c->pos = -1;
    c->parent = NULL;
    return c;
}

/* higtree/src/higtree.c:257-284: child lowpoint = (delta*pp) + lowpoint,
 * highpoint = lowpoint + delta, delta = (hi-lo)/numcells, in that order. */
void hfo_refine_uniform(idgen *g, hcell *c, const int nc[DIM])
{
    Point delta;
    int n = 1;
    for (int i = 0; i < DIM; i++) {
        c->nc[i] = nc[i];
        delta[i] = c->hi[i] - c->lo[i];
        delta[i] = delta[i] / nc[i];
        n *= nc[i];
    }
    c->ch = calloc((size_t)n, sizeof *c->ch);
    for (int j = 0; j < n; j++) {
        int pp[DIM];
        hfo_tobase(j, nc, pp);
        Point lo, hi;
        for (int i = 0; i < DIM; i++) {
            lo[i] = delta[i] * pp[i];
            lo[i] = lo[i] + c->lo[i];
            hi[i] = lo[i] + delta[i];
        }
        hcell *ch = new_cell(g, lo, hi);
        ch->pos = j;
        ch->parent = c;
        c->ch[j] = ch;
    }
}

static int nchildren(const hcell *c)
{
    int n = 1;
    for (int i = 0; i < DIM; i++) n *= c->nc[i];
    return n;
}

void hfo_destroy_tree(hcell *c)
{
    if (!c) return;
    if (c->ch) {
        int n = nchildren(c);
        for (int j = 0; j < n; j++) hfo_destroy_tree(c->ch[j]);
        free(c->ch);
    }
    free(c);
}

/* higtree/src/higtree.c:61-66: center = (lo + hi) / 2 */
void hfo_center(const hcell *c, Point p)
{
    for (int i = 0; i < DIM; i++) {
        p[i] = c->lo[i] + c->hi[i];
        p[i] = p[i] / 2.0;
    }
}

/* higtree/src/higtree.c:68-72 */
void hfo_delta(const hcell *c, Point p)
{
    for (int i = 0; i < DIM; i++) p[i] = c->hi[i] - c->lo[i];
}

/* higtree/src/higtree.c:371-415 */
static void coords_of_point(const hcell *cell, const Point point, int pp[DIM])
{
    for (int i = 0; i < DIM; i++) pp[i] = 0;
    for (int dim = 0; dim < DIM; dim++) {
        int first = 0;
        int last = cell->nc[dim] - 1;
        while (last >= first) {
            pp[dim] = first;
            if (last == first) break;
            int pfirst = hfo_toposition(cell->nc, pp);
            pp[dim] = last;
            int plast = hfo_toposition(cell->nc, pp);
            real delta = (cell->ch[plast]->hi[dim] - cell->ch[pfirst]->lo[dim]) /
                         (last - first + 1);
            int pos = (int)((point[dim] - cell->ch[pfirst]->lo[dim]) / delta) + first;
            if (pos > last) pos = last;
            if (pos < first) pos = first;
            pp[dim] = pos;
            int ppos = hfo_toposition(cell->nc, pp);
            if (FLT_LE(point[dim], cell->ch[ppos]->lo[dim])) {
                last = pos - 1;
            } else if (FLT_GT(point[dim], cell->ch[ppos]->hi[dim])) {
                first = pos + 1;
            } else {
                break;
            }
        }
    }
    for (int dim = 0; dim < DIM; dim++) {
        if (FLT_LT(point[dim], cell->lo[dim])) pp[dim] = -1;
        if (FLT_GT(point[dim], cell->hi[dim])) pp[dim] = cell->nc[dim];
    }
}

/* higtree/src/higtree.c:417-433 */
hcell *hfo_cell_with_point(hcell *top, const Point x)
{
    hcell *cell = top;
    for (int i = 0; i < DIM; i++) {
        if (FLT_GT(cell->lo[i], x[i]) || FLT_LT(cell->hi[i], x[i])) return NULL;
    }
    while (cell->ch != NULL) {
        int pp[DIM];
        coords_of_point(cell, x, pp);
        cell = cell->ch[hfo_toposition(cell->nc, pp)];
    }
    return cell;
}

/* higtree/src/higtree.c:445-455 */
hcell *hfo_cell_with_point_from_another(hcell *c, const Point x)
{
    if (c == NULL) return NULL;
    for (int i = 0; i < DIM; i++) {
        if (FLT_LT(x[i], c->lo[i]) || FLT_GT(x[i], c->hi[i]))
            return hfo_cell_with_point_from_another(c->parent, x);
    }
    return hfo_cell_with_point(c, x);
}

/* higtree/src/higtree.c:435-443 */
hcell *hfo_first_leaf(hcell *c)
{
    while (c->ch != NULL) c = c->ch[0];
    return c;
}

/* higtree/src/higtree.c:173-184 */
int hfo_narrowest_dim(const hcell *c)
{
    real minw = 0.0;
    int dim = -1;
    for (int i = 0; i < DIM; i++) {
        real w = c->hi[i] - c->lo[i];
        if (dim == -1 || w < minw) {
            minw = w;
            dim = i;
        }
    }
    return dim;
}

/* higtree/src/higtree.c:676-687 */
void hfo_facet_center(const hfacet *f, Point p)
{
    hfo_center(f->c, p);
    p[f->dim] = (f->dir == 0) ? f->c->lo[f->dim] : f->c->hi[f->dim];
}

/* higtree/src/higtree.c:710-750 */
bool hfo_facet_with_point(hcell *root, int dim, const Point x, hfacet *f)
{
    hcell *c = hfo_cell_with_point(root, x);
    if (c == NULL) return false;
    int dir = -1;
    if (FLT_EQ(x[dim], c->lo[dim])) {
        dir = 0;
    } else if (FLT_EQ(x[dim], c->hi[dim])) {
        dir = 1;
    }
    if (dir == -1) return false;
    int id = c->fid[dim * 2 + dir];
    if (id > 0) {
        f->c = c;
        f->dim = dim;
        f->dir = dir;
        return true;
    }
    Point x2;
    for (int i = 0; i < DIM; i++) x2[i] = x[i];
    if (dir == 0) x2[dim] -= EPSDELTA;
    else x2[dim] += EPSDELTA;
    hcell *c2 = hfo_cell_with_point_from_another(c, x2);
    if (c2 == NULL) return false;
    f->c = c2;
    f->dim = dim;
    f->dir = 1 - dir;
    return true;
}

/* higtree/src/higtree.c:466-477 */
static bool bbox_inside_cell(const hcell *c, const Point lo, const Point hi)
{
    for (int i = 0; i < DIM; i++) {
        if (FLT_LT(lo[i], c->lo[i]) || FLT_GT(hi[i], c->hi[i])) return false;
    }
    return true;
}

/* higtree/src/higtree.c:479-489 */
bool hfo_intersect_bbox(const hcell *c, const Point lo, const Point hi)
{
    for (int i = 0; i < DIM; i++) {
        if (FLT_GT(lo[i], c->hi[i]) || FLT_LT(hi[i], c->lo[i])) return false;
    }
    return true;
}

/* higtree/src/higtree.c:658-674, 693-708: strict interval overlap */
bool hfo_facet_intersect_bbox(const hfacet *f, const Point lo, const Point hi)
{
    Point fl, fh;
    for (int i = 0; i < DIM; i++) {
        fl[i] = f->c->lo[i];
        fh[i] = f->c->hi[i];
    }
    if (f->dir == 1) fl[f->dim] = f->c->hi[f->dim];
    else fh[f->dim] = f->c->lo[f->dim];
    for (int i = 0; i < DIM; i++) {
        if (!((fl[i] < hi[i]) && (lo[i] < fh[i]))) return false;
    }
    return true;
}

void cellvec_push(cellvec *v, hcell *c)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->v = realloc(v->v, (size_t)v->cap * sizeof *v->v);
    }
    v->v[v->n++] = c;
}

void facetvec_push(facetvec *v, hfacet f)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 64;
        v->v = realloc(v->v, (size_t)v->cap * sizeof *v->v);
    }
    v->v[v->n++] = f;
}

/* higtree/src/higtree-iterator.c:54-99: depth-first, children in position order */
void hfo_leaves(hcell *root, cellvec *out)
{
    if (root->ch == NULL) {
        cellvec_push(out, root);
        return;
    }
    int n = nchildren(root);
    for (int j = 0; j < n; j++) hfo_leaves(root->ch[j], out);
}

/* higtree/src/higtree-iterator.c:151-196: per level the children between the
 * cells holding the box corners are visited by an odometer whose last
 * dimension is fastest; the corner child is always visited first. */
static void bbox_descend(hcell *p, const Point lo, const Point hi, cellvec *out)
{
    if (p->ch == NULL) {
        cellvec_push(out, p);
        return;
    }
    int ppl[DIM], pph[DIM], pps[DIM];
    coords_of_point(p, lo, ppl);
    coords_of_point(p, hi, pph);
    for (int i = 0; i < DIM; i++) {
        if (ppl[i] < 0) ppl[i] = 0;
        if (pph[i] >= p->nc[i]) pph[i] = p->nc[i] - 1;
        pps[i] = ppl[i];
    }
    bbox_descend(p->ch[hfo_toposition(p->nc, pps)], lo, hi, out);
    for (;;) {
        int i = DIM - 1;
        while (i >= 0 && pps[i] >= pph[i]) {
            pps[i] = ppl[i];
            i--;
        }
        if (i < 0) break;
        pps[i]++;
        bbox_descend(p->ch[hfo_toposition(p->nc, pps)], lo, hi, out);
    }
}

/* higtree/src/higtree-iterator.c:220-249 (always called with a root here) */
void hfo_bbox_cells(hcell *root, const Point lo, const Point hi, cellvec *out)
{
    hcell *p = root;
    while (p->parent != NULL && !bbox_inside_cell(p, lo, hi)) p = p->parent;
    if (hfo_intersect_bbox(p, lo, hi)) bbox_descend(p, lo, hi, out);
}

/* higtree/src/higtree-iterator.c:616-747: facets of the box cells in
 * (dim, dir) order, kept when of interest, owned (fid > 0) and strictly
 * intersecting the box. */
void hfo_bbox_facets(hcell *root, const int dimofinterest[DIM], const Point lo,
                     const Point hi, facetvec *out)
{
    cellvec cv = {0};
    hfo_bbox_cells(root, lo, hi, &cv);
    for (int k = 0; k < cv.n; k++) {
        hcell *c = cv.v[k];
        for (int dim = 0; dim < DIM; dim++) {
            for (int dir = 0; dir < 2; dir++) {
                if (!dimofinterest[dim]) continue;
                if (c->fid[dim * 2 + dir] <= 0) continue;
                hfacet f = {c, dim, dir};
                if (!hfo_facet_intersect_bbox(&f, lo, hi)) continue;
                facetvec_push(out, f);
            }
        }
    }
    free(cv.v);
}

/* higtree/src/coord.c:15-22 */
real hfo_distance(const Point a, const Point b)
{
    real sum = 0;
    for (int i = 0; i < DIM; i++) {
        const real d = a[i] - b[i];
        sum += d * d;
    }
    return sqrt(sum);
}

/* higtree/src/coord.c:33-51 */
bool hfo_point_equal(const Point a, const Point b)
{
    for (int i = 0; i < DIM; i++) {
        if (fabs(a[i] - b[i]) >= EPSMACH) return false;
    }
    return true;
}

/* higtree/src/rect.c:44-53 */
bool hfo_rect_contains(const Point lo, const Point hi, const Point x)
{
    for (int i = 0; i < DIM; i++) {
        if (!POS_LE(x[i], hi[i])) return false;
        if (!POS_GE(x[i], lo[i])) return false;
    }
    return true;
}

/* higtree/src/higtree-io.c:269-293 (parse) and :97-153 (build):
 * level 0 refines the root uniformly; each finer level walks its patches in
 * steps of two cells (first dimension fastest) and halves every cell that is
 * still coarser than the level's delta. */
hcell *hfo_read_amr(idgen *g, const char *text)
{
    const char *p = text;
    char *end;
    Point l, h;
#define NEXTD(v)                                                                  \
    do {                                                                          \
        v = strtod(p, &end);                                                      \
        if (end == p) return NULL;                                                \
        p = end;                                                                  \
    } while (0)
#define NEXTI(v)                                                                  \
    do {                                                                          \
        v = (int)strtol(p, &end, 10);                                             \
        if (end == p) return NULL;                                                \
        p = end;                                                                  \
    } while (0)
    for (int d = 0; d < DIM; d++) {
        NEXTD(l[d]);
        NEXTD(h[d]);
    }
    int nlevels;
    NEXTI(nlevels);
    if (nlevels < 1) return NULL;
    Point *ldelta = calloc((size_t)nlevels, sizeof *ldelta);
    int *npatches = calloc((size_t)nlevels, sizeof *npatches);
    int **init = calloc((size_t)nlevels, sizeof *init);
    int **size = calloc((size_t)nlevels, sizeof *size);
    for (int lv = 0; lv < nlevels; lv++) {
        for (int d = 0; d < DIM; d++) NEXTD(ldelta[lv][d]);
        NEXTI(npatches[lv]);
        init[lv] = calloc((size_t)npatches[lv] * DIM, sizeof(int));
        size[lv] = calloc((size_t)npatches[lv] * DIM, sizeof(int));
        for (int pt = 0; pt < npatches[lv]; pt++) {
            for (int d = 0; d < DIM; d++) NEXTI(init[lv][pt * DIM + d]);
            for (int d = 0; d < DIM; d++) NEXTI(size[lv][pt * DIM + d]);
        }
    }
#undef NEXTD
#undef NEXTI
    hcell *root = hfo_create_root(g, l, h);
    hfo_refine_uniform(g, root, &size[0][0]);
    int two[DIM];
    for (int d = 0; d < DIM; d++) two[d] = 2;
    for (int lv = 1; lv < nlevels; lv++) {
        for (int pt = 0; pt < npatches[lv]; pt++) {
            const int *pinit = &init[lv][pt * DIM];
            const int *psize = &size[lv][pt * DIM];
            Point initpatch;
            for (int d = 0; d < DIM; d++) {
                initpatch[d] = ldelta[lv][d] * pinit[d];
                initpatch[d] = initpatch[d] + l[d];
            }
            int k[DIM] = {0};
            for (;;) {
                Point point;
                for (int d = 0; d < DIM; d++) {
                    real off = ldelta[lv][d] * k[d];
                    point[d] = initpatch[d] + off;
                }
                hcell *c = hfo_cell_with_point(root, point);
                if (c) {
                    Point cd;
                    hfo_delta(c, cd);
                    int enough = 1;
                    for (int d = 0; d < DIM; d++) {
                        if (FLT_LT(ldelta[lv][d], cd[d])) {
                            enough = 0;
                            break;
                        }
                    }
                    if (!enough) hfo_refine_uniform(g, c, two);
                }
                int i = 0;
                while (i < DIM && k[i] + 2 >= psize[i] - 1) {
                    k[i] = 0;
                    i++;
                }
                if (i < DIM) k[i] += 2;
                else break;
            }
        }
    }
    for (int lv = 0; lv < nlevels; lv++) {
        free(init[lv]);
        free(size[lv]);
    }
    free(init);
    free(size);
    free(ldelta);
    free(npatches);
    return root;
}

void idmap_init(idmap *m, int maxid)
{
    m->size = maxid + 1;
    m->lid = calloc((size_t)m->size, sizeof(int));
    m->count = 0;
}

int idmap_lookup(const idmap *m, int id)
{
    if (id <= 0 || id >= m->size) return -1;
    return m->lid[id] - 1;
}

/* higtree/src/mapper.c:70-98: assign the next local id unless already mapped */
void idmap_assign(idmap *m, int id)
{
    if (id <= 0 || id >= m->size) return;
    if (m->lid[id] == 0) m->lid[id] = ++m->count;
}
