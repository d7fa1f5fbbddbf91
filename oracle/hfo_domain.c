/*
 * CPU oracle (test infrastructure): simulation domains, boundary conditions,
 * the point-stencil builder and the moving-least-squares interpolation.
 * Restates higtree/src/domain.c, wls.c, pdomain.c (unknown-facet rule) and
 * point-mapper.c.  See hfo_internal.h for the oracle's role.
 */
#include "hfo_internal.h"

static real rmax(real a, real b) { return a > b ? a : b; }

/* ------------------------------------------------------------ stencil */
/* higtree/src/domain.c:1944-2010: accumulate by id, keep first-seen order */
void stn_init(hstencil *s)
{
    s->cap = 100;
    s->n = 0;
    s->ids = malloc(sizeof(int) * (size_t)s->cap);
    s->vals = malloc(sizeof(real) * (size_t)s->cap);
    s->rhs = 0.0;
}

void stn_free(hstencil *s)
{
    free(s->ids);
    free(s->vals);
}

void stn_reset(hstencil *s)
{
    s->n = 0;
    s->rhs = 0.0;
}

void stn_add_elem(hstencil *s, int id, real v)
{
    for (int i = 0; i < s->n; i++) {
        if (s->ids[i] == id) {
            s->vals[i] += v;
            return;
        }
    }
    if (s->n >= s->cap) {
        s->cap *= 2;
        s->ids = realloc(s->ids, sizeof(int) * (size_t)s->cap);
        s->vals = realloc(s->vals, sizeof(real) * (size_t)s->cap);
    }
    s->ids[s->n] = id;
    s->vals[s->n] = v;
    s->n++;
}

/* higtree/src/pdomain.c:1026-1054: value = -rhs + sum w*v in stencil order */
real hfo_interpolate(const real *values, const hstencil *stn)
{
    real res = 0.0;
    res += -stn->rhs;
    for (int i = 0; i < stn->n; i++) {
        real v = values[stn->ids[i]];
        res += v * stn->vals[i];
    }
    return res;
}

/* ---------------------------------------------------------------- WLS */
/* higtree/src/wls.c:207-232 */
static int comb_rep(int n, int k)
{
    int i, c;
    for (i = n, c = 1; i > n - k; i--) c *= i;
    for (i = 1; i <= k; i++) c /= i;
    return c;
}

int hfo_wls_num_min_points(int dim, int ord)
{
    int n = 1;
    for (int i = 1; i <= ord; i++) n += comb_rep(dim + i - 1, i);
    return n;
}

static real **alloc_mat(int m, int n)
{
    real **v = malloc(sizeof(real *) * (size_t)(m + 1));
    for (int i = 0; i <= m; i++) v[i] = calloc((size_t)(n + 1), sizeof(real));
    return v;
}

/* higtree/src/wls.c:333-367 (the random polynomial is unused by the step) */
hwls *hfo_wls_create(int dim, int ord, int maxnumpts)
{
    hwls *w = calloc(1, sizeof *w);
    w->dim = dim;
    w->ord = ord;
    w->maxnumpts = maxnumpts;
    w->numpoly = hfo_wls_num_min_points(dim, ord);
    w->A = alloc_mat(maxnumpts, w->numpoly);
    w->Q = alloc_mat(maxnumpts, maxnumpts);
    w->D = calloc((size_t)maxnumpts + 1, sizeof(real));
    w->b = calloc((size_t)maxnumpts + 1, sizeof(real));
    w->w = calloc((size_t)maxnumpts + 1, sizeof(real));
    return w;
}

void hfo_wls_destroy(hwls *w)
{
    if (!w) return;
    for (int i = 0; i <= w->maxnumpts; i++) {
        free(w->A[i]);
        free(w->Q[i]);
    }
    free(w->A);
    free(w->Q);
    free(w->D);
    free(w->b);
    free(w->w);
    free(w);
}

/* higtree/src/wls.c:236-292: monomials in absolute coordinates, truncated
 * once more than m+1 entries were written */
static void assembly_row(int dim, int ord, int m, real *a, const real *x)
{
    int k = 0;
    if (ord >= 0) {
        a[k] = 1;
        k++;
        if (k > m) return;
    }
    if (ord >= 1) {
        for (int i1 = 0; i1 < dim; i1++) {
            a[k] = x[i1];
            k++;
            if (k > m) return;
        }
    }
    if (ord >= 2) {
        for (int i1 = 0; i1 < dim; i1++)
            for (int i2 = i1; i2 < dim; i2++) {
                a[k] = x[i1] * x[i2];
                k++;
                if (k > m) return;
            }
    }
    if (ord >= 3) {
        for (int i1 = 0; i1 < dim; i1++)
            for (int i2 = i1; i2 < dim; i2++)
                for (int i3 = i2; i3 < dim; i3++) {
                    a[k] = x[i1] * x[i2] * x[i3];
                    k++;
                    if (k > m) return;
                }
    }
}

/* higtree/src/wls.c:15-44 */
static void givens(int m, int n, int i, int k, real **A, real **Q, real s1, real s2)
{
    real s, t;
    if (fabs(s2) + fabs(s1) > 0.0) {
        if (fabs(s2) >= fabs(s1))
            s = sqrt(1.0 + (s1 / s2) * (s1 / s2)) * fabs(s2);
        else
            s = sqrt(1.0 + (s2 / s1) * (s2 / s1)) * fabs(s1);
        s1 = s1 / s;
        s2 = s2 / s;
        for (int j = 0; j < n; j++) {
            s = s1 * A[i][j] + s2 * A[k][j];
            t = -s2 * A[i][j] + s1 * A[k][j];
            A[i][j] = s;
            A[k][j] = t;
        }
        for (int j = 0; j < m; j++) {
            s = s1 * Q[i][j] + s2 * Q[k][j];
            t = -s2 * Q[i][j] + s1 * Q[k][j];
            Q[i][j] = s;
            Q[k][j] = t;
        }
    }
}

/* higtree/src/wls.c:55-73 */
static void wls_qr(int m, int n, real **A, real **Q)
{
    for (int i = 0; i < m; i++) {
        for (int j = 0; j < m; j++) Q[i][j] = 0.0;
        Q[i][i] = 1.0;
    }
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < m; j++) givens(m, n, i, j, A, Q, A[i][i], A[j][i]);
}

/* higtree/src/wls.c:104-124: x = Q^T [R^-T y ; 0], dropping |R_ii| <= eps */
static void wls_solve(int m, int n, real **A, real **Q, real *x, real *y, real eps)
{
    int i, j;
    for (i = 0; i < n; i++) {
        if (fabs(A[i][i]) > eps) {
            for (j = 0, x[i] = y[i]; j < i; j++) x[i] -= A[j][i] * x[j];
            x[i] = x[i] / A[i][i];
        } else {
            x[i] = 0.0;
        }
    }
    for (i = 0; i < n; i++) y[i] = x[i];
    for (i = n; i < m; i++) y[i] = 0.0;
    for (i = 0; i < m; i++)
        for (j = 0, x[i] = 0.0; j < m; j++) x[i] += Q[j][i] * y[j];
}

/* higtree/src/wls.c:495-532: inverse-distance row scaling, Givens QR, solve */
void hfo_wls_calc(hwls *wls, int numpts, const witem *items, const real *x, real *wout)
{
    for (int i = 0; i < numpts; i++)
        assembly_row(wls->dim, wls->ord, numpts, wls->A[i], items[i].x);
    for (int i = 0; i < numpts; i++) {
        real mlsw = 0.0;
        for (int j = 0; j < wls->dim; j++)
            mlsw += (x[j] - items[i].x[j]) * (x[j] - items[i].x[j]);
        wls->D[i] = sqrt(1.0 / (mlsw + EPSMACH));
        for (int j = 0; j < wls->numpoly; j++) wls->A[i][j] *= wls->D[i];
    }
    wls_qr(numpts, wls->numpoly, wls->A, wls->Q);
    assembly_row(wls->dim, wls->ord, numpts, wls->b, x);
    wls_solve(numpts, wls->numpoly, wls->A, wls->Q, wls->w, wls->b, EPSMACH);
    for (int i = 0; i < numpts; i++) wout[i] = wls->D[i] * wls->w[i];
}

/* ------------------------------------------------------------- qsort */
/* glibc 2.35 qsort on small arrays = msort.c msort_with_tmp: top-down merge
 * sort, n1 = n/2, takes the left element when cmp(left,right) <= 0.  With
 * the reference comparator (domain.c:373-377, never 0) equal distances take
 * the right run first. */
static int item_cmp(const witem *a, const witem *b) { return a->dist < b->dist ? -1 : 1; }

static void msort_rec(witem *b, int n, witem *tmp)
{
    if (n <= 1) return;
    int n1 = n / 2, n2 = n - n1;
    witem *b1 = b, *b2 = b + n1;
    msort_rec(b1, n1, tmp);
    msort_rec(b2, n2, tmp);
    witem *t = tmp;
    while (n1 > 0 && n2 > 0) {
        if (item_cmp(b1, b2) <= 0) {
            *t++ = *b1++;
            n1--;
        } else {
            *t++ = *b2++;
            n2--;
        }
    }
    if (n1 > 0) memcpy(t, b1, (size_t)n1 * sizeof *t);
    memcpy(b, tmp, (size_t)(n - n2) * sizeof *b);
}

void hfo_msort_items(witem *a, int n)
{
    if (n <= 1) return;
    witem *tmp = malloc((size_t)n * sizeof *tmp);
    msort_rec(a, n, tmp);
    free(tmp);
}

/* ------------------------------------------------------------- cache */
/* higtree/src/point-mapper.c:3-8 */
static void discretize(const real *x, int64_t *k)
{
    for (int d = 0; d < DIM; d++) k[d] = (int64_t)(x[d] / EPSMACH + 0.5);
}

static unsigned key_hash(const int64_t *k)
{
    uint64_t h = 1469598103934665603ull;
    for (int d = 0; d < DIM; d++) {
        h ^= (uint64_t)k[d];
        h *= 1099511628211ull;
    }
    return (unsigned)(h ^ (h >> 32));
}

static cache_entry *cache_find(pcache *c, const real *x, bool create)
{
    if (!c->buckets) {
        c->nbuckets = 1 << 16;
        c->buckets = calloc((size_t)c->nbuckets, sizeof *c->buckets);
    }
    int64_t k[DIM];
    discretize(x, k);
    unsigned h = key_hash(k) & (unsigned)(c->nbuckets - 1);
    for (cache_entry *e = c->buckets[h]; e; e = e->next) {
        if (memcmp(e->key, k, sizeof k) == 0) return e;
    }
    if (!create) return NULL;
    cache_entry *e = calloc(1, sizeof *e);
    memcpy(e->key, k, sizeof k);
    e->next = c->buckets[h];
    c->buckets[h] = e;
    c->count++;
    return e;
}

static void free_cwls(cached_wls *w)
{
    if (!w) return;
    free(w->ids);
    free(w->w);
    free(w->bcpoints);
    free(w);
}

static void cache_free(pcache *c)
{
    if (!c->buckets) return;
    for (int i = 0; i < c->nbuckets; i++) {
        cache_entry *e = c->buckets[i];
        while (e) {
            cache_entry *n = e->next;
            if (e->without_bc != e->with_bc) free_cwls(e->without_bc);
            free_cwls(e->with_bc);
            free(e);
            e = n;
        }
    }
    free(c->buckets);
    c->buckets = NULL;
}

/* ------------------------------------------------------------ domains */
/* higtree/src/domain.c:102-127 */
static int calc_num_points(int dim, int order)
{
    if (dim == 2) {
        if (order == 3) return 60;
        if (order == 2) return 45;
    } else if (dim == 3) {
        if (order == 3) return 150;
        if (order == 2) return 120;
    }
    return 2 * dim * hfo_wls_num_min_points(dim, order);
}

static void interp_set(interp *in, int dim, int order)
{
    in->dim = dim;
    in->order = order;
    in->maxpts = calc_num_points(dim, order);
    in->numminpoints = hfo_wls_num_min_points(dim, order);
    in->wls = NULL;
}

static hwls *interp_wls(interp *in)
{
    if (!in->wls) in->wls = hfo_wls_create(in->dim, in->order, in->maxpts);
    return in->wls;
}

/* higtree/src/domain.c:129-160 (boundary interpolation: order 1 in DIM-1) */
hdomain *hfo_sd_create(int order)
{
    hdomain *d = calloc(1, sizeof *d);
    interp_set(&d->inter, DIM, order);
    interp_set(&d->bc_inter, DIM - 1, 1);
    d->use_cache = 1;
    return d;
}

void hfo_sd_destroy(hdomain *d)
{
    if (!d) return;
    hfo_wls_destroy(d->inter.wls);
    hfo_wls_destroy(d->bc_inter.wls);
    cache_free(&d->cache);
    free(d->trees);
    free(d->dir);
    free(d->neu);
    free(d->map.lid);
    free(d->cells);
    free(d);
}

void hfo_sd_add_bc(hdomain *d, hboundary *b)
{
    if (b->type == HFO_DIRICHLET) {
        d->dir = realloc(d->dir, sizeof *d->dir * (size_t)(d->ndir + 1));
        d->dir[d->ndir++] = b;
    } else {
        d->neu = realloc(d->neu, sizeof *d->neu * (size_t)(d->nneu + 1));
        d->neu[d->nneu++] = b;
    }
}

/* higtree/src/pdomain.c:154-175: local ids follow the domain cell iterator */
void hfo_sd_build_map(hdomain *d, int maxid)
{
    idmap_init(&d->map, maxid);
    cellvec cv = {0};
    for (int t = 0; t < d->ntrees; t++) hfo_leaves(d->trees[t], &cv);
    for (int k = 0; k < cv.n; k++) idmap_assign(&d->map, cv.v[k]->cid);
    d->ncells = d->map.count;
    d->cells = malloc(sizeof *d->cells * (size_t)(d->ncells > 0 ? d->ncells : 1));
    for (int k = 0; k < cv.n; k++) {
        int lid = idmap_lookup(&d->map, cv.v[k]->cid);
        d->cells[lid] = cv.v[k];
    }
    free(cv.v);
}

/* higtree/src/domain.c:8-34, hig-flow-bc.c:12-25 */
hboundary *hfo_bc_create(hcell *tree, int type, int userid, int maxid)
{
    hboundary *b = calloc(1, sizeof *b);
    b->tree = tree;
    b->type = type;
    b->userid = userid;
    idmap_init(&b->map, maxid);
    cellvec cv = {0};
    hfo_leaves(tree, &cv);
    for (int k = 0; k < cv.n; k++) idmap_assign(&b->map, cv.v[k]->cid);
    b->nvals = b->map.count;
    b->vals = calloc((size_t)b->nvals + 1, sizeof(real));
    b->cells = malloc(sizeof *b->cells * (size_t)(b->nvals + 1));
    for (int k = 0; k < cv.n; k++) b->cells[idmap_lookup(&b->map, cv.v[k]->cid)] = cv.v[k];
    free(cv.v);
    return b;
}

hcell *hfo_sd_cell_with_point(hdomain *d, const Point x)
{
    for (int i = 0; i < d->ntrees; i++) {
        hcell *c = hfo_cell_with_point(d->trees[i], x);
        if (c) return c;
    }
    return NULL;
}

/* higtree/src/domain.c:1703-1722 */
hfacetdomain *hfo_sfd_create(int dim, int order, hcell **trees, int ntrees)
{
    hfacetdomain *f = calloc(1, sizeof *f);
    f->cdom = hfo_sd_create(order);
    f->cdom->sfd = f;
    f->dim = dim;
    f->dimofinterest[dim] = 1;
    f->cdom->trees = malloc(sizeof(hcell *) * (size_t)ntrees);
    memcpy(f->cdom->trees, trees, sizeof(hcell *) * (size_t)ntrees);
    f->cdom->ntrees = ntrees;
    return f;
}

void hfo_sfd_destroy(hfacetdomain *f)
{
    if (!f) return;
    if (f->sfbi_l) {
        for (int t = 0; t < f->cdom->ntrees; t++) {
            free(f->sfbi_l[t]);
            free(f->sfbi_h[t]);
        }
        free(f->sfbi_l);
        free(f->sfbi_h);
    }
    free(f->fmap.lid);
    free(f->facets);
    hfo_sd_destroy(f->cdom);
    free(f);
}

/* higtree/src/domain.c:1440-1521: a facet shared with a same-size neighbour
 * belongs to the left cell (dir 0 gets a negative id); a coarse facet facing
 * smaller cells is disowned (id 0). */
void hfo_sfd_adjust_facet_ids(hfacetdomain *f)
{
    hdomain *sd = f->cdom;
    int dim = f->dim;
    for (int i = 0; i < sd->ntrees; i++) {
        hcell *root = sd->trees[i];
        cellvec cv = {0};
        hfo_leaves(root, &cv);
        for (int k = 0; k < cv.n; k++) {
            hcell *c = cv.v[k];
            Point ofc_p;
            hfo_center(c, ofc_p);
            for (int dir = 0; dir < 2; dir++) {
                ofc_p[dim] = (dir == 0) ? c->lo[dim] - EPSDELTA : c->hi[dim] + EPSDELTA;
                int different = 0, smaller = 0;
                hcell *ofc = hfo_cell_with_point(root, ofc_p);
                for (int j = 0; j < sd->ntrees && ofc == NULL; j++) {
                    if (j != i) ofc = hfo_cell_with_point(sd->trees[j], ofc_p);
                }
                if (ofc != NULL) {
                    for (int kk = 0; kk < DIM; kk++) {
                        if (kk == dim) continue;
                        if (FLT_NE(c->lo[kk], ofc->lo[kk]) || FLT_NE(c->hi[kk], ofc->hi[kk])) {
                            different = 1;
                            if (FLT_GT((c->hi[kk] - c->lo[kk]), (ofc->hi[kk] - ofc->lo[kk])))
                                smaller = 1;
                        }
                    }
                }
                if (different) {
                    if (smaller) c->fid[dim * 2 + dir] = 0;
                } else if (ofc != NULL) {
                    if (dir == 0) c->fid[dim * 2 + dir] = -ofc->fid[dim * 2 + 1];
                }
            }
        }
        free(cv.v);
    }
}

static hcell *child_in_grid(hcell *root, const int pp[DIM])
{
    return root->ch[hfo_toposition(root->nc, pp)];
}

/* higtree/src/pdomain.c:1146-1221: left block facets are never unknowns;
 * right ones are when interior, on a Neumann BC or shared with another tree */
static void block_facet_center(hfacetdomain *f, hcell *root, const int pp[DIM], int dir,
                               Point p)
{
    hcell *c = child_in_grid(root, pp);
    hfo_center(c, p);
    p[f->dim] = (dir == 0) ? c->lo[f->dim] : c->hi[f->dim];
}

static bool in_neumann_bc(hfacetdomain *f, hcell *root, const int pp[DIM], int dir)
{
    Point p;
    block_facet_center(f, root, pp, dir, p);
    for (int i = 0; i < f->cdom->nneu; i++) {
        if (hfo_cell_with_point(f->cdom->neu[i]->tree, p)) return true;
    }
    return false;
}

static bool has_flow(hfacetdomain *f, int t, const int pp[DIM], int dir)
{
    Point p;
    block_facet_center(f, f->cdom->trees[t], pp, dir, p);
    for (int t2 = 0; t2 < f->cdom->ntrees; t2++) {
        if (t2 == t) continue;
        if (hfo_cell_with_point(f->cdom->trees[t2], p)) return true;
    }
    return false;
}

void hfo_sfd_compute_sfbi(hfacetdomain *f)
{
    int nt = f->cdom->ntrees;
    f->sfbi_l = calloc((size_t)nt, sizeof *f->sfbi_l);
    f->sfbi_h = calloc((size_t)nt, sizeof *f->sfbi_h);
    for (int t = 0; t < nt; t++) {
        hcell *root = f->cdom->trees[t];
        int n = 1;
        for (int i = 0; i < DIM; i++) n *= root->nc[i];
        f->sfbi_l[t] = calloc((size_t)n, 1);
        f->sfbi_h[t] = calloc((size_t)n, 1);
        for (int i = 0; i < n; i++) {
            int pp[DIM];
            hfo_tobase(i, root->nc, pp);
            f->sfbi_l[t][i] = in_neumann_bc(f, root, pp, 0);
            f->sfbi_h[t][i] = pp[f->dim] < (root->nc[f->dim] - 1) ||
                              in_neumann_bc(f, root, pp, 1) || has_flow(f, t, pp, 1);
        }
    }
}

/* higtree/src/domain.c:1834-1894: per block, the facets inside the block box
 * shrunk by EPSDELTA, grown across the flagged sides */
void hfo_sfd_domain_facets(hfacetdomain *f, facetvec *out)
{
    for (int t = 0; t < f->cdom->ntrees; t++) {
        hcell *root = f->cdom->trees[t];
        int n = 1;
        for (int i = 0; i < DIM; i++) n *= root->nc[i];
        for (int b = 0; b < n; b++) {
            hcell *c = root->ch[b];
            Point l, h;
            for (int i = 0; i < DIM; i++) {
                l[i] = c->lo[i] + EPSDELTA;
                h[i] = c->hi[i] - EPSDELTA;
            }
            if (f->sfbi_l[t][b]) l[f->dim] -= 2.0 * EPSDELTA;
            if (f->sfbi_h[t][b]) h[f->dim] += 2.0 * EPSDELTA;
            hfo_bbox_facets(root, f->dimofinterest, l, h, out);
        }
    }
}

/* higtree/src/pdomain.c:780-797, mapper.c:84-98 */
void hfo_sfd_build_map(hfacetdomain *f, int maxid)
{
    idmap_init(&f->fmap, maxid);
    facetvec fv = {0};
    hfo_sfd_domain_facets(f, &fv);
    for (int k = 0; k < fv.n; k++)
        idmap_assign(&f->fmap, fv.v[k].c->fid[fv.v[k].dim * 2 + fv.v[k].dir]);
    f->nfacets = f->fmap.count;
    f->facets = malloc(sizeof *f->facets * (size_t)(f->nfacets > 0 ? f->nfacets : 1));
    for (int k = 0; k < fv.n; k++) {
        int lid = idmap_lookup(&f->fmap, fv.v[k].c->fid[fv.v[k].dim * 2 + fv.v[k].dir]);
        f->facets[lid] = fv.v[k];
    }
    free(fv.v);
}

/* higtree/src/domain.c:1927-1938 */
int hfo_sfd_facet_with_point(hfacetdomain *f, const Point x, hfacet *out)
{
    for (int t = 0; t < f->cdom->ntrees; t++) {
        if (hfo_facet_with_point(f->cdom->trees[t], f->dim, x, out)) return 1;
    }
    return 0;
}

/* ------------------------------------------------------ stencil build */
typedef struct {
    witem *ptr;
    int total, numpts;
    real max_dist;
} itemlist;

static itemlist g_items; /* the reference keeps one static list (domain.c:1325) */

/* higtree/src/domain.c:388-412 */
static witem *new_item(int maxpts, itemlist *it, real fdist)
{
    if (fdist >= it->max_dist) {
        if (it->numpts >= maxpts) return NULL;
        it->max_dist = fdist;
    }
    if (it->numpts >= it->total) {
        it->total += maxpts;
        it->ptr = realloc(it->ptr, sizeof *it->ptr * (size_t)it->total);
    }
    witem *n = &it->ptr[it->numpts++];
    n->dist = fdist;
    return n;
}

/* higtree/src/domain.c:415-446 */
static void search_cells_in_tree_box(hcell *root, const idmap *m, const real *x,
                                     const Point lo, const Point hi, int maxpts,
                                     itemlist *items, hboundary *sb, unsigned char type)
{
    cellvec cv = {0};
    hfo_bbox_cells(root, lo, hi, &cv);
    for (int k = 0; k < cv.n; k++) {
        hcell *c = cv.v[k];
        int cid = idmap_lookup(m, c->cid);
        if (cid >= 0) {
            Point center;
            hfo_center(c, center);
            real dist = hfo_distance(x, center);
            witem *n = new_item(maxpts, items, dist);
            if (n) {
                n->id = cid;
                n->bcpoint = sb;
                n->type = type;
                for (int d = 0; d < DIM; d++) n->x[d] = center[d];
            }
        }
    }
    free(cv.v);
}

/* higtree/src/domain.c:1380-1390 (cells), 1523-1562 (facets) */
static void search_box(hdomain *d, const real *x, const Point lo, const Point hi,
                       itemlist *items)
{
    if (!d->sfd) {
        for (int k = 0; k < d->ntrees; k++)
            search_cells_in_tree_box(d->trees[k], &d->map, x, lo, hi, d->inter.maxpts, items,
                                     NULL, 0);
        return;
    }
    hfacetdomain *f = d->sfd;
    for (int k = 0; k < d->ntrees; k++) {
        facetvec fv = {0};
        hfo_bbox_facets(d->trees[k], f->dimofinterest, lo, hi, &fv);
        for (int q = 0; q < fv.n; q++) {
            hfacet *fc = &fv.v[q];
            int fid = idmap_lookup(&f->fmap, fc->c->fid[fc->dim * 2 + fc->dir]);
            if (fid >= 0) {
                Point fcenter;
                hfo_facet_center(fc, fcenter);
                real dist = hfo_distance(x, fcenter);
                witem *n = new_item(d->inter.maxpts, items, dist);
                if (n) {
                    n->id = fid;
                    n->type = 0;
                    n->bcpoint = NULL;
                    for (int dd = 0; dd < DIM; dd++) n->x[dd] = fcenter[dd];
                }
            }
        }
        free(fv.v);
    }
}

/* higtree/src/domain.c:1392-1427 (cells), 1565-1597 (facets) */
static bool find_in_center(hdomain *d, const real *x, real alpha, bool *in_domain,
                           hstencil *stn)
{
    *in_domain = false;
    for (int i = 0; i < d->ntrees; i++) {
        hcell *root = d->trees[i];
        if (!hfo_rect_contains(root->lo, root->hi, x)) continue;
        *in_domain = true;
        if (!d->sfd) {
            hcell *cx = hfo_cell_with_point(root, x);
            if (cx) {
                Point cc;
                hfo_center(cx, cc);
                if (hfo_point_equal(cc, x)) {
                    stn_add_elem(stn, idmap_lookup(&d->map, cx->cid), alpha);
                    return true;
                }
            }
            break;
        } else {
            hfacet fx;
            if (hfo_facet_with_point(root, d->sfd->dim, x, &fx)) {
                Point fc;
                hfo_facet_center(&fx, fc);
                if (hfo_point_equal(fc, x)) {
                    int fid = idmap_lookup(&d->sfd->fmap, fx.c->fid[fx.dim * 2 + fx.dir]);
                    if (fid >= 0) {
                        stn_add_elem(stn, fid, alpha);
                        return true;
                    }
                }
            }
        }
    }
    return false;
}

/* higtree/src/domain.c:1298-1319 */
static bool dirichlet_find_in_center(hdomain *d, const real *x, real alpha, hstencil *stn)
{
    for (int i = 0; i < d->ndir; i++) {
        hboundary *sb = d->dir[i];
        hcell *cx = hfo_cell_with_point(sb->tree, x);
        if (cx) {
            Point cc;
            hfo_center(cx, cc);
            if (hfo_point_equal(cc, x)) {
                real v = sb->vals[idmap_lookup(&sb->map, cx->cid)];
                stn->rhs += -alpha * v;
                return true;
            }
        }
    }
    return false;
}

static void get_stencil(hdomain *d, real delta, const real *x, real alpha, hstencil *stn,
                        bool use_dirichlet, bool use_neumann);

/* Boundary value at a point on (or projected onto) a BC tree: exact at a BC
 * cell centre, else the linear DIM-1 WLS fit of the BC values.  The fit keeps
 * the full-dimensional evaluation point while the samples drop the projected
 * coordinate (domain.c:800-817, 972-989, 1149-1166, 1251-1268). */
static real boundary_value(hdomain *d, hboundary *b, hcell *bcell, const real *px,
                           const real *boxc, real bd, int proj_dir)
{
    Point pc;
    hfo_center(bcell, pc);
    if (hfo_point_equal(pc, px)) return b->vals[idmap_lookup(&b->map, bcell->cid)];
    itemlist *items = &g_items;
    items->numpts = 0;
    items->max_dist = 0.0;
    Point lo, hi;
    for (int i = 0; i < DIM; i++) {
        lo[i] = boxc[i] - 3.0 * bd;
        hi[i] = boxc[i] + 3.0 * bd;
    }
    search_cells_in_tree_box(b->tree, &b->map, px, lo, hi, d->bc_inter.maxpts, items, NULL, 2);
    hfo_msort_items(items->ptr, items->numpts);
    if (items->numpts > d->bc_inter.maxpts) items->numpts = d->bc_inter.maxpts;
    for (int i = 0; i < items->numpts; i++) {
        real *xx = items->ptr[i].x;
        for (int dim = proj_dir + 1; dim < DIM; dim++) xx[dim - 1] = xx[dim];
    }
    real w[512];
    hfo_wls_calc(interp_wls(&d->bc_inter), items->numpts, items->ptr, px, w);
    real v = 0.0;
    for (int i = 0; i < items->numpts; i++) v += w[i] * b->vals[items->ptr[i].id];
    return v;
}

/* higtree/src/domain.c:692-870 */
static bool get_stencil_neumann(hdomain *d, const real *x, real alpha, hstencil *stn)
{
    hboundary *best = NULL;
    hcell *best_cell = NULL;
    Point best_px;
    real best_dist = DBL_MAX;
    int best_pd = -1;
    for (int i = 0; i < d->nneu; i++) {
        hboundary *b = d->neu[i];
        int pd = hfo_narrowest_dim(b->tree);
        Point tc, px;
        hfo_center(b->tree, tc);
        for (int k = 0; k < DIM; k++) px[k] = x[k];
        px[pd] = tc[pd];
        hcell *bc = hfo_cell_with_point(b->tree, px);
        if (!bc) continue;
        real dist = x[pd] - px[pd];
        if (fabs(dist) < fabs(best_dist)) {
            best = b;
            best_cell = bc;
            best_dist = dist;
            best_pd = pd;
            memcpy(best_px, px, sizeof px);
        }
    }
    if (!best) return false;
    {
        real delta = fabs(best_dist);
        for (int i = 0; i < d->ndir; i++) {
            hcell *t = d->dir[i]->tree;
            real acc = 0;
            for (int k = 0; k < DIM; k++) {
                const real dd = rmax(rmax(t->lo[k] - x[k], 0.0), x[k] - t->hi[k]);
                acc += dd * dd;
            }
            if (sqrt(acc) < delta) return false;
        }
    }
    real minus_h;
    {
        hcell *c = hfo_sd_cell_with_point(d, best_px);
        Point cd;
        hfo_delta(c, cd);
        minus_h = copysign(cd[best_pd], best_dist) * -0.5;
        if (d->sfd && d->sfd->dim == best_pd) minus_h *= 2;
    }
    Point inside;
    memcpy(inside, best_px, sizeof inside);
    inside[best_pd] += minus_h;
    Point bdl;
    hfo_delta(best_cell, bdl);
    real bd = 0.0;
    for (int k = 0; k < DIM; k++)
        if (k != best_pd) bd = rmax(bd, bdl[k]);
    get_stencil(d, bd, inside, alpha, stn, true, false);
    real slope = boundary_value(d, best, best_cell, best_px, x, bd, best_pd);
    stn->rhs += alpha * slope * (minus_h + best_dist);
    return true;
}

/* higtree/src/domain.c:877-1029 */
static bool get_stencil_dirichlet(hdomain *d, const real *x, real alpha, hstencil *stn)
{
#ifdef HFO_NO_OUTSIDE_DIRICHLET_GHOST
    /* variant without the outside-Dirichlet ghost rule: such points fall
     * through to WLS with the Dirichlet boundary centres as samples */
    (void)d;
    (void)x;
    (void)alpha;
    (void)stn;
    return false;
#endif
    hboundary *best = NULL;
    hcell *best_cell = NULL;
    Point best_px;
    real best_dist = DBL_MAX;
    int best_pd = -1;
    for (int i = 0; i < d->ndir; i++) {
        hboundary *b = d->dir[i];
        int pd = hfo_narrowest_dim(b->tree);
        Point tc, px;
        hfo_center(b->tree, tc);
        for (int k = 0; k < DIM; k++) px[k] = x[k];
        px[pd] = tc[pd];
        hcell *bc = hfo_cell_with_point(b->tree, px);
        if (!bc) continue;
        real dist = x[pd] - px[pd];
        if (fabs(dist) < fabs(best_dist)) {
            best = b;
            best_cell = bc;
            best_dist = dist;
            best_pd = pd;
            memcpy(best_px, px, sizeof px);
        }
    }
    if (!best) return false;
    real minus_h;
    {
        hcell *c = hfo_sd_cell_with_point(d, best_px);
        Point cd;
        hfo_delta(c, cd);
        minus_h = copysign(cd[best_pd], best_dist) * -0.5;
        if (d->sfd && d->sfd->dim == best_pd) minus_h *= 2;
    }
    Point inside;
    memcpy(inside, best_px, sizeof inside);
    inside[best_pd] += minus_h;
    real ratio = (minus_h - best_dist) / (minus_h);
    Point bdl;
    hfo_delta(best_cell, bdl);
    real bd = 0.0;
    for (int k = 0; k < DIM; k++)
        if (k != best_pd) bd = rmax(bd, bdl[k]);
    get_stencil(d, bd, inside, (1.0 - ratio) * alpha, stn, false, false);
    real bval = boundary_value(d, best, best_cell, best_px, x, bd, best_pd);
    stn->rhs += alpha * bval * (-ratio);
    return true;
}

/* higtree/src/domain.c:1037-1192 */
static bool get_stencil_neumann_boundary(hdomain *d, const real *x, real alpha,
                                         hstencil *stn)
{
    hboundary *cb = NULL;
    hcell *ccell = NULL;
    int pd = -1;
    for (int i = 0; i < d->nneu; i++) {
        hcell *c = hfo_cell_with_point(d->neu[i]->tree, x);
        if (c) {
            cb = d->neu[i];
            ccell = c;
            pd = hfo_narrowest_dim(cb->tree);
        }
    }
    if (!ccell) return false;
    real minus_h;
    {
        int out_sign = 1;
        Point xl, xr;
        memcpy(xl, x, sizeof xl);
        memcpy(xr, x, sizeof xr);
        xl[pd] -= EPSDELTA + EPSMACH;
        xr[pd] += EPSDELTA + EPSMACH;
        for (int i = 0; i < d->ntrees; i++) {
            hcell *root = d->trees[i];
            if (hfo_rect_contains(root->lo, root->hi, xl)) {
                out_sign = 1;
                break;
            }
            if (hfo_rect_contains(root->lo, root->hi, xr)) {
                out_sign = -1;
                break;
            }
        }
        hcell *c = hfo_sd_cell_with_point(d, x);
        Point cd;
        hfo_delta(c, cd);
        minus_h = copysign(cd[pd], (real)out_sign) * -0.5;
        if (d->sfd && d->sfd->dim == pd) minus_h *= 2;
    }
    Point inside;
    memcpy(inside, x, sizeof inside);
    inside[pd] += minus_h;
    Point bdl;
    hfo_delta(ccell, bdl);
    real bd = 0.0;
    for (int k = 0; k < DIM; k++)
        if (k != pd) bd = rmax(bd, bdl[k]);
    get_stencil(d, bd, inside, alpha, stn, false, false);
    real slope = boundary_value(d, cb, ccell, x, x, bd, pd);
    stn->rhs += alpha * slope * minus_h;
    return true;
}

/* higtree/src/domain.c:1200-1294 */
static bool get_stencil_dirichlet_boundary(hdomain *d, const real *x, real alpha,
                                           hstencil *stn)
{
    hboundary *cb = NULL;
    hcell *ccell = NULL;
    int pd = -1;
    for (int i = 0; i < d->ndir; i++) {
        hcell *c = hfo_cell_with_point(d->dir[i]->tree, x);
        if (c) {
            cb = d->dir[i];
            ccell = c;
            pd = hfo_narrowest_dim(cb->tree);
        }
    }
    if (!ccell) return false;
    Point bdl;
    hfo_delta(ccell, bdl);
    real bd = 0.0;
    for (int k = 0; k < DIM; k++)
        if (k != pd) bd = rmax(bd, bdl[k]);
    real bval = boundary_value(d, cb, ccell, x, x, bd, pd);
    stn->rhs += alpha * bval * (-1.0);
    return true;
}

/* higtree/src/domain.c:489-682 */
static void get_stencil_interpolate(hdomain *d, const real *x, real delta, real alpha,
                                    hstencil *stn, bool use_dirichlet)
{
    itemlist *items = &g_items;
    real w[1024];
    cache_entry *cached = NULL;
    if (d->use_cache) cached = cache_find(&d->cache, x, false);
    if (cached && (!use_dirichlet || cached->with_bc) && (use_dirichlet || cached->without_bc)) {
        cached_wls *cw = use_dirichlet ? cached->with_bc : cached->without_bc;
        int tot = cw->numpts + cw->numbcpts;
        if (items->total < tot) {
            items->total = tot + 64;
            items->ptr = realloc(items->ptr, sizeof *items->ptr * (size_t)items->total);
        }
        for (int i = 0; i < cw->numpts; i++) {
            items->ptr[i].id = cw->ids[i];
            w[i] = cw->w[i];
            items->ptr[i].type = 0;
        }
        for (int i = 0; i < cw->numbcpts; i++) {
            items->ptr[i + cw->numpts].id = cw->ids[i + cw->numpts];
            w[i + cw->numpts] = cw->w[i + cw->numpts];
            items->ptr[i + cw->numpts].bcpoint = cw->bcpoints[i];
            items->ptr[i + cw->numpts].type = 1;
        }
        items->numpts = tot;
    } else {
        if (FLT_EQ(delta, 0.0)) {
            for (int i = 0; i < d->ntrees; i++) {
                hcell *c = hfo_cell_with_point(d->trees[i], x);
                if (c) {
                    Point dl;
                    hfo_delta(c, dl);
                    delta = dl[0];
                    for (int k = 1; k < DIM; k++) delta *= dl[k];
                    delta = pow(delta, 1.0 / DIM);
                    break;
                }
            }
            if (FLT_EQ(delta, 0.0)) {
                hcell *c = hfo_first_leaf(d->trees[0]);
                Point dl;
                hfo_delta(c, dl);
                delta = dl[0];
                for (int k = 1; k < DIM; k++) delta *= dl[k];
                delta = pow(delta, 1.0 / DIM);
            }
        }
        items->numpts = 0;
        items->max_dist = 0.0;
        Point lo, hi;
        for (int k = 0; k < DIM; k++) {
            lo[k] = x[k] - 10.0 * delta;
            hi[k] = x[k] + 10.0 * delta;
        }
        search_box(d, x, lo, hi, items);
        if (use_dirichlet) {
            for (int i = 0; i < d->ndir; i++) {
                hboundary *sb = d->dir[i];
                search_cells_in_tree_box(sb->tree, &sb->map, x, lo, hi, d->inter.maxpts, items,
                                         sb, 1);
            }
        }
        if (items->numpts > 0) hfo_msort_items(items->ptr, items->numpts);
        if (items->numpts > d->inter.maxpts) items->numpts = d->inter.maxpts;
        hfo_wls_calc(interp_wls(&d->inter), items->numpts, items->ptr, x, w);
        if (d->use_cache) {
            int nbc = 0, nnon = 0;
            for (int i = 0; i < items->numpts; i++) {
                if (FLT_EQ(w[i], 0.0)) continue;
                if (items->ptr[i].type == 1) nbc++;
                else nnon++;
            }
            cached_wls *cw = calloc(1, sizeof *cw);
            cw->numpts = nnon;
            cw->numbcpts = nbc;
            cw->ids = malloc(sizeof(int) * (size_t)(nnon + nbc + 1));
            cw->w = malloc(sizeof(real) * (size_t)(nnon + nbc + 1));
            cw->bcpoints = malloc(sizeof(hboundary *) * (size_t)(nbc + 1));
            nnon = 0;
            nbc = 0;
            for (int i = 0; i < items->numpts; i++) {
                if (FLT_EQ(w[i], 0.0)) continue;
                if (items->ptr[i].type == 0) {
                    cw->ids[nnon] = items->ptr[i].id;
                    cw->w[nnon] = w[i];
                    nnon++;
                } else if (items->ptr[i].type == 1) {
                    cw->ids[cw->numpts + nbc] = items->ptr[i].id;
                    cw->w[cw->numpts + nbc] = w[i];
                    cw->bcpoints[nbc] = items->ptr[i].bcpoint;
                    nbc++;
                }
            }
            if (cached) {
                if (use_dirichlet) cached->with_bc = cw;
                else cached->without_bc = cw;
            } else {
                cached = cache_find(&d->cache, x, true);
                if (use_dirichlet) {
                    cached->with_bc = cw;
                    cached->without_bc = (nbc == 0) ? cw : NULL;
                } else {
                    cached->with_bc = NULL;
                    cached->without_bc = cw;
                }
            }
        }
    }
    for (int i = 0; i < items->numpts; i++) {
        if (items->ptr[i].type == 1) {
            hboundary *sb = items->ptr[i].bcpoint;
            stn->rhs += -alpha * w[i] * sb->vals[items->ptr[i].id];
        } else {
            stn_add_elem(stn, items->ptr[i].id, alpha * w[i]);
        }
    }
}

/* higtree/src/domain.c:1321-1379 */
static void get_stencil(hdomain *d, real delta, const real *x, real alpha, hstencil *stn,
                        bool use_dirichlet, bool use_neumann)
{
    bool in_domain;
    if (find_in_center(d, x, alpha, &in_domain, stn)) return;
    if (in_domain) {
        if (dirichlet_find_in_center(d, x, alpha, stn)) return;
        if (use_neumann && get_stencil_neumann_boundary(d, x, alpha, stn)) return;
        if (use_dirichlet && get_stencil_dirichlet_boundary(d, x, alpha, stn)) return;
    } else {
        if (use_neumann && get_stencil_neumann(d, x, alpha, stn)) return;
        if (use_dirichlet && get_stencil_dirichlet(d, x, alpha, stn)) return;
    }
    get_stencil_interpolate(d, x, delta, alpha, stn, use_dirichlet);
}

/* higtree/src/domain.c:1429-1433 */
void hfo_sd_get_stencil(hdomain *d, const Point center, const Point x, real alpha,
                        hstencil *stn)
{
    real delta = hfo_distance(center, x);
    get_stencil(d, delta, x, alpha, stn, true, true);
}

/* higtree/src/domain.c:1604-1623 */
void hfo_sfd_get_stencil(hfacetdomain *f, const Point center, const Point x, real alpha,
                         hstencil *stn)
{
    real delta = hfo_distance(center, x);
    if (FLT_EQ(delta, 0.0)) {
        for (int i = 0; i < f->cdom->ntrees; i++) {
            hcell *c = hfo_cell_with_point(f->cdom->trees[i], x);
            if (c) {
                Point dl;
                hfo_delta(c, dl);
                delta = dl[0];
                break;
            }
        }
    }
    get_stencil(f->cdom, delta, x, alpha, stn, true, true);
}
