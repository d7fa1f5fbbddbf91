/*
 * CPU oracle for the HiG-Flow projection time step -- TEST INFRASTRUCTURE.
 *
 * This is a plain-C restatement of the reference algorithm (HiG-Tree mesh,
 * stencil builder, WLS interpolation, Newtonian/viscoelastic projection step,
 * BiCGStab+ILU(0) pressure solve).  It exists only to check the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference root).  DIM is a compile-time constant exactly as in
 * higtree/src/dim.h:28-31; the Makefile builds a 2D and a 3D library.
 *
 * Pinning status (what checks each part, see tests/test_oracle_*.py):
 *   wls_*            pinned: bit-compared with the reference's own wls.c built
 *                    into oracle/_ref, ATF known answers, polynomial
 *                    reproduction.
 *   tree / iterators pinned: leaf and facet counts, unknown-facet set and
 *                    ordering against the reference-held Poiseuille fixture
 *                    (higflow/example2d_Newt/output), brute-force lookups.
 *   stencils         pinned: identity/ghost closed forms, polynomial
 *                    reproduction of every WLS row; candidate tie-breaks
 *                    restate glibc-2.35 msort (parity unpinned for ties
 *                    beyond the reproduction checks).
 *   step             pinned: zero-field fixed point, projection divergence,
 *                    Poiseuille steady state from the reference fixture.
 *   constitutive     pinned: closed-form 2x2 eigen, Oldroyd-B simple-shear
 *                    steady state, log/conformation equivalence.
 */
#ifndef HFO_INTERNAL_H
#define HFO_INTERNAL_H

#include <stdio.h>
#include <stdlib.h>
#include <stdbool.h>
#include <string.h>
#include <math.h>
#include <float.h>
#include <stdint.h>

#ifndef DIM
#define DIM 2
#endif

typedef double real;
typedef real Point[DIM];

/* higtree/src/utils.h:16,25,33-44 */
#define EPSDELTA 1.0e-8
#define EPSMACH 1.0e-12
#define POS_GE(x, y) ((x) > ((y) - EPSDELTA))
#define POS_LE(x, y) ((x) < ((y) + EPSDELTA))
#define FLT_GT(x, y) ((x) > ((y) + EPSMACH))
#define FLT_LT(x, y) ((x) < ((y) - EPSMACH))
#define FLT_LE(x, y) ((x) < ((y) + EPSMACH))
#define FLT_EQ(x, y) (fabs((x) - (y)) < EPSMACH)
#define FLT_NE(x, y) (!FLT_EQ(x, y))

#define HFO_DIRICHLET 0
#define HFO_NEUMANN 1

/* ---------------------------------------------------------------- tree */
/* hig_cell (higtree/src/higtree.h:59-67): box, children grid, ids */
typedef struct hcell {
    Point lo, hi;
    int nc[DIM];
    struct hcell *parent;
    int pos;
    struct hcell **ch;
    int cid;
    int fid[2 * DIM]; /* index dim*2+dir */
} hcell;

typedef struct hfacet {
    hcell *c;
    int dim, dir;
} hfacet;

typedef struct {
    hcell **v;
    int n, cap;
} cellvec;

typedef struct {
    hfacet *v;
    int n, cap;
} facetvec;

typedef struct idgen {
    int next;
} idgen;

int hfo_toposition(const int nc[DIM], const int pp[DIM]);
void hfo_tobase(int pos, const int nc[DIM], int pp[DIM]);
hcell *hfo_create_root(idgen *g, const Point lo, const Point hi);
void hfo_refine_uniform(idgen *g, hcell *c, const int nc[DIM]);
void hfo_destroy_tree(hcell *c);
void hfo_center(const hcell *c, Point p);
void hfo_delta(const hcell *c, Point p);
hcell *hfo_cell_with_point(hcell *top, const Point x);
hcell *hfo_cell_with_point_from_another(hcell *c, const Point x);
hcell *hfo_first_leaf(hcell *c);
int hfo_narrowest_dim(const hcell *c);
void hfo_facet_center(const hfacet *f, Point p);
bool hfo_facet_with_point(hcell *root, int dim, const Point x, hfacet *f);
bool hfo_intersect_bbox(const hcell *c, const Point lo, const Point hi);
bool hfo_facet_intersect_bbox(const hfacet *f, const Point lo, const Point hi);
void cellvec_push(cellvec *v, hcell *c);
void facetvec_push(facetvec *v, hfacet f);
void hfo_leaves(hcell *root, cellvec *out);
void hfo_bbox_cells(hcell *root, const Point lo, const Point hi, cellvec *out);
void hfo_bbox_facets(hcell *root, const int dimofinterest[DIM], const Point lo,
                     const Point hi, facetvec *out);
hcell *hfo_read_amr(idgen *g, const char *text);
real hfo_distance(const Point a, const Point b);
bool hfo_point_equal(const Point a, const Point b);
bool hfo_rect_contains(const Point lo, const Point hi, const Point x);

/* ---------------------------------------------------------------- maps */
typedef struct {
    int *lid; /* indexed by unique id, value lid+1 (0 = undefined) */
    int size;
    int count;
} idmap;

void idmap_init(idmap *m, int maxid);
int idmap_lookup(const idmap *m, int id);
void idmap_assign(idmap *m, int id);

/* --------------------------------------------------------- boundaries */
/* sim_boundary (higtree/src/domain.h:29-38) */
typedef struct {
    hcell *tree;
    int type;
    int userid;
    idmap map;
    int nvals;
    real *vals;
    hcell **cells; /* leaves in map order */
} hboundary;

/* --------------------------------------------------------- stencils */
/* sim_stencil (higtree/src/domain.h:303-309) */
typedef struct {
    int n, cap;
    int *ids;
    real *vals;
    real rhs;
} hstencil;

void stn_init(hstencil *s);
void stn_free(hstencil *s);
void stn_reset(hstencil *s);
void stn_add_elem(hstencil *s, int id, real v);

/* ------------------------------------------------------------- WLS */
typedef struct {
    Point x;
    real dist;
    hboundary *bcpoint;
    int id;
    unsigned char type;
} witem;

typedef struct {
    int dim, ord, maxnumpts, numpoly;
    real **A, **Q;
    real *b, *D, *w;
} hwls;

int hfo_wls_num_min_points(int dim, int ord);
hwls *hfo_wls_create(int dim, int ord, int maxnumpts);
void hfo_wls_destroy(hwls *w);
void hfo_wls_calc(hwls *wls, int numpts, const witem *items, const real *x, real *wout);

typedef struct {
    int dim, order, maxpts, numminpoints;
    hwls *wls;
} interp;

/* point-keyed cache of WLS choices (higtree/src/point-mapper.c, domain.c:489-670) */
typedef struct cached_wls {
    int numpts, numbcpts;
    int *ids;
    real *w;
    hboundary **bcpoints;
} cached_wls;

typedef struct cache_entry {
    int64_t key[DIM];
    cached_wls *with_bc, *without_bc;
    struct cache_entry *next;
} cache_entry;

typedef struct {
    cache_entry **buckets;
    int nbuckets;
    int count;
} pcache;

/* --------------------------------------------------------- domains */
struct hfacetdomain;
/* sim_domain (higtree/src/domain.h:84-107) */
typedef struct hdomain {
    hcell **trees;
    int ntrees;
    hboundary **dir;
    int ndir;
    hboundary **neu;
    int nneu;
    interp inter, bc_inter;
    int use_cache;
    pcache cache;
    idmap map;         /* cell id -> local id */
    hcell **cells;     /* local id -> cell */
    int ncells;
    struct hfacetdomain *sfd; /* non-NULL when this is a facet domain's cdom */
} hdomain;

/* sim_facet_domain (higtree/src/domain.h:109-124) */
typedef struct hfacetdomain {
    hdomain *cdom;
    int dim;
    int dimofinterest[DIM];
    idmap fmap;
    hfacet *facets; /* local id -> facet */
    int nfacets;
    unsigned char **sfbi_l, **sfbi_h; /* per tree, per block */
} hfacetdomain;

hdomain *hfo_sd_create(int order);
void hfo_sd_destroy(hdomain *d);
void hfo_sd_add_bc(hdomain *d, hboundary *b);
void hfo_sd_build_map(hdomain *d, int maxid);
hfacetdomain *hfo_sfd_create(int dim, int order, hcell **trees, int ntrees);
void hfo_sfd_destroy(hfacetdomain *f);
void hfo_sfd_adjust_facet_ids(hfacetdomain *f);
void hfo_sfd_compute_sfbi(hfacetdomain *f);
void hfo_sfd_domain_facets(hfacetdomain *f, facetvec *out);
void hfo_sfd_build_map(hfacetdomain *f, int maxid);
hcell *hfo_sd_cell_with_point(hdomain *d, const Point x);
int hfo_sfd_facet_with_point(hfacetdomain *f, const Point x, hfacet *out);
hboundary *hfo_bc_create(hcell *tree, int type, int userid, int maxid);

void hfo_sd_get_stencil(hdomain *d, const Point center, const Point x, real alpha,
                        hstencil *stn);
void hfo_sfd_get_stencil(hfacetdomain *f, const Point center, const Point x,
                         real alpha, hstencil *stn);
real hfo_interpolate(const real *values, const hstencil *stn);

/* qsort restatement (glibc 2.35 msort semantics) */
void hfo_msort_items(witem *a, int n);

#endif
