/* Broad phase restatement: conservative AABBs, Morton codes, LBVH build
 * (Apetrei 2014 as in lbvh.cpp) and candidate traversal.
 * TEST INFRASTRUCTURE ONLY (see ipc_oracle.h). */
#include "ipco_internal.h"

#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void ipco_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int ipco_get_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- boxes (aabb.cpp:35-126) --------------------------------------------- */
typedef struct box64 {
    double lo[3], hi[3];
} box64;

/* AABB::from_point + conservative_inflation: nextafter(v -/+ r, -/+inf) */
static void point_box(const double *p, double r, box64 *b)
{
    for (int k = 0; k < 3; k++) {
        b->lo[k] = nextafter(p[k] - r, -INFINITY);
        b->hi[k] = nextafter(p[k] + r, INFINITY);
    }
}

static void box_union(const box64 *a, const box64 *b, box64 *r)
{
    for (int k = 0; k < 3; k++) {
        r->lo[k] = b->lo[k] < a->lo[k] ? b->lo[k] : a->lo[k];
        r->hi[k] = a->hi[k] < b->hi[k] ? b->hi[k] : a->hi[k];
    }
}

static box64 *vertex_boxes64(const double *V0, const double *V1, int nV,
                             double r)
{
    box64 *vb = (box64 *)malloc(sizeof(box64) * (size_t)(nV > 0 ? nV : 1));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nV; i++) {
        point_box(V0 + 3 * (size_t)i, r, &vb[i]);
        if (V1) {
            box64 b1;
            point_box(V1 + 3 * (size_t)i, r, &b1);
            box_union(&vb[i], &b1, &vb[i]);
        }
    }
    return vb;
}

static box64 *element_boxes64(const box64 *vb, const int32_t *I, int n,
                              int arity)
{
    box64 *eb = (box64 *)malloc(sizeof(box64) * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        box64 b = vb[I[(size_t)arity * i]];
        for (int j = 1; j < arity; j++)
            box_union(&b, &vb[I[(size_t)arity * i + j]], &b);
        eb[i] = b;
    }
    return eb;
}

/* lbvh.cpp:29-41 assign_inflated_aabb */
static void box_to_f32(const box64 *b, float *f)
{
    for (int k = 0; k < 3; k++) {
        f[k] = nextafterf((float)b->lo[k], -INFINITY);
        f[3 + k] = nextafterf((float)b->hi[k], INFINITY);
    }
}

void ipco_build_boxes_f32(const double *V0, const double *V1, int nV,
                          const int32_t *E, int nE, const int32_t *F, int nF,
                          double inflation, float *vbox, float *ebox,
                          float *fbox)
{
    box64 *vb = vertex_boxes64(V0, V1, nV, inflation);
    if (vbox)
        for (int i = 0; i < nV; i++) box_to_f32(&vb[i], vbox + 6 * (size_t)i);
    if (ebox && nE > 0) {
        box64 *eb = element_boxes64(vb, E, nE, 2);
        for (int i = 0; i < nE; i++) box_to_f32(&eb[i], ebox + 6 * (size_t)i);
        free(eb);
    }
    if (fbox && nF > 0) {
        box64 *fb = element_boxes64(vb, F, nF, 3);
        for (int i = 0; i < nF; i++) box_to_f32(&fb[i], fbox + 6 * (size_t)i);
        free(fb);
    }
    free(vb);
}

/* ---- Morton codes (math/morton.hpp:24-63) -------------------------------- */
static uint64_t expand_bits_2(uint64_t v)
{
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

static double clampd(double v, double lo, double hi)
{
    return v < lo ? lo : (hi < v ? hi : v);
}

static uint64_t morton_3d(double x, double y, double z)
{
    const double scale = 2097152.0; /* 1 << 21 */
    x = clampd(x * scale, 0.0, scale - 1);
    y = clampd(y * scale, 0.0, scale - 1);
    z = clampd(z * scale, 0.0, scale - 1);
    return (expand_bits_2((uint64_t)x) << 2) | (expand_bits_2((uint64_t)y) << 1)
        | expand_bits_2((uint64_t)z);
}

/* stable LSD radix sort of (key, val) by key */
void ipco_sort_u64_pairs(uint64_t *key, int64_t *val, int64_t n)
{
    uint64_t *k2 = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *v2 = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    uint64_t all_or = 0;
    for (int64_t i = 0; i < n; i++) all_or |= key[i];
    for (int shift = 0; shift < 64; shift += 8) {
        if (((all_or >> shift) & 0xFF) == 0) continue;
        int64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (int64_t i = 0; i < n; i++) cnt[((key[i] >> shift) & 0xFF) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < n; i++) {
            const int64_t dst = cnt[(key[i] >> shift) & 0xFF]++;
            k2[dst] = key[i];
            v2[dst] = val[i];
        }
        memcpy(key, k2, sizeof(uint64_t) * (size_t)n);
        memcpy(val, v2, sizeof(int64_t) * (size_t)n);
    }
    free(k2);
    free(v2);
}

/* ---- LBVH (lbvh.cpp:99-330) ---------------------------------------------- */
typedef struct lbvh {
    int n;           /* leaves */
    float *box;      /* (2n-1) x 6 */
    int32_t *left;   /* inner: left child; leaf: primitive id */
    int32_t *right;  /* inner: right child (never 0); leaf: 0 */
    int32_t *rml;    /* rightmost leaf (sorted position) per node */
} lbvh;

static void lbvh_free(lbvh *t)
{
    free(t->box);
    free(t->left);
    free(t->right);
    free(t->rml);
    memset(t, 0, sizeof(*t));
}

static int delta(const uint64_t *codes, int n, int i, uint64_t ci, int j)
{
    if (j < 0 || j >= n) return -1;
    const uint64_t cj = codes[j];
    if (ci == cj) return 32 + __builtin_clz((unsigned)(i ^ j));
    return __builtin_clzll(ci ^ cj);
}

static void lbvh_build(const box64 *boxes, int n, const double *mesh_min,
                       const double *mesh_max, lbvh *t)
{
    memset(t, 0, sizeof(*t));
    t->n = n;
    if (n <= 0) return;
    const int nn = 2 * n - 1;
    t->box = (float *)malloc(sizeof(float) * 6 * (size_t)nn);
    t->left = (int32_t *)calloc((size_t)nn, sizeof(int32_t));
    t->right = (int32_t *)calloc((size_t)nn, sizeof(int32_t));
    t->rml = (int32_t *)malloc(sizeof(int32_t) * (size_t)nn);

    uint64_t *codes = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    double width[3];
    for (int k = 0; k < 3; k++) width[k] = mesh_max[k] - mesh_min[k];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double m[3];
        for (int k = 0; k < 3; k++) {
            const double c = 0.5 * (boxes[i].lo[k] + boxes[i].hi[k]);
            m[k] = (c - mesh_min[k]) / width[k];
        }
        codes[i] = morton_3d(m[0], m[1], m[2]);
        ids[i] = i;
    }
    /* The reference sorts by code only with tbb::parallel_sort (ties in any
     * order); this restatement breaks ties by primitive id (stable sort). */
    ipco_sort_u64_pairs(codes, ids, n);

    const int LEAF = n - 1;
    int *visits = (int *)calloc((size_t)nn, sizeof(int));
    int32_t *lrange = (int32_t *)malloc(sizeof(int32_t) * (size_t)nn);
    int32_t *rrange = (int32_t *)malloc(sizeof(int32_t) * (size_t)nn);
    int root = (n == 1) ? 0 : -1;

#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        box_to_f32(&boxes[ids[i]], t->box + 6 * (size_t)(LEAF + i));
        t->left[LEAF + i] = (int32_t)ids[i];
        t->right[LEAF + i] = 0;
        t->rml[LEAF + i] = i;
        if (n == 1) continue;
        int lk = i, rk = i, cur = LEAF + i;
        for (;;) {
            const int is_a = (lk == 0)
                || (rk != n - 1
                    && delta(codes, n, rk, codes[rk], rk + 1)
                        > delta(codes, n, lk - 1, codes[lk - 1], lk));
            const int parent = is_a ? rk : lk - 1;
            if (is_a) {
                __atomic_store_n(&t->left[parent], cur, __ATOMIC_RELAXED);
                __atomic_store_n(&lrange[parent], lk, __ATOMIC_RELAXED);
            } else {
                __atomic_store_n(&t->right[parent], cur, __ATOMIC_RELAXED);
                __atomic_store_n(&rrange[parent], rk, __ATOMIC_RELAXED);
            }
            if (__atomic_fetch_add(&visits[parent], 1, __ATOMIC_ACQ_REL) == 0)
                break;
            const int l = t->left[parent], r = t->right[parent];
            const float *bl = t->box + 6 * (size_t)l, *br = t->box + 6 * (size_t)r;
            float *bp = t->box + 6 * (size_t)parent;
            for (int k = 0; k < 3; k++) {
                bp[k] = br[k] < bl[k] ? br[k] : bl[k];
                bp[3 + k] = bl[3 + k] < br[3 + k] ? br[3 + k] : bl[3 + k];
            }
            t->rml[parent] = t->rml[l] > t->rml[r] ? t->rml[l] : t->rml[r];
            lk = lrange[parent];
            rk = rrange[parent];
            cur = parent;
            if (lk == 0 && rk == n - 1) {
                root = parent;
                break;
            }
        }
    }
    /* move the root to index 0 (lbvh.cpp:306-329) */
    if (root > 0) {
        float tb[6];
        memcpy(tb, t->box, sizeof(tb));
        memcpy(t->box, t->box + 6 * (size_t)root, sizeof(tb));
        memcpy(t->box + 6 * (size_t)root, tb, sizeof(tb));
        int32_t x;
        x = t->left[0]; t->left[0] = t->left[root]; t->left[root] = x;
        x = t->right[0]; t->right[0] = t->right[root]; t->right[root] = x;
        x = t->rml[0]; t->rml[0] = t->rml[root]; t->rml[root] = x;
        for (int i = 0; i < nn; i++)
            if (t->right[i] != 0 && t->left[i] == 0) t->left[i] = root;
    }
    free(visits);
    free(lrange);
    free(rrange);
    free(codes);
    free(ids);
}

static inline int box_overlap(const float *a, const float *b)
{
    return a[0] <= b[3] && a[1] <= b[4] && a[2] <= b[5] && b[0] <= a[3]
        && b[1] <= a[4] && b[2] <= a[5];
}

/* ---- candidate filters (lbvh.cpp:810-873, collision_filter.hpp) ---------- */
typedef struct filt {
    const int32_t *E, *F, *group;
    const int32_t *vmap, *emap; /* subset -> mesh ids (codim lists) */
} filt;

static inline int can_vv(const filt *f, int a, int b)
{
    return f->group == NULL || f->group[a] != f->group[b];
}

static int can_ee(const filt *f, int ea, int eb)
{
    const int a0 = f->E[2 * (size_t)ea], a1 = f->E[2 * (size_t)ea + 1];
    const int b0 = f->E[2 * (size_t)eb], b1 = f->E[2 * (size_t)eb + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return 0;
    return can_vv(f, a0, b0) || can_vv(f, a0, b1) || can_vv(f, a1, b0)
        || can_vv(f, a1, b1);
}

/* lbvh.cpp:800-807 can_edge_vertex_collide */
static int can_ev(const filt *f, int ei, int vi)
{
    const int e0 = f->E[2 * (size_t)ei], e1 = f->E[2 * (size_t)ei + 1];
    if (vi == e0 || vi == e1) return 0;
    return can_vv(f, vi, e0) || can_vv(f, vi, e1);
}

static int can_fv(const filt *f, int fi, int vi)
{
    const int f0 = f->F[3 * (size_t)fi], f1 = f->F[3 * (size_t)fi + 1],
              f2 = f->F[3 * (size_t)fi + 2];
    if (vi == f0 || vi == f1 || vi == f2) return 0;
    return can_vv(f, vi, f0) || can_vv(f, vi, f1) || can_vv(f, vi, f2);
}

/* lbvh.cpp:836-849 can_edge_face_collide */
static int can_ef(const filt *f, int ei, int fi)
{
    const int e0 = f->E[2 * (size_t)ei], e1 = f->E[2 * (size_t)ei + 1];
    const int f0 = f->F[3 * (size_t)fi], f1 = f->F[3 * (size_t)fi + 1],
              f2 = f->F[3 * (size_t)fi + 2];
    if (e0 == f0 || e0 == f1 || e0 == f2 || e1 == f0 || e1 == f1 || e1 == f2)
        return 0;
    return can_vv(f, e0, f0) || can_vv(f, e0, f1) || can_vv(f, e0, f2)
        || can_vv(f, e1, f0) || can_vv(f, e1, f1) || can_vv(f, e1, f2);
}

typedef struct pairvec {
    int64_t n, cap;
    int32_t *ij;
} pairvec;

static void pv_push(pairvec *v, int32_t a, int32_t b)
{
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->ij = (int32_t *)realloc(v->ij, sizeof(int32_t) * 2 * (size_t)v->cap);
    }
    v->ij[2 * v->n] = a;
    v->ij[2 * v->n + 1] = b;
    v->n++;
}

/* the pair a query leaf `q` and a target leaf `j` report, if any */
static void emit(const filt *f, int kind, int q, int j, pairvec *out)
{
    switch (kind) {
    case 0: /* EE (query, other) */
        if (can_ee(f, q, j)) pv_push(out, q, j);
        break;
    case 1: /* FV (face, vertex) */
        if (can_fv(f, j, q)) pv_push(out, j, q);
        break;
    case 2: /* codim VV (query, other) in mesh ids */
        if (can_vv(f, f->vmap[q], f->vmap[j])) pv_push(out, f->vmap[q], f->vmap[j]);
        break;
    case 3: /* codim EV (edge, vertex) in mesh ids */
        if (can_ev(f, f->emap[j], f->vmap[q])) pv_push(out, f->emap[j], f->vmap[q]);
        break;
    default: /* EF: face queries against the edge tree, (edge, face) */
        if (can_ef(f, j, q)) pv_push(out, j, q);
    }
}

/* traverse_lbvh (lbvh.cpp:368-455): kinds 0 (EE) and 2 (codim VV) are
 * triangular self traversals; 1 (FV) and 3 (codim EV) query vertex leaves
 * against the face / codim-edge tree (swap order); 4 (EF) queries face
 * leaves against the edge tree (lbvh.cpp:771-784). */
static void traverse(const float *qbox, int qprim, int qpos, const lbvh *t,
                     int kind, const filt *f, pairvec *out)
{
    const int nn = 2 * t->n - 1;
    const int tri = kind == 0 || kind == 2;
    if (nn == 1) {
        if (tri) return;
        if (box_overlap(t->box, qbox)) emit(f, kind, qprim, t->left[0], out);
        return;
    }
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    int node = 0;
    do {
        const int l = t->left[node], r = t->right[node];
        int il = box_overlap(t->box + 6 * (size_t)l, qbox);
        int ir = box_overlap(t->box + 6 * (size_t)r, qbox);
        if (tri) {
            if (il && t->rml[l] <= qpos) il = 0;
            if (ir && t->rml[r] <= qpos) ir = 0;
        }
        const int leaf_l = t->right[l] == 0, leaf_r = t->right[r] == 0;
        if (il && leaf_l) emit(f, kind, qprim, t->left[l], out);
        if (ir && leaf_r) emit(f, kind, qprim, t->left[r], out);
        const int tl = il && !leaf_l, tr = ir && !leaf_r;
        if (!tl && !tr) {
            node = stack[--sp];
        } else {
            node = tl ? l : r;
            if (tl && tr) {
                if (sp >= 128) {
                    fprintf(stderr, "ipco: traversal stack overflow\n");
                    abort();
                }
                stack[sp++] = r;
            }
        }
    } while (node != 0);
}

static void detect(const lbvh *src, const lbvh *dst, int kind, const filt *f,
                   ipco_pairs *out)
{
    out->n = 0;
    out->ij = NULL;
    if (src->n == 0 || dst->n == 0) return;
    if ((kind == 0 || kind == 2) && dst->n <= 1) return;
    const int LEAF = src->n - 1;
    int nt = ipco_get_max_threads();
    pairvec *loc = (pairvec *)calloc((size_t)nt, sizeof(pairvec));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < src->n; i++)
            traverse(src->box + 6 * (size_t)(LEAF + i), src->left[LEAF + i], i,
                     dst, kind, f, &loc[tid]);
    }
    int64_t tot = 0;
    for (int k = 0; k < nt; k++) tot += loc[k].n;
    out->n = tot;
    out->ij = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(tot > 0 ? tot : 1));
    int64_t o = 0;
    for (int k = 0; k < nt; k++) {
        if (loc[k].n)
            memcpy(out->ij + 2 * o, loc[k].ij, sizeof(int32_t) * 2 * (size_t)loc[k].n);
        o += loc[k].n;
        free(loc[k].ij);
    }
    free(loc);
}

int ipco_candidates(const double *V0, const double *V1, int nV,
                    const int32_t *E, int nE, const int32_t *F, int nF,
                    double inflation, const int32_t *vertex_group,
                    ipco_pairs *ee, ipco_pairs *fv)
{
    ee->n = fv->n = 0;
    ee->ij = fv->ij = NULL;
    if (nV <= 0) return 0;
    box64 *vb = vertex_boxes64(V0, V1, nV, inflation);
    box64 *eb = nE > 0 ? element_boxes64(vb, E, nE, 2) : NULL;
    box64 *fb = nF > 0 ? element_boxes64(vb, F, nF, 3) : NULL;
    /* broad_phase.cpp:93-125 compute_mesh_aabb over vertex boxes */
    double mn[3] = { INFINITY, INFINITY, INFINITY };
    double mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = 0; i < nV; i++)
        for (int k = 0; k < 3; k++) {
            if (vb[i].lo[k] < mn[k]) mn[k] = vb[i].lo[k];
            if (mx[k] < vb[i].hi[k]) mx[k] = vb[i].hi[k];
        }
    lbvh tv, te, tf;
    lbvh_build(vb, nV, mn, mx, &tv);
    lbvh_build(eb, nE, mn, mx, &te);
    lbvh_build(fb, nF, mn, mx, &tf);
    filt f = { E, F, vertex_group, NULL, NULL };
    detect(&te, &te, 0, &f, ee); /* lbvh.cpp:742-754 */
    detect(&tv, &tf, 1, &f, fv); /* lbvh.cpp:756-769 */
    lbvh_free(&tv);
    lbvh_free(&te);
    lbvh_free(&tf);
    free(vb);
    free(eb);
    free(fb);
    return 0;
}

/* detect_edge_face_candidates (lbvh.cpp:771-784) after a static build
 * (broad_phase.cpp:59-91 boxes, lbvh.cpp:99-131 trees) */
int ipco_edge_face_candidates(const double *V0, int nV, const int32_t *E, int nE,
                              const int32_t *F, int nF, double inflation,
                              const int32_t *vertex_group, ipco_pairs *ef)
{
    ef->n = 0;
    ef->ij = NULL;
    if (nV <= 0 || nE <= 0 || nF <= 0) return 0;
    box64 *vb = vertex_boxes64(V0, NULL, nV, inflation);
    box64 *eb = element_boxes64(vb, E, nE, 2);
    box64 *fb = element_boxes64(vb, F, nF, 3);
    double mn[3] = { INFINITY, INFINITY, INFINITY };
    double mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = 0; i < nV; i++)
        for (int k = 0; k < 3; k++) {
            if (vb[i].lo[k] < mn[k]) mn[k] = vb[i].lo[k];
            if (mx[k] < vb[i].hi[k]) mx[k] = vb[i].hi[k];
        }
    lbvh te, tf;
    lbvh_build(eb, nE, mn, mx, &te);
    lbvh_build(fb, nF, mn, mx, &tf);
    filt f = { E, F, vertex_group, NULL, NULL };
    detect(&tf, &te, 4, &f, ef);
    lbvh_free(&te);
    lbvh_free(&tf);
    free(vb);
    free(eb);
    free(fb);
    return 0;
}

/* ---- codim candidates (candidates.cpp:65-118) ----------------------------
 * Codim vertices: referenced by no edge; codim edges: in no face
 * (collision_mesh.cpp:145-182).  Vertex-vertex pairs among codim vertices
 * (an LBVH over their boxes alone), and (codim edge, codim vertex) pairs.
 * The reference applies mesh.can_collide to indices of its sub-arrays
 * there; this restatement applies it to mesh ids. */
int ipco_codim_lists(int nV, const int32_t *E, int nE, const int32_t *F2E,
                     int nF, int32_t *cv, int32_t *ce, int32_t *counts)
{
    char *used = (char *)calloc((size_t)(nV > 0 ? nV : 1), 1);
    for (int64_t i = 0; i < 2 * (int64_t)nE; i++) used[E[i]] = 1;
    int n = 0;
    for (int i = 0; i < nV; i++)
        if (!used[i]) {
            if (cv) cv[n] = i;
            n++;
        }
    counts[0] = n;
    free(used);
    char *fe = (char *)calloc((size_t)(nE > 0 ? nE : 1), 1);
    for (int64_t i = 0; i < 3 * (int64_t)nF; i++) fe[F2E[i]] = 1;
    n = 0;
    for (int i = 0; i < nE; i++)
        if (!fe[i]) {
            if (ce) ce[n] = i;
            n++;
        }
    counts[1] = n;
    free(fe);
    return 0;
}

int ipco_codim_candidates(const double *V0, const double *V1, int nV,
                          const int32_t *E, int nE, const int32_t *F, int nF,
                          const int32_t *F2E, double inflation,
                          const int32_t *vertex_group, ipco_pairs *vv,
                          ipco_pairs *ev)
{
    vv->n = ev->n = 0;
    vv->ij = ev->ij = NULL;
    int32_t counts[2];
    ipco_codim_lists(nV, E, nE, F2E, nF, NULL, NULL, counts);
    const int ncv = counts[0], nce = counts[1];
    if (ncv == 0) return 0;
    int32_t *cv = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncv);
    int32_t *ce = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nce > 0 ? nce : 1));
    ipco_codim_lists(nV, E, nE, F2E, nF, cv, ce, counts);
    box64 *vb = vertex_boxes64(V0, V1, nV, inflation);
    box64 *sub = (box64 *)malloc(sizeof(box64) * (size_t)ncv);
    double mn[3] = { INFINITY, INFINITY, INFINITY };
    double mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = 0; i < ncv; i++) {
        sub[i] = vb[cv[i]];
        for (int k = 0; k < 3; k++) {
            if (sub[i].lo[k] < mn[k]) mn[k] = sub[i].lo[k];
            if (mx[k] < sub[i].hi[k]) mx[k] = sub[i].hi[k];
        }
    }
    filt f = { E, F, vertex_group, cv, ce };
    lbvh tv;
    lbvh_build(sub, ncv, mn, mx, &tv);
    detect(&tv, &tv, 2, &f, vv);
    if (nce > 0) {
        box64 *eb = (box64 *)malloc(sizeof(box64) * (size_t)nce);
        for (int i = 0; i < nce; i++) {
            const int e = ce[i];
            box_union(&vb[E[2 * (size_t)e]], &vb[E[2 * (size_t)e + 1]], &eb[i]);
            for (int k = 0; k < 3; k++) {
                if (eb[i].lo[k] < mn[k]) mn[k] = eb[i].lo[k];
                if (mx[k] < eb[i].hi[k]) mx[k] = eb[i].hi[k];
            }
        }
        lbvh te;
        lbvh_build(eb, nce, mn, mx, &te);
        detect(&tv, &te, 3, &f, ev);
        lbvh_free(&te);
        free(eb);
    }
    lbvh_free(&tv);
    free(sub);
    free(vb);
    free(cv);
    free(ce);
    return 0;
}

/* ---- sweep and prune (sweep_and_prune.cpp:22-249) ------------------------
 * The reference delegates to scalable-ccd's sort_and_sweep (c80af01, not in
 * the reference tree).  Restated behaviour: FP64 boxes (the conservative
 * boxes above), sorted by (min along the sweep axis, id); a pair is
 * reported when its boxes overlap on all three axes and it passes the
 * adjacency and patch filters (sweep_and_prune.cpp:186-249).  The sweep axis
 * is the axis of largest mesh-box extent (first on ties).  Edge-edge pairs
 * are reported as (earlier, later) in the sorted order; face-vertex pairs
 * as (face, vertex). */
typedef struct skey {
    double key;
    int32_t id;
} skey;

static int cmp_skey(const void *a, const void *b)
{
    const skey *x = (const skey *)a, *y = (const skey *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

static inline int overlap64(const box64 *a, const box64 *b)
{
    return a->lo[0] <= b->hi[0] && a->lo[1] <= b->hi[1] && a->lo[2] <= b->hi[2]
        && b->lo[0] <= a->hi[0] && b->lo[1] <= a->hi[1] && b->lo[2] <= a->hi[2];
}

static skey *sorted_keys(const box64 *b, int n, int axis)
{
    skey *k = (skey *)malloc(sizeof(skey) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        k[i].key = b[i].lo[axis];
        k[i].id = i;
    }
    qsort(k, (size_t)n, sizeof(skey), cmp_skey);
    return k;
}

static void gather_pairs(pairvec *loc, int nt, ipco_pairs *out)
{
    int64_t tot = 0;
    for (int k = 0; k < nt; k++) tot += loc[k].n;
    out->n = tot;
    out->ij = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(tot > 0 ? tot : 1));
    int64_t o = 0;
    for (int k = 0; k < nt; k++) {
        if (loc[k].n)
            memcpy(out->ij + 2 * o, loc[k].ij, sizeof(int32_t) * 2 * (size_t)loc[k].n);
        o += loc[k].n;
        free(loc[k].ij);
    }
    free(loc);
}

int ipco_sap_axis(const double *V0, const double *V1, int nV, double inflation)
{
    box64 *vb = vertex_boxes64(V0, V1, nV, inflation);
    double mn[3] = { INFINITY, INFINITY, INFINITY };
    double mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = 0; i < nV; i++)
        for (int k = 0; k < 3; k++) {
            if (vb[i].lo[k] < mn[k]) mn[k] = vb[i].lo[k];
            if (mx[k] < vb[i].hi[k]) mx[k] = vb[i].hi[k];
        }
    free(vb);
    int axis = 0;
    for (int k = 1; k < 3; k++)
        if (mx[k] - mn[k] > mx[axis] - mn[axis]) axis = k;
    return axis;
}

int ipco_candidates_sap(const double *V0, const double *V1, int nV,
                        const int32_t *E, int nE, const int32_t *F, int nF,
                        double inflation, const int32_t *vertex_group,
                        ipco_pairs *ee, ipco_pairs *fv)
{
    ee->n = fv->n = 0;
    ee->ij = fv->ij = NULL;
    if (nV <= 0) return 0;
    const int axis = ipco_sap_axis(V0, V1, nV, inflation);
    box64 *vb = vertex_boxes64(V0, V1, nV, inflation);
    box64 *eb = element_boxes64(vb, E, nE, 2);
    box64 *fb = element_boxes64(vb, F, nF, 3);
    filt f = { E, F, vertex_group, NULL, NULL };
    const int nt = ipco_get_max_threads();
    /* edge-edge: one list */
    skey *ek = sorted_keys(eb, nE, axis);
    pairvec *loc = (pairvec *)calloc((size_t)nt, sizeof(pairvec));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < nE; i++) {
            const box64 *a = &eb[ek[i].id];
            for (int j = i + 1; j < nE && ek[j].key <= a->hi[axis]; j++)
                if (overlap64(a, &eb[ek[j].id]) && can_ee(&f, ek[i].id, ek[j].id))
                    pv_push(&loc[tid], ek[i].id, ek[j].id);
        }
    }
    gather_pairs(loc, nt, ee);
    /* face-vertex: two lists; a pair is found from the box that starts
     * first along the axis (ties: from the face) */
    skey *fk = sorted_keys(fb, nF, axis);
    skey *vk = sorted_keys(vb, nV, axis);
    loc = (pairvec *)calloc((size_t)nt, sizeof(pairvec));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < nF; i++) {
            const box64 *a = &fb[fk[i].id];
            int lo = 0, hi = nV; /* first vertex with key >= face min */
            while (lo < hi) {
                const int m = (lo + hi) / 2;
                if (vk[m].key < a->lo[axis]) lo = m + 1;
                else hi = m;
            }
            for (int j = lo; j < nV && vk[j].key <= a->hi[axis]; j++)
                if (overlap64(a, &vb[vk[j].id]) && can_fv(&f, fk[i].id, vk[j].id))
                    pv_push(&loc[tid], fk[i].id, vk[j].id);
        }
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < nV; i++) {
            const box64 *a = &vb[vk[i].id];
            int lo = 0, hi = nF; /* first face with key > vertex min */
            while (lo < hi) {
                const int m = (lo + hi) / 2;
                if (fk[m].key <= a->lo[axis]) lo = m + 1;
                else hi = m;
            }
            for (int j = lo; j < nF && fk[j].key <= a->hi[axis]; j++)
                if (overlap64(a, &fb[fk[j].id]) && can_fv(&f, fk[j].id, vk[i].id))
                    pv_push(&loc[tid], fk[j].id, vk[i].id);
        }
    }
    gather_pairs(loc, nt, fv);
    free(ek);
    free(fk);
    free(vk);
    free(vb);
    free(eb);
    free(fb);
    return 0;
}

void ipco_pairs_free(ipco_pairs *p)
{
    free(p->ij);
    p->ij = NULL;
    p->n = 0;
}

int ipco_faces_to_edges(const int32_t *E, int nE, const int32_t *F, int nF,
                        int32_t *F2E)
{
    /* key (min,max) -> edge id via sort + binary search */
    uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nE > 0 ? nE : 1));
    int64_t *id = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nE > 0 ? nE : 1));
    for (int i = 0; i < nE; i++) {
        uint32_t a = (uint32_t)E[2 * i], b = (uint32_t)E[2 * i + 1];
        if (a > b) { uint32_t t = a; a = b; b = t; }
        key[i] = ((uint64_t)a << 32) | b;
        id[i] = i;
    }
    ipco_sort_u64_pairs(key, id, nE);
    int rc = 0;
    for (int f = 0; f < nF; f++)
        for (int j = 0; j < 3; j++) {
            uint32_t a = (uint32_t)F[3 * f + j], b = (uint32_t)F[3 * f + (j + 1) % 3];
            if (a > b) { uint32_t t = a; a = b; b = t; }
            const uint64_t k = ((uint64_t)a << 32) | b;
            int64_t lo = 0, hi = nE;
            while (lo < hi) {
                const int64_t mid = (lo + hi) / 2;
                if (key[mid] < k) lo = mid + 1; else hi = mid;
            }
            if (lo < nE && key[lo] == k) F2E[3 * f + j] = (int32_t)id[lo];
            else { F2E[3 * f + j] = -1; rc = -1; }
        }
    free(key);
    free(id);
    return rc;
}
