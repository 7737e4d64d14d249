/*
 * oracle.c -- CPU restatement of the reference hot path (checker only).
 *
 * Test infrastructure: loaded only by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off -fopenmp). Each function names the reference file:line
 * (relative to the reference repository root) whose behaviour it restates.
 */
#include "oracle.h"
#include "fmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* float3 helpers. Operation order is part of the contract: dot is     */
/* ((x*x + y*y) + z*z), cross is the textbook component formula and    */
/* normalize divides each component by sqrtf(dot(v,v)).                */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, z; } f3;

static inline f3 mk(float x, float y, float z) { f3 r = {x, y, z}; return r; }
static inline f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }
static inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline f3 scl(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline f3 normalize(f3 a)
{
    float l = sqrtf(dot(a, a));
    return mk(a.x / l, a.y / l, a.z / l);
}
static inline f3 fmin3(f3 a, f3 b) { return mk(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
static inline f3 fmax3(f3 a, f3 b) { return mk(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }

/* ------------------------------------------------------------------ */
/* RNG                                                                  */
/* ------------------------------------------------------------------ */

/* src/Device/includes/Hash.h:11-19 (Thomas Wang's 32-bit integer hash). */
uint32_t orc_wang_hash(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

/* src/HostDeviceCommon/Xorshift.h:55-63: Marsaglia xorshift32 (13, 17, 5). */
uint32_t orc_xorshift32(uint32_t* state)
{
    uint32_t x = *state;
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    *state = x;
    return x;
}

/* src/HostDeviceCommon/Xorshift.h:45-50: float(x) / float(0xffffffff), and
 * float(0xffffffff) rounds to 2^32; clamped to 1 - 1e-7 (as binary32). */
float orc_rand01(uint32_t* state)
{
    float a = (float)orc_xorshift32(state) / (float)0xffffffffu;
    const float hi = 1.0f - 1.0e-7f;
    return a < hi ? a : hi;
}

/* src/HostDeviceCommon/Xorshift.h:36-40 */
int32_t orc_random_index(uint32_t* state, int32_t n)
{
    int32_t r = (int32_t)((float)orc_xorshift32(state) / (float)0xffffffffu * (float)n);
    return r < n - 1 ? r : n - 1;
}

/* src/Device/kernels/Megakernel.h:52-57 and CameraRays.h:155-160 */
uint32_t orc_pixel_seed(uint32_t pixel_index, uint32_t sample_number, uint32_t random_number,
                        int32_t freeze_random)
{
    if (freeze_random)
        return orc_wang_hash(pixel_index + 1u);
    return orc_wang_hash((pixel_index + 1u) * (sample_number + 1u) * random_number);
}

void orc_xorshift_fill(uint32_t seed, uint32_t* out, int32_t n)
{
    uint32_t s = seed;
    for (int32_t i = 0; i < n; i++)
        out[i] = orc_xorshift32(&s);
}

void orc_rand01_fill(uint32_t seed, float* out, int32_t n)
{
    uint32_t s = seed;
    for (int32_t i = 0; i < n; i++)
        out[i] = orc_rand01(&s);
}

/* ------------------------------------------------------------------ */
/* Moller-Trumbore, src/Renderer/Triangle.h:20-63                       */
/* ------------------------------------------------------------------ */
static inline int mt(f3 a, f3 e1, f3 e2, f3 o, f3 d, float* tt, float* uu, float* vv)
{
    const float EPS = 0.0000001f;
    f3 h = cross(d, e2);
    float det = dot(e1, h);
    if (det > -EPS && det < EPS)
        return 0;
    float f = 1.0f / det;
    f3 s = sub(o, a);
    float u = f * dot(s, h);
    if (u < 0.0f || u > 1.0f)
        return 0;
    f3 q = cross(s, e1);
    float v = f * dot(d, q);
    if (v < 0.0f || u + v > 1.0f)
        return 0;
    float t = f * dot(e2, q);
    if (!(t > EPS))
        return 0;
    *tt = t;
    *uu = u;
    *vv = v;
    return 1;
}

int32_t orc_mt_intersect(const float* a, const float* b, const float* c, const orc_ray* ray,
                         float* t, float* u, float* v)
{
    f3 A = ld3(a), B = ld3(b), C = ld3(c);
    return mt(A, sub(B, A), sub(C, A), ld3(ray->o), ld3(ray->d), t, u, v);
}

/* Candidate ordering used everywhere in this build: smaller t wins, equal t
 * goes to the smaller primitive index. (The reference keeps the first hit
 * found, BVH.h:174-175, which depends on traversal order.) */
static inline int better(float t, int prim, const orc_hit* best)
{
    if (best->prim < 0)
        return 1;
    return t < best->t || (t == best->t && prim < best->prim);
}

static inline void tri_verts(const float* verts, const int32_t* idx, int32_t i, f3* a, f3* b, f3* c)
{
    *a = ld3(verts + 3 * (int64_t)idx[3 * (int64_t)i + 0]);
    *b = ld3(verts + 3 * (int64_t)idx[3 * (int64_t)i + 1]);
    *c = ld3(verts + 3 * (int64_t)idx[3 * (int64_t)i + 2]);
}

static const orc_hit MISS = {-1, -1.0f, 0.0f, 0.0f};

void orc_trace_brute(const float* verts, const int32_t* idx, int32_t ntri, const orc_ray* rays,
                     orc_hit* hits, int32_t nrays, int32_t any_hit, int32_t nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t r = 0; r < nrays; r++)
    {
        const orc_ray* ray = rays + r;
        f3 o = ld3(ray->o), d = ld3(ray->d);
        orc_hit best = MISS;
        for (int32_t i = 0; i < ntri; i++)
        {
            if (i == ray->last_prim)
                continue;
            f3 a, b, c;
            tri_verts(verts, idx, i, &a, &b, &c);
            float t, u, v;
            if (!mt(a, sub(b, a), sub(c, a), o, d, &t, &u, &v))
                continue;
            if (!(t < ray->tmax))
                continue;
            if (better(t, i, &best))
            {
                best.prim = i; best.t = t; best.u = u; best.v = v;
                if (any_hit)
                    break; /* lowest index wins: triangles are scanned in order */
            }
        }
        hits[r] = best;
    }
}

/* ------------------------------------------------------------------ */
/* Octree of k-DOPs: src/Renderer/BVH.h, BVH.cpp, BoundingVolume.h      */
/* ------------------------------------------------------------------ */
#define NPLANES 7

typedef struct {
    float dn[NPLANES], df[NPLANES];
} kdop;

typedef struct {
    f3 mn, mx;          /* insertion cell (BVH.h:254) */
    kdop bv;            /* bounding volume (BVH.h:255) */
    int32_t is_leaf;
    int32_t child[8];
    int32_t* tris;
    int32_t ntris, cap;
} onode;

struct orc_octree {
    onode* nodes;
    int32_t nnodes, cap;
    f3* tri_a;   /* triangle vertices, by primitive index */
    f3* tri_b;
    f3* tri_c;
    int32_t ntri;
    f3 normals[NPLANES];
    int32_t max_depth, leaf_max;
};

/* BVH.cpp:12-20. The values are computed in binary32 as the reference does. */
static void plane_normals(f3* n)
{
    float s = sqrtf(3.0f) / 3.0f;
    n[0] = mk(1, 0, 0);
    n[1] = mk(0, 1, 0);
    n[2] = mk(0, 0, 1);
    n[3] = mk(s, s, s);
    n[4] = mk(-s, s, s);
    n[5] = mk(-s, -s, s);
    n[6] = mk(s, -s, s);
}

static void kdop_empty(kdop* k)
{
    for (int i = 0; i < NPLANES; i++) { k->dn[i] = INFINITY; k->df[i] = -INFINITY; }
}

/* BoundingVolume.h:30-43 then :45-52 */
static void kdop_extend_tri(kdop* k, const f3* n, f3 a, f3 b, f3 c)
{
    f3 v[3] = {a, b, c};
    for (int i = 0; i < NPLANES; i++)
    {
        float dn = INFINITY, df = -INFINITY;
        for (int j = 0; j < 3; j++)
        {
            float dist = dot(n[i], v[j]);
            dn = dist < dn ? dist : dn;
            df = dist > df ? dist : df;
        }
        k->dn[i] = dn < k->dn[i] ? dn : k->dn[i];
        k->df[i] = df > k->df[i] ? df : k->df[i];
    }
}

static void kdop_extend(kdop* k, const kdop* o)
{
    for (int i = 0; i < NPLANES; i++)
    {
        k->dn[i] = o->dn[i] < k->dn[i] ? o->dn[i] : k->dn[i];
        k->df[i] = o->df[i] > k->df[i] ? o->df[i] : k->df[i];
    }
}

/* BoundingVolume.h:106-131. Note: no test against t_far < 0 -- volumes behind
 * the origin are accepted, as in the reference. */
static int kdop_intersect(const kdop* k, const float* denoms, const float* numers, float* t_near_out)
{
    float t_near = -INFINITY, t_far = INFINITY;
    for (int i = 0; i < NPLANES; i++)
    {
        float denom = denoms[i];
        if (denom == 0.0f)
            continue;
        float a = (k->dn[i] - numers[i]) / denom;
        float b = (k->df[i] - numers[i]) / denom;
        if (denom < 0.0f) { float tmp = a; a = b; b = tmp; }
        t_near = t_near > a ? t_near : a;   /* hippt::max */
        t_far = t_far < b ? t_far : b;      /* hippt::min */
        if (t_far < t_near)
            return 0;
    }
    *t_near_out = t_near;
    return 1;
}

static int32_t new_node(orc_octree* T, f3 mn, f3 mx)
{
    if (T->nnodes == T->cap)
    {
        T->cap = T->cap ? T->cap * 2 : 1024;
        T->nodes = (onode*)realloc(T->nodes, sizeof(onode) * (size_t)T->cap);
    }
    onode* n = T->nodes + T->nnodes;
    memset(n, 0, sizeof(*n));
    n->mn = mn;
    n->mx = mx;
    n->is_leaf = 1;
    kdop_empty(&n->bv);
    for (int i = 0; i < 8; i++) n->child[i] = -1;
    return T->nnodes++;
}

static void push_tri(onode* n, int32_t t)
{
    if (n->ntris == n->cap)
    {
        n->cap = n->cap ? n->cap * 2 : 8;
        n->tris = (int32_t*)realloc(n->tris, sizeof(int32_t) * (size_t)n->cap);
    }
    n->tris[n->ntris++] = t;
}

/* BVH.h:74-86. The octant cells of children 2, 4 and 6 are offset from m_min
 * by the parent's middle (m_min + (0, middle_y, 0) etc.), exactly as the
 * reference writes them; only the subdivision cells are affected, the
 * bounding volumes always come from the triangles. */
static void create_children(orc_octree* T, int32_t ni)
{
    f3 mn = T->nodes[ni].mn, mx = T->nodes[ni].mx;
    float mx_ = (mn.x + mx.x) / 2, my = (mn.y + mx.y) / 2, mz = (mn.z + mx.z) / 2;
    int32_t c[8];
    c[0] = new_node(T, mn, mk(mx_, my, mz));
    c[1] = new_node(T, mk(mx_, mn.y, mn.z), mk(mx.x, my, mz));
    c[2] = new_node(T, add(mn, mk(0, my, 0)), mk(mx_, mx.y, mz));
    c[3] = new_node(T, mk(mx_, my, mn.z), mk(mx.x, mx.y, mz));
    c[4] = new_node(T, add(mn, mk(0, 0, mz)), mk(mx_, my, mx.z));
    c[5] = new_node(T, mk(mx_, mn.y, mz), mk(mx.x, my, mx.z));
    c[6] = new_node(T, add(mn, mk(0, my, mz)), mk(mx_, mx.y, mx.z));
    c[7] = new_node(T, mk(mx_, my, mz), mk(mx.x, mx.y, mx.z));
    for (int i = 0; i < 8; i++) T->nodes[ni].child[i] = c[i];
}

static void insert(orc_octree* T, int32_t ni, int32_t tri, int32_t depth);

/* BVH.h:114-130: octant by the triangle's AABB centre against the middle. */
static void insert_to_children(orc_octree* T, int32_t ni, int32_t tri, int32_t depth)
{
    f3 mn = T->nodes[ni].mn, mx = T->nodes[ni].mx;
    float mx_ = (mn.x + mx.x) / 2, my = (mn.y + mx.y) / 2, mz = (mn.z + mx.z) / 2;
    f3 a = T->tri_a[tri], b = T->tri_b[tri], c = T->tri_c[tri];
    f3 lo = fmin3(a, fmin3(b, c)), hi = fmax3(a, fmax3(b, c));
    f3 cen = add(lo, hi);
    cen = mk(cen.x / 2, cen.y / 2, cen.z / 2);   /* Triangle.cpp:8-11 */
    int oct = 0;
    if (cen.x > mx_) oct += 1;
    if (cen.y > my) oct += 2;
    if (cen.z > mz) oct += 4;
    insert(T, T->nodes[ni].child[oct], tri, depth + 1);
}

/* BVH.h:90-112 */
static void insert(orc_octree* T, int32_t ni, int32_t tri, int32_t depth)
{
    int depth_exceeded = T->max_depth != -1 && depth == T->max_depth;
    if (T->nodes[ni].is_leaf || depth_exceeded)
    {
        push_tri(&T->nodes[ni], tri);
        if (T->nodes[ni].ntris > T->leaf_max && !depth_exceeded)
        {
            T->nodes[ni].is_leaf = 0;
            create_children(T, ni);
            int32_t* moved = T->nodes[ni].tris;
            int32_t nmoved = T->nodes[ni].ntris;
            T->nodes[ni].tris = NULL;
            T->nodes[ni].ntris = T->nodes[ni].cap = 0;
            for (int32_t k = 0; k < nmoved; k++)
                insert_to_children(T, ni, moved[k], depth);
            free(moved);
        }
    }
    else
        insert_to_children(T, ni, tri, depth);
}

/* BVH.h:57-67 */
static kdop compute_volume(orc_octree* T, int32_t ni)
{
    kdop k;
    kdop_empty(&k);
    if (T->nodes[ni].is_leaf)
    {
        for (int32_t i = 0; i < T->nodes[ni].ntris; i++)
        {
            int32_t t = T->nodes[ni].tris[i];
            kdop_extend_tri(&k, T->normals, T->tri_a[t], T->tri_b[t], T->tri_c[t]);
        }
    }
    else
    {
        for (int c = 0; c < 8; c++)
        {
            kdop ck = compute_volume(T, T->nodes[ni].child[c]);
            kdop_extend(&k, &ck);
        }
    }
    T->nodes[ni].bv = k;
    return k;
}

/* BVH.cpp:23-64 */
orc_octree* orc_octree_build(const float* verts, const int32_t* idx, int32_t ntri, int32_t max_depth,
                             int32_t leaf_max)
{
    orc_octree* T = (orc_octree*)calloc(1, sizeof(orc_octree));
    T->ntri = ntri;
    T->max_depth = max_depth;
    T->leaf_max = leaf_max;
    plane_normals(T->normals);
    T->tri_a = (f3*)malloc(sizeof(f3) * (size_t)(ntri > 0 ? ntri : 1));
    T->tri_b = (f3*)malloc(sizeof(f3) * (size_t)(ntri > 0 ? ntri : 1));
    T->tri_c = (f3*)malloc(sizeof(f3) * (size_t)(ntri > 0 ? ntri : 1));
    f3 mn = mk(INFINITY, INFINITY, INFINITY), mx = mk(-INFINITY, -INFINITY, -INFINITY);
    for (int32_t i = 0; i < ntri; i++)
    {
        tri_verts(verts, idx, i, &T->tri_a[i], &T->tri_b[i], &T->tri_c[i]);
        mn = fmin3(mn, fmin3(T->tri_a[i], fmin3(T->tri_b[i], T->tri_c[i])));
        mx = fmax3(mx, fmax3(T->tri_a[i], fmax3(T->tri_b[i], T->tri_c[i])));
    }
    int32_t root = new_node(T, mn, mx);
    for (int32_t i = 0; i < ntri; i++)
        insert(T, root, i, 0);
    compute_volume(T, root);
    return T;
}

void orc_octree_free(orc_octree* T)
{
    if (!T)
        return;
    for (int32_t i = 0; i < T->nnodes; i++)
        free(T->nodes[i].tris);
    free(T->nodes);
    free(T->tri_a);
    free(T->tri_b);
    free(T->tri_c);
    free(T);
}

static void stats_rec(const orc_octree* T, int32_t ni, int32_t depth, int64_t* s)
{
    const onode* n = T->nodes + ni;
    s[0]++;
    if (depth > s[2]) s[2] = depth;
    if (n->is_leaf) { s[1]++; s[3] += n->ntris; return; }
    for (int c = 0; c < 8; c++) stats_rec(T, n->child[c], depth + 1, s);
}

void orc_octree_stats(const orc_octree* T, int64_t* s)
{
    s[0] = s[1] = s[2] = s[3] = 0;
    if (T->nnodes) stats_rec(T, 0, 0, s);
}

/* Small binary min-heap standing in for the std::priority_queue of
 * BVH.h:184-213 (at most 8 entries per interior node). */
typedef struct { int32_t node; float t; } qel;

static void heap_push(qel* h, int* n, qel e)
{
    int i = (*n)++;
    h[i] = e;
    while (i > 0)
    {
        int p = (i - 1) / 2;
        if (!(h[p].t > h[i].t)) break;
        qel tmp = h[p]; h[p] = h[i]; h[i] = tmp;
        i = p;
    }
}

static qel heap_pop(qel* h, int* n)
{
    qel top = h[0];
    h[0] = h[--(*n)];
    int i = 0;
    for (;;)
    {
        int l = 2 * i + 1, r = l + 1, m = i;
        if (l < *n && h[l].t < h[m].t) m = l;
        if (r < *n && h[r].t < h[m].t) m = r;
        if (m == i) break;
        qel tmp = h[m]; h[m] = h[i]; h[i] = tmp;
        i = m;
    }
    return top;
}

/* BVH.h:155-223 (recursive front-to-back traversal with early exit). */
static int node_intersect(const orc_octree* T, int32_t ni, f3 o, f3 d, int32_t last_prim,
                          orc_hit* best, float* t_near, const float* denoms, const float* numers)
{
    const onode* n = T->nodes + ni;
    float trash;
    if (!kdop_intersect(&n->bv, denoms, numers, &trash))
        return 0;
    if (n->is_leaf)
    {
        for (int32_t k = 0; k < n->ntris; k++)
        {
            int32_t tri = n->tris[k];
            f3 a = T->tri_a[tri];
            float t, u, v;
            if (!mt(a, sub(T->tri_b[tri], a), sub(T->tri_c[tri], a), o, d, &t, &u, &v))
                continue;
            if (t < 0.0f)               /* ray.minT = 0 (BVH.h:166) */
                continue;
            if (tri == last_prim)       /* FilterFunction.h:37-45, opaque scene */
                continue;
            if (better(t, tri, best))
            {
                best->prim = tri; best->t = t; best->u = u; best->v = v;
            }
        }
        *t_near = best->t;
        return *t_near > 0;
    }
    qel heap[8];
    int hn = 0;
    for (int c = 0; c < 8; c++)
    {
        float dist;
        if (kdop_intersect(&T->nodes[n->child[c]].bv, denoms, numers, &dist))
        {
            qel e = {n->child[c], dist};
            heap_push(heap, &hn, e);
        }
    }
    int found = 0;
    float closest = 100000000.0f, inter = 100000000.0f;
    while (hn > 0)
    {
        qel top = heap_pop(heap, &hn);
        if (node_intersect(T, top.node, o, d, last_prim, best, &inter, denoms, numers))
        {
            closest = closest < inter ? closest : inter;
            found = 1;
            if (hn == 0 || closest < heap[0].t)
            {
                *t_near = closest;
                return 1;
            }
        }
    }
    if (!found)
        return 0;
    *t_near = closest;
    return 1;
}

void orc_octree_trace(const orc_octree* T, const orc_ray* rays, orc_hit* hits, int32_t nrays,
                      int32_t any_hit, int32_t nthreads)
{
    (void)any_hit; /* any-hit = closest hit then t < tmax, as the CPU path of Intersect.h:297-323 */
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t r = 0; r < nrays; r++)
    {
        const orc_ray* ray = rays + r;
        f3 o = ld3(ray->o), d = ld3(ray->d);
        float denoms[NPLANES], numers[NPLANES];
        for (int i = 0; i < NPLANES; i++)
        {
            denoms[i] = dot(T->normals[i], d);
            numers[i] = dot(T->normals[i], o);
        }
        orc_hit best = MISS;
        float tn;
        if (T->nnodes)
            node_intersect(T, 0, o, d, ray->last_prim, &best, &tn, denoms, numers);
        if (best.prim >= 0 && !(best.t < ray->tmax))
            best = MISS;
        hits[r] = best;
    }
}

/* ------------------------------------------------------------------ */
/* Camera: src/HostDeviceCommon/HIPRTCamera.h:31-52, Math.h:617-634     */
/* ------------------------------------------------------------------ */
static f3 matrix_x_point(const float m[4][4], f3 p)
{
    float xt = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    float yt = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    float zt = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    float wt = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    float inv_w = 1.0f;
    if (!(wt < 1.0e-10f && wt > -1.0e-10f))
        inv_w = 1.0f / wt;
    return mk(xt * inv_w, yt * inv_w, zt * inv_w);
}

void orc_camera_ray(const orc_camera* cam, float x, float y, int32_t resx, int32_t resy, orc_ray* out)
{
    float xn = x / (float)resx * 2.0f - 1.0f;
    float yn = y / (float)resy * 2.0f - 1.0f;
    f3 origin = matrix_x_point(cam->inverse_view, mk(0.0f, 0.0f, 0.0f));
    f3 p_vs = matrix_x_point(cam->inverse_projection, mk(xn, yn, -1.0f));
    f3 p_ws = matrix_x_point(cam->inverse_view, p_vs);
    f3 dir = normalize(sub(p_ws, origin));
    out->o[0] = origin.x; out->o[1] = origin.y; out->o[2] = origin.z;
    out->d[0] = dir.x; out->d[1] = dir.y; out->d[2] = dir.z;
    out->last_prim = -1;
    out->tmax = INFINITY;
}

/* CameraRays.h:155-172: per-pixel seed, optional jitter of two draws. */
void orc_camera_rays(const orc_camera* cam, int32_t w, int32_t h, uint32_t sample_number,
                     uint32_t random_number, int32_t freeze_random, orc_ray* out,
                     uint32_t* out_rng_state)
{
#pragma omp parallel for schedule(static)
    for (int32_t y = 0; y < h; y++)
    {
        for (int32_t x = 0; x < w; x++)
        {
            uint32_t pix = (uint32_t)x + (uint32_t)y * (uint32_t)w;
            uint32_t st = orc_pixel_seed(pix, sample_number, random_number, freeze_random);
            float px = (float)x + 0.5f;
            float py = (float)y + 0.5f;
            if (cam->do_jittering)
            {
                px += orc_rand01(&st) - 0.5f;
                py += orc_rand01(&st) - 0.5f;
            }
            orc_camera_ray(cam, px, py, w, h, out + pix);
            if (out_rng_state)
                out_rng_state[pix] = st;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Hit attributes + fix_backfacing_normals (Intersect.h:142-176)        */
/* ------------------------------------------------------------------ */
static f3 reflect_ray(f3 d, f3 n)   /* Sampling.h:147-150 */
{
    float k = 2.0f * dot(d, n);
    return sub(scl(k, n), d);
}

static void fix_backfacing(f3 view, f3* gn, f3* sn)
{
    if (dot(view, *gn) < 0.0f)
    {
        *gn = scl(-1.0f, *gn);
        *sn = scl(-1.0f, *sn);
    }
    if (dot(view, *sn) < 0.0f)
        *sn = scl(-1.0f, *sn);
    f3 refl = reflect_ray(view, *sn);
    if (dot(refl, *gn) <= 0.0f)
    {
        const float eps = 0.01f;
        refl = sub(refl, normalize(scl(dot(refl, *gn) - eps, *gn)));
        *sn = normalize(add(view, refl));
    }
}

static void attributes(const float* verts, const int32_t* idx, const orc_ray* ray, const orc_hit* hit,
                       f3* p, f3* gn, f3* sn)
{
    f3 a, b, c;
    tri_verts(verts, idx, hit->prim, &a, &b, &c);
    f3 o = ld3(ray->o), d = ld3(ray->d);
    *p = add(o, scl(hit->t, d));
    /* hit.normal = normalize(cross(e1, e2)) (Triangle.h:54), renormalised at
     * Intersect.h:231; no vertex normals -> shading normal = geometric. */
    f3 n = normalize(cross(sub(b, a), sub(c, a)));
    *gn = normalize(n);
    *sn = *gn;
    fix_backfacing(scl(-1.0f, d), gn, sn);
}

void orc_hit_attributes(const float* verts, const int32_t* idx, const orc_ray* ray, const orc_hit* hit,
                        float* inter_point, float* geometric_normal, float* shading_normal)
{
    f3 p, gn, sn;
    attributes(verts, idx, ray, hit, &p, &gn, &sn);
    inter_point[0] = p.x; inter_point[1] = p.y; inter_point[2] = p.z;
    geometric_normal[0] = gn.x; geometric_normal[1] = gn.y; geometric_normal[2] = gn.z;
    shading_normal[0] = sn.x; shading_normal[1] = sn.y; shading_normal[2] = sn.z;
}

/* Sampling.h:180-201 */
static f3 cosine_around_normal(f3 n, uint32_t* st)
{
    float r1 = orc_rand01(st);
    float r2 = 2.0f * orc_rand01(st) - 1.0f;
    if (r1 < 1.0e-8f && r2 < -0.999999f && n.z > 0.999999f)
    {
        r1 += 1.0e-7f;
        r2 += 1.0e-7f;
    }
    float theta = 6.28318530717958647693f * r1;
    float s = sqrtf(1.0f - r2 * r2);
    f3 sp = mk(s * orc_cosf(theta), s * orc_sinf(theta), r2);
    return normalize(add(n, sp));
}

void orc_bounce_rays(const float* verts, const int32_t* idx, const orc_ray* primary, const orc_hit* hits,
                     uint32_t* rng_state, int32_t n, orc_ray* out)
{
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; i++)
    {
        if (hits[i].prim < 0)
        {
            out[i] = primary[i];
            continue;
        }
        f3 p, gn, sn;
        attributes(verts, idx, primary + i, hits + i, &p, &gn, &sn);
        uint32_t st = rng_state[i];
        f3 d = cosine_around_normal(sn, &st);
        rng_state[i] = st;
        out[i].o[0] = p.x; out[i].o[1] = p.y; out[i].o[2] = p.z;
        out[i].d[0] = d.x; out[i].d[1] = d.y; out[i].d[2] = d.z;
        out[i].last_prim = hits[i].prim;
        out[i].tmax = INFINITY;
    }
}

/* Element-wise evaluation of the fixed-sequence math functions of fmath.h,
 * for their accuracy tests: fn 0 = sin, 1 = cos, 2 = pow(x, y), 3 = log2,
 * 4 = exp2, 5 = log, 6 = exp, 7 = atan2(x, y), 8 = asin, 9 = acos. */
void orc_fmath_eval(int32_t fn, const float* x, const float* y, float* out, int32_t n)
{
    for (int32_t i = 0; i < n; i++)
    {
        switch (fn)
        {
        case 0: out[i] = orc_sinf(x[i]); break;
        case 1: out[i] = orc_cosf(x[i]); break;
        case 2: out[i] = orc_powf(x[i], y[i]); break;
        case 3: out[i] = orc_log2f(x[i]); break;
        case 5: out[i] = orc_logf(x[i]); break;
        case 6: out[i] = orc_expf(x[i]); break;
        case 7: out[i] = orc_atan2f(x[i], y[i]); break;
        case 8: out[i] = orc_asinf(x[i]); break;
        case 9: out[i] = orc_acosf(x[i]); break;
        default: out[i] = orc_exp2f(x[i]); break;
        }
    }
}
