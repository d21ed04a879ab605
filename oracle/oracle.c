/*
 * oracle.c -- scalar CPU restatement of the reference traversal paths.
 *
 * TEST INFRASTRUCTURE: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py only, as the checker.  See oracle.h for the
 * arithmetic contract shared with the HIP kernels.
 *
 * Reference files followed:
 *   src/TraversalKernelCWBVH.cu:56-371   CWBVH ray setup, child selection,
 *                                        node decode, slab test, meta decode,
 *                                        triangle loop, pop/finish
 *   src/TraversalKernelBVH2.cu:14-177    BVH2 slab test, while-while loop,
 *                                        postponed leaf, leaf terminator
 *   src/FastDeviceMinMax.h:44-79         int-ordered 3-way min/max
 *   src/GPUBVHConverter.cpp:380-403 +
 *   src/WoopTriangleHelper.h:318-354     Woop transform (cofactor inverse)
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline int32_t f2i(float f) { int32_t i; memcpy(&i, &f, 4); return i; }
static inline float i2f(int32_t i) { float f; memcpy(&f, &i, 4); return f; }

/* int-ordered min/max of float bit patterns (reference vmin/vmax.s32) */
static inline float imin2(float a, float b) { return f2i(a) < f2i(b) ? a : b; }
static inline float imax2(float a, float b) { return f2i(a) > f2i(b) ? a : b; }
static inline float imin3(float a, float b, float c) { return imin2(imin2(a, b), c); }
static inline float imax3(float a, float b, float c) { return imax2(imax2(a, b), c); }

static inline int highest_bit(uint32_t x) { return 31 - __builtin_clz(x); }

typedef struct {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
    float idx, idy, idz;
} RayS;

static inline float safe_rcp(float d) {
    const float ooeps = 0x1p-80f;
    return 1.0f / (fabsf(d) > ooeps ? d : copysignf(ooeps, d));
}

static inline void setup_ray(const or_ray* r, RayS* s) {
    s->ox = r->origin_tmin.x; s->oy = r->origin_tmin.y; s->oz = r->origin_tmin.z;
    s->tmin = r->origin_tmin.w;
    s->dx = r->dir_tmax.x; s->dy = r->dir_tmax.y; s->dz = r->dir_tmax.z;
    s->tmax = r->dir_tmax.w;
    s->idx = safe_rcp(s->dx);
    s->idy = safe_rcp(s->dy);
    s->idz = safe_rcp(s->dz);
}

/* Woop test following the contract in oracle.h.  Returns the computed t in *t
 * always; returns 1 if accepted (given mode / current hit). */
static inline int woop(const or_float4* w, const RayS* r, float tmin, float tmax, int32_t tri_addr,
                       int32_t hit_addr, int any_hit, float* t_out, float* u_out, float* v_out) {
    const or_float4 v0 = w[0], v1 = w[1], v2 = w[2];
    float Oz = fmaf(-r->oz, v0.z, fmaf(-r->oy, v0.y, fmaf(-r->ox, v0.x, v0.w)));
    float Dz = fmaf(r->dz, v0.z, fmaf(r->dy, v0.y, r->dx * v0.x));
    float invDz = 1.0f / Dz;
    float t = Oz * invDz;
    int t_ok;
    if (any_hit) t_ok = t > tmin && t < tmax;
    else t_ok = t > tmin && (t < tmax || (t == tmax && hit_addr >= 0 && tri_addr < hit_addr));
    if (!t_ok) return 0;
    float Ox = fmaf(r->oz, v1.z, fmaf(r->oy, v1.y, fmaf(r->ox, v1.x, v1.w)));
    float Dx = fmaf(r->dz, v1.z, fmaf(r->dy, v1.y, r->dx * v1.x));
    float u = fmaf(t, Dx, Ox);
    if (!(u >= 0.0f && u <= 1.0f)) return 0;
    float Oy = fmaf(r->oz, v2.z, fmaf(r->oy, v2.y, fmaf(r->ox, v2.x, v2.w)));
    float Dy = fmaf(r->dz, v2.z, fmaf(r->dy, v2.y, r->dx * v2.x));
    float v = fmaf(t, Dy, Oy);
    if (!(v >= 0.0f && u + v <= 1.0f)) return 0;
    *t_out = t;
    *u_out = u;
    *v_out = v;
    return 1;
}

int oracle_woop_test(const or_float4* woop3, const or_ray* ray, float tmin, float tmax, float* t, float* u,
                     float* v) {
    RayS s;
    setup_ray(ray, &s);
    float tt = 0, uu = 0, vv = 0;
    int ok = woop(woop3, &s, tmin, tmax, 0, -1, 1, &tt, &uu, &vv);
    if (ok) {
        if (t) *t = tt;
        if (u) *u = uu;
        if (v) *v = vv;
    }
    return ok;
}

static inline void write_hit(or_hit* h, float t, int32_t addr, float u, float v) {
    h->t_triId_u_v.x = t;
    h->t_triId_u_v.y = i2f(addr);
    h->t_triId_u_v.z = u;
    h->t_triId_u_v.w = v;
}

/* ------------------------------------------------------------------ CWBVH */

typedef struct { uint32_t x, y; } U2;

static int trace_cwbvh_one(const or_float4* nodes, int stride, const or_float4* tris, const or_ray* ray, or_hit* hit,
                           int any_hit, uint64_t* n_nodes, uint64_t* n_tris, int cap) {
    RayS r;
    setup_ray(ray, &r);
    uint32_t oct = ((r.dx < 0 ? 1u : 0u) << 2) | ((r.dy < 0 ? 1u : 0u) << 1) | (r.dz < 0 ? 1u : 0u);
    uint32_t octinv = 7u - oct;
    uint32_t octinv4 = octinv * 0x01010101u;
    U2 stack[128];
    int sp = 0;
    U2 ng = {0u, 0x80000000u}, tg = {0u, 0u};
    float tmax = r.tmax, hu = 0.f, hv = 0.f;
    int32_t hit_addr = -1;
    int done = 0;
    for (;;) {
        if (ng.y > 0x00FFFFFFu) {
            uint32_t imask = ng.y;
            int bit = highest_bit(ng.y);
            uint32_t base = ng.x;
            ng.y &= ~(1u << bit);
            if (ng.y > 0x00FFFFFFu) {
                if (sp >= cap) return -1;
                stack[sp++] = ng;
            }
            uint32_t slot = ((uint32_t)bit - 24u) ^ octinv;
            uint32_t rel = (uint32_t)__builtin_popcount(imask & ~(0xFFFFFFFFu << slot));
            const or_float4* nd = nodes + (int64_t)(base + rel) * stride;
            (*n_nodes)++;
            or_float4 n0 = nd[0], n1 = nd[1], n2 = nd[2], n3 = nd[3], n4 = nd[4];
            uint32_t ebits = f2u(n0.w);
            int ex = (int8_t)(ebits & 0xFF), ey = (int8_t)((ebits >> 8) & 0xFF), ez = (int8_t)((ebits >> 16) & 0xFF);
            float adjx = u2f((uint32_t)(ex + 127) << 23) * r.idx;
            float adjy = u2f((uint32_t)(ey + 127) << 23) * r.idy;
            float adjz = u2f((uint32_t)(ez + 127) << 23) * r.idz;
            float origx = -(r.ox - n0.x) * r.idx;
            float origy = -(r.oy - n0.y) * r.idy;
            float origz = -(r.oz - n0.z) * r.idz;
            ng.x = f2u(n1.x);
            tg.x = f2u(n1.y);
            uint32_t hitmask = 0;
            for (int half = 0; half < 2; ++half) {
                uint32_t meta4 = f2u(half ? n1.w : n1.z);
                uint32_t qlx = f2u(half ? n2.y : n2.x), qly = f2u(half ? n2.w : n2.z), qlz = f2u(half ? n3.y : n3.x);
                uint32_t qhx = f2u(half ? n3.w : n3.z), qhy = f2u(half ? n4.y : n4.x), qhz = f2u(half ? n4.w : n4.z);
                uint32_t nx = r.idx < 0 ? qhx : qlx, fx = r.idx < 0 ? qlx : qhx;
                uint32_t ny = r.idy < 0 ? qhy : qly, fy = r.idy < 0 ? qly : qhy;
                uint32_t nz = r.idz < 0 ? qhz : qlz, fz = r.idz < 0 ? qlz : qhz;
                for (int j = 0; j < 4; ++j) {
                    int sh = 8 * j;
                    float tnx = fmaf((float)((nx >> sh) & 0xFF), adjx, origx);
                    float tny = fmaf((float)((ny >> sh) & 0xFF), adjy, origy);
                    float tnz = fmaf((float)((nz >> sh) & 0xFF), adjz, origz);
                    float tfx = fmaf((float)((fx >> sh) & 0xFF), adjx, origx);
                    float tfy = fmaf((float)((fy >> sh) & 0xFF), adjy, origy);
                    float tfz = fmaf((float)((fz >> sh) & 0xFF), adjz, origz);
                    float cmin = imax2(imax3(tnx, tny, tnz), r.tmin);
                    float cmax = imin2(imin3(tfx, tfy, tfz), tmax);
                    if (cmin <= cmax) {
                        uint32_t meta = (meta4 >> sh) & 0xFF;
                        uint32_t is_inner = (meta & (meta << 1)) & 0x10u;
                        uint32_t bit_index = (is_inner ? (meta ^ octinv) : meta) & 0x1Fu;
                        uint32_t child_bits = (meta >> 5) & 0x7u;
                        hitmask |= child_bits << bit_index;
                    }
                }
            }
            (void)octinv4;
            ng.y = (hitmask & 0xFF000000u) | (ebits >> 24);
            tg.y = hitmask & 0x00FFFFFFu;
        } else {
            tg = ng;
            ng.x = 0u;
            ng.y = 0u;
        }
        while (tg.y != 0u) {
            int ti = highest_bit(tg.y);
            int32_t addr = (int32_t)(tg.x * 3u + (uint32_t)ti * 3u);
            (*n_tris)++;
            float t, u, v;
            if (woop(tris + addr, &r, r.tmin, tmax, addr, hit_addr, any_hit, &t, &u, &v)) {
                tmax = t;
                hit_addr = addr;
                hu = u;
                hv = v;
                if (any_hit) {
                    done = 1;
                    break;
                }
            }
            tg.y &= ~(1u << ti);
        }
        if (done) break;
        if (ng.y <= 0x00FFFFFFu) {
            if (sp > 0) ng = stack[--sp];
            else break;
        }
    }
    write_hit(hit, hit_addr >= 0 ? tmax : r.tmax, hit_addr, hit_addr >= 0 ? hu : 0.f, hit_addr >= 0 ? hv : 0.f);
    return 0;
}

/* ------------------------------------------------------------------- BVH2 */

static inline int box2(const float lo[3], const float hi[3], const RayS* r, float ox, float oy, float oz,
                       float tmin, float tmax, float* dist) {
    float lx = fmaf(lo[0], r->idx, -ox), hx = fmaf(hi[0], r->idx, -ox);
    float ly = fmaf(lo[1], r->idy, -oy), hy = fmaf(hi[1], r->idy, -oy);
    float lz = fmaf(lo[2], r->idz, -oz), hz = fmaf(hi[2], r->idz, -oz);
    float smin = imax2(imax3(imin2(lx, hx), imin2(ly, hy), imin2(lz, hz)), tmin);
    float smax = imin2(imin3(imax2(lx, hx), imax2(ly, hy), imax2(lz, hz)), tmax);
    *dist = smin;
    return smin <= smax;
}

static int trace_bvh2_one(const or_float4* nodes, const or_float4* tris, const or_ray* ray, or_hit* hit,
                          int any_hit, uint64_t* n_nodes, uint64_t* n_tris, int cap) {
    const int32_t SENTINEL = 0x76543210;
    RayS r;
    setup_ray(ray, &r);
    float ox = r.ox * r.idx, oy = r.oy * r.idy, oz = r.oz * r.idz;
    int32_t stack[130];
    int sp = 0;
    stack[0] = SENTINEL;
    int32_t node = 0, leaf = 0, hit_addr = -1;
    float hitT = r.tmax, hu = 0.f, hv = 0.f;
    while (node != SENTINEL) {
        leaf = 0;
        while (node != SENTINEL && node >= 0) {
            const or_float4* n = nodes + node;
            (*n_nodes)++;
            float c0lo[3] = {n[0].x, n[0].z, n[2].x}, c0hi[3] = {n[0].y, n[0].w, n[2].y};
            float c1lo[3] = {n[1].x, n[1].z, n[2].z}, c1hi[3] = {n[1].y, n[1].w, n[2].w};
            int32_t ch0 = f2i(n[3].x), ch1 = f2i(n[3].y);
            float d0, d1;
            int t0 = box2(c0lo, c0hi, &r, ox, oy, oz, r.tmin, hitT, &d0);
            int t1 = box2(c1lo, c1hi, &r, ox, oy, oz, r.tmin, hitT, &d1);
            if (!t0 && !t1) {
                node = stack[sp--];
            } else {
                node = t0 ? ch0 : ch1;
                if (t0 && t1) {
                    int32_t far = ch1;
                    if (d1 < d0) {
                        far = node;
                        node = ch1;
                    }
                    if (sp + 1 >= cap) return -1;
                    stack[++sp] = far;
                }
            }
            if (node < 0 && leaf >= 0) {
                leaf = node;
                node = stack[sp--];
            }
            if (!(leaf >= 0)) break; /* the single lane holds a postponed leaf */
        }
        while (leaf < 0) {
            for (int32_t a = ~leaf;; a += 3) {
                if (f2u(tris[a].x) == 0x80000000u) break;
                (*n_tris)++;
                float t, u, v;
                if (woop(tris + a, &r, r.tmin, hitT, a, hit_addr, any_hit, &t, &u, &v)) {
                    hitT = t;
                    hit_addr = a;
                    hu = u;
                    hv = v;
                    if (any_hit) goto finish;
                }
            }
            leaf = node;
            if (node < 0) node = stack[sp--];
        }
    }
finish:
    write_hit(hit, hitT, hit_addr, hit_addr >= 0 ? hu : 0.f, hit_addr >= 0 ? hv : 0.f);
    return 0;
}

/* ------------------------------------------------------------------ brute */

static void trace_brute_one(const or_float4* tris, const int32_t* addrs, int64_t ntri, const or_ray* ray, or_hit* hit,
                            int any_hit) {
    RayS r;
    setup_ray(ray, &r);
    float tmax = r.tmax, hu = 0.f, hv = 0.f;
    int32_t hit_addr = -1;
    for (int64_t i = 0; i < ntri; ++i) {
        int32_t a = addrs[i];
        float t, u, v;
        if (woop(tris + a, &r, r.tmin, tmax, a, hit_addr, any_hit, &t, &u, &v)) {
            tmax = t;
            hit_addr = a;
            hu = u;
            hv = v;
            if (any_hit) break;
        }
    }
    write_hit(hit, hit_addr >= 0 ? tmax : r.tmax, hit_addr, hit_addr >= 0 ? hu : 0.f, hit_addr >= 0 ? hv : 0.f);
}

/* ---------------------------------------------------------------- threads */

typedef struct {
    int kind; /* 0 cwbvh, 1 bvh2, 2 brute */
    const or_float4* nodes;
    int stride;
    const or_float4* tris;
    const int32_t* addrs;
    int64_t ntri;
    const or_ray* rays;
    or_hit* hits;
    int64_t begin, end;
    int any_hit;
    int cap;
    uint64_t nn, nt;
    int status;
} Job;

static void* run_job(void* p) {
    Job* j = (Job*)p;
    for (int64_t i = j->begin; i < j->end; ++i) {
        int rc = 0;
        if (j->kind == 0)
            rc = trace_cwbvh_one(j->nodes, j->stride, j->tris, j->rays + i, j->hits + i, j->any_hit, &j->nn, &j->nt,
                                 j->cap);
        else if (j->kind == 1)
            rc = trace_bvh2_one(j->nodes, j->tris, j->rays + i, j->hits + i, j->any_hit, &j->nn, &j->nt, j->cap);
        else
            trace_brute_one(j->tris, j->addrs, j->ntri, j->rays + i, j->hits + i, j->any_hit);
        if (rc) j->status = rc;
    }
    return NULL;
}

static int run_all(Job proto, int64_t n, int threads, uint64_t* counters) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((int64_t)threads > n) threads = n > 0 ? (int)n : 1;
    Job* jobs = (Job*)calloc((size_t)threads, sizeof(Job));
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    int64_t chunk = (n + threads - 1) / threads;
    for (int k = 0; k < threads; ++k) {
        jobs[k] = proto;
        jobs[k].begin = (int64_t)k * chunk < n ? (int64_t)k * chunk : n;
        jobs[k].end = jobs[k].begin + chunk < n ? jobs[k].begin + chunk : n;
        jobs[k].nn = jobs[k].nt = 0;
        jobs[k].status = 0;
    }
    if (threads == 1) {
        run_job(&jobs[0]);
    } else {
        for (int k = 0; k < threads; ++k) pthread_create(&th[k], NULL, run_job, &jobs[k]);
        for (int k = 0; k < threads; ++k) pthread_join(th[k], NULL);
    }
    int status = 0;
    for (int k = 0; k < threads; ++k) {
        if (counters) {
            counters[0] += jobs[k].nn;
            counters[1] += jobs[k].nt;
        }
        if (jobs[k].status) status = jobs[k].status;
    }
    free(jobs);
    free(th);
    return status;
}

int oracle_trace_cwbvh(const or_float4* nodes, int node_stride, const or_float4* woop, const or_ray* rays, or_hit* hits,
                       int64_t n, int any_hit, uint64_t* counters, int threads, int stack_capacity) {
    if (stack_capacity > 128) stack_capacity = 128;
    Job p;
    memset(&p, 0, sizeof(p));
    p.kind = 0;
    p.nodes = nodes;
    p.stride = node_stride;
    p.tris = woop;
    p.rays = rays;
    p.hits = hits;
    p.any_hit = any_hit;
    p.cap = stack_capacity;
    return run_all(p, n, threads, counters);
}

int oracle_trace_bvh2(const or_float4* nodes, const or_float4* woop, const or_ray* rays, or_hit* hits, int64_t n,
                      int any_hit, uint64_t* counters, int threads, int stack_capacity) {
    if (stack_capacity > 128) stack_capacity = 128;
    Job p;
    memset(&p, 0, sizeof(p));
    p.kind = 1;
    p.nodes = nodes;
    p.tris = woop;
    p.rays = rays;
    p.hits = hits;
    p.any_hit = any_hit;
    p.cap = stack_capacity;
    return run_all(p, n, threads, counters);
}

int oracle_trace_brute(const or_float4* woop, const int32_t* tri_addrs, int64_t ntri, const or_ray* rays, or_hit* hits,
                       int64_t n, int any_hit, int threads) {
    Job p;
    memset(&p, 0, sizeof(p));
    p.kind = 2;
    p.tris = woop;
    p.addrs = tri_addrs;
    p.ntri = ntri;
    p.rays = rays;
    p.hits = hits;
    p.any_hit = any_hit;
    return run_all(p, n, threads, NULL);
}

/* ------------------------------------------------------------ Woop (ref) */

/* Inverse of a 4x4 matrix m (m[r][c]) by the adjugate: inv = adj(m) / det,
 * adj[c][r] = (-1)^(r+c) * minor(r, c).  All in binary32.  Returns 0 when the
 * determinant is zero (the reference then returns the identity). */
static int inverse4(const float m[4][4], float inv[4][4]) {
    float cof[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float s[3][3];
            int ri = 0;
            for (int rr = 0; rr < 4; ++rr) {
                if (rr == r) continue;
                int ci = 0;
                for (int cc = 0; cc < 4; ++cc) {
                    if (cc == c) continue;
                    s[ri][ci++] = m[rr][cc];
                }
                ++ri;
            }
            float minor = s[0][0] * (s[1][1] * s[2][2] - s[1][2] * s[2][1]) -
                          s[0][1] * (s[1][0] * s[2][2] - s[1][2] * s[2][0]) +
                          s[0][2] * (s[1][0] * s[2][1] - s[1][1] * s[2][0]);
            cof[r][c] = ((r + c) & 1) ? -minor : minor;
        }
    float det = 0.f;
    for (int c = 0; c < 4; ++c) det += m[0][c] * cof[0][c];
    if (det == 0.f) return 0;
    float id = 1.0f / det;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) inv[r][c] = cof[c][r] * id;
    return 1;
}

void oracle_woopify_ref(const float* v0, const float* v1, const float* v2, or_float4* out) {
    float e0[3] = {v0[0] - v2[0], v0[1] - v2[1], v0[2] - v2[2]};
    float e1[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
    float n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
    float m[4][4];
    for (int r = 0; r < 3; ++r) {
        m[r][0] = e0[r];
        m[r][1] = e1[r];
        m[r][2] = n[r];
        m[r][3] = v2[r];
    }
    m[3][0] = m[3][1] = m[3][2] = 0.f;
    m[3][3] = 1.f;
    float inv[4][4];
    if (!inverse4(m, inv)) {
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) inv[r][c] = r == c ? 1.f : 0.f;
    }
    out[0].x = inv[2][0]; out[0].y = inv[2][1]; out[0].z = inv[2][2]; out[0].w = -inv[2][3];
    out[1].x = inv[0][0]; out[1].y = inv[0][1]; out[1].z = inv[0][2]; out[1].w = inv[0][3];
    out[2].x = inv[1][0]; out[2].y = inv[1][1]; out[2].z = inv[1][2]; out[2].w = inv[1][3];
}

void oracle_cwbvh_decode(const or_float4* node, double* lo, double* hi) {
    const unsigned char* b0 = (const unsigned char*)&node[0];
    const unsigned char* q = (const unsigned char*)&node[2];
    double p[3] = {node[0].x, node[0].y, node[0].z};
    for (int a = 0; a < 3; ++a) {
        int e = (signed char)b0[12 + a];
        double s = ldexp(1.0, e);
        for (int c = 0; c < 8; ++c) {
            lo[c * 3 + a] = p[a] + q[a * 8 + c] * s;
            hi[c * 3 + a] = p[a] + q[24 + a * 8 + c] * s;
        }
    }
}
