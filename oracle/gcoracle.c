/*
 * gcoracle.c — CPU FP64 restatement of the galaxy-collision force path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see gcoracle.h.
 *
 * Build with -O3 -ffp-contract=off so that every a*b+c stays two roundings,
 * exactly as the reference's x86-64 Release build evaluates it.  Each
 * function names the reference lines whose arithmetic it restates; the
 * restatement keeps the same operation order, the same skip rules and the
 * same traversal order, because bit-level parity with the reference's golden
 * files depends on all three.
 */
#define _GNU_SOURCE /* sincos */
#include "gcoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* direct all-pairs                                                       */
/* ===================================================================== */

/* cpp/direct/direct_solver.cpp:8-23 — r = source - target,
 * s2 = |r|^2 + eps^2 (dot product summed x, y, z), zero when s2 == 0,
 * else r * ((G * m) * inv_r^3) with inv_r = 1 / sqrt(s2). */
void orc_softened_acceleration(const double t[3], const double s[3],
                               double m, double G, double eps, double out[3]) {
    const double rx = s[0] - t[0], ry = s[1] - t[1], rz = s[2] - t[2];
    const double eps2 = eps * eps;
    const double s2 = (rx * rx + ry * ry + rz * rz) + eps2;
    if (s2 == 0.0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    const double inv_r = 1.0 / sqrt(s2);
    const double inv_r3 = inv_r * inv_r * inv_r;
    const double k = G * m * inv_r3;
    out[0] = rx * k;
    out[1] = ry * k;
    out[2] = rz * k;
}

/* cpp/direct/direct_solver.cpp:38-57 — reset, then for i < j scatter
 * +delta*(scale*m_j) to i and -delta*(scale*m_i) to j (Newton's third law),
 * scale = G * inv_r^3, pairs with s2 == 0 skipped. */
void orc_direct_symmetric(size_t n, const double* pos, const double* mass,
                          double G, double eps, double* acc) {
    memset(acc, 0, sizeof(double) * 3 * n);
    const double eps2 = eps * eps;
    for (size_t i = 0; i < n; ++i) {
        const double* pi = pos + 3 * i;
        for (size_t j = i + 1; j < n; ++j) {
            const double* pj = pos + 3 * j;
            const double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
            const double s2 = (dx * dx + dy * dy + dz * dz) + eps2;
            if (s2 == 0.0) continue;
            const double inv_r = 1.0 / sqrt(s2);
            const double inv_r3 = inv_r * inv_r * inv_r;
            const double scale = G * inv_r3;
            const double fj = scale * mass[j];
            const double fi = scale * mass[i];
            acc[3 * i + 0] += dx * fj;
            acc[3 * i + 1] += dy * fj;
            acc[3 * i + 2] += dz * fj;
            acc[3 * j + 0] -= dx * fi;
            acc[3 * j + 1] -= dy * fi;
            acc[3 * j + 2] -= dz * fi;
        }
    }
}

/* One target of cpp/direct/direct_solver.cpp:72-81: ascending j, skip j == i,
 * accumulate softened_acceleration (a zero vector when s2 == 0). */
static void direct_one_target(size_t n, const double* pos, const double* mass,
                              double G, double eps, size_t i, double* out) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    const double* pi = pos + 3 * i;
    double f[3];
    for (size_t j = 0; j < n; ++j) {
        if (j == i) continue;
        orc_softened_acceleration(pi, pos + 3 * j, mass[j], G, eps, f);
        ax += f[0];
        ay += f[1];
        az += f[2];
    }
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
}

/* cpp/direct/direct_solver.cpp:59-83 (the MPI/targets form). */
void orc_direct_targets(size_t n, const double* pos, const double* mass,
                        double G, double eps, size_t begin, size_t end,
                        double* acc) {
    if (end > n) end = n;
    for (size_t i = begin; i < end; ++i)
        direct_one_target(n, pos, mass, G, eps, i, acc + 3 * i);
}

int orc_direct_targets_omp(size_t n, const double* pos, const double* mass,
                           double G, double eps, size_t begin, size_t end,
                           double* acc, int threads) {
    if (end > n) end = n;
    int used = 1;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
    long long b = (long long)begin, e = (long long)end;
#pragma omp parallel num_threads(threads)
    {
#pragma omp single
        used = omp_get_num_threads();
#pragma omp for schedule(dynamic, 16)
        for (long long i = b; i < e; ++i)
            direct_one_target(n, pos, mass, G, eps, (size_t)i, acc + 3 * i);
    }
#else
    (void)threads;
    orc_direct_targets(n, pos, mass, G, eps, begin, end, acc);
#endif
    return used;
}

/* ===================================================================== */
/* integrator                                                             */
/* ===================================================================== */

/* cpp/core/integrator.cpp:7-11 — v += a * dt. */
void orc_kick(size_t n, double* vel, const double* acc, double dt) {
    for (size_t k = 0; k < 3 * n; ++k) vel[k] += acc[k] * dt;
}

/* cpp/core/integrator.cpp:13-17 — x += v * dt. */
void orc_drift(size_t n, double* pos, const double* vel, double dt) {
    for (size_t k = 0; k < 3 * n; ++k) pos[k] += vel[k] * dt;
}

int orc_compute_accelerations(const orc_solver* s, size_t n, const double* pos,
                              const double* mass, double* acc) {
    switch (s->solver) {
    case 0:
        orc_direct_symmetric(n, pos, mass, s->G, s->eps, acc);
        return 0;
    case 1:
        orc_direct_targets(n, pos, mass, s->G, s->eps, 0, n, acc);
        return 0;
    case 2:
        return orc_fmm(n, pos, mass, s->G, s->eps, s->theta, s->leaf_capacity,
                       s->max_depth, s->expansion_order, 0, n, acc, NULL);
    case 3:
        return orc_tree(n, pos, mass, s->G, s->eps, s->theta, s->leaf_capacity,
                        s->max_depth, s->expansion_order, 0, n, acc, NULL);
    default:
        return -1;
    }
}

/* cpp/core/integrator.cpp:33-42 — kick(dt/2), drift(dt), a(x), kick(dt/2). */
int orc_leapfrog(const orc_solver* s, size_t n, double* pos, double* vel,
                 double* acc, const double* mass, double dt, int steps) {
    for (int k = 0; k < steps; ++k) {
        orc_kick(n, vel, acc, 0.5 * dt);
        orc_drift(n, pos, vel, dt);
        int rc = orc_compute_accelerations(s, n, pos, mass, acc);
        if (rc) return rc;
        orc_kick(n, vel, acc, 0.5 * dt);
    }
    return 0;
}

/* ===================================================================== */
/* diagnostics                                                            */
/* ===================================================================== */

/* cpp/core/diagnostics.cpp:7-37. */
void orc_diagnostics(size_t n, const double* pos, const double* vel,
                     const double* mass, double G, double eps, double* out) {
    double M = 0.0, KE = 0.0, PE = 0.0;
    double P[3] = {0, 0, 0}, C[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        const double* x = pos + 3 * i;
        const double* v = vel + 3 * i;
        const double m = mass[i];
        M += m;
        KE += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        P[0] += v[0] * m;
        P[1] += v[1] * m;
        P[2] += v[2] * m;
        C[0] += x[0] * m;
        C[1] += x[1] * m;
        C[2] += x[2] * m;
        const double cx = x[1] * v[2] - x[2] * v[1];
        const double cy = x[2] * v[0] - x[0] * v[2];
        const double cz = x[0] * v[1] - x[1] * v[0];
        L[0] += cx * m;
        L[1] += cy * m;
        L[2] += cz * m;
    }
    if (M > 0.0) {
        C[0] /= M;
        C[1] /= M;
        C[2] /= M;
    }
    const double eps2 = eps * eps;
    for (size_t i = 0; i < n; ++i) {
        const double* pi = pos + 3 * i;
        for (size_t j = i + 1; j < n; ++j) {
            const double* pj = pos + 3 * j;
            const double dx = pj[0] - pi[0], dy = pj[1] - pi[1], dz = pj[2] - pi[2];
            const double d = sqrt((dx * dx + dy * dy + dz * dz) + eps2);
            if (d == 0.0) continue;
            PE -= G * mass[i] * mass[j] / d;
        }
    }
    out[0] = M;
    out[1] = KE;
    out[2] = PE;
    out[3] = KE + PE;
    for (int k = 0; k < 3; ++k) {
        out[4 + k] = P[k];
        out[7 + k] = C[k];
        out[10 + k] = L[k];
    }
    out[13] = out[14] = out[15] = 0.0;
}

/* ===================================================================== */
/* Cartesian multipoles, total degree <= 4 (35 coefficients)              */
/* ===================================================================== */

#define NCOEF 35
#define MAXDEG 4

static int g_exp[NCOEF][3];
static int g_deg[NCOEF];
static int g_index[MAXDEG + 1][MAXDEG + 1][MAXDEG + 1];
/* product table: for (i, j) in row-major loop order with deg_i + deg_j <= 4 */
static int g_mul_i[NCOEF * NCOEF], g_mul_j[NCOEF * NCOEF], g_mul_k[NCOEF * NCOEF];
static int g_mul_count = 0;
static int g_tables_ready = 0;

/* Graded order, descending powers of x then y within each degree
 * (cpp/fmm/multipole.cpp:17-26 lists the same order as a table). */
static void build_tables(void) {
    if (g_tables_ready) return;
    int k = 0;
    for (int d = 0; d <= MAXDEG; ++d)
        for (int a = d; a >= 0; --a)
            for (int b = d - a; b >= 0; --b) {
                g_exp[k][0] = a;
                g_exp[k][1] = b;
                g_exp[k][2] = d - a - b;
                g_deg[k] = d;
                g_index[a][b][d - a - b] = k;
                ++k;
            }
    g_mul_count = 0;
    for (int i = 0; i < NCOEF; ++i)
        for (int j = 0; j < NCOEF; ++j) {
            if (g_deg[i] + g_deg[j] > MAXDEG) continue;
            g_mul_i[g_mul_count] = i;
            g_mul_j[g_mul_count] = j;
            g_mul_k[g_mul_count] = g_index[g_exp[i][0] + g_exp[j][0]]
                                          [g_exp[i][1] + g_exp[j][1]]
                                          [g_exp[i][2] + g_exp[j][2]];
            ++g_mul_count;
        }
    g_tables_ready = 1;
}

void orc_multipole_exponents(int* out) {
    build_tables();
    memcpy(out, g_exp, sizeof(g_exp));
}

/* repeated multiplication starting from 1.0 (multipole.cpp:41-47) */
static double ipow(double v, int k) {
    double r = 1.0;
    for (int t = 0; t < k; ++t) r *= v;
    return r;
}

static int binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    int r = 1;
    for (int t = 1; t <= k; ++t) r = r * (n - k + t) / t;
    return r;
}

/* truncated product, accumulated in (i, j) loop order; zero operands are
 * skipped (multipole.cpp:109-134) */
static void poly_mul(const double* a, const double* b, double* r) {
    for (int t = 0; t < NCOEF; ++t) r[t] = 0.0;
    for (int t = 0; t < g_mul_count; ++t) {
        const int i = g_mul_i[t], j = g_mul_j[t];
        if (a[i] == 0.0 || b[j] == 0.0) continue;
        r[g_mul_k[t]] += a[i] * b[j];
    }
}

/* P2M: multipole.cpp:226-233 — M_alpha += m * offset^alpha, 2 <= |alpha| <= 4 */
void orc_p2m(double* mom, const double off[3], double mass) {
    build_tables();
    for (int i = 0; i < NCOEF; ++i) {
        if (g_deg[i] < 2) continue;
        const double mono = ipow(off[0], g_exp[i][0]) * ipow(off[1], g_exp[i][1]) *
                            ipow(off[2], g_exp[i][2]);
        mom[i] += mass * mono;
    }
}

/* child moment seen by the shift: degree 0 is the mass, degree 1 vanishes
 * about the centre of mass (multipole.cpp:92-101) */
static double shifted_operand(const double* child, int beta, double child_mass) {
    if (g_deg[beta] == 0) return child_mass;
    if (g_deg[beta] == 1) return 0.0;
    return child[beta];
}

/* M2M: multipole.cpp:235-261 with multinomial_shift :74-90 */
void orc_m2m(double* parent, const double* child, const double off[3],
             double child_mass) {
    build_tables();
    for (int a = 0; a < NCOEF; ++a) {
        if (g_deg[a] < 2) continue;
        double value = 0.0;
        for (int b = 0; b < NCOEF; ++b) {
            if (g_deg[b] > g_deg[a]) continue;
            double sf;
            if (g_exp[b][0] > g_exp[a][0] || g_exp[b][1] > g_exp[a][1] ||
                g_exp[b][2] > g_exp[a][2]) {
                sf = 0.0;
            } else {
                const int c = binom(g_exp[a][0], g_exp[b][0]) *
                              binom(g_exp[a][1], g_exp[b][1]) *
                              binom(g_exp[a][2], g_exp[b][2]);
                sf = (double)c * ipow(off[0], g_exp[a][0] - g_exp[b][0]) *
                     ipow(off[1], g_exp[a][1] - g_exp[b][1]) *
                     ipow(off[2], g_exp[a][2] - g_exp[b][2]);
            }
            if (sf == 0.0) continue;
            value += sf * shifted_operand(child, b, child_mass);
        }
        parent[a] += value;
    }
}

/* Taylor polynomial (degree <= 4, in the source offset x) of
 * (h0 + 2 delta.x + |x|^2)^(-3/2): multipole.cpp:161-184.
 * c_n are the binomial-series coefficients of (1+q)^(-3/2). */
static void inv_r3_poly(const double delta[3], double eps, double* out) {
    const double h0 = (delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]) +
                      eps * eps;
    const double base = pow(h0, -1.5);
    double q[NCOEF], series[NCOEF], power[NCOEF], next[NCOEF];
    for (int t = 0; t < NCOEF; ++t) q[t] = series[t] = power[t] = 0.0;
    q[g_index[1][0][0]] = 2.0 * delta[0] / h0;
    q[g_index[0][1][0]] = 2.0 * delta[1] / h0;
    q[g_index[0][0][1]] = 2.0 * delta[2] / h0;
    q[g_index[2][0][0]] = 1.0 / h0;
    q[g_index[0][2][0]] = 1.0 / h0;
    q[g_index[0][0][2]] = 1.0 / h0;
    power[0] = 1.0;
    double c = 1.0; /* (-3/2 choose n) */
    for (int n = 0; n <= MAXDEG; ++n) {
        for (int t = 0; t < NCOEF; ++t) series[t] += power[t] * c;
        poly_mul(power, q, next);
        memcpy(power, next, sizeof(next));
        c = c * (-1.5 - (double)n) / (double)(n + 1);
    }
    for (int t = 0; t < NCOEF; ++t) out[t] = series[t] * base;
}

/* M2P: multipole.cpp:263-296 (component polynomials :152-159, evaluation
 * :186-216). */
void orc_m2p(const double target[3], const double com[3], double mass,
             const double* mom, double G, double eps, int order, double out[3]) {
    build_tables();
    out[0] = out[1] = out[2] = 0.0;
    if (mass <= 0.0) return;
    if (order < 0) order = 0;
    if (order > MAXDEG) order = MAXDEG;
    if (order == 1 || order == 3) ++order;
    const double delta[3] = {com[0] - target[0], com[1] - target[1], com[2] - target[2]};
    const double h0 = (delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]) +
                      eps * eps;
    if (h0 == 0.0) return;
    double inv[NCOEF];
    inv_r3_poly(delta, eps, inv);
    for (int comp = 0; comp < 3; ++comp) {
        /* g = delta_c * inv + x_c * inv (truncated) */
        double g[NCOEF], lin[NCOEF], shifted[NCOEF];
        for (int t = 0; t < NCOEF; ++t) {
            g[t] = inv[t] * delta[comp];
            lin[t] = 0.0;
        }
        lin[1 + comp] = 1.0;
        poly_mul(lin, inv, shifted);
        for (int t = 0; t < NCOEF; ++t) g[t] += shifted[t];
        double value = 0.0;
        for (int t = 0; t < NCOEF; ++t) {
            if (g_deg[t] > order) continue;
            double mv;
            if (g_deg[t] == 0) mv = mass;
            else if (g_deg[t] == 1) mv = 0.0;
            else mv = mom[t];
            value += g[t] * mv;
        }
        out[comp] = G * value;
    }
}

/* ===================================================================== */
/* FMM: octree, P2M/M2M, far/near lists, M2P + P2P                        */
/* ===================================================================== */

typedef struct {
    double c[3], hw;
    double mass, com[3];
    double mom[NCOEF];
    long child0; /* first of 8 consecutive children, -1 for a leaf */
    size_t pbeg, pcnt;
    int depth;
} onode;

typedef struct {
    onode* nodes;
    size_t count, cap;
    size_t* idx; /* particle indices, leaves own contiguous ranges */
    size_t* scratch;
    const double* pos;
    const double* mass;
    size_t leaf_capacity;
    int max_depth;
    int max_depth_reached;
} otree;

static long tree_push(otree* t) {
    if (t->count == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->nodes = (onode*)realloc(t->nodes, t->cap * sizeof(onode));
    }
    memset(&t->nodes[t->count], 0, sizeof(onode));
    t->nodes[t->count].child0 = -1;
    return (long)t->count++;
}

/* fmm_solver.cpp:131-134 — octant from ">=" tests against the centre */
static int octant(const onode* nd, const double* x) {
    return (x[0] >= nd->c[0] ? 1 : 0) + 2 * (x[1] >= nd->c[1] ? 1 : 0) +
           4 * (x[2] >= nd->c[2] ? 1 : 0);
}

/* The reference inserts particles one at a time (fmm_solver.cpp:136-155);
 * the resulting tree depends only on the particle set: a node at depth d is
 * split iff it receives more than leaf_capacity particles and d < max_depth,
 * and every leaf lists its particles in ascending index order.  This builds
 * the same tree top-down with a stable partition. */
static void tree_build(otree* t, long node, size_t beg, size_t cnt, int depth) {
    if (depth > t->max_depth_reached) t->max_depth_reached = depth;
    if (cnt <= t->leaf_capacity || depth >= t->max_depth) {
        t->nodes[node].pbeg = beg;
        t->nodes[node].pcnt = cnt;
        return;
    }
    /* subdivide: fmm_solver.cpp:106-129 */
    const double chw = t->nodes[node].hw * 0.5;
    long first = -1;
    for (int ch = 0; ch < 8; ++ch) {
        long k = tree_push(t);
        if (ch == 0) first = k;
        onode* pn = &t->nodes[node];
        onode* cn = &t->nodes[k];
        const double sx = (ch & 1) ? 1.0 : -1.0;
        const double sy = (ch & 2) ? 1.0 : -1.0;
        const double sz = (ch & 4) ? 1.0 : -1.0;
        cn->c[0] = pn->c[0] + sx * chw;
        cn->c[1] = pn->c[1] + sy * chw;
        cn->c[2] = pn->c[2] + sz * chw;
        cn->hw = chw;
        cn->depth = depth + 1;
    }
    t->nodes[node].child0 = first;
    size_t counts[8] = {0}, starts[8];
    for (size_t p = beg; p < beg + cnt; ++p)
        counts[octant(&t->nodes[node], t->pos + 3 * t->idx[p])]++;
    starts[0] = beg;
    for (int ch = 1; ch < 8; ++ch) starts[ch] = starts[ch - 1] + counts[ch - 1];
    size_t fill[8];
    memcpy(fill, starts, sizeof(fill));
    for (size_t p = beg; p < beg + cnt; ++p) {
        const size_t id = t->idx[p];
        t->scratch[fill[octant(&t->nodes[node], t->pos + 3 * id)]++] = id;
    }
    memcpy(t->idx + beg, t->scratch + beg, cnt * sizeof(size_t));
    for (int ch = 0; ch < 8; ++ch)
        tree_build(t, first + ch, starts[ch], counts[ch], depth + 1);
}

/* fmm_solver.cpp:157-182 — mass and centre of mass, bottom up */
static double tree_mass(otree* t, long node) {
    onode* nd = &t->nodes[node];
    double m = 0.0, w[3] = {0, 0, 0};
    if (nd->child0 < 0) {
        for (size_t p = nd->pbeg; p < nd->pbeg + nd->pcnt; ++p) {
            const size_t id = t->idx[p];
            const double mi = t->mass[id];
            m += mi;
            w[0] += t->pos[3 * id + 0] * mi;
            w[1] += t->pos[3 * id + 1] * mi;
            w[2] += t->pos[3 * id + 2] * mi;
        }
    } else {
        for (int ch = 0; ch < 8; ++ch) {
            const long k = t->nodes[node].child0 + ch;
            const double cm = tree_mass(t, k);
            const onode* cn = &t->nodes[k];
            m += cm;
            w[0] += cn->com[0] * cm;
            w[1] += cn->com[1] * cm;
            w[2] += cn->com[2] * cm;
        }
        nd = &t->nodes[node];
    }
    nd->mass = m;
    for (int d = 0; d < 3; ++d) nd->com[d] = m > 0.0 ? w[d] / m : nd->c[d];
    return m;
}

/* fmm_solver.cpp:184-208 — P2M at leaves, M2M upward, about the COM */
static void tree_moments(otree* t, long node) {
    onode* nd = &t->nodes[node];
    memset(nd->mom, 0, sizeof(nd->mom));
    if (nd->child0 < 0) {
        for (size_t p = nd->pbeg; p < nd->pbeg + nd->pcnt; ++p) {
            const size_t id = t->idx[p];
            const double off[3] = {t->pos[3 * id + 0] - nd->com[0],
                                   t->pos[3 * id + 1] - nd->com[1],
                                   t->pos[3 * id + 2] - nd->com[2]};
            orc_p2m(nd->mom, off, t->mass[id]);
        }
        return;
    }
    for (int ch = 0; ch < 8; ++ch) {
        const long k = t->nodes[node].child0 + ch;
        tree_moments(t, k);
        onode* pn = &t->nodes[node];
        const onode* cn = &t->nodes[k];
        const double off[3] = {cn->com[0] - pn->com[0], cn->com[1] - pn->com[1],
                               cn->com[2] - pn->com[2]};
        orc_m2m(pn->mom, cn->mom, off, cn->mass);
    }
}

typedef struct {
    long* v;
    size_t n, cap;
} lvec;

static void lvec_push(lvec* l, long x) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 256;
        l->v = (long*)realloc(l->v, l->cap * sizeof(long));
    }
    l->v[l->n++] = x;
}

/* fmm_solver.cpp:210-226 — DFS, massless subtrees skipped */
static void collect_leaves(const otree* t, long node, lvec* out) {
    const onode* nd = &t->nodes[node];
    if (nd->mass <= 0.0) return;
    if (nd->child0 < 0) {
        lvec_push(out, node);
        return;
    }
    for (int ch = 0; ch < 8; ++ch) collect_leaves(t, nd->child0 + ch, out);
}

/* fmm_solver.cpp:228-241 — geometric centres: 2(h_t + h_s) / |c_s - c_t| < theta */
static int well_separated(const onode* tg, const onode* src, double theta) {
    if (src->mass <= 0.0) return 0;
    const double dx = src->c[0] - tg->c[0], dy = src->c[1] - tg->c[1],
                 dz = src->c[2] - tg->c[2];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist == 0.0) return 0;
    const double w = 2.0 * (tg->hw + src->hw);
    return w / dist < theta;
}

/* fmm_solver.cpp:256-277 */
static void leaf_lists(const otree* t, long target, long src, double theta,
                       lvec* far, lvec* near) {
    const onode* s = &t->nodes[src];
    if (s->mass <= 0.0) return;
    if (src != target && well_separated(&t->nodes[target], s, theta)) {
        lvec_push(far, src);
        return;
    }
    if (s->child0 < 0) {
        lvec_push(near, src);
        return;
    }
    for (int ch = 0; ch < 8; ++ch) leaf_lists(t, target, s->child0 + ch, theta, far, near);
}

/* fmm_solver.cpp:12-18 (option clamps), :24-98 (build), :279-313
 * (per-particle evaluation: all far cells by M2P in list order, then every
 * near leaf by P2P skipping the target itself). */
int orc_fmm(size_t n, const double* pos, const double* mass, double G,
            double eps, double theta, size_t leaf_capacity, int max_depth,
            int expansion_order, size_t begin, size_t end, double* acc,
            orc_fmm_stats* stats) {
    build_tables();
    if (leaf_capacity < 1) leaf_capacity = 1;
    if (max_depth < 1) max_depth = 1;
    if (theta < 1.0e-6) theta = 1.0e-6;
    if (expansion_order < 0) expansion_order = 0;
    if (expansion_order > 4) expansion_order = 4;
    if (end > n) end = n;
    if (stats) memset(stats, 0, sizeof(*stats));
    for (size_t i = begin; i < end; ++i) acc[3 * i] = acc[3 * i + 1] = acc[3 * i + 2] = 0.0;
    if (n == 0 || begin >= end) return 0;

    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            const double x = pos[3 * i + d];
            if (x < lo[d]) lo[d] = x;
            if (hi[d] < x) hi[d] = x;
        }
    otree t;
    memset(&t, 0, sizeof(t));
    t.pos = pos;
    t.mass = mass;
    t.leaf_capacity = leaf_capacity;
    t.max_depth = max_depth;
    t.idx = (size_t*)malloc(n * sizeof(size_t));
    t.scratch = (size_t*)malloc(n * sizeof(size_t));
    if (!t.idx || !t.scratch) {
        free(t.idx);
        free(t.scratch);
        return -2;
    }
    for (size_t i = 0; i < n; ++i) t.idx[i] = i;
    long root = tree_push(&t);
    {
        onode* r = &t.nodes[root];
        for (int d = 0; d < 3; ++d) r->c[d] = (lo[d] + hi[d]) * 0.5;
        double ext = hi[0] - lo[0];
        if (ext < hi[1] - lo[1]) ext = hi[1] - lo[1];
        if (ext < hi[2] - lo[2]) ext = hi[2] - lo[2];
        double hw = 0.5 * ext;
        if (hw < eps + 1.0e-6) hw = eps + 1.0e-6;
        r->hw = hw * 1.0001;
    }
    tree_build(&t, root, 0, n, 0);
    tree_mass(&t, root);
    tree_moments(&t, root);

    lvec leaves = {0};
    collect_leaves(&t, root, &leaves);
    size_t nfar = 0, nnear = 0;
    /* leaves are independent: each target's sums keep the reference order,
     * so the result does not depend on the thread count */
    const long long nleaves = (long long)leaves.n;
#pragma omp parallel reduction(+ : nfar, nnear)
    {
    lvec far = {0}, near = {0};
#pragma omp for schedule(dynamic, 4)
    for (long long li = 0; li < nleaves; ++li) {
        const long L = leaves.v[li];
        far.n = near.n = 0;
        leaf_lists(&t, L, root, theta, &far, &near);
        nfar += far.n;
        nnear += near.n;
        const onode* leaf = &t.nodes[L];
        for (size_t p = leaf->pbeg; p < leaf->pbeg + leaf->pcnt; ++p) {
            const size_t id = t.idx[p];
            if (id < begin || id >= end) continue;
            const double* x = pos + 3 * id;
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, f[3];
            for (size_t k = 0; k < far.n; ++k) {
                const onode* s = &t.nodes[far.v[k]];
                orc_m2p(x, s->com, s->mass, s->mom, G, eps, expansion_order, f);
                a0 += f[0];
                a1 += f[1];
                a2 += f[2];
            }
            for (size_t k = 0; k < near.n; ++k) {
                const onode* s = &t.nodes[near.v[k]];
                for (size_t q = s->pbeg; q < s->pbeg + s->pcnt; ++q) {
                    const size_t sid = t.idx[q];
                    if (sid == id) continue;
                    orc_softened_acceleration(x, pos + 3 * sid, mass[sid], G, eps, f);
                    a0 += f[0];
                    a1 += f[1];
                    a2 += f[2];
                }
            }
            acc[3 * id + 0] = a0;
            acc[3 * id + 1] = a1;
            acc[3 * id + 2] = a2;
        }
    }
    free(far.v);
    free(near.v);
    }
    if (stats) {
        stats->node_count = t.count;
        stats->leaf_count = leaves.n;
        stats->far_interactions = nfar;
        stats->near_interactions = nnear;
        stats->max_depth_reached = t.max_depth_reached;
    }
    free(leaves.v);
    free(t.nodes);
    free(t.idx);
    free(t.scratch);
    return 0;
}

/* ===================================================================== */
/* Barnes-Hut tree (cpp/fmm/quadtree.cpp)                                 */
/* ===================================================================== */

typedef struct {
    const otree* t;
    double G, eps, theta;
    int order;
    size_t m2p, p2p;
} bh_ctx;

/* quadtree.cpp:193-243 — accumulate_from_node: a massless node gives
 * nothing; a leaf sums its particles (skipping the target) in order; an
 * internal node is one M2P when the target is outside its cube, at a
 * non-zero distance from its centre of mass and width / distance < theta,
 * and otherwise the sum of its children 0..7. */
static void bh_accumulate(bh_ctx* b, long node, size_t target, const double* x, double out[3]) {
    const otree* t = b->t;
    const onode* nd = &t->nodes[node];
    out[0] = out[1] = out[2] = 0.0;
    if (nd->mass <= 0.0) return;
    if (nd->child0 < 0) {
        double f[3];
        for (size_t p = nd->pbeg; p < nd->pbeg + nd->pcnt; ++p) {
            const size_t sid = t->idx[p];
            if (sid == target) continue;
            orc_softened_acceleration(x, t->pos + 3 * sid, t->mass[sid], b->G, b->eps, f);
            out[0] += f[0];
            out[1] += f[1];
            out[2] += f[2];
            b->p2p++;
        }
        return;
    }
    const double dx = nd->com[0] - x[0], dy = nd->com[1] - x[1], dz = nd->com[2] - x[2];
    const double distance = sqrt(dx * dx + dy * dy + dz * dz);
    const double width = 2.0 * nd->hw;
    /* quadtree.cpp:85-90 — contains */
    const int inside = fabs(x[0] - nd->c[0]) <= nd->hw && fabs(x[1] - nd->c[1]) <= nd->hw &&
                       fabs(x[2] - nd->c[2]) <= nd->hw;
    if (!inside && distance > 0.0 && width / distance < b->theta) {
        orc_m2p(x, nd->com, nd->mass, nd->mom, b->G, b->eps, b->order, out);
        b->m2p++;
        return;
    }
    double f[3];
    for (int ch = 0; ch < 8; ++ch) {
        bh_accumulate(b, nd->child0 + ch, target, x, f);
        out[0] += f[0];
        out[1] += f[1];
        out[2] += f[2];
    }
}

/* quadtree.cpp:12-35 (constructor clamps, compute), :37-81 (build: the same
 * root cube and insertion rule as the FMM, so the same tree), :143-191
 * (mass, centre of mass, P2M/M2M about the COM), and
 * compute_tree_accelerations (:245-254, max_depth 32 there).  Every target
 * in [begin, end) gets its own walk from the root. */
int orc_tree(size_t n, const double* pos, const double* mass, double G, double eps,
             double theta, size_t leaf_capacity, int max_depth, int expansion_order,
             size_t begin, size_t end, double* acc, orc_tree_stats* stats) {
    build_tables();
    if (leaf_capacity < 1) leaf_capacity = 1;
    if (max_depth < 1) max_depth = 1;
    if (expansion_order < 0) expansion_order = 0;
    if (expansion_order > 4) expansion_order = 4;
    if (end > n) end = n;
    if (stats) memset(stats, 0, sizeof(*stats));
    for (size_t i = begin; i < end; ++i) acc[3 * i] = acc[3 * i + 1] = acc[3 * i + 2] = 0.0;
    if (n == 0 || begin >= end) return 0;

    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            const double x = pos[3 * i + d];
            if (x < lo[d]) lo[d] = x;
            if (hi[d] < x) hi[d] = x;
        }
    otree t;
    memset(&t, 0, sizeof(t));
    t.pos = pos;
    t.mass = mass;
    t.leaf_capacity = leaf_capacity;
    t.max_depth = max_depth;
    t.idx = (size_t*)malloc(n * sizeof(size_t));
    t.scratch = (size_t*)malloc(n * sizeof(size_t));
    if (!t.idx || !t.scratch) {
        free(t.idx);
        free(t.scratch);
        return -2;
    }
    for (size_t i = 0; i < n; ++i) t.idx[i] = i;
    long root = tree_push(&t);
    {
        onode* r = &t.nodes[root];
        for (int d = 0; d < 3; ++d) r->c[d] = (lo[d] + hi[d]) * 0.5;
        double ext = hi[0] - lo[0];
        if (ext < hi[1] - lo[1]) ext = hi[1] - lo[1];
        if (ext < hi[2] - lo[2]) ext = hi[2] - lo[2];
        double hw = 0.5 * ext;
        if (hw < eps + 1.0e-6) hw = eps + 1.0e-6;
        r->hw = hw * 1.0001;
    }
    tree_build(&t, root, 0, n, 0);
    tree_mass(&t, root);
    tree_moments(&t, root);

    size_t m2p = 0, p2p = 0;
    const long long lo_i = (long long)begin, hi_i = (long long)end;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : m2p, p2p)
    for (long long i = lo_i; i < hi_i; ++i) {
        bh_ctx b = {&t, G, eps, theta, expansion_order, 0, 0};
        bh_accumulate(&b, root, (size_t)i, pos + 3 * i, acc + 3 * i);
        m2p += b.m2p;
        p2p += b.p2p;
    }
    if (stats) {
        size_t leaves = 0;
        for (size_t k = 0; k < t.count; ++k) leaves += t.nodes[k].child0 < 0 && t.nodes[k].pcnt > 0;
        stats->node_count = t.count;
        stats->leaf_count = leaves;
        stats->m2p_evaluations = m2p;
        stats->p2p_pairs = p2p;
        stats->max_depth_reached = t.max_depth_reached;
    }
    free(t.nodes);
    free(t.idx);
    free(t.scratch);
    return 0;
}


/* ===================================================================== */
/* initial conditions                                                     */
/* ===================================================================== */

/* MT19937-64 (Nishimura 2000; the C++ standard's mt19937_64). */
#define MT_N 312
#define MT_M 156
typedef struct {
    uint64_t mt[MT_N];
    int i;
} orc_mt64;

static void mt64_seed(orc_mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int k = 1; k < MT_N; ++k)
        s->mt[k] = 6364136223846793005ULL * (s->mt[k - 1] ^ (s->mt[k - 1] >> 62)) + (uint64_t)k;
    s->i = MT_N;
}

static uint64_t mt64_next(orc_mt64* s) {
    if (s->i >= MT_N) {
        for (int k = 0; k < MT_N; ++k) {
            const uint64_t x = (s->mt[k] & 0xFFFFFFFF80000000ULL) |
                               (s->mt[(k + 1) % MT_N] & 0x7FFFFFFFULL);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
            s->mt[k] = s->mt[(k + MT_M) % MT_N] ^ xa;
        }
        s->i = 0;
    }
    uint64_t y = s->mt[s->i++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

uint64_t orc_mt64_nth(uint64_t seed, size_t n) {
    orc_mt64 s;
    mt64_seed(&s, seed);
    uint64_t y = 0;
    for (size_t k = 0; k < n; ++k) y = mt64_next(&s);
    return y;
}

/* uniform_real_distribution<double>(0, 1) of GNU libstdc++:
 * generate_canonical<double, 53> takes one 64-bit draw, u / 2^64. */
static double mt64_canonical(orc_mt64* s) {
    const double r = (double)mt64_next(s) / 18446744073709551616.0;
    return r >= 1.0 ? nextafter(1.0, 0.0) : r;
}

/* normal_distribution<double>(0, 1) of GNU libstdc++: Marsaglia's polar
 * method, the second value of each accepted pair kept for the next call. */
typedef struct {
    int saved_available;
    double saved;
} orc_polar;

static double polar_normal(orc_polar* p, orc_mt64* s) {
    if (p->saved_available) {
        p->saved_available = 0;
        return p->saved;
    }
    double x, y, r2;
    do {
        x = 2.0 * mt64_canonical(s) - 1.0;
        y = 2.0 * mt64_canonical(s) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double f = sqrt(-2 * log(r2) / r2);
    p->saved = x * f;
    p->saved_available = 1;
    return y * f;
}

static const double kOrcPi = 3.141592653589793238462643383279502884;

/* rotate_x then rotate_z (initial_conditions.cpp:12-26); the reference's
 * GCC build evaluates each cos/sin pair of one angle with sincos. */
static void orient(double orientation, double inclination, double v[3]) {
    double s, c;
    sincos(inclination, &s, &c);
    const double y1 = c * v[1] - s * v[2];
    const double z1 = s * v[1] + c * v[2];
    sincos(orientation, &s, &c);
    const double x2 = c * v[0] - s * y1;
    const double y2 = s * v[0] + c * y1;
    v[0] = x2;
    v[1] = y2;
    v[2] = z1;
}

/* One particle of generate_disk_galaxy (initial_conditions.cpp:28-33,
 * 52-88) from its draws.  g: n, mass, radius, pos[3], vel[3], orientation,
 * thickness, inclination, G, eps. */
static void ref_disk_from_draws(const double* g, double ur, double ut, double nz, double n1,
                                double n2, double n3, double* x, double* v, double* m) {
    const double R = g[2], M = g[1], G = g[12], eps = g[13];
    const double pm = M / g[0];
    const double disp = 0.025 * sqrt(G * M / R);
    const double scale = R / 3.0;
    const double max_cdf = 1.0 - exp(-R / scale);
    const double u = ur < 1.0e-12 ? 1.0e-12 : (1.0 - 1.0e-12 < ur ? 1.0 - 1.0e-12 : ur);
    const double r = -scale * log(1.0 - u * max_cdf);
    const double theta = 2.0 * kOrcPi * ut;
    double st, ct;
    sincos(theta, &st, &ct);
    x[0] = r * ct;
    x[1] = r * st;
    x[2] = g[10] > 0.0 ? g[10] * nz : 0.0;
    double frac = (r * r) / (R * R);
    if (!(frac < 1.0)) frac = 1.0;
    double menc = M * frac;
    if (menc < pm) menc = pm;
    const double s2 = r * r + eps * eps;
    const double vc = sqrt(G * menc * r * r / pow(s2, 1.5));
    v[0] = -st * vc + disp * n1;
    v[1] = ct * vc + disp * n2;
    v[2] = 0.0 * vc + 0.25 * disp * n3;
    orient(g[9], g[11], x);
    orient(g[9], g[11], v);
    for (int d = 0; d < 3; ++d) {
        x[d] = g[3 + d] + x[d];
        v[d] = g[6 + d] + v[d];
    }
    *m = pm;
}

size_t orc_generate_galaxies(const double* gal, int count, double G, double eps, uint64_t seed,
                             double* pos, double* vel, double* mass) {
    orc_mt64 s;
    mt64_seed(&s, seed);
    size_t off = 0;
    for (int k = 0; k < count; ++k) {
        const double* c = gal + 12 * k; /* n, mass, radius, pos3, vel3, orient, thick, incl */
        if (c[0] < 1.0 || c[1] <= 0.0 || c[2] <= 0.0) continue;
        const double g[14] = {c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7],
                              c[8], c[9], c[10], c[11], G, eps};
        orc_polar normal = {0, 0.0};
        for (size_t i = 0; i < (size_t)c[0]; ++i, ++off) {
            const double ur = mt64_canonical(&s);
            const double ut = mt64_canonical(&s);
            const double nz = c[10] > 0.0 ? polar_normal(&normal, &s) : 0.0;
            const double n1 = polar_normal(&normal, &s);
            const double n2 = polar_normal(&normal, &s);
            const double n3 = polar_normal(&normal, &s);
            ref_disk_from_draws(g, ur, ut, nz, n1, n2, n3, pos + 3 * off, vel + 3 * off,
                                mass + off);
        }
    }
    return off;
}

/* Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]}, k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t a = (uint64_t)0xD2511F53u * c[0];
        const uint64_t b = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t o0 = (uint32_t)(b >> 32) ^ c[1] ^ k0;
        const uint32_t o1 = (uint32_t)b;
        const uint32_t o2 = (uint32_t)(a >> 32) ^ c[3] ^ k1;
        const uint32_t o3 = (uint32_t)a;
        c[0] = o0;
        c[1] = o1;
        c[2] = o2;
        c[3] = o3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    memcpy(out, c, sizeof(c));
}

/* Draw block b of particle i: two uniforms of 53 bits in [0, 1). */
static void philox_block(uint64_t seed, uint64_t i, uint32_t stream, uint32_t b, double u[2]) {
    const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)(i >> 32), stream, b};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    u[0] = ldexp((double)(((uint64_t)o[0] << 21) ^ (o[1] >> 11)), -53);
    u[1] = ldexp((double)(((uint64_t)o[2] << 21) ^ (o[3] >> 11)), -53);
}

static void gauss_pair(const double u[2], double* a, double* b) {
    const double r = sqrt(-2.0 * log(1.0 - u[0]));
    *a = r * cos(2.0 * kOrcPi * u[1]);
    *b = r * sin(2.0 * kOrcPi * u[1]);
}

/* The device generators' models on the counter-based stream (blocks 0-2:
 * (u_r, u_angle), normals (z, vx), normals (vy, vz); Plummer: blocks 0-2
 * (u, cos theta, phi, cos theta', phi'), rejection pairs from block 3). */
int orc_ic_generate(int model, const double* p, uint64_t seed, uint32_t stream, size_t n,
                    double* pos, double* vel, double* mass) {
    for (size_t i = 0; i < n; ++i) {
        double b0[2], b1[2], b2[2], n0, n1, n2, n3;
        philox_block(seed, i, stream, 0, b0);
        philox_block(seed, i, stream, 1, b1);
        philox_block(seed, i, stream, 2, b2);
        double* x = pos + 3 * i;
        double* v = vel + 3 * i;
        if (model == 0) { /* reference disk: p as ref_disk_from_draws' g */
            gauss_pair(b1, &n0, &n1);
            gauss_pair(b2, &n2, &n3);
            ref_disk_from_draws(p, b0[0], b0[1], n0, n1, n2, n3, x, v, mass + i);
        } else if (model == 1) {
            /* two-disk galaxy: n, central, md, disk total, h, r_min, r_max,
             * z_sigma, v_sigma, inclination, pos3, vel3, G, eps */
            double px[3] = {0, 0, 0}, pv[3] = {0, 0, 0};
            if (i == 0) {
                mass[i] = p[1];
            } else {
                gauss_pair(b1, &n0, &n1);
                gauss_pair(b2, &n2, &n3);
                const double h = p[4];
                const double ea = exp(-p[5] / h), eb = exp(-p[6] / h);
                const double R = -h * log(ea - b0[0] * (ea - eb));
                const double phi = 2.0 * kOrcPi * b0[1];
                px[0] = R * cos(phi);
                px[1] = R * sin(phi);
                px[2] = p[7] * n0;
                const double cdf = (ea - exp(-R / h)) / (ea - eb);
                const double menc = p[1] + p[3] * cdf;
                const double s2 = R * R + p[17] * p[17];
                const double vc = sqrt(p[16] * menc * R * R / (s2 * sqrt(s2)));
                pv[0] = -sin(phi) * vc + p[8] * n1;
                pv[1] = cos(phi) * vc + p[8] * n2;
                pv[2] = p[8] * n3;
                mass[i] = p[2];
            }
            const double c = cos(p[9]), sn = sin(p[9]);
            const double y0 = c * px[1] - sn * px[2], z0 = sn * px[1] + c * px[2];
            const double y1 = c * pv[1] - sn * pv[2], z1 = sn * pv[1] + c * pv[2];
            x[0] = p[10] + px[0];
            x[1] = p[11] + y0;
            x[2] = p[12] + z0;
            v[0] = p[13] + pv[0];
            v[1] = p[14] + y1;
            v[2] = p[15] + z1;
        } else if (model == 2) { /* Plummer: n, total mass, scale, G */
            const double u = b0[0];
            const double ct = -1.0 + 2.0 * b0[1], ph = 2.0 * kOrcPi * b1[0];
            const double r = 1.0 / sqrt(pow(u, -2.0 / 3.0) - 1.0);
            double st = 1.0 - ct * ct;
            st = sqrt(st > 0.0 ? st : 0.0);
            x[0] = p[2] * r * st * cos(ph);
            x[1] = p[2] * r * st * sin(ph);
            x[2] = p[2] * r * ct;
            double q = 0.0;
            for (uint32_t a = 0; a < 1000; ++a) {
                double w[2];
                philox_block(seed, i, stream, 3 + a, w);
                q = w[0];
                const double g = 0.1 * w[1];
                if (!(g > q * q * pow(1.0 - q * q, 3.5))) break;
            }
            const double ct2 = -1.0 + 2.0 * b1[1], ph2 = 2.0 * kOrcPi * b2[0];
            const double vscale = sqrt(p[3] * p[1] / p[2]);
            const double speed = q * sqrt(2.0) * pow(1.0 + r * r, -0.25);
            double st2 = 1.0 - ct2 * ct2;
            st2 = sqrt(st2 > 0.0 ? st2 : 0.0);
            v[0] = vscale * speed * st2 * cos(ph2);
            v[1] = vscale * speed * st2 * sin(ph2);
            v[2] = vscale * speed * ct2;
            mass[i] = p[1] / p[0];
        } else {
            return 1;
        }
    }
    return 0;
}
