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
