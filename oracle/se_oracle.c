/*
 * se_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, never shipped or called
 * by the product path).  Plain-C restatement of the reference SE hot path.
 * Compile with -ffp-contract=off so that the operation order written here is
 * the operation order executed (the reference is built by g++ for baseline
 * x86-64, which has no FMA).
 *
 * Parity status per function (pins are in tests/test_oracle*.py):
 *   orc_gll, mesh verts/maps, GetEl, GlobalToLocal, ShapeFuncsAtCoord,
 *   InterpGlobal, DetJac, element Laplacian/Mass/FieldRHS, advect (1st/2nd),
 *   GenVel, CalcTimeStep      -- pinned bit-exact against golden vectors from
 *                                the compiled reference (tests/golden) plus
 *                                closed forms.
 *   CSR assembly, lifting     -- pinned against the element blocks above and
 *                                the dense restriction A_full*u_D (closed form).
 *   PCG, solve                -- PETSc KSPSolve is not buildable here: the
 *                                algorithm is the build's definition (Jacobi-
 *                                PCG, x0 = 0, true residual stop); pinned by
 *                                direct dense solves and manufactured solutions.
 *   SVV element operator      -- the literal 2-D construction, pinned on its
 *                                separable closed form (1-D modal filter from
 *                                scipy's Jacobi polynomials), its null space
 *                                and definiteness (tests/test_oracle_svv.py).
 *   SolveQ + DisRHS /         -- pinned on closed forms: pointwise terms with
 *   DissipationRHS               the mass cancelled, periodic copies averaged,
 *                                the viscous term equal to the assembled
 *                                stiffness for continuous gradients; the
 *                                diagonal mass solve is taken as exact
 *                                division (parity unpinned for that one ulp).
 *   QG steps                  -- compositions of the above, pinned on exact
 *                                solutions: Rossby waves (advection, beta, F)
 *                                and damped steady modes (drag, relaxation,
 *                                Ekman drag, viscosity, topography).
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>
#include "se_oracle.h"

/* ------------------------------------------------------------------------- */
/* a1: Legendre polynomials by the Jacobi three-term recurrence              */
/* ------------------------------------------------------------------------- */

/* Jacobi::Gen, Jacobi.cpp:21-36.  c has n+1 entries and starts zeroed. */
static void jac_gen(double* c, int n, double a, const double* m1, int n1,
                    const double* m2, int n2) {
    double c1 = 2 * n * (2 * a + n) * (2 * a + 2 * n - 2);
    double c2 = (2 * a + 2 * n - 1) * (2 * a + 2 * n) * (2 * a + 2 * n - 2);
    double c3 = 2 * (a + n - 1) * (a + n - 1) * (2 * a + 2 * n);
    int i;
    c2 /= c1;
    c3 /= c1;
    for (i = 0; i <= n1; i++) c[i + 1] += c2 * m1[i];
    for (i = 0; i <= n2; i++) c[i] -= c3 * m2[i];
}

/* Jacobi::Eval, Jacobi.cpp:38-46 (parity-definite: every other term). */
double orc_leg_eval(const double* c, int n, double x) {
    double p = 0.0;
    int i;
    for (i = n; i >= 0; i -= 2) p += c[i] * pow(x, i);
    return p;
}

/* Jacobi::EvalDeriv, Jacobi.cpp:58-66 */
double orc_leg_eval_deriv(const double* c, int n, double x) {
    double dp = 0.0;
    int i;
    for (i = n; i >= 1; i -= 2) dp += c[i] * i * pow(x, i - 1);
    return dp;
}

/* Jacobi::Eval2ndDeriv, Jacobi.cpp:68-76 */
double orc_leg_eval_2nd(const double* c, int n, double x) {
    double d2p = 0.0;
    int i;
    for (i = n; i >= 2; i -= 2) d2p += c[i] * i * (i - 1) * pow(x, i - 2);
    return d2p;
}

/* RootFinder, Element.cpp:54-66 (Newton on P_N', tolerance 1e-12).  The
 * iteration cap only guards against a non-converging start; it never
 * triggers for the orders used here. */
static double root_finder(double x0, const double* P, int N) {
    double x = x0, xp;
    int it;
    for (it = 0; it < 100000; it++) {
        xp = x;
        x = xp - orc_leg_eval_deriv(P, N, xp) / orc_leg_eval_2nd(P, N, xp);
        if (fabs(x - xp) < 1.0e-12) return x;
    }
    return x;
}

/* AlreadyFound, Element.cpp:68-77 */
static int already_found(int n, const double* roots, double root) {
    int i;
    for (i = 0; i < n; i++)
        if (fabs(roots[i] - root) < 1.0e-6) return 1;
    return 0;
}

/* SortEntries, Element.cpp:79-106 (selection of the running minimum). */
static void sort_entries(double* vals, int n) {
    int i, j, minInd = 999999999;
    int* done = (int*)calloc((size_t)n, sizeof(int));
    double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
    double minVal;
    for (i = 0; i < n; i++) {
        minVal = 1.0e+99;
        for (j = 0; j < n; j++) {
            if (!done[j] && vals[j] < minVal) {
                minVal = vals[j];
                minInd = j;
            }
        }
        tmp[i] = minVal;
        done[minInd] = 1;
    }
    for (i = 0; i < n; i++) vals[i] = tmp[i];
    free(tmp);
    free(done);
}

/* Legendre::Legendre, Legendre.cpp:15-95.  Outputs abcissa[N+1],
 * weights[N+1], dCij[(N+1)^2] (i = point, j = basis) and the N+1
 * monomial coefficients of P_N.  Returns 0 on success. */
int orc_gll(int N, double* abcissa, double* weights, double* dCij, double* pn) {
    int i, j, endLoop;
    double x0;
    double** polys;
    double* P;
    if (N < 2) return -1; /* Legendre.cpp:58 reads polys[N-2] */
    polys = (double**)malloc(sizeof(double*) * (size_t)(N + 2));
    for (i = 0; i <= N + 1; i++) polys[i] = (double*)calloc((size_t)(i + 1), sizeof(double));
    polys[0][0] = 1.0;
    polys[1][1] = 1.0;
    for (i = 2; i <= N + 1; i++) jac_gen(polys[i], i, 0.0, polys[i - 1], i - 1, polys[i - 2], i - 2);

    abcissa[0] = -1.0;
    endLoop = ((N + 1) % 2 == 1) ? N / 2 : (N + 1) / 2;
    for (i = 1; i < endLoop; i++) {
        x0 = abcissa[i - 1];
        do {
            x0 += 1.0e-3;
            abcissa[i] = root_finder(x0, polys[N], N);
        } while (already_found(i, abcissa, abcissa[i]) || abcissa[i] > 0.0 - 1.0e-3);
        abcissa[N - i] = -abcissa[i];
    }
    abcissa[N] = 1.0;
    if ((N + 1) % 2 == 1) abcissa[(N + 1) / 2] = 0.0;
    sort_entries(abcissa, N + 1);

    weights[0] = 2.0 / ((N + 1) * N);
    for (i = 1; i < N; i++)
        weights[i] = 2.0 / ((N + 1) * N * orc_leg_eval(polys[N], N, abcissa[i]) *
                            orc_leg_eval(polys[N], N, abcissa[i]));
    weights[N] = 2.0 / ((N + 1) * N);

    /* the cardinal-function polynomial P is regenerated from polys[N-1],
     * polys[N-2] exactly as polys[N] was (Legendre.cpp:71-72) */
    P = (double*)calloc((size_t)(N + 1), sizeof(double));
    jac_gen(P, N, 0.0, polys[N - 1], N - 1, polys[N - 2], N - 2);
    for (i = 0; i <= N; i++) {
        for (j = 0; j <= N; j++) {
            double v;
            if (i == j && i == 0) v = -0.25 * N * (N + 1);
            else if (i == j && i == N) v = +0.25 * N * (N + 1);
            else if (i == j) v = 0.0;
            else v = orc_leg_eval(P, N, abcissa[i]) / (orc_leg_eval(P, N, abcissa[j]) * (abcissa[i] - abcissa[j]));
            if (dCij) dCij[i * (N + 1) + j] = v;
        }
    }
    if (pn) memcpy(pn, P, sizeof(double) * (size_t)(N + 1));
    free(P);
    for (i = 0; i <= N + 1; i++) free(polys[i]);
    free(polys);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* a2/a3: mesh and global numbering                                          */
/* ------------------------------------------------------------------------- */

struct orc_mesh {
    int N, nl;                 /* order, (N+1)^2 */
    int nEls[2], nVerts[2], nV, nElsTotal;
    int periodic[2];
    int bc[4];                 /* effective bottom, top, left, right after BCs.cpp:39-44 */
    double min[2], max[2], dx[2];
    double *abcissa, *weights, *dCij, *pn;
    double* verts;             /* nV x 2 */
    int *f2v, *v2f;
    int n, size0;
};

/* BCs::IsBCNode, BCs.cpp:349-367 */
static int is_bc_node(const orc_mesh* m, int node) {
    int t0 = node % m->nVerts[0], t1 = node / m->nVerts[0];
    if (m->bc[0] && t1 == 0) return 1;
    if (m->bc[1] && t1 == m->nVerts[1] - 1) return 1;
    if (m->bc[2] && t0 == 0) return 1;
    if (m->bc[3] && t0 == m->nVerts[0] - 1) return 1;
    return 0;
}

/* BCs::BCs (scalar), BCs.cpp:14-173.  Returns 0, or -1 where the reference
 * exits (inconsistent bcs), or -2 for the combinations the reference leaves
 * unmapped (y-periodic only; x-periodic with a bottom wall only). */
static int build_bcs(orc_mesh* m) {
    int nv0 = m->nVerts[0], nv1 = m->nVerts[1], nV = m->nV;
    int fieldNode, vecNode = 0, t0, t1;
    int* bcb = &m->bc[0]; int* bct = &m->bc[1]; int* bcl = &m->bc[2]; int* bcr = &m->bc[3];
    int size0 = 0;
    if ((*bcb || *bct) && m->periodic[1]) return -1;
    if ((*bcl || *bcr) && m->periodic[0]) return -1;
    if (m->periodic[0]) *bcr = 1;
    if (m->periodic[1]) *bct = 1;
    if (*bcb) size0 += nv0;
    if (*bct || m->periodic[1]) size0 += nv0;
    if (*bcl) size0 += nv1;
    if (*bcr || m->periodic[0]) size0 += nv1;
    if (*bcb && *bcl) size0--;
    if (*bcb && *bcr) size0--;
    if (*bct && *bcl) size0--;
    if (*bct && *bcr) size0--;
    m->size0 = size0;
    m->n = nV - size0;
    m->v2f = (int*)malloc(sizeof(int) * (size_t)(m->n > 0 ? m->n : 1));
    m->f2v = (int*)malloc(sizeof(int) * (size_t)nV);
    if ((!m->periodic[0] && !m->periodic[1]) || (m->periodic[0] && m->periodic[1])) {
        for (fieldNode = 0; fieldNode < nV; fieldNode++) {
            t0 = fieldNode % nv0; t1 = fieldNode / nv0;
            if (!is_bc_node(m, fieldNode)) {
                m->v2f[vecNode] = fieldNode;
                m->f2v[fieldNode] = vecNode++;
            } else if (m->periodic[0] && m->periodic[1] && fieldNode == nV - 1) {
                m->f2v[fieldNode] = 0;
            } else if (m->periodic[0] && t0 == nv0 - 1) {
                m->f2v[fieldNode] = t1 * (nv0 - 1);
            } else if (m->periodic[1] && t1 == nv1 - 1) {
                m->f2v[fieldNode] = t0;
            } else {
                m->f2v[fieldNode] = -1;
            }
        }
    } else if (m->periodic[0] && (*bct && *bcb)) {
        for (fieldNode = 0; fieldNode < nV; fieldNode++) {
            t0 = fieldNode % nv0; t1 = fieldNode / nv0;
            if (!is_bc_node(m, fieldNode)) {
                m->v2f[vecNode] = fieldNode;
                m->f2v[fieldNode] = vecNode++;
            } else if (t1 == 0 || t1 == nv1 - 1) {
                m->f2v[fieldNode] = -1;
            } else if (t0 == nv0 - 1) {
                m->f2v[fieldNode] = (t1 - 1) * (nv0 - 1);
            } else {
                return -1;
            }
        }
    } else if (m->periodic[0] && *bct) {
        for (fieldNode = 0; fieldNode < nV; fieldNode++) {
            t0 = fieldNode % nv0; t1 = fieldNode / nv0;
            if (!is_bc_node(m, fieldNode)) {
                m->v2f[vecNode] = fieldNode;
                m->f2v[fieldNode] = vecNode++;
            } else if (t1 == nv1 - 1) {
                m->f2v[fieldNode] = -1;
            } else if (t0 == nv0 - 1) {
                m->f2v[fieldNode] = t1 * (nv0 - 1);
            } else {
                return -1;
            }
        }
    } else if (m->periodic[0]) {
        for (fieldNode = 0; fieldNode < nV; fieldNode++) {
            t0 = fieldNode % nv0; t1 = fieldNode / nv0;
            if (!is_bc_node(m, fieldNode)) {
                m->v2f[vecNode] = fieldNode;
                m->f2v[fieldNode] = vecNode++;
            } else if (t0 == nv0 - 1) {
                m->f2v[fieldNode] = t1 * (nv0 - 1);
            } else {
                return -1;
            }
        }
    } else {
        return -2; /* y-periodic only: no branch of BCs.cpp:91-172 assigns the map */
    }
    return 0;
}

/* Mesh::Mesh, Mesh.cpp:17-73 (legendre elements) + BCs */
orc_mesh* orc_mesh_create(const int* nEls, int N, const double* min, const double* max,
                          const int* periodic, const int* bc, int* status) {
    orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
    int d, v, rc;
    m->N = N;
    m->nl = (N + 1) * (N + 1);
    m->abcissa = (double*)malloc(sizeof(double) * (size_t)(N + 1));
    m->weights = (double*)malloc(sizeof(double) * (size_t)(N + 1));
    m->dCij = (double*)malloc(sizeof(double) * (size_t)m->nl);
    m->pn = (double*)malloc(sizeof(double) * (size_t)(N + 1));
    if (orc_gll(N, m->abcissa, m->weights, m->dCij, m->pn) != 0) {
        if (status) *status = -3;
        orc_mesh_destroy(m);
        return NULL;
    }
    for (d = 0; d < 2; d++) {
        m->nEls[d] = nEls[d];
        m->nVerts[d] = nEls[d] * N + 1;
        m->min[d] = min[d];
        m->max[d] = max[d];
        m->periodic[d] = periodic[d] ? 1 : 0;
        m->dx[d] = (max[d] - min[d]) / nEls[d];
    }
    for (d = 0; d < 4; d++) m->bc[d] = bc[d] ? 1 : 0;
    m->nElsTotal = m->nEls[0] * m->nEls[1];
    m->nV = m->nVerts[0] * m->nVerts[1];
    m->verts = (double*)malloc(sizeof(double) * 2 * (size_t)m->nV);
    for (v = 0; v < m->nV; v++) {
        int t0 = v % m->nVerts[0], t1 = v / m->nVerts[0];
        int ei = t0 / N, ej = t1 / N;
        double l0 = m->abcissa[t0 % N], l1 = m->abcissa[t1 % N];
        m->verts[2 * v + 0] = m->min[0] + ei * m->dx[0] + ((l0 + 1.0) / 2.0) * m->dx[0];
        m->verts[2 * v + 1] = m->min[1] + ej * m->dx[1] + ((l1 + 1.0) / 2.0) * m->dx[1];
    }
    rc = build_bcs(m);
    if (status) *status = rc;
    if (rc != 0) {
        orc_mesh_destroy(m);
        return NULL;
    }
    return m;
}

void orc_mesh_destroy(orc_mesh* m) {
    if (!m) return;
    free(m->abcissa); free(m->weights); free(m->dCij); free(m->pn);
    free(m->verts); free(m->f2v); free(m->v2f);
    free(m);
}

void orc_mesh_sizes(const orc_mesh* m, int* out) {
    out[0] = m->nVerts[0]; out[1] = m->nVerts[1]; out[2] = m->nV; out[3] = m->nElsTotal;
    out[4] = m->n; out[5] = m->size0; out[6] = m->N;
}

void orc_mesh_verts(const orc_mesh* m, double* verts) {
    memcpy(verts, m->verts, sizeof(double) * 2 * (size_t)m->nV);
}

void orc_mesh_maps(const orc_mesh* m, int* f2v, int* v2f) {
    if (f2v) memcpy(f2v, m->f2v, sizeof(int) * (size_t)m->nV);
    if (v2f) memcpy(v2f, m->v2f, sizeof(int) * (size_t)m->n);
}

void orc_mesh_gll(const orc_mesh* m, double* abcissa, double* weights, double* dCij) {
    if (abcissa) memcpy(abcissa, m->abcissa, sizeof(double) * (size_t)(m->N + 1));
    if (weights) memcpy(weights, m->weights, sizeof(double) * (size_t)(m->N + 1));
    if (dCij) memcpy(dCij, m->dCij, sizeof(double) * (size_t)m->nl);
}

/* Mesh::ElNodes, Mesh.cpp:138-152 */
void orc_el_nodes(const orc_mesh* m, int el, int* nodes) {
    int N = m->N, ii, jj, k = 0;
    int t0 = N * (el % m->nEls[0]), t1 = N * (el / m->nEls[0]);
    int n0 = t1 * m->nVerts[0] + t0;
    for (jj = 0; jj < N + 1; jj++)
        for (ii = 0; ii < N + 1; ii++) nodes[k++] = n0 + jj * m->nVerts[0] + ii;
}

#define NUMAPPROX(var, val) ((var) >= (val) - 1.0e-12 && (var) <= (val) + 1.0e-12)

/* Mesh::GetEl, Mesh.cpp:106-137 */
void orc_get_el(const orc_mesh* m, const double* coord, int* el) {
    int topo[2], d;
    double out, frac, integer;
    for (d = 0; d < 2; d++) {
        if (NUMAPPROX(coord[d] - m->max[d], 0.0)) {
            topo[d] = m->nEls[d] - 1;
        } else {
            out = (coord[d] - m->min[d]) / m->dx[d];
            frac = modf(out, &integer);
            topo[d] = (int)integer;
            if (topo[d] > 0 && NUMAPPROX(frac, 0.0)) topo[d]--;
        }
    }
    *el = topo[1] * m->nEls[0] + topo[0];
}

/* Mesh::GlobalToLocal, Mesh.cpp:163-177 */
void orc_global_to_local(const orc_mesh* m, const double* g, int* el, double* l) {
    int d;
    orc_get_el(m, g, el);
    l[0] = (g[0] - m->min[0] - (*el % m->nEls[0]) * m->dx[0]) * (2.0 / m->dx[0]) - 1.0;
    l[1] = (g[1] - m->min[1] - (*el / m->nEls[0]) * m->dx[1]) * (2.0 / m->dx[1]) - 1.0;
    for (d = 0; d < 2; d++) {
        if (l[d] < -1.0) l[d] = -1;
        if (l[d] > +1.0) l[d] = +1;
    }
}

/* Legendre::IsRoot, Legendre.cpp:291-305 (EPS 1e-8) */
static int is_root(const orc_mesh* m, double x, int* i) {
    int k;
    for (k = 0; k <= m->N; k++) {
        if (fabs(x - m->abcissa[k]) < 1.0e-8) {
            *i = k;
            return 1;
        }
    }
    *i = -1;
    return 0;
}

/* Legendre::ShapeFuncsAtCoord, Legendre.cpp:145-169 */
void orc_shape_funcs_at_coord(const orc_mesh* m, const double* coord, double* Ni) {
    int N = m->N, pt_j, x_i, y_i, x_j, y_j;
    double Nx, Ny;
    for (pt_j = 0; pt_j < m->nl; pt_j++) {
        x_j = pt_j % (N + 1);
        y_j = pt_j / (N + 1);
        if (is_root(m, coord[0], &x_i)) {
            Nx = (x_i == x_j) ? 1.0 : 0.0;
        } else {
            Nx = -((1.0 - coord[0] * coord[0]) /
                   (N * (N + 1) * orc_leg_eval(m->pn, N, m->abcissa[x_j]) * (coord[0] - m->abcissa[x_j]))) *
                 orc_leg_eval_deriv(m->pn, N, coord[0]);
        }
        if (is_root(m, coord[1], &y_i)) {
            Ny = (y_i == y_j) ? 1.0 : 0.0;
        } else {
            Ny = -((1.0 - coord[1] * coord[1]) /
                   (N * (N + 1) * orc_leg_eval(m->pn, N, m->abcissa[y_j]) * (coord[1] - m->abcissa[y_j]))) *
                 orc_leg_eval_deriv(m->pn, N, coord[1]);
        }
        Ni[pt_j] = Nx * Ny;
    }
}

/* Field::InterpLocal, Field.cpp:57-68 (vals is nV x nDofs, node-major) */
static void interp_local(const orc_mesh* m, const double* vals, int nDofs, int el,
                         const double* l, double* out, double* Ni, int* nodes) {
    int node, d;
    orc_shape_funcs_at_coord(m, l, Ni);
    orc_el_nodes(m, el, nodes);
    for (d = 0; d < nDofs; d++) out[d] = 0.0;
    for (node = 0; node < m->nl; node++)
        for (d = 0; d < nDofs; d++) out[d] += Ni[node] * vals[(size_t)nodes[node] * nDofs + d];
}

/* Field::InterpGlobal, Field.cpp:70-76 */
void orc_interp_global(const orc_mesh* m, const double* vals, int nDofs, const double* g, double* out) {
    int el;
    double l[2];
    double* Ni = (double*)malloc(sizeof(double) * (size_t)m->nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)m->nl);
    orc_global_to_local(m, g, &el, l);
    interp_local(m, vals, nDofs, el, l, out, Ni, nodes);
    free(Ni);
    free(nodes);
}

/* Legendre::ShapeFuncDerivs, Legendre.cpp:171-184 (GNix is 2 x nl) */
static void ref_shape_func_derivs(const orc_mesh* m, int pt, double* GNix) {
    int N = m->N, pt_j;
    int x_i = pt % (N + 1), y_i = pt / (N + 1);
    for (pt_j = 0; pt_j < m->nl; pt_j++) {
        int x_j = pt_j % (N + 1), y_j = pt_j / (N + 1);
        GNix[pt_j] = (y_i == y_j) ? m->dCij[x_i * (N + 1) + x_j] : 0.0;
        GNix[m->nl + pt_j] = (x_i == x_j) ? m->dCij[y_i * (N + 1) + y_j] : 0.0;
    }
}

/* Jacobian from vertex coordinates, Mesh.cpp:183-198 / 217-250 */
static void jacobian(const orc_mesh* m, int el, const double* GNix, double J[2][2], int* nodes) {
    int node;
    orc_el_nodes(m, el, nodes);
    J[0][0] = J[0][1] = J[1][0] = J[1][1] = 0.0;
    for (node = 0; node < m->nl; node++) {
        const double* vx = &m->verts[2 * (size_t)nodes[node]];
        J[0][0] += GNix[node] * vx[0];
        J[0][1] += GNix[node] * vx[1];
        J[1][0] += GNix[m->nl + node] * vx[0];
        J[1][1] += GNix[m->nl + node] * vx[1];
    }
}

/* Mesh::DetJac, Mesh.cpp:183-198 */
double orc_det_jac(const orc_mesh* m, int el, int pt) {
    double J[2][2];
    double* GNix = (double*)malloc(sizeof(double) * 2 * (size_t)m->nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)m->nl);
    ref_shape_func_derivs(m, pt, GNix);
    jacobian(m, el, GNix, J, nodes);
    free(GNix);
    free(nodes);
    return J[0][0] * J[1][1] - J[0][1] * J[1][0];
}

/* global derivatives from local ones through the inverted Jacobian,
 * Mesh.cpp:227-249 */
static void global_derivs(const orc_mesh* m, int el, const double* GNix, double* gGNix,
                          double* detJac, int* nodes) {
    double J[2][2], temp;
    int d, dj, node;
    jacobian(m, el, GNix, J, nodes);
    *detJac = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    temp = J[0][0];
    J[0][0] = J[1][1] / (*detJac);
    J[1][1] = temp / (*detJac);
    J[0][1] = -J[0][1] / (*detJac);
    J[1][0] = -J[1][0] / (*detJac);
    for (d = 0; d < 2; d++) {
        for (node = 0; node < m->nl; node++) {
            temp = 0.0;
            for (dj = 0; dj < 2; dj++) temp += GNix[dj * m->nl + node] * J[d][dj];
            gGNix[d * m->nl + node] = temp;
        }
    }
}

/* Mesh::ShapeFuncDerivs, Mesh.cpp:217-250 */
void orc_shape_func_derivs(const orc_mesh* m, int el, int pt, double* gGNix, double* detJac) {
    double* GNix = (double*)malloc(sizeof(double) * 2 * (size_t)m->nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)m->nl);
    ref_shape_func_derivs(m, pt, GNix);
    global_derivs(m, el, GNix, gGNix, detJac, nodes);
    free(GNix);
    free(nodes);
}

/* ------------------------------------------------------------------------- */
/* a5-a7: element operators                                                  */
/* ------------------------------------------------------------------------- */

/* Laplacian::AssembleElement, Operator.cpp:63-84.  Adds into K (nl x nl). */
void orc_elem_laplacian(const orc_mesh* m, int el, double c, int detj_mode, double* K) {
    int nl = m->nl, pt, r, col;
    double* gGNix = (double*)malloc(sizeof(double) * 2 * (size_t)nl);
    for (pt = 0; pt < nl; pt++) {
        int px = pt % (m->N + 1), py = pt / (m->N + 1);
        double weight = m->weights[px] * m->weights[py];
        double detJac, w;
        orc_shape_func_derivs(m, el, pt, gGNix, &detJac);
        if (detj_mode == 0) detJac = 1.0;
        w = detJac * weight * c;
        for (r = 0; r < nl; r++)
            for (col = 0; col < nl; col++)
                K[r * nl + col] += w * (gGNix[r] * gGNix[col] + gGNix[nl + r] * gGNix[nl + col]);
    }
    free(gGNix);
}

/* MassMatrix::AssembleElement, Operator.cpp:157-177 (GLL collocation: the
 * shape functions at quadrature points are Kronecker deltas). */
void orc_elem_mass(const orc_mesh* m, int el, double c, double* M) {
    int nl = m->nl, pt, r, col;
    for (pt = 0; pt < nl; pt++) {
        int px = pt % (m->N + 1), py = pt / (m->N + 1);
        double weight = m->weights[px] * m->weights[py];
        double detJac = orc_det_jac(m, el, pt);
        for (r = 0; r < nl; r++)
            for (col = 0; col < nl; col++) {
                double Nr = (r == pt) ? 1.0 : 0.0, Nc = (col == pt) ? 1.0 : 0.0;
                M[r * nl + col] += detJac * weight * c * Nr * Nc;
            }
    }
}

/* FieldRHS::AssembleElement, RHSOp.cpp:31-50 (nDofs = 1).  Adds into rhs. */
void orc_elem_field_rhs(const orc_mesh* m, int el, double c, const double* field, double* rhs) {
    int nl = m->nl, N = m->N, pt, node;
    double* Ni = (double*)malloc(sizeof(double) * (size_t)nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)nl);
    for (pt = 0; pt < nl; pt++) {
        int px = pt % (N + 1), py = pt / (N + 1);
        double coord[2] = {m->abcissa[px], m->abcissa[py]};
        double weight = m->weights[px] * m->weights[py];
        double detJac = orc_det_jac(m, el, pt), f, w;
        interp_local(m, field, 1, el, coord, &f, Ni, nodes);
        w = detJac * weight * c;
        for (node = 0; node < nl; node++) {
            double Nn = (node == pt) ? 1.0 : 0.0;
            rhs[node] += w * f * Nn;
        }
    }
    free(Ni);
    free(nodes);
}

/* Legendre::ModalBasis (Legendre.cpp:307-338): monomial coefficients of the
 * hierarchical modes, phi[k*(N+1) + i] the coefficient of x^i in phi_k:
 * phi_0 = (1 + x)/2, phi_1 = (1 - x)/2, phi_k = (1 - x^2) P_{k-2} / 4 with
 * P_j the Jacobi(j, 1) polynomials of Jacobi::Gen (Jacobi.cpp:21-36):
 * P_0 = 1, P_1 = 2x. */
static void modal_basis(int N, double* phi) {
    int n1 = N + 1, k, i;
    double* P = (double*)calloc((size_t)n1 * n1, sizeof(double)); /* P[j*n1 + i] */
    memset(phi, 0, sizeof(double) * (size_t)n1 * n1);
    P[0] = 1.0;
    if (N >= 1) P[n1 + 1] = 2.0;
    for (k = 2; k <= N; k++) {
        const double a = 1.0;
        double c1 = 2 * k * (2 * a + k) * (2 * a + 2 * k - 2);
        double c2 = (2 * a + 2 * k - 1) * (2 * a + 2 * k) * (2 * a + 2 * k - 2);
        double c3 = 2 * (a + k - 1) * (a + k - 1) * (2 * a + 2 * k);
        c2 /= c1;
        c3 /= c1;
        for (i = 0; i <= k - 1; i++) P[k * n1 + i + 1] += c2 * P[(k - 1) * n1 + i];
        for (i = 0; i <= k - 2; i++) P[k * n1 + i] -= c3 * P[(k - 2) * n1 + i];
    }
    phi[0] = 0.5;
    phi[1] = 0.5;
    if (N >= 1) {
        phi[n1 + 0] = 0.5;
        phi[n1 + 1] = -0.5;
    }
    for (k = 2; k <= N; k++)
        for (i = 0; i < k - 1; i++) {
            phi[k * n1 + i] += 0.25 * P[(k - 2) * n1 + i];
            phi[k * n1 + i + 2] -= 0.25 * P[(k - 2) * n1 + i];
        }
    free(P);
}

/* Jacobi::EvalTotal (Jacobi.cpp:48-56): highest power first */
static double eval_total(const double* c, int n, double x) {
    double p = 0.0;
    int i;
    for (i = n; i >= 0; i--) p += c[i] * pow(x, i);
    return p;
}

/* dense products of n x n row-major matrices */
static void mat_mul(const double* A, const double* B, double* C, int n) {
    int i, j, k;
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) {
            double t = 0.0;
            for (k = 0; k < n; k++) t += A[i * n + k] * B[k * n + j];
            C[i * n + j] = t;
        }
}
static void mat_tran(const double* A, double* At, int n) {
    int i, j;
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) At[j * n + i] = A[i * n + j];
}
/* inverse by Gauss-Jordan elimination with full pivoting (the method of
 * Inv, LinAlg.cpp:52-110) */
static void mat_inv(const double* A, double* R, int n) {
    int* piv = (int*)calloc((size_t)n, sizeof(int));
    int* rowof = (int*)malloc(sizeof(int) * (size_t)n);
    int* colof = (int*)malloc(sizeof(int) * (size_t)n);
    int i, j, k, r = 0, c = 0;
    memcpy(R, A, sizeof(double) * (size_t)n * n);
    for (i = 0; i < n; i++) {
        double big = 0.0, inv;
        for (j = 0; j < n; j++)
            if (piv[j] != 1)
                for (k = 0; k < n; k++)
                    if (piv[k] == 0 && fabs(R[j * n + k]) >= big) {
                        big = fabs(R[j * n + k]);
                        r = j;
                        c = k;
                    }
        piv[c]++;
        if (r != c)
            for (k = 0; k < n; k++) {
                double t = R[r * n + k];
                R[r * n + k] = R[c * n + k];
                R[c * n + k] = t;
            }
        rowof[i] = r;
        colof[i] = c;
        inv = 1.0 / R[c * n + c];
        R[c * n + c] = 1.0;
        for (k = 0; k < n; k++) R[c * n + k] *= inv;
        for (j = 0; j < n; j++)
            if (j != c) {
                double d = R[j * n + c];
                R[j * n + c] = 0.0;
                for (k = 0; k < n; k++) R[j * n + k] -= R[c * n + k] * d;
            }
    }
    for (i = n - 1; i >= 0; i--)
        if (rowof[i] != colof[i])
            for (k = 0; k < n; k++) {
                double t = R[k * n + rowof[i]];
                R[k * n + rowof[i]] = R[k * n + colof[i]];
                R[k * n + colof[i]] = t;
            }
    free(piv); free(rowof); free(colof);
}

/* SVV::AssembleElement, SVV.cpp:120-176 (nDofs = 1).  Adds into A (nl x nl)
 * c (Lxx + Lyy) with, for the element's quadrature points i and nodes j,
 * (ix, iy) = (i / (N+1), i % (N+1)) as SVV's GetRowCol numbers them:
 *   B_ij = phi_jx(x_ix) phi_jy(x_iy)       nodal values of the modes
 *   W = diag(DetJac w_i), Q = diag(sqrt(exp(-(N-jx)^2/(mN-jx)^2
 *       - (N-jy)^2/(mN-jy)^2))) on the modes with jx, jy > mN (else 0)
 *   Dx_ij = dN_j/dxi at point i (the reference element's derivative)
 *   M = (B^T W B)^-1 B^T W,  F = M^-1 Q M,  Lxx = (F Dx)^T W (F Dx)
 * and Lyy built from the same x terms as Lxx (SVV.cpp:165). */
void orc_elem_svv(const orc_mesh* m, int el, double c, int mN, double* A) {
    const int N = m->N, n1 = N + 1, nl = m->nl;
    const size_t sz = sizeof(double) * (size_t)nl * nl;
    double *phi = (double*)malloc(sizeof(double) * (size_t)n1 * n1), *GNix = (double*)malloc(sizeof(double) * 2 * nl);
    double *B = (double*)malloc(sz), *W = (double*)calloc((size_t)nl * nl, sizeof(double));
    double *Q = (double*)calloc((size_t)nl * nl, sizeof(double)), *Dx = (double*)malloc(sz);
    double *Bt = (double*)calloc((size_t)nl * nl, sizeof(double)), *BtW = (double*)malloc(sz);
    double *BtWB = (double*)malloc(sz);
    double *BtWBinv = (double*)malloc(sz), *M = (double*)malloc(sz), *Minv = (double*)malloc(sz);
    double *T1 = (double*)malloc(sz), *F = (double*)malloc(sz), *FD = (double*)malloc(sz);
    double *FDt = (double*)malloc(sz), *FDtW = (double*)malloc(sz), *L = (double*)malloc(sz);
    int i, j;
    modal_basis(N, phi);
    for (i = 0; i < nl; i++) {
        int ix = i / n1, iy = i % n1;
        double weight = m->weights[i % n1] * m->weights[i / n1];
        ref_shape_func_derivs(m, i, GNix);
        for (j = 0; j < nl; j++) {
            int jx = j / n1, jy = j % n1;
            B[i * nl + j] = eval_total(&phi[jx * n1], N, m->abcissa[ix]) * eval_total(&phi[jy * n1], N, m->abcissa[iy]);
            if (i == j) {
                W[i * nl + j] = orc_det_jac(m, el, i) * weight;
                if (jx > mN && jy > mN) {
                    double qx = -1.0 * ((double)(N - jx) * (N - jx)) / ((mN - jx) * (mN - jx));
                    double qy = -1.0 * ((double)(N - jy) * (N - jy)) / ((mN - jy) * (mN - jy));
                    Q[i * nl + j] = sqrt(exp(qx + qy));
                }
            }
            Dx[i * nl + j] = GNix[j];
        }
    }
    mat_tran(B, Bt, nl);
    mat_mul(Bt, W, BtW, nl);
    mat_mul(BtW, B, BtWB, nl);
    mat_inv(BtWB, BtWBinv, nl);
    mat_mul(BtWBinv, BtW, M, nl);
    mat_inv(M, Minv, nl);
    mat_mul(Minv, Q, T1, nl);
    mat_mul(T1, M, F, nl);
    mat_mul(F, Dx, FD, nl);
    mat_tran(FD, FDt, nl);
    mat_mul(FDt, W, FDtW, nl);
    mat_mul(FDtW, FD, L, nl); /* Lxx, and Lyy the same */
    for (i = 0; i < nl * nl; i++) A[i] += c * (L[i] + L[i]);
    free(phi); free(GNix); free(B); free(W); free(Q); free(Dx); free(Bt); free(BtW); free(BtWB);
    free(BtWBinv); free(M); free(Minv); free(T1); free(F); free(FD); free(FDt); free(FDtW); free(L);
}

/* ------------------------------------------------------------------------- */
/* a8-a10: assembled operator (PETSc AIJ pattern of dense element blocks)    */
/* ------------------------------------------------------------------------- */

struct orc_helm {
    const orc_mesh* m;
    double c_mass, c_lap;
    int detj_mode;
    long nnz;
    long* rowptr;
    int* cols;
    double* vals;
    double* b;        /* Dirichlet lifting, HelmholtzEqn::Assemble */
    double* dinv;     /* Matrix::InverseDiagonal */
    int assembled;
    double svv_c;     /* SVV operator constant (dt) and cutoff; svv_mN = 0: none */
    int svv_mN;
};

/* HelmholtzEqn's third operator, SVV(solField, solField, dt, svvCutoff)
 * (HelmholtzEqn.cpp:71-76); call before orc_helm_assemble */
void orc_helm_set_svv(orc_helm* h, double dt, int mN) {
    h->svv_c = dt;
    h->svv_mN = mN;
}

orc_helm* orc_helm_create(const orc_mesh* m, double c_mass, double c_lap, int detj_mode) {
    orc_helm* h = (orc_helm*)calloc(1, sizeof(orc_helm));
    h->m = m;
    h->c_mass = c_mass;
    h->c_lap = c_lap;
    h->detj_mode = detj_mode;
    return h;
}

void orc_helm_destroy(orc_helm* h) {
    if (!h) return;
    free(h->rowptr); free(h->cols); free(h->vals); free(h->b); free(h->dinv);
    free(h);
}

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* element block: ops applied in the reference order {mass, laplacian, svv}
 * (HelmholtzEqn.cpp:69-81); PoissonEqn has the Laplacian only. */
static void elem_block(const orc_helm* h, int el, double* Ae) {
    memset(Ae, 0, sizeof(double) * (size_t)h->m->nl * (size_t)h->m->nl);
    if (h->c_mass != 0.0) orc_elem_mass(h->m, el, h->c_mass, Ae);
    orc_elem_laplacian(h->m, el, h->c_lap, h->detj_mode, Ae);
    if (h->svv_mN > 0) orc_elem_svv(h->m, el, h->svv_c, h->svv_mN, Ae);
}

/* Matrix::Assemble, Matrix.cpp:47-133 + HelmholtzEqn::Assemble :91-103.
 * The stored pattern is the union of dense element blocks (MatSetValues of the
 * whole block, zeros included); values accumulate in element order. */
void orc_helm_assemble(orc_helm* h, const double* sol_field) {
    const orc_mesh* m = h->m;
    int nl = m->nl, n = m->n, el, i, j;
    int* nodes = (int*)malloc(sizeof(int) * (size_t)nl);
    int* rn = (int*)malloc(sizeof(int) * (size_t)nl);
    double* Ae = (double*)malloc(sizeof(double) * (size_t)nl * (size_t)nl);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)nl);
    long* cnt = (long*)calloc((size_t)n + 1, sizeof(long));
    long k;
    int* tmp;
    free(h->rowptr); free(h->cols); free(h->vals); free(h->b); free(h->dinv);
    /* pattern: count, fill, sort, unique */
    for (el = 0; el < m->nElsTotal; el++) {
        int nv = 0;
        orc_el_nodes(m, el, nodes);
        for (i = 0; i < nl; i++) if (m->f2v[nodes[i]] >= 0) nv++;
        for (i = 0; i < nl; i++) { int r = m->f2v[nodes[i]]; if (r >= 0) cnt[r + 1] += nv; }
    }
    for (i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    tmp = (int*)malloc(sizeof(int) * (size_t)(cnt[n] > 0 ? cnt[n] : 1));
    {
        long* pos = (long*)malloc(sizeof(long) * (size_t)(n + 1));
        memcpy(pos, cnt, sizeof(long) * (size_t)(n + 1));
        for (el = 0; el < m->nElsTotal; el++) {
            orc_el_nodes(m, el, nodes);
            for (i = 0; i < nl; i++) {
                int r = m->f2v[nodes[i]];
                if (r < 0) continue;
                for (j = 0; j < nl; j++) {
                    int cidx = m->f2v[nodes[j]];
                    if (cidx >= 0) tmp[pos[r]++] = cidx;
                }
            }
        }
        free(pos);
    }
    h->rowptr = (long*)malloc(sizeof(long) * (size_t)(n + 1));
    h->rowptr[0] = 0;
    k = 0;
    for (i = 0; i < n; i++) {
        long s = cnt[i], e = cnt[i + 1], t;
        qsort(tmp + s, (size_t)(e - s), sizeof(int), cmp_int);
        for (t = s; t < e; t++)
            if (t == s || tmp[t] != tmp[t - 1]) tmp[k++] = tmp[t];
        h->rowptr[i + 1] = k;
    }
    h->nnz = k;
    h->cols = (int*)malloc(sizeof(int) * (size_t)(k > 0 ? k : 1));
    memcpy(h->cols, tmp, sizeof(int) * (size_t)k);
    free(tmp);
    free(cnt);
    h->vals = (double*)calloc((size_t)(k > 0 ? k : 1), sizeof(double));
    h->b = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    /* values (ADD_VALUES in element order) + lifting (Matrix.cpp:101-124) */
    for (el = 0; el < m->nElsTotal; el++) {
        int addToVec = 0;
        elem_block(h, el, Ae);
        orc_el_nodes(m, el, nodes);
        for (i = 0; i < nl; i++) rn[i] = m->f2v[nodes[i]];
        for (i = 0; i < nl; i++) {
            int r = rn[i];
            long s, e;
            if (r < 0) continue;
            s = h->rowptr[r];
            e = h->rowptr[r + 1];
            for (j = 0; j < nl; j++) {
                int cidx = rn[j];
                long lo = s, hi = e - 1;
                if (cidx < 0) continue;
                while (lo < hi) {
                    long mid = (lo + hi) / 2;
                    if (h->cols[mid] < cidx) lo = mid + 1; else hi = mid;
                }
                h->vals[lo] += Ae[i * nl + j];
            }
        }
        memset(rhs, 0, sizeof(double) * (size_t)nl);
        for (j = 0; j < nl; j++) {
            if (rn[j] == -1) {
                addToVec = 1;
                for (i = 0; i < nl; i++) rhs[i] -= sol_field[nodes[j]] * Ae[i * nl + j];
            }
        }
        if (addToVec)
            for (i = 0; i < nl; i++)
                if (rn[i] >= 0) h->b[rn[i]] += rhs[i];
    }
    /* inverse diagonal, Matrix.cpp:162-184 */
    h->dinv = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    for (i = 0; i < n; i++) {
        long t;
        double d = 0.0;
        for (t = h->rowptr[i]; t < h->rowptr[i + 1]; t++)
            if (h->cols[t] == i) { d = h->vals[t]; break; }
        h->dinv[i] = 1.0 / d;
    }
    h->assembled = 1;
    free(nodes); free(rn); free(Ae); free(rhs);
}

long orc_helm_nnz(const orc_helm* h) { return h->nnz; }

void orc_helm_csr(const orc_helm* h, long* rowptr, int* cols, double* vals) {
    if (rowptr) memcpy(rowptr, h->rowptr, sizeof(long) * (size_t)(h->m->n + 1));
    if (cols) memcpy(cols, h->cols, sizeof(int) * (size_t)h->nnz);
    if (vals) memcpy(vals, h->vals, sizeof(double) * (size_t)h->nnz);
}

void orc_helm_lift(const orc_helm* h, double* b) { memcpy(b, h->b, sizeof(double) * (size_t)h->m->n); }
void orc_helm_inv_diag(const orc_helm* h, double* d) { memcpy(d, h->dinv, sizeof(double) * (size_t)h->m->n); }

/* MatMult_SeqAIJ: y_i = sum over the sorted row of a_ij x_j */
/* Rows are independent: the OpenMP split changes nothing in any row's sum. */
void orc_helm_matvec(const orc_helm* h, const double* x, double* y) {
    int i, n = h->m->n;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) {
        double s = 0.0;
        long t;
        for (t = h->rowptr[i]; t < h->rowptr[i + 1]; t++) s += h->vals[t] * x[h->cols[t]];
        y[i] = s;
    }
}

/* Vector::Assemble of FieldRHS(c = 1) (Vector.cpp:89-121) followed by
 * VecAXPY(f, 1, b) (HelmholtzEqn.cpp:111-118). */
void orc_helm_rhs(const orc_helm* h, const double* rhs_field, double* f) {
    const orc_mesh* m = h->m;
    int nl = m->nl, el, i, n = m->n;
    int* nodes = (int*)malloc(sizeof(int) * (size_t)nl);
    double* re = (double*)malloc(sizeof(double) * (size_t)nl);
    memset(f, 0, sizeof(double) * (size_t)n);
    for (el = 0; el < m->nElsTotal; el++) {
        memset(re, 0, sizeof(double) * (size_t)nl);
        orc_elem_field_rhs(m, el, 1.0, rhs_field, re);
        orc_el_nodes(m, el, nodes);
        for (i = 0; i < nl; i++) {
            int r = m->f2v[nodes[i]];
            if (r >= 0) f[r] += re[i];
        }
    }
    for (i = 0; i < n; i++) f[i] = f[i] + 1.0 * h->b[i];
    free(nodes);
    free(re);
}

/* Sums over fixed blocks of DOT_BLOCK entries, the block sums added in block
 * order: the same result for any number of threads. */
#define DOT_BLOCK 4096
static double dot(const double* a, const double* b, int n) {
    int nb = (n + DOT_BLOCK - 1) / DOT_BLOCK, k;
    double s = 0.0;
    double* part = (double*)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
#pragma omp parallel for schedule(static)
    for (k = 0; k < nb; k++) {
        int i, i1 = (k + 1) * DOT_BLOCK < n ? (k + 1) * DOT_BLOCK : n;
        double t = 0.0;
        for (i = k * DOT_BLOCK; i < i1; i++) t += a[i] * b[i];
        part[k] = t;
    }
    for (k = 0; k < nb; k++) s += part[k];
    free(part);
    return s;
}

static void remove_mean(double* v, int n) {
    double s = 0.0;
    int i;
    for (i = 0; i < n; i++) s += v[i];
    s /= n;
    for (i = 0; i < n; i++) v[i] -= s;
}

/* number of threads of the oracle's parallel loops (0: the OpenMP default) */
void orc_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Jacobi-preconditioned CG, the build's definition of the reference's
 * KSPSolve(cg, jacobi) (parity of PETSc internals unpinned):
 *   sign-normalised to an SPD system (the QG operators are negative definite),
 *   x0 = 0, null space = constant vector (b and every preconditioned residual
 *   projected to zero Euclidean mean), stop when the recursive residual
 *   satisfies ||r|| <= rtol ||b|| AND the recomputed true residual does too
 *   (otherwise the true residual replaces r and iteration continues). */
int orc_helm_pcg(const orc_helm* h, const double* f, double* x, double rtol, int maxit,
                 int null_space, int* iters, double* true_relres, double* hist) {
    int n = h->m->n, i, k, status = 1;
    double sgn = (n > 0 && h->dinv[0] < 0.0) ? -1.0 : 1.0;
    double *bb = (double*)malloc(sizeof(double) * (size_t)n), *r = (double*)malloc(sizeof(double) * (size_t)n);
    double *z = (double*)malloc(sizeof(double) * (size_t)n), *p = (double*)malloc(sizeof(double) * (size_t)n);
    double *q = (double*)malloc(sizeof(double) * (size_t)n);
    double bnorm, rho, rnorm, tnorm = 0.0;
    for (i = 0; i < n; i++) {
        if ((h->dinv[i] < 0.0) != (sgn < 0.0)) { status = 2; goto done_early; }
        bb[i] = sgn * f[i];
        x[i] = 0.0;
    }
    if (null_space) remove_mean(bb, n);
    bnorm = sqrt(dot(bb, bb, n));
    if (!isfinite(bnorm)) { status = 3; goto done_early; }
    if (hist) hist[0] = 1.0;
    if (bnorm == 0.0) { status = 0; k = 0; tnorm = 0.0; goto done; }
    for (i = 0; i < n; i++) { r[i] = bb[i]; z[i] = sgn * h->dinv[i] * r[i]; }
    if (null_space) remove_mean(z, n);
    memcpy(p, z, sizeof(double) * (size_t)n);
    rho = dot(r, z, n);
    for (k = 1; k <= maxit; k++) {
        double pq, alpha, rho_new, beta;
        orc_helm_matvec(h, p, q);
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; i++) q[i] *= sgn;
        pq = dot(p, q, n);
        if (!(pq > 0.0)) { status = isfinite(pq) ? 2 : 3; break; }
        alpha = rho / pq;
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        rnorm = sqrt(dot(r, r, n));
        if (hist) hist[k] = rnorm / bnorm;
        if (!isfinite(rnorm)) { status = 3; break; }
        if (rnorm <= rtol * bnorm) {
            orc_helm_matvec(h, x, q);
            for (i = 0; i < n; i++) q[i] = bb[i] - sgn * q[i];
            if (null_space) remove_mean(q, n);
            tnorm = sqrt(dot(q, q, n));
            if (tnorm <= rtol * bnorm) { status = 0; break; }
            memcpy(r, q, sizeof(double) * (size_t)n);
        }
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; i++) z[i] = sgn * h->dinv[i] * r[i];
        if (null_space) remove_mean(z, n);
        rho_new = dot(r, z, n);
        beta = rho_new / rho;
        rho = rho_new;
#pragma omp parallel for schedule(static)
        for (i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    if (k > maxit) k = maxit;
done:
    if (status != 0 || tnorm == 0.0) {
        orc_helm_matvec(h, x, q);
        for (i = 0; i < n; i++) q[i] = bb[i] - sgn * q[i];
        if (null_space) remove_mean(q, n);
        tnorm = sqrt(dot(q, q, n));
    }
    if (iters) *iters = k;
    if (true_relres) *true_relres = bnorm > 0.0 ? tnorm / bnorm : 0.0;
    free(bb); free(r); free(z); free(p); free(q);
    return status;
done_early:
    if (iters) *iters = 0;
    if (true_relres) *true_relres = NAN;
    free(bb); free(r); free(z); free(p); free(q);
    return status;
}

/* Field::PeriodicUpdate, Field.cpp:232-252 */
void orc_field_periodic_update(const orc_mesh* m, double* field, int nDofs) {
    int nv0 = m->nVerts[0], nv1 = m->nVerts[1], i, d;
    if (m->periodic[0])
        for (i = 0; i < nv1; i++)
            for (d = 0; d < nDofs; d++)
                field[(size_t)(i * nv0 + nv0 - 1) * nDofs + d] = field[(size_t)(i * nv0) * nDofs + d];
    if (m->periodic[1])
        for (i = 0; i < nv0; i++)
            for (d = 0; d < nDofs; d++)
                field[(size_t)(nv0 * (nv1 - 1) + i) * nDofs + d] = field[(size_t)i * nDofs + d];
    if (m->periodic[0] && m->periodic[1])
        for (d = 0; d < nDofs; d++) field[(size_t)(nv0 * nv1 - 1) * nDofs + d] = field[d];
}

/* Vector::UpdateField, Vector.cpp:40-61 */
void orc_update_field(const orc_mesh* m, const double* x, double* field) {
    int i;
    for (i = 0; i < m->n; i++) field[m->v2f[i]] = x[i];
    if (m->periodic[0] || m->periodic[1]) orc_field_periodic_update(m, field, 1);
}

/* Vector::UpdateVec, Vector.cpp:63-81 */
void orc_update_vec(const orc_mesh* m, const double* field, double* x) {
    int v;
    for (v = 0; v < m->nV; v++)
        if (!is_bc_node(m, v)) x[m->f2v[v]] = field[v];
}

/* HelmholtzEqn::Solve / PoissonEqn::Solve (HelmholtzEqn.cpp:105-136) */
int orc_helm_solve(orc_helm* h, const double* rhs_field, double* sol_field, double rtol,
                   int maxit, int null_space, int* iters, double* true_relres) {
    int n = h->m->n, st;
    double* f = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* x = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!h->assembled) orc_helm_assemble(h, sol_field);
    orc_helm_rhs(h, rhs_field, f);
    st = orc_helm_pcg(h, f, x, rtol, maxit, null_space, iters, true_relres, NULL);
    orc_update_field(h->m, x, sol_field);
    free(f);
    free(x);
    return st;
}

/* Matrix-free form of c_mass*M + c_lap*K~ with ideal affine geometry
 * (J = diag(dx/2, dy/2)): per element
 *   A_e u = c_mass*detJ*(w (x) w).u + c_lap*s*[ax W_y (x) G_x + ay G_y (x) W_x] u,
 * G = D^T W D, ax = (2/dx)^2, ay = (2/dy)^2, s = 1 (literal) or detJ (physical).
 * Summed over elements through the numbering map (Dirichlet columns are zero). */
void orc_helm_apply_ideal(const orc_mesh* m, double c_mass, double c_lap, int detj_mode,
                          const double* x, double* y) {
    int N = m->N, n1 = N + 1, nl = m->nl, el, i, j, k, q;
    double* G = (double*)calloc((size_t)n1 * n1, sizeof(double));
    double* u = (double*)malloc(sizeof(double) * (size_t)nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)nl);
    double detJ = 0.25 * m->dx[0] * m->dx[1];
    double ax = (2.0 / m->dx[0]) * (2.0 / m->dx[0]), ay = (2.0 / m->dx[1]) * (2.0 / m->dx[1]);
    double s = detj_mode ? detJ : 1.0;
    for (i = 0; i < n1; i++)
        for (j = 0; j < n1; j++)
            for (q = 0; q < n1; q++) G[i * n1 + j] += m->weights[q] * m->dCij[q * n1 + i] * m->dCij[q * n1 + j];
    memset(y, 0, sizeof(double) * (size_t)m->n);
    for (el = 0; el < m->nElsTotal; el++) {
        orc_el_nodes(m, el, nodes);
        for (i = 0; i < nl; i++) {
            int g = m->f2v[nodes[i]];
            u[i] = g >= 0 ? x[g] : 0.0;
        }
        for (j = 0; j < n1; j++)
            for (i = 0; i < n1; i++) {
                int g = m->f2v[nodes[j * n1 + i]];
                double sx = 0.0, sy = 0.0, v;
                if (g < 0) continue;
                for (k = 0; k < n1; k++) {
                    sx += G[i * n1 + k] * u[j * n1 + k];
                    sy += G[j * n1 + k] * u[k * n1 + i];
                }
                v = c_mass * detJ * m->weights[i] * m->weights[j] * u[j * n1 + i] +
                    c_lap * s * (ax * m->weights[j] * sx + ay * m->weights[i] * sy);
                y[g] += v;
            }
    }
    free(G);
    free(u);
    free(nodes);
}

/* ------------------------------------------------------------------------- */
/* a11-a16: semi-Lagrangian advection                                        */
/* ------------------------------------------------------------------------- */

/* Advector::PeriodicUpdate, Advector.cpp:122-134 */
static void coord_wrap(const orc_mesh* m, double* c) {
    int d;
    for (d = 0; d < 2; d++) {
        if (c[d] < m->min[d]) c[d] = m->periodic[d] ? m->max[d] - m->min[d] + c[d] : m->min[d];
        if (c[d] > m->max[d]) c[d] = m->periodic[d] ? m->min[d] - m->max[d] + c[d] : m->max[d];
    }
}

/* Advector::InterpLagrange, Advector.cpp:248-275 (nDofs <= 3) */
static void interp_lagrange(double x, const double* coords, int nDofs, double vals[4][3], double* psi) {
    int node, d, o, oc, others[3];
    double factor;
    for (d = 0; d < nDofs; d++) psi[d] = 0.0;
    for (node = 0; node < 4; node++) {
        oc = 0;
        for (o = 0; o < 4; o++) if (o != node) others[oc++] = o;
        factor = 1.0;
        for (o = 0; o < 3; o++) factor *= (x - coords[others[o]]) / (coords[node] - coords[others[o]]);
        for (d = 0; d < nDofs; d++) psi[d] += vals[node][d] * factor;
    }
}

/* Advector::Interpolate (bicubic, used when N < 3), Advector.cpp:278-337 */
static void interp_bicubic(const orc_mesh* m, const double* vals, int nDofs, const double* coord, double* psi) {
    int el, x0, y0, xi, yi, d, node;
    int* nodes = (int*)malloc(sizeof(int) * (size_t)m->nl);
    int idx[4][4];
    double px[4], py[4], l[2];
    double ptsX[4][3], ptsY[4][3];
    double ddx = m->dx[0] / m->N, ddy = m->dx[1] / m->N;
    orc_global_to_local(m, coord, &el, l);
    orc_el_nodes(m, el, nodes);
    node = nodes[0];
    x0 = node % m->nVerts[0];
    y0 = (node / m->nVerts[0]) % m->nVerts[1];
    if (coord[0] <= m->verts[2 * (size_t)nodes[1] + 0]) x0--;
    if (coord[1] <= m->verts[2 * (size_t)nodes[3] + 1]) y0--;
    if (coord[0] <= m->min[0] + ddx) x0++;
    else if (coord[0] >= m->max[0] - ddx) x0--;
    if (coord[1] <= m->min[1] + ddy) y0++;
    else if (coord[1] >= m->max[1] - ddy) y0--;
    for (yi = 0; yi < 4; yi++)
        for (xi = 0; xi < 4; xi++) idx[xi][yi] = x0 + xi + (y0 + yi) * m->nVerts[0];
    for (xi = 0; xi < 4; xi++) px[xi] = m->verts[2 * (size_t)idx[xi][0] + 0];
    for (yi = 0; yi < 4; yi++) py[yi] = m->verts[2 * (size_t)idx[0][yi] + 1];
    for (yi = 0; yi < 4; yi++) {
        for (xi = 0; xi < 4; xi++)
            for (d = 0; d < nDofs; d++) ptsX[xi][d] = vals[(size_t)idx[xi][yi] * nDofs + d];
        interp_lagrange(coord[0], px, nDofs, ptsX, ptsY[yi]);
    }
    interp_lagrange(coord[1], py, nDofs, ptsY, psi);
    free(nodes);
}

/* Advector::IntegrateRK2, Advector.cpp:136-148 */
static void integrate_rk2(const orc_mesh* m, double dt, const double* vel0, const double* velH,
                          const double* origin, double* coord) {
    double xh[2], v[2];
    orc_interp_global(m, vel0, 2, origin, v);
    xh[0] = origin[0] - 0.5 * dt * v[0];
    xh[1] = origin[1] - 0.5 * dt * v[1];
    coord_wrap(m, xh);
    orc_interp_global(m, velH, 2, xh, v);
    coord[0] = origin[0] - dt * v[0];
    coord[1] = origin[1] - dt * v[1];
    coord_wrap(m, coord);
}

/* Advector::FirstOrder, Advector.cpp:181-201 (q: nV x nDofs, vel: nV x 2).
 * dep_el / dep_xy (nullable) receive the departure element and point. */
/* Advector::Interpolate at one point (exported for the drop-in check) */
void orc_interp_bicubic(const orc_mesh* m, const double* vals, int nDofs, const double* coord, double* psi) {
    interp_bicubic(m, vals, nDofs, coord, psi);
}

/* Advector::FirstOrder with the interpolation of the advected field chosen
 * by `interp`: 0 the reference's default (bicubic when N < 3, the
 * constructors' cubicInterp), 1 GLL (Field::InterpGlobal), 2 bicubic
 * (cubicInterp set true by the caller, QuasiGeostrophicSpectral.cpp:84). */
void orc_advect_first_interp(const orc_mesh* m, const double* q, int nDofs, const double* vel,
                             double dt, double* qSL, int interp) {
    int v, d;
    double c[2], val[3];
    const int cubic = interp == 2 || (interp == 0 && m->N < 3);
    for (v = 0; v < m->nV; v++) {
        integrate_rk2(m, dt, vel, vel, &m->verts[2 * (size_t)v], c);
        if (cubic) interp_bicubic(m, q, nDofs, c, val);
        else orc_interp_global(m, q, nDofs, c, val);
        for (d = 0; d < nDofs; d++) qSL[(size_t)v * nDofs + d] = val[d];
    }
    if (m->periodic[0] || m->periodic[1]) orc_field_periodic_update(m, qSL, nDofs);
}

void orc_advect_first(const orc_mesh* m, const double* q, int nDofs, const double* vel,
                      double dt, double* qSL, int* dep_el, double* dep_xy) {
    int v, d;
    double c[2], val[3];
    for (v = 0; v < m->nV; v++) {
        integrate_rk2(m, dt, vel, vel, &m->verts[2 * (size_t)v], c);
        if (m->N < 3) interp_bicubic(m, q, nDofs, c, val);
        else orc_interp_global(m, q, nDofs, c, val);
        for (d = 0; d < nDofs; d++) qSL[(size_t)v * nDofs + d] = val[d];
        if (dep_el) orc_get_el(m, c, &dep_el[v]);
        if (dep_xy) { dep_xy[2 * (size_t)v] = c[0]; dep_xy[2 * (size_t)v + 1] = c[1]; }
    }
    if (m->periodic[0] || m->periodic[1]) orc_field_periodic_update(m, qSL, nDofs);
}

/* Interpolation points of the RK2 trajectories (Advector.cpp:136-148 for
 * order 1: K = 2 points per node {x_half, x_dep}; Advector.cpp:203-246 for
 * order 2: K = 4 points {x_half, x1, x_half2, x2}).  pts: nV x K x 2.  Used by
 * the tests to locate departure points where the reference's cardinal-function
 * formula is ill-conditioned (close to, but not within 1e-8 of, a GLL line). */
void orc_sl_points(const orc_mesh* m, int order, const double* vel, const double* velPrev, double dt,
                   double* pts) {
    size_t nv2 = 2 * (size_t)m->nV, i;
    int v, K = order == 2 ? 4 : 2;
    double *vp1 = NULL, *vph = NULL, *vmh = NULL;
    if (order == 2) {
        vp1 = (double*)malloc(sizeof(double) * nv2);
        vph = (double*)malloc(sizeof(double) * nv2);
        vmh = (double*)malloc(sizeof(double) * nv2);
        for (i = 0; i < nv2; i++) {
            vp1[i] = 2.0 * vel[i] - 1.0 * velPrev[i];
            vph[i] = 1.5 * vel[i] - 0.5 * velPrev[i];
            vmh[i] = 0.5 * vel[i] + 0.5 * velPrev[i];
        }
    }
    for (v = 0; v < m->nV; v++) {
        const double* o = &m->verts[2 * (size_t)v];
        double* P = &pts[(size_t)v * K * 2];
        double xh[2], w[2], x1[2], x2[2];
        orc_interp_global(m, order == 2 ? vp1 : vel, 2, o, w);
        xh[0] = o[0] - 0.5 * dt * w[0];
        xh[1] = o[1] - 0.5 * dt * w[1];
        coord_wrap(m, xh);
        orc_interp_global(m, order == 2 ? vph : vel, 2, xh, w);
        x1[0] = o[0] - dt * w[0];
        x1[1] = o[1] - dt * w[1];
        coord_wrap(m, x1);
        P[0] = xh[0]; P[1] = xh[1]; P[2] = x1[0]; P[3] = x1[1];
        if (order == 2) {
            orc_interp_global(m, vel, 2, x1, w);
            xh[0] = x1[0] - 0.5 * dt * w[0];
            xh[1] = x1[1] - 0.5 * dt * w[1];
            coord_wrap(m, xh);
            orc_interp_global(m, vmh, 2, xh, w);
            x2[0] = x1[0] - dt * w[0];
            x2[1] = x1[1] - dt * w[1];
            coord_wrap(m, x2);
            P[4] = xh[0]; P[5] = xh[1]; P[6] = x2[0]; P[7] = x2[1];
        }
    }
    free(vp1); free(vph); free(vmh);
}

/* Advector::ExtrapolateVelocities + SecondOrder, Advector.cpp:203-246,339-349 */
void orc_advect_second(const orc_mesh* m, const double* q, const double* qPrev, int nDofs,
                       const double* vel, const double* velPrev, double dt, double* qSL) {
    size_t nv2 = 2 * (size_t)m->nV, i;
    double* vp1 = (double*)malloc(sizeof(double) * nv2);
    double* vph = (double*)malloc(sizeof(double) * nv2);
    double* vmh = (double*)malloc(sizeof(double) * nv2);
    int v, d;
    for (i = 0; i < nv2; i++) {
        vp1[i] = 2.0 * vel[i] - 1.0 * velPrev[i];
        vph[i] = 1.5 * vel[i] - 0.5 * velPrev[i];
        vmh[i] = 0.5 * vel[i] + 0.5 * velPrev[i];
    }
    for (v = 0; v < m->nV; v++) {
        const double* o = &m->verts[2 * (size_t)v];
        double x1[2], x2[2], xh[2], w[2], p1[3], p2[3];
        orc_interp_global(m, vp1, 2, o, w);
        xh[0] = o[0] - 0.5 * dt * w[0];
        xh[1] = o[1] - 0.5 * dt * w[1];
        coord_wrap(m, xh);
        orc_interp_global(m, vph, 2, xh, w);
        x1[0] = o[0] - dt * w[0];
        x1[1] = o[1] - dt * w[1];
        coord_wrap(m, x1);
        orc_interp_global(m, vel, 2, x1, w);
        xh[0] = x1[0] - 0.5 * dt * w[0];
        xh[1] = x1[1] - 0.5 * dt * w[1];
        coord_wrap(m, xh);
        orc_interp_global(m, vmh, 2, xh, w);
        x2[0] = x1[0] - dt * w[0];
        x2[1] = x1[1] - dt * w[1];
        coord_wrap(m, x2);
        if (m->N < 3) {
            interp_bicubic(m, q, nDofs, x1, p1);
            interp_bicubic(m, qPrev, nDofs, x2, p2);
        } else {
            orc_interp_global(m, q, nDofs, x1, p1);
            orc_interp_global(m, qPrev, nDofs, x2, p2);
        }
        for (d = 0; d < nDofs; d++) qSL[(size_t)v * nDofs + d] = 2.0 * p1[d] - 0.5 * p2[d];
    }
    if (m->periodic[0] || m->periodic[1]) orc_field_periodic_update(m, qSL, nDofs);
    free(vp1); free(vph); free(vmh);
}

/* Legendre::ShapeFuncDerivsAtCoord, Legendre.cpp:187-223 */
static void shape_func_derivs_at_coord(const orc_mesh* m, const double* coord, double* GNix) {
    int N = m->N, nl = m->nl, pt_j, x_i, y_i, x_j, y_j;
    double Nx, Ny, GNx, GNy, a, b, c;
    for (pt_j = 0; pt_j < nl; pt_j++) {
        x_j = pt_j % (N + 1);
        y_j = pt_j / (N + 1);
        if (is_root(m, coord[0], &x_i)) {
            Nx = (x_i == x_j) ? 1.0 : 0.0;
            GNx = m->dCij[x_i * (N + 1) + x_j];
        } else {
            Nx = -((1.0 - coord[0] * coord[0]) /
                   (N * (N + 1) * orc_leg_eval(m->pn, N, m->abcissa[x_j]) * (coord[0] - m->abcissa[x_j]))) *
                 orc_leg_eval_deriv(m->pn, N, coord[0]);
            a = coord[0] * coord[0] + 1.0 - 2.0 * coord[0] * m->abcissa[x_j];
            b = (1.0 - coord[0] * coord[0]) * (coord[0] - m->abcissa[x_j]);
            c = N * (N + 1) * (coord[0] - m->abcissa[x_j]) * (coord[0] - m->abcissa[x_j]) *
                orc_leg_eval(m->pn, N, m->abcissa[x_j]);
            GNx = (a * orc_leg_eval_deriv(m->pn, N, coord[0]) - b * orc_leg_eval_2nd(m->pn, N, coord[0])) / c;
        }
        if (is_root(m, coord[1], &y_i)) {
            Ny = (y_i == y_j) ? 1.0 : 0.0;
            GNy = m->dCij[y_i * (N + 1) + y_j];
        } else {
            Ny = -((1.0 - coord[1] * coord[1]) /
                   (N * (N + 1) * orc_leg_eval(m->pn, N, m->abcissa[y_j]) * (coord[1] - m->abcissa[y_j]))) *
                 orc_leg_eval_deriv(m->pn, N, coord[1]);
            a = coord[1] * coord[1] + 1.0 - 2.0 * coord[1] * m->abcissa[y_j];
            b = (1.0 - coord[1] * coord[1]) * (coord[1] - m->abcissa[y_j]);
            c = N * (N + 1) * (coord[1] - m->abcissa[y_j]) * (coord[1] - m->abcissa[y_j]) *
                orc_leg_eval(m->pn, N, m->abcissa[y_j]);
            GNy = (a * orc_leg_eval_deriv(m->pn, N, coord[1]) - b * orc_leg_eval_2nd(m->pn, N, coord[1])) / c;
        }
        GNix[pt_j] = GNx * Ny;
        GNix[nl + pt_j] = Nx * GNy;
    }
}

/* QuasiGeostrophicEqn::GenVel, QuasiGeostrophicEqn.cpp:207-223, through
 * Field::InterpDerivsGlobal (Field.cpp:96-118) and
 * Mesh::ShapeFuncDerivsAtCoord (Mesh.cpp:252-285).  vel is nV x 2. */
/* Field::InterpDerivsGlobal of a scalar field at global coordinate g:
 * d = (d/dx, d/dy) in the element Mesh::GlobalToLocal picks.  GNix, gGNix
 * (2 nl) and nodes (nl) are work arrays. */
static void interp_derivs_global(const orc_mesh* m, const double* vals, const double* g, double* d,
                                 double* GNix, double* gGNix, int* nodes) {
    int el, node, nl = m->nl;
    double l[2], detJac;
    orc_global_to_local(m, g, &el, l);
    shape_func_derivs_at_coord(m, l, GNix);
    global_derivs(m, el, GNix, gGNix, &detJac, nodes);
    d[0] = d[1] = 0.0;
    for (node = 0; node < nl; node++) {
        d[0] += gGNix[node] * vals[nodes[node]];
        d[1] += gGNix[nl + node] * vals[nodes[node]];
    }
}

void orc_genvel(const orc_mesh* m, const double* psi, double* vel) {
    int nl = m->nl, v;
    double* GNix = (double*)malloc(sizeof(double) * 2 * (size_t)nl);
    double* gGNix = (double*)malloc(sizeof(double) * 2 * (size_t)nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)nl);
    for (v = 0; v < m->nV; v++) {
        double d[2];
        interp_derivs_global(m, psi, &m->verts[2 * (size_t)v], d, GNix, gGNix, nodes);
        vel[2 * (size_t)v + 0] = -1.0 * d[1];
        vel[2 * (size_t)v + 1] = +1.0 * d[0];
    }
    free(GNix);
    free(gGNix);
    free(nodes);
}

/* CalcTimeStep, Utils.cpp:23-40 */
double orc_calc_timestep(const orc_mesh* m, const double* vel, double courant) {
    double minSep = (m->dx[0] < m->dx[1]) ? m->dx[0] : m->dx[1];
    double maxVelSq = 0.0;
    int v;
    for (v = 0; v < m->nV; v++) {
        double s = vel[2 * (size_t)v] * vel[2 * (size_t)v] + vel[2 * (size_t)v + 1] * vel[2 * (size_t)v + 1];
        if (s > maxVelSq) maxVelSq = s;
    }
    return courant * minSep / sqrt(maxVelSq);
}
/* ------------------------------------------------------------------------- */
/* callers: QG steps                                                         */
/* ------------------------------------------------------------------------- */

/* Mesh::LocalToGlobal, Mesh.cpp:178-181 */
static void local_to_global(const orc_mesh* m, const double* l, int el, double* g) {
    g[0] = m->min[0] + (el % m->nEls[0]) * m->dx[0] + 0.5 * (l[0] + 1.0) * m->dx[0];
    g[1] = m->min[1] + (el / m->nEls[0]) * m->dx[1] + 0.5 * (l[1] + 1.0) * m->dx[1];
}

/* SolveQ with its dissipation right-hand side.
 *
 * TwoLayerQGEqn::SolveQ (TwoLayerQGEqn.cpp:396-416) over
 * DissipationRHS::AssembleElement (:437-462), and QuasiGeostrophicEqn::SolveQ
 * (QuasiGeostrophicEqn.cpp:277-301) over DisRHS::AssembleElement (:315-337),
 * which is the same form with rF = 0 and Ei = r.  The matrix is AssembleQMat's
 * a x mass matrix (MassMatrix, Operator.cpp:157-177), diagonal under GLL
 * collocation, so its Dirichlet lifting is zero and KSPSolve is the division
 * by its diagonal (to the solver's tolerance).  The right-hand side is
 * assembled element by element through fieldToVecMap as Vector::Assemble
 * (Vector.cpp:89-120) does: Dirichlet rows are dropped and the two copies of a
 * periodic node add into one unknown.  At each quadrature point of each
 * element, with c = detJ w dt,
 *   rhs_i += detJ w qhat N_i + c rF tau N_i - c Ei omegaHat N_i - c rF tau_e N_i
 *            - c nu (dN_i/dx dW/dx + dN_i/dy dW/dy),
 * where the gradient of W = omegaHat comes from Field::InterpDerivsGlobal at
 * the point's global coordinate (Mesh::LocalToGlobal), i.e. in the element
 * Mesh::GlobalToLocal picks there, not necessarily the element being
 * assembled.  Vector::UpdateField then writes the unknowns and the periodic
 * copies into alpha; alpha's Dirichlet nodes keep what they hold.  tau,
 * omegaHat and tau_e may be NULL (read as zero). */
void orc_qg_solveq(const orc_mesh* m, double a, double dt, const double* qhat, double rF, double Ei, double nu,
                   const double* tau, const double* omegaHat, const double* tau_e, double* alpha) {
    int nl = m->nl, N = m->N, n = m->n, el, pt, node, u;
    double* f = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* diag = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* x = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double* rhs = (double*)malloc(sizeof(double) * (size_t)nl);
    double* Ni = (double*)malloc(sizeof(double) * (size_t)nl);
    double* gGNix = (double*)malloc(sizeof(double) * 2 * (size_t)nl);
    double* dGNix = (double*)malloc(sizeof(double) * 2 * (size_t)nl);
    double* dgGNix = (double*)malloc(sizeof(double) * 2 * (size_t)nl);
    int* nodes = (int*)malloc(sizeof(int) * (size_t)nl);
    int* dnodes = (int*)malloc(sizeof(int) * (size_t)nl);
    const int grad = omegaHat && nu != 0.0;
    for (el = 0; el < m->nElsTotal; el++) {
        for (node = 0; node < nl; node++) rhs[node] = 0.0;
        for (pt = 0; pt < nl; pt++) {
            int px = pt % (N + 1), py = pt / (N + 1);
            double coord[2] = {m->abcissa[px], m->abcissa[py]};
            double weight = m->weights[px] * m->weights[py];
            double detJac, c, q, t = 0.0, w = 0.0, te = 0.0, g[2], gW[2] = {0.0, 0.0};
            orc_shape_func_derivs(m, el, pt, gGNix, &detJac);
            c = detJac * weight * dt;
            local_to_global(m, coord, el, g);
            interp_local(m, qhat, 1, el, coord, &q, Ni, nodes);
            if (tau) interp_local(m, tau, 1, el, coord, &t, Ni, nodes);
            if (omegaHat) interp_local(m, omegaHat, 1, el, coord, &w, Ni, nodes);
            if (tau_e) interp_local(m, tau_e, 1, el, coord, &te, Ni, nodes);
            if (grad) interp_derivs_global(m, omegaHat, g, gW, dGNix, dgGNix, dnodes);
            for (node = 0; node < nl; node++) {
                /* the shape functions at a GLL quadrature point are Kronecker deltas */
                double Nn = (node == pt) ? 1.0 : 0.0;
                rhs[node] += detJac * weight * q * Nn;
                rhs[node] += c * rF * t * Nn;
                rhs[node] -= c * Ei * w * Nn;
                rhs[node] -= c * rF * te * Nn;
                rhs[node] -= c * nu * (gGNix[node] * gW[0] + gGNix[nl + node] * gW[1]);
            }
        }
        orc_el_nodes(m, el, nodes);
        for (node = 0; node < nl; node++) {
            u = m->f2v[nodes[node]];
            if (u >= 0) f[u] += rhs[node];
        }
        /* the diagonal of AssembleQMat's matrix */
        for (pt = 0; pt < nl; pt++) {
            int px = pt % (N + 1), py = pt / (N + 1);
            u = m->f2v[nodes[pt]];
            if (u >= 0) diag[u] += orc_det_jac(m, el, pt) * (m->weights[px] * m->weights[py]) * a;
        }
    }
    for (u = 0; u < n; u++) x[u] = f[u] / diag[u];
    orc_update_field(m, x, alpha);
    free(f); free(diag); free(x); free(rhs); free(Ni); free(gGNix); free(dGNix); free(dgGNix);
    free(nodes); free(dnodes);
}

/* The one-layer QG channel step, QuasiGeostrophicEqn::FirstOrder
 * (QuasiGeostrophicEqn.cpp:117-157) and SecondOrder (:159-205), with the
 * dissipation of DisRHS (linear drag r, viscosity nu) and topography.
 *   1. GenVel (:207-223); alpha = omega - F psi + beta y + topo (GenPV
 *      :225-232); SL advection of alpha: alpha^ (first order), or
 *      2 alpha(x1) - alpha_prev(x2) / 2 from the previous state's PV with
 *      extrapolated velocities (second order, Advector.cpp:203-246).
 *   2. he1: (lap - F) psi^ = alpha^ - a beta y - a topo, and
 *      omega^ = alpha^ + F psi^ - a beta y - a topo (a = 1 first order, 1.5
 *      second order).  he1 was assembled by the constructor (:50-53) while
 *      psi^ was zero, so its Dirichlet lifting is zero; psi^'s wall values are
 *      zero until the second-order path sets psi^ = 1.5 psi (:86-93).
 *   3. SolveQ (orc_qg_solveq) with rF = 0, Ei = r on AssembleQMat(a).
 *   4. he2: (lap - F) psi = alpha - beta y - topo with psi's wall data,
 *      omega = alpha + F psi - beta y - topo.
 * Step 2 only feeds the r and nu terms: with both zero it is skipped (its
 * values are multiplied by zero).  psi / omega are updated in place;
 * psi_prev / omega_prev (NULL for the first-order step) are the state before
 * the previous step; topo may be NULL.  Returns the first failing solve's
 * status. */
static int qg_dirichlet(const orc_mesh* m, int v) { return m->f2v[v] < 0; }

int orc_qg_step(const orc_mesh* m, double F, double beta, double nu, double r, const double* topo, double dt,
                int detj_mode, double* psi, double* omega, const double* psi_prev, const double* omega_prev,
                double rtol, int maxit, int* iters) {
    size_t nV = (size_t)m->nV, v;
    const int second = psi_prev && omega_prev;
    const int diss = nu != 0.0 || r != 0.0;
    const double a = second ? 1.5 : 1.0;
    double* vel = (double*)malloc(sizeof(double) * 2 * nV);
    double* velp = (double*)malloc(sizeof(double) * 2 * nV);
    double* alpha = (double*)malloc(sizeof(double) * nV);
    double* alphap = (double*)malloc(sizeof(double) * nV);
    double* qsl = (double*)malloc(sizeof(double) * nV);
    double* rhs = (double*)malloc(sizeof(double) * nV);
    double* phat = (double*)calloc(nV, sizeof(double));
    double* omh = (double*)calloc(nV, sizeof(double));
    orc_helm* h = orc_helm_create(m, -F, -1.0 * 1.0, detj_mode);
    double rr;
    int st = 0, st1 = 0, it = 0, it1 = 0;
    orc_genvel(m, psi, vel);
    for (v = 0; v < nV; v++)
        alpha[v] = omega[v] - F * psi[v] + beta * m->verts[2 * v + 1] + (topo ? topo[v] : 0.0);
    if (second) {
        orc_genvel(m, psi_prev, velp);
        for (v = 0; v < nV; v++)
            alphap[v] = omega_prev[v] - F * psi_prev[v] + beta * m->verts[2 * v + 1] + (topo ? topo[v] : 0.0);
        orc_advect_second(m, alpha, alphap, 1, vel, velp, dt, qsl);
    } else {
        orc_advect_first(m, alpha, 1, vel, dt, qsl, NULL, NULL);
    }
    if (diss) {
        orc_helm* h1 = orc_helm_create(m, -F, -1.0 * 1.0, detj_mode);
        orc_helm_assemble(h1, phat); /* zero wall data */
        for (v = 0; v < nV; v++) {
            double y = m->verts[2 * v + 1];
            rhs[v] = qsl[v] - a * beta * y - (topo ? a * topo[v] : 0.0);
            if (second && qg_dirichlet(m, (int)v)) phat[v] = 1.5 * psi[v];
        }
        st1 = orc_helm_solve(h1, rhs, phat, rtol, maxit, 0, &it1, &rr);
        for (v = 0; v < nV; v++)
            omh[v] = qsl[v] + F * phat[v] - a * beta * m->verts[2 * v + 1] - (topo ? a * topo[v] : 0.0);
        orc_helm_destroy(h1);
    }
    orc_qg_solveq(m, a, dt, qsl, 0.0, r, nu, NULL, diss ? omh : NULL, NULL, alpha);
    for (v = 0; v < nV; v++) rhs[v] = alpha[v] - beta * m->verts[2 * v + 1] - (topo ? topo[v] : 0.0);
    orc_helm_assemble(h, psi);
    st = orc_helm_solve(h, rhs, psi, rtol, maxit, 0, &it, &rr);
    for (v = 0; v < nV; v++)
        omega[v] = alpha[v] + F * psi[v] - beta * m->verts[2 * v + 1] - (topo ? topo[v] : 0.0);
    if (iters) *iters = it1 + it;
    orc_helm_destroy(h);
    free(vel); free(velp); free(alpha); free(alphap); free(qsl); free(rhs); free(phat); free(omh);
    return st1 != 0 ? st1 : st;
}

/* The two-layer step, TwoLayerQGEqn::FirstOrder (TwoLayerQGEqn.cpp:174-243)
 * and SecondOrder (:245-324), with the dissipation of DissipationRHS
 * (thermal relaxation r toward tau_e, Ekman drag E2 on layer 2, viscosity nu).
 *   1. GenVel (:335-352) and GenPV (:326-333) of both layers,
 *      alpha_i = omega_i + F_i (psi_j - psi_i) + beta y, SL advection to
 *      alpha_i^ (second order as in the one-layer step).
 *   2-4. (lap - (F1+F2)) tau^ = alpha1^ - alpha2^,
 *      lap chi^ = alpha1^ + alpha2^ + (F1-F2) tau^ - 2 a beta y, with the wall
 *      data of tau = psi1 - psi2 and chi = psi1 + psi2 (the constructor's,
 *      :46-53, unscaled in the second-order step);
 *      psi1^ = (tau^ + chi^)/2, psi2^ = (chi^ - tau^)/2,
 *      omega_i^ = alpha_i^ - F_i (psi_j^ - psi_i^) - a beta y.
 *   5. SolveQ (orc_qg_solveq): layer 1 with rF = r F1, Ei = 0 (:78-81);
 *      layer 2 with rF = -r F2, Ei = E2 (:83-86), both with tau^ and tau_e.
 *      FirstOrder calls SolveQ for layer 1 only (:214), so there layer 2's
 *      alpha keeps its GenPV value on every node; SecondOrder solves both
 *      (:291-292; its SolveQ of layer 2 passes layer 1's lifting b1, which is
 *      zero like b2 under the diagonal mass).
 *   6-8. (lap - (F1+F2)) tau = alpha1 - alpha2,
 *      lap chi = alpha1 + alpha2 + (F1-F2) tau - 2 beta y, psi_i from tau and
 *      chi, omega_i = alpha_i - F_i (psi_j - psi_i) - beta y.
 * Steps 2-4 only feed the dissipation terms: with r = E2 = nu = 0 they are
 * skipped (their values are multiplied by zero).  tau_e may be NULL. */
int orc_qg2_step(const orc_mesh* m, double F1, double F2, double beta, double E2, double r, double nu,
                 const double* tau_e, double dt, int detj_mode, double* psi1, double* omega1, double* psi2,
                 double* omega2, const double* psi1p, const double* omega1p, const double* psi2p,
                 const double* omega2p, double rtol, int maxit, int* iters) {
    size_t nV = (size_t)m->nV, v;
    const int second = psi1p && omega1p && psi2p && omega2p;
    const int diss = r != 0.0 || E2 != 0.0 || nu != 0.0;
    const double a = second ? 1.5 : 1.0;
    double* vel1 = (double*)malloc(sizeof(double) * 2 * nV);
    double* vel2 = (double*)malloc(sizeof(double) * 2 * nV);
    double* velp1 = (double*)malloc(sizeof(double) * 2 * nV);
    double* velp2 = (double*)malloc(sizeof(double) * 2 * nV);
    double* a1 = (double*)malloc(sizeof(double) * nV);
    double* a2 = (double*)malloc(sizeof(double) * nV);
    double* ap1 = (double*)malloc(sizeof(double) * nV);
    double* ap2 = (double*)malloc(sizeof(double) * nV);
    double* h1 = (double*)malloc(sizeof(double) * nV);
    double* h2 = (double*)malloc(sizeof(double) * nV);
    double* rhs = (double*)malloc(sizeof(double) * nV);
    double* tau = (double*)malloc(sizeof(double) * nV);
    double* chi = (double*)malloc(sizeof(double) * nV);
    double* th = (double*)malloc(sizeof(double) * nV);
    double* ch = (double*)malloc(sizeof(double) * nV);
    double* om1h = (double*)calloc(nV, sizeof(double));
    double* om2h = (double*)calloc(nV, sizeof(double));
    orc_helm* hh = orc_helm_create(m, -(F1 + F2), -1.0 * 1.0, detj_mode);
    orc_helm* hp = orc_helm_create(m, 0.0, -1.0, detj_mode);
    double rr;
    int st = 0, st2 = 0, sth = 0, sth2 = 0, it1 = 0, it2 = 0, ith1 = 0, ith2 = 0;
    orc_genvel(m, psi1, vel1);
    orc_genvel(m, psi2, vel2);
    for (v = 0; v < nV; v++) {
        double y = m->verts[2 * v + 1];
        a1[v] = omega1[v] + F1 * (psi2[v] - psi1[v]) + beta * y;
        a2[v] = omega2[v] + F2 * (psi1[v] - psi2[v]) + beta * y;
        tau[v] = psi1[v] - psi2[v];
        chi[v] = psi1[v] + psi2[v];
    }
    if (second) {
        orc_genvel(m, psi1p, velp1);
        orc_genvel(m, psi2p, velp2);
        for (v = 0; v < nV; v++) {
            double y = m->verts[2 * v + 1];
            ap1[v] = omega1p[v] + F1 * (psi2p[v] - psi1p[v]) + beta * y;
            ap2[v] = omega2p[v] + F2 * (psi1p[v] - psi2p[v]) + beta * y;
        }
        orc_advect_second(m, a1, ap1, 1, vel1, velp1, dt, h1);
        orc_advect_second(m, a2, ap2, 1, vel2, velp2, dt, h2);
    } else {
        orc_advect_first(m, a1, 1, vel1, dt, h1, NULL, NULL);
        orc_advect_first(m, a2, 1, vel2, dt, h2, NULL, NULL);
    }
    orc_helm_assemble(hh, tau);
    orc_helm_assemble(hp, chi);
    if (diss) {
        memcpy(th, tau, sizeof(double) * nV);
        memcpy(ch, chi, sizeof(double) * nV);
        for (v = 0; v < nV; v++) rhs[v] = h1[v] - h2[v];
        sth = orc_helm_solve(hh, rhs, th, rtol, maxit, 0, &ith1, &rr);
        for (v = 0; v < nV; v++) rhs[v] = h1[v] + h2[v] + (F1 - F2) * th[v] - a * 2.0 * beta * m->verts[2 * v + 1];
        sth2 = orc_helm_solve(hp, rhs, ch, rtol, maxit, 0, &ith2, &rr);
        for (v = 0; v < nV; v++) {
            double y = m->verts[2 * v + 1];
            double p1 = 0.5 * (th[v] + ch[v]), p2 = 0.5 * (ch[v] - th[v]);
            om1h[v] = h1[v] - F1 * (p2 - p1) - a * beta * y;
            om2h[v] = h2[v] - F2 * (p1 - p2) - a * beta * y;
        }
    }
    orc_qg_solveq(m, a, dt, h1, r * F1, 0.0, nu, diss ? th : NULL, diss ? om1h : NULL, diss ? tau_e : NULL, a1);
    if (second)
        orc_qg_solveq(m, a, dt, h2, -r * F2, E2, nu, diss ? th : NULL, diss ? om2h : NULL, diss ? tau_e : NULL, a2);
    for (v = 0; v < nV; v++) rhs[v] = a1[v] - a2[v];
    st = orc_helm_solve(hh, rhs, tau, rtol, maxit, 0, &it1, &rr);
    for (v = 0; v < nV; v++) {
        double y = m->verts[2 * v + 1];
        rhs[v] = a1[v] + a2[v] + (F1 - F2) * tau[v] - 2.0 * beta * y;
    }
    st2 = orc_helm_solve(hp, rhs, chi, rtol, maxit, 0, &it2, &rr);
    for (v = 0; v < nV; v++) {
        double y = m->verts[2 * v + 1];
        psi1[v] = 0.5 * (tau[v] + chi[v]);
        psi2[v] = 0.5 * (chi[v] - tau[v]);
        omega1[v] = a1[v] - F1 * (psi2[v] - psi1[v]) - beta * y;
        omega2[v] = a2[v] - F2 * (psi1[v] - psi2[v]) - beta * y;
    }
    if (iters) *iters = ith1 + ith2 + it1 + it2;
    orc_helm_destroy(hh);
    orc_helm_destroy(hp);
    free(vel1); free(vel2); free(velp1); free(velp2); free(a1); free(a2); free(ap1); free(ap2);
    free(h1); free(h2); free(rhs); free(tau); free(chi); free(th); free(ch); free(om1h); free(om2h);
    if (sth) return sth;
    if (sth2) return sth2;
    return st != 0 ? st : st2;
}
