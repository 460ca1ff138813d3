/*
 * mgm_oracle.c -- CPU FP64 restatement of the reference RBF-FD multigrid path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  Only tests/, smoke() and
 * bench.py's cpu_baseline leg may load this library; the product never does.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off -fno-fast-math [-fopenmp]
 * Without FMA contraction every expression rounds exactly as the reference's
 * gcc -O2 x86-64 build does, so results that depend only on +,-,*,/,sqrt and
 * libm pow are bitwise equal to the reference's.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   gmsh reader, kNN, RCM, normals, d_avg, weights (GJ), transfer rows and
 *   the H/T/F V-cycles in natural order are compared against fixtures written
 *   by the compiled reference (tests/golden/).  The colour-ordered sweep is
 *   pinned by reduction: with the trivial colouring (one node per colour in
 *   index order) it is the natural-order sweep.  LU weights are pinned by
 *   polynomial reproduction and by agreement with GJ.
 */
#include "mgm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

uint64_t orc_splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* ===================================================================== */
/* Monomial exponents: i (x) outer, j (y), k (z) inner (hf/mesh_functions.h:131-170). */
int orc_poly_terms(int dim, int deg, short *px, short *py, short *pz)
{
    int c = 0;
    if (dim == 2) {
        for (int i = 0; i <= deg; i++)
            for (int j = 0; j <= deg - i; j++) {
                if (px) { px[c] = (short)i; py[c] = (short)j; pz[c] = 0; }
                c++;
            }
    } else {
        for (int i = 0; i <= deg; i++)
            for (int j = 0; j <= deg - i; j++)
                for (int k = 0; k <= deg - i - j; k++) {
                    if (px) { px[c] = (short)i; py[c] = (short)j; pz[c] = (short)k; }
                    c++;
                }
    }
    return c;
}

/* ===================================================================== */
/* Gmsh v2 ASCII reader with the reference's tagging and corner removal
 * (hf/mesh_functions.h:172-476). */
int orc_read_gmsh(const char *path, int dim, orc_mesh *o)
{
    memset(o, 0, sizeof(*o));
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    char tok[256];
    int iv, n;
    for (;;) {                                            /* :189-199 */
        if (fscanf(f, "%255s ", tok) == EOF) { fclose(f); return -2; }
        if (strcmp(tok, "$Nodes") == 0) break;
    }
    if (fscanf(f, "%i ", &n) != 1) { fclose(f); return -2; }
    double *x = malloc(sizeof(double) * n), *y = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *nx = calloc(n, sizeof(double)), *ny = calloc(n, sizeof(double)), *nz = calloc(n, sizeof(double));
    uint8_t *bt = calloc(n, 1), *ct = calloc(n, 1);
    double xmin = 1e6, ymin = 1e6, zmin = 1e6, xmax = -1e6, ymax = -1e6, zmax = -1e6;
    for (int i = 0; i < n; i++) {                          /* :205-253 */
        if (fscanf(f, "%i ", &iv) == EOF) goto bad;
        if (fscanf(f, "%lf ", &x[i]) == EOF) goto bad;
        if (fscanf(f, "%lf ", &y[i]) == EOF) goto bad;
        if (fscanf(f, "%lf ", &z[i]) == EOF) goto bad;
        if (x[i] < xmin) xmin = x[i];
        if (y[i] < ymin) ymin = y[i];
        if (z[i] < zmin) zmin = z[i];
        if (x[i] > xmax) xmax = x[i];
        if (y[i] > ymax) ymax = y[i];
        if (z[i] > zmax) zmax = z[i];
    }
    for (;;) {                                            /* :255-265 */
        if (fscanf(f, "%255s ", tok) == EOF) goto bad;
        if (strcmp(tok, "$Elements") == 0) break;
    }
    int ne;
    if (fscanf(f, "%i ", &ne) != 1) goto bad;
    o->num_elem = ne;
    if (dim == 2)                                         /* :274-281 */
        o->d_avg = sqrt((xmax - xmin) * (ymax - ymin) / ne);
    else
        o->d_avg = cbrt((xmax - xmin) * (ymax - ymin) * (zmax - zmin) / ne);

    int et, tg, a, b, c, d;
    for (int ie = 0; ie < ne; ie++) {
        if (fscanf(f, "%i ", &iv) != 1) goto bad;
        if (fscanf(f, "%i ", &et) != 1) goto bad;
        if (dim == 2 && et == 1) {                          /* :293-306 */
            if (fscanf(f, "%i %i %i %i %i ", &iv, &iv, &tg, &a, &b) != 5) goto bad;
            bt[a - 1] = 1; bt[b - 1] = 1;
            double dx = x[b - 1] - x[a - 1], dy = y[b - 1] - y[a - 1];
            nx[a - 1] += -dy; ny[a - 1] += dx;
            nx[b - 1] += -dy; ny[b - 1] += dx;
        } else if (et == 15) {                              /* :307-314, 406-413 */
            if (fscanf(f, "%i %i %i %i ", &iv, &iv, &tg, &a) != 4) goto bad;
            ct[a - 1] = 1;
        } else if (dim == 3 && (et == 2 || et == 3)) {      /* :327-395 */
            int nn = (et == 2) ? 3 : 4;
            int nd[4];
            if (fscanf(f, "%i %i %i ", &iv, &iv, &tg) != 3) goto bad;
            for (int q = 0; q < nn; q++) {
                if (fscanf(f, "%i ", &nd[q]) != 1) goto bad;
                bt[nd[q] - 1] = 1;
            }
            a = nd[0]; b = nd[1]; c = nd[2];
            double dx = x[b - 1] - x[a - 1], dy = y[b - 1] - y[a - 1], dz = z[b - 1] - z[a - 1];
            double dx1 = x[c - 1] - x[a - 1], dy1 = y[c - 1] - y[a - 1], dz1 = z[c - 1] - z[a - 1];
            double t1 = dy * dz1 - dz * dy1;
            for (int q = 0; q < nn; q++) nx[nd[q] - 1] += t1;
            double t2 = dz * dx1 - dx * dz1;
            for (int q = 0; q < nn; q++) ny[nd[q] - 1] += t2;
            double t3 = dx * dy1 - dy * dx1;
            for (int q = 0; q < nn; q++) nz[nd[q] - 1] += t3;
        } else if (dim == 3 && et == 1) {                   /* :396-405 */
            if (fscanf(f, "%i %i %i %i %i ", &iv, &iv, &tg, &a, &b) != 5) goto bad;
            ct[a - 1] = 1; ct[b - 1] = 1;
        } else {
            (void)!fscanf(f, "%*[^\n]\n");                   /* :315-318 */
        }
        (void)d;
    }
    fclose(f);
    int nbt = 0, nc = 0;
    for (int i = 0; i < n; i++) { nbt += bt[i]; nc += ct[i]; }   /* :421-433 */
    /* remove the first num_corners nodes (:434-473) */
    int N = n - nc;
    memmove(x, x + nc, sizeof(double) * N);  memmove(y, y + nc, sizeof(double) * N);
    memmove(z, z + nc, sizeof(double) * N);  memmove(nx, nx + nc, sizeof(double) * N);
    memmove(ny, ny + nc, sizeof(double) * N); memmove(nz, nz + nc, sizeof(double) * N);
    memmove(bt, bt + nc, N); memmove(ct, ct + nc, N);
    o->N = N; o->nb = nbt - nc; o->num_corners_removed = nc;
    o->x = x; o->y = y; o->z = z; o->nx = nx; o->ny = ny; o->nz = nz; o->btag = bt; o->ctag = ct;
    return 0;
bad:
    fclose(f);
    free(x); free(y); free(z); free(nx); free(ny); free(nz); free(bt); free(ct);
    return -3;
}

void orc_free_mesh(orc_mesh *m)
{
    free(m->x); free(m->y); free(m->z); free(m->nx); free(m->ny); free(m->nz);
    free(m->btag); free(m->ctag);
    memset(m, 0, sizeof(*m));
}

/* ===================================================================== */
/* Insertion-ordered kd-tree and range query, restated from hf/kdtree.c.
 * Insert (kdtree.c:167-194): the new node hangs left of a node when its
 * coordinate on that node's axis is strictly smaller, else right; the axis of
 * a new node is (parent axis + 1) % 3, root axis 0.
 * Range query (kdtree.c:262-293, 537-559): preorder, nearer child first,
 * farther child only when |dx| < range; a node is reported when its squared
 * distance is <= range^2.  Reported nodes are prepended (kdtree.c:812-830),
 * so iteration sees them in reverse visit order. */
typedef struct {
    int n;
    double *pos;      /* 3 per node */
    int *left, *right, *data;
    unsigned char *dir;
    int root;
} kdt;

static void kdt_init(kdt *t, int cap)
{
    t->n = 0; t->root = -1;
    t->pos = malloc(sizeof(double) * 3 * (cap > 0 ? cap : 1));
    t->left = malloc(sizeof(int) * (cap > 0 ? cap : 1));
    t->right = malloc(sizeof(int) * (cap > 0 ? cap : 1));
    t->data = malloc(sizeof(int) * (cap > 0 ? cap : 1));
    t->dir = malloc(cap > 0 ? cap : 1);
}
static void kdt_free(kdt *t)
{
    free(t->pos); free(t->left); free(t->right); free(t->data); free(t->dir);
}
static void kdt_insert(kdt *t, double px, double py, double pz, int data)
{
    int id = t->n++;
    t->pos[3 * id] = px; t->pos[3 * id + 1] = py; t->pos[3 * id + 2] = pz;
    t->left[id] = t->right[id] = -1; t->data[id] = data;
    if (t->root < 0) { t->root = id; t->dir[id] = 0; return; }
    const double p[3] = {px, py, pz};
    int cur = t->root;
    for (;;) {
        int d = t->dir[cur];
        if (p[d] < t->pos[3 * cur + d]) {
            if (t->left[cur] < 0) { t->left[cur] = id; t->dir[id] = (unsigned char)((d + 1) % 3); return; }
            cur = t->left[cur];
        } else {
            if (t->right[cur] < 0) { t->right[cur] = id; t->dir[id] = (unsigned char)((d + 1) % 3); return; }
            cur = t->right[cur];
        }
    }
}

/* Returns the number of reported nodes; ids written to out[] in ITERATION
 * order (reverse of visit order).  stack/out must hold t->n entries. */
static int kdt_range(const kdt *t, const double p[3], double range, int *stack, int *out)
{
    int sp = 0, nv = 0;
    if (t->root >= 0) stack[sp++] = t->root;
    const double r2 = range * range;
    while (sp > 0) {
        int nd = stack[--sp];
        const double *q = t->pos + 3 * nd;
        double d2 = 0;
        for (int i = 0; i < 3; i++) d2 += (q[i] - p[i]) * (q[i] - p[i]);
        if (d2 <= r2) out[nv++] = nd;
        int dr = t->dir[nd];
        double dx = p[dr] - q[dr];
        int nearc = dx <= 0.0 ? t->left[nd] : t->right[nd];
        int farc = dx <= 0.0 ? t->right[nd] : t->left[nd];
        if (fabs(dx) < range && farc >= 0) stack[sp++] = farc;
        if (nearc >= 0) stack[sp++] = nearc;
    }
    for (int i = 0, j = nv - 1; i < j; i++, j--) { int tmp = out[i]; out[i] = out[j]; out[j] = tmp; }
    return nv;
}

/* find_neighbours (hf/kdtree_functions.h:61-111): stable insertion of
 * sqrt-distances into k slots (strict <), sentinel 100, radius x1.2 retry. */
static void find_neighbours(const kdt *t, const double p[3], double radius, int k,
                            int *stack, int *lst, double *dist, int *ind)
{
    for (;;) {
        for (int i = 0; i < k; i++) { dist[i] = 100.0; ind[i] = 0; }
        int nv = kdt_range(t, p, radius, stack, lst);
        for (int r = 0; r < nv; r++) {
            const double *q = t->pos + 3 * lst[r];
            double s = 0, df;
            for (int d = 0; d < 3; d++) { df = p[d] - q[d]; s += df * df; }
            double dd = sqrt(s);
            for (int j = 0; j < k; j++) {
                if (dd < dist[j]) {
                    for (int q2 = k - 1; q2 > j; q2--) { dist[q2] = dist[q2 - 1]; ind[q2] = ind[q2 - 1]; }
                    dist[j] = dd; ind[j] = t->data[lst[r]];
                    break;
                }
            }
        }
        if (dist[k - 1] == 100) { radius = 1.2 * radius; continue; }
        return;
    }
}

/* find_cloud_index (hf/kdtree_functions.h:113-151). */
int orc_find_cloud_index(int N, int nb, const double *x, const double *y, const double *z,
                         const uint8_t *btag, const uint8_t *ctag, double d_avg, int m, int *cloud)
{
    kdt t;
    kdt_init(&t, N);
    for (int i = 0; i < N; i++)
        if (!ctag[i]) kdt_insert(&t, x[i], y[i], z[i], i);
    double radius = d_avg * m;
    int *stack = malloc(sizeof(int) * (N + 1)), *lst = malloc(sizeof(int) * (N + 1));
    double *dist = malloc(sizeof(double) * m);
    int *ind = malloc(sizeof(int) * m);
    for (int i = nb; i < N; i++) {
        if (!ctag[i]) {
            double p[3] = {x[i], y[i], z[i]};
            find_neighbours(&t, p, radius, m, stack, lst, dist, ind);
            memcpy(cloud + (size_t)i * m, ind, sizeof(int) * m);
        } else {
            memset(cloud + (size_t)i * m, 0, sizeof(int) * m);
        }
    }
    kdt_free(&t);
    kdt_init(&t, N);
    for (int i = 0; i < N; i++)
        if (!btag[i]) kdt_insert(&t, x[i], y[i], z[i], i);
    for (int i = 0; i < nb; i++) {
        double p[3] = {x[i], y[i], z[i]};
        find_neighbours(&t, p, radius, m, stack, lst, dist, ind);
        cloud[(size_t)i * m] = i;
        for (int j = 1; j < m; j++) cloud[(size_t)i * m + j] = ind[j - 1];
    }
    kdt_free(&t);
    free(stack); free(lst); free(dist); free(ind);
    return 0;
}

/* find_nearest_point (hf/kdtree_functions.h:153-175). */
int orc_find_nearest_point(int N1, const double *x1, const double *y1, const double *z1,
                           const uint8_t *ctag1,
                           int N2, const double *x2, const double *y2, const double *z2,
                           const uint8_t *ctag2, double d_avg2, int *out)
{
    kdt t;
    kdt_init(&t, N2);
    for (int i = 0; i < N2; i++)
        if (!ctag2[i]) kdt_insert(&t, x2[i], y2[i], z2[i], i);
    double radius = d_avg2 * 10;
    int *stack = malloc(sizeof(int) * (N2 + 1)), *lst = malloc(sizeof(int) * (N2 + 1));
    double dist[3];
    int ind[3];
    for (int i = 0; i < N1; i++) {
        if (!ctag1[i]) {
            double p[3] = {x1[i], y1[i], z1[i]};
            find_neighbours(&t, p, radius, 3, stack, lst, dist, ind);
            out[i] = ind[0];
        } else {
            out[i] = 0;
        }
    }
    kdt_free(&t);
    free(stack); free(lst);
    return 0;
}

/* ===================================================================== */
/* BFS reordering from node 0 (hf/mesh_functions.h:774-872).  Permutes
 * coordinates, normals, boundary tags and cloud_index. */
int orc_rcm(int N, int m, double *x, double *y, double *z, double *nx, double *ny, double *nz,
            uint8_t *btag, int *cloud, int *queue1_out)
{
    int *q1 = malloc(sizeof(int) * N), *q2 = malloc(sizeof(int) * N);
    uint8_t *vis = calloc(N, 1);
    int count = 0, head = 0;
    q1[0] = 0; vis[0] = 1; count = 1;
    while (head < count) {
        int cur = q1[head++];
        for (int i = 1; i < m; i++) {
            int nb = cloud[(size_t)cur * m + i];
            if (!vis[nb]) { q1[count++] = nb; vis[nb] = 1; }
        }
    }
    if (count != N) { free(q1); free(q2); free(vis); return -1; }
    for (int i = 0; i < N; i++) q2[q1[i]] = i;
    double *tmp = malloc(sizeof(double) * N);
#define PERMD(a) do { for (int i = 0; i < N; i++) tmp[i] = a[q1[i]]; memcpy(a, tmp, sizeof(double) * N); } while (0)
    PERMD(x); PERMD(y); PERMD(z); PERMD(nx); PERMD(ny); PERMD(nz);
#undef PERMD
    uint8_t *tb = malloc(N);
    for (int i = 0; i < N; i++) tb[i] = btag[q1[i]];
    memcpy(btag, tb, N);
    int *tc = malloc(sizeof(int) * (size_t)N * m);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < m; j++) tc[(size_t)i * m + j] = q2[cloud[(size_t)q1[i] * m + j]];
    memcpy(cloud, tc, sizeof(int) * (size_t)N * m);
    if (queue1_out) memcpy(queue1_out, q1, sizeof(int) * N);
    free(q1); free(q2); free(vis); free(tmp); free(tb); free(tc);
    return 0;
}

/* correct_normal_directions (hf/mesh_functions.h:478-527). */
void orc_correct_normals(int N, int m, const double *x, const double *y, const double *z,
                         double *nx, double *ny, double *nz, const uint8_t *btag, const int *cloud)
{
    for (int i = 0; i < N; i++) {
        if (!btag[i]) continue;
        double xa = 0, ya = 0, za = 0;
        short count = 0;
        for (int j = m - 1; j > 0; j--) {
            int c = cloud[(size_t)i * m + j];
            if (!btag[c]) {
                xa += x[c]; ya += y[c]; za += z[c];
                count++;
                if (count == 1) break;
            }
        }
        xa = xa / count; ya = ya / count; za = za / count;
        double dx = xa - x[i], dy = ya - y[i], dz = za - z[i];
        double dot = dx * nx[i] + dy * ny[i] + dz * nz[i];
        if (dot > 0) { nx[i] = -nx[i]; ny[i] = -ny[i]; nz[i] = -nz[i]; }
        double mag = sqrt(nx[i] * nx[i] + ny[i] * ny[i] + nz[i] * nz[i]);
        nx[i] = nx[i] / mag; ny[i] = ny[i] / mag; nz[i] = nz[i] / mag;
    }
}

/* calculate_avg_dx (hf/mesh_functions.h:895-910). */
double orc_avg_dx(int N, int m, int dim, const double *x, const double *y, const double *z,
                  const uint8_t *btag, const int *cloud)
{
    double s = 0;
    for (int i = 0; i < N; i++) {
        if (btag[i]) continue;
        int c = cloud[(size_t)i * m + 1];
        double dx = x[i] - x[c], dy = y[i] - y[c];
        double dz = (dim == 3) ? z[i] - z[c] : 0;
        s += sqrt(dx * dx + dy * dy + dz * dz);
    }
    return s / N;
}

/* ===================================================================== */
/* calculate_phs_rbf (hf/rbf.h:32-40). */
static double phs_rbf(const double *p1, const double *p2, int phs, int dim)
{
    double sum = 0;
    for (int i = 0; i < dim; i++) sum += pow(p1[i] - p2[i], 2);
    if (sum == 0) return 0;
    return pow(sum, phs * 0.5);
}

/* create_A_matrix_from_cloud_indices (hf/rbf.h:42-77) into a dense mpn x mpn
 * row-major buffer whose zero block is already zero. */
static void build_A(int m, int n, int dim, int phs, const short *px, const short *py, const short *pz,
                    const double *x, const double *y, const double *z, const int *cl, double *A)
{
    int mpn = m + n;
    memset(A, 0, sizeof(double) * mpn * mpn);
    double seed[3] = {x[cl[0]], y[cl[0]], z[cl[0]]};
    double p1[3], p2[3];
    for (int i = 0; i < m; i++) {
        A[i * mpn + i] = 0;
        p1[0] = x[cl[i]]; p1[1] = y[cl[i]]; p1[2] = z[cl[i]];
        for (int j = i + 1; j < m; j++) {
            p2[0] = x[cl[j]]; p2[1] = y[cl[j]]; p2[2] = z[cl[j]];
            A[i * mpn + j] = phs_rbf(p1, p2, phs, dim);
            A[j * mpn + i] = A[i * mpn + j];
        }
    }
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            double v = 1;
            if (px[j] != 0) v *= pow(x[cl[i]] - seed[0], px[j]);
            if (py[j] != 0) v *= pow(y[cl[i]] - seed[1], py[j]);
            if (pz[j] != 0) v *= pow(z[cl[i]] - seed[2], pz[j]);
            A[i * mpn + j + m] = v;
            A[(j + m) * mpn + i] = v;
        }
}

/* matrixInverse_Gauss_Jordan (hf/mat_lib.h:423-491): one bottom-up bubble
 * pass on column 0, then Gauss-Jordan without pivoting, then row scaling.
 * work: order x 2*order doubles; rows are swapped by pointer. */
static void gj_inverse(const double *A, int order, double *inv, double *work, double **rows)
{
    for (int i = 0; i < order; i++) {
        rows[i] = work + (size_t)i * 2 * order;
        for (int j = 0; j < order; j++) rows[i][j] = A[i * order + j];
        for (int j = order; j < 2 * order; j++) rows[i][j] = (j == i + order) ? 1 : 0;
    }
    for (int i = order - 1; i > 0; i--)
        if (rows[i - 1][0] < rows[i][0]) { double *t = rows[i]; rows[i] = rows[i - 1]; rows[i - 1] = t; }
    for (int i = 0; i < order; i++)
        for (int j = 0; j < order; j++)
            if (j != i) {
                double t = rows[j][i] / rows[i][i];
                for (int k = 0; k < 2 * order; k++) rows[j][k] -= rows[i][k] * t;
            }
    for (int i = 0; i < order; i++) {
        double t = rows[i][i];
        for (int j = 0; j < 2 * order; j++) rows[i][j] = rows[i][j] / t;
    }
    for (int i = 0; i < order; i++)
        for (int j = 0; j < order; j++) inv[i * order + j] = rows[i][j + order];
}

/* Textbook LU with partial pivoting; returns A^{-1} (cross-check only). */
static int lu_inverse(const double *A, int n, double *inv, double *LU, int *piv)
{
    memcpy(LU, A, sizeof(double) * n * n);
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(LU[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(LU[i * n + k]) > best) { best = fabs(LU[i * n + k]); p = i; }
        piv[k] = p;
        if (best == 0) return -1;
        if (p != k)
            for (int j = 0; j < n; j++) { double t = LU[k * n + j]; LU[k * n + j] = LU[p * n + j]; LU[p * n + j] = t; }
        for (int i = k + 1; i < n; i++) {
            double l = LU[i * n + k] / LU[k * n + k];
            LU[i * n + k] = l;
            for (int j = k + 1; j < n; j++) LU[i * n + j] -= l * LU[k * n + j];
        }
    }
    for (int c = 0; c < n; c++) {
        double *col = inv + 0;  /* solve into column c, stored row-major */
        double b[1024];
        for (int i = 0; i < n; i++) b[i] = (i == c) ? 1 : 0;
        for (int k = 0; k < n; k++) if (piv[k] != k) { double t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
        for (int i = 1; i < n; i++) { double s = b[i]; for (int j = 0; j < i; j++) s -= LU[i * n + j] * b[j]; b[i] = s; }
        for (int i = n - 1; i >= 0; i--) { double s = b[i]; for (int j = i + 1; j < n; j++) s -= LU[i * n + j] * b[j]; b[i] = s / LU[i * n + i]; }
        for (int i = 0; i < n; i++) col[i * n + c] = b[i];
    }
    return 0;
}

/* Row 0 of the RHS matrices (hf/rbf.h:79-270); only row 0 is ever used
 * (create_full_* keep B1[0][j]).  op: 0 Dx, 1 Dy, 2 Dz, 3 Laplacian. */
static void rhs_row0(int op, int m, int n, int dim, int phs, const short *px, const short *py,
                     const short *pz, const double *x, const double *y, const double *z,
                     const int *cl, double *L)
{
    double p1[3] = {x[cl[0]], y[cl[0]], z[cl[0]]}, p2[3];
    const double *seed = p1;
    for (int j = 0; j < m; j++) {
        p2[0] = x[cl[j]]; p2[1] = y[cl[j]]; p2[2] = z[cl[j]];
        if (op < 3)
            L[j] = phs * (p1[op] - p2[op]) * phs_rbf(p1, p2, phs - 2, dim);
        else
            L[j] = phs * phs * phs_rbf(p1, p2, phs - 2, dim);
    }
    double d0 = p1[0] - seed[0], d1 = p1[1] - seed[1], d2 = p1[2] - seed[2];
    if (op == 0) {
        L[m] = 0;
        for (int j = 1; j < n; j++) {
            if (px[j] == 0) L[j + m] = 0;
            else { double g = px[j] * pow(d0, px[j] - 1); g *= pow(d1, py[j]); g *= pow(d2, pz[j]); L[j + m] = g; }
        }
    } else if (op == 1) {
        L[m] = 0;   /* never written by grady_matrix; create_matrix zero-fills */
        for (int j = 1; j < n; j++) {
            if (py[j] == 0) L[j + m] = 0;
            else { double g = py[j] * pow(d1, py[j] - 1); g = g * pow(d0, px[j]); g = g * pow(d2, pz[j]); L[j + m] = g; }
        }
    } else if (op == 2) {
        L[m] = 0;
        for (int j = 1; j < n; j++) {
            if (pz[j] == 0) L[j + m] = 0;
            else { double g = pz[j] * pow(d2, pz[j] - 1); g = g * pow(d1, py[j]); g = g * pow(d0, px[j]); L[j + m] = g; }
        }
    } else {
        for (int j = 0; j < n; j++) {
            double tx = 0, ty = 0, tz = 0;
            if (px[j] > 1) { tx = px[j] * (px[j] - 1) * pow(d0, px[j] - 2); tx *= pow(d1, py[j]); if (dim == 3) tx *= pow(d2, pz[j]); }
            else if (py[j] > 1) { ty = py[j] * (py[j] - 1) * pow(d1, py[j] - 2); ty *= pow(d0, px[j]); if (dim == 3) ty *= pow(d2, pz[j]); }
            else if (dim == 3 && pz[j] > 1) { tz = pz[j] * (pz[j] - 1) * pow(d2, pz[j] - 2); tz *= pow(d0, px[j]); tz *= pow(d1, py[j]); }
            L[j + m] = (dim == 3) ? tx + ty + tz : tx + ty;
        }
    }
}

/* create_derivative_matrices (hf/rbf.h:401-421).  The reference re-inverts
 * the same A once per operator; the inverse is deterministic, so one inverse
 * per node gives bitwise the same weights. */
int orc_weights(int N, int m, int n, int dim, int phs, const short *px, const short *py,
                const short *pz, const double *x, const double *y, const double *z,
                const uint8_t *ctag, const int *cloud, int method,
                double *Dx, double *Dy, double *Dz, double *lap)
{
    int mpn = m + n, err = 0;
    if (mpn > 1024) return -1;
#pragma omp parallel reduction(|:err)
    {
        double *A = malloc(sizeof(double) * mpn * mpn), *inv = malloc(sizeof(double) * mpn * mpn);
        double *work = malloc(sizeof(double) * mpn * 2 * mpn), **rows = malloc(sizeof(double *) * mpn);
        int *piv = malloc(sizeof(int) * mpn);
        double *L = malloc(sizeof(double) * mpn);
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < N; i++) {
            double *outs[4] = {Dx, Dy, Dz, lap};
            if (ctag && ctag[i]) continue;
            const int *cl = cloud + (size_t)i * m;
            build_A(m, n, dim, phs, px, py, pz, x, y, z, cl, A);
            if (method == 0) gj_inverse(A, mpn, inv, work, rows);
            else if (lu_inverse(A, mpn, inv, work, piv)) err |= 1;
            for (int op = 0; op < 4; op++) {
                if (!outs[op]) continue;
                if (op == 2 && dim != 3) continue;
                rhs_row0(op, m, n, dim, phs, px, py, pz, x, y, z, cl, L);
                for (int j = 0; j < m; j++) {          /* multiply_matrices row 0 (hf/mat_lib.h:214-228) */
                    double c = 0;
                    for (int k = 0; k < mpn; k++) c += L[k] * inv[k * mpn + j];
                    outs[op][(size_t)i * m + j] = c;
                }
            }
        }
        free(A); free(inv); free(work); free(rows); free(piv); free(L);
    }
    return err ? -2 : 0;
}

/* create_laplacian_for_Poisson (hf/rbf.h:382-399). */
void orc_lap_poisson(int N, int m, int dim, const double *Dx, const double *Dy, const double *Dz,
                     const double *lap, const double *nx, const double *ny, const double *nz,
                     const uint8_t *btag, const uint8_t *ctag, double *lapP)
{
    for (int i = 0; i < N; i++) {
        if (!(ctag && ctag[i]))
            for (int j = 0; j < m; j++) lapP[(size_t)i * m + j] = lap[(size_t)i * m + j];
        if (btag[i])
            for (int j = 0; j < m; j++) {
                size_t q = (size_t)i * m + j;
                lapP[q] = Dx[q] * nx[i] + Dy[q] * ny[i];
                if (dim == 3) lapP[q] += Dz[q] * nz[i];
            }
    }
}

/* ===================================================================== */
/* One interpolation row: [phi(x_t - x_j); p(x_t - seed)]^T A^{-1}, first m
 * entries (hf/mesh_functions.h:567-594, hf/mat_lib.h:720-726). */
static void interp_row(int m, int n, int dim, int phs, const short *px, const short *py,
                       const short *pz, const double *xs, const double *ys, const double *zs,
                       const int *cl, const double pt[3], const double *inv, double *tmp, double *row)
{
    int mpn = m + n;
    double p2[3];
    for (int j = 0; j < m; j++) {
        p2[0] = xs[cl[j]]; p2[1] = ys[cl[j]]; p2[2] = zs[cl[j]];
        tmp[j] = phs_rbf(pt, p2, phs, dim);
    }
    double seed[3] = {xs[cl[0]], ys[cl[0]], zs[cl[0]]};
    for (int j = 0; j < n; j++)
        tmp[j + m] = pow(pt[0] - seed[0], px[j]) * pow(pt[1] - seed[1], py[j]) * pow(pt[2] - seed[2], pz[j]);
    for (int i = 0; i < m; i++) {
        double c = 0;
        for (int j = 0; j < mpn; j++) c += tmp[j] * inv[j * mpn + i];
        row[i] = c;
    }
}

static int transfer_rows(int Nt, int nbt, const double *xt, const double *yt, const double *zt,
                         const uint8_t *btag_t, const int *pts, int Ns, int ms, int ns, int dim,
                         int phs, const short *px, const short *py, const short *pz,
                         const double *xs, const double *ys, const double *zs, const int *cloud_s,
                         int row_mode, int method, double *out)
{
    int mpn = ms + ns, err = 0;
    (void)Ns;
    memset(out, 0, sizeof(double) * (size_t)Nt * ms);
#pragma omp parallel reduction(|:err)
    {
        double *A = malloc(sizeof(double) * mpn * mpn), *inv = malloc(sizeof(double) * mpn * mpn);
        double *work = malloc(sizeof(double) * mpn * 2 * mpn), **rows = malloc(sizeof(double *) * mpn);
        int *piv = malloc(sizeof(int) * mpn);
        double *tmp = malloc(sizeof(double) * mpn);
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < Nt; i++) {
            if (row_mode == 0 ? (i < nbt) : (btag_t[i] != 0)) continue;
            const int *cl = cloud_s + (size_t)pts[i] * ms;
            build_A(ms, ns, dim, phs, px, py, pz, xs, ys, zs, cl, A);
            if (method == 0) gj_inverse(A, mpn, inv, work, rows);
            else if (lu_inverse(A, mpn, inv, work, piv)) err |= 1;
            double pt[3] = {xt[i], yt[i], zt[i]};
            interp_row(ms, ns, dim, phs, px, py, pz, xs, ys, zs, cl, pt, inv, tmp, out + (size_t)i * ms);
        }
        free(A); free(inv); free(work); free(rows); free(piv); free(tmp);
    }
    return err ? -2 : 0;
}

/* create_restriction_matrix (hf/mesh_functions.h:533-600): rows live on the
 * coarse level, stencils are fine clouds around restriction_points. */
int orc_restriction(int Nf, int mf, int nf, int dim, int phs, const short *px, const short *py,
                    const short *pz, const double *xf, const double *yf, const double *zf,
                    const int *cloud_f, int Nc, int nbc, const double *xc, const double *yc,
                    const double *zc, const uint8_t *btag_c, const int *restr_pts, int row_mode,
                    int method, double *R)
{
    return transfer_rows(Nc, nbc, xc, yc, zc, btag_c, restr_pts, Nf, mf, nf, dim, phs, px, py, pz,
                         xf, yf, zf, cloud_f, row_mode, method, R);
}

/* create_prolongation_matrix (hf/mesh_functions.h:602-670): rows live on the
 * fine level, stencils are coarse clouds around prolongation_points. */
int orc_prolongation(int Nf, int nbf, const double *xf, const double *yf, const double *zf,
                     const uint8_t *btag_f, const int *prol_pts, int Nc, int mc, int nc, int dim,
                     int phs, const short *px, const short *py, const short *pz,
                     const double *xc, const double *yc, const double *zc, const int *cloud_c,
                     int row_mode, int method, double *P)
{
    return transfer_rows(Nf, nbf, xf, yf, zf, btag_f, prol_pts, Nc, mc, nc, dim, phs, px, py, pz,
                         xc, yc, zc, cloud_c, row_mode, method, P);
}

/* ===================================================================== */
/* V-cycle primitives.  `order` gives the sweep order; consecutive entries of
 * one colour class are independent, so rows of one class may be processed in
 * parallel with identical results. */
static inline void sor_row(orc_level *L, int i, double w, double c)
{
    const int m = L->m;
    const double *Wr = L->Lp + (size_t)i * m;
    const int *cr = L->cloud + (size_t)i * m;
    double sum = 0.0;
    for (int j = 1; j < m; j++) sum += Wr[j] * L->p[cr[j]];
    L->p[i] = w * ((L->s[i] - sum) / Wr[0]) + c * L->p[i];
}

/* One sweep in `order`; with colour classes given, the rows of a class (no two
 * of them neighbours) are updated in parallel -- the same values as in order. */
static void sweep(orc_level *L, float omega)
{
    const int N = L->N;
    const double w = omega, c = (float)(1 - omega);      /* omega is float; (1-omega) is float */
    if (L->order && L->cls) {
        for (int k = 0; k < L->ncls; k++) {
            const int q0 = L->cls[k], q1 = L->cls[k + 1];
#pragma omp parallel for schedule(static)
            for (int q = q0; q < q1; q++) sor_row(L, L->order[q], w, c);
        }
        return;
    }
    for (int q = 0; q < N; q++) sor_row(L, L->order ? L->order[q] : q, w, c);
}

static void pin(orc_level *L)
{
    double pref = L->p[0];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < L->N; i++) L->p[i] = L->p[i] - pref;
}

void orc_set_num_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

/* relaxation: H hf/heat_conduction.h:25-56; T hf/timple_solver.h:235-272;
 * F hf/fractional_step_solver.h:168-196. */
void orc_relax(int variant, orc_level *L, float omega, int num_relax)
{
    const int N = L->N, m = L->m;
    for (int it = 0; it < num_relax; it++) {
        sweep(L, omega);
        if (variant == ORC_VAR_H) continue;
        pin(L);
        const double *W = (variant == ORC_VAR_T) ? L->Lp : L->lap;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < N; i++) {
            if (variant == ORC_VAR_T && L->btag[i]) { L->res[i] = 0; continue; }
            double r = L->s[i];
            for (int j = 0; j < m; j++) r = r - W[(size_t)i * m + j] * L->p[L->cloud[(size_t)i * m + j]];
            L->res[i] = r;
        }
    }
}

/* calculate_residuals: H hf/heat_conduction.h:101-112; T hf/timple_solver.h:274-290;
 * F hf/fractional_step_solver.h:236-246 (uses lap, leaves boundary res). */
void orc_residual(int variant, orc_level *L)
{
    const int N = L->N, m = L->m;
    const double *W = (variant == ORC_VAR_F) ? L->lap : L->Lp;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        if (variant != ORC_VAR_F) L->res[i] = 0.0;
        if (!L->btag[i]) {
            double sum = 0;
            for (int j = 0; j < m; j++) sum += W[(size_t)i * m + j] * L->p[L->cloud[(size_t)i * m + j]];
            L->res[i] = L->s[i] - sum;
        }
    }
}

/* restrict_residuals: H hf/heat_conduction.h:58-77; T hf/timple_solver.h:292-306;
 * F hf/fractional_step_solver.h:198-212 (boundary source left untouched). */
void orc_restrict(int variant, orc_level *f, orc_level *c)
{
    const int mf = f->m;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < c->N; i++) {
        double r = 0.0;
        if (!c->btag[i]) {
            int ir = c->rp[i];
            for (int j = 0; j < mf; j++) r += c->R[(size_t)i * mf + j] * f->res[f->cloud[(size_t)ir * mf + j]];
        } else if (variant == ORC_VAR_F) {
            continue;
        }
        c->s[i] = r;
    }
}

/* prolongate_corrections: H hf/heat_conduction.h:79-99; T hf/timple_solver.h:308-328
 * (zeroes every coarse node); F hf/fractional_step_solver.h:214-234. */
void orc_prolong(int variant, orc_level *f, orc_level *c)
{
    const int mc = c->m;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->N; i++) {
        if (f->btag[i]) continue;
        int ip = f->pp[i];
        double r = 0.0;
        for (int j = 0; j < mc; j++) r += f->P[(size_t)i * mc + j] * c->p[c->cloud[(size_t)ip * mc + j]];
        f->p[i] = f->p[i] + r;
    }
    for (int i = 0; i < c->N; i++)
        if (variant == ORC_VAR_T || !c->btag[i]) c->p[i] = 0.0;
}

/* One weighted-Jacobi sweep (no reference counterpart): sor_row's update of
 * every row with every neighbour from the previous iterate, staged in res. */
static void jacobi_sweep(orc_level *L, float omega)
{
    const int N = L->N, m = L->m;
    const double w = omega, c = (float)(1 - omega);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        const double *Wr = L->Lp + (size_t)i * m;
        const int *cr = L->cloud + (size_t)i * m;
        double sum = 0.0;
        for (int j = 1; j < m; j++) sum += Wr[j] * L->p[cr[j]];
        L->res[i] = w * ((L->s[i] - sum) / Wr[0]) + c * L->p[i];
    }
    memcpy(L->p, L->res, sizeof(double) * (size_t)N);
}

/* n weighted-Jacobi sweeps of a level (no reference counterpart).  The GPU
 * stages the sweeps in the level's residual buffer and one more buffer, so
 * that the last lands in p without a copy; res is left holding the output of
 * the last sweep staged there -- sweep n - 2 of an odd count n >= 3, sweep
 * n - 1 of an even count, the only sweep of n = 1 -- and so is it here. */
static void jacobi_relax(orc_level *L, float omega, int n)
{
    const int keep = n == 1 ? 1 : (n % 2 == 0 ? n - 1 : n - 2);
    const size_t bytes = sizeof(double) * (size_t)L->N;
    double *kept = NULL;
    for (int it = 1; it <= n; it++) {
        jacobi_sweep(L, omega);
        if (it == keep && it != n) {
            kept = (double *)malloc(bytes);
            memcpy(kept, L->res, bytes);
        }
    }
    if (kept) {
        memcpy(L->res, kept, bytes);
        free(kept);
    }
}

/* One V-cycle of the H variant with a smoother kind (0 SOR, 1 Jacobi) and a
 * factor per level (no reference counterpart; with SOR everywhere this is
 * orc_vcycle_levels). */
void orc_vcycle_smooth(int nlev, orc_level *lv, const float *omega, const int *kind, int num_relax)
{
    for (int l = 0; l < nlev; l++) {
        if (kind[l]) jacobi_relax(&lv[l], omega[l], num_relax);
        else orc_relax(ORC_VAR_H, &lv[l], omega[l], num_relax);
        orc_residual(ORC_VAR_H, &lv[l]);
        if (l != nlev - 1) orc_restrict(ORC_VAR_H, &lv[l], &lv[l + 1]);
    }
    for (int l = nlev - 1; l > 0; l--) {
        orc_prolong(ORC_VAR_H, &lv[l - 1], &lv[l]);
        if (l != 1) {
            if (kind[l - 1]) jacobi_relax(&lv[l - 1], omega[l - 1], num_relax);
            else orc_relax(ORC_VAR_H, &lv[l - 1], omega[l - 1], num_relax);
        }
    }
}

/* One V-cycle with a SOR factor per level (no reference counterpart: the
 * reference's cycle below uses one factor; with equal factors this is it). */
void orc_vcycle_levels(int variant, int nlev, orc_level *lv, const float *omega, int num_relax)
{
    for (int l = 0; l < nlev; l++) {
        orc_relax(variant, &lv[l], omega[l], num_relax);
        orc_residual(variant, &lv[l]);
        if (l != nlev - 1) orc_restrict(variant, &lv[l], &lv[l + 1]);
    }
    for (int l = nlev - 1; l > 0; l--) {
        orc_prolong(variant, &lv[l - 1], &lv[l]);
        if (l != 1) orc_relax(variant, &lv[l - 1], omega[l - 1], num_relax);
    }
}

/* One V-cycle: mg_heat_conduction.c:64-75, hf/timple_solver.h:195-207,
 * hf/fractional_step_solver.h:152-163 (no post-smoothing on level 0). */
void orc_vcycle(int variant, int nlev, orc_level *lv, float omega, int num_relax)
{
    for (int l = 0; l < nlev; l++) {
        orc_relax(variant, &lv[l], omega, num_relax);
        orc_residual(variant, &lv[l]);
        if (l != nlev - 1) orc_restrict(variant, &lv[l], &lv[l + 1]);
    }
    for (int l = nlev - 1; l > 0; l--) {
        orc_prolong(variant, &lv[l - 1], &lv[l]);
        if (l != 1) orc_relax(variant, &lv[l - 1], omega, num_relax);
    }
}

/* l2_norm_gen (hf/general_functions.h:181-192). */
double orc_l2_norm_gen(int N, const uint8_t *btag, const double *a, const double *b)
{
    double r = 0.0;
    for (int i = 0; i < N; i++)
        if (btag[i] == 0) { double d = a[i] - (b ? b[i] : 0.0); r += d * d; }
    return sqrt(r / N);
}

/* multiply_sparse_matrix_vector (hf/mat_lib.h:70-79). */
void orc_spmv(int N, int m, const double *W, const int *cloud, const double *f, double *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double r = 0.0;
        for (int j = 0; j < m; j++) r += W[(size_t)i * m + j] * f[cloud[(size_t)i * m + j]];
        out[i] = r;
    }
}

/* ===================================================================== */
/* Greedy colouring in index order on the symmetrised stencil graph: node i
 * takes the smallest colour not held by any j with j in cloud(i) or
 * i in cloud(j).  No reference counterpart (the reference sweeps in index
 * order); this restates the library's colouring so its output is pinned. */
int orc_greedy_colour(int N, int m, const int *cloud, int *colour)
{
    int *deg = calloc(N + 1, sizeof(int));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < m; j++) { int c = cloud[(size_t)i * m + j]; if (c != i) deg[c + 1]++; }
    for (int i = 0; i < N; i++) deg[i + 1] += deg[i];
    int *radj = malloc(sizeof(int) * (deg[N] + 1)), *fill = calloc(N, sizeof(int));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < m; j++) { int c = cloud[(size_t)i * m + j]; if (c != i) radj[deg[c] + fill[c]++] = i; }
    int maxc = 0;
    int *mark = calloc(4096, sizeof(int));
    for (int i = 0; i < N; i++) colour[i] = -1;
    for (int i = 0; i < N; i++) {
        int stamp = i + 1;
        for (int j = 0; j < m; j++) { int c = cloud[(size_t)i * m + j]; if (c != i && colour[c] >= 0) mark[colour[c]] = stamp; }
        for (int k = deg[i]; k < deg[i + 1]; k++) { int c = radj[k]; if (colour[c] >= 0) mark[colour[c]] = stamp; }
        int col = 0;
        while (col < 4095 && mark[col] == stamp) col++;
        colour[i] = col;
        if (col + 1 > maxc) maxc = col + 1;
    }
    free(deg); free(radj); free(fill); free(mark);
    return maxc;
}
