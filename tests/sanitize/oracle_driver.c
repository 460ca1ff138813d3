/* Sanitizer driver for the oracle (oracle/mgm_oracle.c): reader, stencils,
 * reordering, normals, spacing, Gauss-Jordan and LU weights, lap_Poison,
 * transfer rows, colouring and H / T / F V-cycles on two Gmsh levels.  Built
 * with -fsanitize=address,undefined by tests/sanitize/Makefile.
 *   oracle_driver <fine.msh> <coarse.msh> <dim> */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgm_oracle.h"

typedef struct {
    orc_mesh g;
    int m, n;
    short px[64], py[64], pz[64];
    int *cloud, *q1, *col, *rp, *pp, *order;
    double *Dx, *Dy, *Dz, *lap, *lapP, *R, *P, *p, *s, *res;
    double d_avg;
} lev;

static int setup(lev *L, const char *path, int dim)
{
    if (orc_read_gmsh(path, dim, &L->g)) return -1;
    L->n = orc_poly_terms(dim, 2, L->px, L->py, L->pz);
    L->m = 2 * L->n;
    int N = L->g.N, m = L->m;
    L->cloud = calloc((size_t)N * m, sizeof(int));
    L->q1 = calloc(N, sizeof(int));
    if (orc_find_cloud_index(N, L->g.nb, L->g.x, L->g.y, L->g.z, L->g.btag, L->g.ctag, L->g.d_avg, m, L->cloud))
        return -2;
    if (orc_rcm(N, m, L->g.x, L->g.y, L->g.z, L->g.nx, L->g.ny, L->g.nz, L->g.btag, L->cloud, L->q1)) return -3;
    orc_correct_normals(N, m, L->g.x, L->g.y, L->g.z, L->g.nx, L->g.ny, L->g.nz, L->g.btag, L->cloud);
    L->d_avg = orc_avg_dx(N, m, dim, L->g.x, L->g.y, L->g.z, L->g.btag, L->cloud);
    double **outs[5] = {&L->Dx, &L->Dy, &L->Dz, &L->lap, &L->lapP};
    for (int k = 0; k < 5; k++) *outs[k] = calloc((size_t)N * m, sizeof(double));
    for (int method = 0; method <= 1; method++)
        if (orc_weights(N, m, L->n, dim, 3, L->px, L->py, L->pz, L->g.x, L->g.y, L->g.z, L->g.ctag, L->cloud, method,
                        L->Dx, L->Dy, dim == 3 ? L->Dz : NULL, L->lap))
            return -4;
    orc_lap_poisson(N, m, dim, L->Dx, L->Dy, dim == 3 ? L->Dz : NULL, L->lap, L->g.nx, L->g.ny, L->g.nz, L->g.btag,
                    L->g.ctag, L->lapP);
    L->col = calloc(N, sizeof(int));
    int nc = orc_greedy_colour(N, m, L->cloud, L->col);
    L->order = calloc(N, sizeof(int));
    int k = 0;
    for (int c = 0; c < nc; c++)
        for (int i = 0; i < N; i++)
            if (L->col[i] == c) L->order[k++] = i;
    L->p = calloc(N, sizeof(double));
    L->s = calloc(N, sizeof(double));
    L->res = calloc(N, sizeof(double));
    return k == N ? 0 : -5;
}

int main(int argc, char **argv)
{
    if (argc < 4) return 2;
    const int dim = atoi(argv[3]);
    lev F, C;
    memset(&F, 0, sizeof(F));
    memset(&C, 0, sizeof(C));
    int rc = setup(&F, argv[1], dim);
    if (!rc) rc = setup(&C, argv[2], dim);
    if (rc) {
        fprintf(stderr, "oracle_driver: setup failed (%d)\n", rc);
        return 1;
    }
    F.pp = calloc(F.g.N, sizeof(int));
    C.rp = calloc(C.g.N, sizeof(int));
    orc_find_nearest_point(F.g.N, F.g.x, F.g.y, F.g.z, F.g.ctag, C.g.N, C.g.x, C.g.y, C.g.z, C.g.ctag, C.d_avg, F.pp);
    orc_find_nearest_point(C.g.N, C.g.x, C.g.y, C.g.z, C.g.ctag, F.g.N, F.g.x, F.g.y, F.g.z, F.g.ctag, F.d_avg, C.rp);
    C.R = calloc((size_t)C.g.N * F.m, sizeof(double));
    F.P = calloc((size_t)F.g.N * C.m, sizeof(double));
    for (int mode = 0; mode <= 1; mode++) {
        if (orc_restriction(F.g.N, F.m, F.n, dim, 3, F.px, F.py, F.pz, F.g.x, F.g.y, F.g.z, F.cloud, C.g.N, C.g.nb,
                            C.g.x, C.g.y, C.g.z, C.g.btag, C.rp, mode, 1, C.R) ||
            orc_prolongation(F.g.N, F.g.nb, F.g.x, F.g.y, F.g.z, F.g.btag, F.pp, C.g.N, C.m, C.n, dim, 3, C.px, C.py,
                             C.pz, C.g.x, C.g.y, C.g.z, C.cloud, mode, 1, F.P)) {
            fprintf(stderr, "oracle_driver: transfer rows failed\n");
            return 1;
        }
    }
    orc_level lv[2];
    lev *src[2] = {&F, &C};
    for (int k = 0; k < 2; k++) {
        lev *L = src[k];
        orc_level o = {L->g.N, L->m, L->cloud, L->lapP, L->lap, L->g.btag, L->order, L->p, L->s, L->res,
                       L->rp, L->R, L->pp, L->P, NULL, 0};
        lv[k] = o;
    }
    for (int i = 0; i < F.g.N; i++) F.s[i] = (i % 7) - 3.0;
    double norm = 0;
    for (int var = 0; var <= 2; var++) {
        memset(F.p, 0, sizeof(double) * F.g.N);
        orc_vcycle(var, 2, lv, 1.0f, 2);
        norm += orc_l2_norm_gen(F.g.N, F.g.btag, F.p, NULL);
    }
    printf("oracle_driver: ok (%d + %d nodes, norm %.6e)\n", F.g.N, C.g.N, norm);
    return 0;
}
