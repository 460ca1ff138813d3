// ns.hpp -- time-step kernels (TIMPLE and explicit fractional step).
#pragma once

#include "device.hpp"

namespace mgm {

// Device view of level 0 for the momentum / boundary-pressure kernels
// (device node order, sliced-ELL slabs of slice height S).
struct NsDev {
    int m, dim, S;
    const int *col;
    const double *Dx, *Dy, *Dz, *lap;
    const uint8_t *btag;
    const double *nx, *ny, *nz;
    const double *u, *v, *w, *uo, *vo, *wo, *dpdx, *dpdy, *dpdz;
    double *un, *vn, *wn, *pold, *dpdn;
    double rho, mu, dt, facRe, facdt, field_rho;
};

int launch_sum_partials(int n, const double *partial, double *out, hipStream_t s);
int launch_copy(int N, const double *a, double *b, hipStream_t s);
int launch_fs_predict(int N, const double *c, const double *u, const double *v, const double *w,
                      const double *dx, const double *dy, const double *dz, const double *dl,
                      double dt, double nu, double *out, hipStream_t s);
int launch_mass_source(int N, const uint8_t *btag, const double *dx, const double *dy,
                       const double *dz, double rho, double dt, int zero_boundary, double *src,
                       hipStream_t s);
int launch_fs_boundary_dpdn(int N, int dim, const uint8_t *btag, const double *u, const double *v,
                            const double *w, const double *un, const double *vn, const double *wn,
                            const double *nx, const double *ny, const double *nz, double rho,
                            double dt, double *dpdn, hipStream_t s);
int launch_fs_correct(int N, int dim, const uint8_t *btag, const double *un, const double *vn,
                      const double *wn, const double *dx, const double *dy, const double *dz,
                      double dt, double rho, double *u, double *v, double *w, hipStream_t s);
int launch_timple_correct(int N, int dim, const uint8_t *btag, const double *dx, const double *dy,
                          const double *dz, const double *pp, double dt, double rho, double *un,
                          double *vn, double *wn, double *pold, hipStream_t s);
int launch_vel_change(int N, int dim, const double *a0, const double *a1, const double *b0,
                      const double *b1, const double *c0, const double *c1, double *partial,
                      double *out, hipStream_t s);
int launch_momentum_slices(const NsDev &d, const int *slices, int nslices, hipStream_t s);
int launch_momentum_natural(const NsDev &d, int N, const int *iperm, hipStream_t s);
int launch_timple_boundary_pressure(const NsDev &d, int N, hipStream_t s);

}  // namespace mgm
