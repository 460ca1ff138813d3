// mlp.hpp -- launchers of the learned smoother's network kernels (mlp.hip).
#pragma once

#include "device.hpp"

namespace mgm {

constexpr int MLP_H = 64;       // hidden width
constexpr int MLP_MAX_M = 72;   // widest stencil the feature kernels take

int launch_sum_partials(int n, const double *partial, double *out, hipStream_t s);

size_t mlp_weight_count(int in);
// Training (FP32): forward storing hidden activations; loss gradient for
// L = mean (y - t)^2 into G (layout of the weight vector); sum of squared errors.
int mlp_forward(int n, int in, const float *X, const float *W, float *H1, float *H2, float *y, hipStream_t s);
// part: workspace of mlp_gradient_workspace(in) floats (partial sums over sample chunks).
size_t mlp_gradient_workspace(int in);
int mlp_gradient(int n, int in, const float *X, const float *t, const float *W, const float *H1, const float *H2,
                 const float *y, float *dy, float *dZ2, float *dZ1, float *part, float *G, hipStream_t s);
int mlp_sqerr(int n, const float *y, const float *t, double *partial, double *out, hipStream_t s);
int mlp_adam(int nw, float *W, const float *G, float *M, float *V, float lr_t, float b1, float b2, float eps,
             hipStream_t s);
int mlp_to_bf16(int n, const float *in, bf16 *out, hipStream_t s);
int mlp_diff(int n, const double *a, const double *b, double *out, hipStream_t s);   // out = a - b
int mlp_gather_batch(int n, int in, const int *idx, const float *X, const float *t, float *Xb, float *tb,
                     hipStream_t s);
// Feature rows and targets of sampled device rows from residual / error snapshots.
int mlp_features(int n, Ell A, int dim, double h, const double *dx, const double *dy, const double *dz,
                 const int *rows, const int *snap, int Nd, const double *res_snaps, const double *err_snaps,
                 float *X, float *t, hipStream_t s);
// Geometric part of layer 1 per interior row: G holds mlp_geometry_size(Nd)
// floats, 16-row tiles with the hidden units major inside a tile.
size_t mlp_geometry_size(int Nd);
int mlp_geometry(int Nd, Ell A, int dim, double h, const double *dx, const double *dy, const double *dz,
                 const uint8_t *btag, const float *W, float *G, hipStream_t s);
// Raises the LDS limit of the window form of the correction (once per device).
int mlp_device_init();
// The inference image of the network (mlp_net_bytes bytes): the BF16 weight
// fragments in the matrix cores' operand layout, b2 and BF16(w3), built from
// W1rT (the residual columns of W1 transposed, m x MLP_H), W2, b2 and w3.
size_t mlp_net_bytes();
int mlp_pack_net(int m, const bf16 *W1rT, const bf16 *W2, const float *b2, const bf16 *w3, void *out, hipStream_t s);
// p += network correction on interior rows, from the residual res and the
// packed network net; A supplies columns and the FP64 diagonal, and its
// LDS-window tables, where set, select the window form.  BF16 matrix cores:
// BF16 weights and activations, FP32 accumulation.  Instantiated for FP64 and
// FP32 vectors.
template <typename TV>
int mlp_correct(int Nd, Ell A, const uint8_t *btag, const float *G, const void *net, float b3, const TV *res, TV *p,
                hipStream_t s);

}  // namespace mgm
