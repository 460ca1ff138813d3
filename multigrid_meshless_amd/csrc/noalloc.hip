// noalloc.hip -- mgm_spmv_rows_static: mgm_spmv_rows with its device staging
// in a file-scope device array instead of buffers allocated per call.
//
// For callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc).  The array is
// kScratchBytes per GPU, held until the process ends.  Its device address is
// looked up once per device when mgm_init selects that device
// (hipGetSymbolAddress); host code only ever passes that address -- the
// array's own name is a host-side stand-in, not a device pointer.  Every call
// on a GPU shares its one array: calls that overlap on one GPU (two host
// threads, or two streams) race on it.
#pragma clang fp contract(off)

#include "device.hpp"

namespace mgm {

namespace {

constexpr size_t kScratchBytes = 64ull << 20;   // 64 MiB per device
constexpr int kMaxDevices = 64;
constexpr int BLK = 256;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// scratch's device address per device (nullptr until that device's mgm_init)
void *g_scratch[kMaxDevices] = {};

inline size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

// One row per thread over row-major rows: y_i = sum_j D_ij f[c_ij] in stencil
// order from 0.0, each product rounded (the arithmetic of the sliced-ELL
// SpMV, hf/mat_lib.h:70-89), so both entry points agree bit for bit.
__global__ void __launch_bounds__(BLK)
rowmajor_spmv_kernel(int n, int m, const double *__restrict__ D, const int *__restrict__ cl,
                     const double *__restrict__ f, double *__restrict__ out)
{
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const double *w = D + (size_t)i * m;
    const int *c = cl + (size_t)i * m;
    double s = 0.0;
    for (int j = 0; j < m; j++) s += w[j] * f[c[j]];
    out[i] = s;
}

}  // namespace

// Called by mgm_init with `device` current: finds this device's copy of the
// array once.
int noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) MGM_FAIL(-1, "mgm_init: device index beyond the scratch table");
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) MGM_FAIL(-10, "scratch array has no device address");
    g_scratch[device] = p;
    return 0;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

int spmv_rows_static(int device, int n_rows, int n_cols, const double *D, const int32_t *cloud, int ncol,
                     const double *f, double *out, hipStream_t s)
{
    if (n_rows <= 0 || n_cols <= 0 || ncol <= 0 || !D || !cloud || !f || !out)
        MGM_FAIL(-1, "spmv_rows_static: bad arguments");
    if (device < 0 || device >= kMaxDevices || !g_scratch[device])
        MGM_FAIL(-9, "spmv_rows_static: mgm_init has not set up this device");
    // staging: rows (8 B) and columns (4 B) per entry, f, out
    const size_t nnz = (size_t)n_rows * (size_t)n_cols;
    const size_t oD = 0, oC = up256(oD + 8 * nnz), oF = up256(oC + 4 * nnz), oO = up256(oF + 8 * (size_t)ncol);
    const size_t need = oO + 8 * (size_t)n_rows;
    if (need > kScratchBytes)
        MGM_FAIL(-1, "spmv_rows_static: rows, columns and vectors need " + std::to_string(need) +
                         " bytes of staging, more than the " + std::to_string(kScratchBytes) +
                         "-byte scratch array (12 bytes per row entry + 8 per vector entry)");
    for (size_t q = 0; q < nnz; q++)
        if (cloud[q] < 0 || cloud[q] >= ncol) MGM_FAIL(-1, "spmv_rows_static: column index out of range");
    unsigned char *base = static_cast<unsigned char *>(g_scratch[device]);
    double *dD = reinterpret_cast<double *>(base + oD);
    int *dC = reinterpret_cast<int *>(base + oC);
    double *dF = reinterpret_cast<double *>(base + oF);
    double *dO = reinterpret_cast<double *>(base + oO);
    HIP_TRY(hipMemcpyAsync(dD, D, 8 * nnz, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dC, cloud, 4 * nnz, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dF, f, 8 * (size_t)ncol, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rowmajor_spmv_kernel, dim3((n_rows + BLK - 1) / BLK), dim3(BLK), 0, s, n_rows, n_cols,
                       (const double *)dD, (const int *)dC, (const double *)dF, dO);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, dO, 8 * (size_t)n_rows, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

}  // namespace mgm
