// mgm_internal.h -- internal host-side types of libmgm (not part of the ABI).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mgm.h"

namespace mgm {

// Error state shared by every entry point (thread-local last error).
void set_error(const std::string &msg);
const char *last_error();

#define MGM_FAIL(code, msg) do { ::mgm::set_error(msg); return (code); } while (0)

// Points of one level as the Gmsh reader leaves them (after corner removal),
// file order; the reference's read_PointStructure output.
struct PointSet {
    int N = 0, nb = 0, num_elem = 0, num_corners_removed = 0;
    double d_avg = 0;
    std::vector<double> x, y, z, nx, ny, nz;
    std::vector<uint8_t> btag, ctag;
};

int read_gmsh(const char *path, int dim, PointSet &out);
int generate_cloud(int dim, int n, uint64_t seed, double jitter, int with_volume_elements,
                   PointSet &pts, const char *gmsh_path);

uint64_t splitmix64(uint64_t &s);
double uniform01(uint64_t &s);

int poly_terms(int dim, int deg, short *px, short *py, short *pz);

}  // namespace mgm
