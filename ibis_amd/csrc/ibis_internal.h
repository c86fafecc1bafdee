// ibis_internal.h -- shared host-side definitions of libibis_gpu.
#ifndef IBIS_INTERNAL_H
#define IBIS_INTERNAL_H

#include <cstdint>
#include <vector>

#include "../../include/ibis_gpu.h"

struct ibis_grid {
    int dim = 2;
    int64_t N = 0, G = 0, F = 0;
    int ntags = 0, n_lsq = 3;
    std::vector<double> geom;  // 13 x F
    std::vector<int64_t> left, right;
    std::vector<int64_t> cf_off, cf, nbr;
    std::vector<int32_t> outsign;
    std::vector<double> centroid, volume, lsq;
    std::vector<int64_t> tag_off, bfaces, ghosts;
    std::vector<int64_t> cell_new_of_old, face_new_of_old;
};

void ibis_grid_fill_desc(const ibis_grid* g, ibis_mesh_desc* d);

namespace ibis {
// Record a failure message for ibis_last_error() and return the status.
ibis_status fail(ibis_status s, const char* msg);
}  // namespace ibis

#endif
