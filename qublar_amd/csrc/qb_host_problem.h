// qb_host_problem.h -- a problem built on the host side of the library (its arrays in
// host memory), shared by the host builders (qb_scene.cpp) and the device builders
// (qb_build.hip), which copy their results into one.
#pragma once

#include <stdint.h>

#include <vector>

struct qb_host_problem {
    int n_vars = 0, n_rays = 0;
    std::vector<int32_t> row_ptr, ray, nbr_ptr, nbr;
    std::vector<float> a;
    std::vector<double> d, w;
    double lambda = 0.0, kappa = 0.0;
    std::vector<uint8_t> truth;
    std::vector<int32_t> colour;   // empty unless the builder colours its problem
    std::vector<double> field;     // per-bit linear field; empty: lin = kappa
};
