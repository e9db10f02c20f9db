/*
 * oracle_angle.h — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp).
 * Umbrella angle constraint: neighbour pairing (forward.cu:10708-10819) and
 * the two-pass Jacobi projection (forward.cu:4990-5130).
 */
#pragma once
#include <vector>

#include "xpbd_oracle_internal.h"

namespace xo {

// Per directed CSR entry e of centre i (entry e = i -> nbr[e]): the paired
// neighbour (the next one around i in angular order, -1 = none) and the rest
// cosine between the two directions (2 = unused).
struct AnglePairs {
    bool built = false;
    std::vector<int> pairNext;
    std::vector<float> restCos;
};

extern AnglePairs ang;

// Pairing from the positions the graph was uploaded with (the rest positions).
void angle_build(const std::vector<V3>& pos);
// One angle pass: contributions accumulated per node in the serial order
// (centre, entry, role centre -> first -> second), averaged and blended.
void angle_pass(const std::vector<V3>& pin, std::vector<V3>& pout);

}  // namespace xo
