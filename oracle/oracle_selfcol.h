/* oracle_selfcol.h — TEST INFRASTRUCTURE ONLY (see xpbd_oracle.cpp). */
#pragma once
#include <vector>

#include "xpbd_oracle_internal.h"

namespace xo {

struct ObjectState {
    bool valid = false;
    int minComp = -1;
    int K = 0;
    std::vector<int> id, count;
    std::vector<float> c0, params, disp, boxes;
    std::vector<double> restRel, sumQ, shift;
    bool shiftValid = false, boxesValid = false;
    float maxBodyDisp = 0.0f;
};

struct SelfColState {
    float spacing = 0.0f, dc = 0.0f, invCell = 0.0f;
    int table = 0, bits = 0;
    std::vector<unsigned char> active;
    std::vector<unsigned int> keys;
    std::vector<int> order, cellStart, cellEnd, contacts, counts;
    std::vector<V3> accum;
    long long contactsTotal = 0;
    int overflow = 0, activeCount = 0;
};

extern ObjectState obj;
extern SelfColState sc;

unsigned int hash_cell(int cx, int cy, int cz, unsigned int mask);
int cell_of(float v, float invCell);
bool object_ensure_components();
void object_shape_match(std::vector<V3>& p);
void selfcol_build(const std::vector<V3>& p);
void selfcol_iteration(std::vector<V3>& pos);
bool object_contact_push(std::vector<V3>& pos);

}  // namespace xo
