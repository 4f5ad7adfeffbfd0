// ConnStatic's candidate sort for the oracle (TEST INFRASTRUCTURE ONLY; the product's host model
// calls the same standard-library routine). Connections/ConnStatic.cpp:62 sorts a vector of
// {dist, dest_neuron} records with operator< comparing dist only (ConnStatic.h:138-146), so records
// of equal distance end in whatever order std::sort leaves them; calling std::sort on records of
// the same layout with the same comparison reproduces that order.
#include <algorithm>
#include <cstdint>
#include <vector>

#include "bg_oracle.h"

namespace {
struct Cand {
    float dist;
    int dest;
    bool operator<(const Cand &o) const { return dist < o.dist; }
};
}  // namespace

extern "C" void bgo_sort_by_dist(float *dist, int32_t *dest, int32_t n)
{
    std::vector<Cand> v((size_t)(n > 0 ? n : 0));
    for (int32_t i = 0; i < n; i++) v[i] = Cand{dist[i], dest[i]};
    std::sort(v.begin(), v.end());
    for (int32_t i = 0; i < n; i++) {
        dist[i] = v[i].dist;
        dest[i] = v[i].dest;
    }
}
