// kdtree_ref.cpp -- test-only C shim over the reference's kd-tree layer
// (kdtree/kdtree.h + kdtree/nanoflann.hpp, included from the reference tree
// by oracle/Makefile, never copied).  It exposes the reference's exact
// neighbour searches so the oracle's and the GPU's grids can be checked
// against them:
//   radius search with SearchParameters(0, false)   (ppf.cpp:273-274)
//   kNN search                                      (util.cpp:244)
//   greedy sampling over the tree order vAcc_        (util.cpp:107-139, 1 thread)
#include <kdtree.h>

#include <array>
#include <cstring>
#include <vector>

using Cloud  = std::vector<std::array<float, 3>>;
using KDTree = KDTreeVectorOfVectorsAdaptor<Cloud, float>;

static Cloud to_cloud(const float *p, int n) {
    Cloud c(n);
    for (int i = 0; i < n; i++)
        c[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    return c;
}

extern "C" {

// ids in nanoflann traversal order; returns the total count
int kdref_radius(const float *pts, int n, const float *q, float r2, int *ids, int cap) {
    Cloud  c = to_cloud(pts, n);
    KDTree t(c, 10);
    t.index->buildIndex();
    std::vector<nanoflann::ResultItem<int, float>> out;
    t.index->radiusSearch(q, r2, out, nanoflann::SearchParameters(0, false));
    int k = 0;
    for (auto &r : out)
        if (k < cap)
            ids[k++] = r.first;
    return (int)out.size();
}

int kdref_knn(const float *pts, int n, const float *q, int k, int *ids, float *d2) {
    Cloud  c = to_cloud(pts, n);
    KDTree t(c, 10);
    t.index->buildIndex();
    std::vector<int>   ii(k);
    std::vector<float> dd(k);
    int                got = (int)t.index->knnSearch(q, k, ii.data(), dd.data());
    std::memcpy(ids, ii.data(), sizeof(int) * got);
    std::memcpy(d2, dd.data(), sizeof(float) * got);
    return got;
}

// tree order (vAcc_) and the reference's greedy sampling in that order
int kdref_greedy(const float *pts, int n, float step, int *order, int *out) {
    Cloud  c = to_cloud(pts, n);
    KDTree t(c, 10);
    t.index->buildIndex();
    auto &vacc = t.index->vAcc_;
    for (int i = 0; i < n; i++)
        order[i] = vacc[i];
    std::vector<bool> keep(n, true);
    float             radius = step * step;
    int               cnt    = 0;
    for (size_t i = 0; i < vacc.size(); i++) {
        int index = vacc[i];
        if (!keep[index])
            continue;
        out[cnt++] = index;
        std::vector<nanoflann::ResultItem<int, float>> indices;
        nanoflann::RadiusResultSet<float, int>         rs(radius, indices);
        t.index->findNeighbors(rs, &c[index][0], nanoflann::SearchParameters(0, false));
        for (auto &r : indices)
            keep[r.first] = false;
    }
    return cnt;
}
}
