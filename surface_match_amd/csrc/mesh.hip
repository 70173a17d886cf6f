// Mesh models: sampleMesh (util.cpp:20-84), run by trainModel when the
// model has faces (ppf.cpp:63-70).
//
// The reference draws its points from std::mt19937 seeded by
// std::random_device, so its samples differ from run to run; here every
// point j of the sample takes its two uniforms from Philox4x32-10 at counter
// (j, 0, 0, 0) under a caller-given 64-bit key, so the sample is fixed by
// the seed and each point is independent of the others.  Everything else
// follows the reference: per-triangle areas |u x v| / 2 and unit normals in
// float, their total and the area CDF summed in triangle order, nPoints =
// int(area / radius^2), triangle i taking points until the running count
// reaches round(cdf_i * nPoints), and each point a = 1 - sqrt(r1),
// b = sqrt(r1) (1 - r2), c = sqrt(r1) r2 of its triangle's corners.
//
// Work split: the per-triangle areas are independent (one GPU pass); the
// total and the CDF are float sums whose order decides the counts, so they
// run in triangle order on the host over the areas copied back; the points
// are independent again (one GPU thread each, which finds its triangle by
// binary search over the per-triangle first indices and forms the
// triangle's normal itself).
#include "ppf_dev.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

namespace ppfd {
namespace {

// areas of all triangles; a face naming a missing point sets *bad and
// reads no point
__global__ void mesh_area_kernel(const float *P, int nPts, const int *faces, int nTri,
                                 float *area, int *bad) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTri)
        return;
    const int a = faces[3 * (size_t)t], b = faces[3 * (size_t)t + 1], c = faces[3 * (size_t)t + 2];
    if (a < 0 || b < 0 || c < 0 || a >= nPts || b >= nPts || c >= nPts) {
        *bad    = 1;
        area[t] = 0.0f;
        return;
    }
    float n[3];
    area[t] = ppfc::tri_area_normal(P + 3 * (size_t)a, P + 3 * (size_t)b, P + 3 * (size_t)c, n);
}

__global__ void mesh_sample_kernel(const float *P, const int *faces, const int *offs, int nTri,
                                   int nOut, uint32_t k0, uint32_t k1, float *outP, float *outN) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nOut)
        return;
    // the triangle owning point j: the last t with offs[t] <= j
    int lo = 0, hi = nTri - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= j)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int    t  = lo;
    const float *p0 = P + 3 * (size_t)faces[3 * (size_t)t];
    const float *p1 = P + 3 * (size_t)faces[3 * (size_t)t + 1];
    const float *p2 = P + 3 * (size_t)faces[3 * (size_t)t + 2];
    uint32_t     c[4] = {(uint32_t)j, 0u, 0u, 0u};
    ppfc::philox4x32_10(c, k0, k1);
    float p[3], n[3];
    ppfc::mesh_point(p0, p1, p2, ppfc::uniform01(c[0]), ppfc::uniform01(c[1]), p);
    ppfc::tri_area_normal(p0, p1, p2, n);
    for (int k = 0; k < 3; k++) {
        outP[3 * (size_t)j + k] = p[k];
        outN[3 * (size_t)j + k] = n[k];
    }
}

// a mesh resident on the device, with its triangle areas copied back
struct MeshDev {
    DevBuf             dP, dF, dA, dBad;
    std::vector<float> area;
};

void mesh_upload_areas(MeshDev &m, const float *pts, int nPts, const int *faces, int nFaces,
                       hipStream_t s) {
    float *P   = m.dP.as<float>((size_t)3 * std::max(nPts, 1));
    int   *F   = m.dF.as<int>((size_t)3 * nFaces);
    float *A   = m.dA.as<float>(nFaces);
    int   *bad = m.dBad.as<int>(1);
    if (nPts > 0)
        PPF_HIP_CHECK(hipMemcpyAsync(P, pts, sizeof(float) * 3 * nPts, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemcpyAsync(F, faces, sizeof(int) * 3 * (size_t)nFaces, hipMemcpyHostToDevice, s));
    PPF_HIP_CHECK(hipMemsetAsync(bad, 0, sizeof(int), s));
    mesh_area_kernel<<<div_up(nFaces, 256), 256, 0, s>>>(P, nPts, F, nFaces, A, bad);
    PPF_HIP_CHECK(hipGetLastError());
    m.area.resize(nFaces);
    int hbad = 0;
    PPF_HIP_CHECK(hipMemcpyAsync(m.area.data(), A, sizeof(float) * nFaces, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
    if (hbad)
        throw std::runtime_error("Invalid Input: a face refers to a missing point");
}

float mesh_total(const std::vector<float> &area) {
    float total = 0.0f;
    for (float a : area)
        total += a;
    return total;
}

// nPoints = int(total / radius^2) (util.cpp:50)
int mesh_count(float total, float radius) {
    const float q = total / (radius * radius);
    if (!(q >= 0.0f && q < 2147483648.0f))
        throw std::runtime_error("Invalid Input: mesh area or radius out of range in sampleMesh");
    return (int)q;
}

// each triangle's first output index (util.cpp:45-62); returns the count
int mesh_offsets(const std::vector<float> &area, float total, int nPoints,
                 std::vector<int> &offs) {
    const int nFaces = (int)area.size();
    offs.assign(nFaces, 0);
    if (nPoints == 0)  // an empty sample, as the reference's
        return 0;
    size_t idx = 0;
    float  cdf = 0.0f;
    for (int i = 0; i < nFaces; i++) {
        cdf      = i == 0 ? area[0] / total : area[i] / total + cdf;
        if (i == nFaces - 1)
            cdf = 1.f;
        offs[i]  = (int)idx;
        size_t n = (size_t)std::round(cdf * (float)nPoints);
        n        = std::min(n, (size_t)nPoints);  // the reference writes past its buffer here
        idx      = std::max(idx, n);
    }
    return (int)idx;
}

}  // namespace

int sample_mesh_count(const float *pts, int nPts, const int *faces, int nFaces, float radius,
                      hipStream_t s) {
    if (nFaces <= 0)
        return 0;
    MeshDev m;
    mesh_upload_areas(m, pts, nPts, faces, nFaces, s);
    return mesh_count(mesh_total(m.area), radius);
}

void sample_mesh(const float *pts, int nPts, const int *faces, int nFaces, float radius,
                 uint64_t seed, std::vector<float> &outP, std::vector<float> &outN,
                 hipStream_t s) {
    outP.clear();
    outN.clear();
    if (nFaces <= 0)
        return;
    MeshDev m;
    mesh_upload_areas(m, pts, nPts, faces, nFaces, s);
    const float      total = mesh_total(m.area);
    std::vector<int> offs;
    const int        nOut = mesh_offsets(m.area, total, mesh_count(total, radius), offs);
    if (nOut == 0)
        return;
    DevBuf dO, oP, oN;
    int   *O  = dO.as<int>(nFaces);
    float *OP = oP.as<float>((size_t)3 * nOut);
    float *ON = oN.as<float>((size_t)3 * nOut);
    PPF_HIP_CHECK(hipMemcpyAsync(O, offs.data(), sizeof(int) * nFaces, hipMemcpyHostToDevice, s));
    mesh_sample_kernel<<<div_up(nOut, 256), 256, 0, s>>>(m.dP.get<float>(), m.dF.get<int>(), O,
                                                         nFaces, nOut, (uint32_t)seed,
                                                         (uint32_t)(seed >> 32), OP, ON);
    PPF_HIP_CHECK(hipGetLastError());
    outP.resize((size_t)3 * nOut);
    outN.resize((size_t)3 * nOut);
    PPF_HIP_CHECK(hipMemcpyAsync(outP.data(), OP, sizeof(float) * 3 * nOut, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipMemcpyAsync(outN.data(), ON, sizeof(float) * 3 * nOut, hipMemcpyDeviceToHost, s));
    PPF_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace ppfd
