// The facade's host helpers (util.h / helper.h of the reference) on small
// hand-made clouds; no GPU call.  Prints one line per check for the test.
#include <ppf/ppf.h>

#include <cstdio>

int main() {
    ppf::PointCloud pc;
    pc.point  = {ppf::Vec3f(1, 2, 3), ppf::Vec3f(NAN, 0, 0), ppf::Vec3f(-4, 5, 0.5f),
                 ppf::Vec3f(2, -1, 7)};
    pc.normal = {ppf::Vec3f(0, 0, 2), ppf::Vec3f(1, 0, 0), ppf::Vec3f(0, INFINITY, 0),
                 ppf::Vec3f(0, 0, 0)};
    auto v  = ppf::removeNan(pc);
    auto vn = ppf::removeNan(pc, true);
    std::printf("removeNan %zu %d %d %d | %zu %d %d\n", v.size(), v[0], v[1], v[2], vn.size(),
                vn[0], vn[1]);
    auto b = ppf::computeBoundingBox(pc, v);
    std::printf("box %g %g %g %g %g %g diameter %.6f\n", b.min[0], b.min[1], b.min[2], b.max[0],
                b.max[1], b.max[2], b.diameter());
    auto e = ppf::extraIndices(pc, {3, 0});
    std::printf("extra %zu %g %g %g n %g %g %g | %g %g %g box %g %g\n", e.size(), e.point[0][0],
                e.point[0][1], e.point[0][2], e.normal[0][0], e.normal[0][1], e.normal[0][2],
                e.normal[1][0], e.normal[1][1], e.normal[1][2], e.box.min[0], e.box.max[2]);
    ppf::PointCloud q = pc;
    ppf::normalizeNormal(q, true);
    std::printf("invert %g %g %g | %g %g %g\n", q.normal[0][0], q.normal[0][1], q.normal[0][2],
                q.normal[3][0], q.normal[3][1], q.normal[3][2]);
    // a quarter turn about z and a shift
    ppf::Mat4f T;
    T(0, 0) = 0, T(0, 1) = -1, T(1, 0) = 1, T(1, 1) = 0;
    T(0, 3) = 10, T(1, 3) = 20, T(2, 3) = 30;
    ppf::PointCloud two;
    two.point  = {ppf::Vec3f(1, 2, 3), ppf::Vec3f(-1, 0, 4)};
    two.normal = {ppf::Vec3f(1, 0, 0), ppf::Vec3f(0, 1, 0)};
    two.face   = {ppf::Vec3i(0, 1, 1)};
    auto t = ppf::transformPointCloud(two, T);
    std::printf("transform %g %g %g | %g %g %g n %g %g %g faces %zu box %g %g\n", t.point[0][0],
                t.point[0][1], t.point[0][2], t.point[1][0], t.point[1][1], t.point[1][2],
                t.normal[0][0], t.normal[0][1], t.normal[0][2], t.face.size(), t.box.min[0],
                t.box.max[1]);
    auto tn = ppf::transformPointCloud(two, T, false);
    std::printf("no normals %zu\n", tn.normal.size());
    {
        ppf::Timer timer("timer");
    }
    ppf::PointCloud flat;
    std::printf("sampleMesh without faces %zu\n", ppf::sampleMesh(flat, 1.0f).size());
    return 0;
}
