/*
 * ref_capi.cpp -- C entry points around the reference's own geometry code
 * (src/Math/Geometry.cpp, Sphere.cpp, Vector3D.cpp, Tensor3D.cpp and the
 * header-only src/DataStructures/Array3D.tpp), compiled from the reference
 * tree into oracle/_ref/libref_geom.so.  TEST INFRASTRUCTURE ONLY: used by
 * tests/ to pin the oracle's restatement against the reference itself.
 */
#include "Geometry.h"
#include "Sphere.h"
#include "Array3D.h"

extern "C" {

static void load(const double* p, int n, Point3D* t) {
    for (int q = 0; q < n; ++q) t[q] = Point3D(p[3 * q], p[3 * q + 1], p[3 * q + 2]);
}
double ref_hex_volume(const double* p) { Point3D t[8]; load(p, 8, t); return Geometry::computeHexahedronVolume(t); }
void ref_hex_centroid(const double* p, double* o) {
    Point3D t[8]; load(p, 8, t);
    Point3D c = Geometry::computeHexahedronCentroid(t);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}
double ref_quad_area(const double* p) { Point3D t[4]; load(p, 4, t); return Geometry::computeQuadrilateralArea(t); }
void ref_quad_centroid(const double* p, double* o) {
    Point3D t[4]; load(p, 4, t);
    Point3D c = Geometry::computeQuadrilateralCentroid(t);
    o[0] = c.x; o[1] = c.y; o[2] = c.z;
}
int ref_inside_hex(const double* pt, const double* p) {
    Point3D t[8]; load(p, 8, t);
    return Geometry::isInsideHexahedron(Point3D(pt[0], pt[1], pt[2]), t) ? 1 : 0;
}
/* HexaFvmMesh.cpp:725-729: unit(cross(p1-p0, p2-p0)) * quadArea for one face */
void ref_face_area_normal(const double* p, double* o) {
    Point3D t[4]; load(p, 4, t);
    Vector3D n = cross(t[1] - t[0], t[2] - t[0]).unitVector() * Geometry::computeQuadrilateralArea(t);
    o[0] = n.x; o[1] = n.y; o[2] = n.z;
}
void ref_sphere_nearest(double r, const double* c, const double* pt, double* o) {
    Sphere s(r, Point3D(c[0], c[1], c[2]));
    Point3D q = s.nearestIntersect(Point3D(pt[0], pt[1], pt[2]));
    o[0] = q.x; o[1] = q.y; o[2] = q.z;
}
int ref_sphere_inside(double r, const double* c, const double* pt) {
    Sphere s(r, Point3D(c[0], c[1], c[2]));
    return s.isInside(Point3D(pt[0], pt[1], pt[2])) ? 1 : 0;
}
/* Array3D<T>::operator() linear offset for (i,j,k) (Array3D.tpp:40-58) */
long ref_array3d_offset(int nI, int nJ, int nK, int i, int j, int k) {
    Array3D<double> a(nI, nJ, nK);
    return (long)(&a(i, j, k) - &a[0]);
}
}
