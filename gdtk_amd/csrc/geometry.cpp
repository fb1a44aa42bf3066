// geometry.cpp -- host-side cell and face geometry of a structured block.
//
// Computed once per block on the host (as the reference does at start-up) and
// uploaded.  Formulas and operation order follow
//   src/eilmer/fvinterface.d:301-350 (face frame and area),
//   src/geom/elements/properties.d:97-141 (quad_properties), 177-237
//   (xyplane_quad_cell_properties), 293-304 (tetrahedron_volume), 390-429
//   (pyramid volume), 590-680 (hex_cell_properties),
//   src/geom/elements/vector3.d:362-381 (normalize), 616-621 (cross).
#include "geometry.hpp"

#include <cmath>
#include <stdexcept>

namespace gdtk {

namespace {
struct V3 {
    double x, y, z;
};
inline void cross(V3& v3, const V3& v1, const V3& v2) {
    v3.x = v1.y * v2.z - v2.y * v1.z;
    v3.y = v2.x * v1.z - v1.x * v2.z;
    v3.z = v1.x * v2.y - v2.x * v1.y;
}
inline void normalize(V3& v) {
    double m = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (m > 0.0) { v.x /= m; v.y /= m; v.z /= m; } else { v.x = v.y = v.z = 0.0; }
    const double small = 1.0e-30;
    if (std::fabs(v.x) < small) v.x = 0.0;
    if (std::fabs(v.y) < small) v.y = 0.0;
    if (std::fabs(v.z) < small) v.z = 0.0;
}
inline void cross_product(double ax, double ay, double az, double bx, double by, double bz, double& cx, double& cy,
                          double& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}
void quad(const V3& p0, const V3& p1, const V3& p2, const V3& p3, double* g /*10*/) {
    double p01x = p1.x - p0.x + p2.x - p3.x;
    double p01y = p1.y - p0.y + p2.y - p3.y;
    double p01z = p1.z - p0.z + p2.z - p3.z;
    double p03x = p3.x - p0.x + p2.x - p1.x;
    double p03y = p3.y - p0.y + p2.y - p1.y;
    double p03z = p3.z - p0.z + p2.z - p1.z;
    double vx, vy, vz;
    cross_product(p01x, p01y, p01z, p03x, p03y, p03z, vx, vy, vz);
    vx *= 0.25; vy *= 0.25; vz *= 0.25;
    double area = std::sqrt(vx * vx + vy * vy + vz * vz);
    V3 n, t1, t2;
    if (area > 1.0e-20) {
        n = {vx / area, vy / area, vz / area};
        t1 = {p01x, p01y, p01z};
        normalize(t1);
        cross(t2, n, t1);
        normalize(t2);
    } else {
        n = {1.0, 0.0, 0.0}; t1 = {0.0, 1.0, 0.0}; t2 = {0.0, 0.0, 1.0};
    }
    g[0] = n.x; g[1] = n.y; g[2] = n.z;
    g[3] = t1.x; g[4] = t1.y; g[5] = t1.z;
    g[6] = t2.x; g[7] = t2.y; g[8] = t2.z;
    g[9] = area;
}
void face2d(const V3& A, const V3& B, double* g) {
    double xA = A.x, yA = A.y, xB = B.x, yB = B.y;
    double LAB = std::sqrt((xB - xA) * (xB - xA) + (yB - yA) * (yB - yA));
    V3 n, t1, t2;
    double length;
    if (LAB > 1.0e-12) {
        n = {(yB - yA) / LAB, -(xB - xA) / LAB, 0.0};
        t2 = {0.0, 0.0, 1.0};
        cross(t1, n, t2);
        length = LAB;
    } else {
        n = {1.0, 0.0, 0.0}; t2 = {0.0, 0.0, 1.0}; t1 = {0.0, 1.0, 0.0};
        length = 0.0;
    }
    g[0] = n.x; g[1] = n.y; g[2] = n.z;
    g[3] = t1.x; g[4] = t1.y; g[5] = t1.z;
    g[6] = t2.x; g[7] = t2.y; g[8] = t2.z;
    g[9] = length;
}
double tet_volume(const V3& p0, const V3& p1, const V3& p2, const V3& p3) {
    double dx01 = p1.x - p0.x, dy01 = p1.y - p0.y, dz01 = p1.z - p0.z;
    double dx02 = p2.x - p0.x, dy02 = p2.y - p0.y, dz02 = p2.z - p0.z;
    double cx, cy, cz;
    cross_product(dx01, dy01, dz01, dx02, dy02, dz02, cx, cy, cz);
    double dx03 = p3.x - p0.x, dy03 = p3.y - p0.y, dz03 = p3.z - p0.z;
    return (dx03 * cx + dy03 * cy + dz03 * cz) / 6.0;
}
double pyramid_volume(const V3& p0, const V3& p1, const V3& p2, const V3& p3, const V3& p4) {
    V3 pmB{0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y), 0.25 * (p0.z + p1.z + p2.z + p3.z)};
    double v = 0.0;
    v += tet_volume(p0, p1, pmB, p4);
    v += tet_volume(p1, p2, pmB, p4);
    v += tet_volume(p2, p3, pmB, p4);
    v += tet_volume(p3, p0, pmB, p4);
    return v;
}
void hex(const V3* p, double& volume, double& iLen, double& jLen, double& kLen, V3& centroid) {
    const V3 &p0 = p[0], &p1 = p[1], &p2 = p[2], &p3 = p[3], &p4 = p[4], &p5 = p[5], &p6 = p[6], &p7 = p[7];
    V3 c{0.125 * (p0.x + p1.x + p2.x + p3.x + p4.x + p5.x + p6.x + p7.x),
         0.125 * (p0.y + p1.y + p2.y + p3.y + p4.y + p5.y + p6.y + p7.y),
         0.125 * (p0.z + p1.z + p2.z + p3.z + p4.z + p5.z + p6.z + p7.z)};
    V3 pmN{0.25 * (p3.x + p2.x + p6.x + p7.x), 0.25 * (p3.y + p2.y + p6.y + p7.y), 0.25 * (p3.z + p2.z + p6.z + p7.z)};
    V3 pmE{0.25 * (p1.x + p2.x + p6.x + p5.x), 0.25 * (p1.y + p2.y + p6.y + p5.y), 0.25 * (p1.z + p2.z + p6.z + p5.z)};
    V3 pmS{0.25 * (p0.x + p1.x + p5.x + p4.x), 0.25 * (p0.y + p1.y + p5.y + p4.y), 0.25 * (p0.z + p1.z + p5.z + p4.z)};
    V3 pmW{0.25 * (p0.x + p3.x + p7.x + p4.x), 0.25 * (p0.y + p3.y + p7.y + p4.y), 0.25 * (p0.z + p3.z + p7.z + p4.z)};
    V3 pmT{0.25 * (p4.x + p5.x + p6.x + p7.x), 0.25 * (p4.y + p5.y + p6.y + p7.y), 0.25 * (p4.z + p5.z + p6.z + p7.z)};
    V3 pmB{0.25 * (p0.x + p1.x + p2.x + p3.x), 0.25 * (p0.y + p1.y + p2.y + p3.y), 0.25 * (p0.z + p1.z + p2.z + p3.z)};
    double dx, dy, dz;
    dx = pmE.x - pmW.x; dy = pmE.y - pmW.y; dz = pmE.z - pmW.z;
    iLen = std::sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmN.x - pmS.x; dy = pmN.y - pmS.y; dz = pmN.z - pmS.z;
    jLen = std::sqrt(dx * dx + dy * dy + dz * dz);
    dx = pmT.x - pmB.x; dy = pmT.y - pmB.y; dz = pmT.z - pmB.z;
    kLen = std::sqrt(dx * dx + dy * dy + dz * dz);
    volume = 0.0;
    volume += pyramid_volume(p6, p7, p3, p2, c);
    volume += pyramid_volume(p5, p6, p2, p1, c);
    volume += pyramid_volume(p4, p5, p1, p0, c);
    volume += pyramid_volume(p7, p4, p0, p3, c);
    volume += pyramid_volume(p7, p6, p5, p4, c);
    volume += pyramid_volume(p0, p1, p2, p3, c);
    if ((volume < 0.0 && std::fabs(volume) < 1.0e-12) || (volume >= 0.0 && volume < 1.0e-20)) volume = 0.0;
    centroid = c;
}
}  // namespace

void compute_block_geometry(int dims, int nic, int njc, int nkc, const double* vtx, BlockGeometry& out) {
    const int niv = nic + 1, njv = njc + 1, nkv = (dims == 3) ? nkc + 1 : 1;
    auto V = [&](int i, int j, int k) -> V3 {
        size_t n = ((size_t)k * njv + j) * niv + i;
        return V3{vtx[3 * n], vtx[3 * n + 1], vtx[3 * n + 2]};
    };
    const size_t ncell = (size_t)nic * njc * nkc;
    out.volume.assign(ncell, 0.0);
    out.iLen.assign(ncell, 0.0);
    out.jLen.assign(ncell, 0.0);
    out.kLen.assign(ncell, 0.0);
    out.pos.assign(3 * ncell, 0.0);
    out.nf[0][0] = niv; out.nf[0][1] = njc; out.nf[0][2] = nkc;
    out.nf[1][0] = nic; out.nf[1][1] = njv; out.nf[1][2] = nkc;
    out.nf[2][0] = nic; out.nf[2][1] = njc; out.nf[2][2] = (dims == 3) ? nkv : 0;
    for (int d = 0; d < 3; ++d) {
        size_t nface = (size_t)out.nf[d][0] * out.nf[d][1] * out.nf[d][2];
        out.face[d].assign(10 * nface, 0.0);
    }
    auto put = [&](int d, int fi, int fj, int fk, const double* g) {
        size_t nface = (size_t)out.nf[d][0] * out.nf[d][1] * out.nf[d][2];
        size_t f = (size_t)fi + (size_t)out.nf[d][0] * ((size_t)fj + (size_t)out.nf[d][1] * fk);
        for (int q = 0; q < 10; ++q) out.face[d][f + q * nface] = g[q];
    };
    double g[10];
    if (dims == 2) {
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i < nic; ++i) {
                // properties.d:177-237 with p0..p3 = (i,j),(i+1,j),(i+1,j+1),(i,j+1)
                V3 p0 = V(i, j, 0), p1 = V(i + 1, j, 0), p2 = V(i + 1, j + 1, 0), p3 = V(i, j + 1, 0);
                double xA = p1.x, yA = p1.y, xB = p2.x, yB = p2.y, xC = p3.x, yC = p3.y, xD = p0.x, yD = p0.y;
                double area = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) + (xD + xC) * (yD - yC) +
                                     (xA + xD) * (yA - yD));
                if (area < 0.0) throw std::runtime_error("Negative cell volume");
                // centroid of the quadrilateral (properties.d:177-237)
                const double cx = 1.0 / (area * 6.0) *
                                  ((yB - yA) * (xA * xA + xA * xB + xB * xB) + (yC - yB) * (xB * xB + xB * xC + xC * xC) +
                                   (yD - yC) * (xC * xC + xC * xD + xD * xD) + (yA - yD) * (xD * xD + xD * xA + xA * xA));
                const double cy = -1.0 / (area * 6.0) *
                                  ((xB - xA) * (yA * yA + yA * yB + yB * yB) + (xC - xB) * (yB * yB + yB * yC + yC * yC) +
                                   (xD - xC) * (yC * yC + yC * yD + yD * yD) + (xA - xD) * (yD * yD + yD * yA + yA * yA));
                double xN = 0.5 * (xC + xB), yN = 0.5 * (yC + yB);
                double xS = 0.5 * (xD + xA), yS = 0.5 * (yD + yA);
                double xE = 0.5 * (xA + xB), yE = 0.5 * (yA + yB);
                double xW = 0.5 * (xD + xC), yW = 0.5 * (yD + yC);
                double dx = xN - xS, dy = yN - yS;
                size_t c = (size_t)j * nic + i;
                out.jLen[c] = std::sqrt(dx * dx + dy * dy);
                dx = xE - xW; dy = yE - yW;
                out.iLen[c] = std::sqrt(dx * dx + dy * dy);
                out.kLen[c] = 0.0;
                out.volume[c] = area;
                out.pos[c] = cx;
                out.pos[c + ncell] = cy;
            }
        for (int j = 0; j < njc; ++j)
            for (int i = 0; i < niv; ++i) { face2d(V(i, j, 0), V(i, j + 1, 0), g); put(0, i, j, 0, g); }
        for (int j = 0; j < njv; ++j)
            for (int i = 0; i < nic; ++i) { face2d(V(i + 1, j, 0), V(i, j, 0), g); put(1, i, j, 0, g); }
    } else {
        for (int k = 0; k < nkc; ++k)
            for (int j = 0; j < njc; ++j)
                for (int i = 0; i < nic; ++i) {
                    V3 p[8] = {V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k),
                               V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j + 1, k + 1), V(i, j + 1, k + 1)};
                    size_t c = ((size_t)k * njc + j) * nic + i;
                    double vol, iL, jL, kL;
                    V3 cen;
                    hex(p, vol, iL, jL, kL, cen);
                    if (vol <= 0.0) throw std::runtime_error("Invalid cell volume");
                    out.volume[c] = vol; out.iLen[c] = iL; out.jLen[c] = jL; out.kLen[c] = kL;
                    out.pos[c] = cen.x; out.pos[c + ncell] = cen.y; out.pos[c + 2 * ncell] = cen.z;
                }
        for (int k = 0; k < nkc; ++k)
            for (int j = 0; j < njc; ++j)
                for (int i = 0; i < niv; ++i) {
                    quad(V(i, j, k), V(i, j + 1, k), V(i, j + 1, k + 1), V(i, j, k + 1), g);
                    put(0, i, j, k, g);
                }
        for (int k = 0; k < nkc; ++k)
            for (int j = 0; j < njv; ++j)
                for (int i = 0; i < nic; ++i) {
                    quad(V(i, j, k), V(i, j, k + 1), V(i + 1, j, k + 1), V(i + 1, j, k), g);
                    put(1, i, j, k, g);
                }
        for (int k = 0; k < nkv; ++k)
            for (int j = 0; j < njc; ++j)
                for (int i = 0; i < nic; ++i) {
                    quad(V(i, j, k), V(i + 1, j, k), V(i + 1, j + 1, k), V(i, j + 1, k), g);
                    put(2, i, j, k, g);
                }
    }
}

}  // namespace gdtk
