// vgrid.cpp -- icosahedral Voronoi velocity grid and initial angular
// distributions (host, FP64).
//
// The grid must reproduce the reference's directions and weights bit for bit
// (every moment sum depends on them), so the floating-point expressions below
// keep the reference's operand order: patch points (icosahedron.h:74-80),
// circumcentres (Point.h:64-69) and the Voronoi sub-areas accumulated patch by
// patch in the reference's loop order (icosahedron.h:137-170).  The border
// stitching, a lookup table in the reference (icosahedron.h:107-135), is a
// closed-form index function here.
#include "vgrid.h"

#include <cmath>
#include <cstdint>

namespace nuosc {
namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double k, V3 a) { return {k * a.x, k * a.y, k * a.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// circumcentre of the planar triangle abc
V3 circumcentre(V3 a, V3 b, V3 c) {
    const V3 ab = b - a, ac = c - a;
    const V3 n = cross(ab, ac);
    return a + (0.5 / norm2(n)) * cross(n, norm2(ac) * ab - norm2(ab) * ac);
}

// Area of the planar triangle (p, q, o) scaled as the reference does:
// |(o - p) x (q - p) * 0.25|, i.e. half the triangle (p, midpoint(pq), o).
inline double sub_area(V3 p, V3 q, V3 o) { return norm(0.25 * cross(o - p, q - p)); }

// The 10 rhombic patches of the icosahedron, each given by its corners
// (a, b, c, d) among the 12 base vertices: 0 = north pole, 1..5 upper ring,
// 6..10 lower ring, 11 = south pole (icosahedron.h:39-40).
constexpr int kPatch[10][4] = {{1, 0, 2, 6}, {6, 2, 7, 11}, {2, 0, 3, 7}, {7, 3, 8, 11}, {3, 0, 4, 8},
                               {8, 4, 9, 11}, {4, 0, 5, 9}, {9, 5, 10, 11}, {5, 0, 1, 10}, {10, 1, 6, 11}};

class Icosphere {
public:
    explicit Icosphere(unsigned nn) : nn_(nn), N_(10u * nn * nn + 2u), pts_(N_), w_(N_, 0.0) {
        base_vertices();
        patch_points();
        voronoi_weights();
    }
    unsigned size() const { return N_; }
    const V3& point(unsigned k) const { return pts_[k]; }
    double weight(unsigned k) const { return w_[k]; }

private:
    unsigned nn_, N_;
    V3 base_[12];
    std::vector<V3> pts_;
    std::vector<double> w_;

    void base_vertices() {
        const unsigned NP = 5;
        const double lambda = 1.10714871779409;  // icosahedron.h:41
        for (unsigned i = 1; i <= NP; ++i) {
            base_[i] = {std::cos(2 * i * M_PI / NP) * std::sin(lambda), std::sin(2 * i * M_PI / NP) * std::sin(lambda),
                        std::cos(lambda)};
            base_[i + NP] = {std::cos((2 * i + 1) * M_PI / NP) * std::sin(lambda),
                             std::sin((2 * i + 1) * M_PI / NP) * std::sin(lambda), -base_[i].z};
        }
        base_[0] = {0, 0, 1};
        base_[11] = {0, 0, -1};
    }

    // Interior lattice point (i, j), 0 <= i, j < nn, of patch p: bilinear in
    // the patch corners, projected on the sphere.  Id = p*nn^2 + i*nn + j + 1.
    void patch_points() {
        pts_[0] = base_[0];
        pts_[N_ - 1] = base_[11];
        for (unsigned p = 0; p < 10; ++p) {
            const V3 a = base_[kPatch[p][0]], b = base_[kPatch[p][1]];
            const V3 c = base_[kPatch[p][2]], d = base_[kPatch[p][3]];
            for (unsigned i = 0; i < nn_; ++i)
                for (unsigned j = 0; j < nn_; ++j) {
                    const V3 q = (double)i * ((double)j * c + (double)(nn_ - j) * d) +
                                 (double)(nn_ - i) * ((double)j * b + (double)(nn_ - j) * a);
                    pts_[p * nn_ * nn_ + i * nn_ + j + 1] = (1.0 / norm(q)) * q;
                }
        }
    }

    // Unique point id of lattice node (i, j), 0 <= i, j <= nn, of patch p: the
    // nodes on the far borders (i == nn or j == nn) belong to the neighbouring
    // patches or to a pole.
    unsigned node(unsigned p, unsigned i, unsigned j) const {
        const unsigned n = nn_, n2 = nn_ * nn_;
        if (i < n && j < n) return p * n2 + i * n + j + 1;
        if (p % 2 == 0) {  // upper patches
            if (j == n) return (i == 0) ? 0u : ((p + 2) * n2 + (n - i) + 1) % (10 * n2);
            return (p + 1) * n2 + j + 1;  // i == n, j < n
        }
        const unsigned right = (p == 9) ? 0u : (p + 1) * n2;     // patch across the upper-right side
        const unsigned below = (p == 9) ? 2 * n2 : (p + 3) * n2;  // patch across the bottom-right side
        if (j == n && i < n) return right + 1 + n * i;
        if (j == 0) return N_ - 1;  // i == n: south pole
        return below - (n - 1) - n * (j - 1);  // i == n, j >= 1
    }

    void voronoi_weights() {
        for (unsigned p = 0; p < 10; ++p)
            for (unsigned i = 0; i < nn_; ++i)
                for (unsigned j = 0; j < nn_; ++j) {
                    const unsigned ia = node(p, i, j), ic = node(p, i + 1, j + 1);
                    // the two triangles of lattice cell (i, j): b = (i, j+1) then b = (i+1, j)
                    const unsigned ibs[2] = {node(p, i, j + 1), node(p, i + 1, j)};
                    for (unsigned ib : ibs) {
                        const V3 a = pts_[ia], b = pts_[ib], c = pts_[ic];
                        const V3 o = circumcentre(a, b, c);
                        const double fab = sub_area(a, b, o), fbc = sub_area(b, c, o), fca = sub_area(c, a, o);
                        w_[ia] += fab + fca;
                        w_[ib] += fbc + fab;
                        w_[ic] += fca + fbc;
                    }
                }
    }
};

double gauss2(double vx, double vz, double sx, double sz) {
    // nuosc_init.cpp:8-10 with (vx0, vz0) = (1, 1)
    return std::exp(-(vx - 1.0) * (vx - 1.0) / (2.0 * sx * sx) - (vz - 1.0) * (vz - 1.0) / (2.0 * sz * sz));
}

}  // namespace

VGrid make_icosahedral_vgrid(int nn) {
    Icosphere s((unsigned)nn);
    VGrid g;
    const unsigned N = s.size();
    g.vx.resize(N);
    g.vy.resize(N);
    g.vz.resize(N);
    g.vw.resize(N);
    for (unsigned k = 0; k < N; ++k) {
        g.vx[k] = s.point(k).x;
        g.vy[k] = s.point(k).y;
        g.vz[k] = s.point(k).z;
        g.vw[k] = s.weight(k);
    }
    return g;
}

VGrid make_equal_weight_vgrid(int nn) {
    VGrid g = make_icosahedral_vgrid(nn);
    const double w = 4.0 * M_PI / g.size();
    for (auto& x : g.vw) x = w;
    return g;
}

void angular_distributions(const VGrid& g, double alpha, double lnue, double lnueb, double lnuex,
                           double lnuebx, std::vector<double>& G0, std::vector<double>& G0b) {
    const int n = g.size();
    std::vector<double> ng(n), ngb(n);
    double s0 = 0, s1 = 0;
    for (int v = 0; v < n; ++v) {
        ng[v] = gauss2(g.vx[v], g.vz[v], lnuex, lnue);
        ngb[v] = gauss2(g.vx[v], g.vz[v], lnuebx, lnueb);
        s0 += g.vw[v] * ng[v];
        s1 += g.vw[v] * ngb[v];
    }
    const double i0 = 1.0 / s0, i1 = 1.0 / s1;
    G0.resize(n);
    G0b.resize(n);
    for (int v = 0; v < n; ++v) {
        G0[v] = ng[v] * i0;
        G0b[v] = alpha * ngb[v] * i1;
    }
}

}  // namespace nuosc
