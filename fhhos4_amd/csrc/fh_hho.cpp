// HHO discretisation on structured affine meshes + static condensation.
#include "fh_hho.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <unordered_map>

#include "fh_common.h"
#include "fh_hho_dev.h"
#include "fh_kernels.h"

namespace fh {
namespace {

using V3 = std::array<double, 3>;
using I3 = std::array<int, 3>;

// ------------------------------------------------------------- quadrature
// Gauss-Legendre nodes/weights on [0,1]: Newton iteration on the Legendre
// three-term recurrence (Golub & Welsch / textbook method).
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < n; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 1;
        for (int it = 0; it < 100; it++) {
            double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) < 1e-16) break;
        }
        x[i] = 0.5 * (1 - z);
        w[i] = 1.0 / ((1 - z * z) * pp * pp);
    }
}

struct Quad {
    std::vector<V3> p;
    std::vector<double> w;
};

// Reference shapes: simplex (kind 0) with vertices 0, e_1..e_d, box (kind 1) [0,1]^d.
// Simplex rules: collapsed (Duffy / Stroud conical product) Gauss-Legendre.
Quad ref_quad(int dim, int kind, int n) {
    std::vector<double> x, w;
    gauss_legendre01(n, x, w);
    Quad q;
    if (dim == 1) {
        for (int i = 0; i < n; i++) { q.p.push_back({x[i], 0, 0}); q.w.push_back(w[i]); }
    } else if (dim == 2) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                if (kind == 1) { q.p.push_back({x[i], x[j], 0}); q.w.push_back(w[i] * w[j]); }
                else { q.p.push_back({x[i], x[j] * (1 - x[i]), 0}); q.w.push_back(w[i] * w[j] * (1 - x[i])); }
            }
    } else {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                for (int l = 0; l < n; l++) {
                    if (kind == 1) { q.p.push_back({x[i], x[j], x[l]}); q.w.push_back(w[i] * w[j] * w[l]); }
                    else {
                        double u = x[i], v = x[j], t = x[l];
                        q.p.push_back({u, v * (1 - u), t * (1 - u) * (1 - v)});
                        q.w.push_back(w[i] * w[j] * w[l] * (1 - u) * (1 - u) * (1 - v));
                    }
                }
    }
    return q;
}

// ------------------------------------------------------ orthonormal bases
// Monomials of total degree <= p in graded order, orthonormalised in L2 of
// the reference shape (coefficients C = L^-1, M = L L^T).
struct RefBasis {
    int dim = 0, p = 0, n = 0;
    std::vector<I3> e;
    std::vector<double> C;  // n x n, lower triangular
    void monos(const V3& x, double* m, double (*g)[3]) const {
        double pw[3][12], dpw[3][12];
        for (int a = 0; a < 3; a++) {
            pw[a][0] = 1; dpw[a][0] = 0;
            for (int q = 1; q <= p; q++) { pw[a][q] = pw[a][q - 1] * x[a]; dpw[a][q] = q * pw[a][q - 1]; }
        }
        for (int i = 0; i < n; i++) {
            const I3& E = e[i];
            m[i] = pw[0][E[0]] * pw[1][E[1]] * pw[2][E[2]];
            if (g) {
                g[i][0] = dpw[0][E[0]] * pw[1][E[1]] * pw[2][E[2]];
                g[i][1] = pw[0][E[0]] * dpw[1][E[1]] * pw[2][E[2]];
                g[i][2] = pw[0][E[0]] * pw[1][E[1]] * dpw[2][E[2]];
            }
        }
    }
    void eval(const V3& x, double* v, double (*g)[3]) const {
        double m[64], gm[64][3];
        monos(x, m, g ? gm : nullptr);
        for (int i = 0; i < n; i++) {
            double s = 0, g0 = 0, g1 = 0, g2 = 0;
            for (int j = 0; j <= i; j++) {
                double c = C[i * n + j];
                s += c * m[j];
                if (g) { g0 += c * gm[j][0]; g1 += c * gm[j][1]; g2 += c * gm[j][2]; }
            }
            v[i] = s;
            if (g) { g[i][0] = g0; g[i][1] = g1; g[i][2] = g2; }
        }
    }
};

RefBasis make_basis(int dim, int p, const Quad& q) {
    RefBasis B;
    B.dim = dim; B.p = p;
    for (int d = 0; d <= p; d++) {
        if (dim == 1) B.e.push_back({d, 0, 0});
        else if (dim == 2)
            for (int a = d; a >= 0; a--) B.e.push_back({a, d - a, 0});
        else
            for (int a = d; a >= 0; a--)
                for (int b = d - a; b >= 0; b--) B.e.push_back({a, b, d - a - b});
    }
    B.n = (int)B.e.size();
    const int n = B.n;
    FH_CHECK(n <= 64, -2, "polynomial degree too high");
    std::vector<double> M(n * n, 0.0), m(n);
    for (size_t k = 0; k < q.p.size(); k++) {
        B.monos(q.p[k], m.data(), nullptr);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) M[i * n + j] += q.w[k] * m[i] * m[j];
    }
    // Cholesky M = L L^T
    std::vector<double> L(n * n, 0.0);
    for (int j = 0; j < n; j++) {
        double d = M[j * n + j];
        for (int k = 0; k < j; k++) d -= L[j * n + k] * L[j * n + k];
        FH_CHECK(d > 0, -2, "reference mass matrix not SPD");
        L[j * n + j] = std::sqrt(d);
        for (int i = j + 1; i < n; i++) {
            double s = M[i * n + j];
            for (int k = 0; k < j; k++) s -= L[i * n + k] * L[j * n + k];
            L[i * n + j] = s / L[j * n + j];
        }
    }
    // C = L^-1 (lower triangular)
    B.C.assign(n * n, 0.0);
    for (int c = 0; c < n; c++) {
        for (int i = c; i < n; i++) {
            double s = i == c ? 1.0 : 0.0;
            for (int k = c; k < i; k++) s -= L[i * n + k] * B.C[k * n + c];
            B.C[i * n + c] = s / L[i * n + i];
        }
    }
    return B;
}

// ------------------------------------------------------------ small dense
// In-place Cholesky of an SPD n x n (row-major, lower part used).
bool chol(int n, double* A) {
    for (int j = 0; j < n; j++) {
        double d = A[j * n + j];
        for (int k = 0; k < j; k++) d -= A[j * n + k] * A[j * n + k];
        if (!(d > 0)) return false;
        d = std::sqrt(d);
        A[j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double s = A[i * n + j];
            for (int k = 0; k < j; k++) s -= A[i * n + k] * A[j * n + k];
            A[i * n + j] = s / d;
        }
    }
    return true;
}
// Solve L L^T X = B for X (n x m, row-major, ldb = m) in place.
void chol_solve(int n, const double* L, double* B, int m) {
    for (int c = 0; c < m; c++) {
        for (int i = 0; i < n; i++) {
            double s = B[i * m + c];
            for (int k = 0; k < i; k++) s -= L[i * n + k] * B[k * m + c];
            B[i * m + c] = s / L[i * n + i];
        }
        for (int i = n - 1; i >= 0; i--) {
            double s = B[i * m + c];
            for (int k = i + 1; k < n; k++) s -= L[k * n + i] * B[k * m + c];
            B[i * m + c] = s / L[i * n + i];
        }
    }
}

inline V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline double dot3(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline double nrm(const V3& a) { return std::sqrt(dot3(a, a)); }
inline V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

// ------------------------------------------------------ perturbed grids
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

inline double det3(const V3& a, const V3& b, const V3& c) { return dot3(a, cross(b, c)); }

// Cube split of Cube_CartesianTetrahedralMesh (corner code x + 2y + 4z).
const int kTetSplit[6][4] = {{2, 3, 7, 1}, {7, 2, 1, 0}, {0, 1, 7, 5}, {7, 2, 0, 6}, {5, 7, 0, 6}, {5, 0, 4, 6}};

// Smallest cell quality in the background box (bx, by, bz) of a grid of
// spacing h: the Jacobian determinant over h^dim, oriented so that the
// unperturbed cell has quality 1.  Simplices (stetra, stri): the affine
// Jacobian of every cell of the box split.  Bilinear / trilinear cells
// (hex, quad): the Jacobian at each corner, from the corner's edges.
double box_quality(const std::vector<V3>& X, int dim, int nv1, const std::string& mesh, int bx, int by, int bz,
                   double h) {
    auto vid = [&](int ix, int iy, int iz) { return ((size_t)iz * nv1 + iy) * nv1 + ix; };
    V3 P[8];
    const int ncorner = dim == 3 ? 8 : 4;
    for (int cc = 0; cc < ncorner; cc++) P[cc] = X[vid(bx + (cc & 1), by + ((cc >> 1) & 1), dim == 3 ? bz + ((cc >> 2) & 1) : 0)];
    double q = 1e300;
    if (mesh == "stetra") {
        static const double* sgn = [] {
            static double s[6];
            for (int t = 0; t < 6; t++) {
                V3 v[4];
                for (int i = 0; i < 4; i++) v[i] = {double(kTetSplit[t][i] & 1), double((kTetSplit[t][i] >> 1) & 1), double((kTetSplit[t][i] >> 2) & 1)};
                s[t] = det3(sub(v[1], v[0]), sub(v[2], v[0]), sub(v[3], v[0])) > 0 ? 1.0 : -1.0;
            }
            return s;
        }();
        const double h3 = h * h * h;
        for (int t = 0; t < 6; t++) {
            const V3& v0 = P[kTetSplit[t][0]];
            q = std::min(q, sgn[t] * det3(sub(P[kTetSplit[t][1]], v0), sub(P[kTetSplit[t][2]], v0), sub(P[kTetSplit[t][3]], v0)) / h3);
        }
    } else if (mesh == "stri") {
        // (v00, v10, v01) and (v01, v10, v11), both counter-clockwise
        auto cr = [](const V3& a, const V3& b, const V3& c) { return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]); };
        q = std::min(cr(P[0], P[1], P[2]), cr(P[2], P[1], P[3])) / (h * h);
    } else if (dim == 3) {
        const double h3 = h * h * h;
        for (int cc = 0; cc < 8; cc++) {
            V3 e[3];
            for (int d = 0; d < 3; d++) {
                e[d] = sub(P[cc ^ (1 << d)], P[cc]);
                if ((cc >> d) & 1) e[d] = {-e[d][0], -e[d][1], -e[d][2]};
            }
            q = std::min(q, det3(e[0], e[1], e[2]) / h3);
        }
    } else {
        for (int cc = 0; cc < 4; cc++) {
            V3 e0 = sub(P[cc ^ 1], P[cc]), e1 = sub(P[cc ^ 2], P[cc]);
            if (cc & 1) e0 = {-e0[0], -e0[1], 0};
            if (cc & 2) e1 = {-e1[0], -e1[1], 0};
            q = std::min(q, (e0[0] * e1[1] - e0[1] * e1[0]) / (h * h));
        }
    }
    return q;
}

// Vertex positions of the (rn + 1)^dim grid of the unit square / cube with
// every interior vertex moved by perturb * h * u, u uniform in [-1, 1]^dim
// drawn from a counter-based hash of (seed, vertex, coordinate, attempt).
// Vertices are placed one by one in index order; a draw is rejected, and
// the next attempt drawn, when a cell of the vertex's boxes (vertices placed
// so far at their new positions, the others at their grid positions) has a
// quality below kQualityFloor.  After kMaxDraws rejected draws the vertex
// stays at its grid position, which passed that test when its neighbours
// were placed.  Every cell is tested with all its vertices final when its
// last vertex is placed, so no cell of the mesh has a quality below the
// floor (in particular none is inverted).
constexpr double kQualityFloor = 0.1;
constexpr int kMaxDraws = 64;
struct PerturbedGrid {
    std::vector<V3> X;
    int64_t redraws = 0, kept = 0;  // rejected draws; vertices left at their grid position
    double min_quality = 1.0;
};
PerturbedGrid perturbed_grid(int dim, int rn, const std::string& mesh, double perturb, uint64_t seed) {
    PerturbedGrid G;
    const int nv1 = rn + 1;
    const double h = 1.0 / rn;
    const int nz = dim == 3 ? nv1 : 1;
    G.X.resize((size_t)nv1 * nv1 * nz);
    for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < nv1; iy++)
            for (int ix = 0; ix < nv1; ix++) G.X[((size_t)iz * nv1 + iy) * nv1 + ix] = {ix * h, iy * h, dim == 3 ? iz * h : 0.0};
    if (!(perturb > 0)) return G;
    auto quality_around = [&](int ix, int iy, int iz) {
        double q = 1e300;
        for (int bz = (dim == 3 ? iz - 1 : 0); bz <= (dim == 3 ? iz : 0); bz++)
            for (int by = iy - 1; by <= iy; by++)
                for (int bx = ix - 1; bx <= ix; bx++) {
                    if (bx < 0 || by < 0 || bz < 0 || bx >= rn || by >= rn || (dim == 3 && bz >= rn)) continue;
                    q = std::min(q, box_quality(G.X, dim, nv1, mesh, bx, by, bz, h));
                }
        return q;
    };
    for (int iz = 1; iz < (dim == 3 ? rn : 2); iz++) {
        const int z = dim == 3 ? iz : 0;
        for (int iy = 1; iy < rn; iy++)
            for (int ix = 1; ix < rn; ix++) {
                const size_t v = ((size_t)z * nv1 + iy) * nv1 + ix;
                const V3 base = G.X[v];
                const uint64_t gid = ((uint64_t)z * nv1 + (uint64_t)iy) * nv1 + (uint64_t)ix;
                bool ok = false;
                for (int a = 0; a < kMaxDraws && !ok; a++) {
                    V3 p = base;
                    for (int d = 0; d < dim; d++) {
                        uint64_t key = seed * 0x100000001B3ull ^ (gid * 4 + d);
                        if (a > 0) key ^= (uint64_t)a * 0xD1B54A32D192ED03ull;
                        const uint64_t r = splitmix64(key);
                        const double u = (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
                        p[d] += perturb * h * u;
                    }
                    G.X[v] = p;
                    ok = quality_around(ix, iy, z) >= kQualityFloor;
                    if (!ok) G.redraws++;
                }
                if (!ok) {
                    G.X[v] = base;
                    G.kept++;
                }
            }
    }
    for (int bz = 0; bz < (dim == 3 ? rn : 1); bz++)
        for (int by = 0; by < rn; by++)
            for (int bx = 0; bx < rn; bx++) G.min_quality = std::min(G.min_quality, box_quality(G.X, dim, nv1, mesh, bx, by, bz, h));
    return G;
}

}  // namespace

// ------------------------------------------------------------------- Impl
struct HhoProblem::Impl {
    HhoArgs a;
    int dim = 3, kind = 0;   // cell kind: 0 simplex, 1 box (affine), 2 bilinear / trilinear quadrilateral or hexahedron
    int64_t redraws = 0, kept_vertices = 0;  // perturbation: rejected draws, vertices left on the grid
    double min_quality = 1.0;                // smallest cell quality (Jacobian determinant / h^dim)
    int nvc = 0, nfc = 0, nvf = 0;
    std::vector<V3> V;
    std::vector<int> cv;     // cell vertices (nvc per cell), local order of the mesher
    std::vector<int> cf;     // cell faces (nfc per cell)
    std::vector<int> fv;     // face vertices (nvf per face; parametrisation uses 0, 1 and 2 for squares: a, b, d)
    std::vector<int> fcells; // 2 per face
    std::vector<int> funk;   // block row of an interior face, -1 for a Dirichlet face
    std::vector<double> K;   // 9 per cell
    int ncell = 0, nface = 0;
    // reference machinery
    int nc = 0, nr = 0, nfb = 0;
    RefBasis rb, fb;         // reconstruction (cell) and face bases
    Quad cq, fq;             // cell and face rules
    std::vector<double> Shat;  // 9 x nr x nr : int d_a psi_i d_b psi_j
    std::vector<double> cellPsi;  // nq x nr: psi at cell quadrature points
    struct FaceTable {
        std::vector<double> Phi;   // 3 x nr x nfb
        std::vector<double> Psi;   // 3 x nr x nc
        std::vector<double> MFT;   // nfb x nr
    };
    std::map<std::array<int, 9>, int> table_index;
    std::vector<FaceTable> tables;
    std::vector<int> cell_face_table;  // nfc per cell

    V3 ref_vertex(int local) const {
        if (kind == 0) {
            V3 r{0, 0, 0};
            if (local > 0) r[local - 1] = 1;
            return r;
        }
        return {double(local & 1), double((local >> 1) & 1), double((local >> 2) & 1)};
    }

    // "modes": u = sum_m a_m prod_d sin(p_md pi x_d), f = -Lap u.
    struct Mode { double a; int p[3]; };
    std::vector<Mode> modes;
    void make_modes() {
        modes.clear();
        if (a.source != "modes") return;
        FH_CHECK(a.source_modes >= 1 && a.source_modes <= 64, -2, "source_modes must be 1..64");
        uint64_t st = a.source_seed * 0x2545F4914F6CDD1Dull + 0x1234567ull;
        auto nxt = [&st]() {
            uint64_t x = (st += 0x9E3779B97F4A7C15ull);
            x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
            x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
            return x ^ (x >> 31);
        };
        for (int m = 0; m < a.source_modes; m++) {
            Mode md;
            md.a = (double)(nxt() >> 11) * (2.0 / 9007199254740992.0) - 1.0;
            for (int d = 0; d < 3; d++) md.p[d] = 1 + (int)(nxt() % 4);
            modes.push_back(md);
        }
    }
    double source(const V3& x) const {
        if (!modes.empty()) {
            double f = 0;
            for (const Mode& md : modes) {
                double s = 1, k2 = 0;
                for (int d = 0; d < dim; d++) { s *= std::sin(md.p[d] * M_PI * x[d]); k2 += md.p[d] * md.p[d]; }
                f += md.a * M_PI * M_PI * k2 * s;
            }
            return f;
        }
        double w = a.source == "sine" ? 4 * M_PI : M_PI;
        double s = 1;
        for (int d = 0; d < dim; d++) s *= std::sin(w * x[d]);
        return dim * w * w * s;
    }
    double exact(const V3& x) const {
        if (!modes.empty()) {
            double u = 0;
            for (const Mode& md : modes) {
                double s = 1;
                for (int d = 0; d < dim; d++) s *= std::sin(md.p[d] * M_PI * x[d]);
                u += md.a * s;
            }
            return u;
        }
        double w = a.source == "sine" ? 4 * M_PI : M_PI;
        double s = 1;
        for (int d = 0; d < dim; d++) s *= std::sin(w * x[d]);
        return s;
    }

    void build_mesh();
    void build_tables();

    // ---- general cells (kind 2): bilinear quadrilaterals / trilinear
    // hexahedra.  Cell and reconstruction polynomials are tensor Legendre
    // products in the physical coordinates of the cell, scaled to its
    // bounding box; face polynomials are tensor Legendre products in the
    // face's own parameters (s, t) in [-1, 1]^2 (bilinear faces), the basis
    // and order of the reference's Cartesian faces (LegendreBasis2D,
    // FunctionalBasis/Legendre/LegendreBasis.h:80-120; face map
    // Geometry/CartesianShape.h:345-421).  Integrals use tensor Gauss rules
    // of k + 2 points per direction through the cell / face maps.
    std::vector<double> g1x, g1w;        // (k+2)-point Gauss rule on [-1, 1]
    std::vector<double> l1x, l1w;        // (k+4)-point rule (L2 error)
    std::vector<I3> rexp;                // reconstruction exponents, graded (first nc: degree <= k)
    std::vector<std::array<int, 2>> fexp;  // face exponents
    void cell_frame(int c, V3& ctr, V3& hw) const;
    void cell_map(int c, const V3& xi, V3& x, double J[3][3], double& det) const;
    void cell_basis(const V3& ctr, const V3& hw, const V3& x, double* psi, double (*g)[3]) const;
    void face_map(int fid, double s, double t, V3& x, V3& xs, V3& xt) const;
    void face_basis(double s, double t, double* phi) const;
    double face_diameter(int fid) const;
    void local_general(int c, double* Acond, double* bcond, const double* xF, double* recon, double* RF,
                       double* Aloc) const;
    double l2_general(int c, const double* recon, double& u2) const;
    void geometry(int c, V3& v0, double J[3][3], double Ji[3][3], double& det) const;
    // Local condensation; outputs (optional): condensed matrix (m nfb)^2 and rhs;
    // or, for post-processing, the reconstruction coefficients from face values;
    // or the reconstruction-from-faces operator RF (nr x m nfb) without source,
    // p = P [-A_TT^-1 A_TF; I] v_F (ReconstructionFromFacesMatrix,
    // Diff_HHOElement.h:212-232).
    // Aloc: the uncondensed local matrix (nc + m nfb)^2, cell unknowns first.
    void local(int c, double* Acond, double* bcond, const double* xF, double* recon, double* RF = nullptr,
               double* Aloc = nullptr) const;
};

// ------------------------------------------------------------------- mesh
void HhoProblem::Impl::build_mesh() {
    const int n = a.n;
    dim = a.dim;
    const double h = 1.0 / n;
    FH_CHECK(n >= 1 && n <= 1000, -2, "mesh size n out of range");
    const int nv1 = n + 1;
    // vertices (x fastest, then y, then z), perturbed inside the domain.
    // The perturbation lives on the finest grid of a hierarchy (ref_n):
    // coarsened meshes reuse the positions of its vertices.
    int nverts = dim == 3 ? nv1 * nv1 * nv1 : nv1 * nv1;
    V.resize(nverts);
    const int rn = a.ref_n > 0 ? a.ref_n : n;
    FH_CHECK(rn % n == 0, -2, "ref_n must be a multiple of n");
    const int ratio = rn / n;
    const std::string grid_kind = (a.mesh == "hex" || a.mesh == "quad" || a.mesh == "cart") ? "box" : a.mesh;
    FH_CHECK(a.perturb >= 0 && a.perturb < 0.5, -2, "perturb must be in [0, 0.5)");
    PerturbedGrid PG = perturbed_grid(dim, rn, grid_kind, a.perturb, a.seed);
    redraws = PG.redraws;
    kept_vertices = PG.kept;
    min_quality = PG.min_quality;
    for (int iz = 0; iz < (dim == 3 ? nv1 : 1); iz++)
        for (int iy = 0; iy < nv1; iy++)
            for (int ix = 0; ix < nv1; ix++)
                V[(iz * nv1 + iy) * nv1 + ix] =
                    PG.X[((size_t)(iz * ratio) * (rn + 1) + (size_t)(iy * ratio)) * (rn + 1) + (size_t)(ix * ratio)];
    PG.X.clear();
    PG.X.shrink_to_fit();
    auto vid = [&](int ix, int iy, int iz) { return (iz * nv1 + iy) * nv1 + ix; };

    std::unordered_map<uint64_t, int> fmap;
    std::vector<int> interior_list;
    fcells.clear();
    fv.clear();
    auto face_key = [](std::vector<int> v) {
        std::sort(v.begin(), v.end());
        uint64_t k = 0;
        for (int x : v) k = k * 2097152ull + (uint64_t)x;
        return k;
    };
    // Creates the face if it does not exist yet (then a boundary face),
    // otherwise marks it interior; returns the face id.
    auto touch = [&](int cell, const std::vector<int>& vs, uint64_t key) {
        auto it = fmap.find(key);
        if (it == fmap.end()) {
            int id = (int)(fcells.size() / 2);
            fmap.emplace(key, id);
            for (int x : vs) fv.push_back(x);
            fcells.push_back(cell);
            fcells.push_back(-1);
            return id;
        }
        int id = it->second;
        fcells[2 * id + 1] = cell;
        interior_list.push_back(id);
        return id;
    };

    if (a.mesh == "stetra") {
        FH_CHECK(dim == 3, -2, "stetra meshes are 3-D");
        kind = 0; nvc = 4; nfc = 4; nvf = 3;
        ncell = 6 * n * n * n;
        cv.resize((size_t)ncell * 4);
        cf.assign((size_t)ncell * 4, -1);
        std::vector<int> nadded(ncell, 0);
        // Cube split and face creation order of Cube_CartesianTetrahedralMesh::Build
        // (src/Mesher/InHouse/Cube_CartesianTetrahedralMesh.h:44-99).
        static const int T[6][4] = {{2, 3, 7, 1}, {7, 2, 1, 0}, {0, 1, 7, 5}, {7, 2, 0, 6}, {5, 7, 0, 6}, {5, 0, 4, 6}};
        // corner code = x + 2y + 4z: v000=0 v100=1 v010=2 v110=3 v001=4 v101=5 v011=6 v111=7
        struct FC { int t0, t1, a, b, c; };
        static const FC FCALLS[18] = {
            {0, -1, 2, 7, 3}, {0, -1, 2, 3, 1}, {0, 1, 2, 1, 7}, {0, -1, 1, 3, 7}, {1, 3, 7, 2, 0}, {1, 2, 7, 0, 1},
            {1, -1, 0, 2, 1}, {2, -1, 0, 1, 5}, {2, 4, 0, 5, 7}, {2, -1, 5, 1, 7}, {3, -1, 7, 2, 6}, {3, 4, 7, 6, 0},
            {3, -1, 6, 2, 0}, {4, -1, 5, 7, 6}, {4, 5, 5, 6, 0}, {5, -1, 5, 4, 0}, {5, -1, 5, 6, 4}, {5, -1, 6, 0, 4}};
        int cube = 0;
        for (int iz = 0; iz < n; iz++)
            for (int iy = 0; iy < n; iy++)
                for (int ix = 0; ix < n; ix++, cube++) {
                    int corner[8];
                    for (int cc = 0; cc < 8; cc++) corner[cc] = vid(ix + (cc & 1), iy + ((cc >> 1) & 1), iz + ((cc >> 2) & 1));
                    for (int t = 0; t < 6; t++)
                        for (int q = 0; q < 4; q++) cv[(size_t)(6 * cube + t) * 4 + q] = corner[T[t][q]];
                    for (const FC& f : FCALLS) {
                        std::vector<int> vs{corner[f.a], corner[f.b], corner[f.c]};
                        int c0 = 6 * cube + f.t0;
                        int id;
                        if (f.t1 < 0) id = touch(c0, vs, face_key(vs));
                        else {
                            id = (int)(fcells.size() / 2);
                            for (int x : vs) fv.push_back(x);
                            fcells.push_back(c0);
                            fcells.push_back(6 * cube + f.t1);
                            interior_list.push_back(id);
                            fmap.emplace(face_key(vs), id);
                        }
                        int cs[2] = {c0, f.t1 < 0 ? -1 : 6 * cube + f.t1};
                        if (f.t1 < 0 && fcells[2 * id] != c0) cs[0] = c0;  // second touch: this cell
                        for (int s = 0; s < 2; s++) {
                            int cell = cs[s];
                            if (cell < 0) continue;
                            FH_CHECK(nadded[cell] < 4, -2, "tetrahedron with more than 4 faces");
                            cf[(size_t)cell * 4 + nadded[cell]++] = id;
                        }
                    }
                }
    } else if (a.mesh == "stri") {
        // Square_TriangularMesh::Build (src/Mesher/InHouse/Square_TriangularMesh.h:121-239):
        // square (ix, iy) -> lower triangle 2q = (v00, v10, v01) and upper
        // triangle 2q+1 = (v01, v10, v11); interior edges square by square:
        // the oblique edge, then the east and the north interface.
        FH_CHECK(dim == 2, -2, "stri meshes are 2-D");
        kind = 0; nvc = 3; nfc = 3; nvf = 2;
        ncell = 2 * n * n;
        cv.resize((size_t)ncell * 3);
        cf.assign((size_t)ncell * 3, -1);
        std::vector<int> nadded(ncell, 0);
        auto lower = [&](int ix, int iy) { return 2 * (iy * n + ix); };
        for (int iy = 0; iy < n; iy++)
            for (int ix = 0; ix < n; ix++) {
                int c = lower(ix, iy);
                int v00 = vid(ix, iy, 0), v10 = vid(ix + 1, iy, 0), v01 = vid(ix, iy + 1, 0), v11 = vid(ix + 1, iy + 1, 0);
                int lv[3] = {v00, v10, v01}, uv[3] = {v01, v10, v11};
                for (int q = 0; q < 3; q++) { cv[(size_t)c * 3 + q] = lv[q]; cv[(size_t)(c + 1) * 3 + q] = uv[q]; }
            }
        auto add_face = [&](int c0, int c1, int va, int vb) {
            int id = (int)(fcells.size() / 2);
            fv.push_back(va); fv.push_back(vb);
            fcells.push_back(c0); fcells.push_back(c1);
            for (int cc : {c0, c1}) {
                if (cc < 0) continue;
                FH_CHECK(nadded[cc] < 3, -2, "triangle with more than 3 faces");
                cf[(size_t)cc * 3 + nadded[cc]++] = id;
            }
            if (c1 >= 0) interior_list.push_back(id);
        };
        for (int ix = 0; ix < n; ix++) {
            add_face(lower(ix, 0), -1, vid(ix, 0, 0), vid(ix + 1, 0, 0));            // south
            add_face(lower(ix, n - 1) + 1, -1, vid(ix, n, 0), vid(ix + 1, n, 0));    // north
        }
        for (int iy = 0; iy < n; iy++) {
            add_face(lower(0, iy), -1, vid(0, iy + 1, 0), vid(0, iy, 0));            // west
            add_face(lower(n - 1, iy) + 1, -1, vid(n, iy, 0), vid(n, iy + 1, 0));    // east
        }
        for (int iy = 0; iy < n; iy++)
            for (int ix = 0; ix < n; ix++) {
                int lo = lower(ix, iy), up = lo + 1;
                add_face(lo, up, vid(ix, iy + 1, 0), vid(ix + 1, iy, 0));            // oblique
                if (ix != n - 1) add_face(up, lower(ix + 1, iy), vid(ix + 1, iy + 1, 0), vid(ix + 1, iy, 0));
                if (iy != n - 1) add_face(up, lower(ix, iy + 1), vid(ix, iy + 1, 0), vid(ix + 1, iy + 1, 0));
            }
    } else if (a.mesh == "cart" || a.mesh == "hex" || a.mesh == "quad") {
        // "cart": axis-aligned boxes (affine, tabulated on the reference box);
        // "hex" / "quad": the same cells and faces with the vertices free to
        // move (trilinear / bilinear maps, general-cell local operators)
        FH_CHECK(a.mesh != "cart" || a.perturb == 0, -2, "perturbed Cartesian cells are not affine; use hex / quad");
        FH_CHECK(a.mesh != "hex" || dim == 3, -2, "hex meshes are 3-D");
        FH_CHECK(a.mesh != "quad" || dim == 2, -2, "quad meshes are 2-D");
        kind = a.mesh == "cart" ? 1 : 2;
        nvc = dim == 3 ? 8 : 4;
        nfc = 2 * dim;
        nvf = dim == 3 ? 4 : 2;
        ncell = dim == 3 ? n * n * n : n * n;
        cv.resize((size_t)ncell * nvc);
        cf.assign((size_t)ncell * nfc, -1);
        int cell = 0;
        for (int iz = 0; iz < (dim == 3 ? n : 1); iz++)
            for (int iy = 0; iy < n; iy++)
                for (int ix = 0; ix < n; ix++, cell++) {
                    for (int cc = 0; cc < nvc; cc++)
                        cv[(size_t)cell * nvc + cc] = vid(ix + (cc & 1), iy + ((cc >> 1) & 1), dim == 3 ? iz + ((cc >> 2) & 1) : 0);
                    // faces: -x, +x, -y, +y (, -z, +z); vertex order a, b, (c), d with
                    // parametrisation a + s (b - a) + t (d - a)
                    for (int f = 0; f < nfc; f++) {
                        int axis = f / 2, side = f % 2;
                        std::vector<int> vs;
                        if (dim == 2) {
                            int o[2] = {0, 0};
                            o[axis] = side;
                            int t1 = 1 - axis;
                            int p0[2] = {o[0], o[1]}, p1[2] = {o[0], o[1]};
                            p1[t1] = 1;
                            vs = {vid(ix + p0[0], iy + p0[1], 0), vid(ix + p1[0], iy + p1[1], 0)};
                        } else {
                            int t1 = (axis + 1) % 3, t2 = (axis + 2) % 3;
                            int lo = std::min(t1, t2), hi = std::max(t1, t2);
                            int pa[3] = {0, 0, 0};
                            pa[axis] = side;
                            int pb[3] = {pa[0], pa[1], pa[2]}, pd[3] = {pa[0], pa[1], pa[2]}, pc[3] = {pa[0], pa[1], pa[2]};
                            pb[lo] = 1; pd[hi] = 1; pc[lo] = 1; pc[hi] = 1;
                            vs = {vid(ix + pa[0], iy + pa[1], iz + pa[2]), vid(ix + pb[0], iy + pb[1], iz + pb[2]),
                                  vid(ix + pc[0], iy + pc[1], iz + pc[2]), vid(ix + pd[0], iy + pd[1], iz + pd[2])};
                        }
                        uint64_t key = (uint64_t)(*std::min_element(vs.begin(), vs.end())) * 3 + axis;
                        cf[(size_t)cell * nfc + f] = touch(cell, vs, key);
                    }
                }
        // Interior faces in the order of Square_CartesianMesh / Cube_CartesianMesh::Build
        // (src/Mesher/InHouse/Square_CartesianMesh.h:173-200, Cube_CartesianMesh.h:170-205):
        // cell by cell, its +x, +y (, +z) interface.
        interior_list.clear();
        for (int c = 0; c < ncell; c++)
            for (int f = 1; f < nfc; f += 2) {
                int id = cf[(size_t)c * nfc + f];
                if (fcells[2 * id + 1] >= 0) interior_list.push_back(id);
            }
    } else
        throw Error(-2, "unknown mesh '" + a.mesh + "' (stetra | stri | cart | hex | quad)");
    nface = (int)(fcells.size() / 2);
    funk.assign(nface, -1);
    for (size_t i = 0; i < interior_list.size(); i++) funk[interior_list[i]] = (int)i;

    // diffusion tensor per cell
    K.assign((size_t)ncell * 9, 0.0);
    std::mt19937_64 krng(a.seed ^ 0x9E3779B97F4A7C15ull);
    std::normal_distribution<double> N01(0.0, 1.0);
    std::vector<double> sub_kappa;
    int nsub = 0;
    if (a.kappa == "subdomain") {
        // a cell lies in one subdomain, or (on coarse levels of a hierarchy)
        // covers whole subdomains, whose tensors it then averages
        FH_CHECK(a.subdomain % ratio == 0 || ratio % std::max(1, a.subdomain) == 0, -2,
                 "subdomains must be unions of cells or of whole subdomains");
        int ns = std::max(1, rn / std::max(1, a.subdomain));
        nsub = ns;
        sub_kappa.resize((size_t)ns * ns * (dim == 3 ? ns : 1));
        for (auto& s : sub_kappa) s = std::exp(a.sigma * N01(krng));
    }
    // Per-cell lognormal field.  The draws belong to the cells of the finest
    // mesh (ref_n): a coarse level of a hierarchy gives each cell the mean
    // coefficient of the fine cells in its background box.
    std::vector<double> box_kappa;
    const int cpb = kind == 0 ? (dim == 3 ? 6 : 2) : 1;  // cells per background box
    if (a.kappa == "lognormal" && ratio > 1) {
        const size_t nbox = (size_t)n * n * (dim == 3 ? n : 1);
        box_kappa.assign(nbox, 0.0);
        const size_t nfine = (size_t)rn * rn * (dim == 3 ? rn : 1) * cpb;
        for (size_t f = 0; f < nfine; f++) {
            const double kf = std::exp(a.sigma * N01(krng));
            const size_t fb = f / cpb;
            const size_t fx = fb % rn, fy = (fb / rn) % rn, fz = dim == 3 ? fb / ((size_t)rn * rn) : 0;
            box_kappa[((fz / ratio) * n + fy / ratio) * n + fx / ratio] += kf;
        }
        const double per_box = (double)nfine / (double)nbox;
        for (auto& v : box_kappa) v /= per_box;
    }
    for (int c = 0; c < ncell; c++) {
        double kap = 1.0, kz = 1.0;
        if (a.kappa == "lognormal") kap = box_kappa.empty() ? std::exp(a.sigma * N01(krng)) : box_kappa[c / cpb];
        else if (a.kappa == "subdomain") {
            // background cell of the cell (tets: 6 per cube)
            int bc = kind == 0 ? c / (dim == 3 ? 6 : 2) : c;
            int ix = (bc % n) * ratio, iy = ((bc / n) % n) * ratio, iz = (dim == 3 ? bc / (n * n) : 0) * ratio;
            int sd = std::max(1, a.subdomain);
            // subdomains covered by the cell (one unless the cell is coarser)
            const int span = std::max(1, ratio / sd);
            double kxy = 0, kzz = 0;
            int cnt = 0;
            for (int dz = 0; dz < (dim == 3 ? span : 1); dz++)
                for (int dy = 0; dy < span; dy++)
                    for (int dx = 0; dx < span; dx++) {
                        int sx = std::min(ix / sd + dx, nsub - 1), sy = std::min(iy / sd + dy, nsub - 1),
                            sz = std::min(iz / sd + dz, nsub - 1);
                        int s = (dim == 3 ? sz * nsub * nsub : 0) + sy * nsub + sx;
                        const double ks = sub_kappa[s];
                        kxy += ks;
                        kzz += s % 8 == 7 ? ks * a.aniso : ks;  // one subdomain in eight is anisotropic
                        cnt++;
                    }
            if (cnt == 1) {
                const int s0 = (dim == 3 ? std::min(iz / sd, nsub - 1) * nsub * nsub : 0) +
                               std::min(iy / sd, nsub - 1) * nsub + std::min(ix / sd, nsub - 1);
                kap = sub_kappa[s0];
                kz = s0 % 8 == 7 ? a.aniso : 1.0;
            } else {
                kap = kxy / cnt;
                kz = kzz / cnt / kap;
            }
        } else if (a.kappa != "one")
            throw Error(-2, "unknown kappa field '" + a.kappa + "'");
        K[(size_t)c * 9 + 0] = kap;
        K[(size_t)c * 9 + 4] = kap;
        K[(size_t)c * 9 + 8] = dim == 3 ? kap * kz : 1.0;
        if (dim == 2) K[(size_t)c * 9 + 8] = 1.0;
    }
}

void HhoProblem::Impl::geometry(int c, V3& v0, double J[3][3], double Ji[3][3], double& det) const {
    const int* cvv = &cv[(size_t)c * nvc];
    v0 = V[cvv[0]];
    int idx[3];
    if (kind == 0) { idx[0] = 1; idx[1] = 2; idx[2] = 3; }
    else { idx[0] = 1; idx[1] = 2; idx[2] = 4; }
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) J[r][q] = 0;
    for (int q = 0; q < dim; q++) {
        V3 e = sub(V[cvv[idx[q]]], v0);
        for (int r = 0; r < 3; r++) J[r][q] = e[r];
    }
    if (dim == 2) J[2][2] = 1;
    det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    double id = 1.0 / det;
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}

void HhoProblem::Impl::build_tables() {
    const int k = a.k;
    FH_CHECK(k >= 0 && k <= 3, -2, "face degree k must be 0..3");
    if (kind == 2) {
        std::vector<double> x, w;
        gauss_legendre01(k + 2, x, w);
        g1x.resize(x.size()); g1w.resize(x.size());
        for (size_t i = 0; i < x.size(); i++) { g1x[i] = 2 * x[i] - 1; g1w[i] = 2 * w[i]; }
        gauss_legendre01(k + 4, x, w);
        l1x.resize(x.size()); l1w.resize(x.size());
        for (size_t i = 0; i < x.size(); i++) { l1x[i] = 2 * x[i] - 1; l1w[i] = 2 * w[i]; }
        rexp.clear();
        for (int d = 0; d <= k + 1; d++) {
            if (dim == 2)
                for (int aa = d; aa >= 0; aa--) rexp.push_back({aa, d - aa, 0});
            else
                for (int aa = d; aa >= 0; aa--)
                    for (int bb = d - aa; bb >= 0; bb--) rexp.push_back({aa, bb, d - aa - bb});
        }
        nr = (int)rexp.size();
        nc = 0;
        for (auto& e : rexp) if (e[0] + e[1] + e[2] <= k) nc++;
        fexp.clear();
        for (int d = 0; d <= k; d++) {
            if (dim == 2) fexp.push_back({d, 0});
            else
                for (int j = 0; j <= d; j++) fexp.push_back({d - j, j});
        }
        nfb = (int)fexp.size();
        return;
    }
    cq = ref_quad(dim, kind, k + 4);
    fq = ref_quad(dim - 1, kind, k + 3);
    rb = make_basis(dim, k + 1, cq);
    fb = make_basis(dim - 1, k, fq);
    nr = rb.n;
    nfb = fb.n;
    nc = 0;
    for (auto& e : rb.e) if (e[0] + e[1] + e[2] <= k) nc++;
    // stiffness tensors on the reference shape
    Shat.assign((size_t)9 * nr * nr, 0.0);
    cellPsi.assign(cq.p.size() * nr, 0.0);
    std::vector<double> v(nr);
    std::vector<double> g(nr * 3);
    for (size_t q = 0; q < cq.p.size(); q++) {
        rb.eval(cq.p[q], v.data(), (double(*)[3])g.data());
        for (int i = 0; i < nr; i++) cellPsi[q * nr + i] = v[i];
        for (int ax = 0; ax < 3; ax++)
            for (int bx = 0; bx < 3; bx++)
                for (int i = 0; i < nr; i++)
                    for (int j = 0; j < nr; j++)
                        Shat[((size_t)(ax * 3 + bx) * nr + i) * nr + j] += cq.w[q] * g[i * 3 + ax] * g[j * 3 + bx];
    }
    // face tables keyed by the reference position of the face parametrisation
    cell_face_table.assign((size_t)ncell * nfc, -1);
    for (int c = 0; c < ncell; c++) {
        const int* cvv = &cv[(size_t)c * nvc];
        for (int f = 0; f < nfc; f++) {
            int fid = cf[(size_t)c * nfc + f];
            const int* fvv = &fv[(size_t)fid * nvf];
            // a, b, and the second axis vertex (d for squares, c for triangles)
            int pv[3] = {fvv[0], dim >= 2 ? fvv[1] : -1, dim == 3 ? (kind == 1 ? fvv[3] : fvv[2]) : -1};
            std::array<int, 9> key{};
            for (int s = 0; s < dim; s++) {
                int loc = -1;
                for (int l = 0; l < nvc; l++) if (cvv[l] == pv[s]) loc = l;
                FH_CHECK(loc >= 0, -2, "face vertex not in cell");
                V3 r = ref_vertex(loc);
                for (int d = 0; d < 3; d++) key[s * 3 + d] = (int)r[d];
            }
            auto it = table_index.find(key);
            int ti;
            if (it == table_index.end()) {
                ti = (int)tables.size();
                table_index.emplace(key, ti);
                FaceTable T;
                T.Phi.assign((size_t)3 * nr * nfb, 0.0);
                T.Psi.assign((size_t)3 * nr * nc, 0.0);
                T.MFT.assign((size_t)nfb * nr, 0.0);
                V3 ra{double(key[0]), double(key[1]), double(key[2])};
                V3 e1{0, 0, 0}, e2{0, 0, 0};
                if (dim >= 2) e1 = sub(V3{double(key[3]), double(key[4]), double(key[5])}, ra);
                if (dim == 3) e2 = sub(V3{double(key[6]), double(key[7]), double(key[8])}, ra);
                std::vector<double> chi(nfb);
                for (size_t q = 0; q < fq.p.size(); q++) {
                    const V3& eta = fq.p[q];
                    V3 xi{ra[0] + eta[0] * e1[0] + eta[1] * e2[0], ra[1] + eta[0] * e1[1] + eta[1] * e2[1],
                          ra[2] + eta[0] * e1[2] + eta[1] * e2[2]};
                    rb.eval(xi, v.data(), (double(*)[3])g.data());
                    fb.eval(eta, chi.data(), nullptr);
                    double w = fq.w[q];
                    for (int ax = 0; ax < 3; ax++)
                        for (int i = 0; i < nr; i++) {
                            double gi = w * g[i * 3 + ax];
                            for (int j = 0; j < nfb; j++) T.Phi[((size_t)ax * nr + i) * nfb + j] += gi * chi[j];
                            for (int j = 0; j < nc; j++) T.Psi[((size_t)ax * nr + i) * nc + j] += gi * v[j];
                        }
                    for (int j = 0; j < nfb; j++)
                        for (int i = 0; i < nr; i++) T.MFT[(size_t)j * nr + i] += w * chi[j] * v[i];
                }
                tables.push_back(std::move(T));
            } else
                ti = it->second;
            cell_face_table[(size_t)c * nfc + f] = ti;
        }
    }
}

// ------------------------------------------- general (bi/trilinear) cells
namespace {
// Legendre polynomials L_0..L_p at x and their derivatives: Bonnet's
// recurrence and L'_n = L'_{n-2} + (2n - 1) L_{n-1}.
inline void legendre_all(int p, double x, double* L, double* dL) {
    L[0] = 1;
    dL[0] = 0;
    if (p >= 1) { L[1] = x; dL[1] = 1; }
    for (int q = 2; q <= p; q++) {
        L[q] = ((2 * q - 1) * x * L[q - 1] - (q - 1) * L[q - 2]) / q;
        dL[q] = dL[q - 2] + (2 * q - 1) * L[q - 1];
    }
}
}  // namespace

// Centre (vertex mean) and half-widths of the bounding box of cell c.
void HhoProblem::Impl::cell_frame(int c, V3& ctr, V3& hw) const {
    V3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    ctr = {0, 0, 0};
    for (int l = 0; l < nvc; l++) {
        const V3& p = V[cv[(size_t)c * nvc + l]];
        for (int d = 0; d < 3; d++) {
            ctr[d] += p[d] / nvc;
            lo[d] = std::min(lo[d], p[d]);
            hi[d] = std::max(hi[d], p[d]);
        }
    }
    for (int d = 0; d < 3; d++) hw[d] = d < dim ? 0.5 * (hi[d] - lo[d]) : 1.0;
}

// Bilinear / trilinear map of cell c at xi in [-1, 1]^dim (corner code
// x + 2y + 4z), its Jacobian J[r][q] = dx_r / dxi_q and determinant.
void HhoProblem::Impl::cell_map(int c, const V3& xi, V3& x, double J[3][3], double& det) const {
    x = {0, 0, 0};
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) J[r][q] = 0;
    const int ncorner = dim == 3 ? 8 : 4;
    for (int cc = 0; cc < ncorner; cc++) {
        double f[3], df[3];
        for (int d = 0; d < 3; d++) {
            const double sg = ((cc >> d) & 1) ? 1.0 : -1.0;
            f[d] = d < dim ? 0.5 * (1 + sg * xi[d]) : 1.0;
            df[d] = d < dim ? 0.5 * sg : 0.0;
        }
        const V3& P = V[cv[(size_t)c * nvc + cc]];
        const double N = f[0] * f[1] * f[2];
        const double dN[3] = {df[0] * f[1] * f[2], f[0] * df[1] * f[2], f[0] * f[1] * df[2]};
        for (int r = 0; r < 3; r++) {
            x[r] += N * P[r];
            for (int q = 0; q < dim; q++) J[r][q] += dN[q] * P[r];
        }
    }
    if (dim == 2) J[2][2] = 1;
    det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// Reconstruction basis (the first nc functions: the cell basis) at the
// physical point x: products of Legendre polynomials of (x - ctr) / hw.
void HhoProblem::Impl::cell_basis(const V3& ctr, const V3& hw, const V3& x, double* psi, double (*g)[3]) const {
    const int p = a.k + 1;
    double L[3][8], dL[3][8];
    for (int d = 0; d < 3; d++) {
        if (d < dim) legendre_all(p, (x[d] - ctr[d]) / hw[d], L[d], dL[d]);
        else { L[d][0] = 1; dL[d][0] = 0; }
    }
    for (int i = 0; i < nr; i++) {
        const I3& e = rexp[i];
        psi[i] = L[0][e[0]] * L[1][e[1]] * L[2][e[2]];
        if (g) {
            g[i][0] = dL[0][e[0]] / hw[0] * L[1][e[1]] * L[2][e[2]];
            g[i][1] = dim >= 2 ? L[0][e[0]] * dL[1][e[1]] / hw[1] * L[2][e[2]] : 0.0;
            g[i][2] = dim == 3 ? L[0][e[0]] * L[1][e[1]] * dL[2][e[2]] / hw[2] : 0.0;
        }
    }
}

// Face map: segment a + (s + 1)/2 (b - a) (2-D), or the bilinear patch of the
// corners a (-,-), b (+,-), c (+,+), d (-,+) (3-D); xs, xt its tangents.
void HhoProblem::Impl::face_map(int fid, double s, double t, V3& x, V3& xs, V3& xt) const {
    const int* f = &fv[(size_t)fid * nvf];
    if (dim == 2) {
        const V3 &A = V[f[0]], &B = V[f[1]];
        for (int d = 0; d < 3; d++) {
            x[d] = 0.5 * (1 - s) * A[d] + 0.5 * (1 + s) * B[d];
            xs[d] = 0.5 * (B[d] - A[d]);
            xt[d] = 0;
        }
        return;
    }
    const V3 &A = V[f[0]], &B = V[f[1]], &Cc = V[f[2]], &D = V[f[3]];
    for (int d = 0; d < 3; d++) {
        x[d] = 0.25 * ((1 - s) * (1 - t) * A[d] + (1 + s) * (1 - t) * B[d] + (1 + s) * (1 + t) * Cc[d] + (1 - s) * (1 + t) * D[d]);
        xs[d] = 0.25 * ((1 - t) * (B[d] - A[d]) + (1 + t) * (Cc[d] - D[d]));
        xt[d] = 0.25 * ((1 - s) * (D[d] - A[d]) + (1 + s) * (Cc[d] - B[d]));
    }
}

void HhoProblem::Impl::face_basis(double s, double t, double* phi) const {
    double Ls[8], dLs[8], Lt[8], dLt[8];
    legendre_all(a.k, s, Ls, dLs);
    legendre_all(a.k, t, Lt, dLt);
    for (int j = 0; j < nfb; j++) phi[j] = Ls[fexp[j][0]] * (dim == 3 ? Lt[fexp[j][1]] : 1.0);
}

// Longest edge of the face (the side of a square face: CartesianShape's
// diameter is its largest width, Geometry/CartesianShape.h:140).
double HhoProblem::Impl::face_diameter(int fid) const {
    const int* f = &fv[(size_t)fid * nvf];
    if (dim == 2) return nrm(sub(V[f[1]], V[f[0]]));
    double h = 0;
    for (int e = 0; e < 4; e++) h = std::max(h, nrm(sub(V[f[(e + 1) % 4]], V[f[e]])));
    return h;
}

// Local HHO operators of a bilinear / trilinear cell (Diff_HHOElement::InitHHO,
// Diff_HHOElement.h:104-150): the saddle-point reconstruction with the
// mean-value condition (:313-402), consistency P^T S P (:323), the "hho"
// stabilisation with the general projections (:411-451: Dt = M_T^-1 N_T P - I,
// face trace projections Mf^-1 int phi psi), and the static condensation.
// Face integrals use the vector area element x_s x x_t (outward), so the
// integration by parts of the reconstruction is exact on warped faces; the
// stabilisation weight (K n.n)/h_F takes the unit normal at the face centre.
// Outputs as local().
void HhoProblem::Impl::local_general(int c, double* Acond, double* bcond, const double* xF, double* recon, double* RF,
                                     double* Aloc) const {
    const int m = nfc, nh = nc + m * nfb, nF = m * nfb;
    const int nq1 = (int)g1x.size();
    const double* Kc = &K[(size_t)c * 9];
    V3 ctr, hw;
    cell_frame(c, ctr, hw);
    thread_local std::vector<double> S, Lm, MT, BT, B, P, A, Dt, Mf, TF, PF, DF, tmp, W, X;
    S.assign((size_t)nr * nr, 0.0);
    Lm.assign(nr, 0.0);
    MT.assign((size_t)nc * nr, 0.0);
    BT.assign(nc, 0.0);
    B.assign((size_t)nr * nh, 0.0);
    std::vector<double> psi(nr), g(3 * nr), phi(nfb);
    // ---- cell integrals
    const int nqc = dim == 3 ? nq1 * nq1 * nq1 : nq1 * nq1;
    for (int q = 0; q < nqc; q++) {
        const int i0 = q % nq1, i1 = (q / nq1) % nq1, i2 = dim == 3 ? q / (nq1 * nq1) : 0;
        V3 xi{g1x[i0], g1x[i1], dim == 3 ? g1x[i2] : 0.0};
        const double w = g1w[i0] * g1w[i1] * (dim == 3 ? g1w[i2] : 1.0);
        V3 x;
        double J[3][3], det;
        cell_map(c, xi, x, J, det);
        FH_CHECK(det > 0, -6, "cell " + std::to_string(c) + " is inverted at a quadrature point");
        const double wd = w * det;
        cell_basis(ctr, hw, x, psi.data(), (double(*)[3])g.data());
        for (int i = 0; i < nr; i++) {
            const double* gi = &g[3 * i];
            double Kg[3];
            for (int r = 0; r < 3; r++) Kg[r] = Kc[r * 3] * gi[0] + Kc[r * 3 + 1] * gi[1] + Kc[r * 3 + 2] * gi[2];
            for (int j = 0; j < nr; j++) {
                const double* gj = &g[3 * j];
                S[(size_t)i * nr + j] += wd * (Kg[0] * gj[0] + Kg[1] * gj[1] + Kg[2] * gj[2]);
            }
            Lm[i] += wd * psi[i];
        }
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nr; j++) MT[(size_t)i * nr + j] += wd * psi[i] * psi[j];
        if (bcond || recon) {
            const double fw = wd * source(x);
            for (int i = 0; i < nc; i++) BT[i] += fw * psi[i];
        }
    }
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < nc; j++) B[(size_t)i * nh + j] = S[(size_t)i * nr + j];
    // ---- face integrals
    Mf.assign((size_t)m * nfb * nfb, 0.0);
    TF.assign((size_t)m * nfb * nr, 0.0);
    std::vector<double> coef(m);
    const int nqf = dim == 3 ? nq1 * nq1 : nq1;
    for (int f = 0; f < m; f++) {
        const int fid = cf[(size_t)c * nfc + f];
        V3 x0, xs, xt;
        face_map(fid, 0, 0, x0, xs, xt);
        V3 nc0 = dim == 3 ? cross(xs, xt) : V3{xs[1], -xs[0], 0};
        const double sg = dot3(nc0, sub(x0, ctr)) < 0 ? -1.0 : 1.0;  // outward
        const double l0 = nrm(nc0);
        const V3 nu{sg * nc0[0] / l0, sg * nc0[1] / l0, sg * nc0[2] / l0};
        double Kn = 0;
        for (int r = 0; r < 3; r++)
            for (int q = 0; q < 3; q++) Kn += nu[r] * Kc[r * 3 + q] * nu[q];
        coef[f] = Kn / face_diameter(fid);
        double* Mff = &Mf[(size_t)f * nfb * nfb];
        double* TFf = &TF[(size_t)f * nfb * nr];
        for (int q = 0; q < nqf; q++) {
            const int i0 = q % nq1, i1 = dim == 3 ? q / nq1 : 0;
            const double st = g1x[i0], tt = dim == 3 ? g1x[i1] : 0.0;
            const double w = g1w[i0] * (dim == 3 ? g1w[i1] : 1.0);
            V3 x, s1, t1;
            face_map(fid, st, tt, x, s1, t1);
            V3 nv = dim == 3 ? cross(s1, t1) : V3{s1[1], -s1[0], 0};
            for (int d = 0; d < 3; d++) nv[d] *= sg;
            const double da = nrm(nv);
            cell_basis(ctr, hw, x, psi.data(), (double(*)[3])g.data());
            face_basis(st, tt, phi.data());
            for (int i = 0; i < nr; i++) {
                const double* gi = &g[3 * i];
                double Kgn = 0;
                for (int r = 0; r < 3; r++) Kgn += (Kc[r * 3] * gi[0] + Kc[r * 3 + 1] * gi[1] + Kc[r * 3 + 2] * gi[2]) * nv[r];
                const double wk = w * Kgn;
                for (int j = 0; j < nfb; j++) B[(size_t)i * nh + nc + f * nfb + j] += wk * phi[j];
                for (int j = 0; j < nc; j++) B[(size_t)i * nh + j] -= wk * psi[j];
            }
            for (int aa = 0; aa < nfb; aa++) {
                const double wa = w * da * phi[aa];
                for (int bb = 0; bb < nfb; bb++) Mff[aa * nfb + bb] += wa * phi[bb];
                for (int i = 0; i < nr; i++) TFf[(size_t)aa * nr + i] += wa * psi[i];
            }
        }
    }
    for (int j = 0; j < nh; j++) B[j] = 0.0;  // degree-0 row: no equation (Diff_HHOElement.h:366-367)
    // ---- reconstruction: rows 1.. solve S' P' = B'; row 0 from the mean
    // value  sum_i (int psi_i) P_i = [int psi_j (cell), 0 (faces)]
    P.assign((size_t)nr * nh, 0.0);
    const int n1 = nr - 1;
    tmp.assign((size_t)n1 * n1, 0.0);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n1; j++) tmp[(size_t)i * n1 + j] = S[(size_t)(i + 1) * nr + j + 1];
    FH_CHECK(n1 == 0 || chol(n1, tmp.data()), -6, "reconstruction stiffness not SPD");
    for (int i = 1; i < nr; i++)
        for (int j = 0; j < nh; j++) P[(size_t)i * nh + j] = B[(size_t)i * nh + j];
    if (n1) chol_solve(n1, tmp.data(), &P[nh], nh);
    for (int j = 0; j < nh; j++) {
        double v = j < nc ? Lm[j] : 0.0;
        for (int i = 1; i < nr; i++) v -= Lm[i] * P[(size_t)i * nh + j];
        P[j] = v / Lm[0];
    }
    // ---- consistency P^T S P
    W.assign((size_t)nr * nh, 0.0);
    for (int i = 0; i < nr; i++)
        for (int l = 0; l < nr; l++) {
            const double sv = S[(size_t)i * nr + l];
            if (sv == 0.0) continue;
            for (int j = 0; j < nh; j++) W[(size_t)i * nh + j] += sv * P[(size_t)l * nh + j];
        }
    A.assign((size_t)nh * nh, 0.0);
    for (int i = 0; i < nr; i++)
        for (int r = 0; r < nh; r++) {
            const double pv = P[(size_t)i * nh + r];
            if (pv == 0.0) continue;
            for (int j = 0; j < nh; j++) A[(size_t)r * nh + j] += pv * W[(size_t)i * nh + j];
        }
    // ---- Dt = M_T^-1 N_T P - [I 0]   (N_T = MT, M_T = MT[:, :nc])
    std::vector<double> MTc((size_t)nc * nc);
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) MTc[(size_t)i * nc + j] = MT[(size_t)i * nr + j];
    FH_CHECK(chol(nc, MTc.data()), -6, "cell mass matrix not SPD");
    Dt.assign((size_t)nc * nh, 0.0);
    for (int i = 0; i < nc; i++)
        for (int l = 0; l < nr; l++) {
            const double mv = MT[(size_t)i * nr + l];
            for (int j = 0; j < nh; j++) Dt[(size_t)i * nh + j] += mv * P[(size_t)l * nh + j];
        }
    chol_solve(nc, MTc.data(), Dt.data(), nh);
    for (int i = 0; i < nc; i++) Dt[(size_t)i * nh + i] -= 1.0;
    // ---- stabilisation, face by face
    PF.assign((size_t)nfb * nr, 0.0);
    DF.assign((size_t)nfb * nh, 0.0);
    std::vector<double> Lf((size_t)nfb * nfb), MD((size_t)nfb * nh);
    for (int f = 0; f < m; f++) {
        const double* Mff = &Mf[(size_t)f * nfb * nfb];
        std::copy(Mff, Mff + (size_t)nfb * nfb, Lf.begin());
        FH_CHECK(chol(nfb, Lf.data()), -6, "face mass matrix not SPD");
        std::copy(&TF[(size_t)f * nfb * nr], &TF[(size_t)(f + 1) * nfb * nr], PF.begin());
        chol_solve(nfb, Lf.data(), PF.data(), nr);  // ProjF = Mf^-1 int phi psi
        std::fill(DF.begin(), DF.end(), 0.0);
        for (int aa = 0; aa < nfb; aa++) {
            for (int i = 0; i < nr; i++) {
                const double pv = PF[(size_t)aa * nr + i];
                for (int j = 0; j < nh; j++) DF[(size_t)aa * nh + j] += pv * P[(size_t)i * nh + j];
            }
            DF[(size_t)aa * nh + nc + f * nfb + aa] -= 1.0;
            for (int i = 0; i < nc; i++) {
                const double pv = PF[(size_t)aa * nr + i];
                for (int j = 0; j < nh; j++) DF[(size_t)aa * nh + j] -= pv * Dt[(size_t)i * nh + j];
            }
        }
        // A += coef DF^T Mf DF
        std::fill(MD.begin(), MD.end(), 0.0);
        for (int aa = 0; aa < nfb; aa++)
            for (int bb = 0; bb < nfb; bb++) {
                const double mv = Mff[aa * nfb + bb];
                for (int j = 0; j < nh; j++) MD[(size_t)aa * nh + j] += mv * DF[(size_t)bb * nh + j];
            }
        for (int aa = 0; aa < nfb; aa++)
            for (int r = 0; r < nh; r++) {
                const double d = coef[f] * DF[(size_t)aa * nh + r];
                if (d == 0.0) continue;
                for (int j = 0; j < nh; j++) A[(size_t)r * nh + j] += d * MD[(size_t)aa * nh + j];
            }
    }
    if (Aloc) std::copy(A.begin(), A.end(), Aloc);
    // ---- static condensation (Diffusion_HHO.h:624-652, HHOCellSpace.h:170-203)
    if (RF && !bcond && !recon) std::fill(BT.begin(), BT.end(), 0.0);
    tmp.assign((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) tmp[(size_t)i * nc + j] = A[(size_t)i * nh + j];
    FH_CHECK(chol(nc, tmp.data()), -6, "A_TT not SPD");
    if (Acond || bcond) {
        X.assign((size_t)nc * (nF + 1), 0.0);
        for (int i = 0; i < nc; i++) {
            for (int j = 0; j < nF; j++) X[(size_t)i * (nF + 1) + j] = A[(size_t)i * nh + nc + j];
            X[(size_t)i * (nF + 1) + nF] = BT[i];
        }
        chol_solve(nc, tmp.data(), X.data(), nF + 1);
        if (Acond)
            for (int r = 0; r < nF; r++)
                for (int s2 = 0; s2 < nF; s2++) {
                    double v = A[(size_t)(nc + r) * nh + nc + s2];
                    for (int i = 0; i < nc; i++) v -= A[(size_t)i * nh + nc + r] * X[(size_t)i * (nF + 1) + s2];
                    Acond[(size_t)r * nF + s2] = v;
                }
        if (bcond)
            for (int r = 0; r < nF; r++) {
                double v = 0;
                for (int i = 0; i < nc; i++) v -= A[(size_t)i * nh + nc + r] * X[(size_t)i * (nF + 1) + nF];
                bcond[r] = v;
            }
    }
    if (RF) {
        X.assign((size_t)nc * nF, 0.0);
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nF; j++) X[(size_t)i * nF + j] = A[(size_t)i * nh + nc + j];
        chol_solve(nc, tmp.data(), X.data(), nF);
        for (int i = 0; i < nr; i++)
            for (int j = 0; j < nF; j++) {
                double v = P[(size_t)i * nh + nc + j];
                for (int l = 0; l < nc; l++) v -= P[(size_t)i * nh + l] * X[(size_t)l * nF + j];
                RF[(size_t)i * nF + j] = v;
            }
    }
    if (recon) {
        std::vector<double> vT(nc);
        for (int i = 0; i < nc; i++) {
            double v = BT[i];
            for (int j = 0; j < nF; j++) v -= A[(size_t)i * nh + nc + j] * xF[j];
            vT[i] = v;
        }
        chol_solve(nc, tmp.data(), vT.data(), 1);
        for (int i = 0; i < nr; i++) {
            double v = 0;
            for (int j = 0; j < nc; j++) v += P[(size_t)i * nh + j] * vT[j];
            for (int j = 0; j < nF; j++) v += P[(size_t)i * nh + nc + j] * xF[j];
            recon[i] = v;
        }
    }
}

// Integrals of (u - p)^2 (returned) and u^2 (u2) over cell c, p the
// reconstruction with coefficients recon, on the (k+4)-point rule.
double HhoProblem::Impl::l2_general(int c, const double* recon, double& u2) const {
    V3 ctr, hw;
    cell_frame(c, ctr, hw);
    std::vector<double> psi(nr);
    const int n1 = (int)l1x.size();
    const int nq = dim == 3 ? n1 * n1 * n1 : n1 * n1;
    double e2 = 0;
    u2 = 0;
    for (int q = 0; q < nq; q++) {
        const int i0 = q % n1, i1 = (q / n1) % n1, i2 = dim == 3 ? q / (n1 * n1) : 0;
        V3 xi{l1x[i0], l1x[i1], dim == 3 ? l1x[i2] : 0.0};
        const double w = l1w[i0] * l1w[i1] * (dim == 3 ? l1w[i2] : 1.0);
        V3 x;
        double J[3][3], det;
        cell_map(c, xi, x, J, det);
        cell_basis(ctr, hw, x, psi.data(), nullptr);
        double ph = 0;
        for (int i = 0; i < nr; i++) ph += recon[i] * psi[i];
        const double u = exact(x);
        e2 += w * det * (u - ph) * (u - ph);
        u2 += w * det * u * u;
    }
    return e2;
}

// Local HHO operators of cell c (Diff_HHOElement::InitHHO,
// Diff_HHOElement.h:104-150, 312-509) and their static condensation.
void HhoProblem::Impl::local(int c, double* Acond, double* bcond, const double* xF, double* recon, double* RF,
                             double* Aloc) const {
    if (kind == 2) {
        local_general(c, Acond, bcond, xF, recon, RF, Aloc);
        return;
    }
    const int m = nfc, nh = nc + m * nfb;
    V3 v0;
    double J[3][3], Ji[3][3], det;
    geometry(c, v0, J, Ji, det);
    const double adet = std::fabs(det);
    const double* Kc = &K[(size_t)c * 9];
    // G = J^-1 K J^-T
    double G[3][3];
    for (int r = 0; r < 3; r++)
        for (int q = 0; q < 3; q++) {
            double s = 0;
            for (int u = 0; u < 3; u++)
                for (int w = 0; w < 3; w++) s += Ji[r][u] * Kc[u * 3 + w] * Ji[q][w];
            G[r][q] = s;
        }
    thread_local std::vector<double> S, B, P, A, Dt, ProjF, DF, tmp, W;
    S.assign(nr * nr, 0.0);
    for (int ax = 0; ax < dim; ax++)
        for (int bx = 0; bx < dim; bx++) {
            const double gab = G[ax][bx];
            const double* Sh = &Shat[(size_t)(ax * 3 + bx) * nr * nr];
            for (int e = 0; e < nr * nr; e++) S[e] += gab * Sh[e];
        }
    B.assign((size_t)nr * nh, 0.0);
    for (int i = 0; i < nr; i++)
        for (int j = 0; j < nc; j++) B[i * nh + j] = S[i * nr + j];
    ProjF.assign((size_t)m * nfb * nr, 0.0);
    std::vector<double> coef(m);
    // centroid of the cell
    V3 cc{0, 0, 0};
    for (int l = 0; l < nvc; l++)
        for (int d = 0; d < 3; d++) cc[d] += V[cv[(size_t)c * nvc + l]][d] / nvc;
    for (int f = 0; f < m; f++) {
        int fid = cf[(size_t)c * nfc + f];
        const int* fvv = &fv[(size_t)fid * nvf];
        V3 a0 = V[fvv[0]];
        V3 E1 = sub(V[fvv[1]], a0);
        V3 E2 = dim == 3 ? sub(V[kind == 1 ? fvv[3] : fvv[2]], a0) : V3{0, 0, 0};
        double sF, hF;
        V3 nF;
        if (dim == 3) {
            V3 cr = cross(E1, E2);
            sF = nrm(cr);
            nF = {cr[0] / sF, cr[1] / sF, cr[2] / sF};
            if (kind == 0) hF = std::max({nrm(E1), nrm(E2), nrm(sub(V[fvv[2]], V[fvv[1]]))});
            else hF = std::max(nrm(E1), nrm(E2));
        } else {
            sF = nrm(E1);
            hF = sF;
            nF = {E1[1] / sF, -E1[0] / sF, 0};
        }
        V3 fc = a0;
        for (int l = 1; l < nvf; l++)
            for (int d = 0; d < 3; d++) fc[d] += V[fvv[l]][d];
        for (int d = 0; d < 3; d++) fc[d] /= nvf;
        if (dot3(nF, sub(fc, cc)) < 0) nF = {-nF[0], -nF[1], -nF[2]};
        // t = J^-1 K n
        V3 Kn{Kc[0] * nF[0] + Kc[1] * nF[1] + Kc[2] * nF[2], Kc[3] * nF[0] + Kc[4] * nF[1] + Kc[5] * nF[2],
              Kc[6] * nF[0] + Kc[7] * nF[1] + Kc[8] * nF[2]};
        V3 t{Ji[0][0] * Kn[0] + Ji[0][1] * Kn[1] + Ji[0][2] * Kn[2], Ji[1][0] * Kn[0] + Ji[1][1] * Kn[1] + Ji[1][2] * Kn[2],
             Ji[2][0] * Kn[0] + Ji[2][1] * Kn[1] + Ji[2][2] * Kn[2]};
        coef[f] = dot3(Kn, nF) / hF;
        const FaceTable& T = tables[cell_face_table[(size_t)c * nfc + f]];
        const double s1 = std::sqrt(sF / adet), s2 = sF / adet;
        for (int ax = 0; ax < dim; ax++) {
            const double* Phi = &T.Phi[(size_t)ax * nr * nfb];
            const double* Psi = &T.Psi[(size_t)ax * nr * nc];
            const double ta = t[ax];
            for (int i = 0; i < nr; i++) {
                for (int j = 0; j < nfb; j++) B[i * nh + nc + f * nfb + j] += s1 * ta * Phi[i * nfb + j];
                for (int j = 0; j < nc; j++) B[i * nh + j] -= s2 * ta * Psi[i * nc + j];
            }
        }
        for (int e = 0; e < nfb * nr; e++) ProjF[(size_t)f * nfb * nr + e] = s1 * T.MFT[e];
    }
    // Reconstruction: row 0 from the mean-value condition, rows 1.. from S'
    P.assign((size_t)nr * nh, 0.0);
    P[0] = 1.0;
    {
        const int n1 = nr - 1;
        tmp.assign((size_t)n1 * n1, 0.0);
        for (int i = 0; i < n1; i++)
            for (int j = 0; j < n1; j++) tmp[i * n1 + j] = S[(i + 1) * nr + j + 1];
        FH_CHECK(n1 == 0 || chol(n1, tmp.data()), -6, "reconstruction stiffness not SPD");
        for (int i = 0; i < n1; i++)
            for (int j = 0; j < nh; j++) P[(i + 1) * nh + j] = B[(i + 1) * nh + j];
        if (n1) chol_solve(n1, tmp.data(), &P[nh], nh);
    }
    // A = P^T S P + stabilisation
    W.assign((size_t)nr * nh, 0.0);  // S P
    for (int i = 0; i < nr; i++)
        for (int l = 0; l < nr; l++) {
            double s = S[i * nr + l];
            if (s == 0.0) continue;
            for (int j = 0; j < nh; j++) W[i * nh + j] += s * P[l * nh + j];
        }
    A.assign((size_t)nh * nh, 0.0);
    for (int i = 0; i < nr; i++)
        for (int r = 0; r < nh; r++) {
            double p = P[i * nh + r];
            if (p == 0.0) continue;
            for (int j = 0; j < nh; j++) A[r * nh + j] += p * W[i * nh + j];
        }
    Dt.assign((size_t)nc * nh, 0.0);
    for (int i = 0; i < nc; i++) {
        for (int j = 0; j < nh; j++) Dt[i * nh + j] = P[i * nh + j];
        Dt[i * nh + i] -= 1.0;
    }
    DF.assign((size_t)nfb * nh, 0.0);
    for (int f = 0; f < m; f++) {
        const double* PF = &ProjF[(size_t)f * nfb * nr];
        std::fill(DF.begin(), DF.end(), 0.0);
        for (int j = 0; j < nfb; j++) {
            for (int i = 0; i < nr; i++) {
                double p = PF[j * nr + i];
                for (int r = 0; r < nh; r++) DF[j * nh + r] += p * P[i * nh + r];
            }
            DF[j * nh + nc + f * nfb + j] -= 1.0;
            for (int i = 0; i < nc; i++) {
                double p = PF[j * nr + i];
                for (int r = 0; r < nh; r++) DF[j * nh + r] -= p * Dt[i * nh + r];
            }
        }
        const double cf_ = coef[f];
        for (int j = 0; j < nfb; j++)
            for (int r = 0; r < nh; r++) {
                double d = cf_ * DF[j * nh + r];
                if (d == 0.0) continue;
                for (int s = 0; s < nh; s++) A[r * nh + s] += d * DF[j * nh + s];
            }
    }
    if (Aloc) std::copy(A.begin(), A.end(), Aloc);
    // right-hand side on the cell: B_T = int f psi_i
    std::vector<double> BT(nc, 0.0);
    if (bcond || recon) {
        const double sq = std::sqrt(adet);
        for (size_t q = 0; q < cq.p.size(); q++) {
            const V3& xi = cq.p[q];
            V3 x{v0[0] + J[0][0] * xi[0] + J[0][1] * xi[1] + J[0][2] * xi[2],
                 v0[1] + J[1][0] * xi[0] + J[1][1] * xi[1] + J[1][2] * xi[2],
                 v0[2] + J[2][0] * xi[0] + J[2][1] * xi[1] + J[2][2] * xi[2]};
            double fw = cq.w[q] * source(x) * sq;
            for (int i = 0; i < nc; i++) BT[i] += fw * cellPsi[q * nr + i];
        }
    }
    // Static condensation (Diffusion_HHO.h:624-652, HHOCellSpace.h:170-203)
    const int nF = m * nfb;
    if (RF && !bcond && !recon) BT.assign(nc, 0.0);
    tmp.assign((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) tmp[i * nc + j] = A[i * nh + j];
    FH_CHECK(chol(nc, tmp.data()), -6, "A_TT not SPD");
    if (Acond || bcond) {
        std::vector<double> X((size_t)nc * (nF + 1));
        for (int i = 0; i < nc; i++) {
            for (int j = 0; j < nF; j++) X[i * (nF + 1) + j] = A[i * nh + nc + j];
            X[i * (nF + 1) + nF] = BT[i];
        }
        chol_solve(nc, tmp.data(), X.data(), nF + 1);
        if (Acond)
            for (int r = 0; r < nF; r++)
                for (int s = 0; s < nF; s++) {
                    double v = A[(nc + r) * nh + nc + s];
                    for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + s];
                    Acond[r * nF + s] = v;
                }
        if (bcond)
            for (int r = 0; r < nF; r++) {
                double v = 0;
                for (int i = 0; i < nc; i++) v -= A[i * nh + nc + r] * X[i * (nF + 1) + nF];
                bcond[r] = v;
            }
    }
    if (RF) {
        std::vector<double> X((size_t)nc * nF);
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < nF; j++) X[i * nF + j] = A[i * nh + nc + j];
        chol_solve(nc, tmp.data(), X.data(), nF);
        for (int i = 0; i < nr; i++)
            for (int j = 0; j < nF; j++) {
                double s = P[i * nh + nc + j];
                for (int l = 0; l < nc; l++) s -= P[i * nh + l] * X[l * nF + j];
                RF[i * nF + j] = s;
            }
    }
    if (recon) {
        // cell unknowns v_T = A_TT^-1 (B_T - A_TF v_F), then p = P [v_T; v_F]
        std::vector<double> vT(nc);
        for (int i = 0; i < nc; i++) {
            double s = BT[i];
            for (int j = 0; j < nF; j++) s -= A[i * nh + nc + j] * xF[j];
            vT[i] = s;
        }
        chol_solve(nc, tmp.data(), vT.data(), 1);
        for (int i = 0; i < nr; i++) {
            double s = 0;
            for (int j = 0; j < nc; j++) s += P[i * nh + j] * vT[j];
            for (int j = 0; j < nF; j++) s += P[i * nh + nc + j] * xF[j];
            recon[i] = s;
        }
    }
}

// ----------------------------------------------------------------- public
HhoProblem::HhoProblem(const HhoArgs& args) : impl(new Impl) {
    Impl& I = *impl;
    I.a = args;
    FH_CHECK(args.source == "pisine" || args.source == "sine" || args.source == "modes", -2,
             "unknown source '" + args.source + "' (pisine | sine | modes)");
    I.make_modes();
    auto t0 = std::chrono::steady_clock::now();
    I.build_mesh();
    I.build_tables();
    auto t1 = std::chrono::steady_clock::now();
    times.mesh_s = std::chrono::duration<double>(t1 - t0).count();
    nf = I.nfb;
    nb = 0;
    for (int f = 0; f < I.nface; f++) if (I.funk[f] >= 0) nb++;
    // block pattern: faces sharing a cell
    std::vector<std::vector<int>> rows(nb);
    for (int c = 0; c < I.ncell; c++)
        for (int f = 0; f < I.nfc; f++) {
            int r = I.funk[I.cf[(size_t)c * I.nfc + f]];
            if (r < 0) continue;
            for (int g = 0; g < I.nfc; g++) {
                int s = I.funk[I.cf[(size_t)c * I.nfc + g]];
                if (s >= 0) rows[r].push_back(s);
            }
        }
    rowptr.assign(nb + 1, 0);
    for (int r = 0; r < nb; r++) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        rowptr[r + 1] = rowptr[r] + (int)v.size();
    }
    colidx.resize(rowptr[nb]);
    for (int r = 0; r < nb; r++) std::copy(rows[r].begin(), rows[r].end(), colidx.begin() + rowptr[r]);
    rows.clear();
    rows.shrink_to_fit();
    const size_t bsz = (size_t)nf * nf;
    if (!args.host_assembly) {
        // values come from AssembleDevice
        auto t3 = std::chrono::steady_clock::now();
        times.assemble_s = std::chrono::duration<double>(t3 - t1).count();
        return;
    }
    host_assembled = true;
    vals.assign((size_t)rowptr[nb] * bsz, 0.0);
    b.assign((size_t)nb * nf, 0.0);
    auto t2 = std::chrono::steady_clock::now();
    const int nF = I.nfc * nf;
    // Each off-diagonal block receives one cell contribution and each
    // diagonal block two, so atomic accumulation is order independent.
#pragma omp parallel
    {
        std::vector<double> Ac((size_t)nF * nF), bc(nF);
        std::vector<int> rr(I.nfc), pos((size_t)I.nfc * I.nfc);
#pragma omp for schedule(dynamic, 256)
        for (int c = 0; c < I.ncell; c++) {
            I.local(c, Ac.data(), bc.data(), nullptr, nullptr);
            for (int f = 0; f < I.nfc; f++) rr[f] = I.funk[I.cf[(size_t)c * I.nfc + f]];
            for (int f = 0; f < I.nfc; f++) {
                if (rr[f] < 0) continue;
                for (int g = 0; g < I.nfc; g++) {
                    if (rr[g] < 0) continue;
                    const int* beg = &colidx[rowptr[rr[f]]];
                    const int* end = &colidx[rowptr[rr[f] + 1]];
                    int p = (int)(std::lower_bound(beg, end, rr[g]) - &colidx[0]);
                    double* dst = &vals[(size_t)p * bsz];
                    for (int a = 0; a < nf; a++)
                        for (int bb = 0; bb < nf; bb++) {
                            double v = Ac[(size_t)(f * nf + a) * nF + g * nf + bb];
#pragma omp atomic
                            dst[a * nf + bb] += v;
                        }
                }
                for (int a = 0; a < nf; a++) {
#pragma omp atomic
                    b[(size_t)rr[f] * nf + a] += bc[f * nf + a];
                }
            }
        }
    }
    auto t3 = std::chrono::steady_clock::now();
    times.local_s = std::chrono::duration<double>(t3 - t2).count();
    times.assemble_s = std::chrono::duration<double>(t3 - t1).count();
}

void HhoProblem::Uncondensed(Csr& ATT, Csr& ATF, Csr& AFF) const {
    const Impl& I = *impl;
    const int nc = I.nc, nfb = I.nfb, m = I.nfc, nh = nc + m * nfb;
    // Diffusion_HHO::Assemble (Diffusion_HHO.h:410-470): A_T_T blocks, A_T_F
    // blocks on non-Dirichlet faces (A_T_ndF) and A_F_F face pairs of every
    // cell (A_ndF_ndF), through coefficient lists that drop |v| <= 1e-15
    // (NonZeroCoefficients.h:36-40) and sum duplicates.
    struct Tri {
        int i, j;
        double v;
    };
    std::vector<std::vector<Tri>> tt(I.ncell), tf(I.ncell), ff(I.ncell);
#pragma omp parallel
    {
        std::vector<double> Al((size_t)nh * nh);
#pragma omp for schedule(dynamic, 256)
        for (int c = 0; c < I.ncell; c++) {
            I.local(c, nullptr, nullptr, nullptr, nullptr, nullptr, Al.data());
            auto add = [](std::vector<Tri>& t, int i, int j, double v) {
                if (std::fabs(v) > 1e-15) t.push_back({i, j, v});
            };
            for (int a = 0; a < nc; a++)
                for (int b2 = 0; b2 < nc; b2++) add(tt[c], c * nc + a, c * nc + b2, Al[(size_t)a * nh + b2]);
            for (int f = 0; f < m; f++) {
                const int rf = I.funk[I.cf[(size_t)c * m + f]];
                if (rf < 0) continue;
                for (int a = 0; a < nc; a++)
                    for (int b2 = 0; b2 < nfb; b2++)
                        add(tf[c], c * nc + a, rf * nfb + b2, Al[(size_t)a * nh + nc + f * nfb + b2]);
                for (int g = 0; g < m; g++) {
                    const int rg = I.funk[I.cf[(size_t)c * m + g]];
                    if (rg < 0) continue;
                    for (int a = 0; a < nfb; a++)
                        for (int b2 = 0; b2 < nfb; b2++)
                            add(ff[c], rf * nfb + a, rg * nfb + b2,
                                Al[(size_t)(nc + f * nfb + a) * nh + nc + g * nfb + b2]);
                }
            }
        }
    }
    auto fill = [](const std::vector<std::vector<Tri>>& parts, int nr, int ncol, Csr& M) {
        std::vector<std::vector<std::pair<int, double>>> rows(nr);
        for (const auto& p : parts)
            for (const Tri& t : p) rows[t.i].push_back({t.j, t.v});
        M = Csr();
        M.nr = nr;
        M.nc = ncol;
        M.rp.assign(nr + 1, 0);
        for (int i = 0; i < nr; i++) {
            auto& r = rows[i];
            std::stable_sort(r.begin(), r.end(), [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
                return x.first < y.first;
            });
            for (size_t k = 0; k < r.size(); k++) {
                if (k > 0 && r[k].first == M.ci.back())
                    M.v.back() += r[k].second;
                else {
                    M.ci.push_back(r[k].first);
                    M.v.push_back(r[k].second);
                }
            }
            M.rp[i + 1] = (int)M.ci.size();
        }
    };
    fill(tt, I.ncell * nc, I.ncell * nc, ATT);
    fill(tf, I.ncell * nc, nb * nfb, ATF);
    fill(ff, nb * nfb, nb * nfb, AFF);
}

int HhoProblem::cell_unknowns() const { return impl->nc; }

HhoProblem::~HhoProblem() { delete impl; }

void HhoProblem::Prolongation(const HhoProblem& coarse, std::vector<int>& prp, std::vector<int>& pci,
                              std::vector<double>& pv) const {
    const Impl& Fm = *impl;
    const Impl& Cm = *coarse.impl;
    FH_CHECK(Fm.dim == Cm.dim && Fm.kind == Cm.kind && Fm.a.k == Cm.a.k && Fm.a.mesh == Cm.a.mesh, -2,
             "prolongation between problems of different kinds");
    FH_CHECK(Cm.a.n * 2 == Fm.a.n, -2, "the coarse problem must be on the once-coarsened mesh (n/2)");
    const int dim = Fm.dim, n = Fm.a.n, nc_ = Cm.a.n;
    const int nfb = Fm.nfb, nr = Cm.nr, nfc = Cm.nfc, nF = nfc * nfb;
    // fine cell -> coarse cell
    std::vector<int> f2c(Fm.ncell);
    int tet_table[8][6] = {};
    if (Fm.kind == 0 && dim == 2) {
        // Nested triangles of the square split: lower (v00, v10, v01) lies
        // below the anti-diagonal x + y = 1 of its square, upper above it; a
        // fine triangle belongs to the coarse triangle containing its centroid.
        for (int o = 0; o < 4; o++)
            for (int t = 0; t < 2; t++) {
                // centroid of the fine triangle in the coarse square [0,2]^2
                double gx = (o & 1) + (t == 0 ? 1.0 / 3.0 : 2.0 / 3.0);
                double gy = ((o >> 1) & 1) + (t == 0 ? 1.0 / 3.0 : 2.0 / 3.0);
                tet_table[o][t] = gx + gy < 2.0 ? 0 : 1;
            }
    } else if (Fm.kind == 0) {
        // Nested tetrahedra of the cube split: the unperturbed centroid of a
        // fine tetrahedron lies inside exactly one tetrahedron of its coarse cube.
        static const int T[6][4] = {{2, 3, 7, 1}, {7, 2, 1, 0}, {0, 1, 7, 5}, {7, 2, 0, 6}, {5, 7, 0, 6}, {5, 0, 4, 6}};
        auto corner = [](int cc, double s) { return V3{s * (cc & 1), s * ((cc >> 1) & 1), s * ((cc >> 2) & 1)}; };
        for (int o = 0; o < 8; o++)
            for (int t = 0; t < 6; t++) {
                V3 g{0, 0, 0};
                for (int q = 0; q < 4; q++) {
                    V3 p = corner(T[t][q], 1.0);
                    for (int d = 0; d < 3; d++) g[d] += (p[d] + ((o >> d) & 1)) / 4.0;
                }
                int found = -1;
                for (int tc = 0; tc < 6 && found < 0; tc++) {
                    V3 v0 = corner(T[tc][0], 2.0), e1 = sub(corner(T[tc][1], 2.0), v0), e2 = sub(corner(T[tc][2], 2.0), v0),
                       e3 = sub(corner(T[tc][3], 2.0), v0), r = sub(g, v0);
                    double det = dot3(e1, cross(e2, e3));
                    double l1 = dot3(r, cross(e2, e3)) / det, l2 = dot3(e1, cross(r, e3)) / det, l3 = dot3(e1, cross(e2, r)) / det;
                    if (l1 > 1e-9 && l2 > 1e-9 && l3 > 1e-9 && l1 + l2 + l3 < 1 - 1e-9) found = tc;
                }
                FH_CHECK(found >= 0, -2, "non-nested tetrahedral refinement");
                tet_table[o][t] = found;
            }
    }
    for (int c = 0; c < Fm.ncell; c++) {
        const int per = Fm.kind == 0 ? (dim == 3 ? 6 : 2) : 1;
        int bc = c / per;
        int ix = bc % n, iy = (bc / n) % n, iz = dim == 3 ? bc / (n * n) : 0;
        int cb = ((dim == 3 ? (iz / 2) * nc_ : 0) + iy / 2) * nc_ + ix / 2;
        if (Fm.kind == 0) {
            int o = (ix & 1) + 2 * (iy & 1) + 4 * (iz & 1);
            f2c[c] = per * cb + tet_table[o][c % per];
        } else
            f2c[c] = cb;
    }
    // reconstruction-from-faces operator of every coarse cell
    std::vector<double> RFall((size_t)Cm.ncell * nr * nF);
#pragma omp parallel for schedule(dynamic, 256)
    for (int c = 0; c < Cm.ncell; c++) Cm.local(c, nullptr, nullptr, nullptr, nullptr, &RFall[(size_t)c * nr * nF]);
    // rows: fine interior faces
    std::vector<std::vector<int>> rcols(nb);
    std::vector<std::vector<double>> rvals(nb);
    const size_t bsz = (size_t)nfb * nfb;
#pragma omp parallel
    {
        std::vector<double> psi(nr), chi(nfb), Tr((size_t)nfb * nr);
        std::vector<std::pair<int, int>> ent;  // (coarse row, local face of the coarse cell)
#pragma omp for schedule(dynamic, 512)
        for (int f = 0; f < Fm.nface; f++) {
            int row = Fm.funk[f];
            if (row < 0) continue;
            int c1 = Fm.fcells[2 * f], c2 = Fm.fcells[2 * f + 1];
            int C1 = f2c[c1], C2 = c2 >= 0 ? f2c[c2] : -1;
            int cs[2] = {C1, C2 != C1 ? C2 : -1};
            // fine face geometry
            const int* fvv = &Fm.fv[(size_t)f * Fm.nvf];
            V3 a0 = Fm.V[fvv[0]];
            V3 E1 = sub(Fm.V[fvv[1]], a0);
            V3 E2 = dim == 3 ? sub(Fm.V[Fm.kind >= 1 ? fvv[3] : fvv[2]], a0) : V3{0, 0, 0};
            double sF = dim == 3 ? nrm(cross(E1, E2)) : nrm(E1);
            V3 nF_ = dim == 3 ? cross(E1, E2) : V3{E1[1], -E1[0], 0};
            if (Fm.kind == 2) {  // normal at the face centre
                V3 x0, s1, t1;
                Fm.face_map(f, 0, 0, x0, s1, t1);
                nF_ = dim == 3 ? cross(s1, t1) : V3{s1[1], -s1[0], 0};
            }
            double nn = nrm(nF_);
            for (int d = 0; d < 3; d++) nF_[d] /= nn;
            // weights: normal diffusivity (Weight, LevelForHHO.h:743-763)
            double w[2] = {1.0, 0.0};
            if (cs[1] >= 0) {
                double k1 = 0, k2 = 0;
                for (int r = 0; r < 3; r++)
                    for (int q = 0; q < 3; q++) {
                        k1 += nF_[r] * Cm.K[(size_t)cs[0] * 9 + r * 3 + q] * nF_[q];
                        k2 += nF_[r] * Cm.K[(size_t)cs[1] * 9 + r * 3 + q] * nF_[q];
                    }
                w[0] = k1 / (k1 + k2);
                w[1] = k2 / (k1 + k2);
            }
            std::vector<int>& cols = rcols[row];
            std::vector<double>& vals_ = rvals[row];
            for (int s = 0; s < 2; s++) {
                int C = cs[s];
                if (C < 0) continue;
                // trace projection of the coarse reconstruction basis on the fine face
                if (Fm.kind == 2) {
                    // general cells: Tr = Mf^-1 int_F phi psi_C (fine face basis,
                    // coarse cell's physical basis)
                    V3 ctr, hw;
                    Cm.cell_frame(C, ctr, hw);
                    const int nq1 = (int)Fm.g1x.size(), nqf = dim == 3 ? nq1 * nq1 : nq1;
                    std::vector<double> Mf((size_t)nfb * nfb, 0.0);
                    std::fill(Tr.begin(), Tr.end(), 0.0);
                    for (int q = 0; q < nqf; q++) {
                        const int i0 = q % nq1, i1 = dim == 3 ? q / nq1 : 0;
                        const double st = Fm.g1x[i0], tt = dim == 3 ? Fm.g1x[i1] : 0.0;
                        const double wq = Fm.g1w[i0] * (dim == 3 ? Fm.g1w[i1] : 1.0);
                        V3 x, s1, t1;
                        Fm.face_map(f, st, tt, x, s1, t1);
                        const double da = dim == 3 ? nrm(cross(s1, t1)) : nrm(s1);
                        Cm.cell_basis(ctr, hw, x, psi.data(), nullptr);
                        Fm.face_basis(st, tt, chi.data());
                        for (int j = 0; j < nfb; j++) {
                            const double wj = wq * da * chi[j];
                            for (int i = 0; i < nr; i++) Tr[(size_t)j * nr + i] += wj * psi[i];
                            for (int i = 0; i < nfb; i++) Mf[(size_t)j * nfb + i] += wj * chi[i];
                        }
                    }
                    FH_CHECK(chol(nfb, Mf.data()), -6, "face mass matrix not SPD");
                    chol_solve(nfb, Mf.data(), Tr.data(), nr);
                } else {
                V3 v0;
                double J[3][3], Ji[3][3], det;
                Cm.geometry(C, v0, J, Ji, det);
                const double isq = 1.0 / std::sqrt(std::fabs(det)), ssq = std::sqrt(sF);
                std::fill(Tr.begin(), Tr.end(), 0.0);
                for (size_t q = 0; q < Fm.fq.p.size(); q++) {
                    const V3& eta = Fm.fq.p[q];
                    V3 y{a0[0] + eta[0] * E1[0] + eta[1] * E2[0], a0[1] + eta[0] * E1[1] + eta[1] * E2[1],
                         a0[2] + eta[0] * E1[2] + eta[1] * E2[2]};
                    V3 r = sub(y, v0);
                    V3 xi{Ji[0][0] * r[0] + Ji[0][1] * r[1] + Ji[0][2] * r[2], Ji[1][0] * r[0] + Ji[1][1] * r[1] + Ji[1][2] * r[2],
                          Ji[2][0] * r[0] + Ji[2][1] * r[1] + Ji[2][2] * r[2]};
                    if (dim == 2) xi[2] = 0;
                    Cm.rb.eval(xi, psi.data(), nullptr);
                    Fm.fb.eval(eta, chi.data(), nullptr);
                    double wq = Fm.fq.w[q] * ssq * isq;
                    for (int j = 0; j < nfb; j++)
                        for (int i = 0; i < nr; i++) Tr[(size_t)j * nr + i] += wq * chi[j] * psi[i];
                }
                }
                const double* RF = &RFall[(size_t)C * nr * nF];
                for (int lf = 0; lf < nfc; lf++) {
                    int crow = Cm.funk[Cm.cf[(size_t)C * nfc + lf]];
                    if (crow < 0) continue;  // Dirichlet coarse face: zero value
                    int pos = -1;
                    for (size_t t = 0; t < cols.size(); t++) if (cols[t] == crow) pos = (int)t;
                    if (pos < 0) {
                        pos = (int)cols.size();
                        cols.push_back(crow);
                        vals_.resize(vals_.size() + bsz, 0.0);
                    }
                    double* blk = &vals_[(size_t)pos * bsz];
                    for (int j = 0; j < nfb; j++)
                        for (int b2 = 0; b2 < nfb; b2++) {
                            double sum = 0;
                            for (int i = 0; i < nr; i++) sum += Tr[(size_t)j * nr + i] * RF[(size_t)i * nF + lf * nfb + b2];
                            blk[j * nfb + b2] += w[s] * sum;
                        }
                }
            }
        }
    }
    prp.assign(nb + 1, 0);
    for (int r = 0; r < nb; r++) prp[r + 1] = prp[r] + (int)rcols[r].size();
    pci.resize(prp[nb]);
    pv.assign((size_t)prp[nb] * bsz, 0.0);
    for (int r = 0; r < nb; r++) {
        std::vector<int> perm(rcols[r].size());
        for (size_t t = 0; t < perm.size(); t++) perm[t] = (int)t;
        std::sort(perm.begin(), perm.end(), [&](int x, int y) { return rcols[r][x] < rcols[r][y]; });
        for (size_t t = 0; t < perm.size(); t++) {
            pci[prp[r] + t] = rcols[r][perm[t]];
            std::memcpy(&pv[(size_t)(prp[r] + t) * bsz], &rvals[r][(size_t)perm[t] * bsz], bsz * sizeof(double));
        }
    }
}

int HhoProblem::ncells() const { return impl->ncell; }
int64_t HhoProblem::redraws() const { return impl->redraws; }
int64_t HhoProblem::kept_vertices() const { return impl->kept_vertices; }
double HhoProblem::min_quality() const { return impl->min_quality; }
int HhoProblem::nfaces() const { return impl->nface; }

std::vector<double> HhoProblem::CellKappa() const {
    std::vector<double> k(impl->ncell);
    for (int c = 0; c < impl->ncell; c++) k[c] = impl->K[(size_t)c * 9];
    return k;
}

void HhoProblem::FaceCells(std::vector<int>& out) const {
    const Impl& I = *impl;
    out.assign((size_t)nb * 2, -1);
    for (int f = 0; f < I.nface; f++) {
        int r = I.funk[f];
        if (r < 0) continue;
        out[2 * r] = I.fcells[2 * f];
        out[2 * r + 1] = I.fcells[2 * f + 1];
    }
}

double HhoProblem::L2Error(const double* x) const {
    const Impl& I = *impl;
    double err = 0, nrm2 = 0;
#pragma omp parallel reduction(+ : err, nrm2)
    {
        std::vector<double> xF((size_t)I.nfc * nf), rec(I.nr), psi(I.nr);
#pragma omp for schedule(dynamic, 256)
        for (int c = 0; c < I.ncell; c++) {
            for (int f = 0; f < I.nfc; f++) {
                int r = I.funk[I.cf[(size_t)c * I.nfc + f]];
                for (int a = 0; a < nf; a++) xF[f * nf + a] = r >= 0 ? x[(size_t)r * nf + a] : 0.0;
            }
            I.local(c, nullptr, nullptr, xF.data(), rec.data());
            if (I.kind == 2) {
                double u2 = 0;
                err += I.l2_general(c, rec.data(), u2);
                nrm2 += u2;
                continue;
            }
            V3 v0;
            double J[3][3], Ji[3][3], det;
            I.geometry(c, v0, J, Ji, det);
            const double adet = std::fabs(det), isq = 1.0 / std::sqrt(adet);
            for (size_t q = 0; q < I.cq.p.size(); q++) {
                const V3& xi = I.cq.p[q];
                V3 xx{v0[0] + J[0][0] * xi[0] + J[0][1] * xi[1] + J[0][2] * xi[2],
                      v0[1] + J[1][0] * xi[0] + J[1][1] * xi[1] + J[1][2] * xi[2],
                      v0[2] + J[2][0] * xi[0] + J[2][1] * xi[1] + J[2][2] * xi[2]};
                double ph = 0;
                for (int i = 0; i < I.nr; i++) ph += rec[i] * I.cellPsi[q * I.nr + i];
                ph *= isq;
                double u = I.exact(xx);
                double w = I.cq.w[q] * adet;
                err += w * (u - ph) * (u - ph);
                nrm2 += w * u * u;
            }
        }
    }
    return nrm2 > 0 ? std::sqrt(err / nrm2) : std::sqrt(err);
}

// Device description of the problem (geometry, tables and per-cell maps
// uploaded; buffers owned by `keep`).
static HhoDev make_device_problem(Ctx& c, const HhoProblem& H, const HhoProblem::Impl& I, std::vector<DevBuf>& keep) {
    const int m = I.nfc;
    // per-cell maps: block of every (face, face) pair and block row of every face
    std::vector<int> blk((size_t)I.ncell * m * m, -1), frow((size_t)I.ncell * m, -1);
#pragma omp parallel for schedule(static)
    for (int cc = 0; cc < I.ncell; cc++) {
        int rr[8];
        for (int f = 0; f < m; f++) rr[f] = frow[(size_t)cc * m + f] = I.funk[I.cf[(size_t)cc * m + f]];
        for (int f = 0; f < m; f++) {
            if (rr[f] < 0) continue;
            const int* beg = &H.colidx[H.rowptr[rr[f]]];
            const int* end = &H.colidx[H.rowptr[rr[f] + 1]];
            for (int g = 0; g < m; g++)
                if (rr[g] >= 0)
                    blk[((size_t)cc * m + f) * m + g] = (int)(std::lower_bound(beg, end, rr[g]) - &H.colidx[0]);
        }
    }
    std::vector<double> Vf(3 * I.V.size()), cqp(3 * I.cq.p.size()), modes;
    for (size_t v = 0; v < I.V.size(); v++)
        for (int a = 0; a < 3; a++) Vf[3 * v + a] = I.V[v][a];
    for (size_t q = 0; q < I.cq.p.size(); q++)
        for (int a = 0; a < 3; a++) cqp[3 * q + a] = I.cq.p[q][a];
    for (const auto& md : I.modes) modes.insert(modes.end(), {md.a, (double)md.p[0], (double)md.p[1], (double)md.p[2]});
    std::vector<double> Phi, Psi, MFT;
    for (const auto& T : I.tables) {
        Phi.insert(Phi.end(), T.Phi.begin(), T.Phi.end());
        Psi.insert(Psi.end(), T.Psi.begin(), T.Psi.end());
        MFT.insert(MFT.end(), T.MFT.begin(), T.MFT.end());
    }
    auto up = [&](const void* h, size_t bytes) {
        keep.emplace_back(std::max<size_t>(bytes, 8));
        if (bytes) FH_HIP(hipMemcpyAsync(keep.back().p, h, bytes, hipMemcpyHostToDevice, c.stream));
        return keep.back().p;
    };
    HhoDev d{};
    d.dim = I.dim; d.kind = I.kind; d.nvc = I.nvc; d.nfc = m; d.nvf = I.nvf;
    d.nc = I.nc; d.nr = I.nr; d.nfb = I.nfb; d.nq = (int)I.cq.p.size(); d.ncell = I.ncell;
    d.V = (const double*)up(Vf.data(), sizeof(double) * Vf.size());
    d.cv = (const int*)up(I.cv.data(), sizeof(int) * I.cv.size());
    d.cf = (const int*)up(I.cf.data(), sizeof(int) * I.cf.size());
    d.fv = (const int*)up(I.fv.data(), sizeof(int) * I.fv.size());
    d.K = (const double*)up(I.K.data(), sizeof(double) * I.K.size());
    d.cft = (const int*)up(I.cell_face_table.data(), sizeof(int) * I.cell_face_table.size());
    d.Shat = (const double*)up(I.Shat.data(), sizeof(double) * I.Shat.size());
    d.cellPsi = (const double*)up(I.cellPsi.data(), sizeof(double) * I.cellPsi.size());
    d.cqp = (const double*)up(cqp.data(), sizeof(double) * cqp.size());
    d.cqw = (const double*)up(I.cq.w.data(), sizeof(double) * I.cq.w.size());
    d.Phi = (const double*)up(Phi.data(), sizeof(double) * Phi.size());
    d.Psi = (const double*)up(Psi.data(), sizeof(double) * Psi.size());
    d.MFT = (const double*)up(MFT.data(), sizeof(double) * MFT.size());
    d.src = I.a.source == "modes" ? 2 : (I.a.source == "sine" ? 1 : 0);
    d.nmodes = (int)I.modes.size();
    d.modes = (const double*)up(modes.data(), sizeof(double) * modes.size());
    d.blk = (const int*)up(blk.data(), sizeof(int) * blk.size());
    d.frow = (const int*)up(frow.data(), sizeof(int) * frow.size());
    if (I.kind == 2) {
        std::vector<double> g1(I.g1x), l1(I.l1x);
        g1.insert(g1.end(), I.g1w.begin(), I.g1w.end());
        l1.insert(l1.end(), I.l1w.begin(), I.l1w.end());
        std::vector<int> rexp, fexp;
        for (const auto& e : I.rexp) rexp.insert(rexp.end(), {e[0], e[1], e[2]});
        for (const auto& e : I.fexp) fexp.insert(fexp.end(), {e[0], e[1]});
        d.ng1 = (int)I.g1x.size();
        d.nl1 = (int)I.l1x.size();
        d.g1 = (const double*)up(g1.data(), sizeof(double) * g1.size());
        d.l1 = (const double*)up(l1.data(), sizeof(double) * l1.size());
        d.rexp = (const int*)up(rexp.data(), sizeof(int) * rexp.size());
        d.fexp = (const int*)up(fexp.data(), sizeof(int) * fexp.size());
        c.sync();  // staging vectors of this block are released at its end
    }
    c.sync();  // host staging vectors are released on return
    return d;
}

double HhoProblem::L2ErrorDevice(Ctx& c, const double* x_dev) const {
    FH_CHECK(!c.host_only, -9, "device post-processing needs a device context");
    const Impl& I = *impl;
    std::vector<DevBuf> keep;
    HhoDev d = make_device_problem(c, *this, I, keep);
    DevBuf out(sizeof(double) * 2 * std::max(1, I.ncell));
    d.post = 1;
    d.xsol = x_dev;
    d.cell_err = out.as<double>();
    launch_hho_local(c, d);  // synchronises
    std::vector<double> h(2 * (size_t)I.ncell);
    FH_HIP(hipMemcpyAsync(h.data(), out.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    double err = 0, nrm2 = 0;  // summed in cell order
    for (int cc = 0; cc < I.ncell; cc++) {
        err += h[2 * (size_t)cc];
        nrm2 += h[2 * (size_t)cc + 1];
    }
    return nrm2 > 0 ? std::sqrt(err / nrm2) : std::sqrt(err);
}

// Two-colouring of the cells for the element layout: every unknown face
// must lie in exactly two cells, of different colours (the cell graph over
// unknown faces is bipartite).  Returns false (no layout) otherwise, or when
// the cells do not have six faces.
static bool elem_colouring(const HhoProblem::Impl& I, std::vector<int>& pos, int ncol[2]) {
    const int m = I.nfc;
    if (m != 6) return false;
    std::vector<int> fc((size_t)2 * I.nface, -1);
    for (int c = 0; c < I.ncell; c++)
        for (int f = 0; f < m; f++) {
            const int g = I.cf[(size_t)c * m + f];
            if (I.funk[g] < 0) continue;
            int* s = &fc[(size_t)2 * g];
            if (s[0] < 0) s[0] = c;
            else if (s[1] < 0) s[1] = c;
            else return false;  // a face in three cells
        }
    for (int g = 0; g < I.nface; g++)
        if (I.funk[g] >= 0 && fc[(size_t)2 * g + 1] < 0) return false;  // an unknown face in one cell
    std::vector<int> col(I.ncell, -1), queue;
    queue.reserve(I.ncell);
    for (int s = 0; s < I.ncell; s++) {
        if (col[s] >= 0) continue;
        col[s] = 0;
        queue.assign(1, s);
        for (size_t h = 0; h < queue.size(); h++) {
            const int c = queue[h];
            for (int f = 0; f < m; f++) {
                const int g = I.cf[(size_t)c * m + f];
                if (I.funk[g] < 0) continue;
                const int o = fc[(size_t)2 * g] == c ? fc[(size_t)2 * g + 1] : fc[(size_t)2 * g];
                if (col[o] < 0) {
                    col[o] = 1 - col[c];
                    queue.push_back(o);
                } else if (col[o] == col[c]) {
                    return false;  // not bipartite
                }
            }
        }
    }
    pos.assign(I.ncell, -1);
    ncol[0] = ncol[1] = 0;
    for (int c = 0; c < I.ncell; c++) ncol[col[c]]++;
    int next[2] = {0, ncol[0]};
    for (int c = 0; c < I.ncell; c++) pos[c] = next[col[c]]++;
    return true;
}

std::unique_ptr<Bcsr> HhoProblem::AssembleDevice(Ctx& c, double* b_dev, bool keep_host) const {
    FH_CHECK(!c.host_only, -9, "device assembly needs a device context");
    const Impl& I = *impl;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<DevBuf> keep;
    HhoDev d = make_device_problem(c, *this, I, keep);
    // element layout: the condensed cell matrices kept beside the block rows
    std::shared_ptr<ElemLayout> E;
    std::vector<int> pos;
    int ncol[2];
    if (elem_enabled() && nf * I.nfc <= 64 && nf * I.nfc % 2 == 0 && elem_colouring(I, pos, ncol)) {
        E = std::make_shared<ElemLayout>();
        E->nfc = I.nfc;
        E->nf = nf;
        E->m = I.nfc * nf;
        E->stride = elem_stride(E->m);
        E->ncol[0] = ncol[0];
        E->ncol[1] = ncol[1];
        E->prec = Prec::F64;
        std::vector<int> rows((size_t)I.ncell * I.nfc);
        for (int cc = 0; cc < I.ncell; cc++)
            for (int f = 0; f < I.nfc; f++)
                rows[(size_t)pos[cc] * I.nfc + f] = I.funk[I.cf[(size_t)cc * I.nfc + f]];
        E->rows = std::make_shared<DevBuf>(sizeof(int) * rows.size());
        FH_HIP(hipMemcpyAsync(E->rows->p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, c.stream));
        const size_t nv = (size_t)I.ncell * E->stride;
        E->vals.alloc(sizeof(double) * nv);
        FH_HIP(hipMemsetAsync(E->vals.p, 0, sizeof(double) * nv, c.stream));
        keep.emplace_back(sizeof(int) * pos.size());
        FH_HIP(hipMemcpyAsync(keep.back().p, pos.data(), sizeof(int) * pos.size(), hipMemcpyHostToDevice, c.stream));
        d.elem_pos = keep.back().as<int>();
        d.elem_vals = E->vals.as<double>();
        d.elem_stride = E->stride;
        c.sync();  // host staging released
    }
    auto A = std::make_unique<Bcsr>();
    A->nbr = A->nbc = nb;
    A->bs_r = A->bs_c = nf;
    A->nnzb = rowptr[nb];
    A->prec = Prec::F64;
    A->h_rowptr = rowptr;
    A->h_colidx = colidx;
    A->rowptr.alloc(sizeof(int) * rowptr.size());
    A->colidx.alloc(sizeof(int) * std::max<size_t>(1, colidx.size()));
    FH_HIP(hipMemcpyAsync(A->rowptr.p, rowptr.data(), sizeof(int) * rowptr.size(), hipMemcpyHostToDevice, c.stream));
    FH_HIP(hipMemcpyAsync(A->colidx.p, colidx.data(), sizeof(int) * colidx.size(), hipMemcpyHostToDevice, c.stream));
    const size_t nval = (size_t)A->nnzb * nf * nf;
    A->vals.alloc(sizeof(double) * std::max<size_t>(1, nval));
    FH_HIP(hipMemsetAsync(A->vals.p, 0, sizeof(double) * nval, c.stream));
    FH_HIP(hipMemsetAsync(b_dev, 0, sizeof(double) * (size_t)nb * nf, c.stream));
    d.vals = A->vals.as<double>();
    d.b = b_dev;
    launch_hho_local(c, d);  // synchronises
    if (E) {
        A->elem = E;
        set_elem(*A, true);
    }
    if (keep_host) {
        A->h_vals.resize(nval);
        FH_HIP(hipMemcpyAsync(A->h_vals.data(), A->vals.p, sizeof(double) * nval, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    }
    device_assemble_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return A;
}

}  // namespace fh
