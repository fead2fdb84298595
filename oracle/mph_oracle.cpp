// mph_oracle.cpp -- CPU restatement of the reference MPH-I time step.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path and bench.py's CPU
// baseline).  Nothing in the product links or calls this file.
//
// Every function cites the lines of source/main.cpp (Masahiro-Kondo-AIST/MphImplicit)
// whose arithmetic it restates.  Expressions keep the reference's operand order and
// association, because the parity bar is bit equality.  The file is compiled with
// -ffp-contract=off and no -march, like the reference binaries (x86-64 baseline: no FMA).
//
// Pinning (what checks this file):
//   * every phase against dumps of the unmodified reference (oracle/ref_probe.cpp)
//     committed as fixtures under tests/golden/ -- bit equality, serial dot order;
//   * closed forms and brute force in tests/test_oracle_*.py (kernel normalisation
//     integrals, N0 lattice sums, O(N^2) neighbour lists, bitonic network against a
//     sort, dense Galerkin products, CR on a manufactured system).

#include "mph_oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

thread_local std::string g_err;

// Large arrays (neighbour lists, the CSR matrices) are allocated without the serial
// value-initialisation of std::vector and then filled by all threads, as the
// reference's malloc + parallel first writes do: the pages land on the threads'
// memory and no single thread touches gigabytes.  Values are the same as assign().
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U>&) {}
    template <class U>
    void construct(U* p) noexcept {
        ::new ((void*)p) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new ((void*)p) U(std::forward<A>(a)...);
    }
};
template <class T>
using bvec = std::vector<T, NoInitAlloc<T>>;

template <class V, class T>
void par_assign(V& v, size_t n, T val) {
    v.resize(n);
    T* d = v.data();
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)n; ++k) d[k] = val;
}

inline void par_copy(const double* from, long n, double* to) {
#pragma omp parallel for schedule(static)
    for (long k = 0; k < n; ++k) to[k] = from[k];
}

constexpr int kFluidBegin = 0, kFluidEnd = 2, kWallBegin = 2, kWallEnd = 4;  // main.cpp:52-56
inline bool is_fluid(int t) { return kFluidBegin <= t && t < kFluidEnd; }
inline bool is_wall(int t) { return kWallBegin <= t && t < kWallEnd; }

// Mod macro, main.cpp:79
inline double modw(double x, double w) { return x - w * std::floor(x / w); }
// minimum-image difference used by every neighbour loop, e.g. main.cpp:1452
inline double min_image(double xj, double xi, double w) {
    return modw(xj - xi + 0.5 * w, w) - 0.5 * w;
}

struct Level {  // one multigrid level, main.cpp:2650-2746 (structured form)
    int mn[3], mx[3], cnt[3];
    int cs[3];  // stencil width per axis: min(3, cnt)
    int cells, nrow, nshift;
    std::vector<double> A;     // [nrow][nshift], shift order of main.cpp:3318
    std::vector<double> invD;  // [nrow]
    std::vector<double> S, R, Q;
};

struct Scene {
    mph_params p{};
    int dim = 3;
    int dot_mode = 0;
    // constants (initializeWeight/Fluid/Wall/Domain)
    double spacing = 0, volume = 0, dmin[3]{}, dmax[3]{}, dwidth[3]{};
    double radA = 0, radG = 0, radP = 0, radV = 0, maxR = 0;
    double Swa = 1, Swg = 1, Swp = 1, Swv = 1, N0a = 1, N0p = 1, R2g = 1;
    double cofA[4]{}, cofK = 0;
    double wallRot[4][3][3]{}, wallCenter[4][3]{};
    int cellCount[3]{}, totalCells = 0;
    double cellWidth[3]{};
    // time bookkeeping (main.cpp:381-383, 386-472)
    double time = 0, outputNext = 0, vtkOutputNext = 0;
    int iStep = 0;
    // particles
    int np = 0, p2 = 0, p2pow = 0;
    std::vector<int> type, perm;
    std::vector<double> pos, vel, force, mass, kappa, lambda, mu;
    std::vector<double> densA, gcen, pressA, volStrainP, divP, pressP, virialP, virialStress;
    // cells
    std::vector<int> cellIndex, cellParticle;
    std::vector<int> cfb, cfe, cwb, cwe;
    int fb = 0, fe = 0, wb = 0, we = 0;
    // neighbours
    std::vector<int> nCount, nFluid, nCountP;
    std::vector<long> nPtr, nPtrP;
    bvec<int> nInd, nIndP;
    // implicit system
    bvec<long> aPtr, cPtr;
    bvec<int> aInd, cInd;
    bvec<double> aVal, cVal;
    std::vector<double> b, vecP;
    // multigrid
    int depth = 0;
    std::vector<Level> lv;
    std::vector<double> invDiagA;
    // CR summary
    int crIter = 0;
    double rr0 = 0, rr = 0, rs0 = 0, rs = 0;
    // CONVERGENCE_CHECK (main.cpp:46): per iteration (rr, r.s); after the loop the
    // true residual's r.r and r.s
    bool convCheck = false;
    std::vector<double> hist;
    double checkRR = 0, checkRS = 0;
    bool vtkStep = false;
    double diag[4]{};
};

// ---------------------------------------------------------------------------
// weight functions, main.cpp:236-307 (h^2 normalisation in 2D, h^3 in 3D)
inline double hpow(const Scene& s, double h) { return s.dim == 2 ? h * h : h * h * h; }
inline double wa(const Scene& s, double r, double h) {
    return 1.0 / s.Swa * 1.0 / hpow(s, h) * (r / h) * (1.0 - (r / h)) * (1.0 - (r / h));
}
inline double dwadr(const Scene& s, double r, double h) {
    return 1.0 / s.Swa * 1.0 / hpow(s, h) * (1.0 - (r / h)) * (1.0 - 3.0 * (r / h)) * (1.0 / h);
}
inline double wg(const Scene& s, double r, double h) {
    return 1.0 / s.Swg * 1.0 / hpow(s, h) * ((1.0 - r / h) * (1.0 - r / h));
}
inline double dwgdr(const Scene& s, double r, double h) {
    return 1.0 / s.Swg * 1.0 / hpow(s, h) * (-2.0 / h * (1.0 - r / h));
}
inline double wp(const Scene& s, double r, double h) {
    return 1.0 / s.Swp * 1.0 / hpow(s, h) * ((1.0 - r / h) * (1.0 - r / h));
}
inline double dwpdr(const Scene& s, double r, double h) {
    return 1.0 / s.Swp * 1.0 / hpow(s, h) * (-2.0 / h * (1.0 - r / h));
}
inline double dwvdr(const Scene& s, double r, double h) {
    return 1.0 / s.Swv * 1.0 / hpow(s, h) * (-2.0 / h * (1.0 - r / h));
}

inline void rel(const Scene& s, int i, int j, double x[3]) {
    for (int d = 0; d < 3; ++d) x[d] = min_image(s.pos[3 * j + d], s.pos[3 * i + d], s.dwidth[d]);
}
inline double norm2(const double x[3]) { return x[0] * x[0] + x[1] * x[1] + x[2] * x[2]; }

// ---------------------------------------------------------------------------
// initializeWeight, main.cpp:886-985
void init_weight(Scene& s) {
    const double l = s.spacing;
    s.radA = s.p.radius_ratio_a * l;
    s.radG = s.p.radius_ratio_a * l;  // RadiusRatioG = RadiusRatioA (main.cpp:888)
    s.radP = s.p.radius_ratio_p * l;
    s.radV = s.p.radius_ratio_v * l;
    if (s.dim == 2) {
        s.Swa = 1.0 / 2.0 * 2.0 / 15.0 * M_PI / l / l;
        s.Swg = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l / l;
        s.Swp = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l / l;
        s.Swv = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l / l;
        s.R2g = 1.0 / 2.0 * 1.0 / 30.0 * M_PI * s.radG * s.radG / l / l / s.Swg;
    } else {
        s.Swa = 1.0 / 3.0 * 1.0 / 5.0 * M_PI / l / l / l;
        s.Swg = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l / l / l;
        s.Swp = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l / l / l;
        s.Swv = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l / l / l;
        s.R2g = 1.0 / 3.0 * 4.0 / 105.0 * M_PI * s.radG * s.radG / l / l / l / s.Swg;
    }
    auto lattice_sum = [&](double radius, bool useA) {
        const int range = (int)std::ceil(radius / l);
        const int rz = s.dim == 2 ? 0 : range;
        double sum = 0.0;
        for (int ix = -range; ix <= range; ++ix)
            for (int iy = -range; iy <= range; ++iy)
                for (int iz = -rz; iz <= rz; ++iz) {
                    if (ix == 0 && iy == 0 && iz == 0) continue;
                    const double x = l * ((double)ix), y = l * ((double)iy), z = l * ((double)iz);
                    const double r2 = x * x + y * y + z * z;
                    if (r2 <= radius * radius) {
                        const double r = std::sqrt(r2);
                        sum += useA ? wa(s, r, radius) : wp(s, r, radius);
                    }
                }
        return sum;
    };
    s.N0a = lattice_sum(s.radA, true);
    s.N0p = lattice_sum(s.radP, false);
}

// initializeFluid, main.cpp:988-1021 (per-particle coefficients are recomputed
// each step by calculatePhysicalCoefficients)
void init_fluid(Scene& s) {
    double integN, integX;
    if (s.dim == 2) { s.cofK = 0.350778153; integN = 0.024679383; integX = 0.226126699; }
    else            { s.cofK = 0.326976006; integN = 0.021425779; integX = 0.233977488; }
    for (int t = 0; t < 4; ++t)
        s.cofA[t] = s.p.surface_tension[t] / ((s.radG / s.spacing) * (integN + s.cofK * s.cofK * integX));
    for (int i = 0; i < s.np; ++i) {
        const int t = s.type[i];
        s.mass[i] = s.p.density[t] * s.volume;
        s.kappa[i] = s.p.bulk_modulus[t];
        s.lambda[i] = s.p.bulk_viscosity[t];
        s.mu[i] = s.p.shear_viscosity[t];
    }
}

// initializeWall, main.cpp:1023-1062
void init_wall(Scene& s) {
    for (int w = kWallBegin; w < kWallEnd; ++w) {
        const double* om = s.p.wall_omega[w];
        double n[3] = {0.0, 0.0, 0.0};
        const double theta = std::fabs(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        if (theta != 0.0)
            for (int d = 0; d < 3; ++d) n[d] = om[d] / theta;
        const double q0 = n[0] * std::sin(theta * s.p.dt / 2.0), q1 = n[1] * std::sin(theta * s.p.dt / 2.0),
                     q2 = n[2] * std::sin(theta * s.p.dt / 2.0), q3 = std::cos(theta * s.p.dt / 2.0);
        double (&R)[3][3] = s.wallRot[w];
        R[0][0] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
        R[0][1] = 2.0 * (q0 * q1 - q2 * q3);
        R[0][2] = 2.0 * (q0 * q2 + q1 * q3);
        R[1][0] = 2.0 * (q0 * q1 + q2 * q3);
        R[1][1] = -q0 * q0 + q1 * q1 - q2 * q2 + q3 * q3;
        R[1][2] = 2.0 * (q1 * q2 - q0 * q3);
        R[2][0] = 2.0 * (q0 * q2 - q1 * q3);
        R[2][1] = 2.0 * (q1 * q2 + q0 * q3);
        R[2][2] = -q0 * q0 - q1 * q1 + q2 * q2 + q3 * q3;
        for (int d = 0; d < 3; ++d) s.wallCenter[w][d] = s.p.wall_center[w][d];
    }
}

// initializeDomain, main.cpp:1064-1111
void init_domain(Scene& s) {
    double m = 0.0;
    m = (s.radA > m) ? s.radA : m;
    m = (2.0 * s.radP > m) ? 2.0 * s.radP : m;
    m = (s.radV > m) ? s.radV : m;
    s.maxR = m;
    for (int d = 0; d < 3; ++d) s.dwidth[d] = s.dmax[d] - s.dmin[d];
    double cc[3];
    cc[0] = std::floor((s.dmax[0] - s.dmin[0]) / m);
    cc[1] = std::floor((s.dmax[1] - s.dmin[1]) / m);
    cc[2] = s.dim == 2 ? 1 : std::floor((s.dmax[2] - s.dmin[2]) / m);
    for (int d = 0; d < 3; ++d) s.cellCount[d] = (int)cc[d];
    s.totalCells = cc[0] * cc[1] * cc[2];
    for (int d = 0; d < 3; ++d) s.cellWidth[d] = s.dwidth[d] / s.cellCount[d];
    s.cfb.assign(s.totalCells, 0); s.cfe.assign(s.totalCells, 0);
    s.cwb.assign(s.totalCells, 0); s.cwe.assign(s.totalCells, 0);
}

inline int cell_id(const Scene& s, int cx, int cy, int cz) {  // CellId macro, main.cpp:109-113
    const int* C = s.cellCount;
    if (s.dim == 2) return ((cx) % C[0] + C[0]) % C[0] * C[1] + ((cy) % C[1] + C[1]) % C[1];
    return ((cx) % C[0] + C[0]) % C[0] * C[1] * C[2] + ((cy) % C[1] + C[1]) % C[1] * C[2] +
           ((cz) % C[2] + C[2]) % C[2];
}

// ---------------------------------------------------------------------------
// kinematics: calculateConvection 1678-1688, calculateWall 2089-2128,
// calculatePeriodicBoundary 4741-4752, resetForce 1690-1701
void phase_kinematics(Scene& s) {
    const double dt = s.p.dt;
#pragma omp parallel for
    for (int i = s.fb; i < s.fe; ++i)
        for (int d = 0; d < 3; ++d) s.pos[3 * i + d] += s.vel[3 * i + d] * dt;
#pragma omp parallel for
    for (int i = s.wb; i < s.we; ++i)
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] = 0.0;
#pragma omp parallel for
    for (int i = s.wb; i < s.we; ++i) {
        const int t = s.type[i];
        const double (&R)[3][3] = s.wallRot[t];
        const double* w = s.p.wall_omega[t];
        const double* V = s.p.wall_velocity[t];
        const double* c = s.wallCenter[t];
        double r[3] = {s.pos[3 * i] - c[0], s.pos[3 * i + 1] - c[1], s.pos[3 * i + 2] - c[2]};
        r[0] = R[0][0] * r[0] + R[0][1] * r[1] + R[0][2] * r[2];  // in place, as the reference
        r[1] = R[1][0] * r[0] + R[1][1] * r[1] + R[1][2] * r[2];
        r[2] = R[2][0] * r[0] + R[2][1] * r[1] + R[2][2] * r[2];
        s.vel[3 * i + 0] = w[1] * r[2] - w[2] * r[1] + V[0];
        s.vel[3 * i + 1] = w[2] * r[0] - w[0] * r[2] + V[1];
        s.vel[3 * i + 2] = w[0] * r[1] - w[1] * r[0] + V[2];
        s.pos[3 * i + 0] = r[0] + c[0] + V[0] * dt;
        s.pos[3 * i + 1] = r[1] + c[1] + V[1] * dt;
        s.pos[3 * i + 2] = r[2] + c[2] + V[2] * dt;
    }
    for (int w = kWallBegin; w < kWallEnd; ++w)
        for (int d = 0; d < 3; ++d) s.wallCenter[w][d] += s.p.wall_velocity[w][d] * dt;
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i)
        for (int d = 0; d < 3; ++d)
            s.pos[3 * i + d] = modw(s.pos[3 * i + d] - s.dmin[d], s.dwidth[d]) + s.dmin[d];
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i)
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] = 0.0;
}

}  // namespace

// bitonic network of main.cpp:1188-1211: compare on the key only, swap when
// (key[i] > key[i|dist]) == up.  Data-oblivious and not stable.
extern "C" void mpho_bitonic_sort(int power, int* key, int* val) {
    const int n = 1 << power;
    for (int iMain = 0; iMain < power; ++iMain) {
        for (int iSub = 0; iSub <= iMain; ++iSub) {
            const int dist = 1 << (iMain - iSub);
#pragma omp parallel for
            for (int i = 0; i < n; ++i) {
                const bool up = ((i >> iMain) & 2) == 0;
                if (((i & dist) == 0) && ((key[i] > key[i | dist]) == up)) {
                    std::swap(key[i], key[i | dist]);
                    std::swap(val[i], val[i | dist]);
                }
            }
        }
    }
}

namespace {

// calculateCellParticle, main.cpp:1164-1379
void phase_cells(Scene& s) {
    const int T = s.totalCells;
#pragma omp parallel for
    for (int i = 0; i < s.p2; ++i) {
        if (i < s.np) {
            const int cx = ((int)std::floor((s.pos[3 * i] - s.dmin[0]) / s.cellWidth[0])) % s.cellCount[0];
            const int cy = ((int)std::floor((s.pos[3 * i + 1] - s.dmin[1]) / s.cellWidth[1])) % s.cellCount[1];
            const int cz = ((int)std::floor((s.pos[3 * i + 2] - s.dmin[2]) / s.cellWidth[2])) % s.cellCount[2];
            s.cellIndex[i] = cell_id(s, cx, cy, cz);
            if (is_wall(s.type[i])) s.cellIndex[i] += T;
            s.cellParticle[i] = i;
        } else {
            s.cellIndex[i] = 2 * T;
            s.cellParticle[i] = s.np;
        }
    }
    mpho_bitonic_sort(s.p2pow, s.cellIndex.data(), s.cellParticle.data());
    // cell ranges: the boundary detection + up/down-sweep scan of main.cpp:1213-1300
    // yields Begin[c] = #keys < c and End[c] = #keys <= c for every cell (integer
    // counting, order independent); computed here by counting directly.
    std::vector<int> cnt(2 * T + 1, 0);
    for (int i = 0; i < s.np; ++i) cnt[s.cellIndex[i]]++;
    int run = 0;
    std::vector<int> start(2 * T + 1);
    for (int c = 0; c <= 2 * T; ++c) { start[c] = run; if (c < 2 * T) run += cnt[c]; }
    for (int c = 0; c < T; ++c) {
        s.cfb[c] = start[c];
        s.cfe[c] = start[c + 1];
        s.cwb[c] = start[T + c];
        s.cwe[c] = start[T + c + 1];
    }
    s.fb = s.cfb[0]; s.fe = s.cfe[T - 1]; s.wb = s.cwb[0]; s.we = s.cwe[T - 1];
    // gather-reorder, main.cpp:1315-1378
    std::vector<int> ti(s.np);
    std::vector<double> tv(3 * (size_t)s.np);
    for (int i = 0; i < s.np; ++i) ti[i] = s.perm[s.cellParticle[i]];
    s.perm.swap(ti);
    for (int i = 0; i < s.np; ++i) ti[i] = s.type[s.cellParticle[i]];
    std::copy(ti.begin(), ti.end(), s.type.begin());
    for (int i = 0; i < s.np; ++i)
        for (int d = 0; d < 3; ++d) tv[3 * i + d] = s.pos[3 * s.cellParticle[i] + d];
    s.pos.swap(tv);
    for (int i = 0; i < s.np; ++i)
        for (int d = 0; d < 3; ++d) tv[3 * i + d] = s.vel[3 * s.cellParticle[i] + d];
    s.vel.swap(tv);
}

// calculateNeighbor, main.cpp:1382-1667.  The 27-cell (9 in 2D) traversal with the
// rotated start index visits neighbour cells in ascending CellId, so each list
// comes out ascending: fluid block first, then wall block.
template <class Visit>
void scan_cells(const Scene& s, int i, bool fluid, Visit visit) {
    const int* C = s.cellCount;
    const int T = s.totalCells;
    const int key = s.cellIndex[i];
    const int rx = (int)std::ceil(s.maxR / s.cellWidth[0]);
    const int ry = (int)std::ceil(s.maxR / s.cellWidth[1]);
    const int rz = s.dim == 2 ? 0 : (int)std::ceil(s.maxR / s.cellWidth[2]);
    const int icx = (key / (C[1] * C[2])) % T;
    const int icy = (key / C[2]) % C[1];
    const int icz = key % C[2];
    int jx[3], jy[3], jz[3];
    for (int k = 0; k < 2 * rx + 1; ++k) jx[k] = ((icx - rx + k) % C[0] + C[0]) % C[0];
    for (int k = 0; k < 2 * ry + 1; ++k) jy[k] = ((icy - ry + k) % C[1] + C[1]) % C[1];
    for (int k = 0; k < 2 * rz + 1; ++k) jz[k] = ((icz - rz + k) % C[2] + C[2]) % C[2];
    const int bx = (2 * rx) - (icx + rx) % C[0], by = (2 * ry) - (icy + ry) % C[1], bz = (2 * rz) - (icz + rz) % C[2];
    const int x0 = bx > 0 ? bx : 0, y0 = by > 0 ? by : 0, z0 = bz > 0 ? bz : 0;
    for (int a = x0; a < x0 + (2 * rx + 1); ++a)
        for (int b = y0; b < y0 + (2 * ry + 1); ++b)
            for (int c = z0; c < z0 + (2 * rz + 1); ++c) {
                const int jc = cell_id(s, jx[a % (2 * rx + 1)], jy[b % (2 * ry + 1)], jz[c % (2 * rz + 1)]);
                const int lo = fluid ? s.cfb[jc] : s.cwb[jc];
                const int hi = fluid ? s.cfe[jc] : s.cwe[jc];
                for (int j = lo; j < hi; ++j) {
                    double q[3];
                    rel(s, i, j, q);
                    visit(j, norm2(q));
                }
            }
}

int phase_neighbor(Scene& s) {
    const int rx = (int)std::ceil(s.maxR / s.cellWidth[0]);
    const int ry = (int)std::ceil(s.maxR / s.cellWidth[1]);
    const int rz = s.dim == 2 ? 0 : (int)std::ceil(s.maxR / s.cellWidth[2]);
    if (2 * rx + 1 > 3 || 2 * ry + 1 > 3 || 2 * rz + 1 > 3) {  // assert, main.cpp:1402-1404
        g_err = "neighbour range exceeds one cell (assert of main.cpp:1402)";
        return MPH_E_STATE;
    }
    const double R2 = s.maxR * s.maxR, P2 = s.radP * s.radP;
    s.nCount.assign(s.np, 0); s.nFluid.assign(s.np, 0); s.nCountP.assign(s.np, 0);
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        int n = 0, nf = 0, npp = 0;
        scan_cells(s, i, true, [&](int, double q2) {
            if (q2 <= R2) { n++; nf++; if (q2 <= P2) npp++; }
        });
        if (!(s.wb <= i && i < s.we && n == 0))
            scan_cells(s, i, false, [&](int, double q2) {
                if (q2 <= R2) { n++; if (q2 <= P2) npp++; }
            });
        s.nCount[i] = n; s.nFluid[i] = nf; s.nCountP[i] = npp;
    }
    s.nPtr.assign(s.np + 1, 0); s.nPtrP.assign(s.np + 1, 0);
    for (int i = 0; i < s.np; ++i) {
        s.nPtr[i + 1] = s.nPtr[i] + s.nCount[i];
        s.nPtrP[i + 1] = s.nPtrP[i] + s.nCountP[i];
    }
    par_assign(s.nInd, s.nPtr[s.np], 0);
    par_assign(s.nIndP, s.nPtrP[s.np], 0);
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        long k = s.nPtr[i], kp = s.nPtrP[i];
        auto put = [&](int j, double q2) {
            if (q2 <= R2) { s.nInd[k++] = j; if (q2 <= P2) s.nIndP[kp++] = j; }
        };
        scan_cells(s, i, true, put);
        if (!(s.wb <= i && i < s.we && s.nCount[i] == 0)) scan_cells(s, i, false, put);
    }
    return MPH_OK;
}

// densities: calculateDensityA 1737-1764, calculateGravityCenter 1767-1799,
// calculateDensityP 1903-1930
void densities(Scene& s) {
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        double sa = 0.0, sg[3] = {0.0, 0.0, 0.0}, sp = 0.0;
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            const double ratio = s.p.interaction_ratio[s.type[i]][s.type[j]];
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (s.radA * s.radA - r2 >= 0) {
                const double r = std::sqrt(r2);
                const double weight = ratio * wa(s, r, s.radA);
                sa += weight;
            }
            if (s.radG * s.radG - r2 >= 0) {
                const double r = std::sqrt(r2);
                const double weight = ratio * wg(s, r, s.radG);
                for (int d = 0; d < 3; ++d) sg[d] += x[d] * weight / s.R2g * s.radG;
            }
            if (s.radP * s.radP - r2 >= 0) {
                const double r = std::sqrt(r2);
                const double weight = wp(s, r, s.radP);
                sp += weight;
            }
        }
        s.densA[i] = sa;
        for (int d = 0; d < 3; ++d) s.gcen[3 * i + d] = sg[d];
        s.volStrainP[i] = (sp - s.N0p);
    }
}

// calculatePhysicalCoefficients 1704-1735, calculatePressureA 1801-1849,
// calculateDiffuseInterface 1851-1901, calculateGravity 2065-2074,
// calculateAcceleration 2076-2086
void phase_explicit(Scene& s) {
    densities(s);
    const double V = s.volume, l = s.spacing, dt = s.p.dt;
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        const int t = s.type[i];
        s.mass[i] = s.p.density[t] * V;
        s.kappa[i] = s.p.bulk_modulus[t];
        if (s.volStrainP[i] < 0.0) s.kappa[i] = 0.0;
        s.lambda[i] = s.p.bulk_viscosity[t];
        s.mu[i] = s.p.shear_viscosity[t];
        s.pressA[i] = s.cofA[t] * (s.densA[i] - s.N0a) / l;
        if (s.N0a <= s.densA[i]) s.pressA[i] = 0.0;
    }
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            const double rij_ = s.p.interaction_ratio[s.type[i]][s.type[j]];
            const double rji_ = s.p.interaction_ratio[s.type[j]][s.type[i]];
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (s.radA * s.radA - r2 > 0) {
                const double r = std::sqrt(r2);
                const double dwij = rij_ * dwadr(s, r, s.radA);
                const double dwji = rji_ * dwadr(s, r, s.radA);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                for (int d = 0; d < 3; ++d) f[d] += (s.pressA[i] * dwij + s.pressA[j] * dwji) * e[d] * V;
            }
        }
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] += f[d];
    }
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        const double ai = s.cofA[s.type[i]] * (s.cofK) * (s.cofK);
        double f[3] = {0.0, 0.0, 0.0};
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            const double aj = s.cofA[s.type[i]] * (s.cofK) * (s.cofK);  // i's type (main.cpp:1863)
            const double rij_ = s.p.interaction_ratio[s.type[i]][s.type[j]];
            const double rji_ = s.p.interaction_ratio[s.type[j]][s.type[i]];
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (s.radG * s.radG - r2 > 0) {
                const double r = std::sqrt(r2);
                const double wij = rij_ * wg(s, r, s.radG);
                const double wji = rji_ * wg(s, r, s.radG);
                const double* gi = &s.gcen[3 * i];
                const double* gj = &s.gcen[3 * j];
                for (int d = 0; d < 3; ++d)
                    f[d] -= (aj * gj[d] * wji - ai * gi[d] * wij) / s.R2g * s.radG * (V / l);
                const double dwij = rij_ * dwgdr(s, r, s.radG);
                const double dwji = rji_ * dwgdr(s, r, s.radG);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double gr = 0.0;
                for (int d = 0; d < 3; ++d) gr += (aj * gj[d] * dwji - ai * gi[d] * dwij) * x[d];
                for (int d = 0; d < 3; ++d) f[d] -= (gr)*e[d] / s.R2g * s.radG * (V / l);
            }
        }
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] += f[d];
    }
#pragma omp parallel for
    for (int i = s.fb; i < s.fe; ++i)
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] += s.mass[i] * s.p.gravity[d];
#pragma omp parallel for
    for (int i = s.fb; i < s.fe; ++i)
        for (int d = 0; d < 3; ++d) s.vel[3 * i + d] += s.force[3 * i + d] / s.mass[i] * dt;
}

// calculateMatrixA, main.cpp:2137-2337
void phase_matrix_a(Scene& s) {
    const int nf = s.fe - s.fb;
    const long n = 3L * nf;
    par_assign(s.aPtr, n + 1, 0L);
    for (int i = 0; i < nf; ++i)
        for (int r = 0; r < 3; ++r) s.aPtr[3 * i + r + 1] = 3L * s.nFluid[i];
    for (long k = 0; k < n; ++k) s.aPtr[k + 1] += s.aPtr[k];
    const long nnz = s.aPtr[n];
    par_assign(s.aInd, nnz, 0);
    par_assign(s.aVal, nnz, 0.0);
    s.b.assign(n, 0.0);
#pragma omp parallel for
    for (int i = 0; i < nf; ++i)
        for (int jn = 0; jn < s.nFluid[i]; ++jn) {
            const int j = s.nInd[s.nPtr[i] + jn];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) s.aInd[s.aPtr[3 * i + r] + 3 * jn + c] = 3 * j + c;
        }
    const double V = s.volume, dt = s.p.dt;
    const double visc = s.dim == 2 ? 8.0 : 10.0;
#pragma omp parallel for
    for (int i = s.fb; i < s.fe; ++i) {
        int iN = 0;
        double self[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double sumvec[3] = {0.0, 0.0, 0.0};
        for (int jn = 0; jn < s.nCount[i]; ++jn) {
            const int j = s.nInd[s.nPtr[i] + jn];
            if (i == j) { iN = jn; continue; }
            double x[3];
            rel(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (r2 == 0.0) continue;
            if (s.radV * s.radV - r2 > 0) {
                const double dij = std::sqrt(r2);
                const double wdrij = -dwvdr(s, dij, s.radV);
                const double e[3] = {x[0] / dij, x[1] / dij, x[2] / dij};
                const double muij = 2.0 * (s.mu[i] * s.mu[j]) / (s.mu[i] + s.mu[j]);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        const double coef = +visc * muij * e[c] * e[r] * wdrij / dij;
                        self[r][c] += coef;
                        if (is_fluid(s.type[j])) s.aVal[s.aPtr[3 * i + r] + 3 * jn + c] = -coef;
                        else if (is_wall(s.type[j])) sumvec[r] += coef * s.vel[3 * j + c];
                    }
            }
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) s.aVal[s.aPtr[3 * i + r] + 3 * iN + c] = self[r][c];
        for (int r = 0; r < 3; ++r) s.b[3 * i + r] += sumvec[r];
        for (int d = 0; d < 3; ++d) {
            const double coef = s.mass[i] / V / dt;
            s.aVal[s.aPtr[3 * i + d] + 3 * iN + d] += coef;
            s.b[3 * i + d] += coef * s.vel[3 * i + d];
        }
    }
}

// calculateMatrixC, main.cpp:2354-2551
void phase_matrix_c(Scene& s) {
    const int nf = s.fe - s.fb;
    const long n = 3L * nf;
    par_assign(s.cPtr, n + 1, 0L);
    for (int i = 0; i < nf; ++i)
        for (int r = 0; r < 3; ++r) s.cPtr[3 * i + r + 1] = s.nCountP[i];
    for (long k = 0; k < n; ++k) s.cPtr[k + 1] += s.cPtr[k];
    const long nnz = s.cPtr[n];
    par_assign(s.cInd, nnz, 0);
    par_assign(s.cVal, nnz, 0.0);
    s.vecP.assign(s.np, 0.0);
#pragma omp parallel for
    for (int i = 0; i < nf; ++i)
        for (int jn = 0; jn < s.nCountP[i]; ++jn) {
            const int j = s.nIndP[s.nPtrP[i] + jn];
            for (int r = 0; r < 3; ++r) s.cInd[s.cPtr[3 * i + r] + jn] = j;
        }
    const double P2 = s.radP * s.radP;
#pragma omp parallel for
    for (int i = 0; i < nf; ++i) {
        int iN = 0;
        double self[3] = {0.0, 0.0, 0.0};
        for (int jn = 0; jn < s.nCountP[i]; ++jn) {
            const int j = s.nIndP[s.nPtrP[i] + jn];
            if (i == j) { iN = jn; continue; }
            double x[3];
            rel(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (r2 == 0.0) continue;
            if (P2 - r2 > 0) {
                const double dij = std::sqrt(r2);
                const double e[3] = {x[0] / dij, x[1] / dij, x[2] / dij};
                const double wpdrij = -dwpdr(s, dij, s.radP);
                for (int r = 0; r < 3; ++r) {
                    const double coef = e[r] * wpdrij;
                    self[r] += coef;
                    s.cVal[s.cPtr[3 * i + r] + jn] = coef;
                }
            }
        }
        for (int r = 0; r < 3; ++r) s.cVal[s.cPtr[3 * i + r] + iN] = self[r];
    }
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        s.vecP[i] = 0.0;
        if (s.volStrainP[i] > 0.0) s.vecP[i] = s.kappa[i] * s.volStrainP[i];
        double sum = 0.0;
        for (int jn = 0; jn < s.nCountP[i]; ++jn) {
            const int j = s.nIndP[s.nPtrP[i] + jn];
            if (i == j) continue;
            double x[3];
            rel(s, i, j, x);
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            if (r2 == 0.0) continue;
            if (P2 - r2 > 0) {
                const double dij = std::sqrt(r2);
                const double e[3] = {x[0] / dij, x[1] / dij, x[2] / dij};
                const double wpdrij = -dwpdr(s, dij, s.radP);
                for (int c = 0; c < 3; ++c) {
                    if (is_wall(s.type[i])) {
                        const double coef = +e[c] * wpdrij;
                        sum += s.lambda[i] * coef * s.vel[3 * i + c];
                    }
                    if (is_wall(s.type[j])) {
                        const double coef = -e[c] * wpdrij;
                        sum += s.lambda[i] * coef * s.vel[3 * j + c];
                    }
                }
            }
        }
        s.vecP[i] += sum;
    }
}

// multiplyMatrixC, main.cpp:2562-2643: b -= C P, then A += C Lambda C^T by merging
// the two sorted P-neighbour lists of every fluid pair.
void phase_multiply_c(Scene& s) {
    const int nf = s.fe - s.fb;
    const long n = 3L * nf;
#pragma omp parallel for
    for (long row = 0; row < n; ++row)
        for (long k = s.cPtr[row]; k < s.cPtr[row + 1]; ++k) s.b[row] -= s.cVal[k] * s.vecP[s.cInd[k]];
#pragma omp parallel for
    for (int i = 0; i < nf; ++i)
        for (int jn = 0; jn < s.nFluid[i]; ++jn) {
            const int j = s.nInd[s.nPtr[i] + jn];
            int a = 0, c = 0;
            double sum[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
            while (a < s.nCountP[i] && c < s.nCountP[j]) {
                const int ka = s.nIndP[s.nPtrP[i] + a];
                const int kc = s.nIndP[s.nPtrP[j] + c];
                if (ka == kc) {
                    for (int r = 0; r < 3; ++r)
                        for (int q = 0; q < 3; ++q)
                            sum[r][q] += s.cVal[s.cPtr[3 * i + r] + a] * s.lambda[ka] * s.cVal[s.cPtr[3 * j + q] + c];
                    a++; c++;
                } else if (ka < kc) {
                    a++;
                } else {
                    c++;
                }
            }
            for (int r = 0; r < 3; ++r)
                for (int q = 0; q < 3; ++q) s.aVal[s.aPtr[3 * i + r] + 3 * jn + q] += sum[r][q];
        }
}

// ---------------------------------------------------------------------------
// multigrid hierarchy: calculateMultiGridMatrix and its builders, main.cpp:2761-3546.
// The reference stores R, P, P2G, G2P and the level matrices as CSR; their patterns
// are fixed stencils, so here they are applied implicitly with the same summation
// order (children in R order, stencil entries in shift order).

inline int shift_of(const Level& L, int ix, int iy, int iz, int sd) {
    return 3 * (L.cs[2] * L.cs[1] * ix + L.cs[2] * iy + iz) + sd;
}

int mg_build(Scene& s) {
    const int* C = s.cellCount;
    const int T = s.totalCells;
    // calculateMultiGridDepth, 2777-2859
    int xmax = 0, ymax = 0, zmax = 0, xmin = C[0], ymin = C[1], zmin = C[2];
    for (int c = 0; c < T; ++c)
        if (s.cfe[c] - s.cfb[c] > 0) {
            const int cx = (c / (C[1] * C[2])) % C[0], cy = (c / C[2]) % C[1], cz = c % C[2];
            xmax = std::max(xmax, cx); ymax = std::max(ymax, cy); zmax = std::max(zmax, cz);
            xmin = std::min(xmin, cx); ymin = std::min(ymin, cy); zmin = std::min(zmin, cz);
        }
    double mgc = 0.0;
    if ((xmax - xmin) > mgc) mgc = (xmax - xmin + 1);
    if ((ymax - ymin) > mgc) mgc = (ymax - ymin + 1);
    if ((zmax - zmin) > mgc) mgc = (zmax - zmin + 1);
    int depth = 0;
    while ((1 << depth) <= mgc) depth++;
    if (depth < 2) {
        g_err = "multigrid depth < 2: all fluid in one cell (the reference V-cycle reads level -1)";
        return MPH_E_STATE;
    }
    s.depth = depth;
    s.lv.assign(depth, Level());
    const int mn0[3] = {xmin, ymin, zmin}, mx0[3] = {xmax, ymax, zmax};
    for (int l = 0; l < depth; ++l) {
        Level& L = s.lv[l];
        for (int d = 0; d < 3; ++d) {
            L.mn[d] = mn0[d] >> l;
            L.mx[d] = (mx0[d] >> l) + 1;
            L.cnt[d] = L.mx[d] - L.mn[d];
            L.cs[d] = 3 < L.cnt[d] ? 3 : L.cnt[d];
        }
        L.cells = L.cnt[0] * L.cnt[1] * L.cnt[2];
        L.nrow = 3 * L.cells;
        L.nshift = 3 * L.cs[0] * L.cs[1] * L.cs[2];
        L.A.assign((size_t)L.nrow * L.nshift, 0.0);
        L.invD.assign(L.nrow, 0.0);
        L.S.assign(L.nrow, -1.0);
        L.R.assign(L.nrow, -1.0);
        L.Q.assign(L.nrow, -1.0);
    }
    const int nf = s.fe - s.fb;
    const long n = 3L * nf;
    // calculateInvDiagA, 3045-3069
    s.invDiagA.assign(n, 0.0);
#pragma omp parallel for
    for (long row = 0; row < n; ++row)
        for (long k = s.aPtr[row]; k < s.aPtr[row + 1]; ++k)
            if (s.aInd[k] == row && s.aVal[k] != 0.0) s.invDiagA[row] = 1.0 / s.aVal[k];
    // level 0 = sum of particle blocks per cell pair (Galerkin P2G A G2P), 3336-3411
    {
        Level& L = s.lv[0];
#pragma omp parallel for
        for (int row = 0; row < L.nrow; ++row) {
            const int g = row / 3, rd = row % 3;
            const int gx = (g / (L.cnt[1] * L.cnt[2])) % L.cnt[0], gy = (g / L.cnt[2]) % L.cnt[1], gz = g % L.cnt[2];
            std::vector<double> acc(81, 0.0);
            const int cx = gx + L.mn[0], cy = gy + L.mn[1], cz = gz + L.mn[2];
            if (cx < C[0] && cy < C[1] && cz < C[2]) {
                const int ic = cell_id(s, cx, cy, cz);
                for (int ip = s.cfb[ic]; ip < s.cfe[ic]; ++ip) {
                    const long prow = 3L * ip + rd;
                    for (long k = s.aPtr[prow]; k < s.aPtr[prow + 1]; ++k) {
                        const int col = s.aInd[k], jp = col / 3, sd = col % 3;
                        const int key = s.cellIndex[jp];
                        const int jcx = (key / (C[1] * C[2])) % T, jcy = (key / C[2]) % C[1], jcz = key % C[2];
                        const int dx = (jcx - L.mn[0]) - gx, dy = (jcy - L.mn[1]) - gy, dz = (jcz - L.mn[2]) - gz;
                        const int ix = (dx + 1 + L.cnt[0]) % L.cnt[0];
                        const int iy = (dy + 1 + L.cnt[1]) % L.cnt[1];
                        const int iz = (dz + 1 + L.cnt[2]) % L.cnt[2];
                        acc[shift_of(L, ix, iy, iz, sd)] += s.aVal[k];
                    }
                }
            }
            for (int sh = 0; sh < L.nshift; ++sh) L.A[(size_t)row * L.nshift + sh] = acc[sh];
        }
    }
    // coarser levels: R A_{l-1} P by summation, 3414-3507
    for (int l = 1; l < depth; ++l) {
        Level& L = s.lv[l];
        const Level& S = s.lv[l - 1];
#pragma omp parallel for
        for (int row = 0; row < L.nrow; ++row) {
            const int g = row / 3, rd = row % 3;
            const int lx = (g / (L.cnt[1] * L.cnt[2])) % L.cnt[0], ly = (g / L.cnt[2]) % L.cnt[1], lz = g % L.cnt[2];
            const int lcx = lx + L.mn[0], lcy = ly + L.mn[1], lcz = lz + L.mn[2];
            const int fmx = S.mn[0] > 2 * lcx, fmy = S.mn[1] > 2 * lcy, fmz = S.mn[2] > 2 * lcz;
            const int fMx = S.mx[0] < 2 * (lcx + 1), fMy = S.mx[1] < 2 * (lcy + 1), fMz = S.mx[2] < 2 * (lcz + 1);
            const int ncx = 2 - fmx - fMx, ncy = 2 - fmy - fMy, ncz = 2 - fmz - fMz;
            std::vector<double> acc(81, 0.0);
            for (int ax = 0; ax < ncx; ++ax)
                for (int ay = 0; ay < ncy; ++ay)
                    for (int az = 0; az < ncz; ++az) {
                        const int sgx = (2 * lcx + ax + fmx) - S.mn[0];
                        const int sgy = (2 * lcy + ay + fmy) - S.mn[1];
                        const int sgz = (2 * lcz + az + fmz) - S.mn[2];
                        const int srow = 3 * ((S.cnt[1] * S.cnt[2]) * sgx + S.cnt[2] * sgy + sgz) + rd;
                        for (int ex = 0; ex < S.cs[0]; ++ex)
                            for (int ey = 0; ey < S.cs[1]; ++ey)
                                for (int ez = 0; ez < S.cs[2]; ++ez)
                                    for (int sd = 0; sd < 3; ++sd) {
                                        const int jsx = (sgx + ex - 1 + S.cnt[0]) % S.cnt[0];
                                        const int jsy = (sgy + ey - 1 + S.cnt[1]) % S.cnt[1];
                                        const int jsz = (sgz + ez - 1 + S.cnt[2]) % S.cnt[2];
                                        const int jlx = ((jsx + S.mn[0]) / 2) - L.mn[0];
                                        const int jly = ((jsy + S.mn[1]) / 2) - L.mn[1];
                                        const int jlz = ((jsz + S.mn[2]) / 2) - L.mn[2];
                                        const int ix = ((jlx - lx) + 1 + L.cnt[0]) % L.cnt[0];
                                        const int iy = ((jly - ly) + 1 + L.cnt[1]) % L.cnt[1];
                                        const int iz = ((jlz - lz) + 1 + L.cnt[2]) % L.cnt[2];
                                        acc[shift_of(L, ix, iy, iz, sd)] +=
                                            S.A[(size_t)srow * S.nshift + shift_of(S, ex, ey, ez, sd)];
                                    }
                    }
            for (int sh = 0; sh < L.nshift; ++sh) L.A[(size_t)row * L.nshift + sh] = acc[sh];
        }
    }
    // calculateMultiGridInvDiagA, 3510-3546
    for (int l = 0; l < depth; ++l) {
        Level& L = s.lv[l];
        for (int row = 0; row < L.nrow; ++row) {
            double sum = 0.0;
            for (int sh = 0; sh < L.nshift; ++sh) sum += std::fabs(L.A[(size_t)row * L.nshift + sh]);
            L.invD[row] = sum != 0.0 ? 2.0 / sum : 0.0;
        }
    }
    return MPH_OK;
}

// y = beta*y + sum_k (alpha*a_k)*x_k over the level stencil (myDcsrmv, main.cpp:4015-4031)
void mg_spmv(const Level& L, double alpha, const std::vector<double>& x, double beta, std::vector<double>& y) {
#pragma omp parallel for
    for (int row = 0; row < L.nrow; ++row) {
        const int g = row / 3;
        const int gx = (g / (L.cnt[1] * L.cnt[2])) % L.cnt[0], gy = (g / L.cnt[2]) % L.cnt[1], gz = g % L.cnt[2];
        double sum = 0.0;
        const double* a = &L.A[(size_t)row * L.nshift];
        for (int ex = 0; ex < L.cs[0]; ++ex)
            for (int ey = 0; ey < L.cs[1]; ++ey)
                for (int ez = 0; ez < L.cs[2]; ++ez)
                    for (int sd = 0; sd < 3; ++sd) {
                        const int jx = (gx + ex - 1 + L.cnt[0]) % L.cnt[0];
                        const int jy = (gy + ey - 1 + L.cnt[1]) % L.cnt[1];
                        const int jz = (gz + ez - 1 + L.cnt[2]) % L.cnt[2];
                        const int col = 3 * ((L.cnt[1] * L.cnt[2]) * jx + L.cnt[2] * jy + jz) + sd;
                        sum += alpha * a[shift_of(L, ex, ey, ez, sd)] * x[col];
                    }
        y[row] *= beta;
        y[row] += sum;
    }
}

// y = beta*y + alpha*diag*x (myDdmv, main.cpp:4052-4062)
void ddmv(long n, double alpha, const double* diag, const double* x, double beta, double* y) {
#pragma omp parallel for
    for (long k = 0; k < n; ++k) {
        y[k] *= beta;
        y[k] += alpha * diag[k] * x[k];
    }
}

// restriction R (level l-1 -> l), myDcsrmv with alpha=1, beta=0 over main.cpp:3115-3175
void mg_restrict(const Level& S, const Level& L, const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for
    for (int row = 0; row < L.nrow; ++row) {
        const int g = row / 3, rd = row % 3;
        const int lx = (g / (L.cnt[1] * L.cnt[2])) % L.cnt[0], ly = (g / L.cnt[2]) % L.cnt[1], lz = g % L.cnt[2];
        const int lcx = lx + L.mn[0], lcy = ly + L.mn[1], lcz = lz + L.mn[2];
        const int fmx = S.mn[0] > 2 * lcx, fmy = S.mn[1] > 2 * lcy, fmz = S.mn[2] > 2 * lcz;
        const int fMx = S.mx[0] < 2 * (lcx + 1), fMy = S.mx[1] < 2 * (lcy + 1), fMz = S.mx[2] < 2 * (lcz + 1);
        double sum = 0.0;
        for (int ax = 0; ax < 2 - fmx - fMx; ++ax)
            for (int ay = 0; ay < 2 - fmy - fMy; ++ay)
                for (int az = 0; az < 2 - fmz - fMz; ++az) {
                    const int sgx = (2 * lcx + ax + fmx) - S.mn[0];
                    const int sgy = (2 * lcy + ay + fmy) - S.mn[1];
                    const int sgz = (2 * lcz + az + fmz) - S.mn[2];
                    sum += 1.0 * 1.0 * x[3 * ((S.cnt[1] * S.cnt[2]) * sgx + S.cnt[2] * sgy + sgz) + rd];
                }
        y[row] *= 0.0;
        y[row] += sum;
    }
}

// prolongation P (level l -> l-1): y += injection of x, main.cpp:3216-3241 with beta=1
void mg_prolong(const Level& L, const Level& S, const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for
    for (int row = 0; row < S.nrow; ++row) {
        const int g = row / 3, rd = row % 3;
        const int sx = (g / (S.cnt[1] * S.cnt[2])) % S.cnt[0], sy = (g / S.cnt[2]) % S.cnt[1], sz = g % S.cnt[2];
        const int lx = ((sx + S.mn[0]) / 2) - L.mn[0];
        const int ly = ((sy + S.mn[1]) / 2) - L.mn[1];
        const int lz = ((sz + S.mn[2]) / 2) - L.mn[2];
        double sum = 0.0;
        sum += 1.0 * 1.0 * x[3 * ((L.cnt[1] * L.cnt[2]) * lx + L.cnt[2] * ly + lz) + rd];
        y[row] *= 1.0;
        y[row] += sum;
    }
}

inline int grid_of_particle(const Scene& s, const Level& L0, int ip) {
    const int* C = s.cellCount;
    const int key = s.cellIndex[ip];
    const int cx = key / (C[1] * C[2]) % C[0], cy = key / (C[2]) % C[1], cz = key % C[2];
    return (L0.cnt[1] * L0.cnt[2]) * (cx - L0.mn[0]) + L0.cnt[2] * (cy - L0.mn[1]) + (cz - L0.mn[2]);
}

// preconditionWithMultiGrid, main.cpp:4122-4266: s = D^-1 q + G2P V(P2G q)
void precondition(Scene& s, const double* q, double* out, double* buf) {
    const int nf = s.fe - s.fb;
    const long n = 3L * nf;
    par_copy(q, n, buf);
    ddmv(n, 1.0, s.invDiagA.data(), buf, 0.0, out);
    par_copy(q, n, buf);
    const int D = s.depth;
    Level& L0 = s.lv[0];
    const int* C = s.cellCount;
    // P2G (main.cpp:2964-2997): row (g, r) sums q over the cell's fluid particles
#pragma omp parallel for
    for (int row = 0; row < L0.nrow; ++row) {
        const int g = row / 3, rd = row % 3;
        const int gx = (g / (L0.cnt[1] * L0.cnt[2])) % L0.cnt[0], gy = (g / L0.cnt[2]) % L0.cnt[1], gz = g % L0.cnt[2];
        (void)C;
        const int ic = cell_id(s, gx + L0.mn[0], gy + L0.mn[1], gz + L0.mn[2]);
        double sum = 0.0;
        for (int ip = s.cfb[ic]; ip < s.cfe[ic]; ++ip) sum += 1.0 * 1.0 * buf[3 * ip + rd];
        L0.Q[row] *= 0.0;
        L0.Q[row] += sum;
    }
    auto scal0 = [](std::vector<double>& v) { for (double& x : v) x *= 0.0; };
    auto smooth_pre = [&](Level& L, int sweeps) {  // {S += invD R; R = Q - A S}
        for (int it = 0; it < sweeps; ++it) {
            ddmv(L.nrow, 1.0, L.invD.data(), L.R.data(), 1.0, L.S.data());
            L.R = L.Q;
            mg_spmv(L, -1.0, L.S, 1.0, L.R);
        }
    };
    auto smooth_post = [&](Level& L, int sweeps) {  // {R = Q - A S; S += invD R}
        for (int it = 0; it < sweeps; ++it) {
            L.R = L.Q;
            mg_spmv(L, -1.0, L.S, 1.0, L.R);
            ddmv(L.nrow, 1.0, L.invD.data(), L.R.data(), 1.0, L.S.data());
        }
    };
    scal0(L0.S);
    L0.R = L0.Q;
    smooth_pre(L0, 2);
    for (int l = 1; l < D - 1; ++l) {
        mg_restrict(s.lv[l - 1], s.lv[l], s.lv[l - 1].R, s.lv[l].Q);
        scal0(s.lv[l].S);
        s.lv[l].R = s.lv[l].Q;
        smooth_pre(s.lv[l], 3);
    }
    {
        Level& L = s.lv[D - 1];
        mg_restrict(s.lv[D - 2], L, s.lv[D - 2].R, L.Q);
        scal0(L.S);
        L.R = L.Q;
        smooth_pre(L, 4);
        ddmv(L.nrow, 1.0, L.invD.data(), L.R.data(), 1.0, L.S.data());
        mg_prolong(L, s.lv[D - 2], L.S, s.lv[D - 2].S);
    }
    for (int l = D - 2; l > 0; --l) {
        smooth_post(s.lv[l], 3);
        mg_prolong(s.lv[l], s.lv[l - 1], s.lv[l].S, s.lv[l - 1].S);
    }
    smooth_post(L0, 2);
    // G2P (main.cpp:3013-3042): particle row picks its cell
#pragma omp parallel for
    for (int ip = 0; ip < nf; ++ip) {
        const int g = grid_of_particle(s, L0, ip);
        for (int rd = 0; rd < 3; ++rd) {
            double sum = 0.0;
            sum += 1.0 * 1.0 * L0.S[3 * g + rd];
            out[3 * ip + rd] *= 1.0;
            out[3 * ip + rd] += sum;
        }
    }
}

}  // namespace

// dot products: serial order of myDdot (main.cpp:4064-4075), the GPU tree, or OpenMP
extern "C" double mpho_dot(int mode, long n, const double* x, const double* y) {
    if (mode == MPHO_DOT_SERIAL) {
        double sum = 0.0;
        for (long i = 0; i < n; ++i) sum += x[i] * y[i];
        return sum;
    }
    if (mode == MPHO_DOT_OMP) {
        double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
        for (long i = 0; i < n; ++i) sum += x[i] * y[i];
        return sum;
    }
    // MPHO_DOT_TREE.  n = 3*m.  Chunks of 256 particles (768 entries): lane l of
    // chunk c holds t = (x0*y0 + x1*y1) + x2*y2 of particle c*256+l (0.0 past the
    // end); pairwise tree t[l] += t[l+h] for h = 128..1.  The chunk sums are then
    // folded by 256 lanes, lane l summing chunks l, l+256, ... from 0.0, followed by
    // the same pairwise tree.
    const long m = n / 3;
    const long nchunks = (m + 255) / 256;
    std::vector<double> part(nchunks > 0 ? nchunks : 1, 0.0);
#pragma omp parallel for
    for (long c = 0; c < nchunks; ++c) {
        double t[256];
        for (int l = 0; l < 256; ++l) {
            const long i = c * 256 + l;
            if (i < m) {
                const double* a = x + 3 * i;
                const double* b = y + 3 * i;
                t[l] = (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
            } else {
                t[l] = 0.0;
            }
        }
        for (int h = 128; h >= 1; h >>= 1)
            for (int l = 0; l < h; ++l) t[l] = t[l] + t[l + h];
        part[c] = t[0];
    }
    double u[256];
    for (int l = 0; l < 256; ++l) {
        double acc = 0.0;
        for (long c = l; c < nchunks; c += 256) acc = acc + part[c];
        u[l] = acc;
    }
    for (int h = 128; h >= 1; h >>= 1)
        for (int l = 0; l < h; ++l) u[l] = u[l] + u[l + h];
    return u[0];
}

namespace {

// y = beta*y + sum (alpha*a)*x over CSR A (myDcsrmvForA, main.cpp:4033-4050)
void spmv_a(const Scene& s, double alpha, const double* x, double beta, double* y) {
    const long n = 3L * (s.fe - s.fb);
#pragma omp parallel for
    for (long row = 0; row < n; ++row) {
        double sum = 0.0;
        for (long k = s.aPtr[row]; k < s.aPtr[row + 1]; ++k) sum += alpha * s.aVal[k] * x[s.aInd[k]];
        y[row] *= beta;
        y[row] += sum;
    }
}

// solveWithConjugatedGradient, main.cpp:4270-4421 (preconditioned conjugate residual)
void phase_solve(Scene& s) {
    const int nf = s.fe - s.fb;
    const long n = 3L * nf;
    bvec<double> x, r, sv, y, p, q, u, buf;
    par_assign(x, n, 0.0);
    for (bvec<double>* v : {&r, &sv, &y, &p, &q, &u, &buf}) par_assign(*v, n, -1.0);
    const double dt = s.p.dt;
    for (int i = 0; i < nf; ++i)
        for (int d = 0; d < 3; ++d) x[3 * i + d] = s.vel[3 * i + d] - s.force[3 * i + d] / s.mass[i] * dt;
    if (s.p.multigrid)
        for (Level& L : s.lv) {
            std::fill(L.S.begin(), L.S.end(), -1.0);
            std::fill(L.R.begin(), L.R.end(), -1.0);
            std::fill(L.Q.begin(), L.Q.end(), -1.0);
        }
    auto M = [&](const double* in, double* out) {
        if (s.p.multigrid) precondition(s, in, out, buf.data());
        else par_copy(in, n, out);
    };
    auto dot = [&](const double* a, const double* b) { return mpho_dot(s.dot_mode, n, a, b); };
    const double* b = s.b.data();
    double rho = 1.0, rhop = 0.0, tmp, alpha, beta;
    M(b, sv.data());
    s.rs0 = dot(b, sv.data());
    s.rr0 = dot(b, b);
    par_copy(b, n, r.data());
    spmv_a(s, -1.0, x.data(), 1.0, r.data());
    M(r.data(), sv.data());
    const long maxit = s.p.max_iterations > 0 ? s.p.max_iterations : n;
    long iter;
    s.hist.clear();
    for (iter = 0; iter < maxit; ++iter) {
        s.rr = dot(r.data(), r.data());
        if (s.rr / s.rr0 < s.p.cr_tolerance) break;
        if (s.convCheck) {  // main.cpp:4348-4351: rs = r.s, logged with rr
            s.hist.push_back(s.rr);
            s.hist.push_back(dot(r.data(), sv.data()));
        }
        spmv_a(s, 1.0, sv.data(), 0.0, y.data());
        rho = dot(sv.data(), y.data());
        if (iter == 0) {
            par_copy(sv.data(), n, p.data());
            spmv_a(s, 1.0, p.data(), 0.0, q.data());
        } else {  // myDscal + myDaxpy of p and q, element by element
            beta = rho / rhop;
#pragma omp parallel for schedule(static)
            for (long k = 0; k < n; ++k) {
                p[k] *= beta;
                p[k] += 1.0 * sv[k];
                q[k] *= beta;
                q[k] += 1.0 * y[k];
            }
        }
        M(q.data(), u.data());
        tmp = dot(q.data(), u.data());
        alpha = rho / tmp;
#pragma omp parallel for schedule(static)
        for (long k = 0; k < n; ++k) {  // the three myDaxpy of x, r, s
            x[k] += alpha * p[k];
            r[k] += -alpha * q[k];
            sv[k] += -alpha * u[k];
        }
        rhop = rho;
    }
    s.rr = dot(r.data(), r.data());
    s.rs = dot(r.data(), sv.data());
    s.crIter = (int)iter;
    if (s.convCheck) {  // main.cpp:4383-4397: r = b - A x, s = M^-1 r, then r.r and r.s
        par_copy(b, n, r.data());
        spmv_a(s, -1.0, x.data(), 1.0, r.data());
        if (s.p.multigrid) M(r.data(), sv.data());
        else par_copy(b, n, sv.data());  // main.cpp:4391 copies b, not r, without multigrid
        s.checkRR = dot(r.data(), r.data());
        s.checkRS = dot(r.data(), sv.data());
    }
    for (int i = 0; i < nf; ++i)
        for (int d = 0; d < 3; ++d) s.vel[3 * i + d] = x[3 * i + d];
}

// output-step diagnostics: calculateDivergenceP 1932-1969, calculatePressureP
// 1971-2015, calculateViscosityV 2017-2063, calculateVirialStressAtParticle 4494-4738
void phase_diagnostics(Scene& s) {
    const double V = s.volume, l = s.spacing, P2 = s.radP * s.radP;
    const double visc = s.dim == 2 ? 8.0 : 10.0;
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        double sum = 0.0;
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (P2 - r2 >= 0) {
                const double r = std::sqrt(r2);
                const double dw = dwpdr(s, r, s.radP);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double u[3];
                for (int d = 0; d < 3; ++d) u[d] = s.vel[3 * j + d] - s.vel[3 * i + d];
                for (int d = 0; d < 3; ++d) sum -= u[d] * e[d] * dw;
            }
        }
        s.divP[i] = sum;
    }
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        s.pressP[i] = -s.lambda[i] * s.divP[i];
        if (s.volStrainP[i] > 0.0) s.pressP[i] += s.kappa[i] * s.volStrainP[i];
    }
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (P2 - r2 > 0) {
                const double r = std::sqrt(r2);
                const double dw = dwpdr(s, r, s.radP);
                const double gw[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                for (int d = 0; d < 3; ++d) f[d] += (s.pressP[i] + s.pressP[j]) * gw[d] * V;
            }
        }
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] += f[d];
    }
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        double f[3] = {0.0, 0.0, 0.0};
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (s.radV * s.radV - r2 > 0) {
                const double r = std::sqrt(r2);
                const double dwij = -dwvdr(s, r, s.radV);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                double u[3];
                for (int d = 0; d < 3; ++d) u[d] = s.vel[3 * j + d] - s.vel[3 * i + d];
                const double muij = 2.0 * (s.mu[i] * s.mu[j]) / (s.mu[i] + s.mu[j]);
                for (int d = 0; d < 3; ++d)
                    f[d] += visc * muij * (u[0] * e[0] + u[1] * e[1] + u[2] * e[2]) * e[d] * dwij / r * V;
            }
        }
        for (int d = 0; d < 3; ++d) s.force[3 * i + d] += f[d];
    }
    // virial stress: four passes accumulated in the reference's order
#pragma omp parallel for
    for (int i = 0; i < s.np; ++i) {
        double* vs = &s.virialStress[9 * (size_t)i];
        for (int k = 0; k < 9; ++k) vs[k] = 0.0;
        double st[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        for (long k = s.nPtrP[i]; k < s.nPtrP[i] + s.nCountP[i]; ++k) {  // pressureP, 4512-4555
            const int j = s.nIndP[k];
            if (i == j) continue;
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (P2 - r2 > 0) {
                const double r = std::sqrt(r2);
                const double dw = dwpdr(s, r, s.radP);
                const double gw[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                double fij[3];
                for (int d = 0; d < 3; ++d) fij[d] = (s.pressP[i]) * gw[d] * V;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) st[a][c] += 1.0 * fij[a] * x[c] / V;
            }
        }
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) vs[3 * a + c] += st[a][c];
        double sa[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double sv[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        double sg[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        for (long k = s.nPtr[i]; k < s.nPtr[i] + s.nCount[i]; ++k) {
            const int j = s.nInd[k];
            if (i == j) continue;
            double x[3];
            rel(s, i, j, x);
            const double r2 = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            if (r2 == 0.0) continue;
            if (s.radA * s.radA - r2 > 0) {  // pressureA, 4557-4602
                const double ratio = s.p.interaction_ratio[s.type[i]][s.type[j]];
                const double r = std::sqrt(r2);
                const double dw = ratio * dwadr(s, r, s.radA);
                const double gw[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                double fij[3];
                for (int d = 0; d < 3; ++d) fij[d] = (s.pressA[i]) * gw[d] * V;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) sa[a][c] += 1.0 * fij[a] * x[c] / V;
            }
            if (s.radV * s.radV - r2 > 0) {  // viscosity, 4604-4653
                const double r = std::sqrt(r2);
                const double dwij = -dwvdr(s, r, s.radV);
                const double e[3] = {x[0] / r, x[1] / r, x[2] / r};
                const double vij[3] = {s.vel[3 * j] - s.vel[3 * i], s.vel[3 * j + 1] - s.vel[3 * i + 1],
                                       s.vel[3 * j + 2] - s.vel[3 * i + 2]};
                const double muij = 2.0 * (s.mu[i] * s.mu[j]) / (s.mu[i] + s.mu[j]);
                double fij[3];
                for (int d = 0; d < 3; ++d)
                    fij[d] = visc * muij * (vij[0] * e[0] + vij[1] * e[1] + vij[2] * e[2]) * e[d] * dwij / r * V;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) sv[a][c] += 0.5 * fij[a] * x[c] / V;
            }
            if (s.radG * s.radG - r2 > 0) {  // diffuse interface, 4655-4725
                const double a_ = s.cofA[s.type[i]] * (s.cofK) * (s.cofK);
                const double ratio = s.p.interaction_ratio[s.type[i]][s.type[j]];
                const double r = std::sqrt(r2);
                const double weight = ratio * wg(s, r, s.radG);
                const double* gi = &s.gcen[3 * i];
                double fij[3];
                for (int d = 0; d < 3; ++d) fij[d] = -a_ * (-gi[d]) * weight / s.R2g * s.radG * (V / l);
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) sg[a][c] += 1.0 * fij[a] * x[c] / V;
            }
            if (s.radG * s.radG - r2 > 0.0) {
                const double a_ = s.cofA[s.type[i]] * (s.cofK) * (s.cofK);
                const double ratio = s.p.interaction_ratio[s.type[i]][s.type[j]];
                const double r = std::sqrt(r2);
                const double dw = ratio * dwgdr(s, r, s.radG);
                const double gw[3] = {dw * x[0] / r, dw * x[1] / r, dw * x[2] / r};
                const double* gi = &s.gcen[3 * i];
                double gr = 0.0;
                for (int d = 0; d < 3; ++d) gr += (-gi[d]) * x[d];
                double fij[3];
                for (int d = 0; d < 3; ++d) fij[d] = -a_ * (gr)*gw[d] / s.R2g * s.radG * (V / l);
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) sg[a][c] += 1.0 * fij[a] * x[c] / V;
            }
        }
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) vs[3 * a + c] += sa[a][c];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) vs[3 * a + c] += sv[a][c];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) vs[3 * a + c] += sg[a][c];
        if (s.dim == 2) s.virialP[i] = -1.0 / 2.0 * (vs[0] + vs[4]);
        else s.virialP[i] = -1.0 / 3.0 * (vs[0] + vs[4] + vs[8]);
    }
}

int run_phase(Scene& s, int ph) {
    switch (ph) {
        case MPH_PHASE_KINEMATICS: phase_kinematics(s); return MPH_OK;
        case MPH_PHASE_CELLS: phase_cells(s); return MPH_OK;
        case MPH_PHASE_NEIGHBOR: return phase_neighbor(s);
        case MPH_PHASE_EXPLICIT: phase_explicit(s); return MPH_OK;
        case MPH_PHASE_MATRIX_A: phase_matrix_a(s); return MPH_OK;
        case MPH_PHASE_MATRIX_C: phase_matrix_c(s); return MPH_OK;
        case MPH_PHASE_MULTIPLY_C: phase_multiply_c(s); return MPH_OK;
        case MPH_PHASE_MULTIGRID: return s.p.multigrid ? mg_build(s) : MPH_OK;
        case MPH_PHASE_SOLVE: phase_solve(s); return MPH_OK;
        case MPH_PHASE_DIAGNOSTICS: phase_diagnostics(s); return MPH_OK;
        case MPH_PHASE_ADVANCE:
            s.time += s.p.dt;
            s.iStep++;
            return MPH_OK;
        default: g_err = "unknown phase"; return MPH_E_ARG;
    }
}

struct Field {
    void* p;
    size_t count;
    int esize;
};

Field field(Scene& s, const std::string& n) {
    auto F = [](auto& v) { return Field{v.data(), v.size(), (int)sizeof(v[0])}; };
    if (n == "Position") return F(s.pos);
    if (n == "Velocity") return F(s.vel);
    if (n == "Property") return F(s.type);
    if (n == "ParticleIndex") return F(s.perm);
    if (n == "Force") return F(s.force);
    if (n == "Mass") return F(s.mass);
    if (n == "Kappa") return F(s.kappa);
    if (n == "Lambda") return F(s.lambda);
    if (n == "Mu") return F(s.mu);
    if (n == "DensityA") return F(s.densA);
    if (n == "GravityCenter") return F(s.gcen);
    if (n == "PressureA") return F(s.pressA);
    if (n == "VolStrainP") return F(s.volStrainP);
    if (n == "DivergenceP") return F(s.divP);
    if (n == "PressureP") return F(s.pressP);
    if (n == "VirialPressureAtParticle") return F(s.virialP);
    if (n == "VirialStressAtParticle") return F(s.virialStress);
    if (n == "CellIndex") return F(s.cellIndex);
    if (n == "CellParticle") return F(s.cellParticle);
    if (n == "CellFluidParticleBegin") return F(s.cfb);
    if (n == "CellFluidParticleEnd") return F(s.cfe);
    if (n == "CellWallParticleBegin") return F(s.cwb);
    if (n == "CellWallParticleEnd") return F(s.cwe);
    if (n == "NeighborCount") return F(s.nCount);
    if (n == "NeighborFluidCount") return F(s.nFluid);
    if (n == "NeighborCountP") return F(s.nCountP);
    if (n == "NeighborPtr") return F(s.nPtr);
    if (n == "NeighborPtrP") return F(s.nPtrP);
    if (n == "NeighborInd") return F(s.nInd);
    if (n == "NeighborIndP") return F(s.nIndP);
    if (n == "CsrPtrA") return F(s.aPtr);
    if (n == "CsrIndA") return F(s.aInd);
    if (n == "CsrCofA") return F(s.aVal);
    if (n == "VectorB") return F(s.b);
    if (n == "CsrPtrC") return F(s.cPtr);
    if (n == "CsrIndC") return F(s.cInd);
    if (n == "CsrCofC") return F(s.cVal);
    if (n == "VectorP") return F(s.vecP);
    if (n == "InvDiagA") return F(s.invDiagA);
    if (n == "ConvergenceHistory") return F(s.hist);
    if (n.rfind("MGCofA_L", 0) == 0 || n.rfind("MGInvD_L", 0) == 0) {
        const int l = atoi(n.c_str() + 8);
        if (l >= 0 && l < (int)s.lv.size()) return n[2] == 'C' ? F(s.lv[l].A) : F(s.lv[l].invD);
    }
    return Field{nullptr, 0, 0};
}

}  // namespace

struct mpho_scene {
    Scene s;
};

extern "C" {

const char* mpho_last_error(void) { return g_err.c_str(); }

void mpho_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

mpho_scene* mpho_create(const mph_params* p, const mph_grid* g, int dot_mode) {
    if (!p || !g || g->np <= 0 || (p->dim != 2 && p->dim != 3)) {
        g_err = "invalid parameters";
        return nullptr;
    }
    auto* h = new mpho_scene();
    Scene& s = h->s;
    s.p = *p;
    s.dim = p->dim;
    s.dot_mode = dot_mode;
    s.np = g->np;
    s.spacing = g->spacing;
    s.volume = s.dim == 2 ? s.spacing * s.spacing : s.spacing * s.spacing * s.spacing;
    for (int d = 0; d < 3; ++d) { s.dmin[d] = g->domain_min[d]; s.dmax[d] = g->domain_max[d]; }
    s.time = g->time;
    s.type.assign(g->type, g->type + s.np);
    s.pos.assign(g->pos, g->pos + 3 * (size_t)s.np);
    s.vel.assign(g->vel, g->vel + 3 * (size_t)s.np);
    s.perm.resize(s.np);
    for (int i = 0; i < s.np; ++i) s.perm[i] = i;
    const size_t n = s.np;
    s.force.assign(3 * n, 0.0); s.mass.assign(n, 0.0); s.kappa.assign(n, 0.0);
    s.lambda.assign(n, 0.0); s.mu.assign(n, 0.0); s.densA.assign(n, 0.0); s.gcen.assign(3 * n, 0.0);
    s.pressA.assign(n, 0.0); s.volStrainP.assign(n, 0.0); s.divP.assign(n, 0.0); s.pressP.assign(n, 0.0);
    s.virialP.assign(n, 0.0); s.virialStress.assign(9 * n, 0.0);
    // PowerParticleCount: smallest power of two > ParticleCount (main.cpp:640-645)
    s.p2pow = 0;
    while ((s.np >> s.p2pow) != 0) ++s.p2pow;
    s.p2 = 1 << s.p2pow;
    s.cellIndex.assign(s.p2, 0);
    s.cellParticle.assign(s.p2, 0);
    // fluid/wall ranges of the unsorted input (main.cpp:674-687); only the fluid
    // range is read before the first sort recomputes them
    s.fb = s.fe = s.wb = s.we = 0;
    init_weight(s);
    init_fluid(s);
    init_wall(s);
    init_domain(s);
    // init sequence, main.cpp:369-375
    phase_cells(s);
    int rc = phase_neighbor(s);
    if (rc != MPH_OK) { delete h; return nullptr; }
    densities(s);
    s.iStep = (int)(s.time / s.p.dt);
    s.outputNext = s.time;
    s.vtkOutputNext = s.time;
    return h;
}

int mpho_run_phase(mpho_scene* h, int phase) { return run_phase(h->s, phase); }

int mpho_step(mpho_scene* h, int nsteps, mph_step_stats* st) {
    Scene& s = h->s;
    mph_step_stats local{};
    for (int k = 0; k < nsteps; ++k) {
        if (!(s.time < s.p.end_time + 1.0e-5 * s.p.dt)) break;
        if (s.time + 1.0e-5 * s.p.dt >= s.outputNext) s.outputNext += s.p.output_interval;
        for (int ph = MPH_PHASE_KINEMATICS; ph <= MPH_PHASE_SOLVE; ++ph) {
            const int rc = run_phase(s, ph);
            if (rc != MPH_OK) return rc;
        }
        s.vtkStep = false;
        if (s.time + 1.0e-5 * s.p.dt >= s.vtkOutputNext) {
            phase_diagnostics(s);
            s.vtkOutputNext += s.p.vtk_output_interval;
            s.vtkStep = true;
            local.vtk_steps++;
        }
        run_phase(s, MPH_PHASE_ADVANCE);
        local.steps_done++;
        local.cr_iterations = s.crIter;
        local.cr_iterations_total += s.crIter;
        local.rr0 = s.rr0; local.rr = s.rr; local.rs0 = s.rs0; local.rs = s.rs;
    }
    if (st) *st = local;
    return MPH_OK;
}

int mpho_field_size(mpho_scene* h, const char* name, size_t* bytes, int* esize) {
    Field f = field(h->s, name);
    if (!f.p && f.count == 0 && f.esize == 0) { g_err = std::string("unknown field ") + name; return MPH_E_FIELD; }
    if (bytes) *bytes = f.count * f.esize;
    if (esize) *esize = f.esize;
    return MPH_OK;
}

int mpho_read(mpho_scene* h, const char* name, void* out, size_t bytes) {
    Field f = field(h->s, name);
    if (f.esize == 0) { g_err = std::string("unknown field ") + name; return MPH_E_FIELD; }
    if (bytes != f.count * f.esize) { g_err = "size mismatch"; return MPH_E_ARG; }
    if (bytes) std::memcpy(out, f.p, bytes);
    return MPH_OK;
}

int mpho_write(mpho_scene* h, const char* name, const void* in, size_t bytes) {
    Field f = field(h->s, name);
    if (f.esize == 0) { g_err = std::string("unknown field ") + name; return MPH_E_FIELD; }
    if (bytes != f.count * f.esize) { g_err = "size mismatch"; return MPH_E_ARG; }
    if (bytes) std::memcpy(f.p, in, bytes);
    return MPH_OK;
}

int mpho_scalar(mpho_scene* h, const char* name, double* v) {
    Scene& s = h->s;
    const std::string n = name;
    if (n == "Time") *v = s.time;
    else if (n == "ParticleCount") *v = s.np;
    else if (n == "FluidCount") *v = s.fe - s.fb;
    else if (n == "FluidParticleBegin") *v = s.fb;
    else if (n == "FluidParticleEnd") *v = s.fe;
    else if (n == "WallParticleBegin") *v = s.wb;
    else if (n == "WallParticleEnd") *v = s.we;
    else if (n == "N0a") *v = s.N0a;
    else if (n == "N0p") *v = s.N0p;
    else if (n == "Swa") *v = s.Swa;
    else if (n == "Swg") *v = s.Swg;
    else if (n == "Swp") *v = s.Swp;
    else if (n == "Swv") *v = s.Swv;
    else if (n == "R2g") *v = s.R2g;
    else if (n == "MaxRadius") *v = s.maxR;
    else if (n == "TotalCellCount") *v = s.totalCells;
    else if (n == "CellCount0") *v = s.cellCount[0];
    else if (n == "CellCount1") *v = s.cellCount[1];
    else if (n == "CellCount2") *v = s.cellCount[2];
    else if (n == "MultiGridDepth") *v = s.depth;
    else if (n == "CRIterations") *v = s.crIter;
    else if (n == "rr0") *v = s.rr0;
    else if (n == "rr") *v = s.rr;
    else if (n == "rs0") *v = s.rs0;
    else if (n == "rs") *v = s.rs;
    else if (n == "CheckRR") *v = s.checkRR;
    else if (n == "CheckRS") *v = s.checkRS;
    else if (n == "PowerParticleCount") *v = s.p2;
    else if (n == "VtkStep") *v = s.vtkStep;
    else { g_err = "unknown scalar " + n; return MPH_E_FIELD; }
    return MPH_OK;
}

int mpho_set_option(mpho_scene* h, const char* name, double value) {
    const std::string n = name;
    if (n == "ConvergenceCheck") h->s.convCheck = value != 0.0;
    else { g_err = "unknown option " + n; return MPH_E_ARG; }
    return MPH_OK;
}

int mpho_precondition(mpho_scene* h, const double* q, double* out) {
    Scene& s = h->s;
    if (s.lv.empty()) { g_err = "no multigrid hierarchy"; return MPH_E_STATE; }
    const long n = 3L * (s.fe - s.fb);
    std::vector<double> buf(n, -1.0);
    for (Level& L : s.lv) {
        std::fill(L.S.begin(), L.S.end(), -1.0);
        std::fill(L.R.begin(), L.R.end(), -1.0);
        std::fill(L.Q.begin(), L.Q.end(), -1.0);
    }
    precondition(s, q, out, buf.data());
    return MPH_OK;
}

void mpho_destroy(mpho_scene* h) { delete h; }

}  // extern "C"
