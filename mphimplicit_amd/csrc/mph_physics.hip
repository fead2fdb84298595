// mph_physics.hip -- kinematics, explicit particle interactions and implicit assembly.
//
// Replaces calculateConvection/Wall/PeriodicBoundary/resetForce, the explicit force
// loops (source/main.cpp:1678-2128, 4741-4752), calculateMatrixA/C and
// multiplyMatrixC (2137-2643), and the output-step diagnostics (1932-2063,
// 4494-4738).  Each of a particle's sums runs over the neighbour list in list
// order, so every value is bit-identical to the reference.  The list walks of the
// densities, forces and matrix A are entry-parallel: a group of kLanes lanes owns
// the particle, lane g computes the terms of entries g, g + kLanes, ..., and the
// group's first lane adds them, read back through LDS, in list order.
#include "mph_kernels.h"

namespace mph {

namespace {

inline int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

struct Rel {
    double x[3];
    double r2;
};

// both positions are loaded before the first min_image, whose branch would otherwise
// hold each coordinate's load back until the previous one is done
__device__ inline Rel rel(const DevParams& P, const double* __restrict__ pos, int i, int j) {
    Rel o;
    double pj[3], pi[3];
    for (int d = 0; d < 3; ++d) {
        pj[d] = pos[3 * j + d];
        pi[d] = pos[3 * i + d];
    }
    for (int d = 0; d < 3; ++d) o.x[d] = min_image(pj[d], pi[d], P.W[d], P.hW[d]);
    o.r2 = (o.x[0] * o.x[0] + o.x[1] * o.x[1] + o.x[2] * o.x[2]);
    return o;
}

// lanes per particle of the entry-parallel list walks (a workgroup is 256 lanes):
// kLanes for the MaxRadius lists (about 180 entries in 3D), kLanesP for the RadiusP
// lists (about 20).  Measured on the 1M scene: 4 lanes beat 2, 8 and 16 on the long
// lists (k_forces 3.50 against 3.71 / 4.04 / 5.40 ms), 8 beat 4 on the short ones.
constexpr int kLanes = 4, kLanesP = 8;

// calculateConvection, calculateWall, calculatePeriodicBoundary, resetForce
__global__ void k_kinematics(DevParams P, WallKin wk, const int* __restrict__ type, double* __restrict__ pos,
                             double* __restrict__ vel, double* __restrict__ force) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    if (i < P.nf) {
        for (int d = 0; d < 3; ++d) x[d] += vel[3 * i + d] * P.dt;
    } else {
        const int t = type[i];
        const double(&R)[3][3] = wk.R[t];
        const double* w = wk.omega[t];
        const double* V = wk.vel[t];
        const double* c = wk.center[t];
        double r[3] = {x[0] - c[0], x[1] - c[1], x[2] - c[2]};
        r[0] = R[0][0] * r[0] + R[0][1] * r[1] + R[0][2] * r[2];  // in place, main.cpp:2108-2110
        r[1] = R[1][0] * r[0] + R[1][1] * r[1] + R[1][2] * r[2];
        r[2] = R[2][0] * r[0] + R[2][1] * r[1] + R[2][2] * r[2];
        vel[3 * i + 0] = w[1] * r[2] - w[2] * r[1] + V[0];
        vel[3 * i + 1] = w[2] * r[0] - w[0] * r[2] + V[1];
        vel[3 * i + 2] = w[0] * r[1] - w[1] * r[0] + V[2];
        x[0] = r[0] + c[0] + V[0] * P.dt;
        x[1] = r[1] + c[1] + V[1] * P.dt;
        x[2] = r[2] + c[2] + V[2] * P.dt;
    }
    for (int d = 0; d < 3; ++d) {
        const double t = x[d] - P.dmin[d];
        pos[3 * i + d] = (t - P.W[d] * floor(t / P.W[d])) + P.dmin[d];
        force[3 * i + d] = 0.0;
    }
}

// calculateDensityA, calculateGravityCenter, calculateDensityP (three sums, one list walk)
__global__ void __launch_bounds__(256) k_densities(DevParams P, const int* __restrict__ type,
                                                   const double* __restrict__ pos, const int* __restrict__ nCount,
                                                   const long long* __restrict__ nPtr, const int* __restrict__ nInd,
                                                   double* __restrict__ densA, double* __restrict__ gcen,
                                                   double* __restrict__ volStrainP) {
    // terms: DensityA, the three GravityCenter ones, DensityP.  Each lane takes U
    // entries per round (entries base + u*G + g), whose loads are in flight together;
    // the group's first lane adds the round's terms in list order
    constexpr int G = kLanes, K = 5, U = 2;
    __shared__ double tb[256 * U * K];
    __shared__ int tf[256 * U];
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.np;
    const int ti = own ? type[i] : 0;
    double pi[3] = {0.0, 0.0, 0.0};
    if (own)
        for (int d = 0; d < 3; ++d) pi[d] = pos[3 * i + d];
    double sa = 0.0, sg0 = 0.0, sg1 = 0.0, sg2 = 0.0, sp = 0.0;
    const long long b = own ? nPtr[i] : 0;
    const int n = own ? nCount[i] : 0;
    int jnext[U];  // each lane's next entries load one round ahead
    for (int u = 0; u < U; ++u) jnext[u] = g + u * G < n ? nInd[b + g + u * G] : 0;
    for (int base = 0; base < n; base += G * U) {
        int jcur[U];
        for (int u = 0; u < U; ++u) {
            const int jn = base + u * G + g;
            jcur[u] = jnext[u];
            jnext[u] = jn + G * U < n ? nInd[b + jn + G * U] : 0;
        }
        // the round's loads, before any branch (entries past the list read particle 0)
        int tj[U];
        double pj[U][3];
        for (int u = 0; u < U; ++u) {
            tj[u] = type[jcur[u]];
            for (int d = 0; d < 3; ++d) pj[u][d] = pos[3 * jcur[u] + d];
        }
        for (int u = 0; u < U; ++u) {
            const int jn = base + u * G + g;
            int f = 0;
            double t[K] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (jn < n) {
                const int j = jcur[u];
                if (i != j) {
                    const double ratio = P.ratio[ti][tj[u]];
                    Rel q;  // rel() on the loaded positions
                    for (int d = 0; d < 3; ++d) q.x[d] = min_image(pj[u][d], pi[d], P.W[d], P.hW[d]);
                    q.r2 = (q.x[0] * q.x[0] + q.x[1] * q.x[1] + q.x[2] * q.x[2]);
                    if (P.radA2 - q.r2 >= 0) {
                        const double r = sqrt(q.r2);
                        t[0] = ratio * wa_k(P.Ka, r, P.radA);
                        f |= 1;
                    }
                    if (P.radG2 - q.r2 >= 0) {
                        const double r = sqrt(q.r2);
                        const double weight = ratio * wq_k(P.Kg, r, P.radG);
                        t[1] = q.x[0] * weight / P.R2g * P.radG;
                        t[2] = q.x[1] * weight / P.R2g * P.radG;
                        t[3] = q.x[2] * weight / P.R2g * P.radG;
                        f |= 2;
                    }
                    if (P.radP2 - q.r2 >= 0) {
                        const double r = sqrt(q.r2);
                        t[4] = wq_k(P.Kp, r, P.radP);
                        f |= 4;
                    }
                }
            }
            for (int e = 0; e < K; ++e) tb[K * (U * threadIdx.x + u) + e] = t[e];
            tf[U * threadIdx.x + u] = f;
        }
        wave_sync();
        if (g == 0)
            for (int u = 0; u < U; ++u)
                for (int v = 0; v < G && base + u * G + v < n; ++v) {
                    const int sl = U * (threadIdx.x + v) + u;
                    const int fu = tf[sl];
                    const double* tu = tb + K * sl;
                    if (fu & 1) sa += tu[0];
                    if (fu & 2) {
                        sg0 += tu[1];
                        sg1 += tu[2];
                        sg2 += tu[3];
                    }
                    if (fu & 4) sp += tu[4];
                }
        wave_sync();
    }
    if (g == 0 && own) {
        densA[i] = sa;
        gcen[3 * i + 0] = sg0;
        gcen[3 * i + 1] = sg1;
        gcen[3 * i + 2] = sg2;
        volStrainP[i] = (sp - P.N0p);
    }
}

// calculatePhysicalCoefficients (1704-1735) and the PressureA values (1807-1812)
__global__ void k_coefficients(DevParams P, const int* __restrict__ type, const double* __restrict__ densA,
                               const double* __restrict__ volStrainP, double* __restrict__ mass,
                               double* __restrict__ kappa, double* __restrict__ lambda, double* __restrict__ mu,
                               double* __restrict__ pressA) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    const int t = type[i];
    mass[i] = P.density[t] * P.V;
    kappa[i] = volStrainP[i] < 0.0 ? 0.0 : P.bulk[t];
    lambda[i] = P.bulkVisc[t];
    mu[i] = P.shear[t];
    double pa = P.cofA[t] * (densA[i] - P.N0a) / P.l;
    if (P.N0a <= densA[i]) pa = 0.0;
    pressA[i] = pa;
}

// calculatePressureA forces (1814-1848), calculateDiffuseInterface (1851-1901),
// calculateGravity (2065-2074), calculateAcceleration (2076-2086)
// kSurf false: every type's CofA is 0 (no surface tension).  The DiffuseInterface
// terms are then zeros of either sign, and fd, which starts at +0, stays +0 under
// fd -= (+-0); they are not formed, and the force keeps every bit.
template <bool kSurf>
__global__ void __launch_bounds__(256) k_forces(DevParams P, const int* __restrict__ type,
                                                const double* __restrict__ pos, const int* __restrict__ nCount,
                                                const long long* __restrict__ nPtr, const int* __restrict__ nInd,
                                                const double* __restrict__ pressA, const double* __restrict__ gcen,
                                                const double* __restrict__ mass, double* __restrict__ force,
                                                double* __restrict__ vel) {
    // terms: PressureA force, the two DiffuseInterface ones (3 each)
    constexpr int G = kLanes, K = kSurf ? 9 : 3, U = 2;
    __shared__ double tb[256 * U * K];
    __shared__ int tf[256 * U];
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.np;
    const int ti = own ? type[i] : 0;
    const double ai = P.ai[ti];
    const double aj = P.ai[ti];  // the reference uses i's type for aj (main.cpp:1863)
    const double VoL = P.V / P.l;
    const double pai = own ? pressA[i] : 0.0;
    const double gi0 = own ? gcen[3 * i] : 0.0, gi1 = own ? gcen[3 * i + 1] : 0.0, gi2 = own ? gcen[3 * i + 2] : 0.0;
    double fa[3] = {0.0, 0.0, 0.0}, fd[3] = {0.0, 0.0, 0.0};
    const long long b = own ? nPtr[i] : 0;
    const int n = own ? nCount[i] : 0;
    double pi[3] = {0.0, 0.0, 0.0};
    if (own)
        for (int d = 0; d < 3; ++d) pi[d] = pos[3 * i + d];
    // each lane takes U entries per round (base + u*G + g), their loads in flight
    // together and issued one round ahead for the columns; the group's first lane
    // adds the round's terms in list order
    int jnext[U];
    for (int u = 0; u < U; ++u) jnext[u] = g + u * G < n ? nInd[b + g + u * G] : 0;
    for (int base = 0; base < n; base += G * U) {
        int jcur[U], tjs[U];
        double pajs[U], pj[U][3], gjs[U][3];
        for (int u = 0; u < U; ++u) {
            const int jn = base + u * G + g;
            jcur[u] = jnext[u];
            jnext[u] = jn + G * U < n ? nInd[b + jn + G * U] : 0;
        }
        for (int u = 0; u < U; ++u) {  // j's values load with its position (past the list: particle 0)
            const int j = jcur[u];
            tjs[u] = type[j];
            pajs[u] = pressA[j];
            for (int d = 0; d < 3; ++d) {
                pj[u][d] = pos[3 * j + d];
                gjs[u][d] = kSurf ? gcen[3 * j + d] : 0.0;
            }
        }
        for (int u = 0; u < U; ++u) {
            const int jn = base + u * G + g;
            int f = 0;
            double t[K];
            for (int e = 0; e < K; ++e) t[e] = 0.0;
            if (jn < n) {
                const int j = jcur[u];
                const int tj = tjs[u];
                const double paj = pajs[u];
                const double gj[3] = {gjs[u][0], gjs[u][1], gjs[u][2]};
                Rel q;  // rel() on the loaded positions
                for (int d = 0; d < 3; ++d) q.x[d] = min_image(pj[u][d], pi[d], P.W[d], P.hW[d]);
                q.r2 = (q.x[0] * q.x[0] + q.x[1] * q.x[1] + q.x[2] * q.x[2]);
                if (i != j && q.r2 != 0.0) {
                    const double rij_ = P.ratio[ti][tj], rji_ = P.ratio[tj][ti];
                    if (P.radA2 - q.r2 > 0) {
                        const double r = sqrt(q.r2);
                        const double dwij = rij_ * dwadr_k(P.Ka, P.invA, r, P.radA);
                        const double dwji = rji_ * dwadr_k(P.Ka, P.invA, r, P.radA);
                        for (int d = 0; d < 3; ++d) t[d] = (pai * dwij + paj * dwji) * (q.x[d] / r) * P.V;
                        f |= 1;
                    }
                    if (kSurf && P.radG2 - q.r2 > 0) {
                        const double r = sqrt(q.r2);
                        const double wij = rij_ * wq_k(P.Kg, r, P.radG);
                        const double wji = rji_ * wq_k(P.Kg, r, P.radG);
                        const double gi[3] = {gi0, gi1, gi2};
                        for (int d = 0; d < 3; ++d)
                            t[3 + d] = (aj * gj[d] * wji - ai * gi[d] * wij) / P.R2g * P.radG * VoL;
                        const double dwij = rij_ * dwq_k(P.Kg, P.m2G, r, P.radG);
                        const double dwji = rji_ * dwq_k(P.Kg, P.m2G, r, P.radG);
                        double gr = 0.0;
                        for (int d = 0; d < 3; ++d) gr += (aj * gj[d] * dwji - ai * gi[d] * dwij) * q.x[d];
                        for (int d = 0; d < 3; ++d) t[6 + d] = (gr) * (q.x[d] / r) / P.R2g * P.radG * VoL;
                        f |= 2;
                    }
                }
            }
            for (int e = 0; e < K; ++e) tb[K * (U * threadIdx.x + u) + e] = t[e];
            tf[U * threadIdx.x + u] = f;
        }
        wave_sync();
        if (g == 0)
            for (int u = 0; u < U; ++u)
                for (int v = 0; v < G && base + u * G + v < n; ++v) {
                    const int sl = U * (threadIdx.x + v) + u;
                    const int fu = tf[sl];
                    const double* tu = tb + K * sl;
                    if (fu & 1)
                        for (int d = 0; d < 3; ++d) fa[d] += tu[d];
                    if (kSurf && (fu & 2)) {
                        for (int d = 0; d < 3; ++d) fd[d] -= tu[3 + d];
                        for (int d = 0; d < 3; ++d) fd[d] -= tu[6 + d];
                    }
                }
        wave_sync();
    }
    if (g != 0 || !own) return;
    double F[3];
    for (int d = 0; d < 3; ++d) {
        F[d] = force[3 * i + d];
        F[d] += fa[d];
        F[d] += fd[d];
    }
    if (i < P.nf) {
        const double m = mass[i];
        for (int d = 0; d < 3; ++d) {
            F[d] += m * P.gravity[d];
            vel[3 * i + d] += F[d] / m * P.dt;
        }
    }
    for (int d = 0; d < 3; ++d) force[3 * i + d] = F[d];
}

// calculateMatrixA values (2251-2335): viscous blocks, inertia, wall terms of b.
// kBlk: also the factored form's level-0 run sums (blk, from the run starts): while
// the group's first lane adds every entry's block in list order, its second lane adds
// the same fluid entries' -alpha r r^T (the negated block; 0 beyond RadiusV and for
// the self entry) per visited cell, and stores each cell's sum when its run ends
template <bool kBlk>
__global__ void __launch_bounds__(256, 4) k_matrix_a(DevParams P, const int* __restrict__ type,
                                                  const double* __restrict__ pos, const double* __restrict__ vel,
                                                  const double* __restrict__ mass, const double* __restrict__ mu,
                                                  const int* __restrict__ nCount, const int* __restrict__ nFluid,
                                                  const long long* __restrict__ nPtr, const int* __restrict__ nInd,
                                                  const long long* __restrict__ start, int* __restrict__ acol,
                                                  double* __restrict__ aval, int* __restrict__ selfIdx,
                                                  double* __restrict__ b, double* __restrict__ falpha,
                                                  double* __restrict__ fdiag, int* __restrict__ fwrap,
                                                  const unsigned short* __restrict__ runStart,
                                                  double* __restrict__ blk) {
    constexpr int G = kLanes, K = 12;  // terms: the 3x3 coefficients, then a wall neighbour's velocity
    __shared__ double tb[256 * K];
    __shared__ int tf[256];  // bit 0: the self entry, bit 1: inside RadiusV, bit 2: and a wall neighbour
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.nf;
    int iN = 0;
    // s9: the group's first lane sums the diagonal block in it; with kBlk its second
    // lane sums the current visited cell's run of -alpha r r^T (one register array)
    double s9[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double sv0 = 0.0, sv1 = 0.0, sv2 = 0.0;
    const double mui = own ? mu[i] : 0.0;
    const long long b0 = own ? nPtr[i] : 0;
    const int nfl = own ? nFluid[i] : 0, nc = own ? nCount[i] : 0;
    const long long sb = own ? start[i / kSlice] : 0;
    const int lane = i % kSlice;
    // factored form: alpha_ij = visc mu_ij (-w_v'(d)) / d^3, the viscous block being
    // -alpha_ij r r^T (beside the neighbour list; 0 for self and beyond RadiusV)
    auto store_alpha = [&](int k, double a) {
        if (falpha) __builtin_nontemporal_store(a, falpha + b0 + k);
    };
    // the 3x3 blocks are stored unless aval is null (a factored scene that builds its
    // multigrid level 0 from the factored form needs none)
    const bool blocks = aval != nullptr;
    // the run sums: visited cell qr's run of fluid entries is [rsp[qr], rsp[qr+1])
    typedef double dv2 __attribute__((ext_vector_type(2)));
    const bool runner = kBlk && g == 1 && own;
    // the group's run starts, staged in LDS by its four lanes (a flush then reads the
    // next run's end there, not from global memory in the serial loop)
    __shared__ unsigned short rsl[256 / G][kRuns];
    unsigned short* const rsp = rsl[threadIdx.x / G];
    if (kBlk && own)
        for (int k = g; k < kRuns; k += G) rsp[k] = runStart[(long long)i * kRuns + k];
    if (kBlk) wave_sync();
    double* const bk = kBlk ? blk + (long long)(own ? i : 0) * kBlkRuns * kBlkStride : nullptr;
    const int nv = (2 * P.range[0] + 1) * (2 * P.range[1] + 1) * (2 * P.range[2] + 1);
    int qr = 0, qend = runner ? rsp[1] : 0;
    auto flush = [&]() {  // cell qr's sums (upper triangle), then the next cell
        dv2* o = (dv2*)(bk + qr * kBlkStride);
        o[0] = dv2{s9[0], s9[1]};
        o[1] = dv2{s9[2], s9[4]};
        o[2] = dv2{s9[5], s9[8]};
        for (int e9 = 0; e9 < 9; ++e9) s9[e9] = 0.0;
        ++qr;
        qend = qr < nv ? rsp[qr + 1] : 0x7fffffff;
    };
    int jnext = g < nc ? nInd[b0 + g] : 0;  // each lane's next entry loads one round ahead
    for (int base = 0; base < nc; base += G) {
        const int jn = base + g;
        const int jcur = jnext;
        jnext = jn + G < nc ? nInd[b0 + jn + G] : 0;
        int f = 0;
        double t[K] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (jn < nc) {
            const int j = jcur;
            if (i == j) {
                f = 1;
                if (blocks) acol[a_col_index(sb, jn, lane)] = j;
                if (jn < nfl) store_alpha(jn, 0.0);
            } else {
                double alpha = 0.0;
                // j's values load with its position, not after the distance test
                const double muj = mu[j];
                const int tj = type[j];
                const Rel q = rel(P, pos, i, j);
                // factored form: does any pair interact through the periodic boundary
                if (fwrap)
                    for (int d = 0; d < 3; ++d)
                        if (fabs(pos[3 * j + d] - pos[3 * i + d]) > P.hW[d]) atomicOr(fwrap, 1);
                const bool inside = q.r2 != 0.0 && P.radV2 - q.r2 > 0;
                if (inside) {
                    const double dij = sqrt(q.r2);
                    const double wdrij = -dwq_k(P.Kv, P.m2V, dij, P.radV);
                    const double e[3] = {q.x[0] / dij, q.x[1] / dij, q.x[2] / dij};
                    const double muij = 2.0 * (mui * muj) / (mui + muj);
                    if (falpha) alpha = P.visc * muij * wdrij / (dij * dij * dij);
                    for (int r = 0; r < 3; ++r)
                        for (int sdim = 0; sdim < 3; ++sdim)
                            t[3 * r + sdim] = +P.visc * muij * e[sdim] * e[r] * wdrij / dij;
                    f = 2;
                    if (is_wall(tj)) {
                        t[9] = vel[3 * j];
                        t[10] = vel[3 * j + 1];
                        t[11] = vel[3 * j + 2];
                        f |= 4;
                    }
                }
                if (jn < nfl) {
                    store_alpha(jn, alpha);
                    if (blocks) {
                        acol[a_col_index(sb, jn, lane)] = j;
                        for (int e9 = 0; e9 < 9; ++e9) aval[a_val_index(sb, jn, e9, lane)] = inside ? -t[e9] : 0.0;
                    }
                }
            }
        }
        for (int e = 0; e < K; ++e) tb[K * threadIdx.x + e] = t[e];
        tf[threadIdx.x] = f;
        wave_sync();
        // lanes 0 and 1 walk the same entries together: lane 0 adds each block, lane 1
        // (runner) adds its negation, exactly -1.0 times it, per visited cell
        if (g == 0 || runner) {
            const double sgn = g == 0 ? 1.0 : -1.0;
            for (int u = 0; u < G && base + u < nc; ++u) {
                const int sl = threadIdx.x - g + u;
                const int fu = tf[sl];
                const double* c9 = tb + K * sl;
                if (g != 0) {
                    if (base + u >= nfl) break;  // the runs cover the fluid entries
                    while (base + u >= qend) flush();
                }
                if (fu & 2)
                    for (int e9 = 0; e9 < 9; ++e9) s9[e9] += sgn * c9[e9];
                if (g == 0) {
                    if (fu & 1) iN = base + u;
                    if (fu & 4) {
                        const double v0 = c9[9], v1 = c9[10], v2 = c9[11];
                        sv0 += c9[0] * v0; sv0 += c9[1] * v1; sv0 += c9[2] * v2;
                        sv1 += c9[3] * v0; sv1 += c9[4] * v1; sv1 += c9[5] * v2;
                        sv2 += c9[6] * v0; sv2 += c9[7] * v1; sv2 += c9[8] * v2;
                    }
                }
            }
        }
        wave_sync();
    }
    if (runner)
        while (qr < nv) flush();  // the cells after the last fluid entry's run
    if (g != 0 || !own) return;
    const double coef = mass[i] / P.V / P.dt;
    s9[0] += coef;
    s9[4] += coef;
    s9[8] += coef;
    if (blocks)
        for (int e9 = 0; e9 < 9; ++e9) aval[a_val_index(sb, iN, e9, lane)] = s9[e9];
    if (fdiag)
        for (int e9 = 0; e9 < 9; ++e9) fdiag[9 * i + e9] = s9[e9];
    selfIdx[i] = iN;
    const double bv[3] = {0.0 + sv0, 0.0 + sv1, 0.0 + sv2};
    for (int d = 0; d < 3; ++d) b[3 * i + d] = bv[d] + coef * vel[3 * i + d];
}

// calculateMatrixC values (2463-2507): C[3i+r][k] = e_r * (-w'_p).  Also
// D[3i+r][k] = C[3i+r][k] * Lambda[k], the left factor of every product of the
// C Lambda C^T sum (main.cpp:2614 evaluates (C_ik * Lambda_k) * C_jk).
__global__ void __launch_bounds__(256) k_matrix_c(DevParams P, const double* __restrict__ pos,
                                                  const int* __restrict__ nCountP, const long long* __restrict__ nPtrP,
                                                  const int* __restrict__ nIndP, const double* __restrict__ lambda,
                                                  double* __restrict__ cval, double* __restrict__ dval) {
    constexpr int G = kLanesP, K = 3;  // terms: the C block of an entry
    __shared__ double tb[256 * K];
    __shared__ int tf[256];  // bit 0: the self entry, bit 1: inside RadiusP
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.nf;
    int iN = 0;
    double self[3] = {0.0, 0.0, 0.0};
    const long long b0 = own ? nPtrP[i] : 0;
    const int n = own ? nCountP[i] : 0;
    for (int base = 0; base < n; base += G) {
        const int jn = base + g;
        int f = 0;
        double c3[3] = {0.0, 0.0, 0.0};
        if (jn < n) {
            const int j = nIndP[b0 + jn];
            if (i == j) {
                f = 1;
            } else {
                const Rel q = rel(P, pos, i, j);
                if (q.r2 != 0.0 && P.radP2 - q.r2 > 0) {
                    const double dij = sqrt(q.r2);
                    const double e[3] = {q.x[0] / dij, q.x[1] / dij, q.x[2] / dij};
                    const double wpdrij = -dwq_k(P.Kp, P.m2P, dij, P.radP);
                    for (int r = 0; r < 3; ++r) c3[r] = e[r] * wpdrij;
                    f = 2;
                }
                // the self entry's block is written once its sum is complete
                for (int r = 0; r < 3; ++r) cval[3 * (b0 + jn) + r] = c3[r];
                if (dval) {  // D = C Lambda, read only by C Lambda C^T (assembled form)
                    const double lk = lambda[j];
                    for (int r = 0; r < 3; ++r) dval[3 * (b0 + jn) + r] = c3[r] * lk;
                }
            }
        }
        for (int e = 0; e < K; ++e) tb[K * threadIdx.x + e] = c3[e];
        tf[threadIdx.x] = f;
        wave_sync();
        if (g == 0)
            for (int u = 0; u < G && base + u < n; ++u) {
                const int sl = threadIdx.x + u;
                const int fu = tf[sl];
                if (fu & 1) iN = base + u;
                if (fu & 2)
                    for (int r = 0; r < 3; ++r) self[r] += tb[K * sl + r];
            }
        wave_sync();
    }
    if (g != 0 || !own) return;
    for (int r = 0; r < 3; ++r) cval[3 * (b0 + iN) + r] = self[r];
    if (dval) {
        const double li = lambda[i];
        for (int r = 0; r < 3; ++r) dval[3 * (b0 + iN) + r] = self[r] * li;
    }
}

// VectorP, main.cpp:2510-2549
__global__ void k_vector_p(DevParams P, const int* __restrict__ type, const double* __restrict__ pos,
                           const double* __restrict__ vel, const double* __restrict__ kappa,
                           const double* __restrict__ lambda, const double* __restrict__ volStrainP,
                           const int* __restrict__ nCountP, const long long* __restrict__ nPtrP,
                           const int* __restrict__ nIndP, double* __restrict__ vecP) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    double pv = 0.0;
    if (volStrainP[i] > 0.0) pv = kappa[i] * volStrainP[i];
    const bool walli = is_wall(type[i]);
    const double li = lambda[i];
    const double pi[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    double sum = 0.0;
    const long long b0 = nPtrP[i];
    const int n = nCountP[i];
    // U entries at a time: columns, then types and positions, load together (past the
    // end: the last entry again, not visited); a pair adds terms only when i or j is a
    // wall, so fluid pairs stop after the distance test
    constexpr int U = 4;
    for (int jb = 0; jb < n; jb += U) {
        int jj[U], tj[U];
        double pj[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = nIndP[b0 + (jb + u < n ? jb + u : n - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            tj[u] = type[jj[u]];
            for (int d = 0; d < 3; ++d) pj[u][d] = pos[3 * jj[u] + d];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = jj[u];
            if (jb + u >= n || i == j) continue;
            Rel q;  // rel() on the loaded positions
            for (int d = 0; d < 3; ++d) q.x[d] = min_image(pj[u][d], pi[d], P.W[d], P.hW[d]);
            q.r2 = (q.x[0] * q.x[0] + q.x[1] * q.x[1] + q.x[2] * q.x[2]);
            if (q.r2 == 0.0) continue;
            const bool wallj = is_wall(tj[u]);
            if (P.radP2 - q.r2 > 0 && (walli || wallj)) {
                const double dij = sqrt(q.r2);
                const double e[3] = {q.x[0] / dij, q.x[1] / dij, q.x[2] / dij};
                const double wpdrij = -dwq_k(P.Kp, P.m2P, dij, P.radP);
                for (int s = 0; s < 3; ++s) {
                    if (walli) {
                        const double coef = +e[s] * wpdrij;
                        sum += li * coef * vel[3 * i + s];
                    }
                    if (wallj) {
                        const double coef = -e[s] * wpdrij;
                        sum += li * coef * vel[3 * j + s];
                    }
                }
            }
        }
    }
    vecP[i] = pv + sum;
}

// b -= C P (multiplyMatrixC, 2575-2584), accumulated directly into b
__global__ void __launch_bounds__(256) k_b_minus_cp(DevParams P, const int* __restrict__ nCountP,
                                                    const long long* __restrict__ nPtrP,
                                                    const int* __restrict__ nIndP, const double* __restrict__ cval,
                                                    const double* __restrict__ vecP, double* __restrict__ b) {
    constexpr int G = kLanesP, K = 3;  // terms: c_r * P_k
    __shared__ double tb[256 * K];
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.nf;
    double b0 = own ? b[3 * i] : 0.0, b1 = own ? b[3 * i + 1] : 0.0, b2 = own ? b[3 * i + 2] : 0.0;
    const long long p0 = own ? nPtrP[i] : 0;
    const int n = own ? nCountP[i] : 0;
    for (int base = 0; base < n; base += G) {
        const int jn = base + g;
        double t[K] = {0.0, 0.0, 0.0};
        if (jn < n) {
            const double pk = vecP[nIndP[p0 + jn]];
            const double* c = &cval[3 * (p0 + jn)];
            for (int r = 0; r < 3; ++r) t[r] = c[r] * pk;
        }
        for (int e = 0; e < K; ++e) tb[K * threadIdx.x + e] = t[e];
        wave_sync();
        if (g == 0)
            for (int u = 0; u < G && base + u < n; ++u) {
                const double* tu = tb + K * (threadIdx.x + u);
                b0 -= tu[0];
                b1 -= tu[1];
                b2 -= tu[2];
            }
        wave_sync();
    }
    if (g != 0 || !own) return;
    b[3 * i] = b0;
    b[3 * i + 1] = b1;
    b[3 * i + 2] = b2;
}

// A += C Lambda C^T (multiplyMatrixC, 2586-2642): for each fluid i and each of its
// fluid neighbours j, merge the sorted P-neighbour lists of i and j and sum
// D_ik (x) C_jk over the shared k (D = C*Lambda, left to right as the reference),
// then add the 3x3 sum to block (i, j).  Workgroup = one slice: lane = particle,
// wave w takes the blocks k = w, w+4, ...  kMode >= 2: the slice's own index lists
// are staged in LDS (odd row stride, no bank conflicts) and each thread copies j's
// index list into LDS before merging (its loads all in flight at once instead of
// one per merge step); with kMode 0 both are read from global memory;
// kMode 3 in addition first merges without arithmetic, recording the matches, and
// then adds them (the merge's three-way branch no longer carries the 9 products).
// All give the same bits; the fastest is picked at run time.
// The FP32 copy of the finished block is written when aval32 is non-null.
template <int kMode>
__global__ void __launch_bounds__(256) k_cac(DevParams P, const int* __restrict__ nFluid,
                                             const int* __restrict__ nCountP, const long long* __restrict__ nPtrP,
                                             const int* __restrict__ nIndP, const double* __restrict__ cval,
                                             const double* __restrict__ dval, const long long* __restrict__ start,
                                             const int* __restrict__ acol, double* __restrict__ aval,
                                             float* __restrict__ aval32) {
    constexpr int kStride = kMaxP + 1;
    constexpr bool kLdsI = kMode >= 2, kLdsJ = kMode >= 2;  // mode 3: both + two-phase merge
    __shared__ int sk[kLdsI ? kSlice * kStride : 1];
    __shared__ int sj[kLdsJ ? 256 * kStride : 1];
    const int slice = blockIdx.x;
    const int base = slice * kSlice;
    if (kLdsI) {
        for (int t = threadIdx.x; t < kSlice * kMaxP; t += blockDim.x) {
            const int l = t / kMaxP, a = t % kMaxP, i = base + l;
            if (i < P.nf && a < nCountP[i]) sk[l * kStride + a] = nIndP[nPtrP[i] + a];
        }
        __syncthreads();
    }
    const int lane = threadIdx.x % kSlice, wave = threadIdx.x / kSlice;
    const int i = base + lane;
    if (i >= P.nf) return;
    const int nfl = nFluid[i];
    const long long pi = nPtrP[i];
    const int ni = nCountP[i];
    const int* ki = kLdsI ? sk + lane * kStride : nIndP + pi;
    int* kjs = sj + threadIdx.x * kStride;  // this thread's copy of j's list (kLdsJ)
    const double* di = dval + 3 * pi;
    const long long sb = start[slice];
    for (int k = wave; k < nfl; k += 4) {
        const int j = acol[a_col_index(sb, k, lane)];
        const long long pj = nPtrP[j];
        const int nj = nCountP[j];
        const int* kj = nIndP + pj;
        if (kLdsJ) {  // all loads of j's list in flight at once, then merge from LDS
            for (int t = 0; t < nj; ++t) kjs[t] = kj[t];
            kj = kjs;
        }
        double sum[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        auto add = [&](int a, int c) {  // sum += D_i,k (x) C_j,k for the shared k
            const double* cj = &cval[3 * (pj + c)];
            const double c0 = cj[0], c1 = cj[1], c2 = cj[2];
            const double d0 = di[3 * a], d1 = di[3 * a + 1], d2 = di[3 * a + 2];
            sum[0] += d0 * c0; sum[1] += d0 * c1; sum[2] += d0 * c2;
            sum[3] += d1 * c0; sum[4] += d1 * c1; sum[5] += d1 * c2;
            sum[6] += d2 * c0; sum[7] += d2 * c1; sum[8] += d2 * c2;
        };
        if (kMode == 3) {
            // merge without work in the loop: the matches (a, c), ascending, go over
            // the already consumed front of j's list in LDS; then add them in order
            int a = 0, c = 0, n = 0;
            int ka = ni > 0 ? ki[0] : 0;
            int kc = nj > 0 ? kjs[0] : 0;
            while (a < ni && c < nj) {
                const bool eq = ka == kc, le = ka <= kc, ge = ka >= kc;
                if (eq) kjs[n++] = a | (c << 16);
                a += le;
                c += ge;
                if (le && a < ni) ka = ki[a];
                if (ge && c < nj) kc = kjs[c];
            }
            for (int t = 0; t < n; ++t) add(kjs[t] & 0xffff, kjs[t] >> 16);
        } else {
            int a = 0, c = 0;
            int ka = ni > 0 ? ki[0] : 0;
            int kc = nj > 0 ? kj[0] : 0;
            while (a < ni && c < nj) {
                if (ka == kc) {
                    add(a, c);
                    a++;
                    c++;
                    if (a < ni) ka = ki[a];
                    if (c < nj) kc = kj[c];
                } else if (ka < kc) {
                    a++;
                    if (a < ni) ka = ki[a];
                } else {
                    c++;
                    if (c < nj) kc = kj[c];
                }
            }
        }
        for (int e9 = 0; e9 < 9; ++e9) {
            const long long at = a_val_index(sb, k, e9, lane);
            const double v = aval[at] + sum[e9];
            aval[at] = v;
            if (aval32) aval32[a_val32_index(sb, k, e9, lane)] = (float)v;
        }
    }
}


}  // namespace

void launch_kinematics(hipStream_t st, const DevParams& P, const WallKin& wk, const int* type, double* pos,
                       double* vel, double* force) {
    hipLaunchKernelGGL(k_kinematics, dim3(nblk(P.np, 256)), dim3(256), 0, st, P, wk, type, pos, vel, force);
}

void launch_densities(hipStream_t st, const DevParams& P, const int* type, const double* pos, const int* nCount,
                      const long long* nPtr, const int* nInd, double* densA, double* gcen, double* volStrainP) {
    hipLaunchKernelGGL(k_densities, dim3(nblk((long long)kLanes * P.np, 256)), dim3(256), 0, st, P, type, pos, nCount,
                       nPtr, nInd, densA, gcen, volStrainP);
}

void launch_coefficients(hipStream_t st, const DevParams& P, const int* type, const double* densA,
                         const double* volStrainP, double* mass, double* kappa, double* lambda, double* mu,
                         double* pressA) {
    hipLaunchKernelGGL(k_coefficients, dim3(nblk(P.np, 256)), dim3(256), 0, st, P, type, densA, volStrainP, mass,
                       kappa, lambda, mu, pressA);
}

void launch_forces(hipStream_t st, const DevParams& P, const int* type, const double* pos, const int* nCount,
                   const long long* nPtr, const int* nInd, const double* pressA, const double* gcen,
                   const double* mass, double* force, double* vel) {
    const bool surf = !(P.ai[0] == 0.0 && P.ai[1] == 0.0 && P.ai[2] == 0.0 && P.ai[3] == 0.0);
    if (surf)
        hipLaunchKernelGGL(k_forces<true>, dim3(nblk((long long)kLanes * P.np, 256)), dim3(256), 0, st, P, type, pos,
                           nCount, nPtr, nInd, pressA, gcen, mass, force, vel);
    else
        hipLaunchKernelGGL(k_forces<false>, dim3(nblk((long long)kLanes * P.np, 256)), dim3(256), 0, st, P, type, pos,
                           nCount, nPtr, nInd, pressA, gcen, mass, force, vel);
}

void launch_matrix_a(hipStream_t st, const DevParams& P, const int* type, const double* pos, const double* vel,
                     const double* mass, const double* mu, const int* nCount, const int* nFluid,
                     const long long* nPtr, const int* nInd, ASlices A, int* selfIdx, double* b, double* alpha,
                     double* diag, int* wrap, const unsigned short* runStart, double* blk) {
    const dim3 grid(nblk((long long)kLanes * P.nf, 256));
    if (blk && runStart)
        hipLaunchKernelGGL(k_matrix_a<true>, grid, dim3(256), 0, st, P, type, pos, vel, mass, mu, nCount, nFluid, nPtr,
                           nInd, A.start, A.col, A.val, selfIdx, b, alpha, diag, wrap, runStart, blk);
    else
        hipLaunchKernelGGL(k_matrix_a<false>, grid, dim3(256), 0, st, P, type, pos, vel, mass, mu, nCount, nFluid,
                           nPtr, nInd, A.start, A.col, A.val, selfIdx, b, alpha, diag, wrap, nullptr, nullptr);
}

void launch_matrix_c(hipStream_t st, const DevParams& P, const double* pos, const int* nCountP,
                     const long long* nPtrP, const int* nIndP, const double* lambda, double* cval, double* dval) {
    hipLaunchKernelGGL(k_matrix_c, dim3(nblk((long long)kLanesP * P.nf, 256)), dim3(256), 0, st, P, pos, nCountP,
                       nPtrP, nIndP, lambda, cval, dval);
}

void launch_vector_p(hipStream_t st, const DevParams& P, const int* type, const double* pos, const double* vel,
                     const double* kappa, const double* lambda, const double* volStrainP, const int* nCountP,
                     const long long* nPtrP, const int* nIndP, double* vecP) {
    hipLaunchKernelGGL(k_vector_p, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, type, pos, vel, kappa, lambda,
                       volStrainP, nCountP, nPtrP, nIndP, vecP);
}

void launch_b_minus_cp(hipStream_t st, const DevParams& P, const int* nCountP, const long long* nPtrP,
                       const int* nIndP, const double* cval, const double* vecP, double* b) {
    hipLaunchKernelGGL(k_b_minus_cp, dim3(nblk((long long)kLanesP * P.nf, 256)), dim3(256), 0, st, P, nCountP, nPtrP,
                       nIndP, cval, vecP, b);
}

void launch_cac(hipStream_t st, const DevParams& P, const int* nFluid, const int* nCountP, const long long* nPtrP,
                const int* nIndP, const double* cval, const double* dval, ASlices A, int mode) {
    if (A.nslices == 0) return;
#define MPH_CAC(M)                                                                                            \
    hipLaunchKernelGGL(k_cac<M>, dim3(A.nslices), dim3(256), 0, st, P, nFluid, nCountP, nPtrP, nIndP, cval, dval, \
                       A.start, A.col, A.val, A.val32)
    if (mode == 3) MPH_CAC(3);
    else if (mode == 2) MPH_CAC(2);
    else MPH_CAC(0);
#undef MPH_CAC
}

}  // namespace mph
