// mph_factored.hip -- the particle matrix of the implicit solve applied in factored
// form, reading one scalar per stored block instead of nine.
//
// calculateMatrixA and multiplyMatrixC (source/main.cpp:2137-2336, 2562-2644) build,
// for fluid particles i and fluid neighbours j != i,
//   A_ij = -alpha_ij r_ij r_ij^T + sum_k C_ik lambda_k C_jk^T
//   A_ii = D_i + sum_k C_ik lambda_k C_ik^T,  D_i = V_i + m_i/(V dt) I
// where r_ij is the minimum-image displacement pos_j - pos_i, and:
//   - alpha_ij = visc mu_ij (-w_v'(d))/d^3 for 0 < d < RadiusV, 0 otherwise. The
//     viscous coefficient visc mu_ij e_s e_r (-w_v'(d))/d of main.cpp:2279-2288,
//     with e = r/d.
//   - V_i = sum_j alpha_ij r_ij r_ij^T over every neighbour, walls included. This is
//     the viscous self block, main.cpp:2290.
//   - C_ik = beta_ik r_ik, with beta_ik = -w_p'(d)/d for 0 < d < RadiusP. This is
//     C's entry e (-w_p'(d)) of main.cpp:2479-2493.
//   - C_ii = sum_k C_ik (main.cpp:2497-2504).
//   - k runs over the RadiusP neighbours shared by i and j, fluid or wall.
// So A x needs only alpha per A slot, beta per P-list entry, D and C_ii:
//   t_k = lambda_k sum_{j fluid, k in P(j)} C_jk . x_j        (every particle k)
//   (A x)_i = D_i x_i - sum_j alpha_ij r_ij (r_ij . x_j) + C_ii t_i + sum_{k != i} C_ik t_k
// The first pass sums over k's own P list, using C_jk = -beta_kj r_kj.  That holds
// because the RadiusP relation is symmetric up to rounding in the distance.
//
// The product is the assembled one up to rounding: the same terms, associated
// differently.  It reads about 12 bytes per block (a 4-byte column and an 8-byte
// alpha) where the assembled product reads 76.  A neighbour's x, position and t are
// gathered through the caches from planes of 16-byte pairs (FacOp::rec): three
// 16-byte loads per entry, where separate arrays would take six or four 8-byte
// gathers.  Each row is walked by a group of 8 lanes, each with 4 entries in flight,
// so that the gathers of one load instruction fall on consecutive particles of the
// cell-ordered list, and so on consecutive pairs of a plane.  Measured on a jittered
// 100^3 lattice (172 entries per row), every arrangement of this pass with 64-byte
// records per particle (one lane per row, groups of 8 to 64 lanes, with or without
// an XCD-aware block order) landed between 0.94 and 1.3 ms; the planes take it to
// 0.90 ms.  The cache traffic of the gathers, not HBM, bounds it.
//
// Scenes whose multigrid levels allow it assemble no 3x3 blocks at all: level 0
// and InvDiagA come from the same scalars (k_fac_g, k_fac_gal0, k_fac_invdiag).
#include "mph_batch.h"
#include "mph_cr_body.h"

namespace mph {
namespace {

using body::nblk;

// beta beside the P lists (and v = beta r in FP32 for the FP32 form), the extent of
// their fluid entries, C_ii of fluid particles, and the records (positions; x and t
// zero until the first pass)
__global__ void __launch_bounds__(256) k_fac_p(DevParams P, const double* __restrict__ pos,
                                               const int* __restrict__ nCountP, const long long* __restrict__ nPtrP,
                                               const int* __restrict__ nIndP, double* __restrict__ beta,
                                               float* __restrict__ v32, int* __restrict__ runP,
                                               double* __restrict__ selfC, double* __restrict__ rec) {
    // entry-parallel: lane g handles entries g, g + 8, ... (beta, v); the group's
    // first lane adds the C_ii terms in list order and finds the fluid extent
    constexpr int G = 8, K = 3;
    __shared__ double tb[256 * K];
    __shared__ int tf[256];  // bit 0: inside RadiusP (a C_ii term), bit 1: fluid neighbour
    const int g = threadIdx.x & (G - 1);
    const int k = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = k < P.np;
    const long long b0 = own ? nPtrP[k] : 0;
    const int n = own ? nCountP[k] : 0;
    const double pk0 = own ? pos[3 * k] : 0.0, pk1 = own ? pos[3 * k + 1] : 0.0, pk2 = own ? pos[3 * k + 2] : 0.0;
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    int run = 0;
    for (int base = 0; base < n; base += G) {
        const int jn = base + g;
        int f = 0;
        double t[K] = {0.0, 0.0, 0.0};
        if (jn < n) {
            const int j = nIndP[b0 + jn];
            double b = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
            if (j != k) {
                body::fac_disp(P, pos, j, pk0, pk1, pk2, r0, r1, r2);
                const double r2s = (r0 * r0 + r1 * r1 + r2 * r2);
                if (r2s != 0.0 && P.radP2 - r2s > 0) {
                    const double d = sqrt(r2s);
                    b = -dwq_k(P.Kp, P.m2P, d, P.radP) / d;
                    t[0] = b * r0;
                    t[1] = b * r1;
                    t[2] = b * r2;
                    f = 1;
                }
            }
            if (j < P.nf) f |= 2;
            beta[b0 + jn] = b;
            if (v32) {
                typedef float fv4 __attribute__((ext_vector_type(4)));
                ((fv4*)v32)[b0 + jn] = fv4{(float)(b * r0), (float)(b * r1), (float)(b * r2), __int_as_float(j)};
            }
        }
        for (int e = 0; e < K; ++e) tb[K * threadIdx.x + e] = t[e];
        tf[threadIdx.x] = f;
        wave_sync();
        if (g == 0)
            for (int u = 0; u < G && base + u < n; ++u) {
                const int sl = threadIdx.x + u;
                const int fu = tf[sl];
                if (fu & 1) {
                    c0 += tb[K * sl];
                    c1 += tb[K * sl + 1];
                    c2 += tb[K * sl + 2];
                }
                if (fu & 2) run = base + u + 1;
            }
        wave_sync();
    }
    if (g != 0 || !own) return;
    runP[k] = run;
    const body::FacPlanes R = body::fac_planes(P, rec);
    R.a[2 * k] = 0.0;
    R.a[2 * k + 1] = 0.0;
    R.b[2 * k] = 0.0;
    R.b[2 * k + 1] = pk0;
    R.c[2 * k] = pk1;
    R.c[2 * k + 1] = pk2;
    R.t[k] = 0.0;
    if (k < P.nf) {
        selfC[3 * k] = c0;
        selfC[3 * k + 1] = c1;
        selfC[3 * k + 2] = c2;
    }
}

// u = sqrt(alpha) r per fluid-neighbour entry of fluid rows, eight lanes per row
// (coalesced writes beside the list)
__global__ void __launch_bounds__(256) k_fac_u(DevParams P, const int* __restrict__ nFluid,
                                               const long long* __restrict__ nPtr, const int* __restrict__ nInd,
                                               const double* __restrict__ alpha, const double* __restrict__ pos,
                                               float* __restrict__ u32) {
    constexpr int G = 8;
    const int g = threadIdx.x & (G - 1);
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    if (i >= P.nf) return;
    const double pi0 = pos[3 * i], pi1 = pos[3 * i + 1], pi2 = pos[3 * i + 2];
    const long long b0 = nPtr[i];
    const int w = nFluid[i];
    for (int jn = g; jn < w; jn += G) {
        const int j = nInd[b0 + jn];
        double r0, r1, r2;
        body::fac_disp(P, pos, j, pi0, pi1, pi2, r0, r1, r2);
        const double s = sqrt(alpha[b0 + jn]);  // alpha >= 0: visc, mu and -w' are not negative
        typedef float fv4 __attribute__((ext_vector_type(4)));
        ((fv4*)u32)[b0 + jn] = fv4{(float)(s * r0), (float)(s * r1), (float)(s * r2), __int_as_float(j)};
    }
}

// ---------------------------------------------------------------------------
// Multigrid level 0 and the particle-level inverse diagonal from the factored form,
// so a factored scene never assembles the 3x3 blocks.  Level 0 of
// calculateMultiGridCsrA (main.cpp:3336-3410) is, for row cell g and stencil cell h,
//   A0(g, h) = sum_{i in g} sum_{j in h} A_ij
//            = sum_{i in g} [D_i if h = g] - sum_{i in g, j in h, j != i} alpha_ij r r^T
//              + sum_k lambda_k G_gk G_hk^T,     G_gk = sum_{i in g, k in P(i)} C_ik,
// k over every particle with fluid P neighbours in both cells.  A particle's P
// neighbours lie in its own or the adjacent cells (cells are at least MaxRadius >=
// 2 RadiusP wide), so G_k is a 27-slot table around k's cell.

// slot of cell (tx, ty, tz) in the 3x3x3 block around (kx, ky, kz), -1 if outside
__device__ inline int fac_slot(const DevParams& P, int tx, int ty, int tz, int kx, int ky, int kz) {
    int d[3] = {tx - kx, ty - ky, tz - kz};
    for (int a = 0; a < 3; ++a) {
        if (d[a] > 1) d[a] -= P.C[a];
        if (d[a] < -1) d[a] += P.C[a];
        if (d[a] < -1 || d[a] > 1) return -1;
    }
    return 9 * (d[0] + 1) + 3 * (d[1] + 1) + (d[2] + 1);
}

__device__ inline void cell_xyz(const DevParams& P, int key, int& x, int& y, int& z) {
    const int c = key % P.T;  // wall keys carry +T
    x = c / (P.C[1] * P.C[2]);
    y = (c / P.C[2]) % P.C[1];
    z = c % P.C[2];
}

// G_k (27 slots x 3; only the slots in the mask, those where a fluid P neighbour
// lies, are written) and the mask, summed over k's P list in order; C_kk = C_ii of
// fluid k, C_jk = -beta r_kj
__global__ void k_fac_g(DevParams P, const int* __restrict__ cellIndex, const int* __restrict__ runP,
                        const long long* __restrict__ nPtrP, const int* __restrict__ nIndP,
                        const double* __restrict__ beta, const double* __restrict__ pos,
                        const double* __restrict__ selfC, double* __restrict__ gtab, int* __restrict__ gmask) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.np) return;
    double* G = gtab + 81 * (long long)k;
    int kx, ky, kz;
    cell_xyz(P, cellIndex[k], kx, ky, kz);
    const double pk0 = pos[3 * k], pk1 = pos[3 * k + 1], pk2 = pos[3 * k + 2];
    const long long c0 = nPtrP[k];
    const int run = runP[k];
    unsigned mask = 0;
    // the list visits each neighbour cell in one run, so a slot's sum is kept in
    // registers and stored when the run ends; slots outside the mask are never
    // read and stay unwritten (a slot met again would continue from its stored sum)
    int cur = -1;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    // U entries at a time: their columns and beta, then their cell keys and positions,
    // load together (a wall or past-the-end entry loads k's own); the entries are
    // then added one by one in list order
    constexpr int U = 4;
    for (int jb = 0; jb < run; jb += U) {
        int jj[U], key[U];
        double bb[U], q[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long e = c0 + (jb + u < run ? jb + u : run - 1);
            jj[u] = nIndP[e];
            bb[u] = beta[e];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = jj[u] < P.nf ? jj[u] : k;
            key[u] = cellIndex[j];
            for (int d = 0; d < 3; ++d) q[u][d] = pos[3 * j + d];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = jj[u];
            if (jb + u >= run || j >= P.nf) continue;
            int jx, jy, jz;
            cell_xyz(P, key[u], jx, jy, jz);
            const int slot = fac_slot(P, jx, jy, jz, kx, ky, kz);
            if (slot < 0) continue;  // cannot happen: P neighbours are in adjacent cells
            if (slot != cur) {
                if (cur >= 0) {
                    G[3 * cur] = a0;
                    G[3 * cur + 1] = a1;
                    G[3 * cur + 2] = a2;
                }
                cur = slot;
                const bool seen = (mask >> slot) & 1u;
                a0 = seen ? G[3 * slot] : 0.0;
                a1 = seen ? G[3 * slot + 1] : 0.0;
                a2 = seen ? G[3 * slot + 2] : 0.0;
            }
            double c[3];
            if (j == k) {
                c[0] = selfC[3 * k];
                c[1] = selfC[3 * k + 1];
                c[2] = selfC[3 * k + 2];
            } else {  // body::fac_disp on the loaded position
                const double r0 = min_image(q[u][0], pk0, P.W[0], P.hW[0]);
                const double r1 = min_image(q[u][1], pk1, P.W[1], P.hW[1]);
                const double r2 = min_image(q[u][2], pk2, P.W[2], P.hW[2]);
                const double b = bb[u];
                c[0] = -(b * r0);
                c[1] = -(b * r1);
                c[2] = -(b * r2);
            }
            a0 += c[0];
            a1 += c[1];
            a2 += c[2];
            mask |= 1u << slot;
        }
    }
    if (cur >= 0) {
        G[3 * cur] = a0;
        G[3 * cur + 1] = a1;
        G[3 * cur + 2] = a2;
    }
    gmask[k] = (int)mask;
}

// level-0 row cell g (workgroup = one wavefront): the 3x3 blocks A0(g, c) of its
// ncs stencil cells.  Lane l works on stencil cell c = l % ncs in partition
// l / ncs; the 64 / ncs partitions split the row cell's particles (viscous part)
// and the candidate list (C Lambda C^T part) round robin, and their sums are added
// in partition order at the end.  The viscous part and the diagonal blocks come
// from the run sums of the matrix-A pass (blk): a particle's record of each visited
// cell that falls in stencil cell c.  (Walking the runs entry by entry here took
// 3.95 ms of 4.7 on the 1M scene: a lane's particles and entries one after another,
// with the centre cell's runs the longest.)  The candidates are the particles of the
// cells around g whose G table holds g, filtered 64 at a time into an ordered list in
// LDS (flushed when full).  The list and the partition sums share one LDS buffer.
constexpr int kGal0List = 288;
__global__ void __launch_bounds__(64) k_fac_gal0(DevParams P, Level L, FacOp F, const int* __restrict__ cellStart,
                                                 const double* __restrict__ blk, const double* __restrict__ gtab,
                                                 const int* __restrict__ gmask) {
    typedef double dv2 __attribute__((ext_vector_type(2)));
    __shared__ double sbuf[64 * 9];   // 4608 B: the list, then the partition sums
    static_assert(4 * kGal0List * sizeof(int) <= sizeof(sbuf), "list exceeds the shared buffer");
    int* lk = (int*)sbuf;             // candidate particle
    int* lc = lk + kGal0List;         // its cell's offset from g, (dx+1) + 3 (dy+1) + 9 (dz+1)
    int* lm = lc + kGal0List;         // its G mask
    unsigned* lh = (unsigned*)(lm + kGal0List);  // the stencil cells its G mask reaches (a superset)
    const int ncs = L.cs[0] * L.cs[1] * L.cs[2];
    const int nparts = 64 / ncs;
    const int g = blockIdx.x, lane = threadIdx.x;
    const int c = lane % ncs, part = lane / ncs;
    const bool active = part < nparts;
    int gx, gy, gz;
    body::grid_xyz(L, g, gx, gy, gz);
    const int ax = gx + L.mn[0], ay = gy + L.mn[1], az = gz + L.mn[2];
    const int ic = body::cell_id(P, ax, ay, az);
    const int p0 = cellStart[ic], p1 = cellStart[ic + 1];
    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // viscous part and the diagonal blocks: the run sums of the row cell's particles
    // in the visited cells that fall in stencil cell c, then the particle's diagonal
    // block where c holds the row cell itself
    const unsigned vmask = active ? body::galerkin_mask(P, L, cellStart, g, c) : 0u;
    const bool dself = (vmask >> body::galerkin_self(P, L, g)) & 1u;
    if (vmask)
        for (int ip = p0 + part; ip < p1; ip += nparts) {
            const double* rb = blk + (long long)ip * kBlkRuns * kBlkStride;
            for (unsigned m = vmask; m; m &= m - 1) {  // symmetric: the upper triangle twice
                const dv2* o = (const dv2*)(rb + __builtin_ctz(m) * kBlkStride);
                const dv2 v0 = o[0], v1 = o[1], v2 = o[2];
                acc[0] += v0.x;
                acc[1] += v0.y;
                acc[2] += v1.x;
                acc[3] += v0.y;
                acc[4] += v1.y;
                acc[5] += v2.x;
                acc[6] += v1.x;
                acc[7] += v2.x;
                acc[8] += v2.y;
            }
            if (dself)
                for (int e = 0; e < 9; ++e) acc[e] += F.diag[9 * ip + e];
        }
    // C Lambda C^T part: lambda_k G_gk G_hk^T; h is stencil cell c (level-0 counts
    // of at least 3 per axis: the caller checks)
    const int iX = c / (L.cs[1] * L.cs[2]), iY = (c / L.cs[2]) % L.cs[1], iZ = c % L.cs[2];
    const int* C = P.C;
    const int hx = (((L.mn[0] + (gx + (L.cs[0] == 3 ? iX - 1 : 0) + L.cnt[0]) % L.cnt[0]) % C[0]) + C[0]) % C[0];
    const int hy = (((L.mn[1] + (gy + (L.cs[1] == 3 ? iY - 1 : 0) + L.cnt[1]) % L.cnt[1]) % C[1]) + C[1]) % C[1];
    const int hz = (((L.mn[2] + (gz + (L.cs[2] == 3 ? iZ - 1 : 0) + L.cnt[2]) % L.cnt[2]) % C[2]) + C[2]) % C[2];
    const int wx = ((ax % C[0]) + C[0]) % C[0], wy = ((ay % C[1]) + C[1]) % C[1], wz = ((az % C[2]) + C[2]) % C[2];
    int n = 0;  // entries in the list (uniform)
    auto wrapc = [](int v, int cnt) { return v < 0 ? v + cnt : (v >= cnt ? v - cnt : v); };
    auto consume = [&]() {
        // stencil bits of each entry: for every slot s of k's mask, the stencil
        // offset of cell(k) + s from g, wrapped as the stencil wraps the level grid
        // (a superset of the stencil cells whose exact slot test below passes)
        for (int t = lane; t < n; t += 64) {
            const int o = lc[t];
            const int ko[3] = {o % 3 - 1, (o / 3) % 3 - 1, o / 9 - 1};
            unsigned hb = 0;
            for (unsigned m = (unsigned)lm[t]; m; m &= m - 1) {
                const int sl = __builtin_ctz(m);
                const int so[3] = {sl / 9 - 1, (sl / 3) % 3 - 1, sl % 3 - 1};
                int idx = 0;
                bool ok = true;
                for (int a = 0; a < 3; ++a) {
                    int v = ko[a] + so[a];
                    if (L.cs[a] == 3) {
                        if (v > 1) v -= L.cnt[a];
                        else if (v < -1) v += L.cnt[a];
                        ok = ok && v >= -1 && v <= 1;
                        idx = 3 * idx + v + 1;
                    } else {
                        ok = ok && v == 0;
                    }
                }
                if (ok) hb |= 1u << idx;
            }
            lh[t] = hb;
        }
        __syncthreads();
        if (active)
            for (int t = part; t < n; t += nparts) {
                if (!((lh[t] >> c) & 1u)) continue;
                const int o = lc[t];
                const int kx = wrapc(wx + o % 3 - 1, C[0]);
                const int ky = wrapc(wy + (o / 3) % 3 - 1, C[1]);
                const int kz = wrapc(wz + o / 9 - 1, C[2]);
                const int sh = fac_slot(P, hx, hy, hz, kx, ky, kz);
                if (sh < 0 || !(((unsigned)lm[t] >> sh) & 1u)) continue;
                const int sg = fac_slot(P, wx, wy, wz, kx, ky, kz);
                const int k = lk[t];
                const double* G = gtab + 81 * (long long)k;
                const double lv = F.lambda[k];
                const double gg0 = G[3 * sg], gg1 = G[3 * sg + 1], gg2 = G[3 * sg + 2];
                const double gh0 = G[3 * sh], gh1 = G[3 * sh + 1], gh2 = G[3 * sh + 2];
                const double lg[3] = {lv * gg0, lv * gg1, lv * gg2};
                for (int rr = 0; rr < 3; ++rr) {
                    acc[3 * rr] += lg[rr] * gh0;
                    acc[3 * rr + 1] += lg[rr] * gh1;
                    acc[3 * rr + 2] += lg[rr] * gh2;
                }
            }
        __syncthreads();
        n = 0;
    };
    for (int dx = -P.range[0]; dx <= P.range[0]; ++dx)
        for (int dy = -P.range[1]; dy <= P.range[1]; ++dy)
            for (int dz = -P.range[2]; dz <= P.range[2]; ++dz) {
                const int kx = ((ax + dx) % C[0] + C[0]) % C[0];
                const int ky = ((ay + dy) % C[1] + C[1]) % C[1];
                const int kz = ((az + dz) % C[2] + C[2]) % C[2];
                const int sg = fac_slot(P, wx, wy, wz, kx, ky, kz);
                if (sg < 0) continue;
                const int kc = body::cell_id(P, kx, ky, kz);
                for (int half = 0; half < 2; ++half) {  // fluid, then wall particles of the cell
                    const int q0 = cellStart[half * P.T + kc], q1 = cellStart[half * P.T + kc + 1];
                    for (int base = q0; base < q1; base += 64) {
                        const int k = base + lane;
                        const unsigned km = k < q1 ? (unsigned)gmask[k] : 0u;
                        const bool hit = (km >> sg) & 1u;
                        const unsigned long long bal = __ballot(hit);
                        const int before = __popcll(bal & ((1ull << lane) - 1ull));
                        const int cnt = __popcll(bal);
                        if (n + cnt > kGal0List) consume();
                        if (hit) {
                            lk[n + before] = k;
                            lc[n + before] = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                            lm[n + before] = (int)km;
                        }
                        n += cnt;
                        __syncthreads();
                    }
                }
            }
    consume();  // ends with a barrier: the list is dead, sbuf takes the partition sums
    if (active)
        for (int e = 0; e < 9; ++e) sbuf[9 * lane + e] = acc[e];
    __syncthreads();
    if (part == 0) {
        for (int q = 1; q < nparts; ++q)
            for (int e = 0; e < 9; ++e) acc[e] += sbuf[9 * (q * ncs + c) + e];
        for (int rr = 0; rr < 3; ++rr)
            for (int ss = 0; ss < 3; ++ss) L.A[(long long)(3 * c + ss) * L.nrow + 3 * g + rr] = acc[3 * rr + ss];
    }
}

// InvDiagA (calculateInvDiagA, main.cpp:3045-3068): 1/A_ii[r][r] with
// A_ii[r][r] = D_i[r][r] + sum_{k in P(i)} lambda_k C_ik[r]^2
__global__ void __launch_bounds__(256) k_fac_invdiag(DevParams P, FacOp F, double* __restrict__ invDiagA) {
    // entry-parallel: lane g forms the terms of entries g, g + 8, ...; the group's
    // first lane adds them in list order
    constexpr int G = 8, K = 3;
    __shared__ double tb[256 * K];
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.nf;
    const double pi0 = own ? F.pos[3 * i] : 0.0, pi1 = own ? F.pos[3 * i + 1] : 0.0, pi2 = own ? F.pos[3 * i + 2] : 0.0;
    double d[3] = {0.0, 0.0, 0.0};
    if (own) {
        d[0] = F.diag[9 * i];
        d[1] = F.diag[9 * i + 4];
        d[2] = F.diag[9 * i + 8];
    }
    const long long c0 = own ? F.nPtrP[i] : 0;
    const int n = own ? F.nCountP[i] : 0;
    for (int base = 0; base < n; base += G) {
        const int jn = base + g;
        double t[K] = {0.0, 0.0, 0.0};
        if (jn < n) {
            const int k = F.nIndP[c0 + jn];
            double cv[3];
            if (k == i) {
                cv[0] = F.selfC[3 * i];
                cv[1] = F.selfC[3 * i + 1];
                cv[2] = F.selfC[3 * i + 2];
            } else {
                double r0, r1, r2;
                body::fac_disp(P, F.pos, k, pi0, pi1, pi2, r0, r1, r2);
                const double b = F.beta[c0 + jn];
                cv[0] = b * r0;
                cv[1] = b * r1;
                cv[2] = b * r2;
            }
            const double lk = F.lambda[k];
            for (int a = 0; a < 3; ++a) t[a] = lk * cv[a] * cv[a];
        }
        for (int e = 0; e < K; ++e) tb[K * threadIdx.x + e] = t[e];
        wave_sync();
        if (g == 0)
            for (int u = 0; u < G && base + u < n; ++u)
                for (int a = 0; a < 3; ++a) d[a] += tb[K * (threadIdx.x + u) + a];
        wave_sync();
    }
    if (g == 0 && own)
        for (int a = 0; a < 3; ++a) invDiagA[3 * i + a] = d[a] != 0.0 ? 1.0 / d[a] : 0.0;
}

// the wrap flag is uniform, so the branch is taken by the whole launch; the FP32
// form has no displacements to wrap
template <class VT>
__global__ void __launch_bounds__(256) k_fac_t(DevParams P, FacOp F, const double* __restrict__ x) {
    if (cr_halted(P.halt)) return;
    if constexpr (sizeof(VT) == 4) body::fac_t32_chunk(P, F, x, blockIdx.x);
    else if (*F.wrap) body::fac_t_chunk<true>(P, F, x, blockIdx.x);
    else body::fac_t_chunk<false>(P, F, x, blockIdx.x);
}

template <bool kNeg, class VT>
__global__ void __launch_bounds__(256) k_fac_y(DevParams P, FacOp F, const double* __restrict__ x, double beta,
                                               const double* yin, double* y, double* y2, double* p_copy,
                                               const double* __restrict__ dotx, double* __restrict__ part) {
    if (cr_halted(P.halt)) return;
    __shared__ double sh[256];
    __shared__ double rows[3 * 256];
    if constexpr (sizeof(VT) == 4)
        body::fac_y32_chunk<kNeg>(P, F, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh, rows);
    else if (*F.wrap)
        body::fac_y_chunk<kNeg, VT, true>(P, F, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh, rows);
    else body::fac_y_chunk<kNeg, VT, false>(P, F, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh, rows);
}

// ---------------------------------------------------------------------------
// The FP64 second pass with the viscous gathers from LDS.
//
// Every row of one cell visits the same 27 (9 in 2D) cells, so one workgroup per row
// cell stages the records (x and position) of the fluid particles of those cells in
// LDS once, and its rows read their fluid neighbours' records from there instead of
// gathering them through L1 and L2 per entry (each gather instruction of the chunk
// kernel touched about 21 cache lines).  The P part (about a tenth of the entries,
// walls included) keeps the global gathers.  A row's entries, lanes and order of
// summation are those of body::fac_y_chunk, so the results are the chunk kernel's.
// The rows go straight to y; the dot partials of the fused epilogue come from a
// dot-partial launch afterwards (the chunks of 256 rows do not follow cells).
constexpr int kCellThreads = 512;  // 64 rows of kFacLanes lanes per pass
constexpr int kCellSlots = 1536;   // staged records of 48 B: 72 KiB, two workgroups per CU
constexpr int kCellTable = 32;     // visited cells (27 in 3D) padded for the binary search

// element k of y (and y2) from the row's sum sm, as spmv_store: the separate
// roundings of y = beta*yin + sum (this function is compiled without contraction);
// yv and y2v are yin[k] and y2[k], read before
__device__ inline void fac_cell_store1(int k, double sm, double yv, double y2v, double beta, double* y, double* y2) {
    double yy = yv;
    yy *= beta;
    yy += sm;
    if (y2) {
        double zz = y2v;
        zz *= beta;
        zz += sm;
        y2[k] = zz;
    }
    y[k] = yy;
}

typedef double dv2c __attribute__((ext_vector_type(2)));

// the visited cells of a row cell: fluid ranges in visiting order and LDS offsets
struct FacCellTable {
    int vlo[kCellTable], voff[kCellTable], vcnt[kCellTable];
    int vex[kCellTable];  // first staged place of each visited cell (past the last: INT_MAX)
    int total;            // staged records of the visited cells
};
struct FacCellShared {  // the product's LDS: the table and the records
    FacCellTable tab;
    dv2c rec[3 * kCellSlots];
};

// Workgroup b's row cell (b counts the scene's workgroups of the launch; cell
// (x, y, z) of the fluid bounding box, z fastest): its
// fluid rows [rs, rs + nrows) and, in T, the visited cells in the neighbour search's
// order (ascending CellId, so their fluid ranges ascend too) with their LDS offsets,
// filled by wavefront 0.  False (for the whole workgroup) if the cell has no fluid.
// Ends with a workgroup barrier when it returns true.
__device__ inline bool fac_cell_table(const DevParams& P, const FacOp& F, FacCellTable& T, int& rs, int& nrows,
                                      int& nv, int b) {
    const int t = threadIdx.x;
    if (F.cellXcd) {
        // workgroups are dealt to the 8 XCDs in turn: XCD x = b mod 8 takes a contiguous
        // range of cells instead, so that the cells running on one XCD at a time are
        // neighbours and share their visited cells' records in that XCD's L2
        const int nb = F.cellCnt[0] * F.cellCnt[1] * F.cellCnt[2], per = nb / 8, rem = nb % 8, x = b % 8;
        b = x * per + (x < rem ? x : rem) + b / 8;
    }
    const int* C = P.C;
    const int cz = F.cellLo[2] + b % F.cellCnt[2];
    const int cy = F.cellLo[1] + (b / F.cellCnt[2]) % F.cellCnt[1];
    const int cx = F.cellLo[0] + b / (F.cellCnt[2] * F.cellCnt[1]);
    const int c = (cx * C[1] + cy) * C[2] + cz;
    rs = F.cellStart[c];
    nrows = F.cellStart[c + 1] - rs;
    if (nrows <= 0) return false;
    const int nx = 2 * P.range[0] + 1, ny = 2 * P.range[1] + 1, nz = 2 * P.range[2] + 1;
    nv = nx * ny * nz;
    if (t < 64) {
        int lo = 0x7fffffff, cnt = 0;
        if (t < nv) {
            const int bx = 2 * P.range[0] - (cx + P.range[0]) % C[0];
            const int by = 2 * P.range[1] - (cy + P.range[1]) % C[1];
            const int bz = 2 * P.range[2] - (cz + P.range[2]) % C[2];
            const int a = (bx > 0 ? bx : 0) + t / (ny * nz), bb = (by > 0 ? by : 0) + (t / nz) % ny;
            const int cc = (bz > 0 ? bz : 0) + t % nz;
            const int jx = ((cx - P.range[0] + a % nx) % C[0] + C[0]) % C[0];
            const int jy = ((cy - P.range[1] + bb % ny) % C[1] + C[1]) % C[1];
            const int jz = ((cz - P.range[2] + cc % nz) % C[2] + C[2]) % C[2];
            const int jc = (jx * C[1] + jy) * C[2] + jz;
            lo = F.cellStart[jc];
            cnt = F.cellStart[jc + 1] - lo;
        }
        int incl = cnt;  // inclusive scan of the counts over the wavefront
        for (int d = 1; d < 64; d <<= 1) {
            const int u = __shfl_up(incl, d, 64);
            if (t >= d) incl += u;
        }
        if (t < kCellTable) {
            T.vlo[t] = lo;
            T.vcnt[t] = cnt;
            T.voff[t] = (incl - cnt) - (t < nv ? lo : 0);
            T.vex[t] = t < nv ? incl - cnt : 0x7fffffff;
        }
        if (t == 63) T.total = incl;
    }
    __syncthreads();
    return true;
}

// the staged place of fluid particle j: j + voff of the last visited range starting at
// or before j
__device__ inline int fac_cell_slot(const FacCellTable& T, int j) {
    int v = 0;
#pragma unroll
    for (int step = kCellTable / 2; step >= 1; step >>= 1)
        if (T.vlo[v + step] <= j) v += step;
    return j + T.voff[v];
}

// FacOp::slot of every fluid-neighbour entry of the fluid rows of cells whose visited
// cells fit the staging (the product gathers the others' records from global memory)
__global__ void __launch_bounds__(kCellThreads) k_fac_slots(DevParams P, FacOp F) {
    __shared__ FacCellTable T;
    int rs, nrows, nv;
    if (!fac_cell_table(P, F, T, rs, nrows, nv, blockIdx.x)) return;
    if (T.total > kCellSlots) return;
    constexpr int G = body::kFacLanes, U = 4;  // U entries per lane in flight
    for (int r = threadIdx.x / G; r < nrows; r += kCellThreads / G) {
        const int i = rs + r;
        const long long b0 = F.nPtr[i];
        const int w = F.nFluid[i];
        for (int jb = threadIdx.x & (G - 1); jb < w; jb += G * U) {
            int j[U];
#pragma unroll
            for (int u = 0; u < U; ++u) j[u] = jb + u * G < w ? F.nInd[b0 + jb + u * G] : 0;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (jb + u * G < w) F.slot[b0 + jb + u * G] = (unsigned short)fac_cell_slot(T, j[u]);
        }
    }
}

// the viscous list entries of a batch: lane g takes entries jb + g + u*G (u < U).
// kSlot: the staged places (FacOp::slot), else the columns.  Entries past the row's
// end take i's own record (own) with weight 0; their loads read the row's last entry
// (w >= 1), so that the loads need no branches and issue together
template <int U, bool kSlot>
__device__ inline void fac_cell_list(const FacOp& F, const double* __restrict__ alpha, long long b0, int w, int own,
                                     int jb, int g, int* j, double* a) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int jn = jb + g + u * body::kFacLanes;
        const long long e = b0 + (jn < w ? jn : w - 1);
        const int jj = kSlot ? (int)F.slot[e] : F.nInd[e];
        const double aa = alpha[e];
        j[u] = jn < w ? jj : own;
        a[u] = jn < w ? aa : 0.0;
    }
}

// the P-list entries of a batch (branch-free loads as fac_cell_list; n >= 1: self):
// lane g takes entries jb + g + u*G; entries past the list take i with weight 0
template <int UP>
__device__ inline void fac_cell_plist(const FacOp& F, long long c0, int n, int i, int jb, int g, int* k, double* bb) {
#pragma unroll
    for (int u = 0; u < UP; ++u) {
        const int jn = jb + g + u * body::kFacLanes;
        const long long e = c0 + (jn < n ? jn : n - 1);
        const int kk = F.nIndP[e];
        const double be = F.beta[e];
        k[u] = jn < n ? kk : i;
        bb[u] = jn < n ? be : 0.0;
    }
}

// what a row's walk starts from: its position, list extents and staged place (the
// loads that only depend on the row)
struct FacRowHead {
    double pi0, pi1, pi2;
    long long b0, c0;
    int w, n, own;
};

template <bool kSlot>
__device__ inline void fac_cell_row_head(const FacOp& F, const FacCellTable& T, int i, FacRowHead& S) {
    S.pi0 = F.pos[3 * i];
    S.pi1 = F.pos[3 * i + 1];
    S.pi2 = F.pos[3 * i + 2];
    S.b0 = F.nPtr[i];
    S.w = F.nFluid[i];
    S.c0 = F.nPtrP[i];
    S.n = F.nCountP[i];
    S.own = kSlot ? fac_cell_slot(T, i) : i;
}

// The rows of one cell.  kStaged: the fluid records of the visited cells are in LDS
// (L.rec), read through the precomputed places (FacOp::slot) if kSlot, else through
// a search of the table; otherwise every record is gathered from global memory.
// None of this changes which lane takes an entry or the order in which a lane adds
// its terms.  (Loading the next batch's list entries before the current batch's
// records, or 8 entries per batch instead of 4, measured slower: 151.7 and 169.8
// against 125.9 ms per step on the 1M scene; loading them after the records, into a
// second buffer, changed the product by 0.3%; loading the first row's start before
// the staging took 14 more registers and, held to four waves per SIMD, spilled:
// 1.41 against 0.93 ms per product.)
template <bool kNeg, bool kWrap, bool kStaged, bool kSlot>
__device__ inline void fac_y_cell_rows(const DevParams& P, const FacOp& F, const double* __restrict__ x, double beta,
                                       const double* yin, double* y, double* y2, double* p_copy,
                                       const FacCellShared& L, int rs, int nrows) {
#pragma clang fp contract(fast)
    typedef double dv2 __attribute__((ext_vector_type(2)));
    constexpr int G = body::kFacLanes, U = body::kFacBatch, UP = body::kFacBatchP;
    const int t = threadIdx.x;
    const body::FacPlanes R = body::fac_planes(P, F.rec);
    const dv2* __restrict__ ra = (const dv2*)R.a;
    const dv2* __restrict__ rb = (const dv2*)R.b;
    const dv2* __restrict__ rc = (const dv2*)R.c;
    const double* __restrict__ rt = R.t;
    const double* __restrict__ alpha = (const double*)F.alpha;
    const double W0 = P.W[0], W1 = P.W[1], W2 = P.W[2], h0 = P.hW[0], h1 = P.hW[1], h2 = P.hW[2];
    const int g = t & (G - 1);
    const int passes = (nrows + kCellThreads / G - 1) / (kCellThreads / G);
    for (int ps = 0; ps < passes; ++ps) {  // uniform trip count: every lane reaches the group sums
        const int r = ps * (kCellThreads / G) + t / G;
        const bool on = r < nrows;
        const int i = rs + (on ? r : 0);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        double dA = 0.0, dB = 0.0, dC = 0.0, xi0 = 0.0, xi1 = 0.0, xi2 = 0.0, ti = 0.0, sc = 0.0, yv = 0.0, y2v = 0.0;
        if (on) {
            // the row's start; the first P batch stays in flight during the viscous part
            FacRowHead S;
            fac_cell_row_head<kStaged && kSlot>(F, L.tab, i, S);
            const double pi0 = S.pi0, pi1 = S.pi1, pi2 = S.pi2;
            const long long b0 = S.b0, c0 = S.c0;
            const int w = S.w, n = S.n, own = S.own;
            int j[U], k[UP];
            double a[U], bb[UP];
            fac_cell_list<U, kStaged && kSlot>(F, alpha, b0, w, own, 0, g, j, a);
            fac_cell_plist<UP>(F, c0, n, i, 0, g, k, bb);
            // viscous part over the fluid neighbours (alpha 0: self, beyond RadiusV)
            for (int jb = 0; jb < w; jb += G * U) {
                dv2 q0[U], q1[U], q2[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (kStaged) {
                        const int sl = kSlot ? j[u] : fac_cell_slot(L.tab, j[u]);
                        q0[u] = L.rec[3 * sl];
                        q1[u] = L.rec[3 * sl + 1];
                        q2[u] = L.rec[3 * sl + 2];
                    } else {
                        q0[u] = ra[j[u]];
                        q1[u] = rb[j[u]];
                        q2[u] = rc[j[u]];
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double r0 = body::fac_wrap<kWrap>(q1[u][1] - pi0, W0, h0);
                    const double r1 = body::fac_wrap<kWrap>(q2[u][0] - pi1, W1, h1);
                    const double r2 = body::fac_wrap<kWrap>(q2[u][1] - pi2, W2, h2);
                    const double f = a[u] * ((r0 * q0[u][0] + r1 * q0[u][1]) + r2 * q1[u][0]);
                    s0 -= f * r0;
                    s1 -= f * r1;
                    s2 -= f * r2;
                }
                if (jb + G * U < w) fac_cell_list<U, kStaged && kSlot>(F, alpha, b0, w, own, jb + G * U, g, j, a);
            }
            // lanes 0..2 of the row: the loads of their component of the diagonal term
            // and of y, in flight during the P part
            if (g < 3) {
                const double* D = F.diag + 9 * i + 3 * g;
                dA = D[0];
                dB = D[1];
                dC = D[2];
                xi0 = x[3 * i];
                xi1 = x[3 * i + 1];
                xi2 = x[3 * i + 2];
                ti = rt[i];
                sc = F.selfC[3 * i + g];
                yv = yin[3 * i + g];
                if (y2) y2v = y2[3 * i + g];
            }
            // C Lambda C^T part over the P neighbours (beta 0: self)
            for (int jb = 0; jb < n; jb += G * UP) {
                if (jb > 0) fac_cell_plist<UP>(F, c0, n, i, jb, g, k, bb);
                // the position from pos (24 contiguous bytes, mostly one cache line)
                // rather than from the record planes (two lines), and t
                double q0[UP][3], q3[UP];
#pragma unroll
                for (int u = 0; u < UP; ++u) {
                    const long long k3 = 3 * (long long)k[u];
                    q0[u][0] = F.pos[k3];
                    q0[u][1] = F.pos[k3 + 1];
                    q0[u][2] = F.pos[k3 + 2];
                    q3[u] = rt[k[u]];
                }
#pragma unroll
                for (int u = 0; u < UP; ++u) {
                    const double r0 = body::fac_wrap<kWrap>(q0[u][0] - pi0, W0, h0);
                    const double r1 = body::fac_wrap<kWrap>(q0[u][1] - pi1, W1, h1);
                    const double r2 = body::fac_wrap<kWrap>(q0[u][2] - pi2, W2, h2);
                    const double f = bb[u] * q3[u];
                    s0 += f * r0;
                    s1 += f * r1;
                    s2 += f * r2;
                }
            }
        }
        s0 = body::group_sum(s0, G);  // every lane of the group holds the sums
        s1 = body::group_sum(s1, G);
        s2 = body::group_sum(s2, G);
        if (g < 3 && on) {  // lane g: component g, as body::fac_y_chunk computes it
            const double sg = body::fac_diag_add(g == 0 ? s0 : g == 1 ? s1 : s2, dA, dB, dC, xi0, xi1, xi2, sc, ti);
            fac_cell_store1(3 * i + g, kNeg ? -sg : sg, yv, y2v, beta, y, y2);
            if (p_copy) p_copy[3 * i + g] = g == 0 ? xi0 : g == 1 ? xi1 : xi2;
        }
    }
}

template <bool kNeg, bool kWrap>
__device__ inline void fac_y_cell_body(const DevParams& P, const FacOp& F, const double* __restrict__ x, double beta,
                                       const double* yin, double* y, double* y2, double* p_copy, FacCellShared& L,
                                       int b) {
    int rs, nrows, nv;
    if (!fac_cell_table(P, F, L.tab, rs, nrows, nv, b)) return;  // the whole workgroup: no fluid in this cell
    if (L.tab.total > kCellSlots) {  // too many records to stage: gather them from global memory
        fac_y_cell_rows<kNeg, kWrap, false, false>(P, F, x, beta, yin, y, y2, p_copy, L, rs, nrows);
        return;
    }
    {
        // thread t stages places t, t + kCellThreads, ...: every load is issued before
        // the first store, so the staging waits for one round trip to memory.  Place r
        // holds the record of particle vlo[v] + r - vex[v] of the last visited cell v
        // with vex[v] <= r (empty cells share their successor's vex)
        const body::FacPlanes R = body::fac_planes(P, F.rec);
        const dv2c* __restrict__ ra = (const dv2c*)R.a;
        const dv2c* __restrict__ rb = (const dv2c*)R.b;
        const dv2c* __restrict__ rc = (const dv2c*)R.c;
        constexpr int kRounds = kCellSlots / kCellThreads;
        static_assert(kRounds * kCellThreads == kCellSlots, "the staging rounds cover the slots");
        const int t = threadIdx.x, total = L.tab.total;
        dv2c qa[kRounds], qb[kRounds], qc[kRounds];
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            const int r = t + k * kCellThreads;
            if (r < total) {
                int v = 0;
#pragma unroll
                for (int step = kCellTable / 2; step >= 1; step >>= 1)
                    if (L.tab.vex[v + step] <= r) v += step;
                const int j = L.tab.vlo[v] + (r - L.tab.vex[v]);
                qa[k] = ra[j];
                qb[k] = rb[j];
                qc[k] = rc[j];
            }
        }
#pragma unroll
        for (int k = 0; k < kRounds; ++k) {
            const int r = t + k * kCellThreads;
            if (r < total) {
                L.rec[3 * r] = qa[k];
                L.rec[3 * r + 1] = qb[k];
                L.rec[3 * r + 2] = qc[k];
            }
        }
    }
    __syncthreads();
    if (F.slot) fac_y_cell_rows<kNeg, kWrap, true, true>(P, F, x, beta, yin, y, y2, p_copy, L, rs, nrows);
    else fac_y_cell_rows<kNeg, kWrap, true, false>(P, F, x, beta, yin, y, y2, p_copy, L, rs, nrows);
}

template <bool kNeg>
__global__ void __launch_bounds__(kCellThreads) k_fac_y_cell(DevParams P, FacOp F, const double* __restrict__ x,
                                                             double beta, const double* yin, double* y, double* y2,
                                                             double* p_copy) {
    if (cr_halted(P.halt)) return;
    __shared__ FacCellShared L;
    if (*F.wrap) fac_y_cell_body<kNeg, true>(P, F, x, beta, yin, y, y2, p_copy, L, blockIdx.x);
    else fac_y_cell_body<kNeg, false>(P, F, x, beta, yin, y, y2, p_copy, L, blockIdx.x);
}

// The same pass for the scenes of a batch (the scene-table launches of mph_batch.hip):
// workgroup b belongs to the scene s with R.pre[s] <= b < R.pre[s+1] and runs that
// scene's cell b - R.pre[s] with its FP64 factored operator.  The grid covers the
// scenes of the launch's set only (converged scenes have empty ranges).  The scenes
// of `first` also write y2 and p_copy.
template <bool kNeg>
__global__ void __launch_bounds__(kCellThreads) kb_fac_y_cell(const batch::BScene* __restrict__ tab, batch::CellRanges R, int ns,
                                                              batch::SceneSet first, int x, double beta, int yin, int y,
                                                              int y2, int p_copy) {
    int lo = 0, hi = ns;  // the last s with R.pre[s] <= b
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (R.pre[mid] <= (int)blockIdx.x) lo = mid;
        else hi = mid;
    }
    const int s = lo, lb = (int)blockIdx.x - R.pre[lo];
    const batch::BScene& B = tab[s];
    if (batch::halted(B)) return;
    __shared__ FacCellShared L;
    // the scene's parameters as values: uniform loads into scalar registers (read
    // through the table, the compiler keeps them in vector registers and spills)
    const DevParams P = B.P;
    const FacOp F = B.fac[0];
    const bool f = batch::in_set(first, s);
    double* y2p = f && y2 >= 0 ? B.vec[y2] : nullptr;
    double* pc = f && p_copy >= 0 ? B.vec[p_copy] : nullptr;
    if (*F.wrap) fac_y_cell_body<kNeg, true>(P, F, B.vec[x], beta, B.vec[yin], B.vec[y], y2p, pc, L, lb);
    else fac_y_cell_body<kNeg, false>(P, F, B.vec[x], beta, B.vec[yin], B.vec[y], y2p, pc, L, lb);
}

template <bool kNeg, class VT>
void fac_launch(hipStream_t st, const DevParams& P, const FacOp& F, const double* x, double beta, const double* yin,
                double* y, double* y2, double* p_copy, const double* dotx, double* part, int passes) {
    if (passes & 1) hipLaunchKernelGGL(k_fac_t<VT>, dim3(nblk(P.np, 256)), dim3(256), 0, st, P, F, x);
    if (!(passes & 2)) return;
    const long long cells = (long long)F.cellCnt[0] * F.cellCnt[1] * F.cellCnt[2];
    if (sizeof(VT) == 8 && F.cellStart && cells > 0) {
        hipLaunchKernelGGL((k_fac_y_cell<kNeg>), dim3((unsigned)cells), dim3(kCellThreads), 0, st, P, F, x, beta, yin,
                           y, y2, p_copy);
        if (part) launch_dot_partial(st, P, dotx, y, part);
        return;
    }
    hipLaunchKernelGGL((k_fac_y<kNeg, VT>), dim3(nblk(P.nf, kChunk)), dim3(kChunk), 0, st, P, F, x, beta, yin, y, y2,
                       p_copy, dotx, part);
}

}  // namespace

void launch_fac_p(hipStream_t st, const DevParams& P, const double* pos, const int* nCountP, const long long* nPtrP,
                  const int* nIndP, double* beta, float* v32, int* runP, double* selfC, double* rec) {
    if (P.np == 0) return;
    hipLaunchKernelGGL(k_fac_p, dim3(nblk(8LL * P.np, 256)), dim3(256), 0, st, P, pos, nCountP, nPtrP, nIndP, beta,
                       v32, runP, selfC, rec);
}

void launch_fac_level0(hipStream_t st, const DevParams& P, const Level& L0, const FacOp& F, const int* cellIndex,
                       const int* cellStart, const double* blk, double* gtab, int* gmask, double* invDiagA) {
    if (P.nf == 0) return;
    hipLaunchKernelGGL(k_fac_g, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, cellIndex, F.runP, F.nPtrP, F.nIndP,
                       F.beta, F.pos, F.selfC, gtab, gmask);
    hipLaunchKernelGGL(k_fac_invdiag, dim3(nblk(8LL * P.nf, 256)), dim3(256), 0, st, P, F, invDiagA);
    if (L0.cells > 0)
        hipLaunchKernelGGL(k_fac_gal0, dim3(L0.cells), dim3(64), 0, st, P, L0, F, cellStart, blk, gtab, gmask);
}

void launch_fac_u(hipStream_t st, const DevParams& P, const int* nFluid, const long long* nPtr, const int* nInd,
                  const double* alpha, const double* pos, float* u32) {
    if (P.nf == 0) return;
    hipLaunchKernelGGL(k_fac_u, dim3(nblk(8LL * P.nf, 256)), dim3(256), 0, st, P, nFluid, nPtr, nInd, alpha, pos,
                       u32);
}

__global__ void __launch_bounds__(256) k_fac_rec_positions(DevParams P, const double* __restrict__ from,
                                                           double* __restrict__ to) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.np) return;
    const body::FacPlanes F = body::fac_planes(P, const_cast<double*>(from)), T = body::fac_planes(P, to);
    T.b[2 * k + 1] = F.b[2 * k + 1];  // p0
    T.c[2 * k] = F.c[2 * k];          // p1
    T.c[2 * k + 1] = F.c[2 * k + 1];  // p2
}

void launch_fac_rec_positions(hipStream_t st, const DevParams& P, const double* from, double* to) {
    if (P.np == 0) return;
    hipLaunchKernelGGL(k_fac_rec_positions, dim3(nblk(P.np, 256)), dim3(256), 0, st, P, from, to);
}

void launch_fac_slots(hipStream_t st, const DevParams& P, const FacOp& F) {
    const long long cells = (long long)F.cellCnt[0] * F.cellCnt[1] * F.cellCnt[2];
    if (P.nf == 0 || !F.cellStart || !F.slot || cells <= 0) return;
    hipLaunchKernelGGL(k_fac_slots, dim3((unsigned)cells), dim3(kCellThreads), 0, st, P, F);
}

namespace batch {
void spmv_fac_cells(const Tables& T, const SceneSet& on, const SceneSet& first, double alpha, int x, double beta,
                    int yin, int y, int y2, int p_copy) {
    CellRanges R;
    const int* o = T.hoff + (size_t)kOffCell * (T.ns + 1);
    int g = 0;
    for (int s = 0; s < T.ns; ++s) {
        R.pre[s] = g;
        if ((on.w[s >> 6] >> (s & 63)) & 1ull) g += o[s + 1] - o[s];
    }
    R.pre[T.ns] = g;
    if (g <= 0) return;
    if (alpha < 0)
        hipLaunchKernelGGL((kb_fac_y_cell<true>), dim3(g), dim3(kCellThreads), 0, T.st, T.tab, R, T.ns, first, x, beta,
                           yin, y, y2, p_copy);
    else
        hipLaunchKernelGGL((kb_fac_y_cell<false>), dim3(g), dim3(kCellThreads), 0, T.st, T.tab, R, T.ns, first, x,
                           beta, yin, y, y2, p_copy);
}
}  // namespace batch

void launch_fac_spmv(hipStream_t st, const DevParams& P, const FacOp& F, double alpha, const double* x, double beta,
                     const double* yin, double* y, double* y2, double* p_copy, const double* dotx, double* part,
                     bool f32, int passes) {
    // alpha is +1 or -1 at every call site (main.cpp:4337, 4353, 4357)
    if (P.nf == 0) return;
    if (f32) {
        if (alpha < 0) fac_launch<true, float>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part, passes);
        else fac_launch<false, float>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part, passes);
    } else {
        if (alpha < 0) fac_launch<true, double>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part, passes);
        else fac_launch<false, double>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part, passes);
    }
}

}  // namespace mph
