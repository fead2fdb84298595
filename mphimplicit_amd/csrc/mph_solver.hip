// mph_solver.hip -- multigrid hierarchy, V-cycle and conjugate-residual kernels.
//
// Replaces calculateMultiGridMatrix and its builders (source/main.cpp:2761-3546),
// preconditionWithMultiGrid (4122-4266) and solveWithConjugatedGradient (4270-4421).
// The reference materialises P2G, G2P, R, P and the level matrices as CSR; their
// patterns are fixed stencils on the fluid bounding box, so here they are applied
// implicitly (no index arrays), with the reference's per-row summation order.
// Dot products use one fixed tree (chunks of 256 particles, then 256 lanes over the
// chunk sums), independent of launch configuration; the CPU oracle implements the
// same tree so the solve can be compared bit for bit.
#include "mph_cr_body.h"

namespace mph {

namespace {

using body::block_tree_sum;
using body::cell_id;
using body::grid_id;
using body::grid_xyz;
using body::nblk;
using namespace body;  // CR scalar slots

// ---------------------------------------------------------------------------
// hierarchy build

// bounding box of fluid-occupied cells (calculateMultiGridDepth, 2786-2802)
__global__ void k_fluid_bbox(DevParams P, const int* __restrict__ cellStart, int* __restrict__ bbox) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.T) return;
    if (cellStart[c + 1] - cellStart[c] > 0) {
        const int cx = (c / (P.C[1] * P.C[2])) % P.C[0], cy = (c / P.C[2]) % P.C[1], cz = c % P.C[2];
        atomicMax(&bbox[0], cx);
        atomicMax(&bbox[1], cy);
        atomicMax(&bbox[2], cz);
        atomicMin(&bbox[3], cx);
        atomicMin(&bbox[4], cy);
        atomicMin(&bbox[5], cz);
    }
}

// InvDiagA = 1/a_ii (calculateInvDiagA, 3045-3069)
__global__ void k_inv_diag(DevParams P, const int* __restrict__ selfIdx, const long long* __restrict__ start,
                           const double* __restrict__ aval, double* __restrict__ invDiagA) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nf) return;
    const int s = i / kSlice, lane = i % kSlice, k = selfIdx[i];
    for (int r = 0; r < 3; ++r) {
        const double v = aval[a_val_index(start[s], k, 4 * r, lane)];
        invDiagA[3 * i + r] = v != 0.0 ? 1.0 / v : 0.0;
    }
}

// level-0 matrix = sum of the particle blocks of every cell pair, in the order of
// calculateMultiGridCsrA (3351-3410): particles of the row cell ascending, then the
// blocks of each particle ascending.  Workgroup = row cell g; thread = (stencil
// cell c, block entry e), so each thread runs one sequential sum and the cell's
// particles (adjacent lanes of the A slices) share cache lines across the
// workgroup.  The blocks of a particle that fall in one neighbour cell form a
// contiguous run of its sorted fluid-neighbour list; the neighbour search recorded
// where each run starts.
__global__ void __launch_bounds__(256) k_mg_galerkin0(DevParams P, Level L, const int* __restrict__ cellStart,
                                                      const unsigned short* __restrict__ runStart,
                                                      const long long* __restrict__ start,
                                                      const double* __restrict__ aval) {
    const int ncs = L.cs[0] * L.cs[1] * L.cs[2];
    const int g = blockIdx.x;
    if ((int)threadIdx.x >= 9 * ncs) return;
    const int c = threadIdx.x / 9, e = threadIdx.x % 9;
    int gx, gy, gz;
    grid_xyz(L, g, gx, gy, gz);
    const int ic = cell_id(P, gx + L.mn[0], gy + L.mn[1], gz + L.mn[2]);
    const int p0 = cellStart[ic], p1 = cellStart[ic + 1];
    const unsigned mask = galerkin_mask(P, L, cellStart, g, c);
    double acc = 0.0;
    if (mask)
        for (int ip = p0; ip < p1; ++ip) {
            const unsigned short* rs = runStart + (long long)ip * kRuns;
            const int lane = ip % kSlice;
            const long long sb = start[ip / kSlice];
            for (unsigned m = mask; m; m &= m - 1) {
                const int q = __builtin_ctz(m);
                const int k1 = rs[q + 1];
                for (int k = rs[q]; k < k1; k += 8) {  // eight loads in flight, added in order
                    double v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) v[u] = k + u < k1 ? aval[a_val_index(sb, k + u, e, lane)] : 0.0;
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        if (k + u < k1) acc += v[u];
                }
            }
        }
    const int r = e / 3, sd = e % 3;
    L.A[(long long)(3 * c + sd) * L.nrow + 3 * g + r] = acc;
}

// the same sums with the rows of kGalGroup adjacent particles staged in LDS per
// round: a 128-byte line of A holds one entry pair of 8 adjacent particles, so
// loading the group's rows together (one coalesced 64-byte piece per entry pair)
// fetches each line at most twice, where walking one particle at a time across
// thousands of resident workgroups re-fetched it up to eight times.  Every (c, e)
// thread then adds its runs particle by particle from LDS, in the same order.
// Rows longer than kGalRow blocks are not supported (the caller checks).
constexpr int kGalGroup = 4, kGalRow = 216;
__global__ void __launch_bounds__(256) k_mg_galerkin0_lds(DevParams P, Level L, const int* __restrict__ cellStart,
                                                          const unsigned short* __restrict__ runStart,
                                                          const int* __restrict__ nFluid,
                                                          const long long* __restrict__ start,
                                                          const double* __restrict__ aval) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    __shared__ double row[kGalGroup * kGalRow * 9];
    const int ncs = L.cs[0] * L.cs[1] * L.cs[2];
    const int g = blockIdx.x;
    const bool active = (int)threadIdx.x < 9 * ncs;
    const int c = threadIdx.x / 9, e = threadIdx.x % 9;
    int gx, gy, gz;
    grid_xyz(L, g, gx, gy, gz);
    const int ic = cell_id(P, gx + L.mn[0], gy + L.mn[1], gz + L.mn[2]);
    const int p0 = cellStart[ic], p1 = cellStart[ic + 1];
    const unsigned mask = active ? galerkin_mask(P, L, cellStart, g, c) : 0u;
    double acc = 0.0;
    for (int ip0 = p0; ip0 < p1;) {
        // groups aligned to kGalGroup particles never straddle a 64-particle slice
        const int gend = min(p1, (ip0 / kGalGroup + 1) * kGalGroup);
        const int G = gend - ip0;
        const long long sb = start[ip0 / kSlice];
        const int lane0 = ip0 % kSlice;
        int npair = 0;
        for (int p = 0; p < G; ++p) npair = max(npair, (nFluid[ip0 + p] + 1) / 2);
        __syncthreads();  // the previous group has been used
        for (int t = threadIdx.x; t < npair * 9 * G; t += blockDim.x) {
            const int p = t % G, ee = (t / G) % 9, m = t / (G * 9);
            const d2 v = *(const d2*)&aval[a_val_index(sb, 2 * m, ee, lane0 + p)];
            row[(p * kGalRow + 2 * m) * 9 + ee] = v.x;
            row[(p * kGalRow + 2 * m + 1) * 9 + ee] = v.y;
        }
        __syncthreads();
        if (mask)
            for (int p = 0; p < G; ++p) {
                const unsigned short* rs = runStart + (long long)(ip0 + p) * kRuns;
                const double* rp = row + p * kGalRow * 9 + e;
                for (unsigned m = mask; m; m &= m - 1) {
                    const int q = __builtin_ctz(m);
                    const int k1 = rs[q + 1];
                    for (int k = rs[q]; k < k1; ++k) acc += rp[k * 9];
                }
            }
        ip0 = gend;
    }
    if (active) {
        const int r = e / 3, sd = e % 3;
        L.A[(long long)(3 * c + sd) * L.nrow + 3 * g + r] = acc;
    }
}

// coarse level = R A_fine P by summation (3414-3507): thread = (coarse cell, stencil
// cell); children in R order, then fine stencil entries in shift order.
__global__ void k_mg_galerkinL(Level S, Level L) {
    const int ncs = L.cs[0] * L.cs[1] * L.cs[2];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)L.cells * ncs) return;
    const int g = (int)(t / ncs), c = (int)(t % ncs);
    int lx, ly, lz;
    grid_xyz(L, g, lx, ly, lz);
    const int lcx = lx + L.mn[0], lcy = ly + L.mn[1], lcz = lz + L.mn[2];
    const int fmx = S.mn[0] > 2 * lcx, fmy = S.mn[1] > 2 * lcy, fmz = S.mn[2] > 2 * lcz;
    const int fMx = S.mx[0] < 2 * (lcx + 1), fMy = S.mx[1] < 2 * (lcy + 1), fMz = S.mx[2] < 2 * (lcz + 1);
    double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ax = 0; ax < 2 - fmx - fMx; ++ax)
        for (int ay = 0; ay < 2 - fmy - fMy; ++ay)
            for (int az = 0; az < 2 - fmz - fMz; ++az) {
                const int sgx = (2 * lcx + ax + fmx) - S.mn[0];
                const int sgy = (2 * lcy + ay + fmy) - S.mn[1];
                const int sgz = (2 * lcz + az + fmz) - S.mn[2];
                const int srow0 = 3 * grid_id(S, sgx, sgy, sgz);
                // the wraps take arguments in [0, 3 cnt): sg + e - 1 + cnt with sg < cnt,
                // e < 3; jl - l + 1 + cnt with jl, l < cnt (wrap3, no division)
                for (int ex = 0; ex < S.cs[0]; ++ex)
                    for (int ey = 0; ey < S.cs[1]; ++ey)
                        for (int ez = 0; ez < S.cs[2]; ++ez) {
                            const int jsx = wrap3(sgx + ex - 1 + S.cnt[0], S.cnt[0]);
                            const int jsy = wrap3(sgy + ey - 1 + S.cnt[1], S.cnt[1]);
                            const int jsz = wrap3(sgz + ez - 1 + S.cnt[2], S.cnt[2]);
                            const int jlx = ((jsx + S.mn[0]) / 2) - L.mn[0];
                            const int jly = ((jsy + S.mn[1]) / 2) - L.mn[1];
                            const int jlz = ((jsz + S.mn[2]) / 2) - L.mn[2];
                            const int ix = wrap3((jlx - lx) + 1 + L.cnt[0], L.cnt[0]);
                            const int iy = wrap3((jly - ly) + 1 + L.cnt[1], L.cnt[1]);
                            const int iz = wrap3((jlz - lz) + 1 + L.cnt[2], L.cnt[2]);
                            if (L.cs[2] * L.cs[1] * ix + L.cs[2] * iy + iz != c) continue;
                            const int cS = S.cs[2] * S.cs[1] * ex + S.cs[2] * ey + ez;
                            for (int r = 0; r < 3; ++r)
                                for (int sd = 0; sd < 3; ++sd)
                                    acc[3 * r + sd] += S.A[(long long)(3 * cS + sd) * S.nrow + srow0 + r];
                        }
            }
    for (int r = 0; r < 3; ++r)
        for (int sd = 0; sd < 3; ++sd) L.A[(long long)(3 * c + sd) * L.nrow + 3 * g + r] = acc[3 * r + sd];
}

// invD = 2 / sum |a| per row (calculateMultiGridInvDiagA, 3510-3546)
__global__ void k_mg_invd(Level L) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= L.nrow) return;
    double sum = 0.0;
    for (int sh = 0; sh < L.nshift; ++sh) sum += fabs(L.A[(long long)sh * L.nrow + row]);
    L.invD[row] = sum != 0.0 ? 2.0 / sum : 0.0;
}

__global__ void k_fill(double* x, long long n, double v) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = v;
}

__global__ void k_fill_i32(int* x, long long n, int v) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = v;
}

// ---------------------------------------------------------------------------
// V-cycle (preconditionWithMultiGrid).  Every step is myDcsrmv / myDdmv /
// myDcopy / myDscal of the reference: y = beta*y + sum_k (alpha*a_k)*x_k with the
// sum formed first, y = beta*y + (alpha*d)*x.

__global__ void k_mg_p2g(DevParams P, Level L, const int* __restrict__ cellStart, const double* __restrict__ q) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < L.nrow) mg_p2g_row(P, L, cellStart, q, row);
}

// P2G, then begin + update of level 0 on the same row, with kP2gLanes lanes per
// row: lane u loads q of the cell's particles p0 + u, p0 + u + kP2gLanes, ...
// (kP2gK per pass), and the row's first lane adds them in particle order from LDS,
// as mg_p2g_row does.  Level 0 has one row per
// cell and component, too few to fill the GPU with one thread each.
// kPq: the CR loop's q = beta q + y (4364-4366) is formed here as q is loaded and
// written back (every fluid particle lies in exactly one level-0 cell, so each
// element is updated once); q is written only then.
constexpr int kP2gLanes = 8, kP2gK = 8;
template <bool kPq>
__global__ void __launch_bounds__(256) k_mg_p2g_bu_lanes(DevParams P, Level L, const int* __restrict__ cellStart,
                                                         double* q, const double* __restrict__ yv,
                                                         const double* __restrict__ scal) {
    if (cr_halted(L.halt)) return;
    __shared__ double tb[256 * kP2gK];
    const double beta = kPq ? scal[kRHO] / scal[kRHOP] : 0.0;
    const int u = threadIdx.x & (kP2gLanes - 1);
    const int row = (int)(((long long)blockIdx.x * 256 + threadIdx.x) / kP2gLanes);
    const bool on = row < L.nrow;
    int p0 = 0, p1 = 0;
    const int rd = row % 3;
    if (on) {
        int gx, gy, gz;
        grid_xyz(L, row / 3, gx, gy, gz);
        const int ic = cell_id(P, gx + L.mn[0], gy + L.mn[1], gz + L.mn[2]);
        p0 = cellStart[ic];
        p1 = cellStart[ic + 1];
    }
    double sum = 0.0;
    for (int base = p0; base < p1; base += kP2gLanes * kP2gK) {  // the same trips in every lane of the row
        double t[kP2gK];
#pragma unroll
        for (int k = 0; k < kP2gK; ++k) {
            const int ip = base + u + kP2gLanes * k;
            t[k] = ip < p1 ? q[3 * ip + rd] : 0.0;
            if (kPq && ip < p1) {
                t[k] = pq_elem(beta, t[k], yv[3 * ip + rd]);
                q[3 * ip + rd] = t[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kP2gK; ++k) tb[kP2gK * threadIdx.x + k] = t[k];
        wave_sync();
        if (u == 0)
            for (int k = 0; k < kP2gK; ++k)
                for (int v = 0; v < kP2gLanes; ++v)
                    if (base + v + kP2gLanes * k < p1) sum += 1.0 * 1.0 * tb[kP2gK * (threadIdx.x + v) + k];
        wave_sync();
    }
    if (u != 0 || !on) return;
    double y = L.Q[row];
    y *= 0.0;
    y += sum;
    L.Q[row] = y;
    mg_begin_update_row(L, row);
}

// restriction into L, then begin + update of L on the same row
__global__ void k_mg_restrict_bu(Level S, Level L) {
    if (cr_halted(L.halt)) return;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= L.nrow) return;
    mg_restrict_row(S, L, row);
    mg_begin_update_row(L, row);
}

__global__ void k_mg_begin(Level L) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < L.nrow) mg_begin_row(L, row);
}

__global__ void k_mg_update(Level L) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < L.nrow) mg_update_row(L, row);
}

template <class VT>
__global__ void k_mg_residual(Level L) {
    if (cr_halted(L.halt)) return;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < L.nrow) mg_residual_row<VT>(L, row);
}

template <class VT>
__global__ void k_mg_residual_update(Level L) {
    if (cr_halted(L.halt)) return;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < L.nrow) mg_residual_update_row<VT>(L, row);
}

// The same sweeps with kRowLanes lanes per row, for levels too small to fill the
// GPU with one thread per row (a row's 81 loads are then latency-bound).  Lane u
// loads the entries of stencil cells [uQ, (u+1)Q) and forms their terms -1.0 a x;
// the row's first lane adds all terms in stencil order (the order of mg_row_sum),
// read back through LDS, so the sums are those of mg_residual_row and
// mg_residual_update_row.  Measured on a level of 73k rows, one thread per row is
// faster (14.3 against 18.4 us per sweep); on levels of 10k rows and fewer the lanes
// take 20-30% off.  C0 C1 C2 give the stencil at compile time (3x3x3, 3x3x1); 0 0 0
// reads it from the level (other shapes), at the price of runtime index arithmetic.
constexpr int kRowLanes = 4, kRowBlock = 128, kRowQ = 7;  // kRowQ: stencil cells per lane (27 / 4, rounded up)

template <class VT, bool kUpdate, int C0, int C1, int C2>
__global__ void __launch_bounds__(kRowBlock) k_mg_sweep_lanes(Level L) {
    if (cr_halted(L.halt)) return;
    constexpr int T = 3 * kRowQ;
    __shared__ double tb[kRowBlock * T];
    const int u = threadIdx.x & (kRowLanes - 1);
    const int row = (int)(((long long)blockIdx.x * kRowBlock + threadIdx.x) / kRowLanes);
    const bool on = row < L.nrow;
    const int cs1 = C0 ? C1 : L.cs[1], cs2 = C0 ? C2 : L.cs[2];
    const int N = (C0 ? C0 : L.cs[0]) * cs1 * cs2, Q = (N + kRowLanes - 1) / kRowLanes;  // N <= 27
    const VT* __restrict__ A = sizeof(VT) == 8 ? (const VT*)L.A : (const VT*)L.A32;
    const double* __restrict__ x = L.S[L.cur];
    const long long nrow = L.nrow;
    double t[T];
    if (on) {
        int gx, gy, gz;
        grid_xyz(L, row / 3, gx, gy, gz);
        const int c0 = L.cnt[0], c1 = L.cnt[1], c2 = L.cnt[2];
#pragma unroll
        for (int k = 0; k < kRowQ; ++k) {
            const int q = u * Q + k;
            if (k < Q && q < N) {
                const int ex = q / (cs1 * cs2), ey = (q / cs2) % cs1, ez = q % cs2;
                const int jx = wrap3(gx + ex - 1 + c0, c0), jy = wrap3(gy + ey - 1 + c1, c1);
                const int jz = wrap3(gz + ez - 1 + c2, c2);
                const double* xc = x + 3 * ((c1 * c2) * jx + c2 * jy + jz);
#pragma unroll
                for (int sd = 0; sd < 3; ++sd) t[3 * k + sd] = -1.0 * (double)A[(3 * q + sd) * nrow + row] * xc[sd];
            } else {
#pragma unroll
                for (int sd = 0; sd < 3; ++sd) t[3 * k + sd] = 0.0;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < T; ++e) tb[T * threadIdx.x + e] = on ? t[e] : 0.0;
    wave_sync();
    if (u != 0 || !on) return;
    double sum = 0.0;
    for (int v = 0; v < kRowLanes; ++v)
        for (int k = 0; k < Q && v * Q + k < N; ++k)
#pragma unroll
            for (int sd = 0; sd < 3; ++sd) sum += tb[T * (threadIdx.x + v) + 3 * k + sd];
    double y = L.Q[row];
    y *= 1.0;
    y += sum;
    L.R[row] = y;
    if (kUpdate) {
        double sv = x[row];
        sv *= 1.0;
        sv += 1.0 * L.invD[row] * y;
        L.S[L.cur ^ 1][row] = sv;
    }
}

// launches the lane-split sweep on a small level; false for a large one
template <bool kUpdate>
bool launch_sweep_lanes(hipStream_t st, const Level& L) {
    if (!L.lanes || L.cs[0] * L.cs[1] * L.cs[2] > kRowLanes * kRowQ) return false;
    const int nb = nblk((long long)kRowLanes * L.nrow, kRowBlock);
    const bool s3 = L.cs[0] == 3 && L.cs[1] == 3 && L.cs[2] == 3, s2 = L.cs[0] == 3 && L.cs[1] == 3 && L.cs[2] == 1;
#define MPH_SWEEP(VT, A, B, C) hipLaunchKernelGGL((k_mg_sweep_lanes<VT, kUpdate, A, B, C>), dim3(nb), dim3(kRowBlock), 0, st, L)
    if (L.f32) {
        if (s3) MPH_SWEEP(float, 3, 3, 3);
        else if (s2) MPH_SWEEP(float, 3, 3, 1);
        else MPH_SWEEP(float, 0, 0, 0);
    } else {
        if (s3) MPH_SWEEP(double, 3, 3, 3);
        else if (s2) MPH_SWEEP(double, 3, 3, 1);
        else MPH_SWEEP(double, 0, 0, 0);
    }
#undef MPH_SWEEP
    return true;
}

__global__ void k_mg_restrict(Level S, Level L) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < L.nrow) mg_restrict_row(S, L, row);
}

__global__ void k_mg_prolong(Level L, Level S) {
    if (cr_halted(L.halt)) return;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row < S.nrow) mg_prolong_row(L, S, row);
}

// cur: ping-pong bits of levels lc-1..D-1 (bit l), tracked by the host
__global__ void __launch_bounds__(kCoarseThreads) k_mg_coarse(const Level* __restrict__ lvs, int lc, int D, unsigned cur,
                                                   int f32, const int* halt) {
    if (cr_halted(halt)) return;
    mg_coarse_cycle(lvs, lc, D, cur, f32 != 0);
}

__global__ void __launch_bounds__(256) k_mg_g2p(DevParams P, Level L, const int* __restrict__ cellIndex,
                                                const double* __restrict__ invDiagA, const double* __restrict__ q,
                                                double* __restrict__ u, double* __restrict__ part) {
    if (cr_halted(L.halt)) return;
    __shared__ double sh[256];
    mg_g2p_chunk(P, L, cellIndex, invDiagA, q, u, part, blockIdx.x, sh);
}

// ---------------------------------------------------------------------------
// conjugate residual

__global__ void k_cr_init(DevParams P, const double* __restrict__ vel, const double* __restrict__ force,
                          const double* __restrict__ mass, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < P.nf) cr_init_particle(P, vel, force, mass, x, i);
}

template <bool kNeg, class VT, int kUn>
__global__ void __launch_bounds__(256) k_spmv(int nf, const int* __restrict__ nFluid,
                                              const long long* __restrict__ start, const int* __restrict__ col,
                                              const VT* __restrict__ val, const double* __restrict__ x,
                                              double beta, const double* yin, double* y, double* y2,
                                              double* p_copy, const double* __restrict__ dotx,
                                              double* __restrict__ part, const int* halt) {
    if (cr_halted(halt)) return;
    __shared__ double sh[256];
    spmv_chunk<kNeg, VT, kUn>(nf, nFluid, start, col, val, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh);
}

__global__ void __launch_bounds__(256) k_dot_partial(int nf, const double* __restrict__ x,
                                                     const double* __restrict__ y, double* __restrict__ part,
                                                     const int* halt) {
    if (cr_halted(halt)) return;
    __shared__ double sh[256];
    dot_partial_chunk(nf, x, y, part, blockIdx.x, sh);
}

__global__ void __launch_bounds__(256) k_dot_final(const double* __restrict__ part, int nchunks, double* scal,
                                                   int slot, int shift_slot, double* host, const int* halt) {
    if (cr_halted(halt)) return;
    __shared__ double sh[256];
    dot_final_body(part, nchunks, scal, slot, shift_slot, sh, host);
}

// r.s of a CR iteration into scal[kRSIter] and the pair (rr, r.s) into row `it` of
// the record (launch_dot_final_record)
__global__ void __launch_bounds__(256) k_dot_final_record(const double* __restrict__ part, int nchunks, double* scal,
                                                          const int* ctl, long long it_host, double* rec,
                                                          long long cap, const int* halt) {
    if (cr_halted(halt)) return;
    __shared__ double sh[256];
    dot_final_body(part, nchunks, scal, kRSIter, -1, sh);
    if (threadIdx.x != 0) return;
    const long long it = ctl ? (long long)ctl[kCtlIter] : it_host;
    if (it >= 0 && it < cap) {
        rec[2 * it] = scal[kRR];
        rec[2 * it + 1] = scal[kRSIter];
    }
}

// rr and the CR stop test (launch_dot_final_stop); ctl[kCtlHalt] is read by every
// thread before thread 0 writes it, behind the reduction's barriers
__global__ void __launch_bounds__(256) k_dot_final_stop(const double* __restrict__ part, int nchunks, double* scal,
                                                        int* ctl, int step, double tol, long long maxit, int* hctl,
                                                        double* hrr) {
    if (ctl[kCtlHalt]) return;
    __shared__ double sh[256];
    dot_final_body(part, nchunks, scal, kRR, -1, sh, hrr);
    if (threadIdx.x != 0) return;
    const double rr = scal[kRR];
    const int it = ctl[kCtlIter] + step;
    int halt = 0, bad = 0;
    if (it >= maxit) halt = 1;
    else if (!isfinite(rr)) halt = bad = 1;
    else if (rr / scal[kRR0] < tol) halt = 1;
    ctl[kCtlIter] = it;
    ctl[kCtlBad] = bad;
    ctl[kCtlHalt] = halt;
    if (hctl) {  // coherent pinned host memory: the host polls it between iterations
        hctl[kCtlIter] = it;
        hctl[kCtlBad] = bad;
        hctl[kCtlHalt] = halt;
    }
}

__global__ void k_pq_update(int nf, const double* __restrict__ scal, const double* __restrict__ s,
                            const double* __restrict__ y, double* __restrict__ p, double* __restrict__ q,
                            const int* halt) {
    if (cr_halted(halt)) return;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 3LL * nf) pq_update_elem(scal, s, y, p, q, k);
}

template <bool kPq>
__global__ void __launch_bounds__(256) k_xrs_update(int nf, const double* __restrict__ scal,
                                                    double* __restrict__ p, const double* __restrict__ q,
                                                    const double* __restrict__ u, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ s,
                                                    double* __restrict__ part, const int* halt) {
    if (cr_halted(halt)) return;
    __shared__ double sh[256];
    xrs_update_chunk<kPq>(nf, scal, p, q, u, x, r, s, part, blockIdx.x, sh);
}

__global__ void k_store_velocity(int nf, const double* __restrict__ x, double* __restrict__ vel) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 3 * nf) vel[k] = x[k];
}

// ---------------------------------------------------------------------------
// output-step diagnostics: calculateDivergenceP (1932-1969), calculatePressureP
// (1971-2015), calculateViscosityV (2017-2063), calculateVirialStressAtParticle
// (4494-4738)

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

// entry-parallel: a group of four lanes owns particle i, lane g forms the terms of
// entries g, g + 4, ...; the group's first lane subtracts them in list order
__global__ void __launch_bounds__(256) k_divergence(DevParams P, const double* __restrict__ pos,
                                                    const double* __restrict__ vel, const double* __restrict__ lambda,
                                                    const double* __restrict__ kappa,
                                                    const double* __restrict__ volStrainP,
                                                    const int* __restrict__ nCount, const long long* __restrict__ nPtr,
                                                    const int* __restrict__ nInd, double* __restrict__ divP,
                                                    double* __restrict__ pressP) {
    constexpr int G = 4, K = 3;  // terms: u_d e_d dw for d = 0, 1, 2
    __shared__ double tb[256 * K];
    __shared__ int tf[256];
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    const bool own = i < P.np;
    double sum = 0.0;
    const long long b = own ? nPtr[i] : 0;
    const int n = own ? nCount[i] : 0;
    for (int base = 0; base < n; base += G) {
        const int jn = base + g;
        int f = 0;
        double t[K] = {0.0, 0.0, 0.0};
        if (jn < n) {
            const int j = nInd[b + jn];
            if (i != j) {
                const Rel q = rel(P, pos, i, j);
                if (q.r2 != 0.0 && P.radP2 - q.r2 >= 0) {
                    const double r = sqrt(q.r2);
                    const double dw = dwq_k(P.Kp, P.m2P, r, P.radP);
                    const double ev[3] = {q.x[0] / r, q.x[1] / r, q.x[2] / r};
                    for (int d = 0; d < 3; ++d) {
                        const double u = vel[3 * j + d] - vel[3 * i + d];
                        t[d] = u * ev[d] * dw;
                    }
                    f = 1;
                }
            }
        }
        for (int e = 0; e < K; ++e) tb[K * threadIdx.x + e] = t[e];
        tf[threadIdx.x] = f;
        wave_sync();
        if (g == 0)
            for (int u = 0; u < G && base + u < n; ++u)
                if (tf[threadIdx.x + u])
                    for (int d = 0; d < 3; ++d) sum -= tb[K * (threadIdx.x + u) + d];
        wave_sync();
    }
    if (g != 0 || !own) return;
    divP[i] = sum;
    double pp = -lambda[i] * sum;
    if (volStrainP[i] > 0.0) pp += kappa[i] * volStrainP[i];
    pressP[i] = pp;
}

// kSurf false: every type's CofA is 0.  The diffuse-interface virial terms are then
// zeros of either sign, sg stays +0, and v (which starts at +0, so is never -0) keeps
// its bits under v += sg: the terms are not formed.
template <bool kSurf>
__global__ void k_diag_forces(DevParams P, const int* __restrict__ type, const double* __restrict__ pos,
                              const double* __restrict__ vel, const double* __restrict__ mu,
                              const double* __restrict__ pressA, const double* __restrict__ gcen,
                              const double* __restrict__ pressP, const int* __restrict__ nCount,
                              const long long* __restrict__ nPtr, const int* __restrict__ nInd,
                              const int* __restrict__ nCountP, const long long* __restrict__ nPtrP,
                              const int* __restrict__ nIndP, double* __restrict__ force,
                              double* __restrict__ virialStress, double* __restrict__ virialP) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    const int ti = type[i];
    const double V = P.V, VoL = P.V / P.l;
    const double ppi = pressP[i], pai = pressA[i], mui = mu[i];
    double gi[3] = {0.0, 0.0, 0.0};
    if (kSurf)
        for (int d = 0; d < 3; ++d) gi[d] = gcen[3 * i + d];
    const double vi[3] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    const double a_ = P.ai[ti];
    double st1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    {
        const long long b = nPtrP[i], e = b + nCountP[i];
        for (long long k = b; k < e; ++k) {  // pressureP virial (4512-4555)
            const int j = nIndP[k];
            if (i == j) continue;
            const Rel q = rel(P, pos, i, j);
            if (q.r2 == 0.0) continue;
            if (P.radP2 - q.r2 > 0) {
                const double r = sqrt(q.r2);
                const double dw = dwq_k(P.Kp, P.m2P, r, P.radP);
                const double gw[3] = {dw * q.x[0] / r, dw * q.x[1] / r, dw * q.x[2] / r};
                double fij[3];
                for (int d = 0; d < 3; ++d) fij[d] = (ppi) * gw[d] * V;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c) st1[3 * a + c] += 1.0 * fij[a] * q.x[c] / V;
            }
        }
    }
    double fp[3] = {0, 0, 0}, fv[3] = {0, 0, 0};
    double sa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, sg[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const long long b = nPtr[i], e = b + nCount[i];
    for (long long k = b; k < e; ++k) {
        const int j = nInd[k];
        if (i == j) continue;
        const Rel q = rel(P, pos, i, j);
        if (q.r2 == 0.0) continue;
        const double r = sqrt(q.r2);
        if (P.radP2 - q.r2 > 0) {  // pressure force (1983-2014)
            const double dw = dwq_k(P.Kp, P.m2P, r, P.radP);
            const double gw[3] = {dw * q.x[0] / r, dw * q.x[1] / r, dw * q.x[2] / r};
            const double ppj = pressP[j];
            for (int d = 0; d < 3; ++d) fp[d] += (ppi + ppj) * gw[d] * V;
        }
        if (P.radV2 - q.r2 > 0) {  // viscosity force (2036-2056) and virial (4620-4645)
            const double dwij = -dwq_k(P.Kv, P.m2V, r, P.radV);
            const double ev[3] = {q.x[0] / r, q.x[1] / r, q.x[2] / r};
            const double u[3] = {vel[3 * j] - vi[0], vel[3 * j + 1] - vi[1], vel[3 * j + 2] - vi[2]};
            const double muj = mu[j];
            const double muij = 2.0 * (mui * muj) / (mui + muj);
            double fij[3];
            for (int d = 0; d < 3; ++d) {
                fij[d] = P.visc * muij * (u[0] * ev[0] + u[1] * ev[1] + u[2] * ev[2]) * ev[d] * dwij / r * V;
                fv[d] += fij[d];
            }
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) sv[3 * a + c] += 0.5 * fij[a] * q.x[c] / V;
        }
        if (P.radA2 - q.r2 > 0) {  // pressureA virial (4575-4592)
            const double ratio = P.ratio[ti][type[j]];
            const double dw = ratio * dwadr_k(P.Ka, P.invA, r, P.radA);
            const double gw[3] = {dw * q.x[0] / r, dw * q.x[1] / r, dw * q.x[2] / r};
            double fij[3];
            for (int d = 0; d < 3; ++d) fij[d] = (pai) * gw[d] * V;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) sa[3 * a + c] += 1.0 * fij[a] * q.x[c] / V;
        }
        if (kSurf && P.radG2 - q.r2 > 0) {  // diffuse-interface virial, both terms (4672-4716)
            const double ratio = P.ratio[ti][type[j]];
            const double weight = ratio * wq_k(P.Kg, r, P.radG);
            double fij[3];
            for (int d = 0; d < 3; ++d) fij[d] = -a_ * (-gi[d]) * weight / P.R2g * P.radG * VoL;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) sg[3 * a + c] += 1.0 * fij[a] * q.x[c] / V;
            const double dw = ratio * dwq_k(P.Kg, P.m2G, r, P.radG);
            const double gw[3] = {dw * q.x[0] / r, dw * q.x[1] / r, dw * q.x[2] / r};
            double gr = 0.0;
            for (int d = 0; d < 3; ++d) gr += (-gi[d]) * q.x[d];
            for (int d = 0; d < 3; ++d) fij[d] = -a_ * (gr) * gw[d] / P.R2g * P.radG * VoL;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) sg[3 * a + c] += 1.0 * fij[a] * q.x[c] / V;
        }
    }
    for (int d = 0; d < 3; ++d) {
        double F = force[3 * i + d];
        F += fp[d];
        F += fv[d];
        force[3 * i + d] = F;
    }
    double vs[9];
    for (int k = 0; k < 9; ++k) {
        double v = 0.0;
        v += st1[k];
        v += sa[k];
        v += sv[k];
        v += sg[k];
        vs[k] = v;
        virialStress[9 * (long long)i + k] = v;
    }
    virialP[i] = P.dim == 2 ? -1.0 / 2.0 * (vs[0] + vs[4]) : -1.0 / 3.0 * (vs[0] + vs[4] + vs[8]);
}

}  // namespace

void launch_fluid_bbox(hipStream_t st, const DevParams& P, const int* cellStart, int* bbox) {
    hipLaunchKernelGGL(k_fluid_bbox, dim3(nblk(P.T, 256)), dim3(256), 0, st, P, cellStart, bbox);
}
void launch_inv_diag(hipStream_t st, const DevParams& P, const int* selfIdx, ASlices A, double* invDiagA) {
    hipLaunchKernelGGL(k_inv_diag, dim3(nblk(P.nf, 256)), dim3(256), 0, st, P, selfIdx, A.start, A.val, invDiagA);
}
void launch_mg_galerkin0(hipStream_t st, const DevParams& P, const Level& L, const int* cellStart,
                         const unsigned short* runStart, const int* nFluid, ASlices A, int variant) {
    const int threads = 9 * L.cs[0] * L.cs[1] * L.cs[2];  // <= 243
    if (L.cells <= 0) return;
    if (variant == 1)
        hipLaunchKernelGGL(k_mg_galerkin0_lds, dim3(L.cells), dim3(256), 0, st, P, L, cellStart, runStart, nFluid,
                           A.start, A.val);
    else
        hipLaunchKernelGGL(k_mg_galerkin0, dim3(L.cells), dim3((threads + 63) / 64 * 64), 0, st, P, L, cellStart,
                           runStart, A.start, A.val);
}
void launch_mg_galerkinL(hipStream_t st, const Level& S, const Level& L) {
    const long long n = (long long)L.cells * L.cs[0] * L.cs[1] * L.cs[2];
    hipLaunchKernelGGL(k_mg_galerkinL, dim3(nblk(n, 128)), dim3(128), 0, st, S, L);
}
void launch_mg_invd(hipStream_t st, const Level& L) {
    hipLaunchKernelGGL(k_mg_invd, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
}
void launch_fill(hipStream_t st, double* x, long long n, double v) {
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, x, n, v);
}
void launch_fill_i32(hipStream_t st, int* x, long long n, int v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(n, 256)), dim3(256), 0, st, x, n, v);
}
void launch_mg_p2g(hipStream_t st, const DevParams& P, Level& L, const int* cellStart, const double* q) {
    hipLaunchKernelGGL(k_mg_p2g, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, P, L, cellStart, q);
}
void launch_mg_p2g_bu(hipStream_t st, const DevParams& P, Level& L, const int* cellStart, const double* q) {
    // not written without kPq
    hipLaunchKernelGGL(k_mg_p2g_bu_lanes<false>, dim3(nblk((long long)kP2gLanes * L.nrow, 256)), dim3(256), 0, st, P,
                       L, cellStart, const_cast<double*>(q), nullptr, nullptr);
}
void launch_mg_p2g_bu_pq(hipStream_t st, const DevParams& P, Level& L, const int* cellStart, double* q,
                         const double* y, const double* scal) {
    hipLaunchKernelGGL(k_mg_p2g_bu_lanes<true>, dim3(nblk((long long)kP2gLanes * L.nrow, 256)), dim3(256), 0, st, P,
                       L, cellStart, q, y, scal);
}
void launch_mg_restrict_bu(hipStream_t st, const Level& S, Level& L) {
    hipLaunchKernelGGL(k_mg_restrict_bu, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, S, L);
}
void launch_mg_begin(hipStream_t st, Level& L) {
    hipLaunchKernelGGL(k_mg_begin, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
}
void launch_mg_update(hipStream_t st, Level& L) {
    hipLaunchKernelGGL(k_mg_update, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
}
void launch_mg_residual(hipStream_t st, Level& L) {
    if (launch_sweep_lanes<false>(st, L)) return;
    if (L.f32) hipLaunchKernelGGL(k_mg_residual<float>, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
    else hipLaunchKernelGGL(k_mg_residual<double>, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
}
void launch_mg_residual_update(hipStream_t st, Level& L) {
    if (!launch_sweep_lanes<true>(st, L)) {
        if (L.f32) hipLaunchKernelGGL(k_mg_residual_update<float>, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
        else hipLaunchKernelGGL(k_mg_residual_update<double>, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, L);
    }
    L.cur ^= 1;
}
void launch_mg_restrict(hipStream_t st, const Level& S, Level& L) {
    hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(L.nrow, 256)), dim3(256), 0, st, S, L);
}
void launch_mg_prolong(hipStream_t st, const Level& L, Level& S) {
    hipLaunchKernelGGL(k_mg_prolong, dim3(nblk(S.nrow, 256)), dim3(256), 0, st, L, S);
}
void launch_mg_coarse(hipStream_t st, const Level* lvs, int lc, int D, unsigned cur, bool f32, const int* halt) {
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kCoarseThreads), 0, st, lvs, lc, D, cur, f32 ? 1 : 0, halt);
}
void launch_mg_g2p(hipStream_t st, const DevParams& P, const Level& L0, const int* cellIndex,
                   const double* invDiagA, const double* q, double* u, double* part) {
    hipLaunchKernelGGL(k_mg_g2p, dim3(nblk(P.nf, 256)), dim3(256), 0, st, P, L0, cellIndex, invDiagA, q, u, part);
}
void launch_cr_init(hipStream_t st, const DevParams& P, const double* vel, const double* force,
                    const double* mass, double* x) {
    hipLaunchKernelGGL(k_cr_init, dim3(nblk(P.nf, 256)), dim3(256), 0, st, P, vel, force, mass, x);
}
template <bool kNeg, class VT, int kUn>
static void spmv_launch(hipStream_t st, const DevParams& P, const ASlices& A, const int* nFluid, const VT* val,
                        const double* x, double beta, const double* yin, double* y, double* y2, double* p_copy,
                        const double* dotx, double* part) {
    hipLaunchKernelGGL((k_spmv<kNeg, VT, kUn>), dim3(nblk(P.nf, kChunk)), dim3(kChunk), 0, st, P.nf, nFluid, A.start,
                       A.col, val, x, beta, yin, y, y2, p_copy, dotx, part, P.halt);
}

void launch_spmv(hipStream_t st, const DevParams& P, ASlices A, const int* nFluid, double alpha, const double* x,
                 double beta, const double* yin, double* y, double* y2, double* p_copy, const double* dotx,
                 double* part, bool f32, int variant) {
    // alpha is +1 or -1 at every call site (main.cpp:4337, 4353, 4357)
    if (P.nf == 0) return;
#define MPH_SPMV(NEG, VT, UN, VAL) spmv_launch<NEG, VT, UN>(st, P, A, nFluid, VAL, x, beta, yin, y, y2, p_copy, dotx, part)
    if (f32) {
        const int un = variant == 1 ? 2 : 1;
        if (alpha < 0) { if (un == 2) MPH_SPMV(true, float, 2, A.val32); else MPH_SPMV(true, float, 1, A.val32); }
        else { if (un == 2) MPH_SPMV(false, float, 2, A.val32); else MPH_SPMV(false, float, 1, A.val32); }
    } else {
        const int un = variant == 1 ? 3 : variant == 2 ? 4 : 2;
        if (alpha < 0) {
            if (un == 3) MPH_SPMV(true, double, 3, A.val);
            else if (un == 4) MPH_SPMV(true, double, 4, A.val);
            else MPH_SPMV(true, double, 2, A.val);
        } else {
            if (un == 3) MPH_SPMV(false, double, 3, A.val);
            else if (un == 4) MPH_SPMV(false, double, 4, A.val);
            else MPH_SPMV(false, double, 2, A.val);
        }
    }
#undef MPH_SPMV
}

__global__ void k_convert_f32_(const double* __restrict__ in, float* __restrict__ out, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = (float)in[k];
}
void launch_convert_f32(hipStream_t st, const double* in, float* out, long long n) {
    if (n > 0) hipLaunchKernelGGL(k_convert_f32_, dim3(nblk(n, 256)), dim3(256), 0, st, in, out, n);
}

__global__ void k_axpy1_(long long n, const double* __restrict__ x, double* __restrict__ y) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] += 1.0 * x[k];
}
void launch_axpy(hipStream_t st, long long n, const double* x, double* y) {
    if (n > 0) hipLaunchKernelGGL(k_axpy1_, dim3(nblk(n, 256)), dim3(256), 0, st, n, x, y);
}
void launch_dot_partial(hipStream_t st, const DevParams& P, const double* x, const double* y, double* part) {
    hipLaunchKernelGGL(k_dot_partial, dim3(nblk(P.nf, kChunk)), dim3(kChunk), 0, st, P.nf, x, y, part, P.halt);
}void launch_dot_final(hipStream_t st, const DevParams& P, const double* part, double* scal, int slot,
                      int shift_slot, double* host) {
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, st, part, nblk(P.nf, kChunk), scal, slot, shift_slot,
                       host, P.halt);
}
void launch_dot_final_record(hipStream_t st, const DevParams& P, const double* part, double* scal, const int* ctl,
                             long long it, double* rec, long long cap) {
    hipLaunchKernelGGL(k_dot_final_record, dim3(1), dim3(256), 0, st, part, nblk(P.nf, kChunk), scal, ctl, it, rec,
                       cap, P.halt);
}
void launch_dot_final_stop(hipStream_t st, const DevParams& P, const double* part, double* scal, int* ctl, int step,
                           double tol, long long maxit, int* hctl, double* hrr) {
    hipLaunchKernelGGL(k_dot_final_stop, dim3(1), dim3(256), 0, st, part, nblk(P.nf, kChunk), scal, ctl, step, tol,
                       maxit, hctl, hrr);
}
void launch_pq_update(hipStream_t st, const DevParams& P, const double* scal, const double* s, const double* y,
                      double* p, double* q) {
    hipLaunchKernelGGL(k_pq_update, dim3(nblk(3LL * P.nf, 256)), dim3(256), 0, st, P.nf, scal, s, y, p, q, P.halt);
}
void launch_xrs_update(hipStream_t st, const DevParams& P, const double* scal, double* p, const double* q,
                       const double* u, double* x, double* r, double* s, double* part, bool pq) {
    if (pq)
        hipLaunchKernelGGL(k_xrs_update<true>, dim3(nblk(P.nf, kChunk)), dim3(kChunk), 0, st, P.nf, scal, p, q, u, x,
                           r, s, part, P.halt);
    else
        hipLaunchKernelGGL(k_xrs_update<false>, dim3(nblk(P.nf, kChunk)), dim3(kChunk), 0, st, P.nf, scal, p, q, u, x,
                           r, s, part, P.halt);
}
void launch_store_velocity(hipStream_t st, const DevParams& P, const double* x, double* vel) {
    hipLaunchKernelGGL(k_store_velocity, dim3(nblk(3LL * P.nf, 256)), dim3(256), 0, st, P.nf, x, vel);
}
void launch_diagnostics(hipStream_t st, const DevParams& P, const int* type, const double* pos, const double* vel,
                        const double* mu, const double* lambda, const double* kappa, const double* volStrainP,
                        const double* pressA, const double* gcen, const int* nCount, const long long* nPtr,
                        const int* nInd, const int* nCountP, const long long* nPtrP, const int* nIndP, double* divP,
                        double* pressP, double* force, double* virialStress, double* virialP) {
    hipLaunchKernelGGL(k_divergence, dim3(nblk(4LL * P.np, 256)), dim3(256), 0, st, P, pos, vel, lambda, kappa,
                       volStrainP, nCount, nPtr, nInd, divP, pressP);
    const bool surf = !(P.ai[0] == 0.0 && P.ai[1] == 0.0 && P.ai[2] == 0.0 && P.ai[3] == 0.0);
    if (surf)
        hipLaunchKernelGGL(k_diag_forces<true>, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, type, pos, vel, mu, pressA,
                           gcen, pressP, nCount, nPtr, nInd, nCountP, nPtrP, nIndP, force, virialStress, virialP);
    else
        hipLaunchKernelGGL(k_diag_forces<false>, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, type, pos, vel, mu,
                           pressA, gcen, pressP, nCount, nPtr, nInd, nCountP, nPtrP, nIndP, force, virialStress,
                           virialP);
}

}  // namespace mph
