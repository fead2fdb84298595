// mph_cr_body.h -- per-row / per-chunk bodies of the multigrid V-cycle and the
// conjugate-residual loop (source/main.cpp:4122-4421).
//
// The single-scene kernels (mph_solver.hip) and the scene-table kernels of the
// batched executor (mph_batch.hip) both call these functions, so a scene's
// arithmetic is the same instruction sequence whether it runs alone or in a batch.
// Every body keeps the reference's operand order (y = beta*y + sum, with the sum
// formed first, as myDcsrmv / myDdmv / myDcopy / myDscal of main.cpp:4015-4119).
#pragma once

#include "mph_kernels.h"

namespace mph {
namespace body {

__host__ __device__ inline int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// CR scalars: rr0, rs0, rr, rho, rho_prev, tmp (= q.u), rs
enum { kRR0 = 0, kRS0 = 1, kRR = 2, kRHO = 3, kRHOP = 4, kTMP = 5, kRS = 6, kRSIter = 7, kNScal = 8 };

__device__ inline int cell_id(const DevParams& P, int cx, int cy, int cz) {
    const int* C = P.C;
    if (P.dim == 2) return ((cx) % C[0] + C[0]) % C[0] * C[1] + ((cy) % C[1] + C[1]) % C[1];
    return ((cx) % C[0] + C[0]) % C[0] * C[1] * C[2] + ((cy) % C[1] + C[1]) % C[1] * C[2] +
           ((cz) % C[2] + C[2]) % C[2];
}

// a level's S buffer by ping-pong parity, as a select: indexing the two-pointer array
// with a run-time value would keep a Level copied into a local in private memory
__device__ inline double* lv_s(const Level& L, unsigned parity) { return parity ? L.S[1] : L.S[0]; }

__device__ inline void grid_xyz(const Level& L, int g, int& gx, int& gy, int& gz) {
    gx = (g / (L.cnt[1] * L.cnt[2])) % L.cnt[0];
    gy = (g / L.cnt[2]) % L.cnt[1];
    gz = g % L.cnt[2];
}
__device__ inline int grid_id(const Level& L, int gx, int gy, int gz) {
    return (L.cnt[1] * L.cnt[2]) * gx + L.cnt[2] * gy + gz;
}

// pairwise tree over the 256 values of a workgroup (t[l] += t[l+h], h = 128..1);
// the result is valid in thread 0.  Every thread of the workgroup must call it.
__device__ inline double block_tree_sum(double t, double* sh) {
    sh[threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x < 128) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + 128];
    __syncthreads();
    double v = 0.0;
    if (threadIdx.x < 64) {
        v = sh[threadIdx.x] + sh[threadIdx.x + 64];
        for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
    }
    return v;
}

// which of the visited neighbour cells (ascending CellId, as the neighbour search)
// of cell g map to stencil cell c of level 0: bit q = run q of every particle's list
__device__ inline unsigned galerkin_mask(const DevParams& P, const Level& L, const int* __restrict__ cellStart,
                                         int g, int c) {
    int gx, gy, gz;
    grid_xyz(L, g, gx, gy, gz);
    const int cx = gx + L.mn[0], cy = gy + L.mn[1], cz = gz + L.mn[2];
    unsigned mask = 0;
    const int* C = P.C;
    const int rx = P.range[0], ry = P.range[1], rz = P.range[2];
    int jx[3], jy[3], jz[3];
    for (int k = 0; k < 2 * rx + 1; ++k) jx[k] = ((cx - rx + k) % C[0] + C[0]) % C[0];
    for (int k = 0; k < 2 * ry + 1; ++k) jy[k] = ((cy - ry + k) % C[1] + C[1]) % C[1];
    for (int k = 0; k < 2 * rz + 1; ++k) jz[k] = ((cz - rz + k) % C[2] + C[2]) % C[2];
    const int bx = (2 * rx) - (cx + rx) % C[0], by = (2 * ry) - (cy + ry) % C[1], bz = (2 * rz) - (cz + rz) % C[2];
    const int x0 = bx > 0 ? bx : 0, y0 = by > 0 ? by : 0, z0 = bz > 0 ? bz : 0;
    int q = 0;
    for (int a = x0; a < x0 + (2 * rx + 1); ++a)
        for (int b = y0; b < y0 + (2 * ry + 1); ++b)
            for (int cc = z0; cc < z0 + (2 * rz + 1); ++cc, ++q) {
                const int ax = jx[a % (2 * rx + 1)], ay = jy[b % (2 * ry + 1)], az = jz[cc % (2 * rz + 1)];
                const int jc = cell_id(P, ax, ay, az);
                if (cellStart[jc] == cellStart[jc + 1]) continue;  // no fluid there
                const int dx = (ax - L.mn[0]) - gx, dy = (ay - L.mn[1]) - gy, dz = (az - L.mn[2]) - gz;
                const int ix = (dx + 1 + L.cnt[0]) % L.cnt[0];
                const int iy = (dy + 1 + L.cnt[1]) % L.cnt[1];
                const int iz = (dz + 1 + L.cnt[2]) % L.cnt[2];
                if (L.cs[2] * L.cs[1] * ix + L.cs[2] * iy + iz == c) mask |= 1u << q;
            }
    return mask;
}

// galerkin_mask's q of cell g's own cell: the run that holds each of its particles'
// self entry
__device__ inline int galerkin_self(const DevParams& P, const Level& L, int g) {
    int gx, gy, gz;
    grid_xyz(L, g, gx, gy, gz);
    const int cx = gx + L.mn[0], cy = gy + L.mn[1], cz = gz + L.mn[2];
    const int* C = P.C;
    const int rx = P.range[0], ry = P.range[1], rz = P.range[2];
    const int nx = 2 * rx + 1, ny = 2 * ry + 1, nz = 2 * rz + 1;
    const int bx = (2 * rx) - (cx + rx) % C[0], by = (2 * ry) - (cy + ry) % C[1], bz = (2 * rz) - (cz + rz) % C[2];
    const int x0 = bx > 0 ? bx : 0, y0 = by > 0 ? by : 0, z0 = bz > 0 ? bz : 0;
    // the own cell is the one visited with offset index r (jx[r] = cx): a % n == r
    auto at = [](int r, int o, int n) { return ((r - o) % n + n) % n; };
    return (at(rx, x0, nx) * ny + at(ry, y0, ny)) * nz + at(rz, z0, nz);
}

// ---------------------------------------------------------------------------
// multigrid rows (preconditionWithMultiGrid)

// P2G: Q0 = 0*Q0 + sum over the cell's fluid particles of q (2964-2997)
__device__ inline void mg_p2g_row(const DevParams& P, const Level& L, const int* __restrict__ cellStart,
                                  const double* __restrict__ q, int row) {
    const int g = row / 3, rd = row % 3;
    int gx, gy, gz;
    grid_xyz(L, g, gx, gy, gz);
    const int ic = cell_id(P, gx + L.mn[0], gy + L.mn[1], gz + L.mn[2]);
    double sum = 0.0;
    for (int ip = cellStart[ic]; ip < cellStart[ic + 1]; ++ip) sum += 1.0 * 1.0 * q[3 * ip + rd];
    double y = L.Q[row];
    y *= 0.0;
    y += sum;
    L.Q[row] = y;
}

__device__ inline void mg_begin_row(const Level& L, int row) {  // myDscal(0.0, S); myDcopy(Q, R)
    double* S = lv_s(L, L.cur);
    S[row] *= 0.0;
    L.R[row] = L.Q[row];
}

__device__ inline void mg_update_row(const Level& L, int row) {  // myDdmv(1.0, invD, R, 1.0, S)
    double* S = lv_s(L, L.cur);
    double y = S[row];
    y *= 1.0;
    y += 1.0 * L.invD[row] * L.R[row];
    S[row] = y;
}

// mg_begin_row then mg_update_row on the same row (both touch that row only)
__device__ inline void mg_begin_update_row(const Level& L, int row) {
    double* S = lv_s(L, L.cur);
    double y = S[row];
    y *= 0.0;
    const double r = L.Q[row];
    L.R[row] = r;
    y *= 1.0;
    y += 1.0 * L.invD[row] * r;
    S[row] = y;
}

// (v mod n) for 0 <= v < 3n without a division
__device__ inline int wrap3(int v, int n) {
    v -= v >= n ? n : 0;
    v -= v >= n ? n : 0;
    return v;
}

// sum over the 27-point (9 in 2D) block stencil of -1.0*a*x in shift order; the
// wrapped neighbour index (g + e - 1 + cnt) % cnt of the reference's periodic
// stencil (main.cpp:3323-3325) is formed by subtraction (its argument is < 3 cnt).
// Full stencils (3x3x3, 3x3x1) take an unrolled path that issues the loads of nine
// cells at a time before summing them in order; smaller levels take the loop.
template <class VT, int C0, int C1, int C2>
__device__ inline double mg_row_sum_full(const Level& L, const VT* __restrict__ A, const double* __restrict__ x,
                                         int row, int gx, int gy, int gz) {
    constexpr int N = C0 * C1 * C2;
    const int c0 = L.cnt[0], c1 = L.cnt[1], c2 = L.cnt[2];
    const long long nrow = L.nrow;
    int jx[C0], jy[C1], jz[C2];
#pragma unroll
    for (int e = 0; e < C0; ++e) jx[e] = wrap3(gx + e - 1 + c0, c0);
#pragma unroll
    for (int e = 0; e < C1; ++e) jy[e] = wrap3(gy + e - 1 + c1, c1);
#pragma unroll
    for (int e = 0; e < C2; ++e) jz[e] = wrap3(gz + e - 1 + c2, c2);
    double sum = 0.0;
#pragma unroll
    for (int q0 = 0; q0 < N; q0 += 9) {
        VT av[27];
        double xv[27];
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int q = q0 + t;
            if (q < N) {
                const int ex = q / (C1 * C2), ey = (q / C2) % C1, ez = q % C2;
                const double* xc = x + 3 * ((c1 * c2) * jx[ex] + c2 * jy[ey] + jz[ez]);
#pragma unroll
                for (int sd = 0; sd < 3; ++sd) {
                    av[3 * t + sd] = A[(3 * q + sd) * nrow + row];
                    xv[3 * t + sd] = xc[sd];
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 9; ++t)
            if (q0 + t < N)
#pragma unroll
                for (int sd = 0; sd < 3; ++sd) sum += -1.0 * (double)av[3 * t + sd] * xv[3 * t + sd];
    }
    return sum;
}

template <class VT>
__device__ inline double mg_row_sum(const Level& L, const VT* __restrict__ A, const double* __restrict__ x, int row) {
    const int g = row / 3;
    int gx, gy, gz;
    grid_xyz(L, g, gx, gy, gz);
    if (L.cs[0] == 3 && L.cs[1] == 3 && L.cs[2] == 3) return mg_row_sum_full<VT, 3, 3, 3>(L, A, x, row, gx, gy, gz);
    if (L.cs[0] == 3 && L.cs[1] == 3 && L.cs[2] == 1) return mg_row_sum_full<VT, 3, 3, 1>(L, A, x, row, gx, gy, gz);
    const int c0 = L.cnt[0], c1 = L.cnt[1], c2 = L.cnt[2];
    const int nrow = L.nrow;
    const VT* a = A + row;  // entry sh of the row at a[sh * nrow]
    double sum = 0.0;
    for (int ex = 0; ex < L.cs[0]; ++ex) {
        const int jx = wrap3(gx + ex - 1 + c0, c0);
        for (int ey = 0; ey < L.cs[1]; ++ey) {
            const int jy = wrap3(gy + ey - 1 + c1, c1);
            const int bxy = (c1 * c2) * jx + c2 * jy;
            for (int ez = 0; ez < L.cs[2]; ++ez) {
                const int jz = wrap3(gz + ez - 1 + c2, c2);
                const double* xc = x + 3 * (bxy + jz);
                for (int sd = 0; sd < 3; ++sd, a += nrow) sum += -1.0 * (double)a[0] * xc[sd];
            }
        }
    }
    return sum;
}

// myDcopy(Q, R); myDcsrmv(-1.0, A, S, 1.0, R)
template <class VT>
__device__ inline void mg_residual_row(const Level& L, int row) {
    const VT* A = sizeof(VT) == 8 ? (const VT*)L.A : (const VT*)L.A32;
    const double sum = mg_row_sum(L, A, lv_s(L, L.cur), row);
    double y = L.Q[row];
    y *= 1.0;
    y += sum;
    L.R[row] = y;
}

// residual then smoother update of the same row; the new iterate goes to the other
// S buffer so every thread still reads the old one
template <class VT>
__device__ inline void mg_residual_update_row(const Level& L, int row) {
    const double* So = lv_s(L, L.cur);
    const VT* A = sizeof(VT) == 8 ? (const VT*)L.A : (const VT*)L.A32;
    const double sum = mg_row_sum(L, A, So, row);
    double y = L.Q[row];
    y *= 1.0;
    y += sum;
    L.R[row] = y;
    double s = So[row];
    s *= 1.0;
    s += 1.0 * L.invD[row] * y;
    lv_s(L, L.cur ^ 1)[row] = s;
}

// restriction (R of 3115-3175): Q_L = 0*Q_L + sum of the clipped 2x2x2 children of R_S
__device__ inline void mg_restrict_row(const Level& S, const Level& L, int row) {
    const int g = row / 3, rd = row % 3;
    int lx, ly, lz;
    grid_xyz(L, g, lx, ly, lz);
    const int lcx = lx + L.mn[0], lcy = ly + L.mn[1], lcz = lz + L.mn[2];
    const int fmx = S.mn[0] > 2 * lcx, fmy = S.mn[1] > 2 * lcy, fmz = S.mn[2] > 2 * lcz;
    const int fMx = S.mx[0] < 2 * (lcx + 1), fMy = S.mx[1] < 2 * (lcy + 1), fMz = S.mx[2] < 2 * (lcz + 1);
    double sum = 0.0;
    for (int ax = 0; ax < 2 - fmx - fMx; ++ax)
        for (int ay = 0; ay < 2 - fmy - fMy; ++ay)
            for (int az = 0; az < 2 - fmz - fMz; ++az) {
                const int sg = grid_id(S, (2 * lcx + ax + fmx) - S.mn[0], (2 * lcy + ay + fmy) - S.mn[1],
                                       (2 * lcz + az + fmz) - S.mn[2]);
                sum += 1.0 * 1.0 * S.R[3 * sg + rd];
            }
    double y = L.Q[row];
    y *= 0.0;
    y += sum;
    L.Q[row] = y;
}

// prolongation (P of 3216-3241): S_S += injection of the parent's S_L
__device__ inline void mg_prolong_row(const Level& L, const Level& S, int row) {
    const int g = row / 3, rd = row % 3;
    int sx, sy, sz;
    grid_xyz(S, g, sx, sy, sz);
    const int lg = grid_id(L, ((sx + S.mn[0]) / 2) - L.mn[0], ((sy + S.mn[1]) / 2) - L.mn[1],
                           ((sz + S.mn[2]) / 2) - L.mn[2]);
    double sum = 0.0;
    sum += 1.0 * 1.0 * lv_s(L, L.cur)[3 * lg + rd];
    double* Sv = lv_s(S, S.cur);
    double y = Sv[row];
    y *= 1.0;
    y += sum;
    Sv[row] = y;
}

// Levels lc..D-1 of the V-cycle inside ONE workgroup: restriction into lc, the
// down sweep (S = invD R, then two residual-update sweeps and a residual per
// level), the coarsest level (S = invD R, then four residual-update sweeps), the
// up sweep (three residual-update sweeps per level) and the prolongations back
// into lc-1 -- the order of preconditionWithMultiGrid (4160-4255) for those levels.  Each step runs over the level's rows with a
// stride of blockDim.x and ends at a workgroup barrier, so every row sees the
// previous step complete.  `cur` holds the ping-pong state of each level (bit l)
// on entry, lc-1 included; it advances as the separate launches would advance it
// (afterwards every fused level but the coarsest has flipped).  f32: use a level's FP32
// stencil where it has one.
// The workgroup has kCoarseThreads threads: few enough that mg_row_sum's batches of
// nine stencil cells (54 values in flight) fit in registers, so a sweep waits for
// three rounds of loads per row.  (With 1024 threads and three cells per batch it
// waited for nine: 68 us per cycle on the 1M-particle scene, levels 3 and 4.)
constexpr int kCoarseThreads = 256;

__device__ inline void mg_coarse_cycle(const Level* lvs, int lc, int D, unsigned cur, bool f32) {
    auto lev = [&](int l) {
        Level L = lvs[l];
        L.cur = (cur >> l) & 1u;
        return L;
    };
    auto sweep = [&](const Level& L, int row, bool update) {  // mg_residual(_update)_row
        const double* So = lv_s(L, L.cur);
        const double sum = f32 && L.A32 ? mg_row_sum<float>(L, L.A32, So, row) : mg_row_sum<double>(L, L.A, So, row);
        double y = L.Q[row];
        y *= 1.0;
        y += sum;
        L.R[row] = y;
        if (update) {
            double s = So[row];
            s *= 1.0;
            s += 1.0 * L.invD[row] * y;
            lv_s(L, L.cur ^ 1)[row] = s;
        }
    };
    auto ru = [&](int l) {
        const Level L = lev(l);
        for (int row = threadIdx.x; row < L.nrow; row += blockDim.x) sweep(L, row, true);
        __syncthreads();
        cur ^= 1u << l;
    };
    auto restrict_into = [&](int l) {
        const Level L = lev(l);
        for (int row = threadIdx.x; row < L.nrow; row += blockDim.x) mg_restrict_row(lvs[l - 1], L, row);
        __syncthreads();
    };
    auto begin_update = [&](int l) {
        const Level L = lev(l);
        for (int row = threadIdx.x; row < L.nrow; row += blockDim.x) mg_begin_update_row(L, row);
        __syncthreads();
    };
    auto prolong_from = [&](int l) {  // l -> l-1
        const Level L = lev(l), S = lev(l - 1);
        for (int row = threadIdx.x; row < S.nrow; row += blockDim.x) mg_prolong_row(L, S, row);
        __syncthreads();
    };
    for (int l = lc; l < D - 1; ++l) {
        restrict_into(l);
        begin_update(l);
        ru(l);
        ru(l);
        const Level L = lev(l);
        for (int row = threadIdx.x; row < L.nrow; row += blockDim.x) sweep(L, row, false);
        __syncthreads();
    }
    {
        const int c = D - 1;
        restrict_into(c);
        begin_update(c);
        for (int k = 0; k < 4; ++k) ru(c);
        prolong_from(c);
    }
    for (int l = D - 2; l >= lc; --l) {
        for (int k = 0; k < 3; ++k) ru(l);
        prolong_from(l);
    }
}

// u = D^-1 q (4144-4148) + G2P S0 (4260), and the chunk partial of q.u (when part)
__device__ inline void mg_g2p_chunk(const DevParams& P, const Level& L, const int* __restrict__ cellIndex,
                                    const double* __restrict__ invDiagA, const double* __restrict__ q,
                                    double* __restrict__ u, double* __restrict__ part, int chunk, double* sh) {
    const int i = chunk * 256 + threadIdx.x;
    double t = 0.0;
    if (i < P.nf) {
        const int key = cellIndex[i];
        const int cx = key / (P.C[1] * P.C[2]) % P.C[0], cy = key / (P.C[2]) % P.C[1], cz = key % P.C[2];
        const int g = grid_id(L, cx - L.mn[0], cy - L.mn[1], cz - L.mn[2]);
        const double* S0 = lv_s(L, L.cur);
        double uu[3], qq[3];
        for (int r = 0; r < 3; ++r) {
            qq[r] = q[3 * i + r];
            double y = u[3 * i + r];
            y *= 0.0;
            y += 1.0 * invDiagA[3 * i + r] * qq[r];
            double sum = 0.0;
            sum += 1.0 * 1.0 * S0[3 * g + r];
            y *= 1.0;
            y += sum;
            uu[r] = y;
            u[3 * i + r] = y;
        }
        t = (qq[0] * uu[0] + qq[1] * uu[1]) + qq[2] * uu[2];
    }
    if (part) {
        const double v = block_tree_sum(t, sh);
        if (threadIdx.x == 0) part[chunk] = v;
    }
}

// ---------------------------------------------------------------------------
// conjugate residual (solveWithConjugatedGradient, 4270-4421)

// initial guess x = v - F/m dt (4299-4309)
__device__ inline void cr_init_particle(const DevParams& P, const double* __restrict__ vel,
                                        const double* __restrict__ force, const double* __restrict__ mass,
                                        double* __restrict__ x, int i) {
    for (int d = 0; d < 3; ++d) x[3 * i + d] = vel[3 * i + d] - force[3 * i + d] / mass[i] * P.dt;
}

// The tail of every particle-matrix product, given particle i's row sums (s0, s1,
// s2): y = beta*yin + sum (the sum formed first, as myDcsrmvForA), optional second
// output y2 = beta*y2 + sum, optional copy p = x, and the chunk partial of dotx.y.
// Every thread of the workgroup calls it (i >= nf: no row).
__device__ inline void spmv_store(int nf, int i, double s0, double s1, double s2, const double* __restrict__ x,
                                  double beta, const double* yin, double* y, double* y2, double* p_copy,
                                  const double* __restrict__ dotx, double* __restrict__ part, int chunk,
                                  double* sh) {
    double t = 0.0;
    if (i < nf) {
        const double sm[3] = {s0, s1, s2};
        double out[3];
        for (int r = 0; r < 3; ++r) {
            double yy = yin[3 * i + r];
            yy *= beta;
            yy += sm[r];
            out[r] = yy;
        }
        if (y2)
            for (int r = 0; r < 3; ++r) {
                double yy = y2[3 * i + r];
                yy *= beta;
                yy += sm[r];
                y2[3 * i + r] = yy;
            }
        for (int r = 0; r < 3; ++r) y[3 * i + r] = out[r];
        if (p_copy)
            for (int r = 0; r < 3; ++r) p_copy[3 * i + r] = x[3 * i + r];
        if (part) {
            const double d0 = dotx[3 * i], d1 = dotx[3 * i + 1], d2 = dotx[3 * i + 2];
            t = (d0 * out[0] + d1 * out[1]) + d2 * out[2];
        }
    }
    if (part) {
        const double v = block_tree_sum(t, sh);
        if (threadIdx.x == 0) part[chunk] = v;
    }
}

// y = beta*yin + sum_k (alpha*a_k)*x_k over the sliced A (myDcsrmvForA, 4033-4050);
// alpha is +1 or -1, so alpha*a is exact and is applied as a sign.  Optional second
// output y2 = beta*y2 + sum (same sum), optional copy p = x, and the chunk partial
// of dotx.y.  One chunk = 4 slices = 256 particles.  The row sums run over the
// blocks in list order, entries s = 0,1,2 within a block, as the CSR.
template <bool kNeg, class VT, int kUn>
__device__ inline void spmv_chunk(int nf, const int* __restrict__ nFluid, const long long* __restrict__ start,
                                  const int* __restrict__ col, const VT* __restrict__ val,
                                  const double* __restrict__ x, double beta, const double* yin, double* y,
                                  double* y2, double* p_copy, const double* __restrict__ dotx,
                                  double* __restrict__ part, int chunk, double* sh) {
    typedef int it2 __attribute__((ext_vector_type(2)));
    const int i = chunk * 256 + threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (i < nf) {
        const int w = nFluid[i];
        const int lane = i % kSlice;
        const long long sb = start[i / kSlice];
        // the matrix is read once per launch: streaming (nontemporal) loads of whole
        // block pairs (FP64, a_val_index) or quads (FP32, a_val32_index), kUn of them
        // per group with all their loads issued before the first use; columns are
        // pairs (a_col_index)
        const it2* cp = (const it2*)col + (sb / 2) * kSlice + lane;
        auto block = [&](int j, const double* c9) {
            const double x0 = x[3 * j], x1 = x[3 * j + 1], x2 = x[3 * j + 2];
            double c[9];
#pragma unroll
            for (int e = 0; e < 9; ++e) c[e] = kNeg ? -c9[e] : c9[e];
            s0 += c[0] * x0; s0 += c[1] * x1; s0 += c[2] * x2;
            s1 += c[3] * x0; s1 += c[4] * x1; s1 += c[5] * x2;
            s2 += c[6] * x0; s2 += c[7] * x1; s2 += c[8] * x2;
        };
        if constexpr (sizeof(VT) == 8) {
            typedef VT vt2 __attribute__((ext_vector_type(2)));
            const vt2* vp = (const vt2*)val + (sb / 2) * 9 * kSlice + lane;
            auto pair = [&](const it2 j, const vt2* a, int h) {  // blocks 2m, 2m+1 (h of them)
                for (int u = 0; u < h; ++u) {
                    double c9[9];
#pragma unroll
                    for (int e = 0; e < 9; ++e) c9[e] = (double)a[e][u];
                    block(j[u], c9);
                }
            };
            const int wp = w / 2;
            int m = 0;
            for (; m + kUn <= wp; m += kUn) {
                it2 j[kUn];
                vt2 a[kUn][9];
#pragma unroll
                for (int u = 0; u < kUn; ++u) j[u] = __builtin_nontemporal_load(cp + (long long)(m + u) * kSlice);
#pragma unroll
                for (int u = 0; u < kUn; ++u)
#pragma unroll
                    for (int e = 0; e < 9; ++e)
                        a[u][e] = __builtin_nontemporal_load(vp + ((long long)(m + u) * 9 + e) * kSlice);
#pragma unroll
                for (int u = 0; u < kUn; ++u) pair(j[u], a[u], 2);
            }
            for (; m <= wp; ++m) {  // the last full pairs, then the odd last block
                const int h = (m < wp) ? 2 : (w & 1);
                if (h == 0) break;
                const it2 j = __builtin_nontemporal_load(cp + (long long)m * kSlice);
                vt2 a[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) a[e] = __builtin_nontemporal_load(vp + ((long long)m * 9 + e) * kSlice);
                pair(j, a, h);
            }
        } else {
            typedef VT vt4 __attribute__((ext_vector_type(4)));
            const vt4* vp = (const vt4*)val + (sb / 4) * 9 * kSlice + lane;
            auto quad = [&](const it2 j0, const it2 j1, const vt4* a, int h) {  // blocks 4q..4q+h-1
                for (int u = 0; u < h; ++u) {
                    double c9[9];
#pragma unroll
                    for (int e = 0; e < 9; ++e) c9[e] = (double)a[e][u];
                    block(u < 2 ? j0[u] : j1[u - 2], c9);
                }
            };
            const int wq = w / 4;
            int q = 0;
            for (; q + kUn <= wq; q += kUn) {
                it2 j[kUn][2];
                vt4 a[kUn][9];
#pragma unroll
                for (int u = 0; u < kUn; ++u) {
                    j[u][0] = __builtin_nontemporal_load(cp + (long long)(2 * (q + u)) * kSlice);
                    j[u][1] = __builtin_nontemporal_load(cp + (long long)(2 * (q + u) + 1) * kSlice);
                }
#pragma unroll
                for (int u = 0; u < kUn; ++u)
#pragma unroll
                    for (int e = 0; e < 9; ++e)
                        a[u][e] = __builtin_nontemporal_load(vp + ((long long)(q + u) * 9 + e) * kSlice);
#pragma unroll
                for (int u = 0; u < kUn; ++u) quad(j[u][0], j[u][1], a[u], 4);
            }
            for (; q <= wq; ++q) {  // the last full quads, then the partial last one
                const int h = (q < wq) ? 4 : (w & 3);
                if (h == 0) break;
                const it2 j0 = __builtin_nontemporal_load(cp + (long long)(2 * q) * kSlice);
                const it2 j1 = __builtin_nontemporal_load(cp + (long long)(2 * q + 1) * kSlice);
                vt4 a[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) a[e] = __builtin_nontemporal_load(vp + ((long long)q * 9 + e) * kSlice);
                quad(j0, j1, a, h);
            }
        }
    }
    spmv_store(nf, i, s0, s1, s2, x, beta, yin, y, y2, p_copy, dotx, part, chunk, sh);
}

// ---------------------------------------------------------------------------
// the particle matrix in factored form (FacOp; the derivation is at the top of
// mph_factored.hip).  Not bit-identical to the assembled product: the same sums,
// associated differently, with multiply-adds contracted.  Both passes give each particle a group of G lanes
// (kFacLanes) that walk its neighbour list together: the list is in
// cell order, so the group's column and scalar loads are coalesced and its
// gathers of neighbours' records mostly hit consecutive records; each lane keeps a
// batch of entries in flight.  The group's partial sums are combined by a
// butterfly, in the same order for every launch.

// minimum-image displacement pos_j - pos_i (the reference's Mod form, used where
// the scalars are built); j's coordinates are loaded before the first min_image,
// whose branch would otherwise hold each load back until the previous one is done
__device__ inline void fac_disp(const DevParams& P, const double* __restrict__ pos, int j, double pi0, double pi1,
                                double pi2, double& r0, double& r1, double& r2) {
    const double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
    r0 = min_image(x, pi0, P.W[0], P.hW[0]);
    r1 = min_image(y, pi1, P.W[1], P.hW[1]);
    r2 = min_image(z, pi2, P.W[2], P.hW[2]);
}

// the same displacement for the products: positions lie inside the domain, so one
// wrap by the domain width gives the minimum image (equal to the Mod form up to
// rounding).  Only scenes with pairs across the periodic boundary (FacOp::wrap)
// take it; the others use the plain difference (kWrap false).
template <bool kWrap>
__device__ inline double fac_wrap(double r, double w, double hw) {
    if (!kWrap) return r;
    return r > hw ? r - w : (r < -hw ? r + w : r);
}

// sum over the G lanes of each aligned group (G a power of two, <= 64); every lane
// of the wavefront calls it
__device__ inline double group_sum(double v, int G) {
    for (int m = G >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// a row's sum s plus component r of its diagonal term, D_r . x_i + C_ii,r t_i, with
// the fused multiply-adds written out: every kernel of the factored product that
// forms it (whichever lane does) gets the same bits
__device__ inline double fac_diag_add(double s, double dr0, double dr1, double dr2, double x0, double x1, double x2,
                                      double c, double t) {
    double d = __builtin_fma(dr0, x0, dr1 * x1);
    d = __builtin_fma(dr2, x2, d);
    d = __builtin_fma(c, t, d);
    return s + d;
}

// one entry of FacOp::u32 / v32: three floats and the column
struct FacEntry32 {
    float a0, a1, a2;
    int col;
};
__device__ inline FacEntry32 fac_entry32(const float* __restrict__ base, long long e) {
    typedef float fv4 __attribute__((ext_vector_type(4)));
    const fv4 v = ((const fv4*)base)[e];
    return {v[0], v[1], v[2], __float_as_int(v[3])};
}

// lanes per particle in both passes, and the list entries each lane has in flight
// per batch (viscous entries FP64 and FP32; P entries)
constexpr int kFacLanes = 8, kFacBatch = 4, kFacBatch32 = 8, kFacBatchP = 4;
// the first pass (k's P list up to its last fluid entry, about 20 entries): lanes per
// particle and entries in flight per lane (198 us against 200 us with 8 x 4 on the 1M
// scene)
constexpr int kFacLanesT = 4, kFacBatchT = 8;

// the four planes of FacOp::rec
struct FacPlanes {
    double* a;  // [np][2] x0 x1
    double* b;  // [np][2] x2 p0
    double* c;  // [np][2] p1 p2
    double* t;  // [np]
};
__device__ inline FacPlanes fac_planes(const DevParams& P, double* rec) {
    const long long n = P.np;
    return {rec, rec + 2 * n, rec + 4 * n, rec + 6 * n};
}

// first pass over the 256 particles of `chunk` (fluid or wall):
//   t_k = lambda_k (C_kk . x_k [k fluid] - sum_{j fluid, j != k} beta_kj r_kj . x_j),
// written to k's record together with x_k (fluid k).  Every thread calls it.
// (Streaming a stored v = beta r instead of gathering positions measured slower:
// as many load instructions, more bytes.)
template <bool kWrap>
__device__ inline void fac_t_chunk(const DevParams& P, const FacOp& F, const double* __restrict__ x, int chunk) {
#pragma clang fp contract(fast)
    constexpr int G = kFacLanesT, U = kFacBatchT;
    const int g = threadIdx.x & (G - 1);
    const double W0 = P.W[0], W1 = P.W[1], W2 = P.W[2], h0 = P.hW[0], h1 = P.hW[1], h2 = P.hW[2];
    for (int r = threadIdx.x / G; r < 256; r += 256 / G) {
        const int k = chunk * 256 + r;
        double acc = 0.0;
        const int run = k < P.np ? F.runP[k] : 0;
        if (run > 0) {
            const double pk0 = F.pos[3 * k], pk1 = F.pos[3 * k + 1], pk2 = F.pos[3 * k + 2];
            const long long c0 = F.nPtrP[k];
            for (int jb = 0; jb < run; jb += G * U) {
                int jx[U];
                double bb[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    // entries past the end load the run's last entry (run >= 1), so the
                    // list loads need no branches and issue together
                    const int jn = jb + g + u * G;
                    const long long e = c0 + (jn < run ? jn : run - 1);
                    const int je = F.nIndP[e];
                    const double be = F.beta[e];
                    const int j = jn < run ? je : k;
                    const bool use = jn < run && j < P.nf && j != k;
                    jx[u] = use ? j : 0;  // unused entries read particle 0 with weight 0
                    bb[u] = use ? be : 0.0;
                }
                double v[U][6];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const long long j3 = 3 * (long long)jx[u];
                    v[u][0] = F.pos[j3];
                    v[u][1] = F.pos[j3 + 1];
                    v[u][2] = F.pos[j3 + 2];
                    v[u][3] = x[j3];
                    v[u][4] = x[j3 + 1];
                    v[u][5] = x[j3 + 2];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double r0 = fac_wrap<kWrap>(v[u][0] - pk0, W0, h0);
                    const double r1 = fac_wrap<kWrap>(v[u][1] - pk1, W1, h1);
                    const double r2 = fac_wrap<kWrap>(v[u][2] - pk2, W2, h2);
                    acc -= bb[u] * ((r0 * v[u][3] + r1 * v[u][4]) + r2 * v[u][5]);
                }
            }
        }
        acc = group_sum(acc, G);
        if (g == 0 && k < P.np) {
            const FacPlanes R = fac_planes(P, F.rec);
            if (k < P.nf) {
                const double x0 = x[3 * k], x1 = x[3 * k + 1], x2 = x[3 * k + 2];
                acc += (F.selfC[3 * k] * x0 + F.selfC[3 * k + 1] * x1) + F.selfC[3 * k + 2] * x2;
                R.a[2 * k] = x0;
                R.a[2 * k + 1] = x1;
                R.b[2 * k] = x2;
            }
            R.t[k] = run > 0 ? acc * F.lambda[k] : 0.0;
        }
    }
}

// second pass over the 256 fluid rows of `chunk`:
//   y_i = D_i x_i + C_ii t_i - sum_j alpha_ij r_ij (r_ij . x_j) + sum_{k != i} beta_ik r_ik t_k,
// the row sums staged in rows[3*256] (LDS), then the tail of spmv_store.  Every
// thread calls it.
template <bool kNeg, class VT, bool kWrap>
__device__ inline void fac_y_chunk(const DevParams& P, const FacOp& F, const double* __restrict__ x, double beta,
                                   const double* yin, double* y, double* y2, double* p_copy,
                                   const double* __restrict__ dotx, double* __restrict__ part, int chunk,
                                   double* sh, double* rows) {
#pragma clang fp contract(fast)
    typedef double dv2 __attribute__((ext_vector_type(2)));
    constexpr int G = kFacLanes;
    const int g = threadIdx.x & (G - 1);
    const FacPlanes R = fac_planes(P, F.rec);
    const dv2* __restrict__ ra = (const dv2*)R.a;
    const dv2* __restrict__ rb = (const dv2*)R.b;
    const dv2* __restrict__ rc = (const dv2*)R.c;
    const double* __restrict__ rt = R.t;
    const VT* __restrict__ alpha = (const VT*)F.alpha;
    const double W0 = P.W[0], W1 = P.W[1], W2 = P.W[2], h0 = P.hW[0], h1 = P.hW[1], h2 = P.hW[2];
    for (int r = threadIdx.x / G; r < 256; r += 256 / G) {
        const int i = chunk * 256 + r;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (i < P.nf) {
            const double pi0 = F.pos[3 * i], pi1 = F.pos[3 * i + 1], pi2 = F.pos[3 * i + 2];
            // viscous part over the fluid neighbours (alpha 0: self, beyond RadiusV);
            // unused entries of a batch read i's own record with weight 0
            const long long b0 = F.nPtr[i];
            const int w = F.nFluid[i];
            for (int jb = 0; jb < w; jb += G * kFacBatch) {
                constexpr int U = kFacBatch;
                int j[U];
                double a[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jn = jb + g + u * G;
                    j[u] = jn < w ? F.nInd[b0 + jn] : i;
                    a[u] = jn < w ? (double)alpha[b0 + jn] : 0.0;
                }
                dv2 q0[U], q1[U], q2[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    q0[u] = ra[j[u]];
                    q1[u] = rb[j[u]];
                    q2[u] = rc[j[u]];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double r0 = fac_wrap<kWrap>(q1[u][1] - pi0, W0, h0);
                    const double r1 = fac_wrap<kWrap>(q2[u][0] - pi1, W1, h1);
                    const double r2 = fac_wrap<kWrap>(q2[u][1] - pi2, W2, h2);
                    const double f = a[u] * ((r0 * q0[u][0] + r1 * q0[u][1]) + r2 * q1[u][0]);
                    s0 -= f * r0;
                    s1 -= f * r1;
                    s2 -= f * r2;
                }
            }
            // C Lambda C^T part over the P neighbours (beta 0: self)
            const long long c0 = F.nPtrP[i];
            const int n = F.nCountP[i];
            for (int jb = 0; jb < n; jb += G * kFacBatchP) {
                constexpr int U = kFacBatchP;
                int k[U];
                double bb[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jn = jb + g + u * G;
                    k[u] = jn < n ? F.nIndP[c0 + jn] : i;
                    bb[u] = jn < n ? F.beta[c0 + jn] : 0.0;
                }
                dv2 q1[U], q2[U];
                double q3[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    q1[u] = rb[k[u]];
                    q2[u] = rc[k[u]];
                    q3[u] = rt[k[u]];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double r0 = fac_wrap<kWrap>(q1[u][1] - pi0, W0, h0);
                    const double r1 = fac_wrap<kWrap>(q2[u][0] - pi1, W1, h1);
                    const double r2 = fac_wrap<kWrap>(q2[u][1] - pi2, W2, h2);
                    const double f = bb[u] * q3[u];
                    s0 += f * r0;
                    s1 += f * r1;
                    s2 += f * r2;
                }
            }
        }
        s0 = group_sum(s0, G);
        s1 = group_sum(s1, G);
        s2 = group_sum(s2, G);
        if (g == 0 && i < P.nf) {
            const double* D = F.diag + 9 * i;
            const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
            const double ti = rt[i];
            s0 = fac_diag_add(s0, D[0], D[1], D[2], xi0, xi1, xi2, F.selfC[3 * i], ti);
            s1 = fac_diag_add(s1, D[3], D[4], D[5], xi0, xi1, xi2, F.selfC[3 * i + 1], ti);
            s2 = fac_diag_add(s2, D[6], D[7], D[8], xi0, xi1, xi2, F.selfC[3 * i + 2], ti);
            rows[3 * r] = kNeg ? -s0 : s0;
            rows[3 * r + 1] = kNeg ? -s1 : s1;
            rows[3 * r + 2] = kNeg ? -s2 : s2;
        }
    }
    __syncthreads();
    const int i = chunk * 256 + threadIdx.x;
    const int t = threadIdx.x;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (i < P.nf) {
        r0 = rows[3 * t];
        r1 = rows[3 * t + 1];
        r2 = rows[3 * t + 2];
    }
    spmv_store(P.nf, i, r0, r1, r2, x, beta, yin, y, y2, p_copy, dotx, part, chunk, sh);
}

// FP32 form of the two passes (mixed-precision inner solves): the displacements are
// folded into the stored vectors (v_kj = beta_kj r_kj, u_ij = sqrt(alpha_ij) r_ij),
// so a pass gathers only a neighbour's 16-byte record (x, t as floats).  Arithmetic
// is FP64 on the FP32-stored values.
__device__ inline void fac_t32_chunk(const DevParams& P, const FacOp& F, const double* __restrict__ x, int chunk) {
#pragma clang fp contract(fast)
    constexpr int G = kFacLanes, U = kFacBatchP;
    typedef float fv4 __attribute__((ext_vector_type(4)));
    const int g = threadIdx.x & (G - 1);
    for (int r = threadIdx.x / G; r < 256; r += 256 / G) {
        const int k = chunk * 256 + r;
        double acc = 0.0;
        const int run = k < P.np ? F.runP[k] : 0;
        if (run > 0) {
            const long long c0 = F.nPtrP[k];
            for (int jb = 0; jb < run; jb += G * U) {
                int jx[U];
                float v[U][3];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jn = jb + g + u * G;
                    const FacEntry32 en = fac_entry32(F.v32, c0 + (jn < run ? jn : 0));
                    const int j = jn < run ? en.col : k;
                    const bool use = jn < run && j < P.nf && j != k;
                    jx[u] = use ? j : 0;  // unused entries read particle 0 with weight 0
                    v[u][0] = use ? en.a0 : 0.0f;
                    v[u][1] = use ? en.a1 : 0.0f;
                    v[u][2] = use ? en.a2 : 0.0f;
                }
                double xv[U][3];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const long long j3 = 3 * (long long)jx[u];
                    xv[u][0] = x[j3];
                    xv[u][1] = x[j3 + 1];
                    xv[u][2] = x[j3 + 2];
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    acc -= ((double)v[u][0] * xv[u][0] + (double)v[u][1] * xv[u][1]) + (double)v[u][2] * xv[u][2];
            }
        }
        acc = group_sum(acc, G);
        if (g == 0 && k < P.np) {
            fv4 rk = {0.0f, 0.0f, 0.0f, 0.0f};
            if (k < P.nf) {
                const double x0 = x[3 * k], x1 = x[3 * k + 1], x2 = x[3 * k + 2];
                acc += (F.selfC[3 * k] * x0 + F.selfC[3 * k + 1] * x1) + F.selfC[3 * k + 2] * x2;
                rk[0] = (float)x0;
                rk[1] = (float)x1;
                rk[2] = (float)x2;
            }
            rk[3] = run > 0 ? (float)(acc * F.lambda[k]) : 0.0f;
            ((fv4*)F.rec32)[k] = rk;
        }
    }
}

template <bool kNeg>
__device__ inline void fac_y32_chunk(const DevParams& P, const FacOp& F, const double* __restrict__ x, double beta,
                                     const double* yin, double* y, double* y2, double* p_copy,
                                     const double* __restrict__ dotx, double* __restrict__ part, int chunk,
                                     double* sh, double* rows) {
#pragma clang fp contract(fast)
    typedef float fv4 __attribute__((ext_vector_type(4)));
    constexpr int G = kFacLanes, U = kFacBatch32;
    const int g = threadIdx.x & (G - 1);
    const fv4* __restrict__ rec = (const fv4*)F.rec32;
    // the viscous loop is software-pipelined: the columns and u of the next batch
    // (or of the next row's first batch) load while the current batch's records are
    // gathered, so the two dependent round trips of a batch overlap
    struct Batch {
        int j[U];
        float a[U][3];
    };
    auto vload = [&](int i, int w, long long b0, int jb, Batch& b) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jn = jb + g + u * G;
            const bool in = jn < w;
            const FacEntry32 en = fac_entry32(F.u32, b0 + (in ? jn : 0));
            b.j[u] = in ? en.col : i;  // unused entries: i's own record, u = 0
            b.a[u][0] = in ? en.a0 : 0.0f;
            b.a[u][1] = in ? en.a1 : 0.0f;
            b.a[u][2] = in ? en.a2 : 0.0f;
        }
    };
    Batch nb;
    {
        const int i0 = chunk * 256 + threadIdx.x / G;
        if (i0 < P.nf) vload(i0, F.nFluid[i0], F.nPtr[i0], 0, nb);
    }
    for (int r = threadIdx.x / G; r < 256; r += 256 / G) {
        const int i = chunk * 256 + r;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (i < P.nf) {
            // viscous part: -u (u . x_j)
            const long long b0 = F.nPtr[i];
            const int w = F.nFluid[i];
            Batch cb = nb;
            fv4 q[U];
#pragma unroll
            for (int u = 0; u < U; ++u) q[u] = rec[cb.j[u]];
            for (int jb = 0; jb < w; jb += G * U) {
                const bool more = jb + G * U < w;
                const int rn = r + 256 / G, in_ = chunk * 256 + rn;
                if (more) vload(i, w, b0, jb + G * U, nb);
                else if (rn < 256 && in_ < P.nf) vload(in_, F.nFluid[in_], F.nPtr[in_], 0, nb);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double a0 = cb.a[u][0], a1 = cb.a[u][1], a2 = cb.a[u][2];
                    const double f = (a0 * (double)q[u][0] + a1 * (double)q[u][1]) + a2 * (double)q[u][2];
                    s0 -= f * a0;
                    s1 -= f * a1;
                    s2 -= f * a2;
                }
                if (more) {
                    cb = nb;
#pragma unroll
                    for (int u = 0; u < U; ++u) q[u] = rec[cb.j[u]];
                }
            }
            // C Lambda C^T part: v_ik t_k
            const long long c0 = F.nPtrP[i];
            const int n = F.nCountP[i];
            for (int jb = 0; jb < n; jb += G * kFacBatchP) {
                constexpr int U = kFacBatchP;
                int k[U];
                float vv[U][3];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int jn = jb + g + u * G;
                    const bool in = jn < n;
                    const FacEntry32 en = fac_entry32(F.v32, c0 + (in ? jn : 0));
                    k[u] = in ? en.col : i;
                    vv[u][0] = in ? en.a0 : 0.0f;
                    vv[u][1] = in ? en.a1 : 0.0f;
                    vv[u][2] = in ? en.a2 : 0.0f;
                }
                fv4 q[U];
#pragma unroll
                for (int u = 0; u < U; ++u) q[u] = rec[k[u]];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const double t = q[u][3];
                    s0 += (double)vv[u][0] * t;
                    s1 += (double)vv[u][1] * t;
                    s2 += (double)vv[u][2] * t;
                }
            }
        }
        s0 = group_sum(s0, G);
        s1 = group_sum(s1, G);
        s2 = group_sum(s2, G);
        if (g == 0 && i < P.nf) {
            const double* D = F.diag + 9 * i;
            const double xi0 = x[3 * i], xi1 = x[3 * i + 1], xi2 = x[3 * i + 2];
            const double ti = rec[i][3];
            double d0 = (D[0] * xi0 + D[1] * xi1) + D[2] * xi2;
            double d1 = (D[3] * xi0 + D[4] * xi1) + D[5] * xi2;
            double d2 = (D[6] * xi0 + D[7] * xi1) + D[8] * xi2;
            d0 += F.selfC[3 * i] * ti;
            d1 += F.selfC[3 * i + 1] * ti;
            d2 += F.selfC[3 * i + 2] * ti;
            s0 += d0;
            s1 += d1;
            s2 += d2;
            rows[3 * r] = kNeg ? -s0 : s0;
            rows[3 * r + 1] = kNeg ? -s1 : s1;
            rows[3 * r + 2] = kNeg ? -s2 : s2;
        }
    }
    __syncthreads();
    const int i = chunk * 256 + threadIdx.x;
    const int t = threadIdx.x;
    double r0 = 0.0, r1 = 0.0, r2 = 0.0;
    if (i < P.nf) {
        r0 = rows[3 * t];
        r1 = rows[3 * t + 1];
        r2 = rows[3 * t + 2];
    }
    spmv_store(P.nf, i, r0, r1, r2, x, beta, yin, y, y2, p_copy, dotx, part, chunk, sh);
}

// chunk partial of x.y: per particle (x0 y0 + x1 y1) + x2 y2, then the chunk tree
__device__ inline void dot_partial_chunk(int nf, const double* __restrict__ x, const double* __restrict__ y,
                                         double* __restrict__ part, int chunk, double* sh) {
    const int i = chunk * 256 + threadIdx.x;
    double t = 0.0;
    if (i < nf) t = (x[3 * i] * y[3 * i] + x[3 * i + 1] * y[3 * i + 1]) + x[3 * i + 2] * y[3 * i + 2];
    const double v = block_tree_sum(t, sh);
    if (threadIdx.x == 0) part[chunk] = v;
}

// fold the chunk partials: lane l sums chunks l, l+256, ... from 0.0, then the tree.
// scal[shift_slot] receives the previous scal[slot] first (rho -> rho_prev).
__device__ inline void dot_final_body(const double* __restrict__ part, int nchunks, double* scal, int slot,
                                      int shift_slot, double* sh, double* host = nullptr) {
    // lane l adds chunks l, l + 256, ... in that order; eight loads are issued
    // before their adds so the lane does not wait on each in turn
    double acc = 0.0;
    for (int c0 = threadIdx.x; c0 < nchunks; c0 += 256 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = c0 + 256 * u < nchunks ? part[c0 + 256 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (c0 + 256 * u < nchunks) acc = acc + v[u];
    }
    const double v = block_tree_sum(acc, sh);
    if (threadIdx.x == 0) {
        if (shift_slot >= 0) scal[shift_slot] = scal[slot];
        scal[slot] = v;
        if (host) *host = v;  // coherent pinned host memory: the CR stop test reads it
    }
}

// p = beta*p + 1.0*s, q = beta*q + 1.0*y with beta = rho/rho_prev (4360-4364)
// one element of p = beta p + s or q = beta q + y (myDscal then myDaxpy, 4362-4366),
// with the same roundings in every kernel that forms it
__device__ inline double pq_elem(double beta, double v, double w) {
    v *= beta;
    v += 1.0 * w;
    return v;
}

__device__ inline void pq_update_elem(const double* __restrict__ scal, const double* __restrict__ s,
                                      const double* __restrict__ y, double* __restrict__ p, double* __restrict__ q,
                                      long long k) {
    const double beta = scal[kRHO] / scal[kRHOP];
    p[k] = pq_elem(beta, p[k], s[k]);
    q[k] = pq_elem(beta, q[k], y[k]);
}

// alpha = rho/tmp; x += alpha p; r -= alpha q; s -= alpha u (4372-4375) + r.r partial.
// kPq: p = beta p + s first (with s before its update), written back; the CR loop's
// p update is then this kernel's, not a launch of its own
template <bool kPq = false>
__device__ inline void xrs_update_chunk(int nf, const double* __restrict__ scal, double* __restrict__ p,
                                        const double* __restrict__ q, const double* __restrict__ u,
                                        double* __restrict__ x, double* __restrict__ r, double* __restrict__ s,
                                        double* __restrict__ part, int chunk, double* sh) {
    const int i = chunk * 256 + threadIdx.x;
    double t = 0.0;
    if (i < nf) {
        const double alpha = scal[kRHO] / scal[kTMP];
        const double ma = -alpha;
        const double beta = kPq ? scal[kRHO] / scal[kRHOP] : 0.0;
        double rr[3];
        for (int d = 0; d < 3; ++d) {
            const int k = 3 * i + d;
            double pk = p[k];
            if (kPq) {
                pk = pq_elem(beta, pk, s[k]);
                p[k] = pk;
            }
            x[k] += alpha * pk;
            rr[d] = r[k] + ma * q[k];
            r[k] = rr[d];
            s[k] += ma * u[k];
        }
        t = (rr[0] * rr[0] + rr[1] * rr[1]) + rr[2] * rr[2];
    }
    const double v = block_tree_sum(t, sh);
    if (threadIdx.x == 0) part[chunk] = v;
}

}  // namespace body
}  // namespace mph
