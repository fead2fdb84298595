// mph_cells.hip -- cell list, bitonic particle order and neighbour search on gfx950.
//
// Replaces calculateCellParticle (source/main.cpp:1164-1379) and calculateNeighbor
// (1382-1667).  The particle order is the reference's bitonic network, emulated
// exactly (it is not stable, so a radix sort would order ties differently): the
// passes with distance < 2048 run inside LDS, the wider ones as global passes.
// Neighbour lists come out ascending (fluid block, then wall block) because the
// 27 (9 in 2D) cells are visited in ascending CellId, as in the reference.
#include "mph_kernels.h"

namespace mph {

namespace {

constexpr int kSortBlock = 2048;  // keys per LDS block (16 KiB)
constexpr int kSortThreads = 256;

__device__ inline int cell_id(const DevParams& P, int cx, int cy, int cz) {  // CellId, main.cpp:109-113
    const int* C = P.C;
    if (P.dim == 2) return ((cx) % C[0] + C[0]) % C[0] * C[1] + ((cy) % C[1] + C[1]) % C[1];
    return ((cx) % C[0] + C[0]) % C[0] * C[1] * C[2] + ((cy) % C[1] + C[1]) % C[1] * C[2] +
           ((cz) % C[2] + C[2]) % C[2];
}

// key of main.cpp:1168-1186, packed with the particle index for the sort
__global__ void k_hash(DevParams P, const double* __restrict__ pos, const int* __restrict__ type,
                       uint64_t* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.p2) return;
    if (i < P.np) {
        const int cx = ((int)floor((pos[3 * i + 0] - P.dmin[0]) / P.cw[0])) % P.C[0];
        const int cy = ((int)floor((pos[3 * i + 1] - P.dmin[1]) / P.cw[1])) % P.C[1];
        const int cz = ((int)floor((pos[3 * i + 2] - P.dmin[2]) / P.cw[2])) % P.C[2];
        int key = cell_id(P, cx, cy, cz);
        if (is_wall(type[i])) key += P.T;
        keys[i] = ((uint64_t)(uint32_t)key << 32) | (uint32_t)i;
    } else {
        keys[i] = ((uint64_t)(uint32_t)(2 * P.T) << 32) | (uint32_t)P.np;
    }
}

// one compare-exchange of the network of main.cpp:1190-1211 for global index i
__device__ inline void cmp_swap(uint64_t& a, uint64_t& b, bool up) {
    const uint32_t ka = (uint32_t)(a >> 32), kb = (uint32_t)(b >> 32);
    if ((ka > kb) == up) {
        const uint64_t t = a;
        a = b;
        b = t;
    }
}

__device__ inline int pair_lo(int t, int logd) {  // t-th pair (lo index) at distance 2^logd
    return ((t >> logd) << (logd + 1)) | (t & ((1 << logd) - 1));
}

// all stages iMain < log2(n) inside one LDS block of n keys
__global__ void __launch_bounds__(kSortThreads) k_bitonic_block_sort(uint64_t* keys, int n, int logn) {
    __shared__ uint64_t sh[kSortBlock];
    const int base = blockIdx.x * n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) sh[k] = keys[base + k];
    __syncthreads();
    for (int iMain = 0; iMain < logn; ++iMain) {
        for (int iSub = 0; iSub <= iMain; ++iSub) {
            const int logd = iMain - iSub;
            for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
                const int lo = pair_lo(t, logd);
                const int gi = base + lo;
                const bool up = ((gi >> iMain) & 2) == 0;
                cmp_swap(sh[lo], sh[lo | (1 << logd)], up);
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) keys[base + k] = sh[k];
}

// the passes of stage iMain with distance < n, inside one LDS block
__global__ void __launch_bounds__(kSortThreads) k_bitonic_block_merge(uint64_t* keys, int n, int logn, int iMain) {
    __shared__ uint64_t sh[kSortBlock];
    const int base = blockIdx.x * n;
    for (int k = threadIdx.x; k < n; k += blockDim.x) sh[k] = keys[base + k];
    __syncthreads();
    for (int logd = logn - 1; logd >= 0; --logd) {
        for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
            const int lo = pair_lo(t, logd);
            const int gi = base + lo;
            const bool up = ((gi >> iMain) & 2) == 0;
            cmp_swap(sh[lo], sh[lo | (1 << logd)], up);
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) keys[base + k] = sh[k];
}

// one pass at distance 2^logd >= block size
__global__ void k_bitonic_global(uint64_t* keys, int npairs, int logd, int iMain) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const int lo = pair_lo(t, logd);
    const int hi = lo | (1 << logd);
    const bool up = ((lo >> iMain) & 2) == 0;
    uint64_t a = keys[lo], b = keys[hi];
    const uint32_t ka = (uint32_t)(a >> 32), kb = (uint32_t)(b >> 32);
    if ((ka > kb) == up) {
        keys[lo] = b;
        keys[hi] = a;
    }
}

// cellStart[c] = number of sorted keys < c, for c in [0, 2T]; this equals the
// cell Begin/End arrays built by boundary detection and scan at main.cpp:1213-1300
// (integer counts, so any evaluation order gives the same values).
__global__ void k_cell_start(DevParams P, const uint64_t* __restrict__ keys, int* __restrict__ cellStart) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > 2 * P.T) return;
    int lo = 0, hi = P.np;  // first index in [0, np] whose key >= c (keys[np] = 2T)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((int)(keys[mid] >> 32) < c) lo = mid + 1;
        else hi = mid;
    }
    cellStart[c] = lo;
}

// gather-reorder of ParticleIndex/Property/Position/Velocity, main.cpp:1315-1378
__global__ void k_reorder(DevParams P, const uint64_t* __restrict__ keys, const int* __restrict__ type,
                          const int* __restrict__ perm, const double* __restrict__ pos,
                          const double* __restrict__ vel, int* __restrict__ type2, int* __restrict__ perm2,
                          double* __restrict__ pos2, double* __restrict__ vel2, int* __restrict__ cellIndex) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    const int src = (int)(uint32_t)keys[i];
    cellIndex[i] = (int)(keys[i] >> 32);
    type2[i] = type[src];
    perm2[i] = perm[src];
    for (int d = 0; d < 3; ++d) {
        pos2[3 * i + d] = pos[3 * src + d];
        vel2[3 * i + d] = vel[3 * src + d];
    }
}

// The neighbour cells of particle i in the reference's visiting order
// (main.cpp:1410-1445): ascending CellId.  `visit(lo, hi)` receives particle ranges.
template <bool kFluid, class F>
__device__ inline void for_cells(const DevParams& P, const int* __restrict__ cellStart, int key, F&& visit) {
    const int* C = P.C;
    const int rx = P.range[0], ry = P.range[1], rz = P.range[2];
    const int icx = (key / (C[1] * C[2])) % P.T;
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
                const int jc = cell_id(P, jx[a % (2 * rx + 1)], jy[b % (2 * ry + 1)], jz[c % (2 * rz + 1)]);
                const int off = kFluid ? 0 : P.T;
                visit(cellStart[off + jc], cellStart[off + jc + 1]);
            }
}

// visit(j, q2) with q2 the squared minimum-image distance of candidate j to
// (xi, yi, zi), for j = lo..hi-1 in ascending order.  The coordinates of eight
// candidates are loaded together before the first min_image: its branch (the division
// of the wrapped case) would otherwise hold each load back until the previous
// coordinate is done.  Past hi the loads repeat candidate hi-1, which is not visited.
// (Per step on the 1M scene: 3.56 ms with two candidates, 3.35 with four, 3.21 with
// eight, at four waves per SIMD in each case.)
template <class V>
__device__ inline void scan_range(const DevParams& P, const double* __restrict__ pos, double xi, double yi, double zi,
                                  int lo, int hi, V&& visit) {
    constexpr int U = 8;
    for (int j0 = lo; j0 < hi; j0 += U) {
        double c[U][3];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + u < hi ? j0 + u : hi - 1;
            c[u][0] = pos[3 * j];
            c[u][1] = pos[3 * j + 1];
            c[u][2] = pos[3 * j + 2];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (j0 + u >= hi) break;
            const double q0 = min_image(c[u][0], xi, P.W[0], P.hW[0]);
            const double q1 = min_image(c[u][1], yi, P.W[1], P.hW[1]);
            const double q2 = min_image(c[u][2], zi, P.W[2], P.hW[2]);
            visit(j0 + u, q0 * q0 + q1 * q1 + q2 * q2);
        }
    }
}

// count pass, main.cpp:1406-1511; also the largest P-neighbour count of a fluid
// particle (it sizes the LDS staging of the C Lambda C^T kernel), the largest counts
// of any particle (they size the scratch lists of the next search), and, when scrM is
// set, the lists themselves: the hits go in visiting order into per-particle scratch
// of capM / capP entries (runStart as the fill pass writes it), and a particle with
// more raises *overflow, which sends the search to the fill pass.
__global__ void k_neighbor_count(DevParams P, const double* __restrict__ pos, const int* __restrict__ cellIndex,
                                 const int* __restrict__ cellStart, int* __restrict__ nCount,
                                 int* __restrict__ nFluid, int* __restrict__ nCountP, int* __restrict__ maxFluidP,
                                 int* __restrict__ scrM, int capM, int* __restrict__ scrP, int capP,
                                 unsigned short* __restrict__ runStart, int* __restrict__ overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int key = cellIndex[i];
    int n = 0, nfl = 0, npp = 0;
    int* sm = scrM ? scrM + (long long)i * capM : nullptr;
    int* sp = scrM ? scrP + (long long)i * capP : nullptr;
    unsigned short* rs = scrM && i < P.nf ? runStart + (long long)i * kRuns : nullptr;
    int q = 0;
    for_cells<true>(P, cellStart, key, [&](int lo, int hi) {
        if (rs) rs[q] = (unsigned short)n;
        q++;
        scan_range(P, pos, xi, yi, zi, lo, hi, [&](int j, double q2) {
            if (q2 <= P.maxR2) {
                if (sm && n < capM) sm[n] = j;
                n++;
                nfl++;
                if (q2 <= P.radP2) {
                    if (sp && npp < capP) sp[npp] = j;
                    npp++;
                }
            }
        });
    });
    if (rs) rs[q] = (unsigned short)n;
    // walls without fluid neighbours skip the wall scan (main.cpp:1474); walls
    // occupy [nf, np) after every sort
    if (!(i >= P.nf && n == 0))
        for_cells<false>(P, cellStart, key, [&](int lo, int hi) {
            scan_range(P, pos, xi, yi, zi, lo, hi, [&](int j, double q2) {
                if (q2 <= P.maxR2) {
                    if (sm && n < capM) sm[n] = j;
                    n++;
                    if (q2 <= P.radP2) {
                        if (sp && npp < capP) sp[npp] = j;
                        npp++;
                    }
                }
            });
        });
    nCount[i] = n;
    nFluid[i] = nfl;
    nCountP[i] = npp;
    if (i < P.nf) {
        atomicMax(maxFluidP, npp);      // [0]: largest P-neighbour count of a fluid particle
        atomicMax(maxFluidP + 1, nfl);  // [1]: largest fluid-neighbour count (A row length)
    }
    atomicMax(maxFluidP + 2, n);    // [2]: largest neighbour count
    atomicMax(maxFluidP + 3, npp);  // [3]: largest P-neighbour count
    if (sm && (n > capM || npp > capP)) atomicOr(overflow, 1);
}

// the lists from the scratch of the count pass, 32 lanes per particle (coalesced)
__global__ void __launch_bounds__(256) k_neighbor_compact(DevParams P, const int* __restrict__ nCount,
                                                          const int* __restrict__ nCountP,
                                                          const long long* __restrict__ nPtr,
                                                          const long long* __restrict__ nPtrP,
                                                          const int* __restrict__ scrM, int capM,
                                                          const int* __restrict__ scrP, int capP,
                                                          int* __restrict__ nInd, int* __restrict__ nIndP) {
    constexpr int G = 32;
    const int g = threadIdx.x & (G - 1);
    const int i = (int)(((long long)blockIdx.x * blockDim.x + threadIdx.x) / G);
    if (i >= P.np) return;
    const int n = nCount[i], npp = nCountP[i];
    const long long b = nPtr[i], bp = nPtrP[i];
    const int* sm = scrM + (long long)i * capM;
    const int* sp = scrP + (long long)i * capP;
    for (int k = g; k < n; k += G) nInd[b + k] = sm[k];
    for (int k = g; k < npp; k += G) nIndP[bp + k] = sp[k];
}

// fill pass, main.cpp:1567-1666.  For fluid particles it also records where the
// fluid neighbours of each visited cell start in the list (runStart[i][q], q in
// visiting order, runStart[i][ncells] = fluid count): the multigrid level-0 build
// reads these runs directly.
__global__ void k_neighbor_fill(DevParams P, const double* __restrict__ pos, const int* __restrict__ cellIndex,
                                const int* __restrict__ cellStart, const int* __restrict__ nCount,
                                const long long* __restrict__ nPtr, const long long* __restrict__ nPtrP,
                                int* __restrict__ nInd, int* __restrict__ nIndP,
                                unsigned short* __restrict__ runStart) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.np) return;
    const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    const int key = cellIndex[i];
    long long k = nPtr[i], kp = nPtrP[i];
    auto put = [&](int lo, int hi) {
        scan_range(P, pos, xi, yi, zi, lo, hi, [&](int j, double q2) {
            if (q2 <= P.maxR2) {
                nInd[k++] = j;
                if (q2 <= P.radP2) nIndP[kp++] = j;
            }
        });
    };
    unsigned short* rs = i < P.nf ? runStart + (long long)i * kRuns : nullptr;
    int q = 0;
    const long long k0 = k;
    for_cells<true>(P, cellStart, key, [&](int lo, int hi) {
        if (rs) rs[q] = (unsigned short)(k - k0);
        q++;
        put(lo, hi);
    });
    if (rs) rs[q] = (unsigned short)(k - k0);
    if (!(i >= P.nf && nCount[i] == 0)) for_cells<false>(P, cellStart, key, put);
}

// ---- exclusive scan of int counts into int64 offsets (out[0]=0, out[n]=total) ----
constexpr int kScanThreads = 256, kScanItems = 8, kScanTile = kScanThreads * kScanItems;

__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(const int* __restrict__ in, int n,
                                                                 long long* __restrict__ sums) {
    __shared__ long long red[kScanThreads];
    const int base = blockIdx.x * kScanTile;
    long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const int idx = base + threadIdx.x * kScanItems + k;
        if (idx < n) s += in[idx];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kScanThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__device__ inline long long block_exclusive_scan(long long v, long long* sh, long long& total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const long long t = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    total = sh[blockDim.x - 1];
    const long long incl = sh[threadIdx.x];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_sums(long long* sums, int nt) {
    __shared__ long long sh[kScanThreads];
    long long carry = 0;
    for (int base = 0; base < nt; base += kScanThreads) {
        const int idx = base + threadIdx.x;
        const long long v = idx < nt ? sums[idx] : 0;
        long long total;
        const long long ex = block_exclusive_scan(v, sh, total);
        if (idx < nt) sums[idx] = carry + ex;
        carry += total;
    }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const int* __restrict__ in, int n,
                                                             const long long* __restrict__ sums,
                                                             long long* __restrict__ out) {
    __shared__ long long sh[kScanThreads];
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    long long total;
    long long run = block_exclusive_scan(s, sh, total) + sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        run += v[k];
        if (base + k < n) out[base + k + 1] = run;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

__global__ void k_slice_width(int nf, int nslices, const int* __restrict__ nFluid, int* __restrict__ width) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslices) return;
    int w = 0;
    for (int l = 0; l < kSlice; ++l) {
        const int i = s * kSlice + l;
        if (i < nf && nFluid[i] > w) w = nFluid[i];
    }
    width[s] = (w + 3) & ~3;  // block pairs / quads: every slice starts on a multiple of 4
}

inline int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

void launch_hash(hipStream_t st, const DevParams& P, const double* pos, const int* type, uint64_t* keys) {
    hipLaunchKernelGGL(k_hash, dim3(nblk(P.p2, 256)), dim3(256), 0, st, P, pos, type, keys);
}

void launch_bitonic(hipStream_t st, const DevParams& P, uint64_t* keys) {
    const int n = P.p2 < kSortBlock ? P.p2 : kSortBlock;
    int logn = 0;
    while ((1 << logn) < n) ++logn;
    const int nb = P.p2 / n;
    hipLaunchKernelGGL(k_bitonic_block_sort, dim3(nb), dim3(kSortThreads), 0, st, keys, n, logn);
    for (int iMain = logn; iMain < P.p2pow; ++iMain) {
        for (int logd = iMain; logd >= logn; --logd)
            hipLaunchKernelGGL(k_bitonic_global, dim3(nblk(P.p2 / 2, 256)), dim3(256), 0, st, keys, P.p2 / 2, logd,
                               iMain);
        hipLaunchKernelGGL(k_bitonic_block_merge, dim3(nb), dim3(kSortThreads), 0, st, keys, n, logn, iMain);
    }
}

void launch_cell_start(hipStream_t st, const DevParams& P, const uint64_t* keys, int* cellStart) {
    hipLaunchKernelGGL(k_cell_start, dim3(nblk(2 * (long long)P.T + 1, 256)), dim3(256), 0, st, P, keys, cellStart);
}

void launch_reorder(hipStream_t st, const DevParams& P, const uint64_t* keys, const int* type, const int* perm,
                    const double* pos, const double* vel, int* type2, int* perm2, double* pos2, double* vel2,
                    int* cellIndex) {
    hipLaunchKernelGGL(k_reorder, dim3(nblk(P.np, 256)), dim3(256), 0, st, P, keys, type, perm, pos, vel, type2,
                       perm2, pos2, vel2, cellIndex);
}

void launch_neighbor_count(hipStream_t st, const DevParams& P, const double* pos, const int* cellIndex,
                           const int* cellStart, int* nCount, int* nFluid, int* nCountP, int* maxFluidP, int* scrM,
                           int capM, int* scrP, int capP, unsigned short* runStart, int* overflow) {
    hipLaunchKernelGGL(k_neighbor_count, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, pos, cellIndex, cellStart,
                       nCount, nFluid, nCountP, maxFluidP, scrM, capM, scrP, capP, runStart, overflow);
}

void launch_neighbor_compact(hipStream_t st, const DevParams& P, const int* nCount, const int* nCountP,
                             const long long* nPtr, const long long* nPtrP, const int* scrM, int capM,
                             const int* scrP, int capP, int* nInd, int* nIndP) {
    hipLaunchKernelGGL(k_neighbor_compact, dim3(nblk(32LL * P.np, 256)), dim3(256), 0, st, P, nCount, nCountP, nPtr,
                       nPtrP, scrM, capM, scrP, capP, nInd, nIndP);
}

void launch_neighbor_fill(hipStream_t st, const DevParams& P, const double* pos, const int* cellIndex,
                          const int* cellStart, const int* nCount, const long long* nPtr, const long long* nPtrP,
                          int* nInd, int* nIndP, unsigned short* runStart) {
    hipLaunchKernelGGL(k_neighbor_fill, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, pos, cellIndex, cellStart,
                       nCount, nPtr, nPtrP, nInd, nIndP, runStart);
}

int scan_scratch_len(int n) { return nblk(n > 0 ? n : 1, kScanTile); }

void launch_scan_i64(hipStream_t st, const int* in, long long* out, int n, long long* scratch, int scratch_len) {
    const int nt = nblk(n > 0 ? n : 1, kScanTile);
    (void)scratch_len;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(nt), dim3(kScanThreads), 0, st, in, n, scratch);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, st, scratch, nt);
    hipLaunchKernelGGL(k_scan_apply, dim3(nt), dim3(kScanThreads), 0, st, in, n, scratch, out);
}

void launch_slice_width(hipStream_t st, int nf, int nslices, const int* nFluid, int* width) {
    hipLaunchKernelGGL(k_slice_width, dim3(nblk(nslices, 256)), dim3(256), 0, st, nf, nslices, nFluid, width);
}

}  // namespace mph
