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
// gathered through the caches from its 64-byte record (FacOp::rec): three 16-byte
// loads per entry, where separate arrays would take six or four 8-byte gathers.
// Each row is walked by a group of 8 lanes, each with 8 entries in flight, so that
// the gathers of one load instruction fall on consecutive particles of the
// cell-ordered list and enough of them are outstanding.  Measured on a jittered
// 100^3 lattice (172 entries per row), every arrangement of this pass (one lane per
// row, groups of 8 to 64 lanes, with or without an XCD-aware block order) lands
// between 0.94 and 1.3 ms: the cache traffic of the record gathers, not HBM,
// bounds it.
//
// The assembled matrix is still built every step. The multigrid's Galerkin product
// and the diagonal preconditioner read it.
#include "mph_cr_body.h"

namespace mph {
namespace {

using body::nblk;

// beta beside the P lists (and v = beta r in FP32 for the FP32 form), the extent of
// their fluid entries, C_ii of fluid particles, and the records (positions; x and t
// zero until the first pass)
__global__ void k_fac_p(DevParams P, const double* __restrict__ pos, const int* __restrict__ nCountP,
                        const long long* __restrict__ nPtrP, const int* __restrict__ nIndP, double* __restrict__ beta,
                        float* __restrict__ v32, int* __restrict__ runP, double* __restrict__ selfC,
                        double* __restrict__ rec) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P.np) return;
    const long long b0 = nPtrP[k];
    const int n = nCountP[k];
    const double pk0 = pos[3 * k], pk1 = pos[3 * k + 1], pk2 = pos[3 * k + 2];
    double c0 = 0.0, c1 = 0.0, c2 = 0.0;
    int run = 0;
    for (int jn = 0; jn < n; ++jn) {
        const int j = nIndP[b0 + jn];
        double b = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0;
        if (j != k) {
            body::fac_disp(P, pos, j, pk0, pk1, pk2, r0, r1, r2);
            const double r2s = (r0 * r0 + r1 * r1 + r2 * r2);
            if (r2s != 0.0 && P.radP2 - r2s > 0) {
                const double d = sqrt(r2s);
                b = -dwq_k(P.Kp, P.m2P, d, P.radP) / d;
                c0 += b * r0;
                c1 += b * r1;
                c2 += b * r2;
            }
        }
        if (j < P.nf) run = jn + 1;
        beta[b0 + jn] = b;
        if (v32) {
            v32[3 * (b0 + jn)] = (float)(b * r0);
            v32[3 * (b0 + jn) + 1] = (float)(b * r1);
            v32[3 * (b0 + jn) + 2] = (float)(b * r2);
        }
    }
    runP[k] = run;
    double* rk = rec + 8 * (long long)k;
    for (int e = 0; e < 8; ++e) rk[e] = e >= 3 && e < 6 ? pos[3 * k + e - 3] : 0.0;
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
        u32[3 * (b0 + jn)] = (float)(s * r0);
        u32[3 * (b0 + jn) + 1] = (float)(s * r1);
        u32[3 * (b0 + jn) + 2] = (float)(s * r2);
    }
}

// the wrap flag is uniform, so the branch is taken by the whole launch; the FP32
// form has no displacements to wrap
template <class VT>
__global__ void __launch_bounds__(256) k_fac_t(DevParams P, FacOp F, const double* __restrict__ x) {
    if constexpr (sizeof(VT) == 4) body::fac_t32_chunk(P, F, x, blockIdx.x);
    else if (*F.wrap) body::fac_t_chunk<true>(P, F, x, blockIdx.x);
    else body::fac_t_chunk<false>(P, F, x, blockIdx.x);
}

template <bool kNeg, class VT>
__global__ void __launch_bounds__(256) k_fac_y(DevParams P, FacOp F, const double* __restrict__ x, double beta,
                                               const double* yin, double* y, double* y2, double* p_copy,
                                               const double* __restrict__ dotx, double* __restrict__ part) {
    __shared__ double sh[256];
    __shared__ double rows[3 * 256];
    if constexpr (sizeof(VT) == 4)
        body::fac_y32_chunk<kNeg>(P, F, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh, rows);
    else if (*F.wrap)
        body::fac_y_chunk<kNeg, VT, true>(P, F, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh, rows);
    else body::fac_y_chunk<kNeg, VT, false>(P, F, x, beta, yin, y, y2, p_copy, dotx, part, blockIdx.x, sh, rows);
}

template <bool kNeg, class VT>
void fac_launch(hipStream_t st, const DevParams& P, const FacOp& F, const double* x, double beta, const double* yin,
                double* y, double* y2, double* p_copy, const double* dotx, double* part) {
    hipLaunchKernelGGL(k_fac_t<VT>, dim3(nblk(P.np, 256)), dim3(256), 0, st, P, F, x);
    hipLaunchKernelGGL((k_fac_y<kNeg, VT>), dim3(nblk(P.nf, kChunk)), dim3(kChunk), 0, st, P, F, x, beta, yin, y, y2,
                       p_copy, dotx, part);
}

}  // namespace

void launch_fac_p(hipStream_t st, const DevParams& P, const double* pos, const int* nCountP, const long long* nPtrP,
                  const int* nIndP, double* beta, float* v32, int* runP, double* selfC, double* rec) {
    if (P.np == 0) return;
    hipLaunchKernelGGL(k_fac_p, dim3(nblk(P.np, 128)), dim3(128), 0, st, P, pos, nCountP, nPtrP, nIndP, beta, v32,
                       runP, selfC, rec);
}

void launch_fac_u(hipStream_t st, const DevParams& P, const int* nFluid, const long long* nPtr, const int* nInd,
                  const double* alpha, const double* pos, float* u32) {
    if (P.nf == 0) return;
    hipLaunchKernelGGL(k_fac_u, dim3(nblk(8LL * P.nf, 256)), dim3(256), 0, st, P, nFluid, nPtr, nInd, alpha, pos,
                       u32);
}

void launch_fac_spmv(hipStream_t st, const DevParams& P, const FacOp& F, double alpha, const double* x, double beta,
                     const double* yin, double* y, double* y2, double* p_copy, const double* dotx, double* part,
                     bool f32) {
    // alpha is +1 or -1 at every call site (main.cpp:4337, 4353, 4357)
    if (P.nf == 0) return;
    if (f32) {
        if (alpha < 0) fac_launch<true, float>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part);
        else fac_launch<false, float>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part);
    } else {
        if (alpha < 0) fac_launch<true, double>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part);
        else fac_launch<false, double>(st, P, F, x, beta, yin, y, y2, p_copy, dotx, part);
    }
}

}  // namespace mph
