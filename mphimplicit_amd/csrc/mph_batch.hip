// mph_batch.hip -- scene-table kernels of the batched executor: the CR loop and the
// V-cycle of many scenes in one launch each (see mph_batch.h).  The per-row and
// per-chunk arithmetic is the single-scene code of mph_cr_body.h, so every scene
// computes exactly what it computes alone.
#include "mph_batch.h"
#include "mph_cr_body.h"

namespace mph {
namespace batch {

namespace {

using namespace body;

#define MPH_SCENE(row)                                                     \
    const int* o_ = off + (long long)(row) * (ns + 1);                     \
    const int s = find_scene(o_, ns, blockIdx.x);                          \
    if (!in_set(on, s)) return;                                            \
    const int lb = blockIdx.x - o_[s];                                     \
    const BScene& B = tab[s];                                              \
    if (halted(B)) return;

__global__ void kb_fill(const BScene* tab, const int* off, int ns, SceneSet on, int role, double v) {
    MPH_SCENE(kOffElem)
    const long long k = (long long)lb * 256 + threadIdx.x;
    if (k < 3LL * B.P.nf) B.vec[role][k] = v;
}

__global__ void kb_fill_level(const BScene* tab, const int* off, int ns, SceneSet on, int l, double v) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level& L = B.lv[l];
    if (row < L.nrow) {
        L.S[0][row] = v;
        L.S[1][row] = v;
        L.R[row] = v;
        L.Q[row] = v;
    }
}

__global__ void kb_cr_init(const BScene* tab, const int* off, int ns, SceneSet on) {
    MPH_SCENE(kOffChunk)
    const int i = lb * 256 + threadIdx.x;
    if (i < B.P.nf) cr_init_particle(B.P, B.vel, B.force, B.mass, B.vec[kX], i);
}

__global__ void kb_copy(const BScene* tab, const int* off, int ns, SceneSet on, int src, int dst) {
    MPH_SCENE(kOffElem)
    const long long k = (long long)lb * 256 + threadIdx.x;
    if (k < 3LL * B.P.nf) B.vec[dst][k] = B.vec[src][k];
}

__global__ void kb_axpy(const BScene* tab, const int* off, int ns, SceneSet on, int src, int dst) {
    MPH_SCENE(kOffElem)
    const long long k = (long long)lb * 256 + threadIdx.x;
    if (k < 3LL * B.P.nf) B.vec[dst][k] += 1.0 * B.vec[src][k];
}

template <bool kNeg, class VT, int kUn>
__global__ void __launch_bounds__(256) kb_spmv(const BScene* tab, const int* off, int ns, SceneSet on,
                                               SceneSet first, int x, double beta, int yin, int y, int y2,
                                               int p_copy, int dot) {
    __shared__ double sh[256];
    MPH_SCENE(kOffChunk)
    const bool f = in_set(first, s);
    const VT* val = sizeof(VT) == 8 ? (const VT*)B.val : (const VT*)B.val32;
    spmv_chunk<kNeg, VT, kUn>(B.P.nf, B.nFluid, B.start, B.col, val, B.vec[x], beta, B.vec[yin], B.vec[y],
                         f && y2 >= 0 ? B.vec[y2] : nullptr, f && p_copy >= 0 ? B.vec[p_copy] : nullptr,
                         dot >= 0 ? B.vec[dot] : nullptr, dot >= 0 ? B.part : nullptr, lb, sh);
}

template <class VT>
__global__ void __launch_bounds__(256) kb_fac_t(const BScene* tab, const int* off, int ns, SceneSet on, int x) {
    MPH_SCENE(kOffPart)
    const FacOp& F = B.fac[sizeof(VT) == 4 ? 1 : 0];
    if constexpr (sizeof(VT) == 4) fac_t32_chunk(B.P, F, B.vec[x], lb);
    else if (*F.wrap) fac_t_chunk<true>(B.P, F, B.vec[x], lb);
    else fac_t_chunk<false>(B.P, F, B.vec[x], lb);
}

template <bool kNeg, class VT>
__global__ void __launch_bounds__(256) kb_fac_y(const BScene* tab, const int* off, int ns, SceneSet on,
                                                SceneSet first, int x, double beta, int yin, int y, int y2,
                                                int p_copy, int dot) {
    __shared__ double sh[256];
    __shared__ double rows[3 * 256];
    MPH_SCENE(kOffChunk)
    const bool f = in_set(first, s);
    const FacOp& F = B.fac[sizeof(VT) == 4 ? 1 : 0];
    double* y2p = f && y2 >= 0 ? B.vec[y2] : nullptr;
    double* pc = f && p_copy >= 0 ? B.vec[p_copy] : nullptr;
    const double* dx = dot >= 0 ? B.vec[dot] : nullptr;
    double* pt = dot >= 0 ? B.part : nullptr;
    if constexpr (sizeof(VT) == 4)
        fac_y32_chunk<kNeg>(B.P, F, B.vec[x], beta, B.vec[yin], B.vec[y], y2p, pc, dx, pt, lb, sh, rows);
    else if (*F.wrap)
        fac_y_chunk<kNeg, VT, true>(B.P, F, B.vec[x], beta, B.vec[yin], B.vec[y], y2p, pc, dx, pt, lb, sh, rows);
    else fac_y_chunk<kNeg, VT, false>(B.P, F, B.vec[x], beta, B.vec[yin], B.vec[y], y2p, pc, dx, pt, lb, sh, rows);
}

__global__ void __launch_bounds__(256) kb_dot_partial(const BScene* tab, const int* off, int ns, SceneSet on,
                                                      int a, int b) {
    __shared__ double sh[256];
    MPH_SCENE(kOffChunk)
    dot_partial_chunk(B.P.nf, B.vec[a], B.vec[b], B.part, lb, sh);
}

// one workgroup per scene
__global__ void __launch_bounds__(256) kb_dot_final(const BScene* tab, int ns, SceneSet on, int slot,
                                                    int shift_slot) {
    __shared__ double sh[256];
    const int s = blockIdx.x;
    if (s >= ns || !in_set(on, s)) return;
    const BScene& B = tab[s];
    if (halted(B)) return;
    dot_final_body(B.part, nblk(B.P.nf, 256), B.scal, slot, shift_slot, sh);
}

// r.r and the CR stop test of each scene (k_dot_final_stop of mph_solver.hip)
__global__ void __launch_bounds__(256) kb_dot_final_stop(const BScene* tab, int ns, SceneSet on, int step) {
    __shared__ double sh[256];
    const int s = blockIdx.x;
    if (s >= ns || !in_set(on, s)) return;
    const BScene& B = tab[s];
    if (halted(B)) return;
    dot_final_body(B.part, nblk(B.P.nf, 256), B.scal, kRR, -1, sh);
    if (threadIdx.x != 0) return;
    const double rr = B.scal[kRR];
    const int it = B.ctl[kCtlIter] + step;
    int halt = 0, bad = 0;
    if (it >= B.maxit) halt = 1;
    else if (!isfinite(rr)) halt = bad = 1;
    else if (rr / B.scal[kRR0] < B.tol) halt = 1;
    B.ctl[kCtlIter] = it;
    B.ctl[kCtlBad] = bad;
    B.ctl[kCtlHalt] = halt;
    B.hctl[kCtlIter] = it;
    B.hctl[kCtlBad] = bad;
    B.hctl[kCtlHalt] = halt;
}

// the control block of each scene: cleared before the solve, its stop flag lowered
// after the loop (release)
__global__ void kb_ctl(const BScene* tab, int ns, SceneSet on, int release) {
    const int s = blockIdx.x;
    if (s >= ns || !in_set(on, s) || threadIdx.x >= kNCtl) return;
    const BScene& B = tab[s];
    if (!B.ctl) return;
    if (!release || threadIdx.x == kCtlHalt) B.ctl[threadIdx.x] = 0;
    if (!release) B.hctl[threadIdx.x] = 0;
}

__global__ void kb_pq_update(const BScene* tab, const int* off, int ns, SceneSet on) {
    MPH_SCENE(kOffElem)
    const long long k = (long long)lb * 256 + threadIdx.x;
    if (k < 3LL * B.P.nf) pq_update_elem(B.scal, B.vec[kS], B.vec[kY], B.vec[kP], B.vec[kQ], k);
}

__global__ void __launch_bounds__(256) kb_xrs_update(const BScene* tab, const int* off, int ns, SceneSet on,
                                                     int x) {
    __shared__ double sh[256];
    MPH_SCENE(kOffChunk)
    xrs_update_chunk(B.P.nf, B.scal, B.vec[kP], B.vec[kQ], B.vec[kU], B.vec[x], B.vec[kR], B.vec[kS], B.part, lb,
                     sh);
}

__global__ void kb_store_velocity(const BScene* tab, const int* off, int ns, SceneSet on) {
    MPH_SCENE(kOffElem)
    const long long k = (long long)lb * 256 + threadIdx.x;
    if (k < 3LL * B.P.nf) B.vel[k] = B.vec[kX][k];
}

// ---- V-cycle ----

__device__ inline Level level(const BScene& B, int l, const SceneSet& cur, int s) {
    Level L = B.lv[l];
    L.cur = in_set(cur, s) ? 1 : 0;
    return L;
}

__global__ void kb_mg_p2g(const BScene* tab, const int* off, int ns, SceneSet on, int in) {
    MPH_SCENE(kOffLevel)
    const int row = lb * 256 + threadIdx.x;
    const Level& L = B.lv[0];
    if (row < L.nrow) mg_p2g_row(B.P, L, B.cellStart, B.vec[in], row);
}

__global__ void kb_mg_p2g_bu(const BScene* tab, const int* off, int ns, SceneSet on, int in, SceneSet cur) {
    MPH_SCENE(kOffLevel)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, 0, cur, s);
    if (row >= L.nrow) return;
    mg_p2g_row(B.P, L, B.cellStart, B.vec[in], row);
    mg_begin_update_row(L, row);
}

__global__ void kb_mg_restrict_bu(const BScene* tab, const int* off, int ns, SceneSet on, int l, SceneSet cur) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, l, cur, s);
    if (row >= L.nrow) return;
    mg_restrict_row(B.lv[l - 1], L, row);
    mg_begin_update_row(L, row);
}

__global__ void kb_mg_begin(const BScene* tab, const int* off, int ns, SceneSet on, int l, SceneSet cur) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, l, cur, s);
    if (row < L.nrow) mg_begin_row(L, row);
}

__global__ void kb_mg_update(const BScene* tab, const int* off, int ns, SceneSet on, int l, SceneSet cur) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, l, cur, s);
    if (row < L.nrow) mg_update_row(L, row);
}

// FP32 stencil where the level has one and the launch asks for it (mixed inner solve)
__global__ void kb_mg_residual(const BScene* tab, const int* off, int ns, SceneSet on, int l, SceneSet cur,
                               int f32) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, l, cur, s);
    if (row >= L.nrow) return;
    if (f32 && L.A32) mg_residual_row<float>(L, row);
    else mg_residual_row<double>(L, row);
}

__global__ void kb_mg_residual_update(const BScene* tab, const int* off, int ns, SceneSet on, int l, SceneSet cur,
                                      int f32) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, l, cur, s);
    if (row >= L.nrow) return;
    if (f32 && L.A32) mg_residual_update_row<float>(L, row);
    else mg_residual_update_row<double>(L, row);
}

__global__ void kb_mg_restrict(const BScene* tab, const int* off, int ns, SceneSet on, int l) {
    MPH_SCENE(kOffLevel + l)
    const int row = lb * 256 + threadIdx.x;
    const Level& L = B.lv[l];
    if (row < L.nrow) mg_restrict_row(B.lv[l - 1], L, row);
}

__global__ void kb_mg_prolong(const BScene* tab, const int* off, int ns, SceneSet on, int l, SceneSet curL,
                              SceneSet curS) {
    MPH_SCENE(kOffLevel + l - 1)
    const int row = lb * 256 + threadIdx.x;
    const Level L = level(B, l, curL, s);
    const Level S = level(B, l - 1, curS, s);
    if (row < S.nrow) mg_prolong_row(L, S, row);
}

__global__ void __launch_bounds__(256) kb_mg_g2p(const BScene* tab, const int* off, int ns, SceneSet on,
                                                 SceneSet cur0, int in, int out, int part) {
    __shared__ double sh[256];
    MPH_SCENE(kOffChunk)
    const Level L = level(B, 0, cur0, s);
    mg_g2p_chunk(B.P, L, B.cellIndex, B.invDiagA, B.vec[in], B.vec[out], part ? B.part : nullptr, lb, sh);
}

// levels fuse..depth-1 of each scene in its own workgroup; cur.lv[l] holds the
// ping-pong bit of level l per scene
__global__ void __launch_bounds__(kCoarseThreads) kb_mg_coarse(const BScene* tab, int ns, SceneSet on, CurBits cb, int variant,
                                                    int f32) {
    const int s = blockIdx.x;
    if (s >= ns || !in_set(on, s)) return;
    const BScene& B = tab[s];
    if (halted(B)) return;
    unsigned cur = 0;
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) cur |= (in_set(cb.lv[l], s) ? 1u : 0u) << l;
    mg_coarse_cycle(B.lv, B.fuseOf[variant], B.depth, cur, f32 != 0);
}

#undef MPH_SCENE

bool any(const SceneSet& m) {
    for (unsigned long long w : m.w)
        if (w) return true;
    return false;
}

}  // namespace

#define MPH_LAUNCH(kernel, row, ...)                                                                     \
    do {                                                                                                 \
        const int g_ = T.blocks(row);                                                                    \
        if (g_ > 0 && any(on))                                                                           \
            hipLaunchKernelGGL(kernel, dim3(g_), dim3(256), 0, T.st, T.tab, T.off, T.ns, on, __VA_ARGS__); \
    } while (0)

void fill(const Tables& T, const SceneSet& on, int role, double v) { MPH_LAUNCH(kb_fill, kOffElem, role, v); }
void fill_level(const Tables& T, const SceneSet& on, int l, double v) {
    MPH_LAUNCH(kb_fill_level, kOffLevel + l, l, v);
}
void cr_init(const Tables& T, const SceneSet& on) {
    const int g = T.blocks(kOffChunk);
    if (g > 0 && any(on)) hipLaunchKernelGGL(kb_cr_init, dim3(g), dim3(256), 0, T.st, T.tab, T.off, T.ns, on);
}
void copy(const Tables& T, const SceneSet& on, int src, int dst) { MPH_LAUNCH(kb_copy, kOffElem, src, dst); }
void axpy(const Tables& T, const SceneSet& on, int src, int dst) { MPH_LAUNCH(kb_axpy, kOffElem, src, dst); }
void spmv(const Tables& T, const SceneSet& on, const SceneSet& first, bool f32, double alpha, int x, double beta,
          int yin, int y, int y2, int p_copy, int dot, int variant) {
    // alpha is +1 or -1 at every call site (main.cpp:4337, 4353, 4357); variants as
    // launch_spmv: FP64 2/3/4 block pairs per load group, FP32 1/2 block quads
#define MPH_SPMV(NEG, VT, UN) MPH_LAUNCH((kb_spmv<NEG, VT, UN>), kOffChunk, first, x, beta, yin, y, y2, p_copy, dot)
    if (f32) {
        if (alpha < 0) { if (variant == 1) MPH_SPMV(true, float, 2); else MPH_SPMV(true, float, 1); }
        else { if (variant == 1) MPH_SPMV(false, float, 2); else MPH_SPMV(false, float, 1); }
    } else if (alpha < 0) {
        if (variant == 1) MPH_SPMV(true, double, 3);
        else if (variant == 2) MPH_SPMV(true, double, 4);
        else MPH_SPMV(true, double, 2);
    } else {
        if (variant == 1) MPH_SPMV(false, double, 3);
        else if (variant == 2) MPH_SPMV(false, double, 4);
        else MPH_SPMV(false, double, 2);
    }
#undef MPH_SPMV
}
void spmv_fac(const Tables& T, const SceneSet& all, const SceneSet& first, const SceneSet& cells, bool f32,
              double alpha, int x, double beta, int yin, int y, int y2, int p_copy, int dot) {
    SceneSet on = all, staged{};
    for (int k = 0; k < kMaxScenes / 64; ++k) staged.w[k] = f32 ? 0 : all.w[k] & cells.w[k];
    if (f32) MPH_LAUNCH(kb_fac_t<float>, kOffPart, x);
    else MPH_LAUNCH(kb_fac_t<double>, kOffPart, x);
    if (any(staged)) {  // the cell-staged pass, then its dot partials
        spmv_fac_cells(T, staged, first, alpha, x, beta, yin, y, y2, p_copy);
        on = staged;
        if (dot >= 0) MPH_LAUNCH(kb_dot_partial, kOffChunk, dot, y);
        on = all;
        for (int k = 0; k < kMaxScenes / 64; ++k) on.w[k] &= ~staged.w[k];
    }
#define MPH_FAC(NEG, VT) MPH_LAUNCH((kb_fac_y<NEG, VT>), kOffChunk, first, x, beta, yin, y, y2, p_copy, dot)
    if (f32) {
        if (alpha < 0) MPH_FAC(true, float);
        else MPH_FAC(false, float);
    } else {
        if (alpha < 0) MPH_FAC(true, double);
        else MPH_FAC(false, double);
    }
#undef MPH_FAC
}
void dot_partial(const Tables& T, const SceneSet& on, int a, int b) { MPH_LAUNCH(kb_dot_partial, kOffChunk, a, b); }
void dot_final(const Tables& T, const SceneSet& on, int slot, int shift_slot) {
    if (T.ns > 0 && any(on))
        hipLaunchKernelGGL(kb_dot_final, dim3(T.ns), dim3(256), 0, T.st, T.tab, T.ns, on, slot, shift_slot);
}
void dot_final_stop(const Tables& T, const SceneSet& on, int step) {
    if (T.ns > 0 && any(on))
        hipLaunchKernelGGL(kb_dot_final_stop, dim3(T.ns), dim3(256), 0, T.st, T.tab, T.ns, on, step);
}
void ctl_reset(const Tables& T, const SceneSet& on) {
    if (T.ns > 0 && any(on)) hipLaunchKernelGGL(kb_ctl, dim3(T.ns), dim3(64), 0, T.st, T.tab, T.ns, on, 0);
}
void ctl_release(const Tables& T, const SceneSet& on) {
    if (T.ns > 0 && any(on)) hipLaunchKernelGGL(kb_ctl, dim3(T.ns), dim3(64), 0, T.st, T.tab, T.ns, on, 1);
}
void pq_update(const Tables& T, const SceneSet& on) {
    const int g = T.blocks(kOffElem);
    if (g > 0 && any(on)) hipLaunchKernelGGL(kb_pq_update, dim3(g), dim3(256), 0, T.st, T.tab, T.off, T.ns, on);
}
void xrs_update(const Tables& T, const SceneSet& on, int x) { MPH_LAUNCH(kb_xrs_update, kOffChunk, x); }
void store_velocity(const Tables& T, const SceneSet& on) {
    const int g = T.blocks(kOffElem);
    if (g > 0 && any(on))
        hipLaunchKernelGGL(kb_store_velocity, dim3(g), dim3(256), 0, T.st, T.tab, T.off, T.ns, on);
}
void mg_p2g(const Tables& T, const SceneSet& on, int in) { MPH_LAUNCH(kb_mg_p2g, kOffLevel, in); }
void mg_p2g_bu(const Tables& T, const SceneSet& on, int in, const SceneSet& cur) {
    MPH_LAUNCH(kb_mg_p2g_bu, kOffLevel, in, cur);
}
void mg_restrict_bu(const Tables& T, const SceneSet& on, int l, const SceneSet& cur) {
    MPH_LAUNCH(kb_mg_restrict_bu, kOffLevel + l, l, cur);
}
void mg_begin(const Tables& T, const SceneSet& on, int l, const SceneSet& cur) {
    MPH_LAUNCH(kb_mg_begin, kOffLevel + l, l, cur);
}
void mg_update(const Tables& T, const SceneSet& on, int l, const SceneSet& cur) {
    MPH_LAUNCH(kb_mg_update, kOffLevel + l, l, cur);
}
void mg_residual(const Tables& T, const SceneSet& on, int l, const SceneSet& cur, bool f32) {
    MPH_LAUNCH(kb_mg_residual, kOffLevel + l, l, cur, f32 ? 1 : 0);
}
void mg_residual_update(const Tables& T, const SceneSet& on, int l, const SceneSet& cur, bool f32) {
    MPH_LAUNCH(kb_mg_residual_update, kOffLevel + l, l, cur, f32 ? 1 : 0);
}
void mg_restrict(const Tables& T, const SceneSet& on, int l) { MPH_LAUNCH(kb_mg_restrict, kOffLevel + l, l); }
void mg_prolong(const Tables& T, const SceneSet& on, int l, const SceneSet& curL, const SceneSet& curS) {
    MPH_LAUNCH(kb_mg_prolong, kOffLevel + l - 1, l, curL, curS);
}
void mg_g2p(const Tables& T, const SceneSet& on, const SceneSet& cur0, int in, int out, bool part) {
    MPH_LAUNCH(kb_mg_g2p, kOffChunk, cur0, in, out, part ? 1 : 0);
}

void mg_coarse(const Tables& T, const SceneSet& on, const CurBits& cb, int variant, bool f32) {
    if (T.ns > 0 && any(on))
        hipLaunchKernelGGL(kb_mg_coarse, dim3(T.ns), dim3(kCoarseThreads), 0, T.st, T.tab, T.ns, on, cb, variant, f32 ? 1 : 0);
}

#undef MPH_LAUNCH

}  // namespace batch
}  // namespace mph
