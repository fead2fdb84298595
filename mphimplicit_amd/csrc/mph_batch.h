// mph_batch.h -- internal: scene-table launches of the CR loop and the V-cycle for
// the batched executor.  One launch covers every scene of a batch: workgroup b
// belongs to the scene s with off[row][s] <= b < off[row][s+1] and runs that
// scene's rows with that scene's pointers, through the same bodies as the
// single-scene kernels (mph_cr_body.h).  Scenes outside the launch's set return
// at once, so per-scene control flow (converged scenes, different multigrid
// depths, ping-pong buffers) travels as bit sets in the kernel arguments.
#pragma once

#include "mph_kernels.h"

namespace mph {
namespace batch {

constexpr int kMaxScenes = 256;  // bits of a SceneSet

// CR vectors by role
enum { kX = 0, kR, kS, kY, kP, kQ, kU, kB, kROut, kD, kNVec };

// rows of the block-offset table: 256-particle chunks of the fluid particles,
// 256-element blocks of the 3*nf vectors, 256-particle blocks of all particles, the
// fluid bounding box's cells (the cell-staged factored pass; 0 for the other
// scenes), then 256-row blocks of each multigrid level
enum { kOffChunk = 0, kOffElem = 1, kOffPart = 2, kOffCell = 3, kOffLevel = 4, kOffRows = kOffLevel + kMaxLevels };

struct SceneSet {
    unsigned long long w[kMaxScenes / 64];
};

__device__ inline bool in_set(const SceneSet& m, int s) { return (m.w[s >> 6] >> (s & 63)) & 1ull; }

// scene of workgroup b: the last s with o[s] <= b (o[0] = 0, o[ns] = total)
__device__ inline int find_scene(const int* __restrict__ o, int ns, int b) {
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (o[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct CurBits {  // per level, the scenes whose ping-pong bit is set
    SceneSet lv[kMaxLevels];
};

struct BScene {
    DevParams P;
    const int* nFluid;
    const long long* start;
    const int* col;
    const double* val;
    const float* val32;
    const int* cellStart;
    const int* cellIndex;
    const double* invDiagA;
    const double* force;
    const double* mass;
    double* vel;
    double* vec[kNVec];
    double* part;
    double* scal;
    int depth;
    int mg;
    int fuseOf[kFuseVariants];  // first level of the one-workgroup coarse cycle per variant
    int factored;               // MPH_OPERATOR_FACTORED: products through fac, not val/val32
    FacOp fac[2];               // FP64 and FP32 scalars (FP32: mixed precision only)
    Level lv[kMaxLevels];  // .cur is ignored: it comes with each launch
    // the FP64 lockstep solve's stop test on the device (null in the mixed solve):
    // every kernel returns at once for a scene whose ctl[kCtlHalt] is set; hctl is the
    // scene's coherent pinned mirror of ctl, tol and maxit its stop rule
    int* ctl;
    int* hctl;
    double tol;
    long long maxit;
};

// a scene's kernels return at once after its CR loop has stopped on the device
__device__ inline bool halted(const BScene& B) { return B.ctl && B.ctl[kCtlHalt]; }

// device tables of one batch solve; hoff is the host copy of the offsets
struct Tables {
    hipStream_t st;
    const BScene* tab;
    const int* off;
    int ns;
    const int* hoff;
    int blocks(int row) const { return hoff[row * (ns + 1) + ns]; }
};

void fill(const Tables& T, const SceneSet& on, int role, double v);
void fill_level(const Tables& T, const SceneSet& on, int l, double v);
void cr_init(const Tables& T, const SceneSet& on);
void copy(const Tables& T, const SceneSet& on, int src, int dst);
void axpy(const Tables& T, const SceneSet& on, int src, int dst);  // dst += 1.0*src
// y = beta*yin + sum (alpha*a) x with alpha = +-1; y2 and p_copy only for scenes in
// `first`; dot partial of dot.y into the scene's part when dot >= 0
void spmv(const Tables& T, const SceneSet& on, const SceneSet& first, bool f32, double alpha, int x, double beta,
          int yin, int y, int y2, int p_copy, int dot, int variant);
// the same product for scenes of the factored form; `cells`: the scenes whose FP64
// second pass is the cell-staged one (spmv_fac_cells), the others take the chunk
// kernel
void spmv_fac(const Tables& T, const SceneSet& on, const SceneSet& first, const SceneSet& cells, bool f32,
              double alpha, int x, double beta, int yin, int y, int y2, int p_copy, int dot);
// the cell-staged second pass of the FP64 factored product (after the first pass)
// for the scenes of `on`, with no dot partials (mph_factored.hip).  Its grid covers
// the cells of the scenes in `on` only: pre[s] is the first workgroup of scene s
// (scenes outside `on` have empty ranges), passed with the launch.
struct CellRanges {
    int pre[kMaxScenes + 1];
};
void spmv_fac_cells(const Tables& T, const SceneSet& on, const SceneSet& first, double alpha, int x, double beta,
                    int yin, int y, int y2, int p_copy);
void dot_partial(const Tables& T, const SceneSet& on, int a, int b);
void dot_final(const Tables& T, const SceneSet& on, int slot, int shift_slot);
// r.r and the stop test per scene, as launch_dot_final_stop (BScene::ctl, tol, maxit)
void dot_final_stop(const Tables& T, const SceneSet& on, int step);
void ctl_reset(const Tables& T, const SceneSet& on);    // ctl = 0 (before the solve's first kernel)
void ctl_release(const Tables& T, const SceneSet& on);  // ctl[kCtlHalt] = 0 (kernels after the loop run)
void pq_update(const Tables& T, const SceneSet& on);
void xrs_update(const Tables& T, const SceneSet& on, int x);
void store_velocity(const Tables& T, const SceneSet& on);
void mg_p2g(const Tables& T, const SceneSet& on, int in);
void mg_p2g_bu(const Tables& T, const SceneSet& on, int in, const SceneSet& cur);  // P2G, begin, update
void mg_restrict_bu(const Tables& T, const SceneSet& on, int l, const SceneSet& cur);  // l-1 -> l, begin, update
void mg_begin(const Tables& T, const SceneSet& on, int l, const SceneSet& cur);
void mg_update(const Tables& T, const SceneSet& on, int l, const SceneSet& cur);
void mg_residual(const Tables& T, const SceneSet& on, int l, const SceneSet& cur, bool f32);
void mg_residual_update(const Tables& T, const SceneSet& on, int l, const SceneSet& cur, bool f32);
void mg_restrict(const Tables& T, const SceneSet& on, int l);  // level l-1 -> l
void mg_prolong(const Tables& T, const SceneSet& on, int l, const SceneSet& curL, const SceneSet& curS);  // l -> l-1
void mg_g2p(const Tables& T, const SceneSet& on, const SceneSet& cur0, int in, int out, bool part);
// levels fuseOf[variant]..depth-1 of each scene in one workgroup
void mg_coarse(const Tables& T, const SceneSet& on, const CurBits& cb, int variant, bool f32);

}  // namespace batch
}  // namespace mph
