// mph_kernels.h -- launchers of the HIP kernels (one scene, one stream).
//
// Layouts in HBM (all device pointers):
//   particles    AoS double[np][3] for Position/Velocity/Force (as the reference),
//                double[np] scalars, int Property[np], int perm[np]
//   cell keys    uint64 keys[p2] = (CellIndex << 32) | CellParticle while sorting;
//                int cellIndex[np] (sorted keys), int cellStart[2T+1]
//   neighbours   int64 ptr[np+1], int ind[E]; same for the RadiusP lists
//   A            sliced 3x3-block ELL: slice = 64 consecutive fluid particles,
//                width = max fluid-neighbour count in the slice rounded up to
//                a multiple of 4, sliceStart = prefix sum of widths.  Blocks
//                k = 2m, 2m+1 of a particle form a pair: its two columns are
//                adjacent ints and entry e = 3r+s of both blocks are adjacent
//                values (a_col_index / a_val_index), so a wavefront reads 512 B
//                of columns and 1 KiB (FP64) per entry of a block pair
//                with one 8/16-byte load per lane; the FP32 copy groups four
//                blocks per entry (a_val32_index)
//   C            cval[3*(ptrP[i] + jn) + r] for fluid i, P-neighbour jn
//   multigrid    per level: A stored [shift][row], vectors [row]
//   CR vectors   double[3*nf], row 3i+r as the reference
#pragma once

#include "mph_common.h"

namespace mph {

struct WallKin {  // per wall type: rotation, angular and linear velocity, centre
    double R[4][3][3];
    double omega[4][3];
    double vel[4][3];
    double center[4][3];
};

struct Level {  // geometry + storage of one multigrid level
    int mn[3], mx[3], cnt[3], cs[3];
    int cells, nrow, nshift;
    double* A;     // [nshift][nrow]
    double* invD;  // [nrow]
    double* S[2];  // ping-pong smoother iterate
    double* R;
    double* Q;
    int cur;       // which S buffer holds the current iterate
    float* A32;    // FP32 copy of A for the mixed-precision inner solve (or null)
    int f32;       // smoothers read A32 instead of A
    int lanes;     // sweeps split each row over lanes (small levels; set by the host)
    const int* halt;  // the CR stop flag, as DevParams::halt
};

// levels with fewer rows than this split their sweeps' rows over lanes by default
constexpr int kLaneRowsDefault = 32768;

struct ASlices {
    int nslices;
    const int* width;          // [nslices]
    const long long* start;    // [nslices+1]
    int* col;                  // [slots]
    double* val;               // [slots*9]
    float* val32;              // [slots*9] (mixed precision copy) or null
};

// The particle matrix in factored form (a scalar or a vector per neighbour-list
// entry instead of nine values per block; derivation in mph_factored.hip).  The
// values sit beside the neighbour lists, in their order: alpha[nPtr[i] + jn] for the
// fluid neighbours jn < nFluid[i] of fluid i, beta[nPtrP[k] + jn] for the P
// neighbours of every k.
struct FacOp {
    const int* nFluid;        // [np] fluid neighbours (they lead each list)
    const long long* nPtr;    // neighbour lists
    const int* nInd;
    const void* alpha;        // viscous scalar (double, or float in FP32 form)
    const double* diag;       // [nf][9] viscous self block + inertia
    const double* selfC;      // [nf][3] C_ii
    const int* nCountP;       // [np] P entries
    const int* runP;          // [np] P entries up to and including the last fluid one
    const long long* nPtrP;   // P lists
    const int* nIndP;
    const double* beta;       // -w_p'(d)/d per P entry
    const double* lambda;     // [np]
    const double* pos;        // [np][3]
    // [7 np] the operand x (fluid), the position and t = lambda_k C_k^T x of every
    // particle in four planes: [np] pairs x0 x1, [np] pairs x2 p0, [np] pairs p1 p2,
    // then [np] t (body::fac_planes).  The second pass gathers a neighbour's
    // operands with three 16-byte loads, and the eight lanes of a row, which walk
    // consecutive particles of the cell-ordered list, read consecutive 16-byte
    // pairs.  Positions are written with the P slices; x and t by every first pass.
    double* rec;
    // 1 when some neighbour pair interacts through the periodic boundary (then the
    // displacements take the minimum image); set by the matrix-A launch
    const int* wrap;
    // FP32 form (the inner solves of mixed precision), which needs no positions:
    // u = sqrt(alpha) r per fluid-neighbour entry and v = beta r per P entry, each a
    // 16-byte entry beside the lists -- three floats and the column's bits in the
    // fourth (u32[4*(nPtr[i]+jn)+c], v32[4*(nPtrP[k]+jn)+c]; body::fac_entry32), so an
    // entry is one load -- and a 16-byte record x0 x1 x2 t per particle
    const float* u32;
    const float* v32;
    float* rec32;
    // cell-staged FP64 second pass (null cellStart: the chunk kernel): one workgroup
    // per cell of the fluid bounding box cellLo + [0, cellCnt), which stages the
    // records of the row cell's visited cells in LDS for the viscous gathers
    const int* cellStart;
    int cellLo[3], cellCnt[3];
    int cellXcd;  // 1: each XCD takes a contiguous range of cells (else round robin)
    // [E] per fluid-neighbour entry of a fluid row: the neighbour's place in the row
    // cell's staged records (launch_fac_slots), read instead of the column (or null)
    unsigned short* slot;
};

constexpr int kSlice = 64;    // particles per A slice (one wavefront)

// block k of particle lane `lane` of the slice whose first slot is sb (even)
__host__ __device__ inline long long a_col_index(long long sb, int k, int lane) {
    return ((sb / 2 + k / 2) * 64 + lane) * 2 + (k & 1);
}
__host__ __device__ inline long long a_val_index(long long sb, int k, int e, int lane) {
    return (((sb / 2 + k / 2) * 9 + e) * 64 + lane) * 2 + (k & 1);
}
// the FP32 copy of A (mixed precision) groups blocks 4q..4q+3 per entry instead, so
// its loads are 16 bytes per lane as well
__host__ __device__ inline long long a_val32_index(long long sb, int k, int e, int lane) {
    return (((sb / 4 + k / 4) * 9 + e) * 64 + lane) * 4 + (k & 3);
}
constexpr int kChunk = 256;   // particles per dot-product chunk / CR workgroup
constexpr int kRuns = 28;     // run starts per fluid particle: 27 neighbour cells + end
// the factored form's level-0 run sums: per fluid particle and visited cell, the sums
// of -alpha r r^T over that cell's run of fluid entries, a symmetric block kept as its
// upper triangle (xx xy xz yy yz zz) in records of kBlkStride doubles
constexpr int kBlkRuns = kRuns - 1, kBlkStride = 6;
constexpr int kMaxP = 40;     // P-neighbours staged in LDS per particle by the C Lambda C^T kernel
constexpr int kMaxLevels = 20;   // multigrid depth = floor(log2(cells per axis)) + 1
// the V-cycle levels with at most this many rows may run in the one-workgroup
// coarse cycle; the choice among these thresholds is autotuned (0: none)
constexpr int kFuseVariants = 3;
constexpr int kFuseRowsOf[kFuseVariants] = {0, 256, 2048};

// ---- cells (calculateCellParticle) ----
void launch_hash(hipStream_t st, const DevParams& P, const double* pos, const int* type, uint64_t* keys);
void launch_bitonic(hipStream_t st, const DevParams& P, uint64_t* keys);
void launch_cell_start(hipStream_t st, const DevParams& P, const uint64_t* keys, int* cellStart);
void launch_reorder(hipStream_t st, const DevParams& P, const uint64_t* keys, const int* type, const int* perm,
                    const double* pos, const double* vel, int* type2, int* perm2, double* pos2, double* vel2,
                    int* cellIndex);
// ---- neighbours (calculateNeighbor) ----
// maxFluidP[4]: largest P count and fluid count of a fluid particle, largest count and
// P count of any particle.  With scrM set, the lists also go into per-particle
// scratch (capM / capP entries) and runStart is written; *overflow is raised if a
// list does not fit (then launch_neighbor_fill builds the lists instead).
void launch_neighbor_count(hipStream_t st, const DevParams& P, const double* pos, const int* cellIndex,
                           const int* cellStart, int* nCount, int* nFluid, int* nCountP, int* maxFluidP, int* scrM,
                           int capM, int* scrP, int capP, unsigned short* runStart, int* overflow);
// the lists from the count pass's scratch
void launch_neighbor_compact(hipStream_t st, const DevParams& P, const int* nCount, const int* nCountP,
                             const long long* nPtr, const long long* nPtrP, const int* scrM, int capM,
                             const int* scrP, int capP, int* nInd, int* nIndP);
void launch_scan_i64(hipStream_t st, const int* in, long long* out, int n, long long* scratch, int scratch_len);
int scan_scratch_len(int n);
void launch_neighbor_fill(hipStream_t st, const DevParams& P, const double* pos, const int* cellIndex,
                          const int* cellStart, const int* nCount, const long long* nPtr, const long long* nPtrP,
                          int* nInd, int* nIndP, unsigned short* runStart);
void launch_slice_width(hipStream_t st, int nf, int nslices, const int* nFluid, int* width);
// ---- kinematics and explicit forces ----
void launch_kinematics(hipStream_t st, const DevParams& P, const WallKin& wk, const int* type, double* pos,
                       double* vel, double* force);
void launch_densities(hipStream_t st, const DevParams& P, const int* type, const double* pos, const int* nCount,
                      const long long* nPtr, const int* nInd, double* densA, double* gcen, double* volStrainP);
void launch_coefficients(hipStream_t st, const DevParams& P, const int* type, const double* densA,
                         const double* volStrainP, double* mass, double* kappa, double* lambda, double* mu,
                         double* pressA);
void launch_forces(hipStream_t st, const DevParams& P, const int* type, const double* pos, const int* nCount,
                   const long long* nPtr, const int* nInd, const double* pressA, const double* gcen,
                   const double* mass, double* force, double* vel);
// ---- implicit assembly ----
// alpha/diag/wrap (each may be null): the factored form's viscous scalar per
// fluid-neighbour entry (FacOp order), the diagonal block without C Lambda C^T, and
// the flag set when a pair interacts through the periodic boundary (wrap must be 0
// before the launch).  blk (with runStart; may be null): the level-0 run sums
// [nf][kBlkRuns][kBlkStride] of the factored form, each in list order
void launch_matrix_a(hipStream_t st, const DevParams& P, const int* type, const double* pos, const double* vel,
                     const double* mass, const double* mu, const int* nCount, const int* nFluid,
                     const long long* nPtr, const int* nInd, ASlices A, int* selfIdx, double* b,
                     double* alpha = nullptr, double* diag = nullptr, int* wrap = nullptr,
                     const unsigned short* runStart = nullptr, double* blk = nullptr);
void launch_matrix_c(hipStream_t st, const DevParams& P, const double* pos, const int* nCountP,
                     const long long* nPtrP, const int* nIndP, const double* lambda, double* cval, double* dval);
void launch_vector_p(hipStream_t st, const DevParams& P, const int* type, const double* pos, const double* vel,
                     const double* kappa, const double* lambda, const double* volStrainP, const int* nCountP,
                     const long long* nPtrP, const int* nIndP, double* vecP);
void launch_b_minus_cp(hipStream_t st, const DevParams& P, const int* nCountP, const long long* nPtrP,
                       const int* nIndP, const double* cval, const double* vecP, double* b);
// kernel modes: 0 P lists from global memory; 2 both lists staged in LDS; 3 staged,
// and the merge records the matches before any arithmetic.  The tuner chooses
// between modes 3 and 2 (variants 0, 1); mode 0 runs when a fluid particle has more
// than kMaxP P-neighbours.  Same bits in every mode.
constexpr int kCacVariants = 2;
constexpr int kCacModeOf[kCacVariants] = {3, 2};
void launch_cac(hipStream_t st, const DevParams& P, const int* nFluid, const int* nCountP, const long long* nPtrP,
                const int* nIndP, const double* cval, const double* dval, ASlices A, int mode);  // also writes A.val32 when it is non-null
// ---- multigrid ----
void launch_fluid_bbox(hipStream_t st, const DevParams& P, const int* cellStart, int* bbox);
void launch_inv_diag(hipStream_t st, const DevParams& P, const int* selfIdx, ASlices A, double* invDiagA);
// variant 0: per-thread sums straight from A; 1: the rows of four adjacent
// particles staged in LDS per round (rows of at most 216 blocks).  Same bits.
constexpr int kGalVariants = 2, kGalRowMax = 216;
void launch_mg_galerkin0(hipStream_t st, const DevParams& P, const Level& L, const int* cellStart,
                         const unsigned short* runStart, const int* nFluid, ASlices A, int variant);
void launch_mg_galerkinL(hipStream_t st, const Level& S, const Level& L);
void launch_mg_invd(hipStream_t st, const Level& L);
void launch_fill(hipStream_t st, double* x, long long n, double v);
void launch_fill_i32(hipStream_t st, int* x, long long n, int v);  // a kernel, not a memset node
void launch_mg_p2g(hipStream_t st, const DevParams& P, Level& L, const int* cellStart, const double* q);
void launch_mg_p2g_bu(hipStream_t st, const DevParams& P, Level& L, const int* cellStart, const double* q);
// the same after q = beta q + y (beta = rho / rho_prev from scal), written back to q
void launch_mg_p2g_bu_pq(hipStream_t st, const DevParams& P, Level& L, const int* cellStart, double* q,
                         const double* y, const double* scal);
void launch_mg_restrict_bu(hipStream_t st, const Level& S, Level& L);  // restriction, begin, update
void launch_mg_begin(hipStream_t st, Level& L);        // S *= 0; R = Q
void launch_mg_update(hipStream_t st, Level& L);       // S += invD R
void launch_mg_residual(hipStream_t st, Level& L);     // R = Q - A S
void launch_mg_residual_update(hipStream_t st, Level& L);  // R = Q - A S; S += invD R
void launch_mg_restrict(hipStream_t st, const Level& S, Level& L);
void launch_mg_prolong(hipStream_t st, const Level& L, Level& S);
// levels lc..D-1 of the V-cycle in one workgroup; lvs = device copy of the levels,
// cur = their ping-pong bits (bit l, lc-1 included).  Afterwards every level in
// lc..D-2 has flipped (5 sweeps) and the coarsest has not (4 sweeps).  halt: the CR
// stop flag (DevParams::halt), or null.
void launch_mg_coarse(hipStream_t st, const Level* lvs, int lc, int D, unsigned cur, bool f32, const int* halt);
void launch_mg_g2p(hipStream_t st, const DevParams& P, const Level& L0, const int* cellIndex,
                   const double* invDiagA, const double* q, double* u, double* part);
// ---- conjugate residual ----
void launch_cr_init(hipStream_t st, const DevParams& P, const double* vel, const double* force,
                    const double* mass, double* x);
// SpMV variants differ only in how many block pairs are loaded per group (same
// bits): FP64 2, 3 or 4 block pairs; FP32 1 or 2 block quads
constexpr int kSpmvVariants64 = 3, kSpmvVariants32 = 2;
void launch_spmv(hipStream_t st, const DevParams& P, ASlices A, const int* nFluid, double alpha, const double* x,
                 double beta, const double* yin, double* y, double* y2, double* p_copy, const double* dotx,
                 double* part, bool f32 = false, int variant = 0);
// factored form: beta, runP, C_ii and the records' positions (after matrix A has
// written alpha and the diagonal blocks), and y = beta*yin +- A x with the same
// outputs as launch_spmv (two launches: t over every particle, then the fluid rows)
void launch_fac_p(hipStream_t st, const DevParams& P, const double* pos, const int* nCountP, const long long* nPtrP,
                  const int* nIndP, double* beta, float* v32, int* runP, double* selfC, double* rec);
// multigrid level 0 (calculateMultiGridCsrA at level 0) and InvDiagA from the
// factored form, with per-particle scratch gtab [np][27][3] and gmask [np] (level-0
// counts of at least 3 per active axis: the caller checks)
// blk: the run sums launch_matrix_a wrote
void launch_fac_level0(hipStream_t st, const DevParams& P, const Level& L0, const FacOp& F, const int* cellIndex,
                       const int* cellStart, const double* blk, double* gtab, int* gmask, double* invDiagA);
// u32 of the FP32 form from alpha and the positions (after launch_fac_p)
void launch_fac_u(hipStream_t st, const DevParams& P, const int* nFluid, const long long* nPtr, const int* nInd,
                  const double* alpha, const double* pos, float* u32);
// passes: 1 the first pass (t from x) only, 2 the second only (t already in F.rec),
// 3 both
void launch_fac_spmv(hipStream_t st, const DevParams& P, const FacOp& F, double alpha, const double* x, double beta,
                     const double* yin, double* y, double* y2, double* p_copy, const double* dotx, double* part,
                     bool f32, int passes = 3);
// the positions of the FP64 records (planes b odd and c of FacOp::rec, 7*np doubles)
// copied from one record buffer into another; x and t are written by each product
void launch_fac_rec_positions(hipStream_t st, const DevParams& P, const double* from, double* to);
// FacOp::slot of the fluid-neighbour entries (after the lists; the cell-staged pass)
void launch_fac_slots(hipStream_t st, const DevParams& P, const FacOp& F);

void launch_convert_f32(hipStream_t st, const double* in, float* out, long long n);
void launch_axpy(hipStream_t st, long long n, const double* x, double* y);  // y += 1.0*x
void launch_dot_partial(hipStream_t st, const DevParams& P, const double* x, const double* y, double* part);
// host: optional coherent pinned host memory that also receives the value
void launch_dot_final(hipStream_t st, const DevParams& P, const double* part, double* scal, int slot, int shift_slot,
                      double* host = nullptr);
// The CR loop's control block int[kNCtl] in device memory: the stop flag the
// iteration kernels test (DevParams::halt points at ctl[kCtlHalt]), the iterations
// done and a non-finite flag.
enum { kCtlHalt = 0, kCtlIter = 1, kCtlBad = 2, kNCtl = 4 };
// the r.r reduction of the CR loop (scal[2]) followed by the loop's stop test
// (main.cpp:4344-4346) on the device: ctl[kCtlIter] += step; then the loop stops
// (ctl[kCtlHalt] = 1) once ctl[kCtlIter] >= maxit, or on a residual that is not
// finite (also ctl[kCtlBad] = 1), or when rr/rr0 < tol, tested in that order.  A
// no-op once halted.  hctl and hrr: optional coherent pinned host memory that also
// receives ctl[0..2] and rr.
void launch_dot_final_stop(hipStream_t st, const DevParams& P, const double* part, double* scal, int* ctl, int step,
                           double tol, long long maxit, int* hctl, double* hrr);
// The convergence record of a CR iteration (main.cpp:4348-4351): the partials' sum
// r.s into scal[kRSIter], then rec[2 it] = rr (scal[kRR]) and rec[2 it + 1] = r.s
// for it = ctl[kCtlIter] (ctl null: the host's `it`), rows it >= cap skipped.  A
// no-op once halted.
void launch_dot_final_record(hipStream_t st, const DevParams& P, const double* part, double* scal, const int* ctl,
                             long long it, double* rec, long long cap);
void launch_pq_update(hipStream_t st, const DevParams& P, const double* scal, const double* s, const double* y,
                      double* p, double* q);
// pq: p = beta p + s first, written back
void launch_xrs_update(hipStream_t st, const DevParams& P, const double* scal, double* p, const double* q,
                       const double* u, double* x, double* r, double* s, double* part, bool pq = false);
void launch_store_velocity(hipStream_t st, const DevParams& P, const double* x, double* vel);
// ---- output-step diagnostics ----
void launch_diagnostics(hipStream_t st, const DevParams& P, const int* type, const double* pos, const double* vel,
                        const double* mu, const double* lambda, const double* kappa, const double* volStrainP,
                        const double* pressA, const double* gcen, const int* nCount, const long long* nPtr,
                        const int* nInd, const int* nCountP, const long long* nPtrP, const int* nIndP, double* divP,
                        double* pressP, double* force, double* virialStress, double* virialP);

}  // namespace mph
