// mph_scene_impl.h -- internal: the scene record behind the C ABI handle and the
// phase drivers that the batched executor (mph_batch.cpp) shares with the
// single-scene path (mph_scene.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mph_abi.h"
#include "mph_kernels.h"

namespace mph {
namespace host {

// record the message of the failing call (mph_last_error) and return the code
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return ::mph::host::fail(MPH_E_DEVICE, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    int ensure(size_t n) {
        if (n <= cap && p) return MPH_OK;
        size_t want = n + n / 4 + 64;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc((void**)&p, want * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return fail(MPH_E_NOMEM, "device allocation of " + std::to_string(want * sizeof(T)) + " bytes failed");
        }
        cap = want;
        return MPH_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct HostLevel {
    Level L{};
    DBuf<double> A, invD, S0, S1, R, Q;
    DBuf<float> A32;
};

enum { kRR0 = 0, kRS0 = 1, kRR = 2, kRHO = 3, kRHOP = 4, kTMP = 5, kRS = 6, kRSIter = 7, kCheckRR = 8, kCheckRS = 9 };

// Run-time choice among kernel variants that compute identical bits: while
// exploring, launches cycle through the variants with HIP events around them;
// once every variant has `trials` timings, the variant with the fastest single
// trial is kept for good.
struct Autotune {
    int nvar = 1, trials = 1, best = 0;
    std::vector<double> fastest;  // per variant, its fastest trial: a variant's first
                                  // launches can carry one-off costs (code loading,
                                  // cold caches) that a mean would charge to it
    std::vector<int> cnt;
    struct Pend {
        int v;
        hipEvent_t a, b;
    };
    std::vector<Pend> pend;
    std::vector<hipEvent_t> pool;
    void init(int n, int t) {
        nvar = n;
        trials = t;
        best = n == 1 ? 0 : -1;
        fastest.assign(n, 0.0);
        cnt.assign(n, 0);
    }
    int pick() const {  // the variant for the next launch
        if (best >= 0) return best;
        int v = 0;
        for (int k = 1; k < nvar; ++k)
            if (cnt[k] < cnt[v]) v = k;
        return v;
    }
    bool exploring() const { return best < 0; }
    hipEvent_t event() {
        if (pool.empty()) {
            hipEvent_t e = nullptr;
            (void)hipEventCreate(&e);
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    void begin(hipStream_t st, int v) {
        if (!exploring()) return;
        Pend p{v, event(), event()};
        (void)hipEventRecord(p.a, st);
        pend.push_back(p);
    }
    void end(hipStream_t st) {
        if (!exploring() || pend.empty()) return;
        (void)hipEventRecord(pend.back().b, st);
    }
    // the first n recorded trials (all by default), after the stream has been
    // synchronised past their launches; later ones stay pending
    void collect(size_t n = (size_t)-1) {
        if (n > pend.size()) n = pend.size();
        for (size_t k = 0; k < n; ++k) {
            const Pend& p = pend[k];
            float ms = 0;
            if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess && p.v >= 0 && p.v < nvar) {
                if (cnt[p.v] == 0 || ms < fastest[p.v]) fastest[p.v] = ms;
                cnt[p.v]++;
            }
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pend.erase(pend.begin(), pend.begin() + n);
        if (best >= 0) return;
        for (int k = 0; k < nvar; ++k)
            if (cnt[k] < trials) return;
        best = 0;
        for (int k = 1; k < nvar; ++k)
            if (fastest[k] < fastest[best]) best = k;
    }
    void release() {
        for (Pend& p : pend) pool.push_back(p.a), pool.push_back(p.b);
        pend.clear();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

}  // namespace host
}  // namespace mph

struct mph_scene {
    mph_params prm{};
    mph::DevParams P{};
    double dmax[3] = {0, 0, 0};
    int device = 0;
    hipStream_t st = nullptr;
    // host-side state of the reference's globals
    double time = 0, outputNext = 0, vtkOutputNext = 0;
    int iStep = 0;
    bool vtkStep = false;
    mph::WallKin wk{};
    double Swa = 1, Swg = 1, Swp = 1, Swv = 1;
    // particles
    mph::host::DBuf<int> type, type2, perm, perm2;
    mph::host::DBuf<double> pos, pos2, vel, vel2, force, mass, kappa, lambda, mu, densA, gcen, pressA, volStrainP, divP,
        pressP, virialP, virialStress;
    // cells
    mph::host::DBuf<uint64_t> keys;
    mph::host::DBuf<int> cellIndex, cellStart, bbox;
    // neighbours
    mph::host::DBuf<int> nCount, nFluid, nCountP, nInd, nIndP, maxP;
    mph::host::DBuf<long long> nPtr, nPtrP, scanScratch;
    mph::host::DBuf<unsigned short> runStart;
    long long E = 0, EP = 0;
    int hmaxP = 0;
    int hmaxW = 0;  // largest fluid-neighbour count (A row length in blocks)
    // the neighbour search in one scan: per-particle scratch lists sized by the last
    // search's largest counts (hmaxN, hmaxNP) plus a margin; searchCapM = 0 when the
    // current search uses the count-then-fill passes instead
    mph::host::DBuf<int> scrM, scrP, searchOvf;
    int hmaxN = 0, hmaxNP = 0, searchCapM = 0, searchCapP = 0;
    bool oneScan = false;  // the last search built its lists from the scratch
    // A and C
    int nslices = 0;
    mph::host::DBuf<int> width, selfIdx, acol;
    mph::host::DBuf<long long> sliceStart;
    mph::host::DBuf<double> aval, b, cval, dval, vecP;
    mph::host::DBuf<float> aval32;  // FP32 copy of A (mixed precision, assembled operator only)
    long long slots = 0;
    bool assembled = false;  // A (and the factored scalars) match the current particles
    // factored operator (MPH_OPERATOR_FACTORED): scalars beside the neighbour lists
    mph::host::DBuf<int> runP, fwrap, fgmask;
    mph::host::DBuf<double> fgtab;  // level 0 from the factored form: per-particle G tables
    mph::host::DBuf<double> fblk;   // and the viscous run sums [nf][kBlkRuns][kBlkStride]
    bool blocks = true;  // this step assembles the 3x3 blocks (see blocks_wanted)
    mph::host::DBuf<double> falpha, fdiag, selfC, fbeta, frec;  // FacOp::alpha, diag, selfC, beta, rec
    mph::host::DBuf<float> fu32, fv32, frec32;  // FP32 form (mixed precision): FacOp::u32, v32, rec32
    // FP64 second pass staged per cell in LDS: 1 always, 0 never (the chunk kernel),
    // 2 (default) in 3D scenes and in scenes of at most kFacCellsSmall fluid
    // particles.  Measured per product, 2D 130k-188k fluid: chunk 76-88 us, cells
    // 174-241 us (a 2D cell holds ~12 rows, a 512-lane workgroup stages 9 cells for
    // them); 3D 166k-176k: chunk 269-282 us, cells 189-207 us; 2.2k-2.8k fluid, 2D and
    // 3D: chunk 41-68 us, cells 23-27 us.
    int facCells = 2;
    bool facCellXcd = true;  // FacOp::cellXcd
    mph::host::DBuf<unsigned short> fslot;  // FacOp::slot
    bool fslotValid = false;                // fslot matches the current lists
    // multigrid
    int hbbox[6] = {0, 0, 0, 0, 0, 0};
    int depth = 0;
    int laneRowsMax = mph::kLaneRowsDefault;  // levels with fewer rows take the lane-split sweeps
    int fuseOf[mph::kFuseVariants] = {0, 0, 0};  // first level of the one-workgroup coarse cycle
    std::vector<mph::host::HostLevel> lv;
    std::vector<mph::Level> hlv;           // level records as uploaded to lvDev
    mph::host::DBuf<mph::Level> lvDev;
    mph::host::DBuf<double> invDiagA;
    // CR
    mph::host::DBuf<double> x, r, s, y, p, q, u, part, scal, rOut, dvec;
    int irOuter = 0;
    double* hpin = nullptr;  // pinned host scalars
    int crIter = 0;
    double rr0 = 0, rr = 0, rs0 = 0, rs = 0;
    // timing
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double msPhase[12] = {0};
    // per-launch timing of the particle-matrix SpMV (events on the scene stream)
    bool spmvTiming = false;
    std::vector<hipEvent_t> spmvEv;
    std::vector<char> spmvEvF32;  // per event pair: FP32-matrix launch
    int spmvUsed = 0;
    // the factored solve enqueues the next product's first pass before it waits for
    // the convergence scalar: facPref is that pass's operand (null: none pending),
    // facPrefEv the index of its timing start event in spmvEv (-1: none), rrEv the
    // event the host waits on
    const double* facPref = nullptr;
    bool facPrefF32 = false;  // the pass was the FP32 form's
    int facPrefEv = -1;
    hipEvent_t rrEv = nullptr;
    double* hrr = nullptr;  // coherent pinned host memory: rr as the last kRR reduction wrote it
    // The FP64 CR loop with its stop test on the device (crStop 1 or 2): the iteration
    // kernels return at once after the stop (DevParams::halt = ctl), the host enqueues
    // iteration k+1 before it waits for iteration k and reads hctl (int[kNCtl] inside
    // hrr's pinned block).  crStop 2 replays iterations 1, 2, ... from two captured
    // graphs, one per ping-pong parity of the V-cycle, captured anew every solve and
    // applied to the executable graphs by an update where the topology allows.
    // Measured without a profiler, the solve of a 64k-fluid scene takes 15.6 ms with
    // crStop 0, 15.1 ms with 1 and 15.8 ms with 2 (1M: 103.3, 103.9, 104.4 ms).
    int crStop = 1;  // 0: the host tests rr every iteration
    mph::host::DBuf<int> ctl;
    int* hctl = nullptr;
    hipEvent_t crEv[2] = {nullptr, nullptr};
    hipGraphExec_t crExec[2] = {nullptr, nullptr};
    long long crGraphLaunches = 0, crCaptures = 0, crInstantiations = 0;
    double crCaptureMs = 0;  // host time capturing and updating or instantiating
    int crGraphErr = 0;      // the last capture error (the solve then ran its iterations eagerly)
    // CONVERGENCE_CHECK (main.cpp:46, option ConvergenceCheck): the FP64 solve keeps
    // (rr, r.s) of every iteration that did not stop in convRec (the dot r.s adds one
    // reduction per iteration) and, after the loop, r.r and r.s of the true residual
    // r = b - A x with s = M^-1 r (main.cpp:4383-4397)
    int convCheck = 0;
    mph::host::DBuf<double> convRec;
    long long convRows = 0;  // rows of convRec written by the last solve
    double checkRR = 0, checkRS = 0;
    double spmvMs = 0, spmvMs32 = 0;
    long long spmvLaunches = 0, spmvLaunches32 = 0;
    int spmvTimingErr = 0;
    // run-time variant choice (SpMV load groups per precision, C Lambda C^T schedule,
    // V-cycle levels fused into one workgroup)
    mph::host::Autotune tuneSpmv64, tuneSpmv32, tuneCac, tuneFuse, tuneGal;
    int batches = 0;  // batches holding this scene (it cannot be destroyed meanwhile)
};


namespace mph {
namespace host {

bool mixed(const mph_scene& S);
bool factored(const mph_scene& S);
constexpr int kFacCellsSmall = 16384;
// rows of the convergence record (16 bytes each): iterations past this are not kept
constexpr long long kConvRowsMax = 1LL << 22;
bool fac_cells(const mph_scene& S);  // the FP64 second pass is the cell-staged one (facCells)
// whether the step assembles the 3x3 blocks: always for the assembled form; for the
// factored form only when the multigrid's level 0 cannot be built from the factored
// scalars (a level-0 grid under 3 cells along an active axis)
bool blocks_wanted(const mph_scene& S);
mph::FacOp fac_op(mph_scene& S, bool f32);  // the factored form's pointers (FP64 or FP32 scalars)
ASlices aslices(mph_scene& S);
int run_phase(mph_scene& S, int ph);
// calculateNeighbor in two halves around the host readback of the list totals:
// launch = counts, scans and the async copies; finish = read them, grow, fill
int phase_neighbor_launch(mph_scene& S);
int phase_neighbor_finish(mph_scene& S);
// the bookkeeping of one time step around the phases (main.cpp:386-474)
bool step_wanted(const mph_scene& S);
void step_begin(mph_scene& S);
bool step_vtk_due(const mph_scene& S);
void step_vtk_done(mph_scene& S);

}  // namespace host
}  // namespace mph
