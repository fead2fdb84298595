// mph_batch.cpp -- batched executor: many independent scenes advance one time step
// together on one device.
//
// The reference advances one scene per process.  Here a batch holds caller-owned
// scenes (mph_scene_create) and each mph_batch_step runs, for every member:
//   * the phases before the solve on the scene's own stream, interleaved across
//     scenes so small scenes overlap on the GPU, with ONE host synchronisation for
//     all neighbour-list totals of the batch;
//   * the CR solve of all members in lockstep: one launch per kernel covers every
//     scene (scene-table kernels of mph_batch.hip).  In plain FP64 each scene's stop
//     test runs on the device and the host, one iteration ahead, reads the stop flags
//     of the whole batch once per iteration; the mixed-precision solve reads the
//     residuals back instead.  A scene that has converged drops out of the launch
//     sets (its kernels of the one iteration queued after its stop return at once); multigrid levels of scenes with
//     different depths are scheduled level by level so every scene sees its own
//     V-cycle (preconditionWithMultiGrid, main.cpp:4122-4266) in its own order;
//   * diagnostics and the time advance on the scene's stream.
// Every scene runs the same arithmetic as mph_scene_step, so its results are
// bit-identical alone and in any batch.
#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "mph_batch.h"
#include "mph_scene_impl.h"

using namespace mph;
using namespace mph::host;
using namespace mph::batch;

struct mph_batch {
    int device = 0;
    hipStream_t st = nullptr;
    std::vector<mph_scene*> scenes;
    DBuf<BScene> tab;
    DBuf<int> off;
    DBuf<double> scal;        // [ns][8] CR scalars of the solve
    double* hscal = nullptr;  // pinned copy
    size_t hscalCap = 0;
    std::vector<BScene> htab;
    std::vector<int> hoff;
    Autotune tune64, tune32;  // SpMV load-group depth of the scene-table launches
    std::vector<char> fac;      // per scene of the current solve: factored operator
    std::vector<char> cell;     // ... and its FP64 second pass cell-staged
    Autotune tuneFuse;        // V-cycle levels fused into one workgroup per scene
    int* hctl = nullptr;      // [ns][kNCtl] coherent pinned mirrors of the scenes' CR control blocks
    size_t hctlCap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};  // the FP64 loop's last two iterations
};

namespace {

struct Set {
    SceneSet m{};
    void add(int s) { m.w[s >> 6] |= 1ull << (s & 63); }
    bool has(int s) const { return (m.w[s >> 6] >> (s & 63)) & 1ull; }
    bool empty() const {
        for (unsigned long long w : m.w)
            if (w) return false;
        return true;
    }
};

int ensure_pinned(mph_batch& B, size_t n) {
    if (B.hscal && B.hscalCap >= n) return MPH_OK;
    if (B.hscal) (void)hipHostFree(B.hscal);
    B.hscal = nullptr;
    B.hscalCap = 0;
    HIPCHK(hipHostMalloc((void**)&B.hscal, n * sizeof(double), hipHostMallocDefault));
    B.hscalCap = n;
    return MPH_OK;
}

int ensure_ctl(mph_batch& B, size_t n) {
    if (B.hctl && B.hctlCap >= n) return MPH_OK;
    if (B.hctl) (void)hipHostFree(B.hctl);
    B.hctl = nullptr;
    B.hctlCap = 0;
    HIPCHK(hipHostMalloc((void**)&B.hctl, n * sizeof(int), hipHostMallocCoherent));
    B.hctlCap = n;
    return MPH_OK;
}

// scene tables and block offsets of one solve over `sc` (after the multigrid build);
// stop: the FP64 solve with its stop test on the device (BScene::ctl)
int build_tables(mph_batch& B, const std::vector<mph_scene*>& sc, Tables& T, bool stop) {
    const int ns = (int)sc.size();
    int rc;
    if ((rc = B.tab.ensure(ns)) || (rc = B.off.ensure((size_t)kOffRows * (ns + 1))) ||
        (rc = B.scal.ensure((size_t)ns * 8)) || (rc = ensure_pinned(B, (size_t)ns * 8)))
        return rc;
    if (stop && (rc = ensure_ctl(B, (size_t)ns * kNCtl))) return rc;
    B.htab.assign(ns, BScene{});
    B.fac.assign(ns, 0);
    B.cell.assign(ns, 0);
    B.hoff.assign((size_t)kOffRows * (ns + 1), 0);
    for (int s = 0; s < ns; ++s) {
        mph_scene& S = *sc[s];
        if (S.depth > kMaxLevels) return fail(MPH_E_STATE, "multigrid depth above the batch limit");
        if (mixed(S)) {
            const long long n = 3LL * S.P.nf;
            if ((rc = S.rOut.ensure(n > 0 ? n : 1)) || (rc = S.dvec.ensure(n > 0 ? n : 1))) return rc;
        }
        BScene& b = B.htab[s];
        b.P = S.P;
        const ASlices A = aslices(S);
        b.nFluid = S.nFluid.p;
        b.start = A.start;
        b.col = A.col;
        b.val = A.val;
        b.val32 = A.val32;
        b.factored = factored(S) ? 1 : 0;
        B.fac[s] = (char)b.factored;
        if (b.factored) {
            b.fac[0] = fac_op(S, false);
            if (mixed(S)) b.fac[1] = fac_op(S, true);
        }
        b.cellStart = S.cellStart.p;
        b.cellIndex = S.cellIndex.p;
        b.invDiagA = S.invDiagA.p;
        b.force = S.force.p;
        b.mass = S.mass.p;
        b.vel = S.vel.p;
        double* v[kNVec] = {S.x.p, S.r.p, S.s.p, S.y.p, S.p.p, S.q.p, S.u.p, S.b.p, S.rOut.p, S.dvec.p};
        for (int k = 0; k < kNVec; ++k) b.vec[k] = v[k];
        b.part = S.part.p;
        b.scal = B.scal.p + 8 * (size_t)s;
        b.mg = S.prm.multigrid ? 1 : 0;
        b.depth = b.mg ? S.depth : 0;
        for (int v = 0; v < kFuseVariants; ++v) b.fuseOf[v] = b.mg ? S.fuseOf[v] : 0;
        for (int l = 0; l < b.depth; ++l) b.lv[l] = S.lv[l].L;
        if (stop) {
            if ((rc = S.ctl.ensure(kNCtl))) return rc;
            b.ctl = S.ctl.p;
            b.hctl = B.hctl + (size_t)kNCtl * s;
            b.tol = S.prm.cr_tolerance;
            b.maxit = S.prm.max_iterations > 0 ? S.prm.max_iterations : 3LL * S.P.nf;
        }
        auto put = [&](int row, int blocks) { B.hoff[(size_t)row * (ns + 1) + s + 1] = blocks; };
        // the cell-staged FP64 pass: one workgroup per cell of the fluid bounding box
        const FacOp& F = b.fac[0];
        const long long cells = (long long)F.cellCnt[0] * F.cellCnt[1] * F.cellCnt[2];
        B.cell[s] = b.factored && F.cellStart && cells > 0 ? 1 : 0;
        put(kOffCell, B.cell[s] ? (int)cells : 0);
        put(kOffChunk, (S.P.nf + 255) / 256);
        put(kOffElem, (int)((3LL * S.P.nf + 255) / 256));
        put(kOffPart, (S.P.np + 255) / 256);
        for (int l = 0; l < b.depth; ++l) put(kOffLevel + l, (b.lv[l].nrow + 255) / 256);
    }
    for (int row = 0; row < kOffRows; ++row)
        for (int s = 0; s < ns; ++s) B.hoff[(size_t)row * (ns + 1) + s + 1] += B.hoff[(size_t)row * (ns + 1) + s];
    HIPCHK(hipMemcpyAsync(B.tab.p, B.htab.data(), ns * sizeof(BScene), hipMemcpyHostToDevice, B.st));
    HIPCHK(hipMemcpyAsync(B.off.p, B.hoff.data(), B.hoff.size() * sizeof(int), hipMemcpyHostToDevice, B.st));
    T.st = B.st;
    T.tab = B.tab.p;
    T.off = B.off.p;
    T.ns = ns;
    T.hoff = B.hoff.data();
    return MPH_OK;
}

int read_scalars(mph_batch& B, int ns) {
    HIPCHK(hipMemcpyAsync(B.hscal, B.scal.p, (size_t)ns * 8 * sizeof(double), hipMemcpyDeviceToHost, B.st));
    HIPCHK(hipStreamSynchronize(B.st));
    B.tune64.collect();
    B.tune32.collect();
    B.tuneFuse.collect();
    return MPH_OK;
}

// the scene-table SpMV with the batch's tuned load-group depth
void bspmv(mph_batch& B, const Tables& T, const SceneSet& on, const SceneSet& first, bool f32, double alpha, int x,
           double beta, int yin, int y, int y2, int p_copy, int dot) {
    SceneSet onA = on, onF{}, cells{};  // assembled and factored members
    bool anyA = false, anyF = false;
    for (int s = 0; s < T.ns; ++s) {
        if (!((on.w[s >> 6] >> (s & 63)) & 1ull)) continue;
        if (B.fac[s]) {
            onA.w[s >> 6] &= ~(1ull << (s & 63));
            onF.w[s >> 6] |= 1ull << (s & 63);
            if (B.cell[s]) cells.w[s >> 6] |= 1ull << (s & 63);
            anyF = true;
        } else {
            anyA = true;
        }
    }
    if (anyA) {
        Autotune& A = f32 ? B.tune32 : B.tune64;
        const int v = A.pick();
        A.begin(B.st, v);
        spmv(T, onA, first, f32, alpha, x, beta, yin, y, y2, p_copy, dot, v);
        A.end(B.st);
    }
    if (anyF) spmv_fac(T, onF, first, cells, f32, alpha, x, beta, yin, y, y2, p_copy, dot);
}

// the V-cycle of every scene in `on`, level by level; cur[s][l] is the ping-pong
// state of scene s at level l and advances with every smoothing sweep
struct VCycle {
    mph_batch& B;
    const Tables& T;
    const std::vector<mph_scene*>& sc;
    std::vector<std::vector<int>> cur;
    VCycle(mph_batch& b, const Tables& t, const std::vector<mph_scene*>& s) : B(b), T(t), sc(s), cur(s.size()) {
        for (size_t k = 0; k < s.size(); ++k) cur[k].assign(kMaxLevels, 0);
    }
    SceneSet curset(const Set& on, int l) const {
        Set c;
        for (int s = 0; s < (int)sc.size(); ++s)
            if (on.has(s) && cur[s][l]) c.add(s);
        return c.m;
    }
    void ru(const Set& on, int l, bool f32) {
        if (on.empty()) return;
        mg_residual_update(T, on.m, l, curset(on, l), f32);
        for (int s = 0; s < (int)sc.size(); ++s)
            if (on.has(s)) cur[s][l] ^= 1;
    }
    // precondition (main.cpp:4122-4266) of in -> out for the multigrid scenes in `on`
    void apply(const Set& on, int in, int out, bool part, bool f32) {
        if (on.empty()) return;
        const int v = B.tuneFuse.pick();
        B.tuneFuse.begin(T.st, v);
        auto F = [&](int s) { return sc[s]->fuseOf[v]; };
        int maxD = 0;
        for (int s = 0; s < (int)sc.size(); ++s)
            if (on.has(s)) maxD = std::max(maxD, sc[s]->depth);
        mg_p2g_bu(T, on.m, in, curset(on, 0));  // P2G, S = 0, R = Q, S += invD R
        ru(on, 0, f32);                          // pre-smooth x2
        mg_residual(T, on.m, 0, curset(on, 0), f32);
        for (int l = 1; l <= maxD - 1; ++l) {  // levels below each scene's fused part
            Set inner, coarsest, both;
            for (int s = 0; s < (int)sc.size(); ++s) {
                if (!on.has(s) || l >= F(s)) continue;
                const int D = sc[s]->depth;
                if (D - 1 > l) inner.add(s), both.add(s);
                else if (D - 1 == l) coarsest.add(s), both.add(s);
            }
            if (both.empty()) continue;
            mg_restrict_bu(T, both.m, l, curset(both, l));
            ru(both, l, f32);
            ru(both, l, f32);
            if (!inner.empty()) mg_residual(T, inner.m, l, curset(inner, l), f32);  // pre-smooth x3
            ru(coarsest, l, f32);                                                  // coarsest: x4, then
            ru(coarsest, l, f32);                                                  // the prolongation
            if (!coarsest.empty()) mg_prolong(T, coarsest.m, l, curset(coarsest, l), curset(coarsest, l - 1));
        }
        Set fused;  // levels F..depth-1 of every scene in one launch
        CurBits cb{};
        for (int s = 0; s < (int)sc.size(); ++s)
            if (on.has(s) && F(s) < sc[s]->depth) {
                fused.add(s);
                for (int l = F(s) - 1; l < sc[s]->depth; ++l)
                    if (cur[s][l]) cb.lv[l].w[s >> 6] |= 1ull << (s & 63);
            }
        if (!fused.empty()) {
            mg_coarse(T, fused.m, cb, v, f32);
            for (int s = 0; s < (int)sc.size(); ++s)  // five sweeps per fused level, four at the coarsest
                if (fused.has(s))
                    for (int l = F(s); l < sc[s]->depth - 1; ++l) cur[s][l] ^= 1;
        }
        for (int l = maxD - 2; l >= 1; --l) {
            Set up;
            for (int s = 0; s < (int)sc.size(); ++s)
                if (on.has(s) && sc[s]->depth - 2 >= l && l < F(s)) up.add(s);
            if (up.empty()) continue;
            for (int k = 0; k < 3; ++k) ru(up, l, f32);  // post-smooth x3
            mg_prolong(T, up.m, l, curset(up, l), curset(up, l - 1));
        }
        ru(on, 0, f32);  // post-smooth x2
        ru(on, 0, f32);
        mg_g2p(T, on.m, curset(on, 0), in, out, part);
        B.tuneFuse.end(T.st);
    }
};

// apply_m of mph_scene.cpp for a set of scenes: the V-cycle, or a copy without it
void apply_m(VCycle& V, const Set& on, int in, int out, bool part, bool f32) {
    Set mg, plain;
    for (int s = 0; s < (int)V.sc.size(); ++s)
        if (on.has(s)) (V.sc[s]->prm.multigrid ? mg : plain).add(s);
    if (!plain.empty()) {
        copy(V.T, plain.m, in, out);
        if (part) dot_partial(V.T, plain.m, in, out);
    }
    V.apply(mg, in, out, part, f32);
}

void solve_prologue(const Tables& T, const std::vector<mph_scene*>& sc, const Set& all) {
    cr_init(T, all.m);
    for (int role : {kR, kS, kY, kP, kQ, kU}) fill(T, all.m, role, -1.0);
    int maxD = 0;
    for (mph_scene* S : sc) maxD = std::max(maxD, S->prm.multigrid ? S->depth : 0);
    for (int l = 0; l < maxD; ++l) {
        Set at;
        for (int s = 0; s < (int)sc.size(); ++s)
            if (all.has(s) && sc[s]->prm.multigrid && sc[s]->depth > l) at.add(s);
        if (!at.empty()) fill_level(T, at.m, l, -1.0);
    }
}

void finish_levels(const VCycle& V) {
    for (int s = 0; s < (int)V.sc.size(); ++s)
        for (int l = 0; l < V.sc[s]->depth && l < kMaxLevels; ++l) V.sc[s]->lv[l].L.cur = V.cur[s][l];
}

// A member whose residual is not finite (NaN or infinity in its system) would
// never meet the stop rule (rr/rr0 < tol is false for NaN) and would keep the whole
// batch iterating to its cap of 3*Nf (main.cpp:4344-4346): it leaves the solve at
// once and is flagged in `failed`.
bool nonfinite(double v) { return !std::isfinite(v); }

// one iteration of the CR loop body (4352-4378) for the scenes of `on`: y = A s
// (the scenes of `first` also q = y, p = s; the others update p and q), rho = s.y,
// u = M^-1 q, tmp = q.u, x, r and s updated, then r.r and the stop test on the device
void fp64_iteration(mph_batch& B, const Tables& T, VCycle& V, const Set& on, const Set& first) {
    Set rest;
    for (int s = 0; s < T.ns; ++s)
        if (on.has(s) && !first.has(s)) rest.add(s);
    bspmv(B, T, on.m, first.m, false, 1.0, kS, 0.0, kY, kY, kQ, kP, kS);
    dot_final(T, on.m, kRHO, kRHOP);
    pq_update(T, rest.m);
    apply_m(V, on, kQ, kU, true, false);
    dot_final(T, on.m, kTMP, -1);
    xrs_update(T, on.m, kX);
    dot_final_stop(T, on.m, 1);
}

// solveWithConjugatedGradient (4270-4421) for every scene of `sc` in lockstep, with
// each scene's stop test (main.cpp:4344-4346) on the device: the host enqueues
// iteration k+1 for the scenes it has not yet seen stop, then waits for iteration k
// and drops the scenes that stopped; their kernels in iteration k+1 return at once
int solve_fp64(mph_batch& B, const std::vector<mph_scene*>& sc, std::vector<char>& failed) {
    const int ns = (int)sc.size();
    Tables T{};
    int rc;
    if ((rc = build_tables(B, sc, T, true))) return rc;
    for (hipEvent_t& e : B.ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    Set all;
    for (int s = 0; s < ns; ++s) all.add(s);
    VCycle V(B, T, sc);
    ctl_reset(T, all.m);
    solve_prologue(T, sc, all);
    apply_m(V, all, kB, kS, true, false);  // s = M^-1 b; rs0 = b.s; rr0 = b.b
    dot_final(T, all.m, kRS0, -1);
    dot_partial(T, all.m, kB, kB);
    dot_final(T, all.m, kRR0, -1);
    bspmv(B, T, all.m, SceneSet{}, false, -1.0, kX, 1.0, kB, kR, -1, -1, -1);  // r = b - A x; s = M^-1 r
    apply_m(V, all, kR, kS, false, false);
    dot_partial(T, all.m, kR, kR);
    dot_final_stop(T, all.m, 0);
    if ((rc = read_scalars(B, ns))) return rc;
    failed.assign(ns, 0);
    const volatile int* hc = B.hctl;
    Set active;
    for (int s = 0; s < ns; ++s) {
        if (nonfinite(B.hscal[8 * s + kRR0]) || hc[kNCtl * s + kCtlBad]) failed[s] = 1;
        else if (!hc[kNCtl * s + kCtlHalt]) active.add(s);
    }
    if (!active.empty()) {
        // the trials of the autotuners are read as their iterations complete (mark:
        // how many each had recorded at an iteration's event)
        Autotune* tn[] = {&B.tune64, &B.tune32, &B.tuneFuse};
        constexpr int kTn = 3;
        size_t mark[2][kTn];
        fp64_iteration(B, T, V, active, active);
        HIPCHK(hipEventRecord(B.ev[0], B.st));
        for (int t = 0; t < kTn; ++t) mark[0][t] = tn[t]->pend.size();
        for (long long k = 1;; ++k) {
            fp64_iteration(B, T, V, active, Set{});
            HIPCHK(hipEventRecord(B.ev[k & 1], B.st));
            for (int t = 0; t < kTn; ++t) mark[k & 1][t] = tn[t]->pend.size();
            HIPCHK(hipEventSynchronize(B.ev[(k - 1) & 1]));
            for (int t = 0; t < kTn; ++t) {
                const size_t n = mark[(k - 1) & 1][t];
                tn[t]->collect(n);
                mark[k & 1][t] -= n;
            }
            Set now;
            for (int s = 0; s < ns; ++s)
                if (active.has(s) && !hc[kNCtl * s + kCtlHalt]) now.add(s);
            active = now;
            if (active.empty()) break;
        }
    }
    HIPCHK(hipStreamSynchronize(B.st));  // the mirrors are final
    std::vector<long long> iter(ns, 0);
    for (int s = 0; s < ns; ++s) {
        iter[s] = hc[kNCtl * s + kCtlIter];
        if (hc[kNCtl * s + kCtlBad]) failed[s] = 1;
    }
    ctl_release(T, all.m);
    Set ok;  // the members that did not fail
    for (int s = 0; s < ns; ++s)
        if (!failed[s]) ok.add(s);
    if (!ok.empty()) {
        dot_partial(T, ok.m, kR, kS);
        dot_final(T, ok.m, kRS, -1);
        store_velocity(T, ok.m);
    }
    if ((rc = read_scalars(B, ns))) return rc;
    for (int s = 0; s < ns; ++s) {
        mph_scene& S = *sc[s];
        S.crIter = (int)iter[s];
        S.rr0 = B.hscal[8 * s + kRR0];
        S.rs0 = B.hscal[8 * s + kRS0];
        S.rr = B.hscal[8 * s + kRR];
        S.rs = B.hscal[8 * s + kRS];
    }
    finish_levels(V);
    return MPH_OK;
}

// the mixed-precision solve of mph_scene.cpp (FP64 iterative refinement around the
// CR with FP32 matrices) for every scene of `sc`, each at its own point of the
// refinement: per tick a scene runs one inner iteration, or its outer update
int solve_mixed(mph_batch& B, const std::vector<mph_scene*>& sc, std::vector<char>& failed) {
    const int ns = (int)sc.size();
    Tables T{};
    int rc;
    if ((rc = build_tables(B, sc, T, false))) return rc;
    Set all;
    for (int s = 0; s < ns; ++s) all.add(s);
    VCycle V(B, T, sc);
    solve_prologue(T, sc, all);
    dot_partial(T, all.m, kB, kB);
    dot_final(T, all.m, kRR0, -1);
    bspmv(B, T, all.m, SceneSet{}, false, -1.0, kX, 1.0, kB, kROut, -1, -1, -1);
    dot_partial(T, all.m, kROut, kROut);
    dot_final(T, all.m, kRR, -1);
    if ((rc = read_scalars(B, ns))) return rc;
    enum { kCheck, kInner, kUpdate, kDone };
    struct St {
        int state = kCheck;
        long long total = 0, it = 0, maxit = 0;
        int outer = 0;
        double rr0 = 0, rrOut = 0, rri0 = 0, target = 0;
    };
    std::vector<St> z(ns);
    failed.assign(ns, 0);
    for (int s = 0; s < ns; ++s) {
        z[s].rr0 = B.hscal[8 * s + kRR0];
        const long long n = 3LL * sc[s]->P.nf;
        z[s].maxit = sc[s]->prm.max_iterations > 0 ? sc[s]->prm.max_iterations : n;
    }
    for (;;) {
        Set update, restart, inner, rest;
        for (int s = 0; s < ns; ++s) {
            St& q = z[s];
            const double tol = sc[s]->prm.cr_tolerance;
            if (q.state == kUpdate) {  // the outer update of the last tick has been read
                q.outer++;
                q.state = kCheck;
            }
            if (q.state == kCheck) {
                if (q.outer >= 50) {
                    q.state = kDone;
                    continue;
                }
                q.rrOut = B.hscal[8 * s + kRR];
                if (nonfinite(q.rr0) || nonfinite(q.rrOut)) {
                    failed[s] = 1;
                    q.state = kDone;
                    continue;
                }
                if (q.rrOut / q.rr0 < tol || q.total >= q.maxit) {
                    q.state = kDone;
                    continue;
                }
                q.target = tol * q.rr0 / q.rrOut;
                if (q.target < 1.0e-10) q.target = 1.0e-10;
                q.rri0 = q.rrOut;
                q.it = 0;
                q.state = kInner;
                restart.add(s);
                inner.add(s);
                continue;
            }
            if (q.state == kInner) {
                const double rri = B.hscal[8 * s + kRR];
                if (nonfinite(rri)) {
                    failed[s] = 1;
                    q.state = kDone;
                    continue;
                }
                if (rri / q.rri0 < q.target || q.total >= q.maxit) {
                    q.state = kUpdate;
                    update.add(s);
                } else {
                    inner.add(s);
                    rest.add(s);
                }
            }
        }
        if (update.empty() && inner.empty()) break;
        if (!update.empty()) {  // x += d; rOut = b - A64 x; rr = rOut.rOut
            axpy(T, update.m, kD, kX);
            bspmv(B, T, update.m, SceneSet{}, false, -1.0, kX, 1.0, kB, kROut, -1, -1, -1);
            dot_partial(T, update.m, kROut, kROut);
            dot_final(T, update.m, kRR, -1);
        }
        if (!restart.empty()) {  // d = 0; r = rOut; s = M32^-1 r
            fill(T, restart.m, kD, 0.0);
            copy(T, restart.m, kROut, kR);
            apply_m(V, restart, kR, kS, false, true);
        }
        if (!inner.empty()) {
            bspmv(B, T, inner.m, restart.m, true, 1.0, kS, 0.0, kY, kY, kQ, kP, kS);
            dot_final(T, inner.m, kRHO, kRHOP);
            pq_update(T, rest.m);
            apply_m(V, inner, kQ, kU, true, true);
            dot_final(T, inner.m, kTMP, -1);
            xrs_update(T, inner.m, kD);
            dot_final(T, inner.m, kRR, -1);
            for (int s = 0; s < ns; ++s)
                if (inner.has(s)) z[s].it++, z[s].total++;
        }
        if ((rc = read_scalars(B, ns))) return rc;
    }
    Set ok;  // the members that did not fail
    for (int s = 0; s < ns; ++s)
        if (!failed[s]) ok.add(s);
    if (!ok.empty()) store_velocity(T, ok.m);
    HIPCHK(hipStreamSynchronize(B.st));
    for (int s = 0; s < ns; ++s) {
        mph_scene& S = *sc[s];
        S.crIter = (int)z[s].total;
        S.irOuter = z[s].outer;
        S.rr0 = z[s].rr0;
        S.rr = z[s].rrOut;
        S.rs0 = 0;
        S.rs = 0;
    }
    finish_levels(V);
    return MPH_OK;
}

int sync_all(const std::vector<mph_scene*>& sc) {
    for (mph_scene* S : sc) HIPCHK(hipStreamSynchronize(S->st));
    return MPH_OK;
}

int check_launch() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MPH_E_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return MPH_OK;
}

}  // namespace

extern "C" {

int mph_batch_create(int device, mph_batch** out) {
    if (!out) return fail(MPH_E_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(MPH_E_ARG, "no HIP device " + std::to_string(device));
    HIPCHK(hipSetDevice(device));
    mph_batch* B = new mph_batch;
    B->device = device;
    B->tune64.init(kSpmvVariants64, 3);
    B->tune32.init(kSpmvVariants32, 3);
    B->tuneFuse.init(kFuseVariants, 3);
    if (hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return fail(MPH_E_DEVICE, "stream creation failed");
    }
    *out = B;
    return MPH_OK;
}

int mph_batch_add(mph_batch* B, mph_scene* S) {
    if (!B || !S) return fail(MPH_E_ARG, "null argument");
    if (S->device != B->device) return fail(MPH_E_ARG, "scene lives on another device than the batch");
    if (std::find(B->scenes.begin(), B->scenes.end(), S) != B->scenes.end())
        return fail(MPH_E_ARG, "scene is already in the batch");
    if ((int)B->scenes.size() >= kMaxScenes)
        return fail(MPH_E_ARG, "batch is full (" + std::to_string(kMaxScenes) + " scenes)");
    if (S->convCheck)
        return fail(MPH_E_STATE, "the convergence record is kept by standalone steps: turn ConvergenceCheck off "
                                 "before adding the scene to a batch");
    B->scenes.push_back(S);
    S->batches++;
    return MPH_OK;
}

int mph_batch_remove(mph_batch* B, mph_scene* S) {
    if (!B || !S) return fail(MPH_E_ARG, "null argument");
    auto it = std::find(B->scenes.begin(), B->scenes.end(), S);
    if (it == B->scenes.end()) return fail(MPH_E_ARG, "scene is not in the batch");
    B->scenes.erase(it);
    S->batches--;
    return MPH_OK;
}

int mph_batch_size(const mph_batch* B) { return B ? (int)B->scenes.size() : 0; }

int mph_batch_step(mph_batch* B, mph_step_stats* stats) {
    if (!B) return fail(MPH_E_ARG, "null batch");
    HIPCHK(hipSetDevice(B->device));
    const auto t0 = std::chrono::steady_clock::now();
    const int n = (int)B->scenes.size();
    std::vector<mph_scene*> act;
    std::vector<int> idx;
    for (int k = 0; k < n; ++k)
        if (step_wanted(*B->scenes[k])) {
            act.push_back(B->scenes[k]);
            idx.push_back(k);
        }
    std::vector<int> vtk(act.size(), 0);
    int rc = MPH_OK;
    for (mph_scene* S : act) step_begin(*S);
    for (int ph : {MPH_PHASE_KINEMATICS, MPH_PHASE_CELLS})
        for (mph_scene* S : act)
            if ((rc = run_phase(*S, ph))) return rc;
    for (mph_scene* S : act)
        if ((rc = phase_neighbor_launch(*S))) return rc;
    if ((rc = sync_all(act))) return rc;
    for (mph_scene* S : act)
        if ((rc = phase_neighbor_finish(*S))) return rc;
    if ((rc = check_launch())) return rc;
    if ((rc = sync_all(act))) return rc;
    const auto t1 = std::chrono::steady_clock::now();
    for (int ph = MPH_PHASE_EXPLICIT; ph <= MPH_PHASE_MULTIGRID; ++ph)
        for (mph_scene* S : act)
            if ((rc = run_phase(*S, ph))) return rc;
    if ((rc = sync_all(act))) return rc;
    const auto t2 = std::chrono::steady_clock::now();
    std::vector<mph_scene*> plain, mix;
    for (mph_scene* S : act) (mixed(*S) ? mix : plain).push_back(S);
    std::vector<char> badPlain, badMix;
    if (!plain.empty() && (rc = solve_fp64(*B, plain, badPlain))) return rc;
    if (!mix.empty() && (rc = solve_mixed(*B, mix, badMix))) return rc;
    if ((rc = check_launch())) return rc;
    HIPCHK(hipStreamSynchronize(B->st));
    const auto t3 = std::chrono::steady_clock::now();
    std::vector<char> bad(act.size(), 0);  // per active member: its solve failed
    std::string badNames;
    for (size_t k = 0; k < act.size(); ++k) {
        const std::vector<mph_scene*>& grp = mixed(*act[k]) ? mix : plain;
        const std::vector<char>& flags = mixed(*act[k]) ? badMix : badPlain;
        const size_t at = std::find(grp.begin(), grp.end(), act[k]) - grp.begin();
        if (at < flags.size() && flags[at]) {
            bad[k] = 1;
            badNames += (badNames.empty() ? "" : ", ") + std::to_string(idx[k]);
        }
    }
    for (size_t k = 0; k < act.size(); ++k) {
        mph_scene& S = *act[k];
        if (bad[k]) continue;  // not advanced
        if (step_vtk_due(S)) {
            if ((rc = run_phase(S, MPH_PHASE_DIAGNOSTICS))) return rc;
            step_vtk_done(S);
            vtk[k] = 1;
        }
        if ((rc = run_phase(S, MPH_PHASE_ADVANCE))) return rc;
    }
    if ((rc = sync_all(act))) return rc;
    const auto t4 = std::chrono::steady_clock::now();
    auto dms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const double ms = dms(t0, t4);
    if (stats) {
        for (int k = 0; k < n; ++k) {
            stats[k] = mph_step_stats{};
            stats[k].ms_phase[MPH_PHASE_NEIGHBOR] = dms(t0, t1);
            stats[k].ms_phase[MPH_PHASE_MULTIGRID] = dms(t1, t2);
            stats[k].ms_phase[MPH_PHASE_SOLVE] = dms(t2, t3);
            stats[k].ms_phase[MPH_PHASE_ADVANCE] = dms(t3, t4);
        }
        for (size_t k = 0; k < act.size(); ++k) {
            mph_step_stats& st = stats[idx[k]];
            const mph_scene& S = *act[k];
            if (bad[k]) {
                st.status = MPH_E_STATE;
                continue;
            }
            st.steps_done = 1;
            st.cr_iterations = S.crIter;
            st.cr_iterations_total = S.crIter;
            st.vtk_steps = vtk[k];
            st.rr0 = S.rr0;
            st.rr = S.rr;
            st.rs0 = S.rs0;
            st.rs = S.rs;
            st.ms_total = ms;
        }
    }
    if (!badNames.empty())
        return fail(MPH_E_STATE, "non-finite CR residual in batch member(s) " + badNames + " (not advanced)");
    return MPH_OK;
}

int mph_batch_destroy(mph_batch* B) {
    if (!B) return MPH_OK;
    (void)hipSetDevice(B->device);
    if (B->st) (void)hipStreamSynchronize(B->st);
    for (mph_scene* S : B->scenes) S->batches--;
    B->tab.release();
    B->off.release();
    B->scal.release();
    if (B->hscal) (void)hipHostFree(B->hscal);
    if (B->hctl) (void)hipHostFree(B->hctl);
    for (hipEvent_t& e : B->ev)
        if (e) (void)hipEventDestroy(e);
    B->tune64.release();
    B->tune32.release();
    B->tuneFuse.release();
    if (B->st) (void)hipStreamDestroy(B->st);
    delete B;
    return MPH_OK;
}

}  // extern "C"
