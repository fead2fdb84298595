// mph_scene.cpp -- host side of the HIP time stepper: scene state, phase drivers and
// the C ABI of include/mph_abi.h.
//
// One scene = one HIP stream and persistent device buffers that only grow (the
// reference mallocs and frees every CSR, multigrid and neighbour array each step).
// Per step the host reads back one small record (neighbour totals, A slots, fluid
// bounding box) and, inside the solve, one scalar per CR iteration (rr, for the
// stop rule rr/rr0 < tol of source/main.cpp:4346).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mph_scene_impl.h"

using namespace mph;
using namespace mph::host;

namespace mph {
namespace host {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// ---------------------------------------------------------------------------
// constants, as initializeWeight/Fluid/Wall/Domain (source/main.cpp:886-1162)

double weight_prefactor(double Sw, double h, int dim) { return dim == 2 ? 1.0 / Sw * 1.0 / (h * h) : 1.0 / Sw * 1.0 / (h * h * h); }

double lattice_sum(double l, double radius, int dim, double K, bool kernelA) {
    const int range = (int)std::ceil(radius / l);
    const int rz = dim == 2 ? 0 : range;
    double sum = 0.0;
    for (int ix = -range; ix <= range; ++ix)
        for (int iy = -range; iy <= range; ++iy)
            for (int iz = -rz; iz <= rz; ++iz) {
                if (ix == 0 && iy == 0 && iz == 0) continue;
                const double x = l * ((double)ix), y = l * ((double)iy), z = l * ((double)iz);
                const double r2 = x * x + y * y + z * z;
                if (r2 <= radius * radius) {
                    const double r = std::sqrt(r2);
                    sum += kernelA ? wa_k(K, r, radius) : wq_k(K, r, radius);
                }
            }
    return sum;
}

int init_constants(mph_scene& S, const mph_grid& g) {
    const mph_params& p = S.prm;
    DevParams& P = S.P;
    const int dim = p.dim;
    const double l = g.spacing;
    P.dim = dim;
    P.np = g.np;
    P.l = l;
    P.V = dim == 2 ? l * l : l * l * l;
    P.dt = p.dt;
    P.radA = p.radius_ratio_a * l;
    P.radG = p.radius_ratio_a * l;  // RadiusRatioG = RadiusRatioA (main.cpp:888)
    P.radP = p.radius_ratio_p * l;
    P.radV = p.radius_ratio_v * l;
    if (dim == 2) {
        S.Swa = 1.0 / 2.0 * 2.0 / 15.0 * M_PI / l / l;
        S.Swg = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l / l;
        S.Swp = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l / l;
        S.Swv = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l / l;
        P.R2g = 1.0 / 2.0 * 1.0 / 30.0 * M_PI * P.radG * P.radG / l / l / S.Swg;
    } else {
        S.Swa = 1.0 / 3.0 * 1.0 / 5.0 * M_PI / l / l / l;
        S.Swg = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l / l / l;
        S.Swp = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l / l / l;
        S.Swv = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l / l / l;
        P.R2g = 1.0 / 3.0 * 4.0 / 105.0 * M_PI * P.radG * P.radG / l / l / l / S.Swg;
    }
    P.Ka = weight_prefactor(S.Swa, P.radA, dim);
    P.Kg = weight_prefactor(S.Swg, P.radG, dim);
    P.Kp = weight_prefactor(S.Swp, P.radP, dim);
    P.Kv = weight_prefactor(S.Swv, P.radV, dim);
    P.invA = 1.0 / P.radA;
    P.m2G = -2.0 / P.radG;
    P.m2P = -2.0 / P.radP;
    P.m2V = -2.0 / P.radV;
    P.N0a = lattice_sum(l, P.radA, dim, P.Ka, true);
    P.N0p = lattice_sum(l, P.radP, dim, P.Kp, false);
    P.radA2 = P.radA * P.radA;
    P.radG2 = P.radG * P.radG;
    P.radP2 = P.radP * P.radP;
    P.radV2 = P.radV * P.radV;
    // initializeFluid (988-1021)
    double integN, integX;
    if (dim == 2) { P.cofK = 0.350778153; integN = 0.024679383; integX = 0.226126699; }
    else          { P.cofK = 0.326976006; integN = 0.021425779; integX = 0.233977488; }
    for (int t = 0; t < 4; ++t) {
        P.cofA[t] = p.surface_tension[t] / ((P.radG / l) * (integN + P.cofK * P.cofK * integX));
        P.ai[t] = P.cofA[t] * (P.cofK) * (P.cofK);
        P.density[t] = p.density[t];
        P.bulk[t] = p.bulk_modulus[t];
        P.bulkVisc[t] = p.bulk_viscosity[t];
        P.shear[t] = p.shear_viscosity[t];
        for (int u = 0; u < 4; ++u) P.ratio[t][u] = p.interaction_ratio[t][u];
    }
    P.visc = dim == 2 ? 8.0 : 10.0;
    for (int d = 0; d < 3; ++d) P.gravity[d] = p.gravity[d];
    // initializeWall (1023-1062)
    for (int w = 2; w < 4; ++w) {
        const double* om = p.wall_omega[w];
        double n[3] = {0.0, 0.0, 0.0};
        const double theta = std::fabs(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        if (theta != 0.0)
            for (int d = 0; d < 3; ++d) n[d] = om[d] / theta;
        const double q0 = n[0] * std::sin(theta * p.dt / 2.0), q1 = n[1] * std::sin(theta * p.dt / 2.0),
                     q2 = n[2] * std::sin(theta * p.dt / 2.0), q3 = std::cos(theta * p.dt / 2.0);
        double(&R)[3][3] = S.wk.R[w];
        R[0][0] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
        R[0][1] = 2.0 * (q0 * q1 - q2 * q3);
        R[0][2] = 2.0 * (q0 * q2 + q1 * q3);
        R[1][0] = 2.0 * (q0 * q1 + q2 * q3);
        R[1][1] = -q0 * q0 + q1 * q1 - q2 * q2 + q3 * q3;
        R[1][2] = 2.0 * (q1 * q2 - q0 * q3);
        R[2][0] = 2.0 * (q0 * q2 - q1 * q3);
        R[2][1] = 2.0 * (q1 * q2 + q0 * q3);
        R[2][2] = -q0 * q0 - q1 * q1 + q2 * q2 + q3 * q3;
        for (int d = 0; d < 3; ++d) {
            S.wk.omega[w][d] = om[d];
            S.wk.vel[w][d] = p.wall_velocity[w][d];
            S.wk.center[w][d] = p.wall_center[w][d];
        }
    }
    // initializeDomain (1064-1111)
    double m = 0.0;
    m = (P.radA > m) ? P.radA : m;
    m = (2.0 * P.radP > m) ? 2.0 * P.radP : m;
    m = (P.radV > m) ? P.radV : m;
    P.maxR = m;
    P.maxR2 = m * m;
    for (int d = 0; d < 3; ++d) {
        P.dmin[d] = g.domain_min[d];
        P.W[d] = g.domain_max[d] - g.domain_min[d];
        P.hW[d] = 0.5 * P.W[d];
    }
    double cc[3];
    cc[0] = std::floor((g.domain_max[0] - g.domain_min[0]) / m);
    cc[1] = std::floor((g.domain_max[1] - g.domain_min[1]) / m);
    cc[2] = dim == 2 ? 1 : std::floor((g.domain_max[2] - g.domain_min[2]) / m);
    for (int d = 0; d < 3; ++d) P.C[d] = (int)cc[d];
    for (int d = 0; d < (dim == 2 ? 2 : 3); ++d)
        if (P.C[d] < 3)
            return fail(MPH_E_STATE, "domain narrower than 3 neighbour cells along axis " + std::to_string(d) +
                                         " (the reference's 27-cell search would visit cells twice)");
    P.T = cc[0] * cc[1] * cc[2];
    for (int d = 0; d < 3; ++d) P.cw[d] = P.W[d] / P.C[d];
    P.range[0] = (int)std::ceil(P.maxR / P.cw[0]);
    P.range[1] = (int)std::ceil(P.maxR / P.cw[1]);
    P.range[2] = dim == 2 ? 0 : (int)std::ceil(P.maxR / P.cw[2]);
    for (int d = 0; d < 3; ++d)
        if (2 * P.range[d] + 1 > 3)
            return fail(MPH_E_STATE, "neighbour range exceeds one cell (assert of main.cpp:1402)");
    // PowerParticleCount (640-645)
    P.p2pow = 0;
    while ((P.np >> P.p2pow) != 0) ++P.p2pow;
    P.p2 = 1 << P.p2pow;
    return MPH_OK;
}

// ---------------------------------------------------------------------------
// phases

bool mixed(const mph_scene& S) { return S.prm.precision == MPH_PRECISION_MIXED; }
bool factored(const mph_scene& S) { return S.prm.operator_form == MPH_OPERATOR_FACTORED; }

bool fac_cells(const mph_scene& S) {
    return S.facCells == 1 || (S.facCells == 2 && (S.P.dim == 3 || S.P.nf <= kFacCellsSmall));
}

bool blocks_wanted(const mph_scene& S) {
    if (!factored(S)) return true;
    if (!S.prm.multigrid) return false;  // nothing but the solve would read them
    const int* bb = S.hbbox;           // max x,y,z then min x,y,z (cells of the fluid)
    for (int d = 0; d < (S.P.dim == 2 ? 2 : 3); ++d)
        if (bb[d] - bb[3 + d] + 1 < 3) return true;  // level-0 stencil narrower than 3 cells
    return false;
}

ASlices aslices(mph_scene& S) {
    ASlices A;
    A.nslices = S.nslices;
    A.width = S.width.p;
    A.start = S.sliceStart.p;
    A.col = S.blocks ? S.acol.p : nullptr;
    A.val = S.blocks ? S.aval.p : nullptr;
    A.val32 = S.blocks && mixed(S) && !factored(S) ? S.aval32.p : nullptr;
    return A;
}

int phase_kinematics(mph_scene& S) {
    S.assembled = false;
    launch_kinematics(S.st, S.P, S.wk, S.type.p, S.pos.p, S.vel.p, S.force.p);
    for (int w = 2; w < 4; ++w)  // WallCenter += WallVelocity*Dt (main.cpp:2123-2127)
        for (int d = 0; d < 3; ++d) S.wk.center[w][d] += S.wk.vel[w][d] * S.P.dt;
    return MPH_OK;
}

int phase_cells(mph_scene& S) {
    S.assembled = false;
    const DevParams& P = S.P;
    launch_hash(S.st, P, S.pos.p, S.type.p, S.keys.p);
    launch_bitonic(S.st, P, S.keys.p);
    launch_cell_start(S.st, P, S.keys.p, S.cellStart.p);
    launch_reorder(S.st, P, S.keys.p, S.type.p, S.perm.p, S.pos.p, S.vel.p, S.type2.p, S.perm2.p, S.pos2.p, S.vel2.p,
                   S.cellIndex.p);
    std::swap(S.type, S.type2);
    std::swap(S.perm, S.perm2);
    std::swap(S.pos, S.pos2);
    std::swap(S.vel, S.vel2);
    // fluid bounding box for the multigrid depth (main.cpp:2779-2802)
    const int init[6] = {0, 0, 0, P.C[0], P.C[1], P.C[2]};
    HIPCHK(hipMemcpyAsync(S.bbox.p, init, sizeof init, hipMemcpyHostToDevice, S.st));
    launch_fluid_bbox(S.st, P, S.cellStart.p, S.bbox.p);
    return MPH_OK;
}

// scratch of the one-scan neighbour search, per scene
constexpr size_t kSearchScratchMax = size_t(8) << 30;

// The one-scan search's scratch (rows of capM and capP entries per particle, from
// the last search's largest counts plus a margin), allocated if need be.  False
// when there is no count yet, the scratch would exceed kSearchScratchMax, or the
// allocation failed: the scratch is optional, the fill pass then builds the lists.
bool search_scratch(mph_scene& S, int& capM, int& capP) {
    if (S.hmaxN <= 0) return false;
    capM = (S.hmaxN + 16 + 3) & ~3;
    capP = (S.hmaxNP + 8 + 3) & ~3;
    const size_t np = (size_t)S.P.np;
    if (np * (size_t)(capM + capP) * sizeof(int) > kSearchScratchMax) return false;
    if (S.scrM.ensure(np * capM) == MPH_OK && S.scrP.ensure(np * capP) == MPH_OK && S.searchOvf.ensure(1) == MPH_OK)
        return true;
    // A failed hipMalloc leaves its error as the thread's last error, which the launch
    // check after the phase would report; clear it, and give back what was allocated
    // so that the rest of the step has the memory
    (void)hipGetLastError();
    S.scrM.release();
    S.scrP.release();
    g_err.clear();
    return false;
}

int phase_neighbor_launch(mph_scene& S) {
    const DevParams& P = S.P;
    launch_fill_i32(S.st, S.maxP.p, 4, 0);
    // one scan writes the lists into scratch when the last search's largest counts
    // (plus a margin) bound them; else the count pass is followed by the fill pass
    S.searchCapM = S.searchCapP = 0;
    int capM = 0, capP = 0;
    if (search_scratch(S, capM, capP)) {
        S.searchCapM = capM;
        S.searchCapP = capP;
        launch_fill_i32(S.st, S.searchOvf.p, 1, 0);
    }
    const bool scr = S.searchCapM > 0;
    launch_neighbor_count(S.st, P, S.pos.p, S.cellIndex.p, S.cellStart.p, S.nCount.p, S.nFluid.p, S.nCountP.p,
                          S.maxP.p, scr ? S.scrM.p : nullptr, S.searchCapM, scr ? S.scrP.p : nullptr,
                          S.searchCapP, S.runStart.p, scr ? S.searchOvf.p : nullptr);
    launch_scan_i64(S.st, S.nCount.p, S.nPtr.p, P.np, S.scanScratch.p, (int)S.scanScratch.cap);
    launch_scan_i64(S.st, S.nCountP.p, S.nPtrP.p, P.np, S.scanScratch.p, (int)S.scanScratch.cap);
    launch_slice_width(S.st, P.nf, S.nslices, S.nFluid.p, S.width.p);
    launch_scan_i64(S.st, S.width.p, S.sliceStart.p, S.nslices, S.scanScratch.p, (int)S.scanScratch.cap);
    long long* h = (long long*)S.hpin;
    HIPCHK(hipMemcpyAsync(&h[0], S.nPtr.p + P.np, 8, hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipMemcpyAsync(&h[1], S.nPtrP.p + P.np, 8, hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipMemcpyAsync(&h[2], S.sliceStart.p + S.nslices, 8, hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipMemcpyAsync(&h[3], S.bbox.p, 6 * sizeof(int), hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipMemcpyAsync(&h[6], S.maxP.p, 4 * sizeof(int), hipMemcpyDeviceToHost, S.st));
    if (scr) HIPCHK(hipMemcpyAsync(&h[10], S.searchOvf.p, sizeof(int), hipMemcpyDeviceToHost, S.st));
    return MPH_OK;
}

// the stream must have reached the copies of phase_neighbor_launch
int phase_neighbor_finish(mph_scene& S) {
    const DevParams& P = S.P;
    const long long* h = (const long long*)S.hpin;
    S.hmaxP = ((const int*)&h[6])[0];
    S.hmaxW = ((const int*)&h[6])[1];
    S.hmaxN = ((const int*)&h[6])[2];
    S.hmaxNP = ((const int*)&h[6])[3];
    const bool scanned = S.searchCapM > 0 && *(const int*)&h[10] == 0;  // every list fit its scratch
    S.E = h[0];
    S.EP = h[1];
    S.slots = h[2] * kSlice;
    std::memcpy(S.hbbox, &h[3], 6 * sizeof(int));
    S.blocks = blocks_wanted(S);
    S.fslotValid = false;  // the staged places follow the new lists (matrix-C phase)
    int rc;
    if ((rc = S.nInd.ensure(S.E > 0 ? S.E : 1)) || (rc = S.nIndP.ensure(S.EP > 0 ? S.EP : 1))) return rc;
    S.oneScan = scanned;
    if (scanned)
        launch_neighbor_compact(S.st, P, S.nCount.p, S.nCountP.p, S.nPtr.p, S.nPtrP.p, S.scrM.p, S.searchCapM,
                                S.scrP.p, S.searchCapP, S.nInd.p, S.nIndP.p);
    else
        launch_neighbor_fill(S.st, P, S.pos.p, S.cellIndex.p, S.cellStart.p, S.nCount.p, S.nPtr.p, S.nPtrP.p,
                             S.nInd.p, S.nIndP.p, S.runStart.p);
    return MPH_OK;
}

int phase_neighbor(mph_scene& S) {
    int rc;
    if ((rc = phase_neighbor_launch(S))) return rc;
    HIPCHK(hipStreamSynchronize(S.st));
    return phase_neighbor_finish(S);
}

int phase_explicit(mph_scene& S) {
    const DevParams& P = S.P;
    launch_densities(S.st, P, S.type.p, S.pos.p, S.nCount.p, S.nPtr.p, S.nInd.p, S.densA.p, S.gcen.p, S.volStrainP.p);
    launch_coefficients(S.st, P, S.type.p, S.densA.p, S.volStrainP.p, S.mass.p, S.kappa.p, S.lambda.p, S.mu.p,
                        S.pressA.p);
    launch_forces(S.st, P, S.type.p, S.pos.p, S.nCount.p, S.nPtr.p, S.nInd.p, S.pressA.p, S.gcen.p, S.mass.p,
                  S.force.p, S.vel.p);
    return MPH_OK;
}

int phase_matrix_a(mph_scene& S) {
    int rc;
    if (S.blocks &&
        ((rc = S.acol.ensure(S.slots > 0 ? S.slots : 1)) || (rc = S.aval.ensure(S.slots > 0 ? 9 * S.slots : 1))))
        return rc;
    if (factored(S)) {
        const long long ne = S.E > 0 ? S.E : 1;
        if ((rc = S.falpha.ensure(ne)) || (rc = S.fdiag.ensure(9 * (size_t)S.P.nf)) || (rc = S.fwrap.ensure(1)))
            return rc;
        launch_fill_i32(S.st, S.fwrap.p, 1, 0);
    }
    // the level-0 run sums, for a factored level 0 (the multigrid phase reads them)
    const bool runsums = factored(S) && !S.blocks && S.prm.multigrid;
    if (runsums && (rc = S.fblk.ensure((size_t)(S.P.nf > 0 ? S.P.nf : 1) * kBlkRuns * kBlkStride))) return rc;
    launch_matrix_a(S.st, S.P, S.type.p, S.pos.p, S.vel.p, S.mass.p, S.mu.p, S.nCount.p, S.nFluid.p, S.nPtr.p,
                    S.nInd.p, aslices(S), S.selfIdx.p, S.b.p, factored(S) ? S.falpha.p : nullptr,
                    factored(S) ? S.fdiag.p : nullptr, factored(S) ? S.fwrap.p : nullptr,
                    runsums ? S.runStart.p : nullptr, runsums ? S.fblk.p : nullptr);
    return MPH_OK;
}

int phase_matrix_c(mph_scene& S) {
    int rc;
    // D = C Lambda is read only by C Lambda C^T, which a scene without blocks skips
    if ((rc = S.cval.ensure(3 * (S.EP > 0 ? S.EP : 1))) ||
        (S.blocks && (rc = S.dval.ensure(3 * (S.EP > 0 ? S.EP : 1)))))
        return rc;
    launch_matrix_c(S.st, S.P, S.pos.p, S.nCountP.p, S.nPtrP.p, S.nIndP.p, S.lambda.p, S.cval.p,
                    S.blocks ? S.dval.p : nullptr);
    launch_vector_p(S.st, S.P, S.type.p, S.pos.p, S.vel.p, S.kappa.p, S.lambda.p, S.volStrainP.p, S.nCountP.p,
                    S.nPtrP.p, S.nIndP.p, S.vecP.p);
    if (factored(S)) {
        const long long ne = S.EP > 0 ? S.EP : 1;
        if ((rc = S.fbeta.ensure(ne)) || (rc = S.runP.ensure(S.P.np)) || (rc = S.selfC.ensure(3 * (size_t)S.P.nf)) ||
            (rc = S.frec.ensure(7 * (size_t)S.P.np)))
            return rc;
        // one spare entry each: a batch's unused slots read entry 0 of a list
        if (mixed(S) && ((rc = S.fv32.ensure(4 * (ne + 1))) || (rc = S.fu32.ensure(4 * (S.E + 1))) ||
                         (rc = S.frec32.ensure(4 * (size_t)S.P.np))))
            return rc;
        launch_fac_p(S.st, S.P, S.pos.p, S.nCountP.p, S.nPtrP.p, S.nIndP.p, S.fbeta.p, mixed(S) ? S.fv32.p : nullptr,
                     S.runP.p, S.selfC.p, S.frec.p);
        if (mixed(S))  // alpha was written by the matrix-A phase
            launch_fac_u(S.st, S.P, S.nFluid.p, S.nPtr.p, S.nInd.p, S.falpha.p, S.pos.p, S.fu32.p);
        // the staged places of the viscous entries for the cell-staged FP64 pass (a
        // scene alone and in a batch)
        if (fac_cells(S)) {
            if ((rc = S.fslot.ensure(S.E > 0 ? S.E : 1))) return rc;
            S.fslotValid = true;
            launch_fac_slots(S.st, S.P, fac_op(S, false));
        }
    }
    return MPH_OK;
}

int phase_multiply_c(mph_scene& S) {
    int rc;
    if (mixed(S) && !factored(S) && (rc = S.aval32.ensure(S.slots > 0 ? 9 * S.slots : 1))) return rc;
    launch_b_minus_cp(S.st, S.P, S.nCountP.p, S.nPtrP.p, S.nIndP.p, S.cval.p, S.vecP.p, S.b.p);
    if (!S.blocks) {  // factored form: no blocks to add C Lambda C^T to
        S.assembled = true;
        return MPH_OK;
    }
    // C Lambda C^T also writes the FP32 copy of A in mixed precision (aval32) unless
    // the solve applies the factored form
    if (S.hmaxP > kMaxP) {  // P lists too long for the LDS modes
        launch_cac(S.st, S.P, S.nFluid.p, S.nCountP.p, S.nPtrP.p, S.nIndP.p, S.cval.p, S.dval.p, aslices(S), 0);
    } else {
        const int v = S.tuneCac.pick();
        S.tuneCac.begin(S.st, v);
        launch_cac(S.st, S.P, S.nFluid.p, S.nCountP.p, S.nPtrP.p, S.nIndP.p, S.cval.p, S.dval.p, aslices(S),
                   kCacModeOf[v]);
        S.tuneCac.end(S.st);
    }
    S.assembled = true;
    return MPH_OK;
}

// calculateMultiGridDepth (2804-2850) on the bbox read back with the neighbour totals
int phase_multigrid(mph_scene& S) {
    if (!S.prm.multigrid) return MPH_OK;
    const int* bb = S.hbbox;  // max x,y,z then min x,y,z
    double mgc = 0.0;
    if ((bb[0] - bb[3]) > mgc) mgc = (bb[0] - bb[3] + 1);
    if ((bb[1] - bb[4]) > mgc) mgc = (bb[1] - bb[4] + 1);
    if ((bb[2] - bb[5]) > mgc) mgc = (bb[2] - bb[5] + 1);
    int depth = 0;
    while ((1 << depth) <= mgc) depth++;
    if (depth < 2)
        return fail(MPH_E_STATE, "multigrid depth < 2: all fluid in one cell (the reference V-cycle reads level -1)");
    if (depth > kMaxLevels) return fail(MPH_E_STATE, "multigrid depth above " + std::to_string(kMaxLevels));
    S.depth = depth;
    if ((int)S.lv.size() < depth) S.lv.resize(depth);
    int rc;
    for (int l = 0; l < depth; ++l) {
        HostLevel& H = S.lv[l];
        Level& L = H.L;
        for (int d = 0; d < 3; ++d) {
            L.mn[d] = bb[3 + d] >> l;
            L.mx[d] = (bb[d] >> l) + 1;
            L.cnt[d] = L.mx[d] - L.mn[d];
            L.cs[d] = 3 < L.cnt[d] ? 3 : L.cnt[d];
        }
        L.cells = L.cnt[0] * L.cnt[1] * L.cnt[2];
        L.nrow = 3 * L.cells;
        L.nshift = 3 * L.cs[0] * L.cs[1] * L.cs[2];
        if ((rc = H.A.ensure((size_t)L.nrow * L.nshift)) || (rc = H.invD.ensure(L.nrow)) ||
            (rc = H.S0.ensure(L.nrow)) || (rc = H.S1.ensure(L.nrow)) || (rc = H.R.ensure(L.nrow)) ||
            (rc = H.Q.ensure(L.nrow)))
            return rc;
        L.A = H.A.p;
        L.invD = H.invD.p;
        L.S[0] = H.S0.p;
        L.S[1] = H.S1.p;
        L.R = H.R.p;
        L.Q = H.Q.p;
        L.cur = 0;
        L.f32 = 0;
        L.lanes = L.nrow < S.laneRowsMax ? 1 : 0;
        L.A32 = nullptr;
        if (mixed(S) && l + 1 < depth) {  // the coarsest level stays FP64
            if ((rc = H.A32.ensure((size_t)L.nrow * L.nshift))) return rc;
            L.A32 = H.A32.p;
        }
    }
    // the small levels may run in one workgroup (launch_mg_coarse): first fused level
    // per row threshold
    for (int v = 0; v < kFuseVariants; ++v) {
        S.fuseOf[v] = depth;
        for (int l = 1; l < depth; ++l)
            if (S.lv[l].L.nrow <= kFuseRowsOf[v]) {
                S.fuseOf[v] = l;
                break;
            }
    }
    S.hlv.resize(depth);
    for (int l = 0; l < depth; ++l) S.hlv[l] = S.lv[l].L;
    if ((rc = S.lvDev.ensure(kMaxLevels))) return rc;
    HIPCHK(hipMemcpyAsync(S.lvDev.p, S.hlv.data(), depth * sizeof(Level), hipMemcpyHostToDevice, S.st));
    if (!S.blocks) {  // factored form: level 0 and InvDiagA from the factored scalars
        if ((rc = S.fgtab.ensure(81 * (size_t)S.P.np)) || (rc = S.fgmask.ensure(S.P.np))) return rc;
        launch_fac_level0(S.st, S.P, S.lv[0].L, fac_op(S, false), S.cellIndex.p, S.cellStart.p, S.fblk.p,
                          S.fgtab.p, S.fgmask.p, S.invDiagA.p);
    } else {
        launch_inv_diag(S.st, S.P, S.selfIdx.p, aslices(S), S.invDiagA.p);
        // the LDS variant holds one particle's row: its slice width must fit
        int v = S.tuneGal.pick();
        if (v == 1 && S.hmaxW > kGalRowMax) {
            launch_mg_galerkin0(S.st, S.P, S.lv[0].L, S.cellStart.p, S.runStart.p, S.nFluid.p, aslices(S), 0);
        } else {
            S.tuneGal.begin(S.st, v);
            launch_mg_galerkin0(S.st, S.P, S.lv[0].L, S.cellStart.p, S.runStart.p, S.nFluid.p, aslices(S), v);
            S.tuneGal.end(S.st);
        }
    }
    for (int l = 1; l < depth; ++l) launch_mg_galerkinL(S.st, S.lv[l - 1].L, S.lv[l].L);
    for (int l = 0; l < depth; ++l) launch_mg_invd(S.st, S.lv[l].L);
    if (mixed(S))
        for (int l = 0; l + 1 < depth; ++l)
            launch_convert_f32(S.st, S.lv[l].L.A, S.lv[l].L.A32, (long long)S.lv[l].L.nrow * S.lv[l].L.nshift);
    return MPH_OK;
}

// preconditionWithMultiGrid (4122-4266): out = D^-1 in + G2P V(P2G in); `part`
// receives the chunk partials of in.out when non-null.  pqy non-null: in is S.q, and
// the P2G forms the CR update q = beta q + pqy (4364-4366) as it reads q
void precondition(mph_scene& S, const double* in, double* out, double* part, const double* pqy = nullptr) {
    hipStream_t st = S.st;
    const int v = S.tuneFuse.pick();
    const int D = S.depth, F = S.fuseOf[v];  // levels F..D-1 run fused (F = D: none)
    S.tuneFuse.begin(st, v);
    Level& L0 = S.lv[0].L;
    if (pqy) launch_mg_p2g_bu_pq(st, S.P, L0, S.cellStart.p, S.q.p, pqy, S.scal.p);
    else launch_mg_p2g_bu(st, S.P, L0, S.cellStart.p, in);  // P2G, S = 0, R = Q, S += invD R
    launch_mg_residual_update(st, L0);                  // pre-smooth x2
    launch_mg_residual(st, L0);
    for (int l = 1; l < D - 1 && l < F; ++l) {
        Level& L = S.lv[l].L;
        launch_mg_restrict_bu(st, S.lv[l - 1].L, L);  // pre-smooth x3
        launch_mg_residual_update(st, L);
        launch_mg_residual_update(st, L);
        launch_mg_residual(st, L);
    }
    if (F < D) {
        unsigned cur = 0;
        for (int l = F - 1; l < D; ++l) cur |= (unsigned)S.lv[l].L.cur << l;
        launch_mg_coarse(st, S.lvDev.p, F, D, cur, S.lv[0].L.f32 != 0, S.P.halt);
        for (int l = F; l < D - 1; ++l) S.lv[l].L.cur ^= 1;  // five sweeps each; the coarsest four
    } else {
        Level& L = S.lv[D - 1].L;
        launch_mg_restrict_bu(st, S.lv[D - 2].L, L);  // S = invD R, then four sweeps
        for (int k = 0; k < 4; ++k) launch_mg_residual_update(st, L);
        launch_mg_prolong(st, L, S.lv[D - 2].L);
    }
    for (int l = (D - 2 < F - 1 ? D - 2 : F - 1); l > 0; --l) {
        Level& L = S.lv[l].L;
        for (int k = 0; k < 3; ++k) launch_mg_residual_update(st, L);  // post-smooth x3
        launch_mg_prolong(st, L, S.lv[l - 1].L);
    }
    for (int k = 0; k < 2; ++k) launch_mg_residual_update(st, L0);  // post-smooth x2
    launch_mg_g2p(st, S.P, L0, S.cellIndex.p, S.invDiagA.p, in, out, part);
    S.tuneFuse.end(st);
}

// without MULTIGRID_SOLVER the preconditioner is a copy (main.cpp:4331, 4341, 4369)
int apply_m(mph_scene& S, const double* in, double* out, double* part) {
    if (S.prm.multigrid) {
        precondition(S, in, out, part);
    } else {
        HIPCHK(hipMemcpyAsync(out, in, sizeof(double) * 3 * (size_t)S.P.nf, hipMemcpyDeviceToDevice, S.st));
        if (part) launch_dot_partial(S.st, S.P, in, out, part);
    }
    return MPH_OK;
}

// The rest of a CR iteration once rho = s.y is reduced (4362-4377): later iterations
// p = beta p + s and q = beta q + y; u = M^-1 q, tmp = q.u; then xv += alpha p,
// r -= alpha q, s -= alpha u with the r.r partials in S.part.  With the multigrid
// preconditioner the p and q updates are formed by the kernels that read those
// vectors next (the P2G of q; the xv update, which reads p): the same arithmetic
// with one launch and two vector passes fewer.
int cr_tail(mph_scene& S, bool first, double* xv) {
    const DevParams& P = S.P;
    hipStream_t st = S.st;
    const bool fuse = !first && S.prm.multigrid;
    int rc;
    if (!first && !fuse) launch_pq_update(st, P, S.scal.p, S.s.p, S.y.p, S.p.p, S.q.p);
    if (fuse) precondition(S, S.q.p, S.u.p, S.part.p, S.y.p);  // u = M^-1 q; tmp = q.u
    else if ((rc = apply_m(S, S.q.p, S.u.p, S.part.p))) return rc;
    launch_dot_final(st, P, S.part.p, S.scal.p, kTMP, -1);
    launch_xrs_update(st, P, S.scal.p, S.p.p, S.q.p, S.u.p, xv, S.r.p, S.s.p, S.part.p, fuse);
    return MPH_OK;
}

void collect_tuning(mph_scene& S) {  // the stream is idle
    S.tuneSpmv64.collect();
    S.tuneSpmv32.collect();
    S.tuneCac.collect();
    S.tuneFuse.collect();
    S.tuneGal.collect();
}

int read_scalar(mph_scene& S, int slot, double* v) {
    HIPCHK(hipMemcpyAsync(S.hpin + 8, S.scal.p + slot, 8, hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipStreamSynchronize(S.st));
    *v = S.hpin[8];
    collect_tuning(S);
    return MPH_OK;
}

void prefetch_fac_t(mph_scene& S, const double* x, bool f32);

// read_scalar for the CR loop of the factored form: the first pass of the next
// product on x is enqueued behind the copy and runs while the host waits for the
// copy alone (the 8-byte round trip would otherwise leave the GPU idle every
// iteration).  If the loop then stops, that pass only wrote the product's scratch t.
// The kRR reductions that precede it wrote rr straight into S.hrr (rr_host), so the
// host waits on an event behind that kernel and no copy is enqueued.
int read_scalar_prefetch(mph_scene& S, int slot, double* v, const double* x, bool f32) {
    if (!factored(S) || !S.hrr || slot != kRR) return read_scalar(S, slot, v);
    if (!S.rrEv) HIPCHK(hipEventCreateWithFlags(&S.rrEv, hipEventDisableTiming));
    HIPCHK(hipEventRecord(S.rrEv, S.st));
    prefetch_fac_t(S, x, f32);
    HIPCHK(hipEventSynchronize(S.rrEv));
    *v = S.hrr[0];
    collect_tuning(S);
    return MPH_OK;
}

// where the kRR reductions of a factored CR loop also put rr (null: copy it back)
double* rr_host(mph_scene& S) { return factored(S) ? S.hrr : nullptr; }

FacOp fac_op(mph_scene& S, bool f32) {
    FacOp F;
    F.nFluid = S.nFluid.p;
    F.nPtr = S.nPtr.p;
    F.nInd = S.nInd.p;
    F.alpha = S.falpha.p;
    F.diag = S.fdiag.p;
    F.selfC = S.selfC.p;
    F.nCountP = S.nCountP.p;
    F.runP = S.runP.p;
    F.nPtrP = S.nPtrP.p;
    F.nIndP = S.nIndP.p;
    F.beta = S.fbeta.p;
    F.lambda = S.lambda.p;
    F.pos = S.pos.p;
    F.rec = S.frec.p;
    F.wrap = S.fwrap.p;
    F.u32 = f32 ? S.fu32.p : nullptr;
    F.v32 = f32 ? S.fv32.p : nullptr;
    F.rec32 = f32 ? S.frec32.p : nullptr;
    // the cell-staged second pass over the fluid bounding box (max x,y,z then min)
    const bool cells = fac_cells(S);
    F.cellStart = cells ? S.cellStart.p : nullptr;
    F.slot = cells && S.fslot.p && S.fslotValid ? S.fslot.p : nullptr;
    F.cellXcd = S.facCellXcd ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
        F.cellLo[d] = S.hbbox[3 + d];
        F.cellCnt[d] = cells ? S.hbbox[d] - S.hbbox[3 + d] + 1 : 0;
    }
    return F;
}

// the particle-matrix SpMV, bracketed by events when per-launch timing is on
// room for one more SpMV event pair at S.spmvUsed
void spmv_event_room(mph_scene& S) {
    if ((int)S.spmvEv.size() < S.spmvUsed + 2) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreateWithFlags(&e0, hipEventDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&e1, hipEventDefault);
        if (e != hipSuccess) S.spmvTimingErr = (int)e;
        S.spmvEv.push_back(e0);
        S.spmvEv.push_back(e1);
        S.spmvEvF32.push_back(0);
    }
}

// the first pass of the next product on x, enqueued ahead (factored form)
void prefetch_fac_t(mph_scene& S, const double* x, bool f32) {
    if (!factored(S) || S.P.nf == 0) return;
    if (S.spmvTiming) {
        spmv_event_room(S);
        S.spmvEvF32[S.spmvUsed / 2] = f32 ? 1 : 0;
        hipError_t e = hipEventRecord(S.spmvEv[S.spmvUsed], S.st);
        if (e != hipSuccess) S.spmvTimingErr = (int)e;
        S.facPrefEv = S.spmvUsed;
    }
    launch_fac_spmv(S.st, S.P, fac_op(S, f32), 1.0, x, 0.0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, f32,
                    1);
    S.facPref = x;
    S.facPrefF32 = f32;
}

void spmv(mph_scene& S, double alpha, const double* x, double beta, const double* yin, double* y, double* y2,
          double* p_copy, const double* dotx, double* part, bool f32 = false) {
    // a first pass enqueued ahead for this operand: only the second pass is left
    const bool pref = factored(S) && x != nullptr && S.facPref == x && S.facPrefF32 == f32;
    if (S.spmvTiming) {
        spmv_event_room(S);
        S.spmvEvF32[S.spmvUsed / 2] = f32 ? 1 : 0;
        if (!pref || S.facPrefEv != S.spmvUsed) {  // else the pair starts at the first pass
            hipError_t e = hipEventRecord(S.spmvEv[S.spmvUsed], S.st);
            if (e != hipSuccess) S.spmvTimingErr = (int)e;
        }
    }
    S.facPref = nullptr;
    S.facPrefEv = -1;
    if (factored(S)) {
        launch_fac_spmv(S.st, S.P, fac_op(S, f32), alpha, x, beta, yin, y, y2, p_copy, dotx, part, f32,
                        pref ? 2 : 3);
    } else {
        Autotune& T = f32 ? S.tuneSpmv32 : S.tuneSpmv64;
        const int v = T.pick();
        T.begin(S.st, v);
        launch_spmv(S.st, S.P, aslices(S), S.nFluid.p, alpha, x, beta, yin, y, y2, p_copy, dotx, part, f32, v);
        T.end(S.st);
    }
    if (S.spmvTiming) {
        hipError_t e = hipEventRecord(S.spmvEv[S.spmvUsed + 1], S.st);
        if (e != hipSuccess) S.spmvTimingErr = (int)e;
        S.spmvUsed += 2;
    }
}

// collect the SpMV event pairs recorded since the last call (stream must be idle)
void collect_spmv_timing(mph_scene& S) {
    for (int k = 0; k + 1 < S.spmvUsed; k += 2) {
        float ms = 0;
        hipError_t e = hipEventSynchronize(S.spmvEv[k + 1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, S.spmvEv[k], S.spmvEv[k + 1]);
        if (e == hipSuccess) {
            if (S.spmvEvF32[k / 2]) {
                S.spmvMs32 += ms;
                S.spmvLaunches32++;
            } else {
                S.spmvMs += ms;
                S.spmvLaunches++;
            }
        } else {
            S.spmvTimingErr = (int)e;
        }
    }
    // a pending first pass keeps its start event: it becomes pair 0
    if (S.facPrefEv > 0) {
        std::swap(S.spmvEv[0], S.spmvEv[S.facPrefEv]);
        S.spmvEvF32[0] = S.spmvEvF32[S.facPrefEv / 2];
    }
    if (S.facPrefEv >= 0) S.facPrefEv = 0;
    S.spmvUsed = 0;
}

// A residual that is not finite (NaN or infinity in b or in the iterates) never
// meets the stop rule: rr/rr0 < tol is false for NaN, so the reference iterates to
// its cap of 3*Nf (main.cpp:4344-4346).  The solve fails at once instead; the scene
// state is then unspecified.
void set_halt(mph_scene& S, const int* halt);
int solve_nonfinite(mph_scene& S, const char* what, double v) {
    S.facPref = nullptr;  // a first pass enqueued ahead is never used
    S.facPrefEv = -1;
    set_halt(S, nullptr);
    (void)hipStreamSynchronize(S.st);
    if (S.spmvTiming) collect_spmv_timing(S);
    collect_tuning(S);
    char msg[160];
    std::snprintf(msg, sizeof msg, "non-finite CR residual (%s = %g) at step %d", what, v, S.iStep);
    return fail(MPH_E_STATE, msg);
}

// the CR iteration kernels test the stop flag halt (null: they run unconditionally)
void set_halt(mph_scene& S, const int* halt) {
    S.P.halt = halt;
    for (int l = 0; l < S.depth; ++l) S.lv[l].L.halt = halt;
}

// The loop body of solveWithConjugatedGradient (4352-4378): y = A s (the first
// iteration also p = s, q = y), rho = s.y, p and q updated (later iterations),
// u = M^-1 q, tmp = q.u, x, r and s updated, rr = r.r.  dev: rr goes to the
// device-side stop test.  With ConvergenceCheck the iteration first keeps rr and
// r.s (main.cpp:4348-4351) in row `it` of the record (dev: row ctl[kCtlIter]).
int cr_iteration(mph_scene& S, bool first, bool dev, long long maxit, long long it = 0) {
    const DevParams& P = S.P;
    hipStream_t st = S.st;
    int rc;
    if (S.convCheck) {
        launch_dot_partial(st, P, S.r.p, S.s.p, S.part.p);
        launch_dot_final_record(st, P, S.part.p, S.scal.p, dev ? S.ctl.p : nullptr, it, S.convRec.p,
                                (long long)(S.convRec.cap / 2));
    }
    if (first) {  // y = A s; p = s; q = A p (same sum); rho = s.y
        spmv(S, 1.0, S.s.p, 0.0, S.y.p, S.y.p, S.q.p, S.p.p, S.s.p, S.part.p);
        launch_dot_final(st, P, S.part.p, S.scal.p, kRHO, kRHOP);
    } else {
        spmv(S, 1.0, S.s.p, 0.0, S.y.p, S.y.p, nullptr, nullptr, S.s.p, S.part.p);
        launch_dot_final(st, P, S.part.p, S.scal.p, kRHO, kRHOP);
    }
    if ((rc = cr_tail(S, first, S.x.p))) return rc;
    if (dev) launch_dot_final_stop(st, P, S.part.p, S.scal.p, S.ctl.p, 1, S.prm.cr_tolerance, maxit, S.hctl, S.hrr);
    else launch_dot_final(st, P, S.part.p, S.scal.p, kRR, -1, rr_host(S));
    return MPH_OK;
}

// One later iteration (not the first) captured into the executable graph of slot k:
// an update of the existing one where HIP accepts it, else a new instantiation.
// Nothing runs; the host's ping-pong state advances as if it had.
hipError_t capture_iteration(mph_scene& S, int k, long long maxit) {
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    const int rc = cr_iteration(S, false, true, maxit);
    e = hipStreamEndCapture(S.st, &g);
    if (e == hipSuccess && rc) e = hipErrorUnknown;
    if (e == hipSuccess && S.crExec[k]) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
        if (hipGraphExecUpdate(S.crExec[k], g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(S.crExec[k]);
            S.crExec[k] = nullptr;
        }
    }
    if (e == hipSuccess && !S.crExec[k]) {
        e = hipGraphInstantiate(&S.crExec[k], g, nullptr, nullptr, 0);
        if (e != hipSuccess) S.crExec[k] = nullptr;
        else S.crInstantiations++;
    }
    if (g) (void)hipGraphDestroy(g);
    if (e == hipSuccess) S.crCaptures++;
    return e;
}

// the graphs hold fixed launches: no variant exploration or per-launch timing
bool graphs_allowed(const mph_scene& S) {
    return S.crStop >= 2 && !S.spmvTiming && (!S.prm.multigrid || !S.tuneFuse.exploring()) &&
           (factored(S) || !S.tuneSpmv64.exploring());
}

// The FP64 CR loop with the stop test on the device.  The prologue's stop test has
// completed (the stream is idle).  Iteration k+1 is enqueued before the host waits
// for iteration k; once the host reads the stop, the one iteration still queued
// returns at once in every kernel.  Iterations 1, 2, ... replay two captured graphs
// (the V-cycle's ping-pong parity alternates per iteration) when graphs_allowed.
int cr_loop_device(mph_scene& S, long long maxit) {
    const volatile int* hc = S.hctl;
    if (hc[kCtlHalt]) return MPH_OK;
    for (hipEvent_t& e : S.crEv)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int rc;
    if ((rc = cr_iteration(S, true, true, maxit))) return rc;
    HIPCHK(hipEventRecord(S.crEv[0], S.st));
    bool graphs = graphs_allowed(S);
    if (graphs) {  // captured while the GPU runs iteration 0
        std::vector<int> cur0(S.depth);
        for (int l = 0; l < S.depth; ++l) cur0[l] = S.lv[l].L.cur;
        const auto t0 = std::chrono::steady_clock::now();
        hipError_t e = capture_iteration(S, 0, maxit);
        if (e == hipSuccess) e = capture_iteration(S, 1, maxit);
        S.crCaptureMs += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (e != hipSuccess) {
            S.crGraphErr = (int)e;
            (void)hipGetLastError();
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(S.st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                hipGraph_t g = nullptr;  // end a capture left open by a failed launch
                (void)hipStreamEndCapture(S.st, &g);
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
            }
            // nothing captured ran: the eager iterations below continue from the
            // ping-pong state before the captures
            for (int l = 0; l < S.depth; ++l) S.lv[l].L.cur = cur0[l];
            graphs = false;
        }
    }
    // the variant trials of the eager iterations are read as their iterations complete
    // (mark: how many trials each autotuner had recorded at an iteration's event)
    Autotune* tn[] = {&S.tuneSpmv64, &S.tuneSpmv32, &S.tuneCac, &S.tuneFuse, &S.tuneGal};
    constexpr int kTn = sizeof(tn) / sizeof(tn[0]);
    size_t mark[2][kTn];
    for (int t = 0; t < kTn; ++t) mark[0][t] = tn[t]->pend.size();
    for (long long k = 1;; ++k) {
        if (graphs) {
            HIPCHK(hipGraphLaunch(S.crExec[(k - 1) & 1], S.st));
            S.crGraphLaunches++;
        } else if ((rc = cr_iteration(S, false, true, maxit))) {
            return rc;
        }
        HIPCHK(hipEventRecord(S.crEv[k & 1], S.st));
        for (int t = 0; t < kTn; ++t) mark[k & 1][t] = tn[t]->pend.size();
        HIPCHK(hipEventSynchronize(S.crEv[(k - 1) & 1]));
        for (int t = 0; t < kTn; ++t) {
            const size_t n = mark[(k - 1) & 1][t];
            tn[t]->collect(n);
            mark[k & 1][t] -= n;
        }
        if (hc[kCtlHalt]) break;
    }
    HIPCHK(hipStreamSynchronize(S.st));  // hctl final
    return MPH_OK;
}

// solveWithConjugatedGradient (4270-4421)
int phase_solve(mph_scene& S) {
    const DevParams& P = S.P;
    hipStream_t st = S.st;
    const long long n = 3LL * P.nf;
    int rc;
    const bool dev = S.crStop > 0;
    struct HaltOff {  // no kernel outside this solve tests the flag, on any return
        mph_scene& S;
        ~HaltOff() { set_halt(S, nullptr); }
    } haltOff{S};
    if (dev) {  // the control block starts at zero; every CR kernel tests its flag
        if ((rc = S.ctl.ensure(kNCtl))) return rc;
        launch_fill_i32(st, S.ctl.p, kNCtl, 0);
        set_halt(S, S.ctl.p);
    }
    launch_cr_init(st, P, S.vel.p, S.force.p, S.mass.p, S.x.p);
    for (double* v : {S.r.p, S.s.p, S.y.p, S.p.p, S.q.p, S.u.p}) launch_fill(st, v, n, -1.0);
    if (S.prm.multigrid)
        for (int l = 0; l < S.depth; ++l) {
            Level& L = S.lv[l].L;
            for (double* v : {L.S[0], L.S[1], L.R, L.Q}) launch_fill(st, v, L.nrow, -1.0);
            L.cur = 0;
        }
    // s = M^-1 b; rs0 = b.s; rr0 = b.b
    if ((rc = apply_m(S, S.b.p, S.s.p, S.part.p))) return rc;
    launch_dot_final(st, P, S.part.p, S.scal.p, kRS0, -1);
    launch_dot_partial(st, P, S.b.p, S.b.p, S.part.p);
    launch_dot_final(st, P, S.part.p, S.scal.p, kRR0, -1);
    // r = b - A x; s = M^-1 r
    spmv(S, -1.0, S.x.p, 1.0, S.b.p, S.r.p, nullptr, nullptr, nullptr, nullptr);
    if ((rc = apply_m(S, S.r.p, S.s.p, nullptr))) return rc;
    launch_dot_partial(st, P, S.r.p, S.r.p, S.part.p);
    const long long maxit = S.prm.max_iterations > 0 ? S.prm.max_iterations : n;
    S.convRows = 0;
    if (S.convCheck && (rc = S.convRec.ensure(2 * (size_t)std::min(std::max(maxit, 1LL), kConvRowsMax)))) return rc;
    if (dev) launch_dot_final_stop(st, P, S.part.p, S.scal.p, S.ctl.p, 0, S.prm.cr_tolerance, maxit, S.hctl, S.hrr);
    else launch_dot_final(st, P, S.part.p, S.scal.p, kRR, -1, rr_host(S));
    if ((rc = read_scalar(S, kRR0, &S.rr0))) return rc;
    if (!std::isfinite(S.rr0)) return solve_nonfinite(S, "b.b", S.rr0);
    long long iter;
    if (dev) {
        if ((rc = cr_loop_device(S, maxit))) return rc;
        set_halt(S, nullptr);  // the epilogue runs after the stop
        const volatile int* hc = S.hctl;
        iter = hc[kCtlIter];
        if (hc[kCtlBad]) return solve_nonfinite(S, "r.r", S.hrr[0]);
    } else {
        for (iter = 0; iter < maxit; ++iter) {
            if ((rc = read_scalar_prefetch(S, kRR, &S.rr, S.s.p, false))) return rc;  // both products below are on s
            if (S.spmvTiming) collect_spmv_timing(S);
            if (!std::isfinite(S.rr)) return solve_nonfinite(S, "r.r", S.rr);
            if (S.rr / S.rr0 < S.prm.cr_tolerance) break;
            if ((rc = cr_iteration(S, iter == 0, false, maxit, iter))) return rc;
        }
        S.facPref = nullptr;  // a first pass enqueued ahead of the stop is not used
        S.facPrefEv = -1;
    }
    launch_dot_partial(st, P, S.r.p, S.s.p, S.part.p);
    launch_dot_final(st, P, S.part.p, S.scal.p, kRS, -1);
    launch_store_velocity(st, P, S.x.p, S.vel.p);
    if (S.convCheck) {  // main.cpp:4383-4397: r = b - A x; s = M^-1 r; r.r and r.s
        spmv(S, -1.0, S.x.p, 1.0, S.b.p, S.r.p, nullptr, nullptr, nullptr, nullptr);
        if (S.prm.multigrid) {
            if ((rc = apply_m(S, S.r.p, S.s.p, nullptr))) return rc;
        } else {  // main.cpp:4391 copies b, not r, without the multigrid solver
            HIPCHK(hipMemcpyAsync(S.s.p, S.b.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        launch_dot_partial(st, P, S.r.p, S.r.p, S.part.p);
        launch_dot_final(st, P, S.part.p, S.scal.p, kCheckRR, -1);
        launch_dot_partial(st, P, S.r.p, S.s.p, S.part.p);
        launch_dot_final(st, P, S.part.p, S.scal.p, kCheckRS, -1);
    }
    HIPCHK(hipMemcpyAsync(S.hpin + 16, S.scal.p, 16 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (S.spmvTiming) collect_spmv_timing(S);
    collect_tuning(S);
    S.crIter = (int)iter;
    S.rr0 = S.hpin[16 + kRR0];
    S.rs0 = S.hpin[16 + kRS0];
    S.rr = S.hpin[16 + kRR];
    S.rs = S.hpin[16 + kRS];
    if (S.convCheck) {
        S.convRows = std::min(iter, (long long)(S.convRec.cap / 2));
        S.checkRR = S.hpin[16 + kCheckRR];
        S.checkRS = S.hpin[16 + kCheckRS];
    }
    return MPH_OK;
}

// FP32 stencils on every level but the coarsest, whose smoothing stays FP64
void set_levels_f32(mph_scene& S, int on) {
    for (int l = 0; l + 1 < S.depth; ++l) S.lv[l].L.f32 = on;
}

// Mixed precision: FP64 iterative refinement around the same preconditioned CR run
// with FP32-stored particle and multigrid matrices (the coarsest multigrid level,
// vectors, accumulation and dot products stay FP64).  The outer residual r = b - A x is always formed with the
// FP64 matrix and decides convergence: ||r||/||b|| below sqrt(cr_tolerance).
int phase_solve_mixed(mph_scene& S) {
    const DevParams& P = S.P;
    hipStream_t st = S.st;
    const long long n = 3LL * P.nf;
    int rc;
    if ((rc = S.rOut.ensure(n > 0 ? n : 1)) || (rc = S.dvec.ensure(n > 0 ? n : 1))) return rc;
    launch_cr_init(st, P, S.vel.p, S.force.p, S.mass.p, S.x.p);
    for (double* v : {S.r.p, S.s.p, S.y.p, S.p.p, S.q.p, S.u.p}) launch_fill(st, v, n, -1.0);
    if (S.prm.multigrid)
        for (int l = 0; l < S.depth; ++l) {
            Level& L = S.lv[l].L;
            for (double* v : {L.S[0], L.S[1], L.R, L.Q}) launch_fill(st, v, L.nrow, -1.0);
            L.cur = 0;
        }
    launch_dot_partial(st, P, S.b.p, S.b.p, S.part.p);
    launch_dot_final(st, P, S.part.p, S.scal.p, kRR0, -1);
    spmv(S, -1.0, S.x.p, 1.0, S.b.p, S.rOut.p, nullptr, nullptr, nullptr, nullptr);
    launch_dot_partial(st, P, S.rOut.p, S.rOut.p, S.part.p);
    launch_dot_final(st, P, S.part.p, S.scal.p, kRR, -1);
    if ((rc = read_scalar(S, kRR0, &S.rr0))) return rc;
    if (!std::isfinite(S.rr0)) return solve_nonfinite(S, "b.b", S.rr0);
    const long long maxit = S.prm.max_iterations > 0 ? S.prm.max_iterations : n;
    long long total = 0;
    int outer = 0;
    double rrOut = 0;
    for (; outer < 50; ++outer) {
        if ((rc = read_scalar(S, kRR, &rrOut))) return rc;
        if (S.spmvTiming) collect_spmv_timing(S);
        if (!std::isfinite(rrOut)) return solve_nonfinite(S, "outer r.r", rrOut);
        if (rrOut / S.rr0 < S.prm.cr_tolerance || total >= maxit) break;
        // inner solve A32 d = rOut from d = 0 down to the remaining reduction
        double target = S.prm.cr_tolerance * S.rr0 / rrOut;
        if (target < 1.0e-10) target = 1.0e-10;
        set_levels_f32(S, 1);
        launch_fill(st, S.dvec.p, n, 0.0);
        HIPCHK(hipMemcpyAsync(S.r.p, S.rOut.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        if ((rc = apply_m(S, S.r.p, S.s.p, nullptr))) return rc;
        const double rri0 = rrOut;
        double rri = rrOut;
        for (long long it = 0; total < maxit; ++it, ++total) {
            if (it > 0) {
                if ((rc = read_scalar_prefetch(S, kRR, &rri, S.s.p, true))) return rc;
                if (S.spmvTiming) collect_spmv_timing(S);
                if (!std::isfinite(rri)) {
                    set_levels_f32(S, 0);
                    return solve_nonfinite(S, "inner r.r", rri);
                }
            }
            if (rri / rri0 < target) break;
            if (it == 0) {
                spmv(S, 1.0, S.s.p, 0.0, S.y.p, S.y.p, S.q.p, S.p.p, S.s.p, S.part.p, true);
                launch_dot_final(st, P, S.part.p, S.scal.p, kRHO, kRHOP);
            } else {
                spmv(S, 1.0, S.s.p, 0.0, S.y.p, S.y.p, nullptr, nullptr, S.s.p, S.part.p, true);
                launch_dot_final(st, P, S.part.p, S.scal.p, kRHO, kRHOP);
            }
            if ((rc = cr_tail(S, it == 0, S.dvec.p))) return rc;
            launch_dot_final(st, P, S.part.p, S.scal.p, kRR, -1, rr_host(S));
        }
        S.facPref = nullptr;  // a first pass enqueued ahead of the inner stop is not used
        S.facPrefEv = -1;
        set_levels_f32(S, 0);
        // x += d; rOut = b - A64 x
        launch_axpy(st, n, S.dvec.p, S.x.p);
        spmv(S, -1.0, S.x.p, 1.0, S.b.p, S.rOut.p, nullptr, nullptr, nullptr, nullptr);
        launch_dot_partial(st, P, S.rOut.p, S.rOut.p, S.part.p);
        launch_dot_final(st, P, S.part.p, S.scal.p, kRR, -1);
    }
    launch_store_velocity(st, P, S.x.p, S.vel.p);
    HIPCHK(hipStreamSynchronize(st));
    if (S.spmvTiming) collect_spmv_timing(S);
    collect_tuning(S);
    S.crIter = (int)total;
    S.irOuter = outer;
    S.rr = rrOut;
    S.rs0 = 0;
    S.rs = 0;
    return MPH_OK;
}

int phase_diagnostics(mph_scene& S) {
    launch_diagnostics(S.st, S.P, S.type.p, S.pos.p, S.vel.p, S.mu.p, S.lambda.p, S.kappa.p, S.volStrainP.p,
                       S.pressA.p, S.gcen.p, S.nCount.p, S.nPtr.p, S.nInd.p, S.nCountP.p, S.nPtrP.p, S.nIndP.p,
                       S.divP.p, S.pressP.p, S.force.p, S.virialStress.p, S.virialP.p);
    return MPH_OK;
}

int run_phase(mph_scene& S, int ph) {
    if (S.timing) (void)hipEventRecord(S.ev[0], S.st);
    int rc;
    switch (ph) {
        case MPH_PHASE_KINEMATICS: rc = phase_kinematics(S); break;
        case MPH_PHASE_CELLS: rc = phase_cells(S); break;
        case MPH_PHASE_NEIGHBOR: rc = phase_neighbor(S); break;
        case MPH_PHASE_EXPLICIT: rc = phase_explicit(S); break;
        case MPH_PHASE_MATRIX_A: rc = phase_matrix_a(S); break;
        case MPH_PHASE_MATRIX_C: rc = phase_matrix_c(S); break;
        case MPH_PHASE_MULTIPLY_C: rc = phase_multiply_c(S); break;
        case MPH_PHASE_MULTIGRID: rc = phase_multigrid(S); break;
        case MPH_PHASE_SOLVE: rc = mixed(S) ? phase_solve_mixed(S) : phase_solve(S); break;
        case MPH_PHASE_DIAGNOSTICS: rc = phase_diagnostics(S); break;
        case MPH_PHASE_ADVANCE:
            S.time += S.P.dt;
            S.iStep++;
            rc = MPH_OK;
            break;
        default: return fail(MPH_E_ARG, "unknown phase " + std::to_string(ph));
    }
    if (rc) return rc;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MPH_E_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    if (S.timing) {
        (void)hipEventRecord(S.ev[1], S.st);
        HIPCHK(hipEventSynchronize(S.ev[1]));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, S.ev[0], S.ev[1]);
        if (ph >= 0 && ph < 12) S.msPhase[ph] += ms;
    }
    return MPH_OK;
}

bool step_wanted(const mph_scene& S) { return S.time < S.prm.end_time + 1.0e-5 * S.P.dt; }

void step_begin(mph_scene& S) {
    if (S.time + 1.0e-5 * S.P.dt >= S.outputNext) S.outputNext += S.prm.output_interval;
    S.vtkStep = false;
}

bool step_vtk_due(const mph_scene& S) { return S.time + 1.0e-5 * S.P.dt >= S.vtkOutputNext; }

void step_vtk_done(mph_scene& S) {
    S.vtkOutputNext += S.prm.vtk_output_interval;
    S.vtkStep = true;
}

int allocate(mph_scene& S) {
    const DevParams& P = S.P;
    const size_t np = P.np, nf = P.nf;
    int rc;
#define ENS(buf, n) \
    if ((rc = (buf).ensure((n) > 0 ? (n) : 1))) return rc
    ENS(S.type, np); ENS(S.type2, np); ENS(S.perm, np); ENS(S.perm2, np);
    ENS(S.pos, 3 * np); ENS(S.pos2, 3 * np); ENS(S.vel, 3 * np); ENS(S.vel2, 3 * np); ENS(S.force, 3 * np);
    ENS(S.mass, np); ENS(S.kappa, np); ENS(S.lambda, np); ENS(S.mu, np); ENS(S.densA, np); ENS(S.gcen, 3 * np);
    ENS(S.pressA, np); ENS(S.volStrainP, np); ENS(S.divP, np); ENS(S.pressP, np); ENS(S.virialP, np);
    ENS(S.virialStress, 9 * np);
    ENS(S.keys, (size_t)P.p2); ENS(S.cellIndex, np); ENS(S.cellStart, 2 * (size_t)P.T + 2); ENS(S.bbox, 8);
    ENS(S.nCount, np); ENS(S.nFluid, np); ENS(S.nCountP, np); ENS(S.nPtr, np + 1); ENS(S.nPtrP, np + 1);
    ENS(S.maxP, 4); ENS(S.runStart, nf * kRuns);
    ENS(S.scanScratch, (size_t)scan_scratch_len((int)np + 1) + 64);
    S.nslices = (int)((nf + kSlice - 1) / kSlice);
    ENS(S.width, (size_t)S.nslices); ENS(S.sliceStart, (size_t)S.nslices + 1); ENS(S.selfIdx, nf);
    ENS(S.b, 3 * nf); ENS(S.vecP, np); ENS(S.invDiagA, 3 * nf);
    ENS(S.x, 3 * nf); ENS(S.r, 3 * nf); ENS(S.s, 3 * nf); ENS(S.y, 3 * nf); ENS(S.p, 3 * nf); ENS(S.q, 3 * nf);
    ENS(S.u, 3 * nf); ENS(S.part, nf / kChunk + 2); ENS(S.scal, 16);
#undef ENS
    HIPCHK(hipHostMalloc((void**)&S.hpin, 64 * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&S.hrr, 8 * sizeof(double), hipHostMallocCoherent));
    static_assert(kNCtl * sizeof(int) <= 4 * sizeof(double), "the CR control mirror fits hrr[4..7]");
    S.hctl = (int*)(S.hrr + 4);
    for (int k = 0; k < kNCtl; ++k) S.hctl[k] = 0;
    // zero-fill the per-particle state the reference allocates uninitialised
    for (DBuf<double>* d : {&S.force, &S.mass, &S.kappa, &S.lambda, &S.mu, &S.densA, &S.gcen, &S.pressA,
                            &S.volStrainP, &S.divP, &S.pressP, &S.virialP, &S.virialStress, &S.b, &S.vecP})
        HIPCHK(hipMemsetAsync(d->p, 0, d->cap * sizeof(double), S.st));
    return MPH_OK;
}

void destroy(mph_scene* S) {
    if (!S) return;
    if (S->st) (void)hipStreamSynchronize(S->st);
    for (DBuf<int>* b : {&S->type, &S->type2, &S->perm, &S->perm2, &S->cellIndex, &S->cellStart, &S->bbox,
                         &S->nCount, &S->nFluid, &S->nCountP, &S->nInd, &S->nIndP, &S->width, &S->selfIdx, &S->acol,
                         &S->maxP, &S->runP, &S->fwrap, &S->fgmask, &S->scrM, &S->scrP, &S->searchOvf, &S->ctl})
        b->release();
    for (hipGraphExec_t& g : S->crExec)
        if (g) (void)hipGraphExecDestroy(g);
    for (hipEvent_t& e : S->crEv)
        if (e) (void)hipEventDestroy(e);
    S->fgtab.release();
    S->fblk.release();
    S->runStart.release();
    for (DBuf<double>* b : {&S->falpha, &S->fdiag, &S->selfC, &S->fbeta, &S->frec}) b->release();
    S->fu32.release();
    S->fslot.release();
    S->fv32.release();
    S->frec32.release();
    for (DBuf<double>* b : {&S->pos, &S->pos2, &S->vel, &S->vel2, &S->force, &S->mass, &S->kappa, &S->lambda,
                            &S->mu, &S->densA, &S->gcen, &S->pressA, &S->volStrainP, &S->divP, &S->pressP,
                            &S->virialP, &S->virialStress, &S->aval, &S->b, &S->cval, &S->dval, &S->vecP,
                            &S->invDiagA, &S->rOut, &S->dvec, &S->x,
                            &S->r, &S->s, &S->y, &S->p, &S->q, &S->u, &S->part, &S->scal, &S->convRec})
        b->release();
    for (DBuf<long long>* b : {&S->nPtr, &S->nPtrP, &S->scanScratch, &S->sliceStart}) b->release();
    S->keys.release();
    for (HostLevel& H : S->lv) {
        for (DBuf<double>* b : {&H.A, &H.invD, &H.S0, &H.S1, &H.R, &H.Q}) b->release();
        H.A32.release();
    }
    S->aval32.release();
    S->lvDev.release();
    S->tuneSpmv64.release();
    S->tuneSpmv32.release();
    S->tuneCac.release();
    S->tuneFuse.release();
    S->tuneGal.release();
    if (S->hpin) (void)hipHostFree(S->hpin);
    if (S->hrr) (void)hipHostFree(S->hrr);
    for (hipEvent_t& e : S->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : S->spmvEv) (void)hipEventDestroy(e);
    if (S->rrEv) (void)hipEventDestroy(S->rrEv);
    if (S->st) (void)hipStreamDestroy(S->st);
    delete S;
}

// ---------------------------------------------------------------------------
// field access (arrays in the reference's layouts, current sorted order)

template <class T>
int d2h(mph_scene& S, std::vector<T>& out, const T* src, size_t n) {
    out.resize(n);
    if (n) HIPCHK(hipMemcpyAsync(out.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, S.st));
    HIPCHK(hipStreamSynchronize(S.st));
    return MPH_OK;
}

struct Blob {
    std::vector<char> bytes;
    int esize = 0;
};

template <class T>
void put(Blob& b, const std::vector<T>& v) {
    b.esize = sizeof(T);
    b.bytes.resize(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(b.bytes.data(), v.data(), b.bytes.size());
}

struct DevField {
    void* p;
    size_t count;
    int esize;
};

DevField device_field(mph_scene& S, const std::string& n) {
    const size_t np = S.P.np, nf = S.P.nf;
    auto F = [](auto* p, size_t c) { return DevField{(void*)p, c, (int)sizeof(*p)}; };
    if (n == "Position") return F(S.pos.p, 3 * np);
    if (n == "Velocity") return F(S.vel.p, 3 * np);
    if (n == "Force") return F(S.force.p, 3 * np);
    if (n == "Property") return F(S.type.p, np);
    if (n == "ParticleIndex") return F(S.perm.p, np);
    if (n == "Mass") return F(S.mass.p, np);
    if (n == "Kappa") return F(S.kappa.p, np);
    if (n == "Lambda") return F(S.lambda.p, np);
    if (n == "Mu") return F(S.mu.p, np);
    if (n == "DensityA") return F(S.densA.p, np);
    if (n == "GravityCenter") return F(S.gcen.p, 3 * np);
    if (n == "PressureA") return F(S.pressA.p, np);
    if (n == "VolStrainP") return F(S.volStrainP.p, np);
    if (n == "DivergenceP") return F(S.divP.p, np);
    if (n == "PressureP") return F(S.pressP.p, np);
    if (n == "VirialPressureAtParticle") return F(S.virialP.p, np);
    if (n == "VirialStressAtParticle") return F(S.virialStress.p, 9 * np);
    if (n == "NeighborCount") return F(S.nCount.p, np);
    if (n == "NeighborFluidCount") return F(S.nFluid.p, np);
    if (n == "NeighborCountP") return F(S.nCountP.p, np);
    if (n == "NeighborPtr") return F(S.nPtr.p, np + 1);
    if (n == "NeighborPtrP") return F(S.nPtrP.p, np + 1);
    if (n == "NeighborInd") return F(S.nInd.p, (size_t)S.E);
    if (n == "NeighborIndP") return F(S.nIndP.p, (size_t)S.EP);
    if (n == "VectorB") return F(S.b.p, 3 * nf);
    if (n == "VectorP") return F(S.vecP.p, np);
    if (n == "InvDiagA") return F(S.invDiagA.p, S.prm.multigrid ? 3 * nf : 0);  // built by the multigrid phase only
    // (rr, r.s) per iteration of the last FP64 solve with ConvergenceCheck on
    if (n == "ConvergenceHistory") return F(S.convRec.p, S.convRec.p ? 2 * (size_t)S.convRows : 0);
    return DevField{nullptr, 0, 0};
}

// derived fields assembled on the host from the device layouts
int host_field(mph_scene& S, const std::string& n, Blob& out, bool& found) {
    found = true;
    const int np = S.P.np, nf = S.P.nf, T = S.P.T;
    int rc;
    if (n == "CellIndex" || n == "CellParticle") {
        std::vector<uint64_t> k;
        if ((rc = d2h(S, k, S.keys.p, (size_t)S.P.p2))) return rc;
        std::vector<int> v(k.size());
        for (size_t i = 0; i < k.size(); ++i) v[i] = n == "CellIndex" ? (int)(k[i] >> 32) : (int)(uint32_t)k[i];
        put(out, v);
        return MPH_OK;
    }
    if (n.rfind("Cell", 0) == 0 &&
        (n.find("ParticleBegin") != std::string::npos || n.find("ParticleEnd") != std::string::npos)) {
        std::vector<int> cs;
        if ((rc = d2h(S, cs, S.cellStart.p, 2 * (size_t)T + 1))) return rc;
        std::vector<int> v(T);
        const bool wall = n.find("Wall") != std::string::npos, end = n.find("End") != std::string::npos;
        for (int c = 0; c < T; ++c) v[c] = cs[(wall ? T : 0) + c + (end ? 1 : 0)];
        put(out, v);
        return MPH_OK;
    }
    if (n == "CsrPtrA" || n == "CsrIndA" || n == "CsrCofA" || n == "AVal" || n == "ACol") {
        if (!S.blocks) return fail(MPH_E_STATE, "the factored form assembled no 3x3 blocks this step");
        std::vector<int> nfl, col;
        std::vector<long long> start;
        std::vector<double> val;
        if ((rc = d2h(S, nfl, S.nFluid.p, nf)) || (rc = d2h(S, start, S.sliceStart.p, (size_t)S.nslices + 1)))
            return rc;
        if ((rc = d2h(S, col, S.acol.p, (size_t)S.slots)) || (rc = d2h(S, val, S.aval.p, 9 * (size_t)S.slots)))
            return rc;
        if (n == "AVal") { put(out, val); return MPH_OK; }
        if (n == "ACol") { put(out, col); return MPH_OK; }
        std::vector<long long> ptr(3 * (size_t)nf + 1, 0);
        for (int i = 0; i < nf; ++i)
            for (int r = 0; r < 3; ++r) ptr[3 * i + r + 1] = ptr[3 * i + r] + 3LL * nfl[i];
        if (n == "CsrPtrA") { put(out, ptr); return MPH_OK; }
        const size_t nnz = ptr.back();
        std::vector<int> ind(nnz);
        std::vector<double> cof(nnz);
        for (int i = 0; i < nf; ++i) {
            const int sl = i / kSlice, lane = i % kSlice;
            for (int k = 0; k < nfl[i]; ++k) {
                const long long slot = a_col_index(start[sl], k, lane);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        const long long at = ptr[3 * i + r] + 3LL * k + c;
                        ind[at] = 3 * col[slot] + c;
                        cof[at] = val[a_val_index(start[sl], k, 3 * r + c, lane)];
                    }
            }
        }
        if (n == "CsrIndA") put(out, ind);
        else put(out, cof);
        return MPH_OK;
    }
    if (n == "CsrPtrC" || n == "CsrIndC" || n == "CsrCofC") {
        std::vector<int> cnt, ind;
        std::vector<long long> ptrP;
        std::vector<double> cv;
        if ((rc = d2h(S, cnt, S.nCountP.p, np)) || (rc = d2h(S, ptrP, S.nPtrP.p, (size_t)np + 1)) ||
            (rc = d2h(S, ind, S.nIndP.p, (size_t)S.EP)) || (rc = d2h(S, cv, S.cval.p, 3 * (size_t)S.EP)))
            return rc;
        std::vector<long long> ptr(3 * (size_t)nf + 1, 0);
        for (int i = 0; i < nf; ++i)
            for (int r = 0; r < 3; ++r) ptr[3 * i + r + 1] = ptr[3 * i + r] + cnt[i];
        if (n == "CsrPtrC") { put(out, ptr); return MPH_OK; }
        std::vector<int> ci(ptr.back());
        std::vector<double> cc(ptr.back());
        for (int i = 0; i < nf; ++i)
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < cnt[i]; ++k) {
                    ci[ptr[3 * i + r] + k] = ind[ptrP[i] + k];
                    cc[ptr[3 * i + r] + k] = cv[3 * (ptrP[i] + k) + r];
                }
        if (n == "CsrIndC") put(out, ci);
        else put(out, cc);
        return MPH_OK;
    }
    if (n.rfind("MGCofA_L", 0) == 0 || n.rfind("MGInvD_L", 0) == 0) {
        const int l = atoi(n.c_str() + 8);
        if (l < 0 || l >= S.depth) return fail(MPH_E_FIELD, "no multigrid level " + std::to_string(l));
        const Level& L = S.lv[l].L;
        if (n[2] == 'I') {
            std::vector<double> v;
            if ((rc = d2h(S, v, L.invD, (size_t)L.nrow))) return rc;
            put(out, v);
            return MPH_OK;
        }
        std::vector<double> t;
        if ((rc = d2h(S, t, L.A, (size_t)L.nrow * L.nshift))) return rc;
        std::vector<double> v(t.size());
        for (int row = 0; row < L.nrow; ++row)
            for (int sh = 0; sh < L.nshift; ++sh) v[(size_t)row * L.nshift + sh] = t[(size_t)sh * L.nrow + row];
        put(out, v);
        return MPH_OK;
    }
    found = false;
    return MPH_OK;
}

}  // namespace host
}  // namespace mph

// ===========================================================================
// C ABI

// lets the host-only I/O translation unit report its failures through mph_last_error
void mph_internal_set_error(const char* msg) { g_err = msg ? msg : ""; }

// ---------------------------------------------------------------------------
// scratch of the allocation-free entry point (mph_scene_apply_matrix_noalloc): the
// records of the factored product (x and t, written by the product, and the
// positions, copied in from the scene) live in this file-scope device array
// instead of the scene's record buffer.  Each device holds its own copy from the
// library's load until the process ends.  In host code `scratch` only names the
// symbol: its device address comes from hipGetSymbolAddress, once per device with
// that device current, and only that address is passed to kernels.  Every call on
// one device shares the array: calls that overlap on a device race.
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(96) << 20;  // 56 B per particle in FP64: up to 1,797,558
constexpr int kNoAllocMaxDevices = 64;
__device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
std::mutex g_scratchMu;
double* g_scratchPtr[kNoAllocMaxDevices] = {};

// the device address of `scratch` on `device` (current); looked up on first use
int noalloc_scratch(int device, double** out) {
    if (device < 0 || device >= kNoAllocMaxDevices) return fail(MPH_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPH_E_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratchPtr[device] = (double*)p;
    }
    *out = g_scratchPtr[device];
    return MPH_OK;
}

// bytes of scratch one product of the scene needs: the FP64 records (7 doubles per
// particle) or the FP32 ones (4 floats); the assembled product needs none
size_t noalloc_scratch_bytes(const mph_scene& S, bool f32) {
    if (!factored(S)) return 0;
    return (size_t)S.P.np * (f32 ? 4 * sizeof(float) : 7 * sizeof(double));
}
}  // namespace

extern "C" {

const char* mph_last_error(void) { return g_err.c_str(); }

void mph_params_default(mph_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    p->dim = 3;
    p->dt = 1.0e100;
    p->cr_tolerance = 1.0e-12;
    p->multigrid = 1;
    p->precision = MPH_PRECISION_FP64;
    p->max_iterations = 0;
    p->operator_form = MPH_OPERATOR_ASSEMBLED;
}

int mph_scene_create(const mph_params* p, const mph_grid* g, int device, mph_scene** out) {
    if (!p || !g || !out) return fail(MPH_E_ARG, "null argument");
    *out = nullptr;
    if (p->dim != 2 && p->dim != 3) return fail(MPH_E_ARG, "dim must be 2 or 3");
    if (g->np <= 0 || !g->type || !g->pos || !g->vel) return fail(MPH_E_ARG, "empty particle set");
    if (p->precision != MPH_PRECISION_FP64 && p->precision != MPH_PRECISION_MIXED)
        return fail(MPH_E_ARG, "unknown precision mode");
    if (p->operator_form != MPH_OPERATOR_ASSEMBLED && p->operator_form != MPH_OPERATOR_FACTORED)
        return fail(MPH_E_ARG, "unknown operator form");
    int nf = 0;
    for (int i = 0; i < g->np; ++i) {
        if (g->type[i] < 0 || g->type[i] > 3) return fail(MPH_E_ARG, "particle type outside 0..3");
        nf += g->type[i] < 2;
    }
    if (nf == 0) return fail(MPH_E_STATE, "no fluid particle");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(MPH_E_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(MPH_E_ARG, "device index out of range");
    HIPCHK(hipSetDevice(device));
    mph_scene* S = new mph_scene();
    S->prm = *p;
    S->device = device;
    int rc = init_constants(*S, *g);
    if (rc) { delete S; return rc; }
    S->P.nf = nf;
    if (hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking) != hipSuccess) {
        delete S;
        return fail(MPH_E_DEVICE, "stream creation failed");
    }
    (void)hipEventCreate(&S->ev[0]);
    (void)hipEventCreate(&S->ev[1]);
    if ((rc = allocate(*S))) { destroy(S); return rc; }
    const size_t np = g->np;
    std::vector<int> perm(np);
    for (size_t i = 0; i < np; ++i) perm[i] = (int)i;
    if (hipMemcpyAsync(S->type.p, g->type, np * sizeof(int), hipMemcpyHostToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(S->perm.p, perm.data(), np * sizeof(int), hipMemcpyHostToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(S->pos.p, g->pos, 3 * np * sizeof(double), hipMemcpyHostToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(S->vel.p, g->vel, 3 * np * sizeof(double), hipMemcpyHostToDevice, S->st) != hipSuccess ||
        hipStreamSynchronize(S->st) != hipSuccess) {
        destroy(S);
        return fail(MPH_E_DEVICE, "initial copy failed");
    }
    S->time = g->time;
    for (int d = 0; d < 3; ++d) S->dmax[d] = g->domain_max[d];
    // init sequence, main.cpp:369-375 (the VTK of output.vtk is the caller's)
    if ((rc = run_phase(*S, MPH_PHASE_CELLS)) || (rc = run_phase(*S, MPH_PHASE_NEIGHBOR))) {
        destroy(S);
        return rc;
    }
    launch_densities(S->st, S->P, S->type.p, S->pos.p, S->nCount.p, S->nPtr.p, S->nInd.p, S->densA.p, S->gcen.p,
                     S->volStrainP.p);
    if (hipStreamSynchronize(S->st) != hipSuccess) {
        destroy(S);
        return fail(MPH_E_DEVICE, "init sequence failed");
    }
    // the first step's one-scan search sizes its scratch by these lists: allocate it
    // now, so that the scene's device bytes include it from the start
    int capM = 0, capP = 0;
    (void)search_scratch(*S, capM, capP);
    S->iStep = (int)(S->time / S->P.dt);
    S->tuneSpmv64.init(kSpmvVariants64, 3);
    S->tuneSpmv32.init(kSpmvVariants32, 3);
    S->tuneCac.init(kCacVariants, 1);
    S->tuneFuse.init(kFuseVariants, 3);
    S->tuneGal.init(kGalVariants, 1);
    S->outputNext = S->time;
    S->vtkOutputNext = S->time;
    *out = S;
    return MPH_OK;
}

int mph_scene_run_phase(mph_scene* S, int phase) {
    if (!S) return fail(MPH_E_ARG, "null scene");
    HIPCHK(hipSetDevice(S->device));
    int rc = run_phase(*S, phase);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(S->st));
    if (phase == MPH_PHASE_DIAGNOSTICS) S->vtkStep = true;
    return MPH_OK;
}

int mph_scene_step(mph_scene* S, int nsteps, mph_step_stats* st) {
    if (!S) return fail(MPH_E_ARG, "null scene");
    HIPCHK(hipSetDevice(S->device));
    mph_step_stats local{};
    for (int k = 0; k < 12; ++k) S->msPhase[k] = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, S->st);
    int rc = MPH_OK;
    for (int k = 0; k < nsteps && rc == MPH_OK; ++k) {
        if (!step_wanted(*S)) break;
        step_begin(*S);
        for (int ph = MPH_PHASE_KINEMATICS; ph <= MPH_PHASE_SOLVE && rc == MPH_OK; ++ph) rc = run_phase(*S, ph);
        if (rc) break;
        if (step_vtk_due(*S)) {
            if ((rc = run_phase(*S, MPH_PHASE_DIAGNOSTICS))) break;
            step_vtk_done(*S);
            local.vtk_steps++;
        }
        run_phase(*S, MPH_PHASE_ADVANCE);
        local.steps_done++;
        local.cr_iterations = S->crIter;
        local.cr_iterations_total += S->crIter;
    }
    (void)hipEventRecord(t1, S->st);
    (void)hipEventSynchronize(t1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    local.ms_total = ms;
    local.rr0 = S->rr0; local.rr = S->rr; local.rs0 = S->rs0; local.rs = S->rs;
    for (int k = 0; k < 12; ++k) local.ms_phase[k] = S->msPhase[k];
    if (st) *st = local;
    return rc;
}

int mph_scene_field_size(mph_scene* S, const char* field, size_t* bytes, int* esize) {
    if (!S || !field) return fail(MPH_E_ARG, "null argument");
    DevField f = device_field(*S, field);
    if (f.esize) {
        if (bytes) *bytes = f.count * f.esize;
        if (esize) *esize = f.esize;
        return MPH_OK;
    }
    Blob b;
    bool found;
    int rc = host_field(*S, field, b, found);
    if (rc) return rc;
    if (!found) return fail(MPH_E_FIELD, std::string("unknown field ") + field);
    if (bytes) *bytes = b.bytes.size();
    if (esize) *esize = b.esize;
    return MPH_OK;
}

int mph_scene_read(mph_scene* S, const char* field, void* host, size_t bytes) {
    if (!S || !field || (!host && bytes)) return fail(MPH_E_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    DevField f = device_field(*S, field);
    if (f.esize) {
        if (bytes != f.count * f.esize) return fail(MPH_E_ARG, std::string("size mismatch for ") + field);
        if (bytes) HIPCHK(hipMemcpyAsync(host, f.p, bytes, hipMemcpyDeviceToHost, S->st));
        HIPCHK(hipStreamSynchronize(S->st));
        return MPH_OK;
    }
    Blob b;
    bool found;
    int rc = host_field(*S, field, b, found);
    if (rc) return rc;
    if (!found) return fail(MPH_E_FIELD, std::string("unknown field ") + field);
    if (bytes != b.bytes.size()) return fail(MPH_E_ARG, std::string("size mismatch for ") + field);
    if (bytes) std::memcpy(host, b.bytes.data(), bytes);
    return MPH_OK;
}

int mph_scene_write(mph_scene* S, const char* field, const void* host, size_t bytes) {
    if (!S || !field || (!host && bytes)) return fail(MPH_E_ARG, "null argument");
    HIPCHK(hipSetDevice(S->device));
    const std::string n = field;
    DevField f = device_field(*S, n);
    if (!f.esize) return fail(MPH_E_FIELD, std::string("field is not writable: ") + field);
    if (n == "Property") {  // types decide the fluid count; they must not change it
        const int* t = (const int*)host;
        int nf = 0;
        for (size_t i = 0; i < bytes / sizeof(int); ++i) {
            if (t[i] < 0 || t[i] > 3) return fail(MPH_E_ARG, "particle type outside 0..3");
            nf += t[i] < 2;
        }
        if (nf != S->P.nf) return fail(MPH_E_ARG, "writing Property may not change the fluid count");
    }
    if (bytes != f.count * f.esize) return fail(MPH_E_ARG, std::string("size mismatch for ") + field);
    if (bytes) HIPCHK(hipMemcpyAsync(f.p, host, bytes, hipMemcpyHostToDevice, S->st));
    HIPCHK(hipStreamSynchronize(S->st));
    return MPH_OK;
}

int mph_scene_scalar(mph_scene* S, const char* name, double* v) {
    if (!S || !name || !v) return fail(MPH_E_ARG, "null argument");
    const std::string n = name;
    const DevParams& P = S->P;
    if (n == "Time") *v = S->time;
    else if (n == "Dt") *v = P.dt;
    else if (n == "iStep") *v = S->iStep;
    else if (n == "OutputNext") *v = S->outputNext;
    else if (n == "VtkOutputNext") *v = S->vtkOutputNext;
    else if (n == "VtkStep") *v = S->vtkStep;
    else if (n == "ParticleCount") *v = P.np;
    else if (n == "FluidCount") *v = P.nf;
    else if (n == "PowerParticleCount") *v = P.p2;
    else if (n == "TotalCellCount") *v = P.T;
    else if (n == "CellCount0") *v = P.C[0];
    else if (n == "CellCount1") *v = P.C[1];
    else if (n == "CellCount2") *v = P.C[2];
    else if (n == "N0a") *v = P.N0a;
    else if (n == "N0p") *v = P.N0p;
    else if (n == "Swa") *v = S->Swa;
    else if (n == "Swg") *v = S->Swg;
    else if (n == "Swp") *v = S->Swp;
    else if (n == "Swv") *v = S->Swv;
    else if (n == "R2g") *v = P.R2g;
    else if (n == "MaxRadius") *v = P.maxR;
    else if (n == "NeighborIndCount") *v = (double)S->E;
    else if (n == "NeighborIndCountP") *v = (double)S->EP;
    else if (n == "ASlots") *v = (double)S->slots;
    else if (n == "AssembledBlocks") *v = S->blocks ? 1.0 : 0.0;
    else if (n == "NeighborOneScan") *v = S->oneScan ? 1.0 : 0.0;
    else if (n == "MultiGridDepth") *v = S->depth;
    else if (n == "RefinementSteps") *v = S->irOuter;
    else if (n == "SpmvMs") *v = S->spmvMs;
    else if (n == "SpmvLaunches") *v = (double)S->spmvLaunches;
    else if (n == "SpmvMsF32") *v = S->spmvMs32;
    else if (n == "TunedSpmv64") *v = S->tuneSpmv64.best;
    else if (n == "TunedSpmv32") *v = S->tuneSpmv32.best;
    else if (n == "TunedCac") *v = S->tuneCac.best;
    else if (n == "TunedFuse") *v = S->tuneFuse.best;
    else if (n == "TunedGalerkin") *v = S->tuneGal.best;
    else if (n == "FactoredCellsUsed") *v = factored(*S) && fac_cells(*S) ? 1.0 : 0.0;
    else if (n == "FactoredBytes" || n == "FactoredBytesF32") {
        // algorithmic bytes of one factored product, every array read or written once:
        // per list entry a column and the stored value (FP64: alpha or beta, 8 B;
        // FP32: u or v, 12 B) -- first pass: P entries up to the last fluid one,
        // second pass: all entries of fluid rows; per particle the counts, lambda,
        // the diagonal block, C_ii, the record written (FP64: x and t, 32 B, and the
        // position read; FP32: 16 B), and the CR vectors as the assembled product (x
        // read, y read and written: 72 B per fluid particle).  Gathers of neighbours'
        // records are cache reads (in the cell-staged FP64 pass, LDS reads of records
        // staged per row cell).  That pass reads a 2-byte staged place per viscous entry
        // in place of the 4-byte column.
        if (!factored(*S)) return fail(MPH_E_STATE, "the scene applies the assembled operator");
        const bool f64 = n == "FactoredBytes";
        const double ab = f64 ? 8.0 : 12.0, vb = f64 ? 8.0 : 12.0, recb = f64 ? 32.0 : 16.0;
        // the cell-staged FP64 pass reads the 2-byte place instead of the column
        const double colb = f64 && fac_cells(*S) && S->fslotValid ? 2.0 : 4.0;
        const double posb = f64 ? 24.0 : 0.0;
        std::vector<int> nfl, ncp, run;
        int rc;
        if ((rc = d2h(*S, nfl, S->nFluid.p, (size_t)P.nf)) || (rc = d2h(*S, ncp, S->nCountP.p, (size_t)P.np)) ||
            (rc = d2h(*S, run, S->runP.p, (size_t)P.np)))
            return rc;
        double t = 0;
        for (int k = 0; k < P.np; ++k) {
            t += 4.0 + recb;  // run count, record written
            if (run[k] > 0) t += (4.0 + vb) * run[k] + 8.0 + posb;  // entries, lambda, position
        }
        for (int i = 0; i < P.nf; ++i)
            t += (colb + ab) * nfl[i] + (4.0 + vb) * ncp[i] + 4.0 + 4.0 + 72.0 + 24.0 + 72.0;
        *v = t;
    }
    else if (n == "SpmvLaunchesF32") *v = (double)S->spmvLaunches32;
    else if (n == "SpmvTimingError") *v = S->spmvTimingErr;
    else if (n == "ABlocks" || n == "CBlocks") {
        std::vector<int> c;
        int rc = d2h(*S, c, n[0] == 'A' ? S->nFluid.p : S->nCountP.p, (size_t)P.nf);
        if (rc) return rc;
        double t = 0;
        for (int k : c) t += k;
        *v = t;
    }
    else if (n == "CrGraphLaunches") *v = (double)S->crGraphLaunches;
    else if (n == "CrCaptures") *v = (double)S->crCaptures;
    else if (n == "CrInstantiations") *v = (double)S->crInstantiations;
    else if (n == "CrCaptureMs") *v = S->crCaptureMs;
    else if (n == "CrGraphError") *v = S->crGraphErr;
    else if (n == "CRIterations") *v = S->crIter;
    else if (n == "CheckRR") *v = S->checkRR;
    else if (n == "CheckRS") *v = S->checkRS;
    else if (n == "rr0") *v = S->rr0;
    else if (n == "rr") *v = S->rr;
    else if (n == "rs0") *v = S->rs0;
    else if (n == "rs") *v = S->rs;
    else return fail(MPH_E_FIELD, "unknown scalar " + n);
    return MPH_OK;
}

int mph_scene_domain(mph_scene* S, double* spacing, double dmin[3], double dmax[3]) {
    if (!S || !spacing || !dmin || !dmax) return fail(MPH_E_ARG, "null argument");
    *spacing = S->P.l;
    for (int d = 0; d < 3; ++d) {
        dmin[d] = S->P.dmin[d];
        dmax[d] = S->dmax[d];
    }
    return MPH_OK;
}

int mph_scene_set_timing(mph_scene* S, int on) {
    if (!S) return fail(MPH_E_ARG, "null scene");
    S->timing = (on & 1) != 0;
    S->spmvTiming = (on & 2) != 0;
    S->spmvMs = 0;
    S->spmvLaunches = 0;
    return MPH_OK;
}

int mph_scene_set_option(mph_scene* S, const char* name, double value) {
    if (!S || !name) return fail(MPH_E_ARG, "null argument");
    const std::string n = name;
    if (n == "LaneRowsMax") {
        if (!(value >= 0.0 && value <= 2147483647.0)) return fail(MPH_E_ARG, "LaneRowsMax out of range");
        S->laneRowsMax = (int)value;
        return MPH_OK;
    }
    if (n == "FactoredCells") {
        if (value != 0.0 && value != 1.0 && value != 2.0) return fail(MPH_E_ARG, "FactoredCells is 0, 1 or 2");
        S->facCells = (int)value;
        return MPH_OK;
    }
    if (n == "FactoredCellsXcd") {
        if (value != 0.0 && value != 1.0) return fail(MPH_E_ARG, "FactoredCellsXcd is 0 or 1");
        S->facCellXcd = value != 0.0;
        return MPH_OK;
    }
    if (n == "CrStop") {
        if (value != 0.0 && value != 1.0 && value != 2.0) return fail(MPH_E_ARG, "CrStop is 0, 1 or 2");
        S->crStop = (int)value;
        return MPH_OK;
    }
    if (n == "ConvergenceCheck") {
        if (value != 0.0 && value != 1.0) return fail(MPH_E_ARG, "ConvergenceCheck is 0 or 1");
        if (value != 0.0 && mixed(*S))
            return fail(MPH_E_STATE, "the convergence record follows the FP64 CR loop; the scene solves in mixed "
                                     "precision");
        if (value != 0.0 && S->batches > 0)
            return fail(MPH_E_STATE, "the convergence record is kept by standalone steps; the scene is in a batch");
        S->convCheck = (int)value;
        if (!S->convCheck) S->convRows = 0;
        return MPH_OK;
    }
    return fail(MPH_E_FIELD, "unknown option " + n);
}

int mph_scene_apply_matrix(mph_scene* S, int precision, const double* x, double* y) {
    if (!S || !x || !y) return fail(MPH_E_ARG, "null argument");
    if (precision != MPH_PRECISION_FP64 && precision != MPH_PRECISION_MIXED)
        return fail(MPH_E_ARG, "unknown precision mode");
    if (!S->assembled) return fail(MPH_E_STATE, "no particle matrix for the current particles");
    const bool f32 = precision == MPH_PRECISION_MIXED;
    if (f32 && !mixed(*S)) return fail(MPH_E_STATE, "FP32 values exist in mixed-precision scenes only");
    HIPCHK(hipSetDevice(S->device));
    const size_t n = 3 * (size_t)S->P.nf;
    HIPCHK(hipMemcpyAsync(S->s.p, x, n * sizeof(double), hipMemcpyHostToDevice, S->st));
    launch_fill(S->st, S->y.p, (long long)n, 0.0);
    spmv(*S, 1.0, S->s.p, 0.0, S->y.p, S->y.p, nullptr, nullptr, nullptr, nullptr, f32);
    HIPCHK(hipMemcpyAsync(y, S->y.p, n * sizeof(double), hipMemcpyDeviceToHost, S->st));
    HIPCHK(hipStreamSynchronize(S->st));
    if (S->spmvTiming) collect_spmv_timing(*S);
    collect_tuning(*S);
    return MPH_OK;
}

int mph_scene_apply_matrix_noalloc(mph_scene* S, int precision, const double* x, double* y) {
    if (!S || !x || !y) return fail(MPH_E_ARG, "null argument");
    if (precision != MPH_PRECISION_FP64 && precision != MPH_PRECISION_MIXED)
        return fail(MPH_E_ARG, "unknown precision mode");
    const bool f32 = precision == MPH_PRECISION_MIXED;
    if (f32 && !mixed(*S)) return fail(MPH_E_STATE, "FP32 values exist in mixed-precision scenes only");
    const size_t need = noalloc_scratch_bytes(*S, f32);
    if (need > kNoAllocScratchBytes)
        return fail(MPH_E_ARG, "the product's records need " + std::to_string(need) + " bytes, more than the " +
                                   std::to_string(kNoAllocScratchBytes) + "-byte scratch of the allocation-free entry "
                                   "point (at most " + std::to_string(kNoAllocScratchBytes / (f32 ? 16 : 56)) +
                                   " particles)");
    if (!S->assembled) return fail(MPH_E_STATE, "no particle matrix for the current particles");
    HIPCHK(hipSetDevice(S->device));
    double* scr = nullptr;
    int rc;
    if ((rc = noalloc_scratch(S->device, &scr))) return rc;
    S->facPref = nullptr;  // a first pass enqueued ahead by the solve is not this product's
    S->facPrefEv = -1;
    const DevParams& P = S->P;
    launch_fill(S->st, y, 3LL * P.nf, 0.0);  // y = 0*y + A x reads y
    if (factored(*S)) {
        FacOp F = fac_op(*S, f32);
        if (f32) {
            F.rec32 = (float*)scr;
        } else {
            F.rec = scr;
            launch_fac_rec_positions(S->st, P, S->frec.p, scr);
        }
        launch_fac_spmv(S->st, P, F, 1.0, x, 0.0, y, y, nullptr, nullptr, nullptr, nullptr, f32, 3);
    } else {
        const Autotune& T = f32 ? S->tuneSpmv32 : S->tuneSpmv64;
        launch_spmv(S->st, P, aslices(*S), S->nFluid.p, 1.0, x, 0.0, y, y, nullptr, nullptr, nullptr, nullptr, f32,
                    T.best >= 0 ? T.best : 0);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MPH_E_DEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(S->st));
    return MPH_OK;
}

int mph_scene_device_bytes(mph_scene* S, unsigned long long* bytes) {
    if (!S || !bytes) return fail(MPH_E_ARG, "null argument");
    unsigned long long n = 0;
    for (DBuf<int>* b : {&S->type, &S->type2, &S->perm, &S->perm2, &S->cellIndex, &S->cellStart, &S->bbox,
                         &S->nCount, &S->nFluid, &S->nCountP, &S->nInd, &S->nIndP, &S->width, &S->selfIdx, &S->acol,
                         &S->maxP, &S->runP, &S->fwrap, &S->fgmask, &S->scrM, &S->scrP, &S->searchOvf})
        n += b->cap * sizeof(int);
    n += (S->fgtab.cap + S->fblk.cap) * sizeof(double);
    for (DBuf<double>* b : {&S->falpha, &S->fdiag, &S->selfC, &S->fbeta, &S->frec}) n += b->cap * sizeof(double);
    n += (S->fu32.cap + S->fv32.cap + S->frec32.cap) * sizeof(float);
    n += S->fslot.cap * sizeof(unsigned short);
    n += S->runStart.cap * sizeof(unsigned short);
    for (DBuf<double>* b : {&S->pos, &S->pos2, &S->vel, &S->vel2, &S->force, &S->mass, &S->kappa, &S->lambda,
                            &S->mu, &S->densA, &S->gcen, &S->pressA, &S->volStrainP, &S->divP, &S->pressP,
                            &S->virialP, &S->virialStress, &S->aval, &S->b, &S->cval, &S->dval, &S->vecP,
                            &S->invDiagA, &S->rOut, &S->dvec, &S->x, &S->r, &S->s, &S->y, &S->p, &S->q, &S->u,
                            &S->part, &S->scal, &S->convRec})
        n += b->cap * sizeof(double);
    for (DBuf<long long>* b : {&S->nPtr, &S->nPtrP, &S->scanScratch, &S->sliceStart}) n += b->cap * sizeof(long long);
    n += S->keys.cap * sizeof(uint64_t);
    for (HostLevel& H : S->lv) {
        for (DBuf<double>* b : {&H.A, &H.invD, &H.S0, &H.S1, &H.R, &H.Q}) n += b->cap * sizeof(double);
        n += H.A32.cap * sizeof(float);
    }
    n += S->aval32.cap * sizeof(float);
    *bytes = n;
    return MPH_OK;
}

int mph_scene_destroy(mph_scene* S) {
    if (!S) return MPH_OK;
    if (S->batches > 0) return fail(MPH_E_STATE, "scene is still a member of a batch: remove it first");
    (void)hipSetDevice(S->device);
    destroy(S);
    return MPH_OK;
}

}  // extern "C"
