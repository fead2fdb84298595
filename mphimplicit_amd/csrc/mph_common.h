// mph_common.h -- device-side constants and arithmetic shared by the HIP kernels.
//
// All arithmetic that feeds a parity-checked result keeps the operand order and
// association of source/main.cpp (Masahiro-Kondo-AIST/MphImplicit).  Every HIP file
// is compiled with -ffp-contract=off, so each multiply and add rounds separately,
// exactly like the x86-64 reference build (no FMA).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPH_HD __host__ __device__ inline

namespace mph {

// Run-time constants of one scene: the reference's globals that never change after
// initialisation (initializeWeight/Fluid/Wall/Domain, source/main.cpp:886-1162),
// plus prefactors whose values are identical to what the reference recomputes
// inside every weight-function call.
struct DevParams {
    int dim;      // 2 or 3
    int np;       // ParticleCount
    int nf;       // fluid count (= FluidParticleEnd after every sort)
    int p2;       // PowerParticleCount
    int p2pow;    // ParticleCountPower
    int C[3];     // CellCount
    int T;        // TotalCellCount
    int range[3]; // neighbour cell range per axis (main.cpp:1393-1399)
    double cw[3];   // CellWidth
    double dmin[3]; // DomainMin
    double W[3];    // DomainWidth
    double hW[3];   // 0.5*DomainWidth
    double l, V, dt;
    double radA, radG, radP, radV, maxR;
    double radA2, radG2, radP2, radV2, maxR2;
    double N0a, N0p, R2g;
    // weight prefactors 1.0/Sw * 1.0/(h^d), as evaluated in wa..dwvdr (main.cpp:236-307)
    double Ka, Kg, Kp, Kv;
    double invA;         // 1.0/RadiusA   (dwadr)
    double m2G, m2P, m2V; // -2.0/h        (dwgdr, dwpdr, dwvdr)
    double cofA[4], cofK;
    double ai[4];        // CofA[t]*CofK*CofK
    double visc;         // 10.0 (3D) or 8.0 (2D), main.cpp:2285-2287
    double gravity[3];
    double density[4], bulk[4], bulkVisc[4], shear[4];
    double ratio[4][4];  // InteractionRatio
    // the CR stop flag: the kernels of a CR iteration return at once while *halt is
    // nonzero (null outside the CR loop, and in the solve with the host stop test)
    const int* halt;
};

// a CR iteration kernel's entry test (uniform: the whole grid returns or none of it)
__device__ inline bool cr_halted(const int* halt) { return halt && *halt; }

// orders the LDS writes and reads of the lanes of one wavefront: the entry-parallel
// walks hand terms to a group's first lane, which adds them in order (the lanes of a
// group are in one wavefront and run the same iterations)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

MPH_HD bool is_fluid(int t) { return 0 <= t && t < 2; }
MPH_HD bool is_wall(int t) { return 2 <= t && t < 4; }

// Mod(x,w) of main.cpp:79.  When 0 < x <= 0.75w the quotient rounds into (0, 1),
// so floor(x/w) = +0 and the reference's result x - w*(+0) equals x - w*0.0: the
// division is skipped.  Every other x (zeros of either sign included) takes the
// reference's formula.
__device__ inline double modw(double x, double w) {
    if (x > 0.0 && x <= 0.75 * w) return x - w * 0.0;
    return x - w * floor(x / w);
}

// minimum-image difference of every neighbour loop, e.g. main.cpp:1452
__device__ inline double min_image(double xj, double xi, double w, double hw) {
    return modw(xj - xi + hw, w) - hw;
}

// weight functions with the prefactor precomputed (same bits as the reference)
MPH_HD double wa_k(double K, double r, double h) { return K * (r / h) * (1.0 - (r / h)) * (1.0 - (r / h)); }
MPH_HD double dwadr_k(double K, double invh, double r, double h) {
    return K * (1.0 - (r / h)) * (1.0 - 3.0 * (r / h)) * invh;
}
MPH_HD double wq_k(double K, double r, double h) { return K * ((1.0 - r / h) * (1.0 - r / h)); }
MPH_HD double dwq_k(double K, double m2h, double r, double h) { return K * (m2h * (1.0 - r / h)); }

}  // namespace mph
