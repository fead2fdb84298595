// ref_probe.cpp -- test instrument, never shipped, never linked into the product.
//
// Builds the reference solver's own translation unit (source/main.cpp of
// Masahiro-Kondo-AIST/MphImplicit, compiled unmodified from where it lies) into a
// driver that calls the reference's phase functions in the order of its time loop
// (source/main.cpp:386-474) and dumps the full-precision state after each phase.
// Nothing of the reference's arithmetic is changed.  The only additions are:
//   * `main` is renamed so this file can supply the driver;
//   * `log_printf` is routed through probe_log() which still prints the text but
//     also captures the CR summary (iter, rr0, rr, rs0, rs) at full precision
//     (source/main.cpp:4381);
//   * ParticleIndex is set to 0..N-1 after reading the grid (the reference
//     allocates it but never initialises it, source/main.cpp:586), so the dumps
//     carry the sorted-slot -> input-particle permutation.  It is never read by
//     any arithmetic in the reference.
//
// The reference source path is given by -DMPH_REFERENCE_MAIN="\".../main.cpp\"".
// Output: <outdir>/stepNNN/<array>.bin (raw little-endian) + manifest.txt lines
//   "<array> <dtype> <count>".

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdarg>
#include <string>
#include <sys/stat.h>
#include <time.h>

static int    probe_cr_iter = -1;
static double probe_cr_rr0 = 0, probe_cr_rr = 0, probe_cr_rs0 = 0, probe_cr_rs = 0;

int probe_log(const char* format, ...);

#define main mph_reference_main
#define log_printf probe_log
#include MPH_REFERENCE_MAIN
#undef main
#undef log_printf
#undef malloc
#undef fopen

int probe_log(const char* format, ...)
{
    // capture the CR summary line (source/main.cpp:4381 / 3970)
    if (strncmp(format, "line:%d, iter,=%d, rr0=,%e, rr=,%e, rs0=,%e, rs=,%e", 48) == 0) {
        va_list a;
        va_start(a, format);
        (void)va_arg(a, int);
        probe_cr_iter = va_arg(a, int);
        probe_cr_rr0 = va_arg(a, double);
        probe_cr_rr = va_arg(a, double);
        probe_cr_rs0 = va_arg(a, double);
        probe_cr_rs = va_arg(a, double);
        va_end(a);
    }
    if (getenv("REF_PROBE_QUIET")) return 0;
    va_list b;
    va_start(b, format);
    int r = vfprintf(stderr, format, b);
    va_end(b);
    return r;
}

static std::string g_out;
static bool g_time = false;  // REF_PROBE_TIME: phase timing, no dumps
static std::string g_dir;
static FILE* g_manifest = nullptr;

static void open_step(const char* tag)
{
    if (g_time) return;
    if (g_manifest) fclose(g_manifest);
    g_dir = g_out + "/" + tag;
    mkdir(g_dir.c_str(), 0755);
    g_manifest = ::fopen((g_dir + "/manifest.txt").c_str(), "w");
}

// REF_PROBE_TIME=1: no dumps; the wall time of every phase of every step is printed
// to stdout instead ("phase <step> <name> <seconds>"), for timing the reference
// phase by phase (built with -fopenmp as ref_probe_3d_omp)
static double probe_now()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
static double g_t0 = 0;
static void phase_done(int step, const char* name)
{
    if (!g_time) return;
    const double t = probe_now();
    printf("phase %d %s %.6f\n", step, name, t - g_t0);
    fflush(stdout);
    g_t0 = probe_now();
}

static void dump(const char* name, const void* p, const char* dtype, long count, size_t esize)
{
    if (g_time) return;
    std::string path = g_dir + "/" + name + ".bin";
    FILE* f = ::fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    if (count > 0) fwrite(p, esize, (size_t)count, f);
    fclose(f);
    fprintf(g_manifest, "%s %s %ld\n", name, dtype, count);
    fflush(g_manifest);
}
#define DUMP_I32(n, p, c) dump(n, p, "i4", (long)(c), 4)
#define DUMP_I64(n, p, c) dump(n, p, "i8", (long)(c), 8)
#define DUMP_F64(n, p, c) dump(n, p, "f8", (long)(c), 8)

// what to dump (bit mask, argv[5])
enum {
    D_CELLS = 1, D_NEIGH = 2, D_EXPL = 4, D_MATA = 8, D_MATC = 16, D_MULT = 32,
    D_MG = 64, D_SOLVE = 128, D_VTK = 256, D_PRECOND = 512, D_FILES = 1024
};

static void dump_particles(const char* suffix)
{
    char n[128];
    snprintf(n, sizeof n, "Position%s", suffix);  DUMP_F64(n, Position, 3L * ParticleCount);
    snprintf(n, sizeof n, "Velocity%s", suffix);  DUMP_F64(n, Velocity, 3L * ParticleCount);
    snprintf(n, sizeof n, "Property%s", suffix);  DUMP_I32(n, Property, ParticleCount);
}

static void dump_mg()
{
    int d = MultiGridDepth;
    DUMP_I32("MultiGridDepth", &MultiGridDepth, 1);
    DUMP_I32("MultiGridCellMin", MultiGridCellMin, 3L * d);
    DUMP_I32("MultiGridCellMax", MultiGridCellMax, 3L * d);
    DUMP_I32("MultiGridCount", MultiGridCount, 3L * d);
    DUMP_I32("MultiGridOffset", MultiGridOffset, d + 1L);
    const int fluid = FluidParticleEnd - FluidParticleBegin;
    DUMP_F64("InvDiagA", InvDiagA, 3L * fluid);
    // per-level A values and inverse diagonals in level order 0..d-1
    for (int l = 0; l < d; ++l) {
        const int offPtr = MultiGridOffset[d - l - 1] * OneGridSizeCsrPtrA + (d - l - 1);
        const int offInd = MultiGridOffset[d - l - 1] * OneGridSizeCsrIndA;
        const int offInv = MultiGridOffset[d - l - 1] * OneGridSizeInvDiagA;
        const int* gc = MultiGridCount[l];
        const int nrow = 3 * gc[0] * gc[1] * gc[2];
        char n[64];
        snprintf(n, sizeof n, "MGPtrA_L%d", l);  DUMP_I32(n, &MultiGridCsrPtrA[offPtr], nrow + 1);
        const int nnz = MultiGridCsrPtrA[offPtr + nrow];
        snprintf(n, sizeof n, "MGIndA_L%d", l);  DUMP_I32(n, &MultiGridCsrIndA[offInd], nnz);
        snprintf(n, sizeof n, "MGCofA_L%d", l);  DUMP_F64(n, &MultiGridCsrCofA[offInd], nnz);
        snprintf(n, sizeof n, "MGInvD_L%d", l);  DUMP_F64(n, &MultiGridInvDiagA[offInv], nrow);
    }
}

int main(int argc, char** argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s data grid outdir nsteps [dumpmask] [vtkEvery]\n", argv[0]);
        return 2;
    }
    g_out = argv[3];
    mkdir(g_out.c_str(), 0755);
    const int nsteps = atoi(argv[4]);
    const int mask = argc > 5 ? atoi(argv[5]) : 0xffff;
    g_time = getenv("REF_PROBE_TIME") != nullptr;

    log_open((g_out + "/probe.log").c_str());
    readDataFile(argv[1]);
    readGridFile(argv[2]);
    for (int i = 0; i < ParticleCount; ++i) ParticleIndex[i] = i;
    initializeWeight();
    initializeFluid();
    initializeWall();
    initializeDomain();

    open_step("init");
    {
        double consts[12] = {N0a, N0p, Swa, Swg, Swp, Swv, R2g, MaxRadius, RadiusA, RadiusP, RadiusV, RadiusG};
        DUMP_F64("consts", consts, 12);
        DUMP_I32("CellCount", CellCount, 3);
        DUMP_F64("CellWidth", CellWidth, 3);
        int p2[2] = {PowerParticleCount, ParticleCountPower};
        DUMP_I32("Power", p2, 2);
        DUMP_F64("CofA", CofA, 4);
        DUMP_F64("WallRotation", WallRotation, 4 * 9);
    }
    dump_particles("_in");
    // the init sequence of source/main.cpp:369-375 (VTK write of output.vtk skipped)
    calculateCellParticle();
    calculateNeighbor();
    calculateDensityA();
    calculateGravityCenter();
    calculateDensityP();
    if (mask & D_CELLS) {
        DUMP_I32("ParticleIndex", ParticleIndex, ParticleCount);
        DUMP_I32("CellIndex", CellIndex, PowerParticleCount);
        DUMP_I32("CellParticle", CellParticle, PowerParticleCount);
    }
    if (mask & D_NEIGH) {
        DUMP_I32("NeighborCount", NeighborCount, ParticleCount);
        DUMP_I64("NeighborPtr", NeighborPtr, ParticleCount + 1L);
        DUMP_I32("NeighborInd", NeighborInd, NeighborIndCount);
    }
    dump_particles("_sorted");
    freeNeighbor();

    int iStep = (int)(Time / Dt);
    OutputNext = Time;
    VtkOutputNext = Time;
    int done = 0;
    while (Time < EndTime + 1.0e-5 * Dt && done < nsteps) {
        char tag[32];
        snprintf(tag, sizeof tag, "step%03d", done);
        open_step(tag);
        double tinfo[2] = {Time, Dt};
        DUMP_F64("Time", tinfo, 2);
        dump_particles("_begin");

        if (Time + 1.0e-5 * Dt >= OutputNext) {
            if (mask & D_FILES) {
                char fn[1024];
                snprintf(fn, sizeof fn, "%s/ref.prof", g_dir.c_str());
                writeProfFile(fn);
            }
            OutputNext += OutputInterval;
        }

        g_t0 = probe_now();
        calculateConvection();
        calculateWall();
        calculatePeriodicBoundary();
        resetForce();
        phase_done(done, "kinematics");
        dump_particles("_kin");

        calculateCellParticle();
        phase_done(done, "cells");
        if (mask & D_CELLS) {
            DUMP_I32("ParticleIndex", ParticleIndex, ParticleCount);
            DUMP_I32("CellIndex", CellIndex, PowerParticleCount);
            DUMP_I32("CellParticle", CellParticle, PowerParticleCount);
            DUMP_I32("CellFluidParticleBegin", CellFluidParticleBegin, TotalCellCount);
            DUMP_I32("CellFluidParticleEnd", CellFluidParticleEnd, TotalCellCount);
            DUMP_I32("CellWallParticleBegin", CellWallParticleBegin, TotalCellCount);
            DUMP_I32("CellWallParticleEnd", CellWallParticleEnd, TotalCellCount);
            int r[4] = {FluidParticleBegin, FluidParticleEnd, WallParticleBegin, WallParticleEnd};
            DUMP_I32("Ranges", r, 4);
            dump_particles("_sort");
        }
        calculateNeighbor();
        phase_done(done, "neighbor");
        if (mask & D_NEIGH) {
            DUMP_I32("NeighborCount", NeighborCount, ParticleCount);
            DUMP_I32("NeighborFluidCount", NeighborFluidCount, ParticleCount);
            DUMP_I32("NeighborCountP", NeighborCountP, ParticleCount);
            DUMP_I64("NeighborPtr", NeighborPtr, ParticleCount + 1L);
            DUMP_I64("NeighborPtrP", NeighborPtrP, ParticleCount + 1L);
            DUMP_I32("NeighborInd", NeighborInd, NeighborIndCount);
            DUMP_I32("NeighborIndP", NeighborIndP, NeighborIndCountP);
        }

        calculateDensityA();
        calculateGravityCenter();
        calculateDensityP();
        calculatePhysicalCoefficients();
        calculatePressureA();
        calculateDiffuseInterface();
        calculateGravity();
        calculateAcceleration();
        phase_done(done, "explicit");
        if (mask & D_EXPL) {
            DUMP_F64("DensityA", DensityA, ParticleCount);
            DUMP_F64("GravityCenter", GravityCenter, 3L * ParticleCount);
            DUMP_F64("VolStrainP", VolStrainP, ParticleCount);
            DUMP_F64("PressureA", PressureA, ParticleCount);
            DUMP_F64("Mass", Mass, ParticleCount);
            DUMP_F64("Kappa", Kappa, ParticleCount);
            DUMP_F64("Lambda", Lambda, ParticleCount);
            DUMP_F64("Mu", Mu, ParticleCount);
            DUMP_F64("Force", Force, 3L * ParticleCount);
            DUMP_F64("VelocityExpl", Velocity, 3L * ParticleCount);
        }

        calculateMatrixA();
        phase_done(done, "matrix_a");
        const int fluid = FluidParticleEnd - FluidParticleBegin;
        const long n = 3L * fluid;
        if (mask & D_MATA) {
            DUMP_I64("CsrPtrA", CsrPtrA, n + 1);
            DUMP_I32("CsrIndA", CsrIndA, NonzeroCountA);
            DUMP_F64("CsrCofA_visc", CsrCofA, NonzeroCountA);
            DUMP_F64("VectorB_visc", VectorB, n);
        }
        calculateMatrixC();
        phase_done(done, "matrix_c");
        if (mask & D_MATC) {
            DUMP_I64("CsrPtrC", CsrPtrC, n + 1);
            DUMP_I32("CsrIndC", CsrIndC, NonzeroCountC);
            DUMP_F64("CsrCofC", CsrCofC, NonzeroCountC);
            DUMP_F64("VectorP", VectorP, ParticleCount);
        }
        multiplyMatrixC();
        freeMatrixC();
        phase_done(done, "multiply_c");
        if (mask & D_MULT) {
            DUMP_F64("CsrCofA", CsrCofA, NonzeroCountA);
            DUMP_F64("VectorB", VectorB, n);
        }
        calculateMultiGridMatrix();
        phase_done(done, "multigrid");
        if (mask & D_MG) dump_mg();
        if (mask & D_PRECOND) {
            // M^-1 applied to a fixed vector e_k = 1 + ((k*7919) % 13) / 8
            double* q = (double*)std::malloc(sizeof(double) * n);
            double* s = (double*)std::malloc(sizeof(double) * n);
            double* buf = (double*)std::malloc(sizeof(double) * n);
            for (long k = 0; k < n; ++k) {
                q[k] = 1.0 + (double)((k * 7919) % 13) / 8.0;
                s[k] = -1.0;
                buf[k] = -1.0;
            }
            for (int k = 0; k < AllGridSizeVecS; ++k) MultiGridVecS[k] = -1.0;
            for (int k = 0; k < AllGridSizeVecR; ++k) MultiGridVecR[k] = -1.0;
            for (int k = 0; k < AllGridSizeVecQ; ++k) MultiGridVecQ[k] = -1.0;
            preconditionWithMultiGrid(q, s, buf);
            DUMP_F64("PrecondIn", q, n);
            DUMP_F64("PrecondOut", s, n);
            std::free(q); std::free(s); std::free(buf);
        }
        probe_cr_iter = -1;
        g_t0 = probe_now();
        solveWithConjugatedGradient();
        freeMultiGridMatrix();
        freeMatrixA();
        phase_done(done, "solve");
        if (mask & D_SOLVE) {
            DUMP_F64("VelocitySolved", Velocity, 3L * ParticleCount);
            double cr[5] = {(double)probe_cr_iter, probe_cr_rr0, probe_cr_rr, probe_cr_rs0, probe_cr_rs};
            DUMP_F64("CR", cr, 5);
        }

        if (Time + 1.0e-5 * Dt >= VtkOutputNext) {
            calculateDivergenceP();
            calculatePressureP();
            calculateViscosityV();
            calculateVirialStressAtParticle();
            if (mask & D_VTK) {
                DUMP_F64("DivergenceP", DivergenceP, ParticleCount);
                DUMP_F64("PressureP", PressureP, ParticleCount);
                DUMP_F64("VirialPressureAtParticle", VirialPressureAtParticle, ParticleCount);
                DUMP_F64("VirialStressAtParticle", VirialStressAtParticle, 9L * ParticleCount);
                DUMP_F64("ForceVtk", Force, 3L * ParticleCount);
                DUMP_I32("NeighborCountVtk", NeighborCount, ParticleCount);
            }
            if (mask & D_FILES) {
                char fn[1024];
                snprintf(fn, sizeof fn, "%s/ref.vtk", g_dir.c_str());
                writeVtkFile(fn);
            }
            VtkOutputNext += VtkOutputInterval;
        }
        freeNeighbor();
        Time += Dt;
        iStep++;
        done++;
        dump_particles("_end");
    }
    if (g_manifest) fclose(g_manifest);
    log_close();
    return 0;
}
