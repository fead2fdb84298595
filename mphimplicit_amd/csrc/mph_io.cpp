// mph_io.cpp -- the reference's text formats: *.data, *.grid/*.prof and legacy VTK.
//
// readDataFile (source/main.cpp:523-563) and readGridFile (565-694) are keyword /
// column parsers over sscanf; the patterns below are the same patterns, so a file
// parses to the same values.  writeProfFile (696-720) and writeVtkFile (722-883)
// are reproduced byte for byte: %e formatting, (float) casts of the VTK fields,
// Property and NeighborCount printed with %d.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mph_abi.h"

void mph_internal_set_error(const char* msg);  // mph_scene.cpp

namespace {
int io_fail(int code, const std::string& what) {
    mph_internal_set_error(what.c_str());
    return code;
}
}  // namespace

extern "C" {

int mph_read_data_file(const char* path, mph_params* p) {
    if (!path || !p) return io_fail(MPH_E_ARG, "null argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return io_fail(MPH_E_IO, std::string("cannot open ") + path);
    char buf[1024];
    while (!std::feof(fp) && !std::ferror(fp)) {
        if (std::fgets(buf, sizeof buf, fp) == nullptr) continue;
        double(&w2c)[3] = p->wall_center[2];
        double(&w2v)[3] = p->wall_velocity[2];
        double(&w2o)[3] = p->wall_omega[2];
        double(&w3c)[3] = p->wall_center[3];
        double(&w3v)[3] = p->wall_velocity[3];
        double(&w3o)[3] = p->wall_omega[3];
        double* d = p->density;
        double* k = p->bulk_modulus;
        double* bv = p->bulk_viscosity;
        double* sv = p->shear_viscosity;
        double* st = p->surface_tension;
        double(*ir)[4] = p->interaction_ratio;
        if (buf[0] == '#') {
        } else if (std::sscanf(buf, " Dt %lf", &p->dt) == 1) {
        } else if (std::sscanf(buf, " OutputInterval %lf", &p->output_interval) == 1) {
        } else if (std::sscanf(buf, " VtkOutputInterval %lf", &p->vtk_output_interval) == 1) {
        } else if (std::sscanf(buf, " EndTime %lf", &p->end_time) == 1) {
        } else if (std::sscanf(buf, " RadiusRatioA %lf", &p->radius_ratio_a) == 1) {
        } else if (std::sscanf(buf, " RadiusRatioP %lf", &p->radius_ratio_p) == 1) {
        } else if (std::sscanf(buf, " RadiusRatioV %lf", &p->radius_ratio_v) == 1) {
        } else if (std::sscanf(buf, " Density %lf %lf %lf %lf", &d[0], &d[1], &d[2], &d[3]) == 4) {
        } else if (std::sscanf(buf, " BulkModulus %lf %lf %lf %lf", &k[0], &k[1], &k[2], &k[3]) == 4) {
        } else if (std::sscanf(buf, " BulkViscosity %lf %lf %lf %lf", &bv[0], &bv[1], &bv[2], &bv[3]) == 4) {
        } else if (std::sscanf(buf, " ShearViscosity %lf %lf %lf %lf", &sv[0], &sv[1], &sv[2], &sv[3]) == 4) {
        } else if (std::sscanf(buf, " SurfaceTension %lf %lf %lf %lf", &st[0], &st[1], &st[2], &st[3]) == 4) {
        } else if (std::sscanf(buf, " InteractionRatio(Type0) %lf %lf %lf %lf", &ir[0][0], &ir[0][1], &ir[0][2],
                               &ir[0][3]) == 4) {
        } else if (std::sscanf(buf, " InteractionRatio(Type1) %lf %lf %lf %lf", &ir[1][0], &ir[1][1], &ir[1][2],
                               &ir[1][3]) == 4) {
        } else if (std::sscanf(buf, " InteractionRatio(Type2) %lf %lf %lf %lf", &ir[2][0], &ir[2][1], &ir[2][2],
                               &ir[2][3]) == 4) {
        } else if (std::sscanf(buf, " InteractionRatio(Type3) %lf %lf %lf %lf", &ir[3][0], &ir[3][1], &ir[3][2],
                               &ir[3][3]) == 4) {
        } else if (std::sscanf(buf, " Gravity %lf %lf %lf", &p->gravity[0], &p->gravity[1], &p->gravity[2]) == 3) {
        } else if (std::sscanf(buf, " Wall2  Center %lf %lf %lf Velocity %lf %lf %lf Omega %lf %lf %lf", &w2c[0],
                               &w2c[1], &w2c[2], &w2v[0], &w2v[1], &w2v[2], &w2o[0], &w2o[1], &w2o[2]) == 9) {
        } else if (std::sscanf(buf, " Wall3  Center %lf %lf %lf Velocity %lf %lf %lf Omega %lf %lf %lf", &w3c[0],
                               &w3c[1], &w3c[2], &w3v[0], &w3v[1], &w3v[2], &w3o[0], &w3o[1], &w3o[2]) == 9) {
        } else {
            std::fprintf(stderr, "Invalid line in data file \"%s\"\n", buf);
        }
    }
    std::fclose(fp);
    return MPH_OK;
}

int mph_read_grid_header(const char* path, int* np) {
    if (!path || !np) return io_fail(MPH_E_ARG, "null argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return io_fail(MPH_E_IO, std::string("cannot open ") + path);
    char buf[1024];
    int rc = MPH_E_IO;
    if (std::fgets(buf, sizeof buf, fp) && std::fgets(buf, sizeof buf, fp) && std::sscanf(buf, "%d", np) == 1)
        rc = *np >= 0 ? MPH_OK : MPH_E_IO;
    std::fclose(fp);
    return rc ? io_fail(rc, std::string("malformed grid header in ") + path) : rc;
}

int mph_read_grid_file(const char* path, int np, double* time, double* spacing, double dmin[3], double dmax[3],
                       int* type, double* pos, double* vel) {
    if (!path || !time || !spacing || !dmin || !dmax || (np > 0 && (!type || !pos || !vel)))
        return io_fail(MPH_E_ARG, "null argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return io_fail(MPH_E_IO, std::string("cannot open ") + path);
    char buf[1024];
    int count = 0;
    if (!std::fgets(buf, sizeof buf, fp) || std::sscanf(buf, "%lf", time) != 1 || !std::fgets(buf, sizeof buf, fp) ||
        std::sscanf(buf, "%d  %lf  %lf %lf %lf  %lf %lf %lf", &count, spacing, &dmin[0], &dmax[0], &dmin[1], &dmax[1],
                    &dmin[2], &dmax[2]) != 8 ||
        count != np) {
        std::fclose(fp);
        return io_fail(MPH_E_IO, std::string("malformed grid header in ") + path);
    }
    for (int i = 0; i < np; ++i) {
        if (!std::fgets(buf, sizeof buf, fp) ||
            std::sscanf(buf, "%d  %lf %lf %lf  %lf %lf %lf", &type[i], &pos[3 * i], &pos[3 * i + 1], &pos[3 * i + 2],
                        &vel[3 * i], &vel[3 * i + 1], &vel[3 * i + 2]) != 7) {
            std::fclose(fp);
            return io_fail(MPH_E_IO, std::string("malformed particle line ") + std::to_string(i) + " in " + path);
        }
    }
    std::fclose(fp);
    return MPH_OK;
}

// scalars and fields of the scene used by the writers
int mph_scene_scalar(mph_scene* s, const char* name, double* value);
int mph_scene_field_size(mph_scene* s, const char* field, size_t* bytes, int* elem_bytes);
int mph_scene_read(mph_scene* s, const char* field, void* host, size_t bytes);
int mph_scene_domain(mph_scene* s, double* spacing, double dmin[3], double dmax[3]);

}  // extern "C"

namespace {

template <class T>
int fetch(mph_scene* s, const char* name, std::vector<T>& v) {
    size_t bytes = 0;
    int es = 0;
    int rc = mph_scene_field_size(s, name, &bytes, &es);
    if (rc) return rc;
    v.resize(bytes / sizeof(T));
    return mph_scene_read(s, name, v.data(), bytes);
}

}  // namespace

extern "C" {

// writeProfFile, main.cpp:696-720, over host arrays
int mph_write_prof_arrays(const char* path, double time, int n, double spacing, const double dmin[3],
                          const double dmax[3], const int* type, const double* pos, const double* vel) {
    if (!path || n < 0 || (n > 0 && (!type || !pos || !vel))) return MPH_E_ARG;
    FILE* fp = std::fopen(path, "w");
    if (!fp) return MPH_E_IO;
    std::fprintf(fp, "%e\n", time);
    std::fprintf(fp, "%d %e %e %e %e %e %e %e\n", n, spacing, dmin[0], dmax[0], dmin[1], dmax[1], dmin[2], dmax[2]);
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%d %e %e %e  %e %e %e\n", type[i], pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], vel[3 * i],
                     vel[3 * i + 1], vel[3 * i + 2]);
    std::fflush(fp);
    std::fclose(fp);
    return MPH_OK;
}

// writeVtkFile, main.cpp:722-883, over host arrays
int mph_write_vtk_arrays(const char* path, int n, const int* type, const double* pos, const double* vel,
                         const double* force, const double* volStrainP, const double* divergenceP,
                         const double* pressureP, const double* virialPressure, const int* neighborCount) {
    if (!path || n < 0) return MPH_E_ARG;
    if (n > 0 && (!type || !pos || !vel || !force || !volStrainP || !divergenceP || !pressureP || !virialPressure ||
                  !neighborCount))
        return MPH_E_ARG;
    FILE* fp = std::fopen(path, "w");
    if (!fp) return MPH_E_IO;
    std::fprintf(fp, "# vtk DataFile Version 2.0\n");
    std::fprintf(fp, "Unstructured Grid Example\n");
    std::fprintf(fp, "ASCII\n");
    std::fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
    std::fprintf(fp, "POINTS %d float\n", n);
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%e %e %e\n", (float)pos[3 * i], (float)pos[3 * i + 1], (float)pos[3 * i + 2]);
    std::fprintf(fp, "CELLS %d %d\n", n, 2 * n);
    for (int i = 0; i < n; ++i) std::fprintf(fp, "1 %d ", i);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "CELL_TYPES %d\n", n);
    for (int i = 0; i < n; ++i) std::fprintf(fp, "1 ");
    std::fprintf(fp, "\n");
    std::fprintf(fp, "\n");
    std::fprintf(fp, "POINT_DATA %d\n", n);
    std::fprintf(fp, "SCALARS label float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%d\n", type[i]);
    std::fprintf(fp, "\n");
    const char* names[4] = {"VolStrainP", "DivergenceP", "PressureP", "VirialPressureAtParticle"};
    const double* vals[4] = {volStrainP, divergenceP, pressureP, virialPressure};
    for (int f = 0; f < 4; ++f) {
        std::fprintf(fp, "SCALARS %s float 1\n", names[f]);
        std::fprintf(fp, "LOOKUP_TABLE default\n");
        for (int i = 0; i < n; ++i) std::fprintf(fp, "%e\n", (float)vals[f][i]);
        std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "SCALARS neighbor float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%d\n", neighborCount[i]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "VECTORS velocity float\n");
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%e %e %e\n", (float)vel[3 * i], (float)vel[3 * i + 1], (float)vel[3 * i + 2]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "VECTORS force float\n");
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%e %e %e\n", (float)force[3 * i], (float)force[3 * i + 1], (float)force[3 * i + 2]);
    std::fprintf(fp, "\n");
    std::fflush(fp);
    std::fclose(fp);
    return MPH_OK;
}

int mph_scene_write_prof(mph_scene* s, const char* path) {
    if (!s || !path) return MPH_E_ARG;
    double t = 0, spacing = 0, dmin[3], dmax[3];
    int rc = mph_scene_scalar(s, "Time", &t);
    if (rc || (rc = mph_scene_domain(s, &spacing, dmin, dmax))) return rc;
    std::vector<int> type;
    std::vector<double> pos, vel;
    if ((rc = fetch(s, "Property", type)) || (rc = fetch(s, "Position", pos)) || (rc = fetch(s, "Velocity", vel)))
        return rc;
    return mph_write_prof_arrays(path, t, (int)type.size(), spacing, dmin, dmax, type.data(), pos.data(), vel.data());
}

int mph_scene_write_vtk(mph_scene* s, const char* path) {
    if (!s || !path) return MPH_E_ARG;
    std::vector<int> type, ncount;
    std::vector<double> pos, vel, force, vs, dv, pp, vp;
    int rc;
    if ((rc = fetch(s, "Property", type)) || (rc = fetch(s, "Position", pos)) || (rc = fetch(s, "Velocity", vel)) ||
        (rc = fetch(s, "Force", force)) || (rc = fetch(s, "VolStrainP", vs)) || (rc = fetch(s, "DivergenceP", dv)) ||
        (rc = fetch(s, "PressureP", pp)) || (rc = fetch(s, "VirialPressureAtParticle", vp)) ||
        (rc = fetch(s, "NeighborCount", ncount)))
        return rc;
    return mph_write_vtk_arrays(path, (int)type.size(), type.data(), pos.data(), vel.data(), force.data(), vs.data(),
                                dv.data(), pp.data(), vp.data(), ncount.data());
}

}  // extern "C"
