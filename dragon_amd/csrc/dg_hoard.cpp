// dg_hoard.cpp -- frame and restart files in the reference's HDF5 format
// (DRAGONHOARD/HDF5Output.cpp, HDF5Restart.cpp, HDF5_Attrs.hpp), written from
// and read into x-fastest host arrays that dg_snapshot / dg_restore move to
// and from the device state.  Host code, HDF5 C library; built without FMA
// contraction so the derived datasets (E, momentum) round as the reference's
// PrimitiveState::energy() and ConservativeState(PrimitiveState) do.
#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dragon_hoard.h"

namespace {

// HDF5_Attrs.hpp:14-59
const char* const kFluid = "/fluid";
const char* const kRho = "/fluid/rho";
const char* const kP = "/fluid/p";
const char* const kE = "/fluid/E";
const char* const kV = "/fluid/v";
const char* const kVx = "/fluid/v/x";
const char* const kVy = "/fluid/v/y";
const char* const kVz = "/fluid/v/z";
const char* const kMom = "/fluid/momentum";
const char* const kPx = "/fluid/momentum/x";
const char* const kPy = "/fluid/momentum/y";
const char* const kPz = "/fluid/momentum/z";
const char* const kB = "/fluid/B";
const char* const kBx = "/fluid/B/x";
const char* const kBy = "/fluid/B/y";
const char* const kBz = "/fluid/B/z";
const char* const kBf = "/fluid/Bf";
const char* const kBfx = "/fluid/Bf/x";
const char* const kBfy = "/fluid/Bf/y";
const char* const kBfz = "/fluid/Bf/z";
const char* const kFmt = "format_version";
const char* const kWrtOpt = "write_option";
const char* const kDim = "dim";
const char* const kMhd = "MHD";
const char* const kNx = "nx";
const char* const kNy = "ny";
const char* const kNz = "nz";
const char* const kNg = "ng";
const char* const kDx = "dx";
const char* const kDy = "dy";
const char* const kDz = "dz";
const char* const kCyc = "cycle";
const char* const kTime = "time";

struct Err {
    char* buf;
    int n;
    int set(const std::string& m) {
        if (buf && n > 0) {
            std::strncpy(buf, m.c_str(), n - 1);
            buf[n - 1] = 0;
        }
        return 1;
    }
};

// An HDF5 id closed on scope exit.
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    bool ok() const { return id >= 0; }
};

// extents of the written box per dimension (cells or faces), x first
struct Box {
    int dim;
    long n[3];
    long count() const { return n[0] * (dim >= 2 ? n[1] : 1) * (dim >= 3 ? n[2] : 1); }
};
Box cell_box(const dgh_frame& f) {
    Box b{f.dim, {f.nx + 2L * f.ng, f.ny + 2L * f.ng, f.nz + 2L * f.ng}};
    return b;
}
Box face_box(const dgh_frame& f) {
    Box b{f.dim, {f.nx + 1 + 2L * f.ng, f.ny + 1 + 2L * f.ng, f.nz + 1 + 2L * f.ng}};
    return b;
}

// writeArray (HDF5Output.cpp:46-90): dims {nz, ny, nx}, chunks of at most
// 1024 (1D), 32 x 32 (2D), 16 x 16 x 16 (3D), deflate at the compression level.
template <class T>
bool write_array(hid_t file, const char* name, const T* data, const Box& b, int level) {
    hsize_t dims[3], chunk[3];
    const int r = b.dim;
    const hsize_t cap = r == 1 ? 1024 : (r == 2 ? 32 : 16);
    for (int q = 0; q < r; q++) {
        dims[q] = (hsize_t)b.n[r - 1 - q];
        chunk[q] = dims[q] < cap ? dims[q] : cap;
    }
    Hid space(H5Screate_simple(r, dims, nullptr), H5Sclose);
    Hid plist(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    if (!space.ok() || !plist.ok()) return false;
    if (level > 0) {
        if (H5Pset_chunk(plist.id, r, chunk) < 0 || H5Pset_deflate(plist.id, (unsigned)level) < 0) return false;
    }
    const hid_t type = sizeof(T) == 4 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE;
    Hid ds(H5Dcreate2(file, name, type, space.id, H5P_DEFAULT, plist.id, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return false;
    return H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
}

template <class T>
bool write_attr(hid_t file, const char* name, T value) {
    const hid_t type = std::is_same<T, int>::value ? H5T_NATIVE_INT : H5T_NATIVE_DOUBLE;
    Hid space(H5Screate(H5S_SCALAR), H5Sclose);
    Hid at(H5Acreate2(file, name, type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return at.ok() && H5Awrite(at.id, type, &value) >= 0;
}

bool make_group(hid_t file, const char* name) {
    Hid g(H5Gcreate2(file, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    return g.ok();
}

// PrimitiveState::energy() (FluidElement.cpp:53-59), the vec3 dot of
// Arithmetic.cpp:92-94: (x*x + y*y) + z*z.
double energy(const double* c, long n, long N, double gamma, bool mhd) {
    const double rho = c[0 * N + n], p = c[1 * N + n];
    const double vx = c[2 * N + n], vy = c[3 * N + n], vz = c[4 * N + n];
    double E = p / (gamma - 1.0) + (rho / 2) * (vx * vx + vy * vy + vz * vz);
    if (mhd) {
        const double bx = c[5 * N + n], by = c[6 * N + n], bz = c[7 * N + n];
        E += (bx * bx + by * by + bz * bz) * (0.125 / M_PI);
    }
    return E;
}

bool read_int_attr(hid_t file, const char* name, int* v) {
    Hid at(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
    return at.ok() && H5Aread(at.id, H5T_NATIVE_INT, v) >= 0;
}
bool read_double_attr(hid_t file, const char* name, double* v) {
    Hid at(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
    return at.ok() && H5Aread(at.id, H5T_NATIVE_DOUBLE, v) >= 0;
}
// readArray (HDF5Restart.cpp:24-37): the whole dataset as double; false if absent
bool read_array(hid_t file, const char* name, std::vector<double>& out) {
    if (H5Lexists(file, name, H5P_DEFAULT) <= 0) return false;
    Hid ds(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return false;
    Hid space(H5Dget_space(ds.id), H5Sclose);
    const int rank = H5Sget_simple_extent_ndims(space.id);
    if (rank < 1 || rank > 3) return false;
    hsize_t dims[3];
    H5Sget_simple_extent_dims(space.id, dims, nullptr);
    size_t total = 1;
    for (int q = 0; q < rank; q++) total *= (size_t)dims[q];
    out.assign(total, 0.0);
    return H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
}

// HDF5 prints its error stack by default; errors here are reported through err.
struct QuietHdf5 {
    H5E_auto2_t fn = nullptr;
    void* data = nullptr;
    QuietHdf5() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~QuietHdf5() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

}  // namespace

extern "C" {

long dgh_cell_count(const dgh_frame* f) { return f ? cell_box(*f).count() : -1; }
long dgh_face_count(const dgh_frame* f) { return f ? face_box(*f).count() : -1; }

int dgh_write(const char* path, const dgh_frame* fp, const double* cells, const double* faces, char* err, int nerr) {
    Err e{err, nerr};
    if (!path || !fp || !cells) return e.set("dgh_write: null argument");
    const dgh_frame& f = *fp;
    if (f.dim < 1 || f.dim > 3) return e.set("IO attempted with unsupported Grid type");
    const int opt = f.write_option;
    // HDF5Output.cpp:21-30
    const bool prims = (opt & DGH_WRITE_PRIMITIVE) != 0;
    const bool consv = (opt & (DGH_WRITE_CONSERVATIVE - 2)) != 0;
    const bool redundants = f.redundant != DGH_WRITE_OMIT;
    const bool write_E = (opt & 2) && (consv || redundants);
    if (!(prims || consv))
        return e.set("Specified write option does not write a sufficient description of the fluid. Please reconfigure HDF5_WRITE_OPTION in Config.h");
    const bool mhd = f.mhd != 0;
    const bool faces_out = mhd && f.dim >= 2;
    if (faces_out && !faces) return e.set("dgh_write: face array required for 2D/3D MHD");
    const Box cb = cell_box(f), fb = face_box(f);
    const long N = cb.count();
    QuietHdf5 quiet;
    Hid file(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return e.set(std::string("cannot create ") + path);
    const hid_t h = file.id;
    // metadata (HDF5Output.cpp:138-155, 250-269, 393-413)
    bool ok = write_attr(h, kFmt, 1) && write_attr(h, kWrtOpt, opt) && write_attr(h, kDim, f.dim) &&
              write_attr(h, kMhd, mhd ? 1 : 0) && write_attr(h, kNg, f.ng) && write_attr(h, kNx, f.nx);
    if (f.dim >= 2) ok = ok && write_attr(h, kNy, f.ny);
    if (f.dim >= 3) ok = ok && write_attr(h, kNz, f.nz);
    ok = ok && write_attr(h, kDx, f.dx);
    if (f.dim >= 2) ok = ok && write_attr(h, kDy, f.dy);
    if (f.dim >= 3) ok = ok && write_attr(h, kDz, f.dz);
    ok = ok && write_attr(h, kCyc, f.cycle) && write_attr(h, kTime, f.time);
    if (!ok) return e.set("attribute write failed");
    const int lv = f.compression;
    auto field = [&](int q) { return cells + (long)q * N; };
    ok = make_group(h, kFluid) && write_array(h, kRho, field(0), cb, lv);
    if (prims) {
        ok = ok && make_group(h, kV) && write_array(h, kVx, field(2), cb, lv) && write_array(h, kVy, field(3), cb, lv) &&
             write_array(h, kVz, field(4), cb, lv) && write_array(h, kP, field(1), cb, lv);
    }
    if (ok && write_E) {
        // E is PrimitiveState::energy() (equal to ConservativeState(w).E); double
        // with conservative values or the DOUBLE option, else float
        std::vector<double> E(N);
        for (long n = 0; n < N; n++) E[n] = energy(cells, n, N, f.gamma, mhd);
        if (consv || f.redundant == DGH_WRITE_DOUBLE) {
            ok = write_array(h, kE, E.data(), cb, lv);
        } else {
            std::vector<float> Ef(E.begin(), E.end());
            ok = write_array(h, kE, Ef.data(), cb, lv);
        }
        if (ok && consv) {  // ConservativeState(w).mom = rho * v (FluidElement.cpp:35-42)
            std::vector<double> m(N);
            ok = make_group(h, kMom);
            const char* keys[3] = {kPx, kPy, kPz};
            for (int c = 0; c < 3 && ok; c++) {
                for (long n = 0; n < N; n++) m[n] = field(2 + c)[n] * field(0)[n];
                ok = write_array(h, keys[c], m.data(), cb, lv);
            }
        }
    }
    if (ok && mhd) {
        const char* keys[3] = {kBx, kBy, kBz};
        if (f.dim == 1) {  // 1D: cell B always, in double (HDF5Output.cpp:176-232)
            ok = make_group(h, kB);
            for (int c = 0; c < 3 && ok; c++) ok = write_array(h, keys[c], field(5 + c), cb, lv);
        } else {
            if (redundants) {  // cell B: float or double (HDF5Output.cpp:290-300, 364-370)
                ok = make_group(h, kB);
                for (int c = 0; c < 3 && ok; c++) {
                    if (f.redundant == DGH_WRITE_FLOAT) {
                        std::vector<float> b(field(5 + c), field(5 + c) + N);
                        ok = write_array(h, keys[c], b.data(), cb, lv);
                    } else {
                        ok = write_array(h, keys[c], field(5 + c), cb, lv);
                    }
                }
            }
            const long NF = fb.count();
            const char* fk[3] = {kBfx, kBfy, kBfz};
            ok = ok && make_group(h, kBf);
            for (int c = 0; c < 3 && ok; c++) ok = write_array(h, fk[c], faces + (long)c * NF, fb, lv);
        }
    }
    if (!ok) return e.set(std::string("dataset write failed: ") + path);
    if (H5Fflush(h, H5F_SCOPE_LOCAL) < 0) return e.set(std::string("flush failed: ") + path);
    return 0;
}

int dgh_read_meta(const char* path, dgh_frame* f, char* err, int nerr) {
    Err e{err, nerr};
    if (!path || !f) return e.set("dgh_read_meta: null argument");
    QuietHdf5 quiet;
    Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return e.set(std::string("cannot open ") + path);
    const hid_t h = file.id;
    std::memset(f, 0, sizeof *f);
    int fmt = 0;
    bool ok = read_int_attr(h, kFmt, &fmt) && read_int_attr(h, kDim, &f->dim) && read_int_attr(h, kMhd, &f->mhd) &&
              read_int_attr(h, kNx, &f->nx) && read_int_attr(h, kNg, &f->ng) && read_double_attr(h, kDx, &f->dx) &&
              read_double_attr(h, kTime, &f->time) && read_int_attr(h, kCyc, &f->cycle) &&
              read_int_attr(h, kWrtOpt, &f->write_option);
    if (ok && f->dim >= 2) ok = read_int_attr(h, kNy, &f->ny) && read_double_attr(h, kDy, &f->dy);
    if (ok && f->dim >= 3) ok = read_int_attr(h, kNz, &f->nz) && read_double_attr(h, kDz, &f->dz);
    if (!ok) return e.set(std::string("missing attributes in ") + path);
    if (fmt > 1) return e.set("Please update to the latest version of DRAGON to read this file");
    return 0;
}

int dgh_read(const char* path, const dgh_frame* gp, dgh_frame* got, double* cells, double* faces, char* err, int nerr) {
    Err e{err, nerr};
    if (!path || !gp || !got || !cells) return e.set("dgh_read: null argument");
    const dgh_frame& grid = *gp;
    char sub[256] = {0};
    if (dgh_read_meta(path, got, sub, sizeof sub)) return e.set(sub);
    // HDF5Restart.cpp:86-110, 164-192, 267-296: the same checks and messages
    if (got->dim != grid.dim) {
        const char* word = grid.dim == 1 ? "dimension" : "dimensions";
        return e.set("Expected " + std::to_string(grid.dim) + " " + word + " but got " + std::to_string(got->dim));
    }
    if (got->mhd && !grid.mhd) return e.set("File was saved using MHD. Please enable MHD in Config.h");
    if (got->nx != grid.nx || (grid.dim >= 2 && got->ny != grid.ny) || (grid.dim >= 3 && got->nz != grid.nz))
        return e.set("File size does not match the grid size");
    if (got->ng > grid.ng) got->ng = grid.ng;
    got->gamma = grid.gamma;
    const Box cb = cell_box(*got), fb = face_box(*got);
    const long N = cb.count(), NF = fb.count();
    QuietHdf5 quiet;
    Hid file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!file.ok()) return e.set(std::string("cannot open ") + path);
    const hid_t h = file.id;
    std::vector<double> a;
    auto take = [&](const char* key, double* dst, long count, bool required) -> int {
        if (!read_array(h, key, a)) {
            if (required) return e.set(std::string("dataset missing: ") + key);
            std::fill(dst, dst + count, 0.0);  // the reference's try/catch: zeros
            return 0;
        }
        if ((long)a.size() < count) return e.set(std::string("dataset too small: ") + key);
        // a file written with more ghosts than the grid holds: its arrays are
        // larger; the reference indexes them with the clamped ng (HDF5Restart.cpp:101-120)
        std::copy(a.begin(), a.begin() + count, dst);
        return 0;
    };
    if (take(kRho, cells, N, true)) return 1;
    const bool mhd_grid = grid.mhd != 0;
    for (int c = 0; c < 3; c++) {
        double* dst = cells + (long)(5 + c) * N;
        const char* keys[3] = {kBx, kBy, kBz};
        if (mhd_grid && got->mhd) {
            if (take(keys[c], dst, N, grid.dim == 1)) return 1;
        } else {
            std::fill(dst, dst + N, 0.0);
        }
        if (faces) {
            const char* fk[3] = {kBfx, kBfy, kBfz};
            double* fd = faces + (long)c * NF;
            if (mhd_grid && got->mhd && grid.dim >= 2) {
                if (take(fk[c], fd, NF, false)) return 1;
            } else {
                std::fill(fd, fd + NF, 0.0);
            }
        }
    }
    if (got->write_option & DGH_WRITE_PRIMITIVE) {
        if (take(kP, cells + 1 * N, N, true) || take(kVx, cells + 2 * N, N, true) || take(kVy, cells + 3 * N, N, true) ||
            take(kVz, cells + 4 * N, N, true))
            return 1;
    } else {
        // PrimitiveState(ConservativeState) (FluidElement.cpp:44-51, 60-66)
        std::vector<double> E(N), px(N), py(N), pz(N);
        if (take(kE, E.data(), N, true) || take(kPx, px.data(), N, true) || take(kPy, py.data(), N, true) ||
            take(kPz, pz.data(), N, true))
            return 1;
        for (long n = 0; n < N; n++) {
            const double rho = cells[n];
            const double vx = px[n] / rho, vy = py[n] / rho, vz = pz[n] / rho;
            double Et = E[n] - (px[n] * px[n] + py[n] * py[n] + pz[n] * pz[n]) / (2 * rho);
            if (mhd_grid) {
                const double bx = cells[5 * N + n], by = cells[6 * N + n], bz = cells[7 * N + n];
                Et -= (bx * bx + by * by + bz * bz) * (0.125 / M_PI);
            }
            cells[1 * N + n] = (grid.gamma - 1.0) * Et;
            cells[2 * N + n] = vx;
            cells[3 * N + n] = vy;
            cells[4 * N + n] = vz;
        }
    }
    return 0;
}

}  // extern "C"
