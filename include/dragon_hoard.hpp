// dragon_hoard.hpp -- C++ mirror of the reference's frame / restart API
// (DRAGONHOARD/DragonHoard.hpp: writeToFile, loadFromFile, cycle_string) for
// grids that live on the GPU (header-only; link libdragon_gpu and
// libdragon_hoard).
//
//   DRAGONHOARD::writeToFile(grid, t, cycle, name)   -> dragon_gpu::hoard::writeToFile(grid, t, cycle, name, opts)
//   DRAGONHOARD::loadFromFile(grid, t, cycle, name)  -> dragon_gpu::hoard::loadFromFile(grid, t, cycle, name, opts)
//
// The DRAGONHOARD_Config.h switches are the fields of Options, with the
// reference's defaults.  The grid's state is copied to the host in the
// file's x-fastest order on the GPU (dg_snapshot) and written by the HDF5 C
// library (dgh_write); a restart reads the file (dgh_read) and copies it
// into the grid (dg_restore), together with its dx, dy, dz.  Failures throw
// HoardError with the reference's message where it has one.
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "dragon_gpu.hpp"
#include "dragon_hoard.h"

namespace dragon_gpu {
namespace hoard {

// DRAGONHOARD_Config.h:16-37
struct Options {
    std::string output_dir;                             // output_dir
    int write_option = DGH_WRITE_PRIMITIVE_AND_ENERGY;  // HDF5_WRITE_OPTION
    int redundant = DGH_WRITE_FLOAT;                    // HDF5_REDUNDANT_VALS_OPTION
    int compression = 4;                                // HDF5_COMPRESSION_LEVEL
    bool write_ghosts = false;                          // WRITE_GHOSTS_TO_FILE
};

class HoardError : public std::runtime_error {
public:
    explicit HoardError(const std::string& m) : std::runtime_error(m) {}
};

// HDF5Output.cpp:33-36: the cycle number zero-padded to five digits
inline std::string cycle_string(int cycle) {
    char b[32];
    std::snprintf(b, sizeof b, "%05d", cycle);
    return b;
}
// HDF5Restart.cpp:54-58: ".h5" appended unless the name already ends in it
inline std::string check_extension(const std::string& name) {
    const std::string ext = ".h5";
    if (name.size() > ext.size() && name.compare(name.size() - ext.size(), ext.size(), ext) == 0) return name;
    return name + ext;
}

inline dgh_frame frame_of(const GpuGrid& g, double t, int cycle, int ng, const Options& o) {
    dgh_frame f{};
    f.dim = g.dim();
    f.mhd = g.config().mhd ? 1 : 0;
    f.nx = g.getSizeX();
    f.ny = g.getSizeY();
    f.nz = g.getSizeZ();
    f.ng = ng;
    f.dx = g.spacing(0);
    f.dy = g.spacing(1);
    f.dz = g.spacing(2);
    f.gamma = g.config().gamma;
    f.time = t;
    f.cycle = cycle;
    f.write_option = o.write_option;
    f.redundant = o.redundant;
    f.compression = o.compression;
    return f;
}

// DRAGONHOARD::writeToFile (HDF5Output.cpp:103-525): output_dir/filename.h5
inline void writeToFile(GpuGrid& grid, double t, int cycle, const std::string& filename, const Options& o = Options()) {
    const dgh_frame f = frame_of(grid, t, cycle, o.write_ghosts ? grid.getGhosts() : 0, o);
    std::vector<double> cells(8 * (size_t)dgh_cell_count(&f));
    std::vector<double> faces(f.mhd && f.dim >= 2 ? 3 * (size_t)dgh_face_count(&f) : 0);
    int rc = dg_snapshot(grid.handle(), f.ng, cells.data(), faces.empty() ? nullptr : faces.data());
    if (!rc) rc = dg_snapshot_wait(grid.handle());
    if (rc) throw GpuError(rc, dg_last_error(grid.handle()));
    char err[512] = {0};
    const std::string path = o.output_dir + "/" + filename + ".h5";
    if (dgh_write(path.c_str(), &f, cells.data(), faces.empty() ? nullptr : faces.data(), err, (int)sizeof err))
        throw HoardError(err);
}

// DRAGONHOARD::loadFromFile (HDF5Restart.cpp:61-374): the grid takes the
// file's cells (and face B) and its dx, dy, dz; t and cycle are returned.
inline void loadFromFile(GpuGrid& grid, double& t, int& cycle, const std::string& filename, const Options& o = Options()) {
    const std::string path = o.output_dir + "/" + check_extension(filename);
    char err[512] = {0};
    dgh_frame meta{};
    if (dgh_read_meta(path.c_str(), &meta, err, (int)sizeof err)) throw HoardError(err);
    const dgh_frame want = frame_of(grid, 0.0, 0, grid.getGhosts(), o);
    dgh_frame probe = want;
    probe.ng = meta.ng < want.ng ? meta.ng : want.ng;  // the reader keeps min(file ghosts, grid ghosts)
    std::vector<double> cells(8 * (size_t)dgh_cell_count(&probe));
    std::vector<double> faces(want.mhd && want.dim >= 2 ? 3 * (size_t)dgh_face_count(&probe) : 0);
    dgh_frame got{};
    if (dgh_read(path.c_str(), &want, &got, cells.data(), faces.empty() ? nullptr : faces.data(), err, (int)sizeof err))
        throw HoardError(err);
    grid.set_spacing(got.dx, got.dy, got.dz);
    const int rc = dg_restore(grid.handle(), got.ng, cells.data(), faces.empty() ? nullptr : faces.data());
    if (rc) throw GpuError(rc, dg_last_error(grid.handle()));
    t = got.time;
    cycle = got.cycle;
}

}  // namespace hoard
}  // namespace dragon_gpu
