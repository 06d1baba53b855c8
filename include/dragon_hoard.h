/* dragon_hoard.h -- C ABI of the frame / restart files (the reference's
 * DRAGONHOARD module) for grids that live on the GPU.
 *
 * The file format is the reference's, key for key and type for type
 * (DRAGONHOARD/HDF5_Attrs.hpp:14-59, HDF5Output.cpp:123-525,
 * HDF5Restart.cpp:81-374): attributes format_version, write_option, dim, MHD,
 * ng, nx[, ny[, nz]], dx[, dy[, dz]], cycle, time on the root group;
 * datasets /fluid/rho, /fluid/v/{x,y,z}, /fluid/p, /fluid/E,
 * /fluid/momentum/{x,y,z}, /fluid/B/{x,y,z} (cell B), /fluid/Bf/{x,y,z}
 * (face B, 2D/3D), each x-fastest with the dataspace dims {nz, ny, nx},
 * chunked and deflated as HDF5Output.cpp:46-90 does.
 *
 * The arrays handed to these functions are in that file order: field f of
 * element n at base[f * count + n], n = ((k - k0) * NY + (j - j0)) * NX + (i - i0)
 * over the cells (or faces) written, i0 = -ng.  Cells carry 8 fields (rho, p,
 * vx, vy, vz, Bx, By, Bz: the PrimitiveState order, FluidElement.hpp:34-43),
 * faces 3 (Bx, By, Bz on the (n+1)-sized index space; 2D .z is the
 * cell-centred Bz).  dg_snapshot / dg_restore (dragon_gpu.h) move them between
 * these arrays and the device state.
 *
 * Host code only (HDF5 1.10 C library); no device calls.  Every function
 * returns 0 or a nonzero error with its text in err.
 */
#ifndef DRAGON_HOARD_H
#define DRAGON_HOARD_H

#ifdef __cplusplus
extern "C" {
#endif

/* HDF5_WRITE_OPTION (DRAGONHOARD_Config.h:25-29) */
enum { DGH_WRITE_PRIMITIVE = 1, DGH_WRITE_CONSERVATIVE = 6, DGH_WRITE_PRIMITIVE_AND_ENERGY = 3,
       DGH_WRITE_PRIMITIVE_AND_CONSERVATIVE = 7 };
/* HDF5_REDUNDANT_VALS_OPTION (DRAGONHOARD_Config.h:34-38): cell B (2D/3D) and E (primitive + energy) */
enum { DGH_WRITE_OMIT = 0, DGH_WRITE_FLOAT = 1, DGH_WRITE_DOUBLE = 2 };

typedef struct dgh_frame {
    int dim;            /* 1, 2, 3 */
    int mhd;            /* #define MHD */
    int nx, ny, nz;     /* interior cells */
    int ng;             /* ghost layers in the arrays and the file (WRITE_GHOSTS_TO_FILE; 0 = interior only) */
    double dx, dy, dz;
    double gamma;       /* Constants.h:18; E = PrimitiveState::energy() (FluidElement.cpp:53-59) */
    double time;
    int cycle;
    int write_option;   /* DGH_WRITE_* */
    int redundant;      /* DGH_WRITE_OMIT / FLOAT / DOUBLE */
    int compression;    /* HDF5_COMPRESSION_LEVEL (DRAGONHOARD_Config.h:22); <= 0: no chunking or deflate */
} dgh_frame;

/* element counts of the cell and face arrays for the frame's dim, sizes and ng */
long dgh_cell_count(const dgh_frame* f);
long dgh_face_count(const dgh_frame* f);

/* DRAGONHOARD::writeToFile(grid, t, cycle, filename) (HDF5Output.cpp:103-525),
 * the file path given in full (output_dir + "/" + name + ".h5"). */
int dgh_write(const char* path, const dgh_frame* f, const double* cells, const double* faces, char* err, int nerr);

/* The file's metadata: dim, MHD, sizes, ng, dx.., time, cycle, write_option
 * (HDF5Restart.cpp:86-110). */
int dgh_read_meta(const char* path, dgh_frame* f, char* err, int nerr);

/* DRAGONHOARD::loadFromFile(grid, t, cycle, filename) (HDF5Restart.cpp:61-374):
 * `grid` holds the receiving grid (dim, mhd, sizes, ng = its ghost count,
 * gamma); the reference's checks and messages apply.  On return `got` holds
 * the file's metadata with got->ng clamped to grid->ng; cells / faces (sized
 * for got->ng) hold rho, the cell and face B (zeros where the file has none or
 * was written without MHD), and p and v from the primitive datasets or, for
 * files with only conservative values, from E and the momentum
 * (PrimitiveState(ConservativeState), FluidElement.cpp:44-51). */
int dgh_read(const char* path, const dgh_frame* grid, dgh_frame* got, double* cells, double* faces, char* err, int nerr);

#ifdef __cplusplus
}
#endif
#endif
