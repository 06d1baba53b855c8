// dragon_gpu.hpp -- C++ mirror of the reference's Grid / Boundary / Config.h
// interface over the C ABI of dragon_gpu.h (header-only).
//
// What a reference user writes stays the same, with the Gpu prefix:
//   Grid2D g(nx, ny, dx, dy, ghosts)           -> dragon_gpu::GpuGrid2D g(nx, ny, dx, dy, ghosts, cfg)
//   g[i, j] = PrimitiveState{...}               -> g(i, j) = dragon_gpu::State{...}; g.push();
//   g.boundary = Outflow(X) + Periodic(Y)      -> g.boundary = Outflow(X) + Periodic(Y)
//   g.initialize_B_fields(); g.advance(dt)      -> same names
// (Hydro/Grid.hpp:19-144, Boundary/Boundary.hpp:160-229, BoundaryList.hpp,
// Hydro/CFL/CFL.cpp:151-238).  Failures throw std::runtime_error carrying the
// reference's message, as unsplit_step does (Unsplit.cpp:97, 267; Flux.cpp:94-95;
// CFL.cpp:127); advance() halves dt on failure like advance_unsplit.
//
// The state lives on the GPU.  Host access goes through a mirror: pull()
// downloads it, push() uploads edits (the reference reads grid[i,j] directly).
#pragma once
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "dragon_gpu.h"

namespace dragon_gpu {

// Config.h switches (production defaults, Config.h:27-90).
struct Config {
    int mhd = 1;
    double gamma = 5.0 / 3.0;
    int solver = DG_HLLX;
    int limiter = DG_MC;
    int muscl = 1;
    int ctu = 1;
    int ct_energy = DG_CT_BETA_GATED;
    double ct_beta = 0.1;
    int cfl_mode = DG_CFL_ADD;
    double cfl_coeff = 0.3;
    int verify_fallback = -1;
    int fallback_try_hlle = 0;
    double fallback_param = 1.0;
    int hlld_safety = 1;
    double timestep_tol = 1e-14;
    int precision = DG_FP64;
    int dimension_unsplit = 1;  // DIMENSION_UNSPLIT: advance() runs advance_unsplit, else advance_split (hydro)
};

// PrimitiveState record (FluidElement.hpp:34-43).
struct State {
    double rho = 0, p = 0, vx = 0, vy = 0, vz = 0, Bx = 0, By = 0, Bz = 0;
};
struct Vec3 {
    double x = 0, y = 0, z = 0;
};

class GpuError : public std::runtime_error {
public:
    GpuError(int code, const std::string& m) : std::runtime_error(m), code(code) {}
    int code;
};

// Face masks (Boundary.hpp:139-149).
enum Faces { X_negative = 1, X_positive = 2, Y_negative = 4, Y_positive = 8, Z_negative = 16, Z_positive = 32,
             X = 3, Y = 12, Z = 48 };

// face_mask("XY+", "LR", ...) (Boundary/GhostFill.cpp:14-49).
inline int face_mask(const std::string& s) {
    int m = 0;
    for (size_t q = 0; q < s.size(); q++) {
        const char c = s[q], nx = q + 1 < s.size() ? s[q + 1] : ' ';
        if (c == 'X' || c == 'Y' || c == 'Z' || c == 'x' || c == 'y' || c == 'z') {
            const int sh = 2 * ((c | 32) - 'x');
            m |= nx == '+' ? (2 << sh) : nx == '-' ? (1 << sh) : (3 << sh);
        } else if (c == 'L' || c == 'l') m |= X_negative;
        else if (c == 'R' || c == 'r') m |= X_positive;
        else if (c == 'U' || c == 'u') m |= Y_positive;
        else if (c == 'D' || c == 'd') m |= Y_negative;
    }
    return m;
}

// Boundary elements and lists (Boundary.hpp:160-229, BoundaryList.hpp:22-128).
struct GhostFill {
    dg_bc bc{};
    int get_faces() const { return bc.faces; }
};
inline GhostFill make_fill(int kind, int faces, bool corners, bool gated = false) {
    GhostFill g;
    g.bc.kind = kind;
    g.bc.faces = kind == DG_BC_PERIODIC ? (((faces & 3) ? 3 : 0) | ((faces & 12) ? 12 : 0) | ((faces & 48) ? 48 : 0)) : faces;
    g.bc.corners = corners ? 1 : 0;
    g.bc.gated = gated ? 1 : 0;
    return g;
}
struct Outflow : GhostFill {
    explicit Outflow(int faces = X | Y | Z, bool corner_ghosts = true, bool gated = false) {
        bc = make_fill(DG_BC_OUTFLOW, faces, corner_ghosts, gated).bc;
    }
    explicit Outflow(const std::string& f, bool corner_ghosts = true, bool gated = false) : Outflow(face_mask(f), corner_ghosts, gated) {}
    static Outflow Gated(int faces = X | Y | Z, bool corner_ghosts = true) { return Outflow(faces, corner_ghosts, true); }
    static Outflow Gated(const std::string& f, bool corner_ghosts = true) { return Outflow(face_mask(f), corner_ghosts, true); }
};
struct Periodic : GhostFill {
    explicit Periodic(int faces = X | Y | Z, bool corner_ghosts = true) { bc = make_fill(DG_BC_PERIODIC, faces, corner_ghosts).bc; }
    explicit Periodic(const std::string& f, bool corner_ghosts = true) : Periodic(face_mask(f), corner_ghosts) {}
};
struct Reflective : GhostFill {
    explicit Reflective(int faces = X | Y | Z, bool corner_ghosts = true) { bc = make_fill(DG_BC_REFLECTIVE, faces, corner_ghosts).bc; }
    explicit Reflective(const std::string& f, bool corner_ghosts = true) : Reflective(face_mask(f), corner_ghosts) {}
};
struct Fixed : GhostFill {
    explicit Fixed(const State& w, int faces = X | Y | Z, bool corner_ghosts = true) {
        bc = make_fill(DG_BC_FIXED, faces, corner_ghosts).bc;
        const double s[8] = {w.rho, w.p, w.vx, w.vy, w.vz, w.Bx, w.By, w.Bz};
        for (int q = 0; q < 8; q++) bc.state[q] = s[q];
    }
    Fixed(const State& w, const std::string& f, bool corner_ghosts = true) : Fixed(w, face_mask(f), corner_ghosts) {}
};
struct Ignore : GhostFill {
    explicit Ignore(int faces = X | Y | Z) { bc = make_fill(DG_BC_IGNORE, faces, true).bc; }
    explicit Ignore(const std::string& f) : Ignore(face_mask(f)) {}
};
class GpuGrid;
// A boundary element evaluated on the host: the counterpart of a user
// GhostFill subclass (Boundary/GhostFill.hpp:16-30; e.g. the reference's Jet,
// Examples/CollidingFlows/Jets/Jets.hpp:12-39).  apply() edits the ghost
// entries of grid.cell(...) / grid.face(...); the grid downloads its state
// before and uploads it after, keeping the list's order with the device
// elements around it.  Not usable with the split step (its boundary passes
// run on the device between sweeps).
struct HostGhostFill {
    int faces;
    bool corners;
    explicit HostGhostFill(int faces_ = X | Y | Z, bool corner_ghosts = true) : faces(faces_), corners(corner_ghosts) {}
    virtual ~HostGhostFill() = default;
    int get_faces() const { return faces; }
    virtual void apply(GpuGrid& grid) = 0;
};
struct BoundaryList {
    std::vector<dg_bc> items;
    std::vector<std::shared_ptr<HostGhostFill>> host;  // per item: null for a device element
    BoundaryList() = default;
    template <class... T>
    explicit BoundaryList(const T&... e) { (add(e), ...); }
    BoundaryList& operator+=(const GhostFill& e) { add(e); return *this; }
    BoundaryList& operator+=(std::shared_ptr<HostGhostFill> e) { add(std::move(e)); return *this; }
    BoundaryList& operator+=(const BoundaryList& o) {
        items.insert(items.end(), o.items.begin(), o.items.end());
        host.insert(host.end(), o.host.begin(), o.host.end());
        return *this;
    }
    BoundaryList(const GhostFill& e) { add(e); }  // NOLINT: implicit, as the reference's
    bool has_host() const {
        for (const auto& h : host)
            if (h) return true;
        return false;
    }

private:
    void add(const GhostFill& e) { items.push_back(e.bc); host.push_back(nullptr); }
    void add(std::shared_ptr<HostGhostFill> e) {
        dg_bc b{};
        b.kind = DG_BC_IGNORE;
        b.faces = e->faces;
        items.push_back(b);
        host.push_back(std::move(e));
    }
};
inline BoundaryList operator+(BoundaryList a, const GhostFill& b) { a += b; return a; }
inline BoundaryList operator+(BoundaryList a, std::shared_ptr<HostGhostFill> b) { a += std::move(b); return a; }
inline BoundaryList operator+(const GhostFill& a, const GhostFill& b) { BoundaryList l(a); l += b; return l; }

// Grid::advance / advance_unsplit / unsplit_step / on_step_fail / initialize_B_fields.
class GpuGrid {
public:
    BoundaryList boundary;

    GpuGrid(int dim, const int n[3], const double d[3], int ghosts, const Config& c) : dim_(dim) {
        dg_desc desc{};
        desc.dim = dim;
        desc.nx = n[0]; desc.ny = n[1]; desc.nz = n[2];
        desc.dx = d[0]; desc.dy = d[1]; desc.dz = d[2];
        desc.ghosts = ghosts;
        desc.mhd = c.mhd; desc.gamma = c.gamma; desc.solver = c.solver; desc.limiter = c.limiter; desc.muscl = c.muscl;
        desc.ctu = c.ctu; desc.ct_energy = c.ct_energy; desc.ct_beta = c.ct_beta; desc.cfl_mode = c.cfl_mode;
        desc.cfl_coeff = c.cfl_coeff; desc.verify_fallback = c.verify_fallback; desc.fallback_try_hlle = c.fallback_try_hlle;
        desc.fallback_param = c.fallback_param; desc.hlld_safety = c.hlld_safety; desc.timestep_tol = c.timestep_tol;
        desc.precision = c.precision;
        dg_grid* h = nullptr;
        const int rc = dg_create(&desc, &h);
        if (rc) throw GpuError(rc, "dg_create failed: invalid grid description or device error");
        h_.reset(h);
        g_ = dg_ghosts(h);
        for (int q = 0; q < 3; q++) n_[q] = q < dim ? n[q] : 1;
        for (int q = 0; q < 3; q++) d_[q] = q < dim ? d[q] : 1.0;
        cfg_ = c;
        has_B_ = c.mhd && dim >= 2;
        w_.assign(dg_cell_count(h) * 8, 0.0);
        if (has_B_) B_.assign(dg_face_count(h) * 3, 0.0);
        tol_ = c.timestep_tol;
        unsplit_ = c.dimension_unsplit != 0;
    }
    virtual ~GpuGrid() = default;

    int getGhosts() const { return g_; }
    int getSizeX() const { return n_[0]; }
    int getSizeY() const { return n_[1]; }
    int getSizeZ() const { return n_[2]; }
    int dim() const { return dim_; }
    // cell spacing per dimension (grid.dx / dy / dz)
    double spacing(int q) const { return d_[q]; }
    // grid.dx = ... (as loadFromFile sets them)
    void set_spacing(double dx, double dy, double dz) {
        check(dg_set_spacing(h_.get(), dx, dy, dz));
        d_[0] = dx;
        if (dim_ >= 2) d_[1] = dy;
        if (dim_ >= 3) d_[2] = dz;
    }
    const Config& config() const { return cfg_; }
    dg_grid* handle() const { return h_.get(); }

    // host mirror of grid[i,j,k] and _B()[i,j,k] (ghost-offset indices)
    State& cell(int i, int j = 0, int k = 0) { return *reinterpret_cast<State*>(&w_[8 * lin(i, j, k, 0)]); }
    Vec3& face(int i, int j = 0, int k = 0) { return *reinterpret_cast<Vec3*>(&B_[3 * lin(i, j, k, 1)]); }
    void pull() {
        check(dg_download(h_.get(), w_.data(), 8, has_B_ ? B_.data() : nullptr));
    }
    void push() {
        check(dg_upload(h_.get(), w_.data(), 8, has_B_ ? B_.data() : nullptr));
    }

    // BoundaryList::apply (BoundaryList.hpp:79-104): implicit Outflow on the
    // uncovered faces, then the elements in order
    void apply_boundary() {
        push_bc();
        check(dg_apply_boundary(h_.get()));  // with host elements: the callback below runs apply_list()
    }

protected:
    // The list with host elements, in the reference's order; every boundary
    // pass of the library calls it through the boundary callback (the advance
    // loops, dg_init_B, the pass before each directional sweep of the split
    // step, which then acts on that sweep's input state).
    void apply_list() {
        int covered = 0;
        for (const auto& b : boundary.items) covered |= b.faces;
        bool fresh = false, dirty = false;  // host mirror current / edited since the download
        auto device = [&](const dg_bc& b) {
            if (dirty) { push(); dirty = false; }
            fresh = false;
            check(dg_apply_boundary_element(h_.get(), &b));
        };
        device(Outflow(63 & ~covered).bc);
        for (size_t q = 0; q < boundary.items.size(); q++) {
            if (boundary.host[q]) {
                if (!fresh) { pull(); fresh = true; }
                boundary.host[q]->apply(*this);
                dirty = true;
            } else {
                device(boundary.items[q]);
            }
        }
        if (dirty) push();
    }
    static int bc_callback(void* user) {  // no exception may cross the C library
        GpuGrid* g = static_cast<GpuGrid*>(user);
        try {
            g->apply_list();
            return DG_OK;
        } catch (const GpuError& e) {
            g->cb_error_ = std::current_exception();
            return e.code ? e.code : DG_ERR_ARG;
        } catch (...) {
            g->cb_error_ = std::current_exception();
            return DG_ERR_ARG;
        }
    }

public:
    double cfl_time() {
        push_bc();
        double dt = 0;
        dg_status st{};
        check(dg_cfl_dt(h_.get(), &dt, &st), &st);
        return dt;
    }
    void initialize_B_fields() {
        push_bc();
        check(dg_init_B(h_.get()));  // its boundary pass runs the list (host elements through the callback)
    }
    // Grid::unsplit_step(dt): throws on failure, commits nothing
    virtual void unsplit_step(double dt) {
        dg_status st{};
        check(dg_step(h_.get(), dt, &st), &st);
    }
    // Grid::on_step_fail: true = retry with half the step (the reference's default)
    virtual bool on_step_fail(const std::exception&) { return true; }
    // Grid2D/3D::split_step(dt) (Splitting.cpp:118-216): Strang-split sweeps, hydro only
    virtual void split_step(double dt) {
        dg_status st{};
        check(dg_split_step(h_.get(), dt, &st), &st);
    }
    // Grid::advance_unsplit(dt, check_cfl) (CFL.cpp:213-238); returns accepted steps
    long advance_unsplit(double dt, bool check_cfl = true) { return advance_loop(dt, check_cfl, false); }
    // Grid::advance_split(dt, check_cfl) (CFL.cpp:160-184)
    long advance_split(double dt, bool check_cfl = true) { return advance_loop(dt, check_cfl, true); }
    // Grid::advance (CFL.cpp:151-157)
    long advance(double dt, bool check_cfl = true) { return unsplit_ ? advance_unsplit(dt, check_cfl) : advance_split(dt, check_cfl); }

protected:
    long advance_loop(double dt, bool check_cfl, bool split) {
        push_bc();
        long steps = 0;
        while (dt > tol_) {
            apply_boundary();
            double t1 = dt;
            if (check_cfl) {
                const double c = cfl_time();
                t1 = c < dt ? c : dt;
            }
            while (true) {
                try {
                    if (split) split_step(t1);
                    else unsplit_step(t1);
                    break;
                } catch (const GpuError& e) {
                    if (e.code == DG_ERR_HIP || e.code == DG_ERR_ARG || e.code == DG_ERR_UNSUPPORTED || !on_step_fail(e)) throw;
                    t1 *= 0.5;
                    if (t1 < tol_) throw GpuError(DG_ERR_DT_UNDERFLOW, "Timestep has fallen below minimum. Exiting");
                }
            }
            dt -= t1;
            steps++;
        }
        return steps;
    }
    long lin(int i, int j, int k, int extra) const {
        const long ex = n_[0] + extra + 2L * g_;
        const long ey = dim_ >= 2 ? n_[1] + extra + 2L * g_ : 1, ez = dim_ >= 3 ? n_[2] + extra + 2L * g_ : 1;
        (void)ex;
        return ((long)(i + g_) * ey + (dim_ >= 2 ? j + g_ : 0)) * ez + (dim_ >= 3 ? k + g_ : 0);
    }
    void push_bc() {
        if (boundary.has_host()) {  // the library's boundary passes call back into apply_list()
            dg_bc ign{};
            ign.kind = DG_BC_IGNORE;
            ign.faces = 63;
            check(dg_set_boundary(h_.get(), &ign, 1));
            check(dg_set_boundary_callback(h_.get(), &GpuGrid::bc_callback, this));
            return;
        }
        check(dg_set_boundary_callback(h_.get(), nullptr, nullptr));
        check(dg_set_boundary(h_.get(), boundary.items.data(), (int)boundary.items.size()));
    }
    // A failed call raises the exception a host element threw during it, if
    // one did, else GpuError with the library's message.
    void check(int rc, const dg_status* st = nullptr) {
        if (cb_error_) {
            std::exception_ptr e = cb_error_;
            cb_error_ = nullptr;
            std::rethrow_exception(e);
        }
        if (!rc) return;
        std::string m = st && st->msg[0] ? std::string(st->msg) : std::string(dg_last_error(h_.get()));
        throw GpuError(rc, m);
    }
    struct Del {
        void operator()(dg_grid* h) const { dg_destroy(h); }
    };
    std::unique_ptr<dg_grid, Del> h_;
    int dim_ = 1, g_ = 0, n_[3] = {1, 1, 1};
    double d_[3] = {1, 1, 1};
    Config cfg_;
    std::exception_ptr cb_error_;
    bool has_B_ = false;
    double tol_ = 1e-14;
    bool unsplit_ = true;
    std::vector<double> w_, B_;
};

struct GpuGrid1D : GpuGrid {
    GpuGrid1D(int nx, double dx, int ghosts = 1, const Config& c = Config())
        : GpuGrid(1, std::vector<int>{nx, 1, 1}.data(), std::vector<double>{dx, 1, 1}.data(), ghosts, c) {}
};
struct GpuGrid2D : GpuGrid {
    GpuGrid2D(int nx, int ny, double dx, double dy, int ghosts = 1, const Config& c = Config())
        : GpuGrid(2, std::vector<int>{nx, ny, 1}.data(), std::vector<double>{dx, dy, 1}.data(), ghosts, c) {}
    State& operator()(int i, int j) { return cell(i, j); }
};
struct GpuGrid3D : GpuGrid {
    GpuGrid3D(int nx, int ny, int nz, double dx, double dy, double dz, int ghosts = 1, const Config& c = Config())
        : GpuGrid(3, std::vector<int>{nx, ny, nz}.data(), std::vector<double>{dx, dy, dz}.data(), ghosts, c) {}
    State& operator()(int i, int j, int k) { return cell(i, j, k); }
};

}  // namespace dragon_gpu
