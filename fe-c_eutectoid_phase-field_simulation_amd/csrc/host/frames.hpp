// HDF5 frames, checkpoints and their XMF descriptions in the reference
// program's format (write_frame / read_checkpoint / write_initial_phases,
// pearl_serial.cpp:1843-2020, :2075-2115), so that each program restarts
// from the other's checkpoints and post-processing reads either's frames.
//
// A frame holds the five fields as [ny][nx] datasets (chunked, Fletcher-32
// checksummed) and scalar attributes.  The attribute list below is shared
// by the writer and the reader: names, order and the state field each one
// carries.  Files are written under a ".part" name and renamed when
// complete.
#pragma once

#include <hdf5.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "settings.hpp"

namespace pearl_host {

// Host copy of a state: planar phi [3][ny][nx], carbon and mu [ny][nx], and
// the State scalars in the C ABI's layout.
struct HostState {
    int nx = 0, ny = 0;
    std::vector<double> phi, carbon, mu;
    pearl_scalars sc{};
    std::size_t cells() const { return static_cast<std::size_t>(nx) * ny; }
    void allocate(int x, int y) {
        nx = x;
        ny = y;
        phi.assign(3 * cells(), 0.0);
        carbon.assign(cells(), 0.0);
        mu.assign(cells(), 0.0);
    }
    const double* plane(int a) const { return phi.data() + a * cells(); }
};

namespace h5 {

inline void ok(herr_t status, const char* what) {
    if (status < 0) throw std::runtime_error(std::string("HDF5: ") + what);
}

// An HDF5 identifier closed on scope exit.
class Handle {
public:
    Handle(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {
        if (id < 0) throw std::runtime_error("HDF5 open/create failure");
    }
    ~Handle() { close_(id_); }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    operator hid_t() const { return id_; }

private:
    hid_t id_;
    herr_t (*close_)(hid_t);
};

inline void put_number(hid_t file, const char* name, double v) {
    Handle space(H5Screate(H5S_SCALAR), H5Sclose);
    Handle attr(H5Acreate2(file, name, H5T_NATIVE_DOUBLE, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    ok(H5Awrite(attr, H5T_NATIVE_DOUBLE, &v), "write scalar");
}

inline double get_number(hid_t file, const char* name) {
    Handle attr(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
    double v = 0;
    ok(H5Aread(attr, H5T_NATIVE_DOUBLE, &v), "read scalar");
    return v;
}

inline void put_text(hid_t file, const char* name, const std::string& v) {
    Handle type(H5Tcopy(H5T_C_S1), H5Tclose);
    ok(H5Tset_size(type, v.size() + 1), "string size");
    Handle space(H5Screate(H5S_SCALAR), H5Sclose);
    Handle attr(H5Acreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    ok(H5Awrite(attr, type, v.c_str()), "string");
}

inline std::string get_text(hid_t file, const char* name) {
    Handle attr(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
    Handle type(H5Aget_type(attr), H5Tclose);
    std::vector<char> buf(H5Tget_size(type) + 1, 0);
    ok(H5Aread(attr, type, buf.data()), "read string");
    return buf.data();
}

inline void put_field(hid_t file, const char* name, const double* v, int nx, int ny) {
    const hsize_t dims[2] = {static_cast<hsize_t>(ny), static_cast<hsize_t>(nx)};
    const hsize_t chunk[2] = {static_cast<hsize_t>(std::min(ny, 64)), static_cast<hsize_t>(std::min(nx, 64))};
    Handle space(H5Screate_simple(2, dims, nullptr), H5Sclose);
    Handle props(H5Pcreate(H5P_DATASET_CREATE), H5Pclose);
    ok(H5Pset_chunk(props, 2, chunk), "chunk");
    ok(H5Pset_fletcher32(props), "checksum");
    Handle set(H5Dcreate2(file, name, H5T_IEEE_F64LE, space, H5P_DEFAULT, props, H5P_DEFAULT), H5Dclose);
    ok(H5Dwrite(set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v), "dataset write");
}

inline void get_field(hid_t file, const char* name, double* out, int nx, int ny) {
    Handle set(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
    Handle space(H5Dget_space(set), H5Sclose);
    if (H5Sget_simple_extent_ndims(space) != 2) throw std::runtime_error("dataset is not 2D");
    hsize_t dims[2] = {0, 0};
    H5Sget_simple_extent_dims(space, dims, nullptr);
    if (dims[0] != static_cast<hsize_t>(ny) || dims[1] != static_cast<hsize_t>(nx))
        throw std::runtime_error("checkpoint dimensions mismatch");
    ok(H5Dread(set, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out), "dataset read");
}

// Writes `path` through a ".part" file that is renamed when `fill` is done.
template <class F>
void write_atomically(const std::filesystem::path& path, F&& fill) {
    std::filesystem::path part = path;
    part += ".part";
    {
        Handle file(H5Fcreate(part.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        fill(static_cast<hid_t>(file));
        ok(H5Fflush(file, H5F_SCOPE_GLOBAL), "flush");
    }
    std::filesystem::rename(part, path);
}

}  // namespace h5

// The XMF companion of an HDF5 file: a 2D co-rectilinear grid whose node
// values are the listed datasets.
struct XmfGrid {
    std::string name;
    bool timed = false;
    double time = 0, origin_y = 0;
    std::vector<const char*> fields;
};

inline void write_xmf(const std::filesystem::path& h5file, const XmfGrid& g, int nx, int ny, double dx, double dy,
                      bool atomic, const char* failure) {
    std::filesystem::path xmf = h5file;
    xmf.replace_extension(".xmf");
    std::filesystem::path target = xmf;
    if (atomic) target += ".part";
    std::ofstream out(target);
    if (!out) throw std::runtime_error("cannot write XMF");
    out << std::setprecision(17) << "<?xml version=\"1.0\"?>\n<Xdmf Version=\"3.0\"><Domain><Grid Name=\"" << g.name
        << "\" GridType=\"Uniform\">\n";
    if (g.timed) out << "<Time Value=\"" << g.time << "\"/>\n";
    out << "<Topology TopologyType=\"2DCoRectMesh\" Dimensions=\"" << ny << " " << nx << "\"/>\n"
        << "<Geometry GeometryType=\"ORIGIN_DXDY\"><DataItem Dimensions=\"2\" Format=\"XML\">" << g.origin_y << " "
        << .5 * dx << "</DataItem><DataItem Dimensions=\"2\" Format=\"XML\">" << dy << " " << dx
        << "</DataItem></Geometry>\n";
    const std::string source = h5file.filename().string();
    for (const char* f : g.fields)
        out << "<Attribute Name=\"" << f << "\" AttributeType=\"Scalar\" Center=\"Node\"><DataItem Dimensions=\"" << ny
            << " " << nx << "\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">" << source << ":/" << f
            << "</DataItem></Attribute>\n";
    out << "</Grid></Domain></Xdmf>\n";
    out.close();
    if (!out) throw std::runtime_error(failure);
    if (atomic) std::filesystem::rename(target, xmf);
}

constexpr const char* kFrameFormat = "pearlite-v0.4-5fields";
constexpr const char* kFieldNames[5] = {"alpha", "theta", "gamma", "carbon", "mu"};

// The run's scalar attributes of a frame, in file order, with the state
// field each one holds (mu_reference is the run's, not the state's).
struct FrameScalars {
    pearl_scalars sc{};
    double mu_reference = 0;
};

struct ScalarAttr {
    const char* name;
    double (*get)(const FrameScalars&);
    void (*set)(FrameScalars&, double);
};

inline const std::array<ScalarAttr, 12>& frame_attributes() {
    static const std::array<ScalarAttr, 12> table = {{
        {"time", [](const FrameScalars& f) { return f.sc.time; }, [](FrameScalars& f, double v) { f.sc.time = v; }},
        {"dt_next", [](const FrameScalars& f) { return f.sc.dt; }, [](FrameScalars& f, double v) { f.sc.dt = v; }},
        {"step", [](const FrameScalars& f) { return static_cast<double>(f.sc.step); },
         [](FrameScalars& f, double v) { f.sc.step = static_cast<long long>(v); }},
        {"dx", nullptr, nullptr},
        {"dy", nullptr, nullptr},
        {"nx", nullptr, nullptr},
        {"ny", nullptr, nullptr},
        {"shift_distance", [](const FrameScalars& f) { return f.sc.shift_distance; },
         [](FrameScalars& f, double v) { f.sc.shift_distance = v; }},
        {"shifts", [](const FrameScalars& f) { return static_cast<double>(f.sc.shifts); },
         [](FrameScalars& f, double v) { f.sc.shifts = static_cast<long long>(v); }},
        {"initial_carbon", [](const FrameScalars& f) { return f.sc.initial_carbon; },
         [](FrameScalars& f, double v) { f.sc.initial_carbon = v; }},
        {"carbon_added", [](const FrameScalars& f) { return f.sc.carbon_added; },
         [](FrameScalars& f, double v) { f.sc.carbon_added = v; }},
        {"carbon_removed", [](const FrameScalars& f) { return f.sc.carbon_removed; },
         [](FrameScalars& f, double v) { f.sc.carbon_removed = v; }},
    }};
    return table;
}

// The five fields, the physics signature and the scalars (write_frame).
inline void write_frame(const std::filesystem::path& path, const HostState& st, const Settings& cfg,
                        double mu_reference) {
    if (std::filesystem::exists(path)) throw std::runtime_error("refuse overwriting frame " + path.string());
    const FrameScalars fs{st.sc, mu_reference};
    h5::write_atomically(path, [&](hid_t file) {
        for (int a = 0; a < 3; ++a) h5::put_field(file, kFieldNames[a], st.plane(a), st.nx, st.ny);
        h5::put_field(file, kFieldNames[3], st.carbon.data(), st.nx, st.ny);
        h5::put_field(file, kFieldNames[4], st.mu.data(), st.nx, st.ny);
        h5::put_text(file, "physics_signature", cfg.signature());
        h5::put_text(file, "format", kFrameFormat);
        h5::put_text(file, "thermo_mode", cfg.thermo_mode);
        h5::put_number(file, "temperature_K", cfg.T);
        if (cfg.thermo_mode == "eutectoid_reference") {
            h5::put_number(file, "T_eutectoid_K", cfg.T_eutectoid);
            h5::put_number(file, "undercooling_K", cfg.undercooling);
        }
        const double geometry[4] = {cfg.dx, cfg.dy, static_cast<double>(cfg.nx), static_cast<double>(cfg.ny)};
        int g = 0;
        for (const ScalarAttr& a : frame_attributes()) h5::put_number(file, a.name, a.get ? a.get(fs) : geometry[g++]);
        h5::put_number(file, "mu_reference", mu_reference);
        h5::put_number(file, "rejected_steps", static_cast<double>(st.sc.rejected));
    });
    write_xmf(path, {"pearlite", true, st.sc.time, .5 * cfg.dy + st.sc.shift_distance, {kFieldNames, kFieldNames + 5}},
              cfg.nx, cfg.ny, cfg.dx, cfg.dy, true, "XMF write failure");
}

// The three phase fields of an initialization stage (write_initial_phases).
inline void write_phases(const std::filesystem::path& path, const HostState& st, const Settings& cfg,
                         const std::string& stage, int pseudo_step) {
    if (std::filesystem::exists(path)) throw std::runtime_error("initial snapshot exists");
    std::filesystem::path part = path;
    part += ".part";
    {
        h5::Handle file(H5Fcreate(part.c_str(), H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        for (int a = 0; a < 3; ++a) h5::put_field(file, kFieldNames[a], st.plane(a), st.nx, st.ny);
        h5::put_text(file, "format", "pearlite-preprocessing-phases-only");
        h5::put_text(file, "stage", stage);
        h5::put_number(file, "pseudo_step", pseudo_step);
        h5::put_number(file, "dx", cfg.dx);
        h5::put_number(file, "dy", cfg.dy);
        h5::ok(H5Fflush(file, H5F_SCOPE_GLOBAL), "flush initial snapshot");
    }
    std::filesystem::rename(part, path);
    write_xmf(path, {stage, false, 0.0, .5 * cfg.dy, {kFieldNames, kFieldNames + 3}}, cfg.nx, cfg.ny, cfg.dx, cfg.dy,
              false, "initial XMF write failure");
}

// Fields and scalars of a checkpoint of this run's physics.  The format,
// signature, shapes and the scalars' finiteness and ranges are checked here;
// the state's own checks (model range, closure, carbon ledger) follow once
// it is on the device.  Returns the run's mu_reference.
inline double read_frame(const std::string& path, const Settings& cfg, HostState& st) {
    st.allocate(cfg.nx, cfg.ny);
    FrameScalars fs{};
    double rejected = 0;
    {
        h5::Handle file(H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        if (h5::get_text(file, "format") != kFrameFormat || h5::get_text(file, "physics_signature") != cfg.signature())
            throw std::runtime_error("checkpoint physics signature mismatch");
        for (int a = 0; a < 3; ++a) h5::get_field(file, kFieldNames[a], st.phi.data() + a * st.cells(), cfg.nx, cfg.ny);
        h5::get_field(file, kFieldNames[3], st.carbon.data(), cfg.nx, cfg.ny);
        h5::get_field(file, kFieldNames[4], st.mu.data(), cfg.nx, cfg.ny);
        std::vector<std::pair<const ScalarAttr*, double>> values;
        for (const ScalarAttr& a : frame_attributes())
            if (a.set) values.emplace_back(&a, h5::get_number(file, a.name));
        fs.mu_reference = h5::get_number(file, "mu_reference");
        rejected = h5::get_number(file, "rejected_steps");
        bool finite = std::isfinite(fs.mu_reference) && std::isfinite(rejected);
        for (const auto& v : values) finite = finite && std::isfinite(v.second);
        if (!finite) throw std::runtime_error("nonfinite checkpoint metadata");
        const double time = values[0].second, dt = values[1].second, step = values[2].second;
        if (time < 0 || dt <= 0 || step < 0 || step != std::floor(step) || step > 9e15)
            throw std::runtime_error("bad checkpoint time/step");
        for (const auto& v : values) v.first->set(fs, v.second);
    }
    st.sc = fs.sc;
    st.sc.dt = std::min(cfg.dt_max, fs.sc.dt);
    st.sc.rejected = static_cast<long long>(rejected);
    return fs.mu_reference;
}

}  // namespace pearl_host
