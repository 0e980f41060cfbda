// mflbm_run -- the simulation driver over the C ABI of include/mflbm.h.
//
// Runs a colour-gradient simulation from the reference's input files in the current
// directory, with the reference's main-loop semantics (src/main.cpp:30-366):
//   input/job_status.txt             new_simulation | continue_simulation
//   input/simulation_control.txt     key/value control file (src/IO_multiphase.cpp:16-246)
//   input/Geometry_File_Path.txt     geometry .dat path (src/Misc.cpp:17-51, 139-186)
// and writes the reference's outputs under results/: monitor series (src/Monitor.cpp:17-277),
// time.dat, info.txt, legacy VTK files (src/IO_multiphase.cpp:500-716, 800-837), the
// checkpoint results/out2.checkpoint/id0000 (src/IO_multiphase.cpp:252-305) and
// ./job_status.txt.
//
// Differences from the reference, each on purpose:
//   * a resumed run starts at the step stored in the checkpoint (the reference restarts one
//     step early, with the AA parity of the step already taken);
//   * the end-of-run checkpoint and VTK files hold the final state (the reference's host
//     copy may be from the last monitor step);
//   * VTK scalars are written in the type the header names (the reference declares
//     "float" but writes T_P when output_fieldData_precision_cmd is 0);
//   * the device state is copied to the host only for the outputs that need it;
//   * precision is a run-time option (--precision fp64|fp32) rather than a build switch;
//     a checkpoint of either precision can be resumed by either;
//   * so is the collision model (--mrt 1-4, the reference's `mrt` macro of
//     includes/preprocessor.h, default 2).
//   * the convective-outlet tail of the checkpoint is written when that outlet is active
//     (outlet_BC = 1 on an open z axis); the reference also writes it, unused, when
//     outlet_BC = 1 on a closed or periodic z axis.  Files with or without it load.
// Precomputed boundary files (geometry_preprocess_cmd = 1) stop with the reference's own
// error: its loader is not implemented either.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "mflbm.h"

namespace {

struct Fatal : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------------------
// Control file: first non-comment line whose first token is the key (src/utils.cpp:17-154)
// ---------------------------------------------------------------------------------------
class KeyFile {
public:
    explicit KeyFile(const std::string& path) : path_(path) {
        std::ifstream in(path);
        if (!in.good()) throw Fatal("Missing file " + path + "! Exiting program!");
        std::string line;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            lines_.push_back(line);
        }
    }
    std::string str(const std::string& key) const {
        for (const std::string& line : lines_) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ss(line);
            std::string k, v;
            ss >> k;
            if (k != key) continue;
            if (!(ss >> v)) throw Fatal("no value for (" + key + ") in " + path_);
            return v;
        }
        throw Fatal("ERROR! Couldn't find the variable named: (" + key + ") in the file " + path_);
    }
    bool has(const std::string& key) const {
        for (const std::string& line : lines_) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream ss(line);
            std::string k;
            ss >> k;
            if (k == key) return true;
        }
        return false;
    }
    long long i64(const std::string& key) const { return std::stoll(str(key)); }
    int i32(const std::string& key) const { return std::stoi(str(key)); }
    double real(const std::string& key) const { return std::stod(str(key)); }

private:
    std::string path_;
    std::vector<std::string> lines_;
};

struct Control {
    int option, benchmark, extreme_large, breakthrough_check, steady_option, field_precision;
    int modify_geometry, external_geometry, dims_type_size, geometry_preprocess, porous_plate;
    int change_inlet_phase, n_exclude_inlet, n_exclude_outlet;
    int wall[6];  // x min/max, y min/max, z min/max
    int iper, jper, kper, inlet_bc, outlet_bc;
    int max_step, ntime_visual, ntime_animation, ntime_monitor, monitor_profile_ratio;
    int ntime_clock_sum, ntime_display, rho_in_new, rho_out_bc, z_porous_plate;
    long long nx, ny, nz;
    double convergence, nu1, nu2, gamma, theta, beta, sa_inject, interface_z0, body_force, sa_target;
    uint64_t random_seed;
    double checkpoint_timer, checkpoint2_timer, duration_timer;
    double capillary_number, target_inject_pore_volume, d_vol_animation, d_vol_detail, d_vol_monitor;
    bool open_z() const { return kper == 0 && wall[4] == 0 && wall[5] == 0; }
    bool velocity_inlet() const { return open_z() && inlet_bc == 1; }
    bool convective_outlet() const { return open_z() && outlet_bc == 1; }
};

Control read_control(const std::string& path) {
    const KeyFile f(path);
    Control c{};
    c.option = f.i32("initial_fluid_distribution_option");
    c.benchmark = f.i32("benchmark_cmd");
    c.extreme_large = f.i32("extreme_large_sim_cmd");
    c.breakthrough_check = f.i32("breakthrough_check");
    c.steady_option = f.i32("steady_state_option");
    c.convergence = f.real("convergence_criteria");
    c.field_precision = f.i32("output_fieldData_precision_cmd");
    c.modify_geometry = f.i32("modify_geometry_cmd");
    c.external_geometry = f.i32("external_geometry_read_cmd");
    c.dims_type_size = f.i32("geometry_dims_type_size");
    if (c.dims_type_size != 4 && c.dims_type_size != 8) throw Fatal("Incorrect geometry_dims_type_size!");
    c.geometry_preprocess = f.i32("geometry_preprocess_cmd");
    c.porous_plate = f.i32("porous_plate_cmd");
    c.z_porous_plate = f.i32("Z_porous_plate");
    c.change_inlet_phase = f.i32("change_inlet_fluid_phase_cmd");
    c.nx = f.i64("nxGlobal");
    c.ny = f.i64("nyGlobal");
    c.nz = f.i64("nzGlobal");
    c.n_exclude_inlet = f.i32("n_exclude_inlet");
    c.n_exclude_outlet = f.i32("n_exclude_outlet");
    c.wall[0] = f.i32("domain_wall_status_x_min");
    c.wall[1] = f.i32("domain_wall_status_x_max");
    c.wall[2] = f.i32("domain_wall_status_y_min");
    c.wall[3] = f.i32("domain_wall_status_y_max");
    c.wall[4] = f.i32("domain_wall_status_z_min");
    c.wall[5] = f.i32("domain_wall_status_z_max");
    c.iper = f.i32("iper");
    c.jper = f.i32("jper");
    c.kper = f.i32("kper");
    c.nu1 = f.real("fluid1_viscosity");
    c.nu2 = f.real("fluid2_viscosity");
    c.gamma = f.real("surface_tension");
    c.theta = f.real("theta");
    c.beta = f.real("RK_beta");
    c.inlet_bc = f.i32("inlet_BC");
    c.outlet_bc = f.i32("outlet_BC");
    c.sa_inject = f.real("saturation_injection");
    c.interface_z0 = f.real("initial_interface_position");
    c.sa_target = f.real("target_fluid1_saturation");
    // option 6's generator seed: a key of this build (the reference seeds from the clock);
    // absent, seed 1
    c.random_seed = f.has("random_seed") ? std::stoull(f.str("random_seed")) : 1;
    c.body_force = f.real("body_force_0");
    c.max_step = f.i32("max_time_step");
    c.ntime_visual = f.i32("ntime_visual");
    c.ntime_animation = f.i32("ntime_animation");
    c.ntime_monitor = f.i32("monitor_timer");
    c.monitor_profile_ratio = f.i32("monitor_profile_timer_ratio");
    c.ntime_clock_sum = f.i32("computation_time_timer");
    c.ntime_display = f.i32("display_steps_timer");
    c.checkpoint_timer = f.real("checkpoint_save_timer");
    c.checkpoint2_timer = f.real("checkpoint_2rd_save_timer");
    c.duration_timer = f.real("simulation_duration_timer");
    c.rho_in_new = f.i32("rho_in_new");
    c.rho_out_bc = f.i32("rho_out_BC");
    c.capillary_number = f.real("capillary_number");
    c.target_inject_pore_volume = f.real("target_inject_pore_volume");
    c.d_vol_animation = f.real("d_vol_animation");
    c.d_vol_detail = f.real("d_vol_detail");
    c.d_vol_monitor = f.real("d_vol_monitor");
    return c;
}

// Input checks of src/IO_multiphase.cpp:200-233, plus the options this build rejects.
void check_control(const Control& c) {
    std::vector<std::string> err;
    if (c.theta > 90.0) err.push_back("contact angle is larger than 90 degrees");
    if (c.iper == 1 || c.wall[1] == 0 || c.wall[0] == 0)
        err.push_back("X direction periodic BC enabled or non-slip BC not applied at x = xmin or x = xmax");
    if (c.jper == 0 && (c.wall[3] == 0 || c.wall[2] == 0))
        err.push_back("non-slip BC not applied at y = ymin or y = ymax while y direction periodic BC not enabled");
    if (c.jper == 1 && (c.wall[3] == 1 || c.wall[2] == 1))
        err.push_back("non-slip BC applied at y = ymin or y = ymax while y direction periodic BC enabled");
    if (c.kper == 1 && (c.wall[5] == 1 || c.wall[4] == 1))
        err.push_back("non-slip BC applied at z = zmin or z = zmax while z direction periodic BC enabled");
    if (c.outlet_bc == 1 && c.inlet_bc == 2) err.push_back("Inlet pressure + outlet convective BC is not supported");
    if (c.porous_plate < 0 || c.porous_plate > 2) err.push_back("porous_plate_cmd must be 0, 1 or 2");
    // the reference's loader of precomputed boundary info stops with this error itself
    // (src/Geometry_preprocessing.cpp:413-429); the boundary info is computed on the device
    if (c.geometry_preprocess == 1) err.push_back("geometry_preprocessing_load is not implemented yet");
    else if (c.geometry_preprocess != 0) err.push_back("Wrong value of geometry_preprocess_cmd");
    if (c.extreme_large != 0)
        err.push_back("Extreme large simulation domain (extreme_large_sim_cmd = 1) is not supported yet");
    if (c.option < 1 || c.option > 6) err.push_back("Input parameter initial_fluid_distribution_option error");
    if (c.option == 6 && !(c.sa_target >= 0.0 && c.sa_target <= 1.0))
        err.push_back("target_fluid1_saturation must lie in [0, 1]");
    if (c.ntime_monitor <= 0 || c.ntime_clock_sum <= 0 || c.ntime_display <= 0 || c.ntime_animation <= 0 ||
        c.ntime_visual <= 0 || c.monitor_profile_ratio <= 0)
        err.push_back("timers must be positive");
    if (!err.empty()) {
        std::string m;
        for (const auto& e : err) m += "Error: " + e + "!\n";
        throw Fatal(m + "Exit Program!");
    }
}

// ---------------------------------------------------------------------------------------
// Step constants (src/IO_multiphase.cpp:200-206, src/Init_multiphase.cpp:173-214)
// ---------------------------------------------------------------------------------------
constexpr double PI = 3.14159265358979323846;

mflbm_params derive(const Control& c, int precision, int mrt) {
    mflbm_params p{};
    p.mrt = mrt;
    p.nx = c.nx; p.ny = c.ny; p.nz = c.nz;
    p.iper = c.iper; p.jper = c.jper; p.kper = c.kper;
    p.wall_x_min = c.wall[0]; p.wall_x_max = c.wall[1];
    p.wall_y_min = c.wall[2]; p.wall_y_max = c.wall[3];
    p.wall_z_min = c.wall[4]; p.wall_z_max = c.wall[5];
    p.inlet_bc = c.inlet_bc; p.outlet_bc = c.outlet_bc;
    p.lbm_gamma = c.gamma;
    p.la_nui1 = 1.0 / c.nu1;
    p.la_nui2 = 1.0 / c.nu2;
    p.lbm_beta = c.beta;
    p.phi_inlet = 2.0 * c.sa_inject - 1.0;
    p.sa_inject = c.sa_inject;
    const double theta = (180.0 - c.theta) * PI / 180.0;  // measured through the defending phase
    p.cos_theta = std::cos(theta);
    p.force_z = c.body_force;
    p.rho_out = 1.0;
    p.rho_in = 0.0;
    p.porous_plate = c.porous_plate;
    p.z_porous_plate = c.z_porous_plate;
    p.uin_avg = 0.0;
    if (c.velocity_inlet()) {  // src/Init_multiphase.cpp:195-199, 262-263
        p.force_z = 0.0;
        p.uin_avg = c.capillary_number * c.gamma / c.nu1;
    }
    if (c.open_z() && c.inlet_bc == 2) {
        p.force_z = 0.0;
        const double p_gradient = -c.body_force / 3.0;
        if (c.rho_out_bc) p.rho_out = 1.0 - p_gradient * (double)c.nz;
        else p.rho_in = p.rho_out - p_gradient * (double)c.nz;
    }
    p.relaxation = 1.0;
    p.precision = precision;
    p.flags = 0;
    return p;
}

// ---------------------------------------------------------------------------------------
// Files and directories
// ---------------------------------------------------------------------------------------
std::string read_whole(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.good()) throw Fatal("Missing job status file! Exiting program!");
    std::ostringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

void write_status(const std::string& status) {
    std::ofstream out("./job_status.txt");
    if (!out.good()) throw Fatal("Could not open ./job_status.txt");
    out << status << std::endl;
}

void make_dir(const std::string& path) {
    if (mkdir(path.c_str(), 0755) != 0 && errno != EEXIST) throw Fatal("Could not Create " + path + " folder !");
}

// Geometry path rule (src/Misc.cpp:24-43): drop trailing non-digits, add ".dat".
std::string dat_path(std::string v) {
    while (!v.empty() && !std::isdigit((unsigned char)v.back())) v.pop_back();
    return v + ".dat";
}

// .dat file: three dimensions (int32 or int64) then nx*ny*nz bytes (1 solid), x fastest.
// Returns the voxel array of the whole domain; a sample smaller than the domain is padded
// with solid when both x_max and y_max walls are set (src/Misc.cpp:175-184).
// modify_geometry_cmd = 1 (src/Misc.cpp:105-136): the reference's sample edit, a sphere of
// radius ny/4 at the domain centre and everything outside a tube of radius ny/2 between the
// 10-layer inlet and outlet reservoirs made solid, before the domain walls are added.
void modify_geometry(const Control& c, std::vector<int8_t>& geo) {
    const double xc = 0.5 * double(c.nx + 1), yc = 0.5 * double(c.ny + 1), zc = 0.5 * double(c.nz + 1);
    const double r1 = 0.25 * (double)c.ny, r2 = (double)c.ny * 0.5;
    const long long buffer = 10;
    for (long long k = 1; k <= c.nz; ++k)
        for (long long j = 1; j <= c.ny; ++j)
            for (long long i = 1; i <= c.nx; ++i) {
                const double d2 = std::pow(i - xc, 2) + std::pow(j - yc, 2) + std::pow(k - zc, 2);
                int8_t& w = geo[(size_t)(i - 1) + c.nx * ((size_t)(j - 1) + c.ny * (size_t)(k - 1))];
                if (d2 < std::pow(r1, 2)) w = 1;
                if (d2 > std::pow(r2, 2) && k > buffer && k < c.nz - buffer + 1) w = 1;
            }
    std::cout << "Internal geometry modified!" << std::endl;
}

std::vector<int8_t> read_geometry_file(const Control& c);

std::vector<int8_t> read_geometry(const Control& c) {
    std::vector<int8_t> geo = read_geometry_file(c);
    if (c.modify_geometry == 1) modify_geometry(c, geo);
    return geo;
}

std::vector<int8_t> read_geometry_file(const Control& c) {
    const size_t n = (size_t)c.nx * c.ny * c.nz;
    std::vector<int8_t> geo(n, 0);
    if (c.external_geometry != 1) return geo;  // an empty duct
    const KeyFile paths("input/Geometry_File_Path.txt");
    const std::string file = dat_path(paths.str("geo_file_path_"));
    FILE* fp = std::fopen(file.c_str(), "rb");
    if (!fp) throw Fatal("Could not open the geometry file " + file + " !");
    long long dims[3];
    for (int a = 0; a < 3; ++a) {
        if (c.dims_type_size == 4) {
            int32_t v;
            if (std::fread(&v, 4, 1, fp) != 1) { std::fclose(fp); throw Fatal("Could not load from the geometry file!"); }
            dims[a] = v;
        } else {
            int64_t v;
            if (std::fread(&v, 8, 1, fp) != 1) { std::fclose(fp); throw Fatal("Could not load from the geometry file!"); }
            dims[a] = v;
        }
    }
    std::cout << "Porous media sample size: nx = " << dims[0] << ", ny = " << dims[1] << ", nz = " << dims[2] << std::endl;
    if (c.nx < dims[0] || c.ny < dims[1] || c.nz < dims[2]) {
        std::fclose(fp);
        throw Fatal("Error! Domain size is smaller than porous media sample size! Exiting program!");
    }
    std::vector<int8_t> sample((size_t)dims[0] * dims[1] * dims[2]);
    const size_t got = std::fread(sample.data(), 1, sample.size(), fp);
    std::fclose(fp);
    if (got != sample.size()) throw Fatal("Could not load from the geometry file!");
    for (long long k = 0; k < dims[2]; ++k)
        for (long long j = 0; j < dims[1]; ++j)
            for (long long i = 0; i < dims[0]; ++i)
                geo[(size_t)i + c.nx * ((size_t)j + c.ny * k)] = sample[(size_t)i + dims[0] * ((size_t)j + dims[1] * k)] ? 1 : 0;
    if (c.wall[1] == 1 && c.wall[3] == 1)
        for (long long k = 0; k < c.nz; ++k)
            for (long long j = 0; j < c.ny; ++j)
                for (long long i = 0; i < c.nx; ++i)
                    if (j + 1 >= dims[1] || i + 1 >= dims[0]) geo[(size_t)i + c.nx * ((size_t)j + c.ny * k)] = 1;
    return geo;
}

template <typename T>
void put_be(std::ofstream& out, T v) {
    unsigned char b[sizeof(T)];
    std::memcpy(b, &v, sizeof(T));
    for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
    out.write(reinterpret_cast<const char*>(b), sizeof(T));
}

// ---------------------------------------------------------------------------------------
// The simulation
// ---------------------------------------------------------------------------------------
class Simulation {
public:
    Simulation(const Control& c, int precision, int mrt) : c_(c), prec_(precision), mrt_(mrt) {
        nx_ = c.nx; ny_ = c.ny; nz_ = c.nz;
        s1_ = (nx_ + 2) * (ny_ + 2) * (nz_ + 2);
        s2_ = (nx_ + 4) * (ny_ + 4) * (nz_ + 4);
        s4_ = (nx_ + 8) * (ny_ + 8) * (nz_ + 8);
        es_ = precision == MFLBM_FP64 ? sizeof(double) : sizeof(float);
        ntime_monitor_profile_ = c.monitor_profile_ratio * c.ntime_monitor;
    }
    ~Simulation() {
        if (writer_.joinable()) writer_.join();
        if (h_) mflbm_destroy(h_);
    }

    int run(const std::string& job_status);

private:
    // index helpers of the reference layouts (1-based interior)
    int64_t i_s1(int64_t i, int64_t j, int64_t k) const { return i + (nx_ + 2) * (j + (ny_ + 2) * k); }
    int64_t i_s2(int64_t i, int64_t j, int64_t k) const { return (i + 1) + (nx_ + 4) * ((j + 1) + (ny_ + 4) * (k + 1)); }
    int64_t i_s4(int64_t i, int64_t j, int64_t k) const { return (i + 3) + (nx_ + 8) * ((j + 3) + (ny_ + 8) * (k + 3)); }

    void check(int rc, const char* what) const {
        if (rc != 0) throw Fatal(std::string(what) + " failed: " + (h_ ? mflbm_last_error(h_) : "no handle"));
    }
    void create(const mflbm_params& p, const std::vector<int8_t>& geo) {
        if (mflbm_create(&p, geo.data(), &h_) != 0) {
            const std::string m = h_ ? mflbm_last_error(h_) : "device initialisation failed";
            throw Fatal("mflbm_create failed: " + m);
        }
    }
    std::vector<double> download_real(int field, int64_t n) const;
    void upload_real(int field, const std::vector<double>& v);
    void geometry_info();
    void injected_volume_timers();
    void change_inlet_fluid_phase();
    void write_info() const;
    void load_checkpoint(double* force_z, double* rho_in, std::vector<double>* pdf, std::vector<double>* phi,
                         std::vector<double>* conv);
    void save_checkpoint(int where);
    void monitor();
    void monitor_phase_field();
    void monitor_capillary_pressure();
    void monitor_breakthrough();
    void vtk(int type);
    void vtk_walls() const;
    bool first_record() const { return ntime0_ == 1 && ntime_ == c_.ntime_monitor; }
    std::ofstream series(const std::string& name) const {
        std::ofstream f("results/out1.output/" + name, first_record() ? std::ios::out : std::ios::app);
        if (!f.good()) throw Fatal("Could not open results/out1.output/" + name);
        return f;
    }

    // File output runs on one background thread: the device state is downloaded at the output
    // step, then formatted and written while the next steps run.  One writer at a time, in
    // output order; a writer's error surfaces at the next output or at the end of the run.
    void drain() {
        if (writer_.joinable()) writer_.join();
        if (writer_err_) {
            const std::exception_ptr e = writer_err_;
            writer_err_ = nullptr;
            std::rethrow_exception(e);
        }
    }
    void background(std::function<void()> job) {
        drain();
        writer_ = std::thread([this, job = std::move(job)]() {
            try {
                job();
            } catch (...) {
                writer_err_ = std::current_exception();
            }
        });
    }
    std::thread writer_;
    std::exception_ptr writer_err_;

    Control c_;  // a copy: the velocity inlet rescales the timers by the injected volume
    const int prec_;
    const int mrt_;  // collision model (--mrt)
    mflbm_handle* h_ = nullptr;
    mflbm_params p_{};
    int64_t nx_, ny_, nz_, s1_, s2_, s4_;
    size_t es_;
    int ntime0_ = 1, ntime_ = 1, end_ = 0;
    int ntime_monitor_profile_;
    std::vector<int8_t> walls_;       // s2
    std::vector<int> pore_profile_;   // pore nodes per z slice
    long long pore_sum_ = 0, pore_sum_effective_ = 0, n_solid_boundary_ = 0, n_fluid_boundary_ = 0;
    double a_xy_ = 0.0;
    double monitor_previous_ = 0.0, monitor_current_ = 0.0;
    std::vector<double> phi_old_;
};

std::vector<double> Simulation::download_real(int field, int64_t n) const {
    std::vector<double> out((size_t)n);
    if (prec_ == MFLBM_FP64) {
        check(mflbm_download(h_, field, out.data(), n * sizeof(double)), "download");
    } else {
        std::vector<float> tmp((size_t)n);
        check(mflbm_download(h_, field, tmp.data(), n * sizeof(float)), "download");
        for (int64_t i = 0; i < n; ++i) out[i] = tmp[i];
    }
    return out;
}

void Simulation::upload_real(int field, const std::vector<double>& v) {
    if (prec_ == MFLBM_FP64) {
        check(mflbm_upload(h_, field, v.data(), v.size() * sizeof(double)), "upload");
    } else {
        std::vector<float> tmp(v.begin(), v.end());
        check(mflbm_upload(h_, field, tmp.data(), tmp.size() * sizeof(float)), "upload");
    }
}

// pore profile and boundary-node counts (src/Misc.cpp:195-219, src/Init_multiphase.cpp:125-160)
void Simulation::geometry_info() {
    walls_.assign((size_t)s2_, 0);
    check(mflbm_download(h_, MFLBM_WALLS, walls_.data(), walls_.size()), "download walls");
    std::vector<int8_t> wtype((size_t)s4_);
    check(mflbm_download(h_, MFLBM_WALLS_TYPE, wtype.data(), wtype.size()), "download walls_type");
    pore_profile_.assign((size_t)nz_, 0);
    for (int64_t k = 1; k <= nz_; ++k)
        for (int64_t j = 1; j <= ny_; ++j)
            for (int64_t i = 1; i <= nx_; ++i) {
                if (walls_[i_s2(i, j, k)] <= 0) pore_profile_[k - 1]++;
                const int t = wtype[i_s4(i, j, k)];
                if (t == 2) n_solid_boundary_++;
                if (t == -1) n_fluid_boundary_++;
            }
    for (int64_t k = 1; k <= nz_; ++k) pore_sum_ += pore_profile_[k - 1];
    for (int64_t k = 1 + c_.n_exclude_inlet; k <= nz_ - c_.n_exclude_outlet; ++k) pore_sum_effective_ += pore_profile_[k - 1];
    const double la_x = (double)(nx_ - 1) - 0.5 - 0.5, la_y = (double)(ny_ - 1) - 0.5 - 0.5;
    a_xy_ = la_x * la_y;
}

// Timers of a velocity-inlet run set by the injected volume, in pore volumes, over the inlet
// flow rate uin_avg * A_xy, rounded up to even steps (src/Init_multiphase.cpp:218-248, 286-292);
// computed in the run's precision as the reference computes them in T_P.
void Simulation::injected_volume_timers() {
    if (!c_.velocity_inlet()) return;
    auto steps = [&](double vol) -> int {
        int t;
        if (prec_ == MFLBM_FP64) {
            const double flowrate = p_.uin_avg * a_xy_;
            t = (int)((double)(vol * (double)pore_sum_) / flowrate);
        } else {
            const float la_x = (float)(nx_ - 1) - 0.5f - 0.5f, la_y = (float)(ny_ - 1) - 0.5f - 0.5f;
            const float uin = (float)c_.capillary_number * (float)c_.gamma / (float)c_.nu1;
            const float flowrate = uin * (la_x * la_y);
            t = (int)((float)((float)vol * (float)pore_sum_) / flowrate);
        }
        return t % 2 == 1 ? t + 1 : t;
    };
    std::cout << "Inlet average velocity = " << p_.uin_avg << std::endl;
    std::cout << "Inlet flowrate = " << p_.uin_avg * a_xy_ << std::endl;
    if (c_.target_inject_pore_volume > 0) {
        c_.max_step = steps(c_.target_inject_pore_volume);
        std::cout << "Maximum time step is modified based on target inject volume! ntime_max = " << c_.max_step
                  << std::endl;
    }
    if (c_.d_vol_animation > 0) {
        c_.ntime_animation = steps(c_.d_vol_animation);
        std::cout << "Animation VTK files timer is modified based on injected volume: " << c_.ntime_animation << std::endl;
    }
    if (c_.d_vol_detail > 0) {
        c_.ntime_visual = steps(c_.d_vol_detail);
        std::cout << "Full VTK files timer is modified based on injected volume: " << c_.ntime_visual << std::endl;
    }
    if (c_.d_vol_monitor > 0) {
        c_.ntime_monitor = steps(c_.d_vol_monitor);
        ntime_monitor_profile_ = c_.monitor_profile_ratio * c_.ntime_monitor;
        std::cout << "Monitor timer is modified based on injected volume: " << c_.ntime_monitor << std::endl;
    }
    if (c_.ntime_animation <= 0 || c_.ntime_visual <= 0 || c_.ntime_monitor <= 0)
        throw Fatal("Error: an injected-volume timer is below one step (raise d_vol_* or lower the flow rate)!");
}

// change_inlet_fluid_phase_cmd 1 / 2 (src/Misc.cpp:277-383, called from src/main.cpp:82-85
// after the new or loaded state): below interface_z0 every population of the other colour
// is moved into fluid 1 / fluid 2 (over i = 0 .. nx - 1, j = 0 .. ny + 1, k = 0 .. nz + 1,
// the reference's ranges); phi is left as it is until the next collision.
void Simulation::change_inlet_fluid_phase() {
    std::vector<double> pdf = download_real(MFLBM_PDF, 38 * s1_);
    const int to = c_.change_inlet_phase == 1 ? 0 : 1, from = 1 - to;
    for (int64_t k = 0; k <= nz_ + 1; ++k) {
        if (!((double)k <= c_.interface_z0)) continue;
        for (int64_t j = 0; j <= ny_ + 1; ++j)
            for (int64_t i = 0; i <= nx_ - 1; ++i)
                for (int e = 0; e < 19; ++e) {
                    double& a = pdf[(size_t)(e + 19 * to) * s1_ + i_s1(i, j, k)];
                    double& b = pdf[(size_t)(e + 19 * from) * s1_ + i_s1(i, j, k)];
                    a += b;
                    b = 0.0;
                }
    }
    upload_real(MFLBM_PDF, pdf);
    std::cout << "Inlet fluid phase was changed (1 to fluid1; 2 to fluid2): " << c_.change_inlet_phase << std::endl;
}

void Simulation::write_info() const {
    std::ofstream f("results/out1.output/info.txt");
    if (!f.good()) throw Fatal("Couldn't open the file results/out1.output/info.txt !");
    const double porosity_full = (double)pore_sum_ / (double)((nx_ - 2) * (ny_ - 2) * nz_);
    const double porosity_eff = (double)pore_sum_effective_ / (double)((nx_ - 2) * (ny_ - 2)) /
                                (double)(nz_ - c_.n_exclude_outlet - c_.n_exclude_inlet);
    f << " Grid information:" << std::endl;
    f << "nxGlobal = " << nx_ << " , nyGlobal = " << ny_ << " , nzGlobal = " << nz_ << std::endl;
    f << "Inlet open cross sectional area = " << a_xy_ << std::endl;
    f << " Pore information:" << std::endl;
    f << "Total number of pore nodes = " << pore_sum_ << std::endl;
    f << "Total number of solid boundary nodes = " << n_solid_boundary_ << std::endl;
    f << "Total number of fluid boundary nodes = " << n_fluid_boundary_ << std::endl;
    f << "Total number of effective pore nodes (excluding inlet/outlet) = " << pore_sum_effective_ << std::endl;
    f << "Full domain porosity = " << porosity_full << std::endl;
    f << "Effective domain porosity  (excluding inlet/outlet) = " << porosity_eff << std::endl;
    std::cout << "Total number of pore nodes = " << pore_sum_ << std::endl;
    std::cout << "Full domain porosity = " << porosity_full << std::endl;
}

// Checkpoint: int32 (ntime + 1) | T force_z | T rho_in | pdf (38 s1 T) | phi (s4 T)
// [| f_convec, g_convec (19 X1 Y1 T each) | phi_convec (X1 Y1 T) with the convective outlet],
// native endian, T the run's precision (src/IO_multiphase.cpp:252-287).  Either precision loads.
void Simulation::load_checkpoint(double* force_z, double* rho_in, std::vector<double>* pdf,
                                 std::vector<double>* phi, std::vector<double>* conv) {
    const std::string path = "results/out2.checkpoint/id0000";
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw Fatal("Checkpoint data not found! Exiting program!");
    std::fseek(fp, 0, SEEK_END);
    const long long bytes = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    const long long tail = 39 * (long long)(nx_ + 2) * (ny_ + 2), base = 2 + 38 * (long long)s1_ + s4_;
    long long body = 0;
    size_t es = 0;
    for (long long b : {base, base + tail})
        for (size_t e : {sizeof(double), sizeof(float)})
            if (bytes == 4 + (long long)e * b) { body = b; es = e; }
    if (!es) { std::fclose(fp); throw Fatal("Checkpoint size does not match the domain"); }
    const bool has_tail = body > base;
    if (c_.convective_outlet() && !has_tail)
        { std::fclose(fp); throw Fatal("Could not load from data check_point_file file! (no convective-outlet state)"); }
    int32_t nt;
    bool ok = std::fread(&nt, 4, 1, fp) == 1;
    std::vector<double> all((size_t)body);
    if (es == sizeof(double)) {
        ok = ok && std::fread(all.data(), sizeof(double), all.size(), fp) == all.size();
    } else {
        std::vector<float> tmp(all.size());
        ok = ok && std::fread(tmp.data(), sizeof(float), tmp.size(), fp) == tmp.size();
        for (size_t i = 0; i < tmp.size(); ++i) all[i] = tmp[i];
    }
    std::fclose(fp);
    if (!ok) throw Fatal("Could not load from data check_point_file file!");
    *force_z = all[0];
    *rho_in = all[1];
    pdf->assign(all.begin() + 2, all.begin() + 2 + 38 * s1_);
    phi->assign(all.begin() + 2 + 38 * s1_, all.begin() + 2 + 38 * s1_ + s4_);
    conv->clear();
    if (has_tail && c_.convective_outlet()) conv->assign(all.begin() + 2 + 38 * s1_ + s4_, all.end());
    ntime0_ = nt;  // the first step not yet taken
}

void Simulation::save_checkpoint(int where) {
    drain();  // one set of output buffers at a time
    const std::string path = where == 0 ? "results/out2.checkpoint/id0000" : "results/out2.checkpoint/2rd_backup/id0000";
    const int32_t nt = ntime_ + 1;
    std::vector<char> pdf((size_t)(38 * s1_) * es_), phi((size_t)s4_ * es_), conv;
    check(mflbm_download(h_, MFLBM_PDF, pdf.data(), pdf.size()), "download pdf");
    check(mflbm_download(h_, MFLBM_PHI, phi.data(), phi.size()), "download phi");
    if (c_.convective_outlet()) {
        conv.resize((size_t)(39 * (nx_ + 2) * (ny_ + 2)) * es_);
        check(mflbm_download(h_, MFLBM_CONVEC, conv.data(), conv.size()), "download convective outlet state");
    }
    const double fz = p_.force_z, ri = p_.rho_in;
    const size_t es = es_;
    const bool fp64 = prec_ == MFLBM_FP64;
    background([=, pdf = std::move(pdf), phi = std::move(phi), conv = std::move(conv)]() {
        FILE* fp = std::fopen(path.c_str(), "wb+");
        if (!fp) throw Fatal("Could not create save_checkpoint file!");
        bool ok = std::fwrite(&nt, 4, 1, fp) == 1;
        auto put = [&](const void* data, size_t n) { ok = ok && std::fwrite(data, es, n, fp) == n; };
        if (fp64) {
            put(&fz, 1);
            put(&ri, 1);
        } else {
            const float f = (float)fz, r = (float)ri;
            put(&f, 1);
            put(&r, 1);
        }
        put(pdf.data(), pdf.size() / es);
        put(phi.data(), phi.size() / es);
        if (!conv.empty()) put(conv.data(), conv.size() / es);
        std::fclose(fp);
        if (!ok) throw Fatal("Could not write the checkpoint file " + path);
        if (where == 0) {
            std::cout << "Saving checkpoint data completed!" << std::endl;
            write_status("continue_simulation");
        } else {
            std::cout << "Saving secondary checkpoint data completed!" << std::endl;
        }
    });
}

// Monitor of an unsteady (displacement) run: src/Monitor.cpp:17-277, sums from the device.
void Simulation::monitor() {
    std::vector<double> fl1(nz_), fl2(nz_), vol1(nz_), vol2(nz_), mass1(nz_), mass2(nz_), pre(nz_);
    mflbm_monitor_out m{};
    m.fl1 = fl1.data(); m.fl2 = fl2.data(); m.vol1 = vol1.data(); m.vol2 = vol2.data();
    m.mass1 = mass1.data(); m.mass2 = mass2.data(); m.pre = pre.data();
    check(mflbm_monitor(h_, &m), "monitor");
    const double ke1 = 0.5 * m.usq1, ke2 = 0.5 * m.usq2;
    const double umax = std::sqrt(m.umax2);

    double mass1_sum = 0, mass2_sum = 0, vol1_sum = 0, vol2_sum = 0;
    for (int64_t k = c_.n_exclude_inlet + 1; k <= nz_ - c_.n_exclude_outlet; ++k) {
        mass1_sum += mass1[k - 1]; mass2_sum += mass2[k - 1];
        vol1_sum += vol1[k - 1]; vol2_sum += vol2[k - 1];
    }
    const double saturation = vol1_sum / (vol1_sum + vol2_sum);
    series("saturation.dat") << ntime_ << " " << saturation << " " << vol1_sum << " " << vol2_sum << " " << mass1_sum
                             << " " << mass2_sum << std::endl;
    double t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    for (int64_t k = 1; k <= nz_; ++k) { t1 += mass1[k - 1]; t2 += mass2[k - 1]; t3 += vol1[k - 1]; t4 += vol2[k - 1]; }
    const double saturation_full = t3 / (t3 + t4);

    double fl1_whole = 0, fl2_whole = 0, fl1_avg = 0, fl2_avg = 0;
    for (int64_t k = 1; k <= nz_; ++k) { fl1_whole += fl1[k - 1]; fl2_whole += fl2[k - 1]; }
    fl1_whole /= (double)nz_;
    fl2_whole /= (double)nz_;
    for (int64_t k = c_.n_exclude_inlet + 1; k <= nz_ - c_.n_exclude_outlet; ++k) { fl1_avg += fl1[k - 1]; fl2_avg += fl2[k - 1]; }
    const double n_eff = (double)(nz_ - c_.n_exclude_outlet - c_.n_exclude_inlet);
    fl1_avg /= n_eff;
    fl2_avg /= n_eff;
    const double fl_avg = fl1_avg + fl2_avg;
    const double ca = fl_avg / a_xy_ * c_.nu1 / c_.gamma;  // Darcy velocity of the bulk fluid
    series("Ca_number.dat") << ntime_ << " " << ca << " " << umax << " " << ke1 << " " << ke2 << std::endl;
    series("flowrate_time.dat") << ntime_ << " " << fl1_whole + fl2_whole << " " << fl1_whole << " " << fl2_whole << " "
                                << fl_avg << " " << fl1_avg << " " << fl2_avg << std::endl;
    series("saturation_full_domain.dat") << ntime_ << " " << saturation_full << " " << t3 << " " << t4 << " " << t1 << " "
                                         << t2 << std::endl;
    if (c_.kper == 0 && c_.wall[4] == 0 && c_.wall[5] == 0) {
        const int64_t ki = 1 + c_.n_exclude_inlet, ko = nz_ - c_.n_exclude_outlet;
        const double p1 = pre[ki - 1] / pore_profile_[ki - 1], p2 = pre[ko - 1] / pore_profile_[ko - 1];
        series("pre.dat") << ntime_ << " " << p1 << " " << p2 << " " << p1 - p2 << std::endl;
    }
    if (ntime_ % ntime_monitor_profile_ == 0) {
        std::ostringstream name;
        name << "results/out1.output/profile/monitor " << std::setfill('0') << std::setw(8) << ntime_;
        std::ofstream f(name.str());
        const double eps = 1.1920928955078125e-07;  // FLT_EPSILON (includes/Module.h:10-16)
        for (int64_t k = 0; k < nz_; ++k)
            f << (double)k << " " << fl1[k] << " " << fl2[k] << " " << mass1[k] / (mass1[k] + mass2[k]) << " "
              << pre[k] / ((double)pore_profile_[k] + eps) << std::endl;
    }
    if (std::isnan(saturation_full) || std::isnan(ca)) {
        std::cout << "Simulation failed due to NAN of saturation or capillary number" << std::endl;
        end_ = 3;
    } else if (umax > 0.5) {
        std::cout << "Simulation failed due to maximum velocity larger than 0.5 !" << std::endl;
        end_ = 3;
    }
    if (c_.steady_option == 3) {
        const double eps = 1.1920928955078125e-07;
        monitor_previous_ = monitor_current_;
        monitor_current_ = saturation_full;
        const double rel = std::fabs(monitor_current_ - monitor_previous_) / (std::fabs(monitor_current_) + eps);
        series("steady_monitor_saturation_error.dat") << ntime_ << " " << rel << std::endl;
        if (rel < c_.convergence && ntime_ > c_.ntime_monitor) {
            std::cout << "Simulation converged based on saturation! Relative error = " << rel
                      << " ; maximum velocity = " << umax << std::endl;
            end_ = 1;
        }
    }
}

// Steady state on the largest local change of phi between monitor steps
// (src/Monitor.cpp:279-352).
void Simulation::monitor_phase_field() {
    mflbm_monitor_out m{};
    check(mflbm_monitor(h_, &m), "monitor");
    const std::vector<double> phi = download_real(MFLBM_PHI, s4_);
    double dmax = 0.0;
    if (phi_old_.empty()) phi_old_.assign((size_t)s4_, 0.0);
    for (int64_t k = 1; k <= nz_; ++k)
        for (int64_t j = 1; j <= ny_; ++j)
            for (int64_t i = 1; i <= nx_; ++i) {
                const int64_t c = i_s4(i, j, k);
                const double d = std::fabs(phi[c] - phi_old_[c]) * (1 - walls_[i_s2(i, j, k)]);
                if (dmax < d) dmax = d;
            }
    phi_old_ = phi;
    if (std::isnan(m.umax2) || std::isnan(dmax)) {
        std::cout << "Simulation failed due to NAN!" << std::endl;
        end_ = 3;
        return;
    }
    const double umax = std::sqrt(m.umax2);
    series("steady_monitor_max_phi_change.dat") << ntime_ << " " << dmax << " " << umax << std::endl;
    if (umax < 0.5) {
        if (dmax < c_.convergence && ntime_ > c_.ntime_monitor) {
            std::cout << "Simulation converged based on maximum local change of phi! Relative error =    " << dmax
                      << "; maximum velocity =  " << umax << std::endl;
            end_ = 1;
        }
    } else {
        std::cout << "Maximum velocity larger than 0.5, simulation failed!!!" << std::endl;
        end_ = 3;
    }
}

// Steady state on the relative change of the capillary pressure, the difference of the
// mean densities of the bulk phases over 3 (src/Monitor.cpp:354-445).
void Simulation::monitor_capillary_pressure() {
    std::vector<double> rho((size_t)s1_), u((size_t)s1_), v((size_t)s1_), w((size_t)s1_);
    check(mflbm_macro(h_, rho.data(), u.data(), v.data(), w.data()), "macro");
    const std::vector<double> phi = download_real(MFLBM_PHI, s4_);
    double umax = 0.0, pre_w = 0.0, pre_nw = 0.0;
    long long n_w = 0, n_nw = 0;
    for (int64_t k = 1; k <= nz_; ++k)
        for (int64_t j = 1; j <= ny_; ++j)
            for (int64_t i = 1; i <= nx_; ++i) {
                const int wall = walls_[i_s2(i, j, k)];
                const int64_t c1 = i_s1(i, j, k);
                const double t = (u[c1] * u[c1] + v[c1] * v[c1] + w[c1] * w[c1]) * (1 - wall);
                if (umax < t) umax = t;
                if (wall == 0) {
                    const double ph = phi[i_s4(i, j, k)];
                    if (ph < -0.99) { pre_w += rho[c1]; n_w++; }
                    if (ph > 0.99) { pre_nw += rho[c1]; n_nw++; }
                }
            }
    const double eps = 1.1920928955078125e-07;
    umax = std::sqrt(umax);
    pre_w = pre_w / ((double)n_w + eps);
    pre_nw = pre_nw / ((double)n_nw + eps);
    const double pc = (pre_nw - pre_w) / 3.0;
    monitor_previous_ = monitor_current_;
    monitor_current_ = pc;
    const double rel = std::fabs(monitor_current_ - monitor_previous_) / (std::fabs(monitor_current_) + eps);
    series("steady_monitor_capillary_pressure_error.dat") << ntime_ << " " << rel << " " << pc << " " << pre_w << " "
                                                          << pre_nw << " " << rel << " " << umax << std::endl;
    if (std::isnan(umax) || std::isnan(pc)) {
        std::cout << "Simulation failed due to NAN!" << std::endl;
        end_ = 3;
    } else if (umax < 0.5) {
        if (rel < c_.convergence && ntime_ > c_.ntime_monitor) {
            std::cout << "Simulation converged based on change of capillary pressure! Relative error =   " << rel
                      << " ; maximum velocity = " << umax << std::endl;
            end_ = 1;
        }
    } else {
        std::cout << "maximum velocity larger than 0.5, simulation failed!!!" << std::endl;
        end_ = 3;
    }
}

// Fluid 1 seen at the plane next to the outlet (src/Monitor.cpp:447-470).
void Simulation::monitor_breakthrough() {
    const std::vector<double> phi = download_real(MFLBM_PHI, s4_);
    const int64_t k = nz_ - 1;
    for (int64_t j = 1; j <= ny_; ++j)
        for (int64_t i = 1; i <= nx_; ++i)
            if (walls_[i_s2(i, j, k)] == 0 && phi[i_s4(i, j, k)] > 0.0) {
                end_ = 1;
                std::cout << "Breakthrough point reached! Exiting program!" << std::endl;
                return;
            }
}

// Legacy binary VTK, big endian (src/IO_multiphase.cpp:500-716).  Type 1: phi, density and
// velocity; type 2: phi (float) with phi = 0 at solid nodes.
void Simulation::vtk(int type) {
    drain();  // one set of output buffers at a time
    std::ostringstream name;
    name << std::setfill('0') << std::setw(10) << ntime_ << ".vtk";
    const std::string path = type == 1 ? "results/out3.field_data/full_flow_field/full_flow_field_" + name.str()
                                       : "results/out3.field_data/phase_distribution/small_" + name.str();
    std::vector<double> phi = download_real(MFLBM_PHI, s4_), rho, u, v, w;
    if (type == 1) {
        rho.resize((size_t)s1_); u.resize((size_t)s1_); v.resize((size_t)s1_); w.resize((size_t)s1_);
        check(mflbm_macro(h_, rho.data(), u.data(), v.data(), w.data()), "macro");
    }
    // walls_ and the sizes do not change after set-up, so the writer reads them in place
    background([this, type, path, phi = std::move(phi), rho = std::move(rho), u = std::move(u), v = std::move(v),
                w = std::move(w)]() {
        std::ofstream out(path, std::ios::out | std::ios::binary);
        if (!out.good()) throw Fatal("Could not open vtk outputl file!");
        const int64_t n = nx_ * ny_ * nz_;
        out << "# vtk DataFile Version 3.0" << std::endl << "vtk output" << std::endl << "BINARY" << std::endl
            << "DATASET STRUCTURED_POINTS" << std::endl
            << "DIMENSIONS " << nx_ << " " << ny_ << " " << nz_ << std::endl
            << "ORIGIN 1 1 1" << std::endl << "SPACING 1 1 1" << std::endl << "POINT_DATA " << n << std::endl;
        if (type == 2) {
            out << "SCALARS phi float" << std::endl << "LOOKUP_TABLE default" << std::endl;
            for (int64_t k = 1; k <= nz_; ++k)
                for (int64_t j = 1; j <= ny_; ++j)
                    for (int64_t i = 1; i <= nx_; ++i) {
                        const int wl = walls_[i_s2(i, j, k)];
                        put_be(out, (float)(0.0 * wl + phi[i_s4(i, j, k)] * (1 - wl)));
                    }
            return;
        }
        const bool dbl = c_.field_precision != 0;
        const char* fmt = dbl ? "double" : "float";
        auto block = [&](const char* label, auto value) {
            out << "SCALARS " << label << " " << fmt << std::endl << "LOOKUP_TABLE default" << std::endl;
            for (int64_t k = 1; k <= nz_; ++k)
                for (int64_t j = 1; j <= ny_; ++j)
                    for (int64_t i = 1; i <= nx_; ++i) {
                        const double x = value(i, j, k);
                        if (dbl) put_be(out, x);
                        else put_be(out, (float)x);
                    }
        };
        block("phi", [&](int64_t i, int64_t j, int64_t k) { return phi[i_s4(i, j, k)]; });
        block("density", [&](int64_t i, int64_t j, int64_t k) { return rho[i_s1(i, j, k)]; });
        block("velocity_X", [&](int64_t i, int64_t j, int64_t k) { return u[i_s1(i, j, k)]; });
        block("velocity_Y", [&](int64_t i, int64_t j, int64_t k) { return v[i_s1(i, j, k)]; });
        block("velocity_Z", [&](int64_t i, int64_t j, int64_t k) { return w[i_s1(i, j, k)]; });
        if (!out.good()) throw Fatal("Could not write " + path);
    });
}

void Simulation::vtk_walls() const {
    std::ofstream out("results/out3.field_data/walls_bin.vtk", std::ios::out | std::ios::binary);
    if (!out.good()) throw Fatal("Could not create 'results/out3.field_data/walls_bin.vtk' file!");
    out << "# vtk DataFile Version 3.0" << std::endl << "vtk output" << std::endl << "BINARY" << std::endl
        << "DATASET STRUCTURED_POINTS" << std::endl
        << "DIMENSIONS " << nx_ << " " << ny_ << " " << nz_ << std::endl
        << "ORIGIN 1 1 1" << std::endl << "SPACING 1 1 1" << std::endl << "POINT_DATA " << nx_ * ny_ * nz_ << std::endl
        << "SCALARS walls int" << std::endl << "LOOKUP_TABLE default" << std::endl;
    for (int64_t k = 1; k <= nz_; ++k)
        for (int64_t j = 1; j <= ny_; ++j)
            for (int64_t i = 1; i <= nx_; ++i) put_be(out, (int32_t)walls_[i_s2(i, j, k)]);
}

int Simulation::run(const std::string& job_status) {
    for (const char* d : {"results/", "results/out2.checkpoint/", "results/out2.checkpoint/2rd_backup/",
                          "results/out1.output/", "results/out1.output/profile/", "results/out3.field_data/",
                          "results/out3.field_data/phase_distribution/", "results/out3.field_data/full_flow_field/"})
        make_dir(d);
    const std::vector<int8_t> geo = read_geometry(c_);
    p_ = derive(c_, prec_, mrt_);
    std::vector<double> ck_pdf, ck_phi, ck_conv;
    if (job_status == "continue_simulation") {
        std::cout << "loading checkpoint data ... ";
        double fz, ri;
        load_checkpoint(&fz, &ri, &ck_pdf, &ck_phi, &ck_conv);
        p_.force_z = fz;
        p_.rho_in = ri;
        std::cout << "Complete" << std::endl;
        // a new inlet density from the control file overrides the stored one (src/main.cpp:121-132)
        if (c_.kper == 0 && c_.wall[4] == 0 && c_.wall[5] == 0 && c_.inlet_bc == 2 && c_.rho_in_new) {
            const double p_gradient = -c_.body_force / 3.0;
            if (c_.rho_out_bc) p_.rho_out = 1.0 - p_gradient * (double)nz_;
            else p_.rho_in = p_.rho_out - p_gradient * (double)nz_;
        }
    }
    std::cout << "Fluid 1 relaxation time = " << 3.0 * c_.nu1 + 0.5 << std::endl;
    std::cout << "Fluid 2 relaxation time = " << 3.0 * c_.nu2 + 0.5 << std::endl;
    create(p_, geo);
    geometry_info();
    injected_volume_timers();
    write_info();
    if (job_status == "new_simulation") {
        ntime0_ = 1;
        if (c_.option == 6) check(mflbm_init_phi_random(h_, c_.sa_target, c_.random_seed), "init_phi_random");
        else check(mflbm_init_phi(h_, c_.option, c_.interface_z0), "init_phi");
        check(mflbm_init_equilibrium(h_, nullptr, nullptr, nullptr, nullptr), "init_equilibrium");
    } else {
        upload_real(MFLBM_PDF, ck_pdf);
        upload_real(MFLBM_PHI, ck_phi);
        if (!ck_conv.empty()) upload_real(MFLBM_CONVEC, ck_conv);
        monitor_previous_ = -1.0;  // no convergence test on the first monitor after a restart
        monitor_current_ = -1.0;
    }
    if (c_.change_inlet_phase == 1 || c_.change_inlet_phase == 2) change_inlet_fluid_phase();
    check(mflbm_color_gradient(h_), "color_gradient");
    {
        mflbm_monitor_out m{};
        std::vector<double> v1(nz_), v2(nz_);
        m.vol1 = v1.data(); m.vol2 = v2.data();
        check(mflbm_monitor(h_, &m), "monitor");
        double a = 0, b = 0;
        for (int64_t k = 0; k < nz_; ++k) { a += v1[k]; b += v2[k]; }
        std::cout << "Initial saturation: " << a / (a + b + 1.1920928955078125e-07) << std::endl;
    }
    ntime_ = ntime0_;
    if (job_status == "new_simulation" && c_.benchmark == 0) {
        vtk(2);
        vtk_walls();
    }

    std::cout << "************************** Entering main loop *********************************" << std::endl;
    using clock = std::chrono::steady_clock;
    const auto t_loop = clock::now();
    auto t_clock = t_loop, t_display = t_loop;
    int counter1 = 1, counter2 = 1;
    const int last = ntime0_ + c_.max_step;
    int nt = ntime0_;
    auto next_multiple = [](int t, int m) { return ((t + m - 1) / m) * m; };
    while (nt <= last) {
        // run up to the next step that has an output, as one graph-replayed batch
        int stop = last;
        for (int m : {c_.ntime_clock_sum, c_.ntime_monitor, c_.ntime_display, c_.ntime_animation, c_.ntime_visual})
            stop = std::min(stop, next_multiple(nt, m));
        check(mflbm_run(h_, nt, stop - nt + 1), "run");
        ntime_ = stop;
        nt = stop + 1;
        if (ntime_ % c_.ntime_clock_sum == 0) {
            std::ofstream f("results/out1.output/time.dat",
                            (ntime0_ == 1 && ntime_ == c_.ntime_clock_sum) ? std::ios::out : std::ios::app);
            if (!f.good()) throw Fatal("Could not open results/out1.output/time.dat");
            const auto now = clock::now();
            f << ntime_ << " " << std::chrono::duration<double>(now - t_clock).count() << " sec" << std::endl;
            t_clock = now;
        }
        if (ntime_ % c_.ntime_monitor == 0) {
            if (c_.steady_option == 0 || c_.steady_option == 3) {
                monitor();
                if (c_.breakthrough_check == 1) monitor_breakthrough();
            } else if (c_.steady_option == 1) {
                monitor_capillary_pressure();
            } else if (c_.steady_option == 2) {
                monitor_phase_field();
            }
        }
        if (ntime_ % c_.ntime_display == 0) {
            const auto now = clock::now();
            const double sec = std::chrono::duration<double>(now - t_display).count();
            std::cout << "ntime =\t" << ntime_ << std::endl;
            std::cout << "simulation speed: " << (double)(nx_ * ny_ * nz_) * c_.ntime_display / sec * 1e-6 << " MLUPS"
                      << std::endl;
            t_display = now;
        }
        if (ntime_ % c_.ntime_animation == 0) vtk(2);
        if (ntime_ % c_.ntime_visual == 0) vtk(1);
        if (ntime_ % c_.ntime_clock_sum == 0) {
            const double hours = std::chrono::duration<double>(clock::now() - t_loop).count() / 3600.0;
            std::cout << "simulation has run\t" << hours << "\thours" << std::endl;
            if (hours >= c_.duration_timer) {
                std::cout << "Time to save checkpoint data and exit the program!" << std::endl;
                end_ = 2;
            }
            if (hours >= counter1 * c_.checkpoint_timer) {
                std::cout << "Time to save checkpoint data!" << std::endl;
                save_checkpoint(0);
                counter1++;
            }
            if (hours >= counter2 * c_.checkpoint2_timer) {
                std::cout << "Time to save secondary checkpoint data!" << std::endl;
                save_checkpoint(1);
                counter2++;
            }
        }
        if (end_ > 0) break;
    }
    std::cout << "************************** Exiting main iteration *********************************" << std::endl;
    const double total = std::chrono::duration<double>(clock::now() - t_loop).count();
    std::cout << "Code speed:\t" << (double)(nx_ * ny_ * nz_) * (double)(ntime_ - ntime0_ + 1) / (1e6 * total) << " MLUPS"
              << std::endl;
    if (end_ == 0) {
        save_checkpoint(0);
        vtk(2);
        vtk(1);
        drain();
        std::cout << "Simulation ended after\t" << ntime_ << " iterations which reached the maximum time step!" << std::endl;
        write_status("simulation_reached_max_step");
    } else if (end_ == 1) {
        save_checkpoint(0);
        vtk(2);
        vtk(1);
        drain();
        std::cout << "Simulation ended successfully after\t" << ntime_ << " iterations!" << std::endl;
        write_status("simulation_done");
    } else if (end_ == 2) {
        save_checkpoint(0);
        drain();
        std::cout << "Simulation data saved but not finished after\t" << ntime_ << " iterations!" << std::endl;
        write_status("continue_simulation");
    } else {
        vtk(1);
        drain();
        std::cout << "Simulation failed after\t" << ntime_ << " iterations!" << std::endl;
        write_status("simulation_failed");
    }
    std::cout << std::endl << "Code Finished " << std::endl;
    return end_ == 3 ? 1 : 0;
}

void usage() {
    std::cout << "usage: mflbm_run [--precision fp64|fp32] [--mrt 1|2|3|4] [--check]\n"
                 "Runs the simulation described by input/job_status.txt, input/simulation_control.txt\n"
                 "and input/Geometry_File_Path.txt in the current directory; outputs go to results/.\n"
                 "--mrt selects the collision model, the reference's compile-time mrt switch: 1 MRT\n"
                 "bounce-back optimised, 2 MRT original (default), 3 SRT, 4 MRT advection optimised.\n"
                 "--check reads and validates the input files without touching the GPU.\n";
}

}  // namespace

int main(int argc, char** argv) {
    int precision = MFLBM_FP64;
    int mrt = 2;
    bool check_only = false;
    for (int a = 1; a < argc; ++a) {
        const std::string arg = argv[a];
        if (arg == "--precision" && a + 1 < argc) {
            const std::string v = argv[++a];
            if (v == "fp64") precision = MFLBM_FP64;
            else if (v == "fp32") precision = MFLBM_FP32;
            else { usage(); return 2; }
        } else if (arg == "--mrt" && a + 1 < argc) {
            const std::string v = argv[++a];
            if (v.size() != 1 || v[0] < '1' || v[0] > '4') { usage(); return 2; }
            mrt = v[0] - '0';
        } else if (arg == "--check") {
            check_only = true;
        } else if (arg == "--help" || arg == "-h") {
            usage();
            return 0;
        } else {
            usage();
            return 2;
        }
    }
    try {
        std::cout << "MF-LBM colour-gradient solver, MI355X build (" << (precision == MFLBM_FP64 ? "double" : "single")
                  << " precision)" << std::endl;
        const std::string status = read_whole("input/job_status.txt");
        if (status != "new_simulation" && status != "continue_simulation")
            throw Fatal("Wrong simlation status! Exiting program!");
        std::cout << "Simulation status is: " << status << std::endl;
        const Control c = read_control("input/simulation_control.txt");
        check_control(c);
        std::cout << "Everything looks good!" << std::endl;
        if (check_only) {
            if (c.external_geometry == 1) (void)read_geometry(c);
            return 0;
        }
        Simulation sim(c, precision, mrt);
        return sim.run(status);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
