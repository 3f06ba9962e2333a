// fhsl_gpu.hpp -- C++ host mirror of eQ's diffusion-solver surface over the
// C ABI of libeq_field.so (include/eq_field.h).
//
// The reference host (src/simulation.cpp) creates one fenicsInterface per
// HSL species (src/simulation.cpp:181-394) and touches these members:
//   initDiffusion(params&)                       src/fHSL.cpp:37-53
//   stepDiffusion()                              src/fHSL.cpp:98-161
//   shell->mesh->num_vertices(), shell->mesh_coords,
//   shell->dof_from_vertex                       src/simulation.cpp:298-308, src/fHSL.h:82-88
//   D11 / D22 / D12                              src/simulation.cpp:311-313, 503-505
//   solution_vector                              src/simulation.cpp:492, 501
//   totalBoundaryFlux, getBoundaryFlux()          src/simulation.cpp:546, 583; src/eQ.h:327
//   setBoundaryValues(bval)                       src/simulation.cpp:602; src/fHSL.cpp:601-603
//   writeDiffusionFiles (+ channel series), writeDataFiles,
//   the data-recording constructor, finalize      src/simulation.cpp:536, 569, 578, 683
// eqgpu::fenicsInterface keeps those names and meanings.  All species of a
// trap share one device batch (eqgpu::trapField), so the solve of every
// species is one set of launches, and the controller's cell loop
// (eQabm::updateCells, src/abm/eQabm.cpp:234-449) becomes one call,
// trapField::coupleCells.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eq_field.h"
#include "field_io.hpp"

namespace eqgpu {

inline void check(int rc, const char* what)
{
    if (rc != EQ_OK) throw std::runtime_error(std::string(what) + ": " + eq_last_error());
}

// One data file of the controller's data-recording grid (eQ::data::record,
// src/eQ.h:95-104): a file name and the eQ::data::tensor it samples.  The
// tensor's grids are natural-order arrays of (y*n + 1) x (x*n + 1) nodes at n
// nodes per micron (src/eQ.h:147-157); a vector tensor has a second grid.
struct dataRecord {
    std::string fileName;
    const std::vector<double>* grid0 = nullptr;
    const std::vector<double>* grid1 = nullptr; // non-null: VECTOR rank
    size_t nodesWide = 0, nodesHigh = 0;
    double nodesPerMicron = 1.0;
    // tensor::eval / evalVector: the grid value at ij_from_xy(x, y, n)
    // (src/eQ.h:119-124, 162-175); indices past the grid are clamped to it
    double at(const std::vector<double>& g, double x, double y) const
    {
        size_t j = (size_t)std::llround(std::max(0.0, x * nodesPerMicron));
        size_t i = (size_t)std::llround(std::max(0.0, y * nodesPerMicron));
        j = std::min(j, nodesWide - 1);
        i = std::min(i, nodesHigh - 1);
        return g[j + i * nodesWide];
    }
};

// eQ::diffusionSolver::params (src/eQ.h:305-321) without the MPI and JSON
// handles, which a single-GPU process does not need
struct solverParams {
    size_t uniqueID = 0;
    double dt = 0.1;
    double D_HSL = 0.0;
    std::string filePath;
    std::string filePathTopChannel;    // channel file series (MICROFLUIDIC_TRAP, src/fHSL.cpp:613-627)
    std::string filePathBottomChannel;
    std::vector<dataRecord> dataFiles; // data-recording grid (src/fHSL.cpp:288-307)
    double trapHeightMicrons = 0.0;
    double trapWidthMicrons = 0.0;
    double nodesPerMicron = 1.0;
    double trapChannelVelocity = 0.0;
};

// One trap, all of its species, on the device.
class trapField {
public:
    // walls in eQ::boundaryCondition::wall order: left, right, top, bottom
    trapField(const solverParams& base, const std::vector<double>& D, const std::vector<double>& gamma,
              const eq_bc walls[4], const eq_solver_desc* solver = nullptr)
    {
        // nodesW/H = ceil(W*npm) cells, +1 vertices (src/fHSL.cpp:240-243, 285-287)
        nx_ = (int)std::ceil(base.trapWidthMicrons * base.nodesPerMicron);
        ny_ = (int)std::ceil(base.trapHeightMicrons * base.nodesPerMicron);
        npm_ = base.nodesPerMicron;
        width_ = base.trapWidthMicrons;
        height_ = base.trapHeightMicrons;
        std::vector<eq_field_desc> descs(D.size());
        for (size_t s = 0; s < D.size(); ++s) {
            eq_field_desc& d = descs[s];
            d.nx = nx_;
            d.ny = ny_;
            d.width = base.trapWidthMicrons;
            d.height = base.trapHeightMicrons;
            d.dt = base.dt;
            d.D = D[s];
            d.gamma = gamma.empty() ? 0.0 : gamma[s];
            d.f = 0.0; // f is bound to Constant(0) in the reference (src/fHSL.h:201)
            for (int w = 0; w < 4; ++w) d.wall[w] = walls[w];
        }
        eq_batch* b = nullptr;
        check(eq_batch_create(descs.data(), (int32_t)descs.size(), (int32_t)descs.size(), solver, &b),
              "eq_batch_create");
        batch_.reset(b);
    }

    int32_t species() const { return eq_batch_num_problems(batch_.get()); }
    int nx() const { return nx_; }
    int ny() const { return ny_; }
    double width() const { return width_; }
    double height() const { return height_; }
    size_t nodes() const { return (size_t)(nx_ + 1) * (size_t)(ny_ + 1); }
    eq_batch* handle() const { return batch_.get(); }

    void upload(int s, const std::vector<double>& u) { check(eq_batch_upload(batch_.get(), s, u.data(), u.size()), "upload"); }
    void download(int s, std::vector<double>& u)
    {
        u.resize(nodes());
        check(eq_batch_download(batch_.get(), s, u.data(), u.size()), "download");
    }
    eq_step_stats step()
    {
        eq_step_stats st{};
        check(eq_batch_step(batch_.get(), &st), "eq_batch_step");
        last_ = st;
        return st;
    }
    // sampled and counts may be null; q holds nc x species amounts
    void coupleCells(const std::vector<eq_cell>& cells, const std::vector<double>& q, std::vector<double>* sampled,
                     std::vector<int32_t>* counts = nullptr)
    {
        const size_t nc = cells.size();
        if (sampled) sampled->resize(nc * species());
        if (counts) counts->resize(nc);
        check(eq_batch_couple(batch_.get(), 0, cells.data(), nc, npm_, q.data(), sampled ? sampled->data() : nullptr,
                              counts ? counts->data() : nullptr, EQ_COUPLE_SAMPLE | EQ_COUPLE_DEPOSIT),
              "eq_batch_couple");
    }
    double boundaryFlux(int s)
    {
        double v = 0.0;
        check(eq_batch_boundary_flux(batch_.get(), s, &v), "boundary flux");
        return v;
    }
    // setBoundaryValues of one species (stream ordered on the device)
    void setBoundaryValue(int s, double v) { check(eq_batch_set_boundary_value(batch_.get(), s, v), "boundary value"); }
    // Simulation::computeBoundaryWell run on the device after every step
    // (DIRICHLET_UPDATE): wellScaling, boundaryDecayRate, starting value
    void setBoundaryWell(int s, double wellScaling, double decayRate, double w0 = 0.0)
    {
        const eq_well_desc w{wellScaling, decayRate, w0};
        check(eq_batch_set_well(batch_.get(), s, &w), "boundary well");
    }
    double boundaryWellConcentration(int s)
    {
        double w = 0.0;
        check(eq_batch_well_state(batch_.get(), s, &w, nullptr, nullptr), "well state");
        return w;
    }
    // flow channels above/below the trap (MICROFLUIDIC_TRAP); null: no channel
    void setChannels(int s, const eq_channel_desc* top, const eq_channel_desc* bottom)
    {
        check(eq_batch_set_channels(batch_.get(), s, top, bottom), "channels");
        if ((int)chan_.size() <= s) chan_.resize(s + 1, 0);
        chan_[s] = (top ? 1 : 0) | (bottom ? 2 : 0);
    }
    int channels(int s) const { return s < (int)chan_.size() ? chan_[s] : 0; }
    // topChannelData / bottomChannelData (nx + 1 values each)
    void channelValues(int s, std::vector<double>& top, std::vector<double>& bottom)
    {
        top.resize((size_t)nx_ + 1);
        bottom.resize((size_t)nx_ + 1);
        check(eq_batch_channel_values(batch_.get(), s, top.data(), bottom.data()), "channel values");
    }

    // Per-species host vectors registered by the fenicsInterface views.  The
    // first species to step in a time step uploads every registered vector,
    // solves all species in one batch and downloads every result; the other
    // species' calls of the same time step find their vector already solved.
    void attach(int s, std::vector<double>* v)
    {
        if ((int)host_.size() <= s) {
            host_.resize(s + 1, nullptr);
            solved_.resize(s + 1, 0);
        }
        host_[s] = v;
    }
    // Per-species tensor vectors (D11, D22, D12 of the fenicsInterface views,
    // src/fHSL.cpp:313-323): sent to the device before a solve whenever they
    // differ from what the device holds; all (1, 1, 0) means constant
    // coefficients, the reference's default.
    void attachTensor(int s, const std::shared_ptr<std::vector<double>>* d11,
                      const std::shared_ptr<std::vector<double>>* d22, const std::shared_ptr<std::vector<double>>* d12)
    {
        if ((int)tens_.size() <= s) tens_.resize(s + 1);
        tens_[s].src[0] = d11;
        tens_[s].src[1] = d22;
        tens_[s].src[2] = d12;
    }
    void syncTensors()
    {
        for (size_t k = 0; k < tens_.size(); ++k) {
            Tens& t = tens_[k];
            if (!t.src[0] || !*t.src[0] || !*t.src[1] || !*t.src[2]) continue;
            const std::vector<double>& a = **t.src[0];
            const std::vector<double>& b = **t.src[1];
            const std::vector<double>& c = **t.src[2];
            if (a == t.sent[0] && b == t.sent[1] && c == t.sent[2]) continue;
            bool iso = true;
            for (size_t q = 0; q < a.size() && iso; ++q) iso = a[q] == 1.0 && b[q] == 1.0 && c[q] == 0.0;
            if (iso)
                check(eq_batch_set_tensor(batch_.get(), (int)k, nullptr, nullptr, nullptr, 0), "set tensor");
            else
                check(eq_batch_set_tensor(batch_.get(), (int)k, a.data(), b.data(), c.data(), a.size()), "set tensor");
            t.sent[0] = a;
            t.sent[1] = b;
            t.sent[2] = c;
        }
    }
    void stepSpecies(int s)
    {
        if (!solved_[s]) {
            for (size_t k = 0; k < host_.size(); ++k)
                if (host_[k]) upload((int)k, *host_[k]);
            syncTensors();
            step();
            for (size_t k = 0; k < host_.size(); ++k)
                if (host_[k]) {
                    download((int)k, *host_[k]);
                    solved_[k] = 1;
                }
        }
        solved_[s] = 0;
    }
    const eq_step_stats& lastStats() const { return last_; }

private:
    struct Tens {
        const std::shared_ptr<std::vector<double>>* src[3] = {nullptr, nullptr, nullptr};
        std::vector<double> sent[3]; // last values sent (empty: constant coefficients)
    };
    std::vector<Tens> tens_;
    std::vector<int> chan_;
    std::vector<std::vector<double>*> host_;
    std::vector<char> solved_;
    eq_step_stats last_{};
    struct deleter {
        void operator()(eq_batch* b) const { eq_batch_destroy(b); }
    };
    std::unique_ptr<eq_batch, deleter> batch_;
    int nx_ = 0, ny_ = 0;
    double npm_ = 1.0, width_ = 0.0, height_ = 0.0;
};

// The per-species surface the reference host uses.
class fenicsInterface {
public:
    struct meshView {
        size_t n = 0;
        size_t num_vertices() const { return n; }
    };
    struct shellView {
        std::shared_ptr<meshView> mesh = std::make_shared<meshView>();
        std::vector<double> mesh_coords;  // 2N, vertex order
        std::vector<int> dof_from_vertex; // N (identity)
    };

    fenicsInterface(std::shared_ptr<trapField> field, int species) : field_(std::move(field)), s_(species) {}

    // Data recording only (src/fHSL.cpp:27-33, 288-307): no field and no
    // solve; a P1 mesh of ceil(W*n) x ceil(H*n) cells at the params' nodes per
    // micron and one .pvd series per data file at filePath + fileName.
    explicit fenicsInterface(const solverParams& p) : params_(p), recording_(true)
    {
        nx_ = (int)std::ceil(p.trapWidthMicrons * p.nodesPerMicron);
        ny_ = (int)std::ceil(p.trapHeightMicrons * p.nodesPerMicron);
        const size_t n = (size_t)(nx_ + 1) * (ny_ + 1);
        shell = std::make_shared<shellView>();
        shell->mesh->n = n;
        for (const dataRecord& d : p.dataFiles)
            data_.emplace_back(p.filePath + d.fileName, nx_, ny_, p.trapWidthMicrons, p.trapHeightMicrons,
                               d.grid1 ? "v" : "u");
    }

    void initDiffusion(const solverParams& p)
    {
        params_ = p;
        const size_t n = field_->nodes();
        shell = std::make_shared<shellView>();
        shell->mesh->n = n;
        shell->mesh_coords.resize(2 * n);
        shell->dof_from_vertex.resize(n);
        check(eq_batch_vertex_map(field_->handle(), s_, shell->mesh_coords.data(), shell->dof_from_vertex.data()),
              "vertex map");
        solution_vector.assign(n, 0.0);
        field_->attach(s_, &solution_vector);
        // isotropic default tensor (src/fHSL.cpp:313-323)
        D11 = std::make_shared<std::vector<double>>(n, 1.0);
        D22 = std::make_shared<std::vector<double>>(n, 1.0);
        D12 = std::make_shared<std::vector<double>>(n, 0.0);
        field_->attachTensor(s_, &D11, &D22, &D12);
        // dolfin::File(comm, filePath, "compressed") (src/fHSL.cpp:606-613)
        if (!p.filePath.empty())
            files_ = eqio::pvdSeries(p.filePath, field_->nx(), field_->ny(), field_->width(), field_->height());
        nx_ = field_->nx();
        ny_ = field_->ny();
    }

    // u0 <- solution_vector; solve; solution_vector <- u; boundary flux
    // (src/fHSL.cpp:104-108, 156-160).  Call once per species per time step.
    void stepDiffusion()
    {
        field_->stepSpecies(s_);
        totalBoundaryFlux = field_->boundaryFlux(s_);
    }
    // one .vtu of solution_vector per call, listed in the .pvd at timeStamp,
    // then the top and bottom channel fields when this species has channels
    // and the params name their files (src/fHSL.cpp:613-634)
    void writeDiffusionFiles(double timeStamp)
    {
        if (files_.open()) files_.write(solution_vector.data(), timeStamp);
        const int ch = field_->channels(s_);
        if (!ch) return;
        const bool top = (ch & 1) && !params_.filePathTopChannel.empty();
        const bool bottom = (ch & 2) && !params_.filePathBottomChannel.empty();
        if (!top && !bottom) return;
        std::vector<double> t, b;
        field_->channelValues(s_, t, b);
        if (top) {
            if (!topFiles_.open())
                topFiles_ = eqio::pvdSeries(params_.filePathTopChannel, field_->nx(), 0, field_->width(), 0.0);
            topFiles_.write(t.data(), timeStamp);
        }
        if (bottom) {
            if (!bottomFiles_.open())
                bottomFiles_ = eqio::pvdSeries(params_.filePathBottomChannel, field_->nx(), 0, field_->width(), 0.0);
            bottomFiles_.write(b.data(), timeStamp);
        }
    }
    // setBoundaryValues(bval) (src/fHSL.cpp:601-603): every Dirichlet wall of
    // this species takes bval from the next solve on
    void setBoundaryValues(double bval) { field_->setBoundaryValue(s_, bval); }
    // getBoundaryFlux (src/eQ.h:327): at least "totalFlux"
    std::map<std::string, double> getBoundaryFlux() const { return {{"totalFlux", totalBoundaryFlux}}; }
    // writeDataFiles (src/fHSL.cpp:637-654): each data tensor interpolated on
    // the mesh vertices (x_j, y_i) = (j W/nx, i H/ny), one .vtu per file per call
    void writeDataFiles(double t)
    {
        if (!recording_) return;
        const size_t n = (size_t)(nx_ + 1) * (ny_ + 1);
        std::vector<double> a(n), b;
        for (size_t k = 0; k < params_.dataFiles.size(); ++k) {
            const dataRecord& d = params_.dataFiles[k];
            if (d.grid1) b.resize(n);
            for (int i = 0; i <= ny_; ++i)
                for (int j = 0; j <= nx_; ++j) {
                    const double x = 0.0 + ((double)j * (params_.trapWidthMicrons - 0.0)) / (double)nx_;
                    const double y = 0.0 + ((double)i * (params_.trapHeightMicrons - 0.0)) / (double)ny_;
                    const size_t q = (size_t)j + (size_t)i * (nx_ + 1);
                    a[q] = d.at(*d.grid0, x, y);
                    if (d.grid1) b[q] = d.at(*d.grid1, x, y);
                }
            data_[k].write(a.data(), t, d.grid1 ? b.data() : nullptr);
        }
    }
    void finalize() {}

    std::shared_ptr<shellView> shell;
    std::vector<double> solution_vector;
    std::shared_ptr<std::vector<double>> D11, D22, D12;
    double totalBoundaryFlux = 0.0;

private:
    std::shared_ptr<trapField> field_;
    int s_ = 0;
    solverParams params_;
    bool recording_ = false;
    int nx_ = 0, ny_ = 0;
    eqio::pvdSeries files_, topFiles_, bottomFiles_;
    std::vector<eqio::pvdSeries> data_;
};

// diffusionPETSc (diffuclass.h:45-108): the finite-difference alternative
// behind the same eQ::diffusionSolver surface.  One species per object, as in
// the reference; the grid is (W/h + 1) x (H/h + 1) nodes with h = 1/npm
// (diffuclass.cpp:356-358) and the walls default to Dirichlet 0 (the
// DIRICHLET_0 set-up, diffuclass.cpp:70-89).  solution_vector is in natural
// order, as the reference's (96-101).
class diffusionPETSc {
public:
    diffusionPETSc() = default;

    void initDiffusion(const solverParams& p, const eq_bc* walls = nullptr, const eq_solver_desc* solver = nullptr)
    {
        params_ = p;
        const double h = 1.0 / p.nodesPerMicron;
        eq_field_desc d{};
        d.nx = (int32_t)(p.trapWidthMicrons / h);
        d.ny = (int32_t)(p.trapHeightMicrons / h);
        d.width = d.nx * h;
        d.height = d.ny * h;
        d.dt = p.dt;
        d.D = p.D_HSL;
        d.scheme = EQ_SCHEME_FD;
        for (int w = 0; w < 4; ++w) {
            if (walls) {
                d.wall[w] = walls[w];
            } else {
                d.wall[w] = eq_bc{};
                d.wall[w].kind = EQ_BC_DIRICHLET;
            }
        }
        eq_batch* b = nullptr;
        check(eq_batch_create(&d, 1, 1, solver, &b), "eq_batch_create (FD)");
        batch_.reset(b);
        solution_vector.assign((size_t)(d.nx + 1) * (d.ny + 1), 0.0);
        nx_ = d.nx;
        ny_ = d.ny;
        h_ = h;
        step_ = 0;
        // output folder filePath + "petsc", objects named "grid" (diffuclass.cpp:62-66)
        if (!p.filePath.empty()) {
            dir_ = p.filePath + "petsc";
            std::filesystem::create_directories(dir_);
        }
    }
    // TimeStep: boundary values into the right-hand side, solve, read back
    // (diffuclass.cpp:107-117, 405-413)
    void stepDiffusion()
    {
        check(eq_batch_upload(batch_.get(), 0, solution_vector.data(), solution_vector.size()), "upload");
        eq_step_stats st{};
        check(eq_batch_step(batch_.get(), &st), "eq_batch_step (FD)");
        last_ = st;
        check(eq_batch_download(batch_.get(), 0, solution_vector.data(), solution_vector.size()), "download");
        ++step_;
    }
    double getDiffusionConstant() const { return params_.D_HSL; }
    const eq_step_stats& lastStats() const { return last_; }
    // RecordData: "<filePath>petsc/grid%04d.vtr" numbered by the steps taken
    // (diffuclass.cpp:168-172, 401-408, 560-583)
    void writeDiffusionFiles(double)
    {
        if (!dir_.empty()) eqio::write_vtr(dir_, "grid", step_, nx_, ny_, h_, solution_vector.data());
    }
    void finalize() {}

    std::vector<double> solution_vector;

private:
    struct deleter {
        void operator()(eq_batch* b) const { eq_batch_destroy(b); }
    };
    std::unique_ptr<eq_batch, deleter> batch_;
    solverParams params_;
    eq_step_stats last_{};
    std::string dir_;
    int nx_ = 0, ny_ = 0, step_ = 0;
    double h_ = 1.0;
};

} // namespace eqgpu
