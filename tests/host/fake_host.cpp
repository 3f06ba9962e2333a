// fake_host.cpp -- a stand-in for eQ's C++ host driving the GPU field step
// through eqgpu::fenicsInterface / eqgpu::trapField (eq_amd/host/fhsl_gpu.hpp)
// in the order of Simulation::stepSimulation (src/simulation.cpp:414-510).
//
//   usage: fake_host <input.bin> <output.bin> [<files dir>]
//
// input (little-endian): int32 mode, nx, ny, nspecies, ncells, nsteps;
//   double W, H, npm, dt; double D[ns], gamma[ns]; 4 x {int32 kind, int32 pad,
//   double value, r, s}; eq_cell cells[nc]; double q[nc*ns]; double u0[ns*N].
// mode 0 (host vectors, the reference's MPI flow): each step the host adds
//   q_c to node (c mod N) of species s's solution_vector, then calls
//   stepDiffusion() on every species.
// mode 1 (device resident): each step trapField::coupleCells, then one batch
//   step for all species; the fields are downloaded at the end.
// mode 2: mode 0 with a per-node diffusion tensor: double tensor[ns][3][N]
//   (D11, D22, D12) follows u0 and is written into each species' D11/D22/D12
//   vectors before the first step (src/simulation.cpp:311-313).
// mode 3: mode 0 on eqgpu::diffusionPETSc objects (the finite-difference
//   solver, one per species) with the given walls.
// mode 4: mode 0 with the DIRICHLET_UPDATE boundary well computed by the host
//   after every step, as Simulation::computeBoundaryWell does
//   (src/simulation.cpp:581-604), and sent with setBoundaryValues; input adds
//   double wellScaling, decayRate, w0.  With a files dir, a data-recording
//   fenicsInterface (src/fHSL.cpp:27-33) also writes <dir>/data_s.pvd (scalar:
//   species 0's field) and <dir>/data_v.pvd (vector: species 0 and 1) per step.
// mode 5: mode 1 with the same well run on the device (trapField::setBoundaryWell).
// mode 6: mode 0 with flow channels above and below the trap: input adds two
//   eq_channel_desc (top, bottom); with a files dir writeDiffusionFiles also
//   writes <dir>/top<s>.pvd and <dir>/bottom<s>.pvd.
// output: double fields[ns*N], double sampled[nc*ns] (mode 1, last step),
//   double flux[ns] (modes 0-2, 4-6), int32 iterations_max; then modes 4 and 5:
//   double w[ns] (the well after the last step); mode 6: double top[ns][nx+1],
//   bottom[ns][nx+1].
// With a files dir (modes 0, 2, 3) every step also writes the solvers' files
// (writeDiffusionFiles: <dir>/hsl<s>.pvd + .vtu, or <dir>/fd<s>_petsc/grid%04d.vtr)
// and a frame of the controller's JSON log (<dir>/test-0_0-jsonData.json)
// whose externalHSL holds the species fields after the step.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../eq_amd/host/fhsl_gpu.hpp"

template <typename T>
static void rd(FILE* f, T* p, size_t n)
{
    if (fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input");
}

int main(int argc, char** argv)
{
    if (argc != 3 && argc != 4) {
        std::fprintf(stderr, "usage: fake_host <input.bin> <output.bin> [<files dir>]\n");
        return 2;
    }
    const std::string files = argc == 4 ? argv[3] : "";
    try {
        FILE* in = std::fopen(argv[1], "rb");
        if (!in) throw std::runtime_error("cannot open input");
        int32_t hdr[6];
        rd(in, hdr, 6);
        const int mode = hdr[0], nx = hdr[1], ny = hdr[2], ns = hdr[3], nc = hdr[4], nsteps = hdr[5];
        double g4[4];
        rd(in, g4, 4);
        std::vector<double> D(ns), gamma(ns);
        rd(in, D.data(), ns);
        rd(in, gamma.data(), ns);
        eq_bc walls[4];
        rd(in, walls, 4);
        std::vector<eq_cell> cells(nc);
        rd(in, cells.data(), nc);
        std::vector<double> q((size_t)nc * ns);
        rd(in, q.data(), q.size());
        const size_t N = (size_t)(nx + 1) * (ny + 1);
        std::vector<std::vector<double>> u0(ns, std::vector<double>(N));
        for (int s = 0; s < ns; ++s) rd(in, u0[s].data(), N);
        std::vector<std::vector<double>> tens;
        if (mode == 2) {
            tens.assign((size_t)ns * 3, std::vector<double>(N));
            for (auto& t : tens) rd(in, t.data(), N);
        }
        double well[3] = {0.0, 0.0, 0.0};
        if (mode == 4 || mode == 5) rd(in, well, 3);
        eq_channel_desc chans[2];
        if (mode == 6) rd(in, chans, 2);
        std::fclose(in);

        eqgpu::solverParams p;
        p.trapWidthMicrons = g4[0];
        p.trapHeightMicrons = g4[1];
        p.nodesPerMicron = g4[2];
        p.dt = g4[3];
        eqio::jsonRecording rec(0, 0);
        std::vector<eqio::frameCell> fcells(nc);
        for (int c = 0; c < nc; ++c) {
            fcells[c].id = (uint64_t)c + 1;
            fcells[c].data = {cells[c].x, cells[c].y, cells[c].angle, cells[c].length, 0.0};
        }
        if (!files.empty()) rec.init(files, "test", 0, "{\"dt\":" + eqio::json_number(p.dt) + "}");
        if (mode == 3) {
            std::vector<eqgpu::diffusionPETSc> fd(ns);
            int32_t itmax = 0;
            for (int s = 0; s < ns; ++s) {
                eqgpu::solverParams ps = p;
                ps.D_HSL = D[s];
                if (!files.empty()) ps.filePath = files + "/fd" + std::to_string(s) + "_";
                fd[s].initDiffusion(ps, walls);
                if (fd[s].solution_vector.size() != N) throw std::runtime_error("FD grid size mismatch");
                fd[s].solution_vector = u0[s];
            }
            for (int k = 0; k < nsteps; ++k) {
                for (int c = 0; c < nc; ++c)
                    for (int s = 0; s < ns; ++s) fd[s].solution_vector[(size_t)c % N] += q[(size_t)c * ns + s];
                for (int s = 0; s < ns; ++s) {
                    fd[s].stepDiffusion();
                    itmax = std::max(itmax, fd[s].lastStats().iterations_max);
                }
                if (!files.empty()) {
                    std::vector<std::vector<double>> hsl;
                    for (int s = 0; s < ns; ++s) {
                        fd[s].writeDiffusionFiles((k + 1) * p.dt);
                        hsl.push_back(fd[s].solution_vector);
                    }
                    rec.recordFrame((k + 1) * p.dt, fcells, hsl);
                }
            }
            rec.finalize();
            FILE* fo = std::fopen(argv[2], "wb");
            if (!fo) throw std::runtime_error("cannot open output");
            for (int s = 0; s < ns; ++s) std::fwrite(fd[s].solution_vector.data(), sizeof(double), N, fo);
            std::vector<double> zeros((size_t)nc * ns + ns, 0.0);
            std::fwrite(zeros.data(), sizeof(double), zeros.size(), fo);
            std::fwrite(&itmax, sizeof(int32_t), 1, fo);
            std::fclose(fo);
            return 0;
        }
        auto field = std::make_shared<eqgpu::trapField>(p, D, gamma, walls);
        if ((size_t)field->nodes() != N) throw std::runtime_error("grid size mismatch");
        std::vector<eqgpu::fenicsInterface> species;
        species.reserve(ns);
        for (int s = 0; s < ns; ++s) {
            species.emplace_back(field, s);
            eqgpu::solverParams ps = p;
            if (!files.empty() && mode != 1 && mode != 5) ps.filePath = files + "/hsl" + std::to_string(s) + ".pvd";
            if (!files.empty() && mode == 6) {
                ps.filePathTopChannel = files + "/top" + std::to_string(s) + ".pvd";
                ps.filePathBottomChannel = files + "/bottom" + std::to_string(s) + ".pvd";
            }
            if (mode == 6) field->setChannels(s, &chans[0], &chans[1]);
            species.back().initDiffusion(ps);
            if (species.back().shell->mesh->num_vertices() != N) throw std::runtime_error("num_vertices");
            species.back().solution_vector = u0[s];
            if (mode == 2) {
                *species.back().D11 = tens[(size_t)s * 3 + 0];
                *species.back().D22 = tens[(size_t)s * 3 + 1];
                *species.back().D12 = tens[(size_t)s * 3 + 2];
            }
        }
        std::vector<double> sampled;
        if (mode == 1 || mode == 5)
            for (int s = 0; s < ns; ++s) field->upload(s, u0[s]);
        // the host's boundary well (Simulation::initBoundaryWell / computeBoundaryWell)
        std::vector<double> wellConc(ns, well[2]);
        if (mode == 4)
            for (int s = 0; s < ns; ++s) species[s].setBoundaryValues(wellConc[s]);
        if (mode == 5)
            for (int s = 0; s < ns; ++s) field->setBoundaryWell(s, well[0], well[1], well[2]);
        // data recording on the controller (src/simulation.cpp:557-571)
        std::unique_ptr<eqgpu::fenicsInterface> recorder;
        if (mode == 4 && !files.empty()) {
            eqgpu::solverParams pd = p;
            pd.filePath = files + "/";
            eqgpu::dataRecord sc{"data_s.pvd", &species[0].solution_vector, nullptr, (size_t)nx + 1, (size_t)ny + 1,
                                 p.nodesPerMicron};
            eqgpu::dataRecord vc{"data_v.pvd", &species[0].solution_vector, &species[ns > 1 ? 1 : 0].solution_vector,
                                 (size_t)nx + 1, (size_t)ny + 1, p.nodesPerMicron};
            pd.dataFiles = {sc, vc};
            recorder = std::make_unique<eqgpu::fenicsInterface>(pd);
        }
        for (int k = 0; k < nsteps; ++k) {
            if (mode == 4) {
                for (int c = 0; c < nc; ++c)
                    for (int s = 0; s < ns; ++s) species[s].solution_vector[(size_t)c % N] += q[(size_t)c * ns + s];
                for (int s = 0; s < ns; ++s) {
                    species[s].stepDiffusion();
                    const double flux = species[s].getBoundaryFlux().at("totalFlux");
                    wellConc[s] += flux / well[0];
                    wellConc[s] -= p.dt * well[1] * wellConc[s];
                    if (wellConc[s] < 0.0) wellConc[s] = 0.0;
                    species[s].setBoundaryValues(wellConc[s]);
                }
                if (recorder) recorder->writeDataFiles((k + 1) * p.dt);
            } else if (mode == 0 || mode == 2 || mode == 6) {
                // controller: updateCells writes into its copy of each field,
                // then the copies travel to the solvers (src/simulation.cpp:447-463)
                for (int c = 0; c < nc; ++c)
                    for (int s = 0; s < ns; ++s) species[s].solution_vector[(size_t)c % N] += q[(size_t)c * ns + s];
                for (int s = 0; s < ns; ++s) species[s].stepDiffusion();
                if (!files.empty()) {
                    std::vector<std::vector<double>> hsl;
                    for (int s = 0; s < ns; ++s) {
                        species[s].writeDiffusionFiles((k + 1) * p.dt);
                        hsl.push_back(species[s].solution_vector);
                    }
                    rec.recordFrame((k + 1) * p.dt, fcells, hsl);
                }
            } else {
                field->coupleCells(cells, q, &sampled);
                field->step();
            }
        }
        rec.finalize();
        std::vector<double> out;
        for (int s = 0; s < ns; ++s) {
            std::vector<double> u;
            if (mode == 1 || mode == 5) field->download(s, u);
            else u = species[s].solution_vector;
            out.insert(out.end(), u.begin(), u.end());
        }
        FILE* fo = std::fopen(argv[2], "wb");
        if (!fo) throw std::runtime_error("cannot open output");
        std::fwrite(out.data(), sizeof(double), out.size(), fo);
        sampled.resize((size_t)nc * ns, 0.0);
        std::fwrite(sampled.data(), sizeof(double), sampled.size(), fo);
        for (int s = 0; s < ns; ++s) {
            const double fl = field->boundaryFlux(s);
            std::fwrite(&fl, sizeof(double), 1, fo);
        }
        const int32_t it = field->lastStats().iterations_max;
        std::fwrite(&it, sizeof(int32_t), 1, fo);
        if (mode == 5)
            for (int s = 0; s < ns; ++s) wellConc[s] = field->boundaryWellConcentration(s);
        if (mode == 4 || mode == 5) std::fwrite(wellConc.data(), sizeof(double), ns, fo);
        if (mode == 6) {
            std::vector<std::vector<double>> tb(2 * ns);
            for (int s = 0; s < ns; ++s) field->channelValues(s, tb[s], tb[ns + s]);
            for (auto& v : tb) std::fwrite(v.data(), sizeof(double), v.size(), fo);
        }
        std::fclose(fo);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fake_host: %s\n", e.what());
        return 1;
    }
    return 0;
}
