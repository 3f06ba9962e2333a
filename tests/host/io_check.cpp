// io_check.cpp -- drives the file-format writers of eq_amd/host/field_io.hpp
// on fields given in a file, so a test can read the files back and compare.
//
//   usage: io_check <input.bin> <dir>
//
// input (little-endian): int32 nx, ny, nsteps, ncells; double W, H, dt;
//   double u[nsteps][(nx+1)(ny+1)]; double cells[ncells][5].
// writes: <dir>/z.pvd + z000000.vtu.. (compressed), <dir>/r.pvd + r*.vtu
// (raw binary), <dir>/g%04d.vtr (one per step, raw), and
// <dir>/io-3_7-jsonData.json with one frame per step whose externalHSL is
// {u[k], 2 u[k]} and whose divisions hold one record per cell.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../eq_amd/host/field_io.hpp"

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: io_check <input.bin> <dir>\n");
        return 2;
    }
    try {
        FILE* in = std::fopen(argv[1], "rb");
        if (!in) throw std::runtime_error("cannot open input");
        int32_t hdr[4];
        double g[3];
        if (std::fread(hdr, 4, 4, in) != 4 || std::fread(g, 8, 3, in) != 3) throw std::runtime_error("short header");
        const int nx = hdr[0], ny = hdr[1], nsteps = hdr[2], nc = hdr[3];
        const size_t N = (size_t)(nx + 1) * (ny + 1);
        std::vector<double> u((size_t)nsteps * N), cells((size_t)nc * 5);
        if (std::fread(u.data(), 8, u.size(), in) != u.size() || std::fread(cells.data(), 8, cells.size(), in) != cells.size())
            throw std::runtime_error("short input");
        std::fclose(in);
        const std::string dir = argv[2];
        eqio::pvdSeries z(dir + "/z.pvd", nx, ny, g[0], g[1], "u", true);
        eqio::pvdSeries r(dir + "/r", nx, ny, g[0], g[1], "u", false);
        eqio::jsonRecording rec(3, 7);
        rec.init(dir, "io", 1234567890123L, "{\"dt\":" + eqio::json_number(g[2]) + ",\"name\":\"io \\\"check\\\"\"}");
        std::vector<eqio::frameCell> fc(nc);
        std::vector<std::vector<double>> div(nc);
        for (int c = 0; c < nc; ++c) {
            fc[c].id = (uint64_t)c * 1000003u;
            fc[c].strainType = (uint64_t)(c % 3);
            fc[c].data.assign(cells.begin() + 5 * c, cells.begin() + 5 * c + 5);
            div[c] = {g[2] * c, (double)c, cells[5 * c + 3], (double)(c + 1), 0.5 * cells[5 * c + 3]};
        }
        for (int k = 0; k < nsteps; ++k) {
            const double* uk = u.data() + (size_t)k * N;
            const double t = (k + 1) * g[2];
            z.write(uk, t);
            r.write(uk, t);
            eqio::write_vtr(dir, "g", k + 1, nx, ny, g[0] / nx, uk);
            std::vector<double> twice(uk, uk + N);
            for (double& v : twice) v *= 2.0;
            rec.recordFrame(t, fc, {std::vector<double>(uk, uk + N), twice}, k == 0 ? div : std::vector<std::vector<double>>{});
        }
        rec.finalize();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "io_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
