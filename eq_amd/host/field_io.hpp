// field_io.hpp -- the on-disk and wire formats around the field step, for the
// host mirror (fhsl_gpu.hpp) and any other caller of libeq_field.so.
//
// Three writers, each following the file the reference produces:
//   pvdSeries      -- dolfin::File(comm, path, "compressed") << (u, t)
//                     (src/fHSL.cpp:606-636): a ParaView collection (.pvd)
//                     naming one VTK XML unstructured grid (.vtu) per call,
//                     the P1 field on the "right"-diagonal triangulation
//                     (RectangleMesh(..., "right"), src/fHSL.cpp:164-173),
//                     point data zlib-compressed and base64-encoded.
//   write_vtr      -- diffusionPETSc::RecordData (diffuclass.cpp:560-583):
//                     "<dir>/<name>%04d.vtr", a VTK XML rectilinear grid of
//                     the finite-difference nodes.
//   jsonRecording  -- the controller's frame log (src/main.cpp:81-187): one
//                     JSON document of frames, each with the cells and
//                     "externalHSL", the species fields in grid order
//                     (src/main.cpp:148-162).  Frames are streamed to the
//                     file as they are recorded instead of being held in
//                     memory until finalize (a 33.5 M-node field is about
//                     0.7 GB of text per species and frame); the finished
//                     file has the same keys in the same (sorted) order as
//                     the reference's nlohmann::json output.
// Every field is in natural order j + i (nx + 1), the order eq_*_download
// returns and, with the identity vertex map, the reference's DOF order.
#pragma once

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace eqio {

inline std::string base64(const unsigned char* p, size_t n)
{
    static const char* tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve(4 * ((n + 2) / 3));
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
        out += tab[v >> 18];
        out += tab[(v >> 12) & 63];
        out += tab[(v >> 6) & 63];
        out += tab[v & 63];
    }
    if (i < n) {
        const uint32_t v = (uint32_t)p[i] << 16 | (i + 1 < n ? (uint32_t)p[i + 1] << 8 : 0u);
        out += tab[v >> 18];
        out += tab[(v >> 12) & 63];
        out += (i + 1 < n) ? tab[(v >> 6) & 63] : '=';
        out += '=';
    }
    return out;
}

// Payload of a binary VTK XML DataArray with UInt32 headers.  Compressed
// (vtkZLibDataCompressor, one block): header [1, n, n, compressed size],
// base64-encoded on its own, then the compressed block.  Raw: header [n] and
// the data, base64-encoded together.
inline std::string vtk_payload(const void* data, size_t nbytes, bool compressed)
{
    if (nbytes > 0xffffffffu) throw std::runtime_error("vtk_payload: array over 4 GiB needs UInt64 headers");
    if (!compressed) {
        std::vector<unsigned char> buf(4 + nbytes);
        const uint32_t n = (uint32_t)nbytes;
        std::memcpy(buf.data(), &n, 4);
        if (nbytes) std::memcpy(buf.data() + 4, data, nbytes);
        return base64(buf.data(), buf.size());
    }
    uLongf clen = compressBound((uLong)nbytes);
    std::vector<unsigned char> z(clen);
    if (compress2(z.data(), &clen, (const Bytef*)data, (uLong)nbytes, Z_DEFAULT_COMPRESSION) != Z_OK)
        throw std::runtime_error("vtk_payload: zlib compress2 failed");
    const uint32_t hdr[4] = {1u, (uint32_t)nbytes, (uint32_t)nbytes, (uint32_t)clen};
    return base64((const unsigned char*)hdr, sizeof(hdr)) + base64(z.data(), clen);
}

inline void put_array(std::ofstream& f, const char* type, const char* name, int ncomp, const void* data,
                      size_t nbytes, bool compressed)
{
    f << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\"" << ncomp
      << "\" format=\"binary\">" << vtk_payload(data, nbytes, compressed) << "</DataArray>\n";
}

inline const char* vtk_header(bool compressed)
{
    return compressed ? " version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\" "
                        "compressor=\"vtkZLibDataCompressor\""
                      : " version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\"";
}

// P1 field of an nx x ny cell trap W x H on the "right" triangulation: cell
// (i, j) splits along its bottom-left -> top-right diagonal into
// (v00, v01, v11) and (v00, v11, v10), 2 nx ny triangles; points are the
// (nx + 1)(ny + 1) nodes in natural order at (j W / nx, i H / ny, 0).
// uy non-null: a vector field (ux, uy), written as 3 components with z = 0,
// the layout dolfin::File uses for a 2-D vector function.
inline void write_vtu(const std::string& path, int nx, int ny, double W, double H, const double* u,
                      const std::string& name, bool compressed = true, const double* uy = nullptr)
{
    if (nx < 1 || ny < 1) throw std::runtime_error("write_vtu: empty grid");
    const size_t np = (size_t)(nx + 1) * (ny + 1), nc = 2 * (size_t)nx * ny;
    std::vector<double> pts(3 * np);
    for (int i = 0; i <= ny; ++i)
        for (int j = 0; j <= nx; ++j) {
            const size_t k = (size_t)j + (size_t)i * (nx + 1);
            pts[3 * k] = W * j / nx;
            pts[3 * k + 1] = H * i / ny;
            pts[3 * k + 2] = 0.0;
        }
    std::vector<int64_t> conn(3 * nc), offs(nc);
    std::vector<uint8_t> types(nc, 5); // VTK_TRIANGLE
    size_t c = 0;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nx; ++j) {
            const int64_t v00 = j + (int64_t)i * (nx + 1), v01 = v00 + 1, v10 = v00 + nx + 1, v11 = v10 + 1;
            const int64_t tri[2][3] = {{v00, v01, v11}, {v00, v11, v10}};
            for (int t = 0; t < 2; ++t, ++c) {
                for (int q = 0; q < 3; ++q) conn[3 * c + q] = tri[t][q];
                offs[c] = 3 * (int64_t)(c + 1);
            }
        }
    std::ofstream f(path, std::ios::trunc);
    if (!f) throw std::runtime_error("write_vtu: cannot open " + path);
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\"" << vtk_header(compressed) << ">\n"
      << "<UnstructuredGrid>\n<Piece NumberOfPoints=\"" << np << "\" NumberOfCells=\"" << nc << "\">\n<Points>\n";
    put_array(f, "Float64", "Points", 3, pts.data(), pts.size() * 8, compressed);
    f << "</Points>\n<Cells>\n";
    put_array(f, "Int64", "connectivity", 1, conn.data(), conn.size() * 8, compressed);
    put_array(f, "Int64", "offsets", 1, offs.data(), offs.size() * 8, compressed);
    put_array(f, "UInt8", "types", 1, types.data(), types.size(), compressed);
    if (uy) {
        std::vector<double> v(3 * np);
        for (size_t k = 0; k < np; ++k) {
            v[3 * k] = u[k];
            v[3 * k + 1] = uy[k];
            v[3 * k + 2] = 0.0;
        }
        f << "</Cells>\n<PointData Vectors=\"" << name << "\">\n";
        put_array(f, "Float64", name.c_str(), 3, v.data(), v.size() * 8, compressed);
    } else {
        f << "</Cells>\n<PointData Scalars=\"" << name << "\">\n";
        put_array(f, "Float64", name.c_str(), 1, u, np * 8, compressed);
    }
    f << "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    if (!f) throw std::runtime_error("write_vtu: write failed for " + path);
}

// A 1-D P1 field on the interval mesh of a flow channel (fenicsChannel's
// IntervalMesh over the trap's columns, src/fHSL.h:240-332): nx + 1 points at
// (j W / nx, 0, 0) and nx line cells.
inline void write_vtu_line(const std::string& path, int nx, double W, const double* u, const std::string& name,
                           bool compressed = true)
{
    if (nx < 1) throw std::runtime_error("write_vtu_line: empty mesh");
    const size_t np = (size_t)nx + 1, nc = (size_t)nx;
    std::vector<double> pts(3 * np, 0.0);
    for (size_t j = 0; j < np; ++j) pts[3 * j] = W * (double)j / nx;
    std::vector<int64_t> conn(2 * nc), offs(nc);
    std::vector<uint8_t> types(nc, 3); // VTK_LINE
    for (size_t c = 0; c < nc; ++c) {
        conn[2 * c] = (int64_t)c;
        conn[2 * c + 1] = (int64_t)c + 1;
        offs[c] = 2 * (int64_t)(c + 1);
    }
    std::ofstream f(path, std::ios::trunc);
    if (!f) throw std::runtime_error("write_vtu_line: cannot open " + path);
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\"" << vtk_header(compressed) << ">\n"
      << "<UnstructuredGrid>\n<Piece NumberOfPoints=\"" << np << "\" NumberOfCells=\"" << nc << "\">\n<Points>\n";
    put_array(f, "Float64", "Points", 3, pts.data(), pts.size() * 8, compressed);
    f << "</Points>\n<Cells>\n";
    put_array(f, "Int64", "connectivity", 1, conn.data(), conn.size() * 8, compressed);
    put_array(f, "Int64", "offsets", 1, offs.data(), offs.size() * 8, compressed);
    put_array(f, "UInt8", "types", 1, types.data(), types.size(), compressed);
    f << "</Cells>\n<PointData Scalars=\"" << name << "\">\n";
    put_array(f, "Float64", name.c_str(), 1, u, np * 8, compressed);
    f << "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
    if (!f) throw std::runtime_error("write_vtu_line: write failed for " + path);
}

// A .pvd collection with one .vtu per write, named <stem>NNNNNN.vtu beside
// it (stem = the .pvd path without its extension), as dolfin::File does for
// a ".pvd" path.  The collection file is rewritten after every write, so it
// is complete whenever the run stops.
class pvdSeries {
public:
    pvdSeries() = default;
    pvdSeries(std::string path, int nx, int ny, double W, double H, std::string name = "u", bool compressed = true)
        : path_(std::move(path)), name_(std::move(name)), nx_(nx), ny_(ny), W_(W), H_(H), compressed_(compressed)
    {
        const size_t dot = path_.rfind(".pvd");
        stem_ = dot == std::string::npos ? path_ : path_.substr(0, dot);
        path_ = stem_ + ".pvd";
    }
    bool open() const { return !path_.empty(); }
    const std::string& path() const { return path_; }

    // ny = 0: a 1-D channel field (write_vtu_line); uy non-null: a vector field
    std::string write(const double* u, double t, const double* uy = nullptr)
    {
        char num[16];
        std::snprintf(num, sizeof(num), "%06d", (int)steps_.size());
        const std::string vtu = stem_ + num + ".vtu";
        if (ny_ == 0)
            write_vtu_line(vtu, nx_, W_, u, name_, compressed_);
        else
            write_vtu(vtu, nx_, ny_, W_, H_, u, name_, compressed_, uy);
        const size_t slash = vtu.find_last_of('/');
        steps_.emplace_back(t, slash == std::string::npos ? vtu : vtu.substr(slash + 1));
        std::ofstream f(path_, std::ios::trunc);
        if (!f) throw std::runtime_error("pvdSeries: cannot open " + path_);
        f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n<Collection>\n";
        char ts[40];
        for (const auto& e : steps_) {
            std::snprintf(ts, sizeof(ts), "%.17g", e.first);
            f << "<DataSet timestep=\"" << ts << "\" part=\"0\" file=\"" << e.second << "\" />\n";
        }
        f << "</Collection>\n</VTKFile>\n";
        return vtu;
    }

private:
    std::string path_, stem_, name_;
    int nx_ = 0, ny_ = 0;
    double W_ = 0.0, H_ = 0.0;
    bool compressed_ = true;
    std::vector<std::pair<double, std::string>> steps_;
};

// Finite-difference nodes (nx + 1) x (ny + 1) at spacing h as a VTK XML
// rectilinear grid, raw binary; the file name follows RecordData's
// "%s/%s%04d.vtr" (diffuclass.cpp:569-570).
inline std::string write_vtr(const std::string& dir, const std::string& name, int step, int nx, int ny, double h,
                             const double* u)
{
    char fn[64];
    std::snprintf(fn, sizeof(fn), "%04d.vtr", step);
    const std::string path = dir + "/" + name + fn;
    std::vector<double> xs(nx + 1), ys(ny + 1), zs(1, 0.0);
    for (int j = 0; j <= nx; ++j) xs[j] = j * h;
    for (int i = 0; i <= ny; ++i) ys[i] = i * h;
    std::ofstream f(path, std::ios::trunc);
    if (!f) throw std::runtime_error("write_vtr: cannot open " + path);
    f << "<?xml version=\"1.0\"?>\n<VTKFile type=\"RectilinearGrid\"" << vtk_header(false) << ">\n"
      << "<RectilinearGrid WholeExtent=\"0 " << nx << " 0 " << ny << " 0 0\">\n<Piece Extent=\"0 " << nx << " 0 "
      << ny << " 0 0\">\n<Coordinates>\n";
    put_array(f, "Float64", "x", 1, xs.data(), xs.size() * 8, false);
    put_array(f, "Float64", "y", 1, ys.data(), ys.size() * 8, false);
    put_array(f, "Float64", "z", 1, zs.data(), 8, false);
    f << "</Coordinates>\n<PointData Scalars=\"" << name << "\">\n";
    put_array(f, "Float64", name.c_str(), 1, u, (size_t)(nx + 1) * (ny + 1) * 8, false);
    f << "</PointData>\n</Piece>\n</RectilinearGrid>\n</VTKFile>\n";
    if (!f) throw std::runtime_error("write_vtr: write failed for " + path);
    return path;
}

// Shortest text that reads back to the same double.
inline std::string json_number(double v)
{
    if (v != v || v - v != 0.0) return "null"; // NaN and infinities, as nlohmann::json writes them
    char buf[40];
    for (int prec = 15; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    // integral values keep a ".0" (as nlohmann::json writes them), so -0.0
    // and large integers read back as doubles
    std::string s = buf;
    if (s.find_first_of(".e") == std::string::npos) s += ".0";
    return s;
}

inline std::string json_string(const std::string& s)
{
    std::string o = "\"";
    for (const unsigned char ch : s) {
        if (ch == '"' || ch == '\\') {
            o += '\\';
            o += (char)ch;
        } else if (ch < 0x20) {
            char b[8];
            std::snprintf(b, sizeof(b), "\\u%04x", ch);
            o += b;
        } else {
            o += (char)ch;
        }
    }
    return o + "\"";
}

// One cell of a frame: {[id, strainType], [x, y, angle, length, spring, hsl...]}
// (src/main.cpp:114-146).
struct frameCell {
    uint64_t id = 0;
    uint64_t strainType = 0;
    std::vector<double> data;
};

// The controller's JSON log, "jsonSimVersion" v00_02 (src/main.cpp:88-103).
// The file is "<compileTime>-<nodeID>_<simNumber>-jsonData.json"
// (src/main.cpp:170-174); parameters is the experiment's JSON text.
class jsonRecording {
public:
    jsonRecording(int nodeID, int simNumber) : nodeID_(nodeID), simNumber_(simNumber) {}

    std::string init(const std::string& dir, const std::string& compileTime, long timeSinceEpoch,
                     std::string parametersJson = "{}")
    {
        time_ = timeSinceEpoch;
        params_ = std::move(parametersJson);
        path_ = (dir.empty() ? std::string() : dir + "/") + compileTime + "-" + std::to_string(nodeID_) + "_" +
                std::to_string(simNumber_) + "-jsonData.json";
        f_.open(path_, std::ios::trunc);
        if (!f_) throw std::runtime_error("jsonRecording: cannot open " + path_);
        f_ << "{\"cellDataFormat\":[[\"id\",\"strainType\"],[\"x\",\"y\",\"angle\",\"length\",\"spring\",\"vargs\"]],"
              "\"frames\":[";
        frames_ = 0;
        return path_;
    }

    // externalHSL: one natural-order vector per species (empty: no signalling);
    // divisions: the (time, id, length, id, length) records of the frame
    // (src/abm/eQabm.h:281, src/abm/eQabm.cpp:462-476)
    void recordFrame(double simTime, const std::vector<frameCell>& cells,
                     const std::vector<std::vector<double>>& externalHSL,
                     const std::vector<std::vector<double>>& divisions = {})
    {
        if (!f_.is_open()) throw std::runtime_error("jsonRecording: recordFrame before init");
        std::string s;
        s += frames_++ ? ",{" : "{";
        s += "\"cells\":[";
        for (size_t c = 0; c < cells.size(); ++c) {
            if (c) s += ',';
            s += "[[" + std::to_string(cells[c].id) + "," + std::to_string(cells[c].strainType) + "],[";
            for (size_t k = 0; k < cells[c].data.size(); ++k) {
                if (k) s += ',';
                s += json_number(cells[c].data[k]);
            }
            s += "]]";
        }
        s += "],\"divisions\":";
        s += array2(divisions);
        s += ",\"externalHSL\":";
        f_ << s;
        write_fields(externalHSL);
        f_ << ",\"simTime\":" << json_number(simTime) << "}";
        f_.flush();
    }

    void finalize()
    {
        if (!f_.is_open()) return;
        f_ << "],\"jsonSimVersion\":\"v00_02\",\"nodeID\":" << nodeID_ << ",\"parameters\":" << params_
           << ",\"simNumber\":" << simNumber_ << ",\"timeSinceEpoch\":" << time_ << "}\n";
        f_.close();
    }
    ~jsonRecording() { finalize(); }
    const std::string& path() const { return path_; }

private:
    static std::string array2(const std::vector<std::vector<double>>& a)
    {
        std::string s = "[";
        for (size_t r = 0; r < a.size(); ++r) {
            if (r) s += ',';
            s += '[';
            for (size_t k = 0; k < a[r].size(); ++k) {
                if (k) s += ',';
                s += json_number(a[r][k]);
            }
            s += ']';
        }
        return s + "]";
    }
    void write_fields(const std::vector<std::vector<double>>& a)
    {
        f_ << '[';
        std::string chunk;
        for (size_t r = 0; r < a.size(); ++r) {
            f_ << (r ? ",[" : "[");
            chunk.clear();
            for (size_t k = 0; k < a[r].size(); ++k) {
                if (k) chunk += ',';
                chunk += json_number(a[r][k]);
                if (chunk.size() > (1u << 20)) {
                    f_ << chunk;
                    chunk.clear();
                }
            }
            f_ << chunk << ']';
        }
        f_ << ']';
    }

    int nodeID_, simNumber_;
    long time_ = 0;
    std::string params_, path_;
    std::ofstream f_;
    size_t frames_ = 0;
};

} // namespace eqio
