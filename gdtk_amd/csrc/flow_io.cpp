// flow_io.cpp -- the reference's legacy flow-solution files of a structured block, on the
// host: the header, the cell records as text (gzip-compressed) or raw binary, and the
// reader.  The device side (packing the flow-state planes into cell records and back)
// lives in flow_io.hip; sim.cpp composes the two for gdtk_hip_block_write_flow/read_flow.
//
//   variable list      src/eilmer/fluidblockio_old.d:107-151 (build_flow_variable_list)
//   writer             src/eilmer/fluidblockio_old.d:388-447 (write_solution, SFluidBlock)
//   cell record        src/eilmer/fluidblockio_old.d:1487-1691 (cell_data_as_string,
//                      cell_data_to_raw_binary)
//   reader             src/eilmer/fluidblockio_old.d:156-273 (read_legacy_solution, SFluidBlock)
//
// The configuration of this path fixes the list: no MHD, no quality, no radiation, a laminar
// (no turbulence variables) single-temperature gas without local time stepping, so a cell
// record is pos.x pos.y pos.z volume rho vel.x vel.y vel.z p a mu k mu_t k_t S, the mass
// fractions (one, 1.0, for a single-species gas), dt_chem for a multi-species gas, u, T.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "flow_io.hpp"

namespace gdtk {

std::vector<std::string> flow_variable_list(int nsp, const std::vector<std::string>& species) {
    if ((int)species.size() != nsp) throw std::runtime_error("flow variables: one species name per species");
    std::vector<std::string> v = {"pos.x", "pos.y", "pos.z", "volume", "rho", "vel.x", "vel.y", "vel.z",
                                  "p",     "a",     "mu",    "k",      "mu_t", "k_t",  "S"};
    // massfName (fluidblockio_old.d:98-102): white space inside a name becomes dashes
    for (int i = 0; i < nsp; ++i) {
        std::string name;
        bool in_space = false;
        for (char ch : species[i]) {
            if (ch == ' ' || ch == '\t') {
                if (!in_space) name += '-';
                in_space = true;
            } else {
                name += ch;
                in_space = false;
            }
        }
        v.push_back("massf[" + std::to_string(i) + "]-" + name);
    }
    if (nsp > 1) v.push_back("dt_chem");
    v.push_back("u");
    v.push_back("T");
    return v;
}

std::vector<std::string> split_names(const char* names) {
    std::vector<std::string> out;
    if (!names) return out;
    const char* p = names;
    while (*p) {
        while (*p == ' ' || *p == '\t' || *p == '\n') ++p;
        if (!*p) break;
        const char* q = p;
        while (*q && *q != ' ' && *q != '\t' && *q != '\n') ++q;
        out.emplace_back(p, q);
        p = q;
    }
    return out;
}

namespace {

struct GzFile {
    gzFile f = nullptr;
    GzFile(const std::string& fn, const char* mode) {
        f = gzopen(fn.c_str(), mode);
        if (!f) throw std::runtime_error("cannot open " + fn);
    }
    ~GzFile() {
        if (f) gzclose(f);
    }
    void write(const std::string& s) {
        if (s.empty()) return;
        if (gzwrite(f, s.data(), (unsigned)s.size()) != (int)s.size()) throw std::runtime_error("gzip write failed");
    }
    void close() {
        if (f && gzclose(f) != Z_OK) {
            f = nullptr;
            throw std::runtime_error("gzip close failed");
        }
        f = nullptr;
    }
};

struct RawFile {
    FILE* f = nullptr;
    RawFile(const std::string& fn, const char* mode) {
        f = std::fopen(fn.c_str(), mode);
        if (!f) throw std::runtime_error("cannot open " + fn);
    }
    ~RawFile() {
        if (f) std::fclose(f);
    }
    void write(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) throw std::runtime_error("write failed");
    }
    void read(void* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) throw std::runtime_error("flow file ends early");
    }
    void close() {
        if (f && std::fclose(f) != 0) {
            f = nullptr;
            throw std::runtime_error("close failed");
        }
        f = nullptr;
    }
};

// One gzip member written from text pieces compressed on several threads: each piece is
// an independent raw deflate stream primed with the last 32 KiB of the text before it and
// ended with a sync flush (the last one with the final block), so the pieces concatenate
// into one deflate stream; the CRC-32 of the whole text comes from the pieces' CRCs
// (crc32_combine).  Any gzip reader, zlib's included, reads it as the single-stream file
// the reference writes.
class ParallelGzip {
public:
    explicit ParallelGzip(RawFile& f) : f_(f) {
        const unsigned char header[10] = {0x1f, 0x8b, 8, 0, 0, 0, 0, 0, 0, 3};
        f_.write(header, sizeof header);
    }
    // Compresses the pieces in parallel and writes them in order; `last` ends the stream.
    void add_parallel(const std::vector<std::string>& pieces, bool last) {
        const size_t n = pieces.size();
        std::vector<std::string> out(n);
        std::vector<unsigned long> crc(n);
        std::vector<std::string> dict(n);
        for (size_t t = 0; t < n; ++t) {
            dict[t] = tail_;
            if (!pieces[t].empty()) tail_ = tail_of(tail_, pieces[t]);
        }
        std::vector<std::thread> th;
        for (size_t t = 0; t < n; ++t)
            th.emplace_back([&, t] {
                out[t] = deflate_piece(pieces[t], dict[t], last && t + 1 == n);
                crc[t] = crc32(0L, reinterpret_cast<const Bytef*>(pieces[t].data()), (uInt)pieces[t].size());
            });
        for (auto& x : th) x.join();
        for (size_t t = 0; t < n; ++t) {
            f_.write(out[t].data(), out[t].size());
            crc_ = crc32_combine(crc_, crc[t], (z_off_t)pieces[t].size());
            size_ += pieces[t].size();
        }
        finished_ = last;
    }
    // One piece that does not end the stream (the header).
    void add_text(const std::string& text) { add_parallel(std::vector<std::string>{text}, false); }
    void finish() {
        if (!finished_) add_parallel(std::vector<std::string>{std::string()}, true);
        unsigned char trailer[8];
        for (int i = 0; i < 4; ++i) trailer[i] = (unsigned char)((crc_ >> (8 * i)) & 0xff);
        for (int i = 0; i < 4; ++i) trailer[4 + i] = (unsigned char)((size_ >> (8 * i)) & 0xff);
        f_.write(trailer, sizeof trailer);
    }

private:
    RawFile& f_;
    std::string tail_;
    unsigned long crc_ = crc32(0L, Z_NULL, 0);
    uint64_t size_ = 0;
    bool finished_ = false;

    static std::string tail_of(const std::string& before, const std::string& text) {
        const size_t W = 32768;
        if (text.size() >= W) return text.substr(text.size() - W);
        std::string t = before + text;
        return t.size() > W ? t.substr(t.size() - W) : t;
    }
    static std::string deflate_piece(const std::string& in, const std::string& dict, bool last) {
        z_stream z{};
        if (deflateInit2(&z, Z_DEFAULT_COMPRESSION, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK)
            throw std::runtime_error("deflateInit2 failed");
        if (!dict.empty())
            deflateSetDictionary(&z, reinterpret_cast<const Bytef*>(dict.data()), (uInt)dict.size());
        std::string out;
        out.resize(deflateBound(&z, (uLong)in.size()) + 64);
        z.next_in = reinterpret_cast<Bytef*>(const_cast<char*>(in.data()));
        z.avail_in = (uInt)in.size();
        z.next_out = reinterpret_cast<Bytef*>(&out[0]);
        z.avail_out = (uInt)out.size();
        const int rc = deflate(&z, last ? Z_FINISH : Z_SYNC_FLUSH);
        const bool ok = last ? (rc == Z_STREAM_END) : (rc == Z_OK && z.avail_in == 0);
        out.resize(out.size() - z.avail_out);
        deflateEnd(&z);
        if (!ok) throw std::runtime_error("deflate failed");
        return out;
    }
};

int hw_threads() {
    const unsigned n = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(n, 16u));
}

// " v0 v1 ... vN\n" per cell, each value as %.18e (fluidblockio_old.d:434-440, 1534-1573).
void format_records(const double* rec, int64_t c0, int64_t c1, int nv, std::string& out) {
    char buf[40];
    out.clear();
    out.reserve((size_t)(c1 - c0) * (size_t)nv * 26 + 16);
    for (int64_t c = c0; c < c1; ++c) {
        const double* r = rec + c * nv;
        for (int v = 0; v < nv; ++v) {
            const int n = std::snprintf(buf, sizeof buf, " %.18e", r[v]);
            out.append(buf, (size_t)n);
        }
        out.push_back('\n');
    }
}

}  // namespace

void write_flow_file(const std::string& fn, int format, const std::string& label, double sim_time, int dims, int nic,
                     int njc, int nkc, const std::vector<std::string>& vars, const double* rec, int64_t ncell) {
    const int nv = (int)vars.size();
    if (format == FLOW_RAWBINARY) {
        RawFile f(fn, "wb");
        const char header[] = "structured_grid_flow 1.0";
        f.write(header, std::strlen(header));
        const int32_t lab = (int32_t)label.size();
        f.write(&lab, 4);
        f.write(label.data(), label.size());
        f.write(&sim_time, 8);
        const int32_t n = nv;
        f.write(&n, 4);
        for (const auto& s : vars) {
            const int32_t l = (int32_t)s.size();
            f.write(&l, 4);
            f.write(s.data(), s.size());
        }
        const int32_t d4[4] = {dims, nic, njc, nkc};
        f.write(d4, 16);
        f.write(rec, (size_t)ncell * nv * sizeof(double));
        f.close();
        return;
    }
    if (format != FLOW_GZIPTEXT) throw std::runtime_error("unknown flow format");
    char buf[128];
    std::string h = "structured_grid_flow 1.0\n";
    h += "label: " + label + "\n";
    std::snprintf(buf, sizeof buf, "sim_time: %.18e\n", sim_time);
    h += buf;
    h += "variables: " + std::to_string(nv) + "\n";
    for (const auto& s : vars) h += " \"" + s + "\"";
    h += "\n";
    h += "dimensions: " + std::to_string(dims) + "\n";
    h += "nicell: " + std::to_string(nic) + "\n";
    h += "njcell: " + std::to_string(njc) + "\n";
    h += "nkcell: " + std::to_string(nkc) + "\n";
    RawFile f(fn, "wb");
    ParallelGzip gz(f);
    gz.add_text(h);
    // Format and compress chunks of cells on the host threads, write them in order.
    const int nt = hw_threads();
    const int64_t chunk = 32768;
    std::vector<std::string> parts(nt);
    for (int64_t base = 0; base < ncell; base += chunk * nt) {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) {
            const int64_t c0 = std::min(ncell, base + t * chunk), c1 = std::min(ncell, c0 + chunk);
            th.emplace_back([&, t, c0, c1] { format_records(rec, c0, c1, nv, parts[t]); });
        }
        for (auto& x : th) x.join();
        gz.add_parallel(parts, base + chunk * nt >= ncell);
    }
    gz.finish();
    f.close();
}

namespace {

// Reads one line (without its end-of-line) from a gzip or plain text file.
bool gz_line(gzFile f, std::string& line) {
    line.clear();
    char buf[4096];
    while (true) {
        if (!gzgets(f, buf, sizeof buf)) return !line.empty();
        line += buf;
        if (!line.empty() && line.back() == '\n') {
            line.pop_back();
            if (!line.empty() && line.back() == '\r') line.pop_back();
            return true;
        }
    }
}

std::string after_prefix(const std::string& line, const char* prefix) {
    const size_t n = std::strlen(prefix);
    if (line.compare(0, n, prefix) != 0) throw std::runtime_error(std::string("flow file: expected '") + prefix + "'");
    return line.substr(n);
}

// Column of each variable of `want` in the file's list; every one must be present
// (fluidblockio_old.d:222-232: "Could not find variable").
std::vector<int> columns(const std::vector<std::string>& have, const std::vector<std::string>& want) {
    std::vector<int> col(want.size(), -1);
    for (size_t i = 0; i < want.size(); ++i) {
        for (size_t j = 0; j < have.size(); ++j)
            if (have[j] == want[i]) {
                col[i] = (int)j;
                break;
            }
        if (col[i] < 0) throw std::runtime_error("Could not find variable: " + want[i]);
    }
    return col;
}

void check_size(int dims_read, int nic_read, int njc_read, int nkc_read, int dims, int nic, int njc, int nkc) {
    if (dims_read != dims) throw std::runtime_error("dimensions found: " + std::to_string(dims_read));
    if (nic_read != nic || njc_read != njc || nkc_read != ((dims == 3) ? nkc : 1))
        throw std::runtime_error("mismatch in solution size: have read nic=" + std::to_string(nic_read) +
                                 " njc=" + std::to_string(njc_read) + " nkc=" + std::to_string(nkc_read));
}

}  // namespace

double read_flow_file(const std::string& fn, int format, int dims, int nic, int njc, int nkc,
                      const std::vector<std::string>& vars, std::vector<double>& rec) {
    const int64_t ncell = (int64_t)nic * njc * nkc;
    const int nv = (int)vars.size();
    rec.assign((size_t)ncell * nv, 0.0);
    double sim_time = 0.0;
    if (format == FLOW_RAWBINARY) {
        RawFile f(fn, "rb");
        const char expected[] = "structured_grid_flow 1.0";
        char found[sizeof expected] = {0};
        f.read(found, std::strlen(expected));
        if (std::strcmp(found, expected) != 0) throw std::runtime_error("unexpected header in raw binary flow file");
        int32_t lab = 0;
        f.read(&lab, 4);
        if (lab < 0 || lab > (1 << 20)) throw std::runtime_error("flow file: bad label length");
        std::string label((size_t)lab, '\0');
        f.read(&label[0], (size_t)lab);
        f.read(&sim_time, 8);
        int32_t n = 0;
        f.read(&n, 4);
        if (n < 0 || n > 4096) throw std::runtime_error("flow file: bad variable count");
        std::vector<std::string> names((size_t)n);
        for (auto& s : names) {
            int32_t l = 0;
            f.read(&l, 4);
            if (l < 0 || l > 4096) throw std::runtime_error("flow file: bad variable name");
            s.assign((size_t)l, '\0');
            f.read(&s[0], (size_t)l);
        }
        int32_t d4[4];
        f.read(d4, 16);
        check_size(d4[0], d4[1], d4[2], d4[3], dims, nic, njc, nkc);
        // The raw binary records have the writer's fixed layout (read_values_from_raw_binary
        // takes them in that order), one double per listed variable.
        const std::vector<int> col = columns(names, vars);
        std::vector<double> row((size_t)n);
        for (int64_t c = 0; c < ncell; ++c) {
            f.read(row.data(), (size_t)n * 8);
            for (int v = 0; v < nv; ++v) rec[(size_t)c * nv + v] = row[(size_t)col[v]];
        }
        return sim_time;
    }
    if (format != FLOW_GZIPTEXT) throw std::runtime_error("unknown flow format");
    GzFile f(fn, "rb");
    std::string line;
    auto next = [&]() -> const std::string& {
        if (!gz_line(f.f, line)) throw std::runtime_error("flow file ends early");
        return line;
    };
    if (after_prefix(next(), "structured_grid_flow ") != "1.0") throw std::runtime_error("flow file: format version");
    after_prefix(next(), "label:");
    sim_time = std::strtod(after_prefix(next(), "sim_time:").c_str(), nullptr);
    const int n = std::atoi(after_prefix(next(), "variables:").c_str());
    std::vector<std::string> names;
    for (auto& s : split_names(next().c_str())) {
        std::string t = s;
        t.erase(std::remove(t.begin(), t.end(), '"'), t.end());
        names.push_back(t);
    }
    if ((int)names.size() != n) throw std::runtime_error("flow file: variable count does not match the list");
    const std::vector<int> col = columns(names, vars);
    const int dims_read = std::atoi(after_prefix(next(), "dimensions:").c_str());
    const int nic_read = std::atoi(after_prefix(next(), "nicell:").c_str());
    const int njc_read = std::atoi(after_prefix(next(), "njcell:").c_str());
    const int nkc_read = std::atoi(after_prefix(next(), "nkcell:").c_str());
    check_size(dims_read, nic_read, njc_read, nkc_read, dims, nic, njc, nkc);
    // The cell lines: decompressed in one piece, then parsed on the host threads.
    std::string body;
    {
        std::vector<char> buf(1 << 22);
        int got;
        while ((got = gzread(f.f, buf.data(), (unsigned)buf.size())) > 0) body.append(buf.data(), (size_t)got);
        if (got < 0) throw std::runtime_error("flow file: gzip read failed");
    }
    std::vector<size_t> line_start;
    line_start.reserve((size_t)ncell + 1);
    for (size_t pos = 0; pos < body.size() && (int64_t)line_start.size() < ncell;) {
        line_start.push_back(pos);
        const size_t e = body.find('\n', pos);
        pos = (e == std::string::npos) ? body.size() : e + 1;
    }
    if ((int64_t)line_start.size() < ncell) throw std::runtime_error("flow file ends early");
    const int nt = hw_threads();
    std::vector<std::string> errs((size_t)nt);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            std::vector<double> row((size_t)n);
            const int64_t c0 = ncell * t / nt, c1 = ncell * (t + 1) / nt;
            for (int64_t c = c0; c < c1 && errs[(size_t)t].empty(); ++c) {
                const char* p = body.c_str() + line_start[(size_t)c];
                for (int j = 0; j < n; ++j) {
                    while (*p == ' ' || *p == '\t') ++p;
                    char* e = nullptr;
                    if (*p != '\n' && *p != '\r' && *p != '\0') row[(size_t)j] = std::strtod(p, &e);
                    if (!e || e == p) {
                        errs[(size_t)t] = "flow file: cell " + std::to_string(c) + " has too few values";
                        break;
                    }
                    p = e;
                }
                for (int v = 0; v < nv; ++v) rec[(size_t)c * nv + v] = row[(size_t)col[v]];
            }
        });
    for (auto& x : th) x.join();
    for (const auto& e : errs)
        if (!e.empty()) throw std::runtime_error(e);
    return sim_time;
}

}  // namespace gdtk
