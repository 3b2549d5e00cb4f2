// host_model.cpp -- host camera model, BAL text I/O and error state.
//
// snavely_residual_host follows the reader's model (reference
// src/bal_reader.cpp:100-128: zero-order branch for theta <= 1e-12);
// total_squared_error sums in observation order (:130-139).  The BAL reader
// has the reference's contract (:12-98): whole-file read, .bz2 streamed
// through bzcat, strict number parsing, positive header, index range checks,
// no trailing garbage; any violation is an error.
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gba_c.h"
#include "gba_internal.hpp"
#include "snavely_model.hpp"

namespace gba_amd {

namespace {
thread_local std::string g_last_error;

bool ends_with_bz2(const std::string& p) {
    return p.size() >= 4 && p.compare(p.size() - 4, 4, ".bz2") == 0;
}

std::string read_file(const std::string& path) {
    const bool bz = ends_with_bz2(path);
    FILE* fp = nullptr;
    if (bz) {
        std::string cmd = "bzcat -- \"" + path + "\"";
        fp = popen(cmd.c_str(), "r");
    } else {
        fp = std::fopen(path.c_str(), "rb");
    }
    if (!fp) throw Error("cannot open " + path);
    std::string buf;
    std::vector<char> chunk(1 << 20);
    size_t n;
    while ((n = std::fread(chunk.data(), 1, chunk.size(), fp)) > 0) buf.append(chunk.data(), n);
    const int rc = bz ? pclose(fp) : std::fclose(fp);
    if (rc != 0) throw Error("read failed for " + path);
    return buf;
}

class Tokens {
public:
    explicit Tokens(const std::string& s) : p_(s.data()), end_(s.data() + s.size()) {}
    template <typename T>
    T next() {
        skip();
        T v{};
        auto r = std::from_chars(p_, end_, v);
        if (r.ec != std::errc()) throw Error("malformed BAL number");
        p_ = r.ptr;
        return v;
    }
    bool done() {
        skip();
        return p_ == end_;
    }

private:
    void skip() {
        while (p_ != end_ && (*p_ == ' ' || *p_ == '\n' || *p_ == '\t' || *p_ == '\r')) ++p_;
    }
    const char* p_;
    const char* end_;
};

struct Parsed {
    int C = 0, P = 0, O = 0;
    std::vector<double> cams, pts, px;
    std::vector<int> oc, op;
};

Parsed parse_bal(const std::string& path, bool header_only) {
    const std::string text = read_file(path);
    Tokens t(text);
    Parsed r;
    r.C = t.next<int>();
    r.P = t.next<int>();
    r.O = t.next<int>();
    if (r.C <= 0 || r.P <= 0 || r.O <= 0) throw Error("bad BAL header in " + path);
    if (header_only) return r;
    r.oc.resize(r.O);
    r.op.resize(r.O);
    r.px.resize(2 * static_cast<size_t>(r.O));
    for (int i = 0; i < r.O; ++i) {
        r.oc[i] = t.next<int>();
        r.op[i] = t.next<int>();
        r.px[2 * static_cast<size_t>(i)] = t.next<double>();
        r.px[2 * static_cast<size_t>(i) + 1] = t.next<double>();
        if (r.oc[i] < 0 || r.oc[i] >= r.C || r.op[i] < 0 || r.op[i] >= r.P)
            throw Error("observation index out of range in " + path);
    }
    r.cams.resize(9 * static_cast<size_t>(r.C));
    for (auto& v : r.cams) v = t.next<double>();
    r.pts.resize(3 * static_cast<size_t>(r.P));
    for (auto& v : r.pts) v = t.next<double>();
    if (!t.done()) throw Error("trailing garbage in " + path);
    return r;
}

}  // namespace

int set_error(const std::string& msg) {
    g_last_error = msg;
    return 1;
}

const char* last_error() { return g_last_error.c_str(); }

void snavely_residual_host(const double* cam, const double* X, const double* obs, double* r) {
    const double th = std::sqrt(cam[0] * cam[0] + cam[1] * cam[1] + cam[2] * cam[2]);
    double P[3];
    if (th > 1e-12) {
        // R X = cos(th) X + sin(th) (w x X) + (1 - cos(th)) (w . X) w, w = a / th
        const double w[3] = {cam[0] / th, cam[1] / th, cam[2] / th};
        const double cs = std::cos(th), sn = std::sin(th);
        const double wx[3] = {w[1] * X[2] - w[2] * X[1], w[2] * X[0] - w[0] * X[2],
                              w[0] * X[1] - w[1] * X[0]};
        const double q = w[0] * X[0] + w[1] * X[1] + w[2] * X[2];
        for (int i = 0; i < 3; ++i) P[i] = X[i] * cs + wx[i] * sn + w[i] * q * (1.0 - cs) + cam[3 + i];
    } else {
        for (int i = 0; i < 3; ++i) P[i] = X[i] + cam[3 + i];
    }
    const double u = -P[0] / P[2], v = -P[1] / P[2];
    const double s = u * u + v * v;
    const double d = 1.0 + cam[7] * s + cam[8] * s * s;
    r[0] = cam[6] * d * u - obs[0];
    r[1] = cam[6] * d * v - obs[1];
}

int bal_header(const char* path, int* C, int* P, int* O) {
    Parsed r = parse_bal(path, true);
    *C = r.C;
    *P = r.P;
    *O = r.O;
    return 0;
}

int load_bal_into(const char* path, gba_problem_view* out) {
    Parsed r = parse_bal(path, false);
    if (out->num_cameras != r.C || out->num_points != r.P || out->num_observations != r.O)
        return set_error("gba_load_bal: view sizes do not match the file header");
    std::memcpy(out->cameras, r.cams.data(), sizeof(double) * r.cams.size());
    std::memcpy(out->points, r.pts.data(), sizeof(double) * r.pts.size());
    std::memcpy(out->obs_camera, r.oc.data(), sizeof(int) * r.oc.size());
    std::memcpy(out->obs_point, r.op.data(), sizeof(int) * r.op.size());
    std::memcpy(out->obs_pixel, r.px.data(), sizeof(double) * r.px.size());
    return 0;
}

int write_bal(const char* path, const gba_problem_view* p) {
    FILE* fp = std::fopen(path, "w");
    if (!fp) return set_error(std::string("cannot write ") + path);
    std::fprintf(fp, "%d %d %d\n", p->num_cameras, p->num_points, p->num_observations);
    for (int i = 0; i < p->num_observations; ++i)
        std::fprintf(fp, "%d %d %.17g %.17g\n", p->obs_camera[i], p->obs_point[i],
                     p->obs_pixel[2 * static_cast<size_t>(i)], p->obs_pixel[2 * static_cast<size_t>(i) + 1]);
    for (size_t i = 0; i < 9 * static_cast<size_t>(p->num_cameras); ++i)
        std::fprintf(fp, "%.17g\n", p->cameras[i]);
    for (size_t i = 0; i < 3 * static_cast<size_t>(p->num_points); ++i)
        std::fprintf(fp, "%.17g\n", p->points[i]);
    if (std::fclose(fp) != 0) return set_error(std::string("write failed for ") + path);
    return 0;
}

}  // namespace gba_amd
