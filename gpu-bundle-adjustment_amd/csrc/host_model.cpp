// host_model.cpp -- host camera model, BAL text I/O and error state.
//
// snavely_residual_host follows the reader's model (reference
// src/bal_reader.cpp:100-128: zero-order branch for theta <= 1e-12);
// total_squared_error sums in observation order (:130-139).
//
// BAL text format (Agarwal et al., "Bundle Adjustment in the Large"): three
// counts "C P O", then O records "camera point x y", then 9 C camera values,
// then 3 P point coordinates, all separated by white space.  The loader keeps
// the reference's contract (bal_problem.hpp:32-36): .bz2 input is read
// through bzcat, and any malformed number, a non-positive count, an index out
// of range, a missing value or extra data after the points is an error.
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gba_c.h"
#include "gba_internal.hpp"
#include "snavely_model.hpp"

namespace gba_amd {

namespace {
thread_local std::string g_last_error;

// Pulls white-space separated tokens out of a BAL stream through a fixed
// buffer, so a file is never held in memory as text.
class BalStream {
public:
    explicit BalStream(const std::string& path) : path_(path) {
        piped_ = path.size() > 4 && path.compare(path.size() - 4, 4, ".bz2") == 0;
        if (piped_) {
            std::string quoted = "'";
            for (char ch : path) {
                if (ch == '\'') quoted += "'\\''";
                else quoted += ch;
            }
            quoted += "'";
            fp_ = popen(("bzcat -- " + quoted).c_str(), "r");
        } else {
            fp_ = std::fopen(path.c_str(), "rb");
        }
        if (!fp_) throw Error("BAL: cannot open '" + path + "'");
    }
    ~BalStream() {
        if (fp_) (void)(piped_ ? pclose(fp_) : std::fclose(fp_));
    }
    BalStream(const BalStream&) = delete;
    BalStream& operator=(const BalStream&) = delete;

    // Next token into tok_; false at the end of the stream.
    bool token() {
        tok_.clear();
        int ch;
        do {
            ch = get();
        } while (ch != EOF && std::isspace(ch));
        while (ch != EOF && !std::isspace(ch)) {
            tok_.push_back(static_cast<char>(ch));
            ch = get();
        }
        if (!tok_.empty()) ++count_;
        return !tok_.empty();
    }
    double real(const char* what) {
        need(what);
        if (!numeric_text(tok_)) bad(what);
        char* end = nullptr;
        const double v = std::strtod(tok_.c_str(), &end);
        if (end != tok_.c_str() + tok_.size()) bad(what);
        return v;
    }
    int integer(const char* what) {
        need(what);
        bool digits = !tok_.empty();
        for (size_t i = (tok_[0] == '-') ? 1 : 0; i < tok_.size(); ++i) digits = digits && std::isdigit(
            static_cast<unsigned char>(tok_[i]));
        if (!digits || tok_ == "-") bad(what);
        errno = 0;
        char* end = nullptr;
        const long v = std::strtol(tok_.c_str(), &end, 10);
        if (errno != 0 || v < INT32_MIN || v > INT32_MAX) bad(what);
        return static_cast<int>(v);
    }
    // Ends the read: extra tokens are an error, and so is a failed bzcat.
    void finish() {
        if (token()) throw Error("BAL: '" + path_ + "' has data after the last point");
        const int rc = piped_ ? pclose(fp_) : std::fclose(fp_);
        fp_ = nullptr;
        if (rc != 0) throw Error("BAL: reading '" + path_ + "' failed");
    }
    const std::string& path() const { return path_; }

private:
    int get() {
        if (pos_ == len_) {
            len_ = std::fread(buf_, 1, sizeof buf_, fp_);
            pos_ = 0;
            if (len_ == 0) return EOF;
        }
        return static_cast<unsigned char>(buf_[pos_++]);
    }
    void need(const char* what) {
        if (!token()) throw Error("BAL: '" + path_ + "' ends before its " + what);
    }
    [[noreturn]] void bad(const char* what) {
        throw Error("BAL: '" + path_ + "': token " + std::to_string(count_) + " ('" + tok_ +
                    "') is not a valid " + what);
    }
    // Decimal notation only ([+-] digits [. digits] [e [+-] digits]) or a
    // signed nan / inf / infinity; hexadecimal floats are refused.
    static bool numeric_text(const std::string& t) {
        size_t i = (t[0] == '+' || t[0] == '-') ? 1 : 0;
        std::string rest;
        for (size_t k = i; k < t.size(); ++k) rest.push_back(static_cast<char>(std::tolower(
            static_cast<unsigned char>(t[k]))));
        if (rest == "nan" || rest == "inf" || rest == "infinity") return true;
        bool mant = false;
        while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) i++, mant = true;
        if (i < t.size() && t[i] == '.') {
            ++i;
            while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) i++, mant = true;
        }
        if (!mant) return false;
        if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
            ++i;
            if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
            bool ex = false;
            while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) i++, ex = true;
            if (!ex) return false;
        }
        return i == t.size();
    }

    std::string path_;
    FILE* fp_ = nullptr;
    bool piped_ = false;
    char buf_[1 << 16];
    size_t pos_ = 0, len_ = 0;
    std::string tok_;
    long long count_ = 0;
};

struct BalCounts {
    int C, P, O;
};

BalCounts read_counts(BalStream& in) {
    BalCounts n{in.integer("camera count"), in.integer("point count"), in.integer("observation count")};
    if (n.C <= 0 || n.P <= 0 || n.O <= 0)
        throw Error("BAL: '" + in.path() + "' must start with three positive counts");
    return n;
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
    BalStream in(path);
    const BalCounts n = read_counts(in);
    *C = n.C;
    *P = n.P;
    *O = n.O;
    return 0;  // the rest of the stream is not read (a bzcat pipe just closes)
}

int load_bal_into(const char* path, gba_problem_view* out) {
    BalStream in(path);
    const BalCounts n = read_counts(in);
    if (out->num_cameras != n.C || out->num_points != n.P || out->num_observations != n.O)
        return set_error("gba_load_bal: view sizes do not match the file header");
    // parsed into local arrays first: the caller's view changes only when the
    // whole file is valid
    const size_t O = static_cast<size_t>(n.O);
    std::vector<int> oc(O), op(O);
    std::vector<double> px(2 * O), cams(9 * static_cast<size_t>(n.C)), pts(3 * static_cast<size_t>(n.P));
    for (size_t e = 0; e < O; ++e) {
        oc[e] = in.integer("camera index");
        op[e] = in.integer("point index");
        px[2 * e] = in.real("pixel coordinate");
        px[2 * e + 1] = in.real("pixel coordinate");
        if (oc[e] < 0 || oc[e] >= n.C || op[e] < 0 || op[e] >= n.P)
            throw Error("BAL: '" + in.path() + "': observation " + std::to_string(e) +
                        " names a camera or point that does not exist");
    }
    for (double& v : cams) v = in.real("camera parameter");
    for (double& v : pts) v = in.real("point coordinate");
    in.finish();
    std::memcpy(out->cameras, cams.data(), sizeof(double) * cams.size());
    std::memcpy(out->points, pts.data(), sizeof(double) * pts.size());
    std::memcpy(out->obs_camera, oc.data(), sizeof(int) * O);
    std::memcpy(out->obs_point, op.data(), sizeof(int) * O);
    std::memcpy(out->obs_pixel, px.data(), sizeof(double) * px.size());
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
