// Amber restart reader and writer (see include/md/IO/Restart.h).
#include "IO/Restart.h"

#include <algorithm>
#include <cctype>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace md::io {

namespace {

bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// The numbers of one line: 12-column fields when each nonblank field parses
// and together they hold at least as many numbers as the blank-separated
// runs, else the runs.
void line_values(const std::string& line, std::vector<double>& out) {
    const char* s = line.c_str();
    const size_t n = line.size(), start = out.size();
    bool fixed_ok = true;
    size_t nfixed = 0;
    char buf[13];
    for (size_t i = 0; i < n; i += 12) {
        const size_t len = std::min<size_t>(12, n - i);
        std::memcpy(buf, s + i, len);
        buf[len] = '\0';
        char* p = buf;
        while (*p && is_blank(*p)) ++p;
        if (!*p) continue;
        char* end = nullptr;
        const double v = std::strtod(p, &end);
        while (end && *end && is_blank(*end)) ++end;
        if (end == p || (end && *end)) {
            fixed_ok = false;
            break;
        }
        out.push_back(v);
        ++nfixed;
    }
    size_t nruns = 0;
    for (size_t i = 0; i < n;) {
        if (is_blank(s[i])) {
            ++i;
            continue;
        }
        ++nruns;
        while (i < n && !is_blank(s[i])) ++i;
    }
    if (fixed_ok && nfixed >= nruns) return;
    out.resize(start);
    // blank-separated runs, each of which must be one number
    for (size_t i = 0; i < n;) {
        if (is_blank(s[i])) {
            ++i;
            continue;
        }
        size_t j = i;
        while (j < n && !is_blank(s[j])) ++j;
        const std::string t(s + i, j - i);
        char* end = nullptr;
        const double v = std::strtod(t.c_str(), &end);
        if (end != t.c_str() + t.size()) throw std::runtime_error("rst7: unreadable numbers in line: " + line);
        out.push_back(v);
        i = j;
    }
}

}  // namespace

Restart read_rst7(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("rst7: cannot open " + path);
    Restart r;
    std::string line;
    if (!std::getline(in, r.title)) throw std::runtime_error("rst7: empty file " + path);
    while (!r.title.empty() && (r.title.back() == '\r' || r.title.back() == ' ')) r.title.pop_back();
    if (!std::getline(in, line)) throw std::runtime_error("rst7: no atom-count line in " + path);
    {
        std::istringstream ss(line);
        long n = -1;
        if (!(ss >> n) || n <= 0) throw std::runtime_error("rst7: bad atom count in " + path);
        r.n_atoms = static_cast<int>(n);
        double t;
        if (ss >> t) {
            r.has_time = true;
            r.time = t;
        }
    }
    std::vector<double> v;
    v.reserve(6 * static_cast<size_t>(r.n_atoms) + 6);
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        line_values(line, v);
    }
    const size_t n3 = 3 * static_cast<size_t>(r.n_atoms), m = v.size();
    size_t nbox = 0;
    bool vel = false;
    if (m == n3 || m == n3 + 3 || m == n3 + 6) {
        nbox = m - n3;
    } else if (m == 2 * n3 || m == 2 * n3 + 3 || m == 2 * n3 + 6) {
        vel = true;
        nbox = m - 2 * n3;
    } else {
        throw std::runtime_error("rst7: " + std::to_string(m) + " values fit neither 3N nor 6N (+3 or +6) for N = " +
                                 std::to_string(r.n_atoms));
    }
    r.coords.assign(v.begin(), v.begin() + n3);
    if (vel) {
        r.velocities.assign(v.begin() + n3, v.begin() + 2 * n3);
        for (double& x : r.velocities) x *= kRestartVelocityScale;
    }
    if (nbox) {
        r.has_box = true;
        const size_t b0 = m - nbox;
        for (size_t k = 0; k < nbox; ++k) r.box[k] = v[b0 + k];
    }
    return r;
}

void write_rst7(const std::string& path, const Restart& r) {
    const size_t n3 = 3 * static_cast<size_t>(r.n_atoms);
    if (r.n_atoms <= 0 || r.coords.size() != n3 || (!r.velocities.empty() && r.velocities.size() != n3))
        throw std::runtime_error("rst7: coordinates / velocities do not match the atom count");
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("rst7: cannot write " + path);
    auto put6 = [&](const double* p, size_t n, double scale) {
        for (size_t k = 0; k < n; ++k) {
            const double x = p[k] * scale;
            if (!(std::fabs(x) < 9999.99999995)) {
                std::fclose(f);
                throw std::runtime_error("rst7: a value does not fit F12.7");
            }
            std::fprintf(f, "%12.7f", x);
            if (k % 6 == 5 || k + 1 == n) std::fputc('\n', f);
        }
    };
    std::fprintf(f, "%-80.80s\n", r.title.c_str());
    if (r.has_time) std::fprintf(f, "%6d%15.7E\n", r.n_atoms, r.time);
    else std::fprintf(f, "%6d\n", r.n_atoms);
    put6(r.coords.data(), n3, 1.0);
    if (!r.velocities.empty()) put6(r.velocities.data(), n3, 1.0 / kRestartVelocityScale);
    if (r.has_box) put6(r.box.data(), 6, 1.0);
    if (std::fclose(f) != 0) throw std::runtime_error("rst7: write failed for " + path);
}

}  // namespace md::io
