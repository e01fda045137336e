// On-disk formats of the reference's iteration loop (host):
//   restart files restart_file_{1,2}.neu written in turn, one line per node
//     "u\tv\tw\trho\tp\tt\te\tmu\titer" (Tecplot_output_files.cpp:115-127), read back by picking the
//     newer complete file (presolver.cpp:86-181, Main.cpp:5688-5740);
//   Tecplot ASCII point files "nodefile_<iter>.dat" with the reference's variable list and a
//     FEBRICK zone (Tecplot_output_files.cpp:92-113).
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

std::string slot_path(const char *dir, int slot)
{
    std::string p = dir && *dir ? std::string(dir) : std::string(".");
    if (p.back() != '/') p += '/';
    return p + "restart_file_" + std::to_string(slot) + ".neu";
}

// Reads a restart file: n lines of 9 columns with one iteration number. Returns the line count
// read with that iteration (n when complete), or -1 when the file cannot be opened.
int64_t read_restart(const std::string &path, int64_t n, cfdx_flow *flow, int64_t *iter)
{
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return -1;
    char line[1024];
    int64_t count = 0, it0 = -1;
    while (count < n && std::fgets(line, sizeof line, f)) {
        double v[8];
        long long it = 0;
        char *p = line, *q = nullptr;
        bool ok = true;
        for (int c = 0; c < 8 && ok; ++c) {
            v[c] = std::strtod(p, &q);
            ok = q != p;
            p = q;
        }
        if (ok) {
            it = std::strtoll(p, &q, 10);
            ok = q != p;
        }
        if (!ok) break;                 // a truncated or empty line ends the file
        if (it0 < 0) it0 = it;
        if (it != it0) break;           // lines of another iteration: the write was interrupted
        if (flow) {
            cfdx_flow &g = flow[count];
            g.u = v[0]; g.v = v[1]; g.w = v[2]; g.rho = v[3]; g.p = v[4]; g.t = v[5]; g.e = v[6]; g.mu = v[7];
            g.a = 0.0;
        }
        ++count;
    }
    std::fclose(f);
    if (iter) *iter = it0;
    return count;
}

}  // namespace

extern "C" int cfdx_write_restart(const char *dir, int32_t slot, int64_t n, const cfdx_flow *flow, int64_t iter,
                                  int32_t digits)
{
    if ((slot != 1 && slot != 2) || n < 0 || (!flow && n > 0) || digits < 1 || digits > 17)
        return cfdx::fail(CFDX_E_INVALID, "write_restart: slot must be 1 or 2, digits 1..17");
    const std::string path = slot_path(dir, slot);
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return cfdx::fail(CFDX_E_INVALID, "write_restart: cannot open %s: %s", path.c_str(), std::strerror(errno));
    for (int64_t i = 0; i < n; ++i) {
        const cfdx_flow &g = flow[i];
        std::fprintf(f, "%.*g\t%.*g\t%.*g\t%.*g\t%.*g\t%.*g\t%.*g\t%.*g\t%" PRId64 "\n", digits, g.u, digits, g.v, digits, g.w,
                     digits, g.rho, digits, g.p, digits, g.t, digits, g.e, digits, g.mu, iter);
    }
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return cfdx::fail(CFDX_E_INVALID, "write_restart: write error on %s", path.c_str());
    return CFDX_OK;
}

extern "C" int cfdx_restart_latest(const char *dir, int64_t n, int32_t *slot, int64_t *iter)
{
    if (!slot || !iter || n < 1) return cfdx::fail(CFDX_E_INVALID, "restart_latest: null argument or n < 1");
    int64_t it[3] = {-1, -1, -1};
    bool complete[3] = {false, false, false};
    for (int s = 1; s <= 2; ++s) {
        int64_t i = -1;
        const int64_t c = read_restart(slot_path(dir, s), n, nullptr, &i);
        it[s] = i;
        complete[s] = c == n;
    }
    // the newer complete file; an interrupted newer file falls back to the other one
    int best = 0;
    for (int s = 1; s <= 2; ++s)
        if (complete[s] && (best == 0 || it[s] > it[best])) best = s;
    *slot = best;
    *iter = best ? it[best] : -1;
    return CFDX_OK;
}

extern "C" int cfdx_read_restart(const char *dir, int32_t slot, int64_t n, cfdx_flow *flow, int64_t *iter)
{
    if ((slot != 1 && slot != 2) || n < 1 || !flow || !iter) return cfdx::fail(CFDX_E_INVALID, "read_restart: bad argument");
    const std::string path = slot_path(dir, slot);
    const int64_t c = read_restart(path, n, flow, iter);
    if (c < 0) return cfdx::fail(CFDX_E_INVALID, "read_restart: cannot open %s", path.c_str());
    if (c != n) return cfdx::fail(CFDX_E_INVALID, "read_restart: %s holds %lld complete lines, expected %lld", path.c_str(),
                                  (long long)c, (long long)n);
    return CFDX_OK;
}

extern "C" int cfdx_write_tecplot(const char *path, int32_t ni, int32_t nj, int32_t nk, const double *xyz,
                                  const cfdx_flow *flow, const int32_t *loc, double solution_time)
{
    if (!path || ni < 1 || nj < 1 || nk < 1 || !xyz || !flow) return cfdx::fail(CFDX_E_INVALID, "write_tecplot: bad argument");
    FILE *f = std::fopen(path, "w");
    if (!f) return cfdx::fail(CFDX_E_INVALID, "write_tecplot: cannot open %s: %s", path, std::strerror(errno));
    const int64_t n = (int64_t)ni * nj * nk;
    const int64_t ne = (int64_t)(ni > 1 ? ni - 1 : 0) * (nj > 1 ? nj - 1 : 0) * (nk > 1 ? nk - 1 : 0);
    std::fprintf(f, "TITLE = Node file\n");
    std::fprintf(f, "VARIABLES = X, Y, Z, U, V, W, RHO, P, T, e, mu, loc, Proc, DUCROS,\n");
    std::fprintf(f, "ZONE NODES= %lld ELEMENTS= %lld DATAPACKING=POINT, ZONETYPE=FEBRICK, SOLUTIONTIME= %.17g\n", (long long)n,
                 (long long)ne, solution_time);
    for (int64_t i = 0; i < n; ++i) {
        const cfdx_flow &g = flow[i];
        std::fprintf(f, "%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%.17g\t%d\t0\t0\n", xyz[3 * i],
                     xyz[3 * i + 1], xyz[3 * i + 2], g.u, g.v, g.w, g.rho, g.p, g.t, g.e, g.mu, loc ? loc[i] : 0);
    }
    std::fprintf(f, "\n\n\n");
    // hexahedra of the structured block (no wrap-around cells), 1-based node ids in Tecplot
    // brick order: the k face counter-clockwise, then the k+1 face
    auto id = [&](int32_t i, int32_t j, int32_t k) { return (long long)(1 + i + (int64_t)ni * (j + (int64_t)nj * k)); };
    for (int32_t k = 0; k + 1 < nk; ++k)
        for (int32_t j = 0; j + 1 < nj; ++j)
            for (int32_t i = 0; i + 1 < ni; ++i)
                std::fprintf(f, "%lld\t%lld\t%lld\t%lld\t%lld\t%lld\t%lld\t%lld\n", id(i, j, k), id(i + 1, j, k), id(i + 1, j + 1, k),
                             id(i, j + 1, k), id(i, j, k + 1), id(i + 1, j, k + 1), id(i + 1, j + 1, k + 1), id(i, j + 1, k + 1));
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return cfdx::fail(CFDX_E_INVALID, "write_tecplot: write error on %s", path);
    return CFDX_OK;
}
