// writers.hip — the reference's catalog files, byte for byte:
//   PlanetesimalList<3>::OutputPlanetesimalsInfo (src/tree.hpp:4958-5150),
//     the peaks table, clumps by ascending mass;
//   PlanetesimalList<3>::OutputSinglePlanetesimal (src/tree.hpp:5191-5222),
//     one particle list per clump;
//   PlanetesimalList<3>::WriteBasicResults (src/tree.hpp:4935-4953),
//     one summary line per snapshot.
// The member positions come from the device in one gather (sinks expanded
// into the particles they merged, in sorted order); the arithmetic runs on
// the host with the reference's operation order and the same libm calls, and
// the text goes through printf conversions equal to the iostream settings the
// reference uses (setprecision(16) general = %.16g, scientific = %.16e).

#include "ctx.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

namespace plan {

/* sink members (1 for a plain particle) of each listed tree particle */
__global__ void k_member_counts(const uint32_t *__restrict__ tidx, uint32_t m, uint32_t n,
                                const uint32_t *__restrict__ sinkcnt, uint32_t *__restrict__ cnt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t t = tidx[i];
    cnt[i] = t < n ? sinkcnt[t] : 0u;
}

/* tree particle of each listed member (pos + mass, vel + ath_density) and
 * the positions of the particles it stands for, at off[i].. */
__global__ void k_member_gather(const uint32_t *__restrict__ tidx, uint32_t m, uint32_t n,
                                const uint32_t *__restrict__ off, const uint32_t *__restrict__ sinkcnt,
                                const uint32_t *__restrict__ sinkoff, const uint4 *__restrict__ skeys,
                                const double4 *__restrict__ pm, const double4 *__restrict__ va,
                                const double4 *__restrict__ ppos, double4 *__restrict__ opm, double4 *__restrict__ ova,
                                double4 *__restrict__ oexp)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t t = tidx[i];
    if (t >= n) return;
    opm[i] = pm[t];
    ova[i] = va[t];
    const uint32_t s = sinkoff[t], c = sinkcnt[t], o = off[i];
    for (uint32_t j = 0; j < c; ++j) oexp[o + j] = ppos[skeys[s + j].w];
}

namespace {

struct FileCloser {
    void operator()(FILE *f) const
    {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

}  // namespace

void catalog_members(const DevTree &t, const DevParticles &ps, const plan_catalog &cat, Arena &a, hipStream_t s,
                     CatalogMembers &out)
{
    const uint32_t m = cat.num_members_total;
    out.pm.assign(m, double4{});
    out.va.assign(m, double4{});
    out.off.assign(m + 1, 0);
    out.exp.clear();
    if (m == 0) return;
    for (uint32_t i = 0; i < m; ++i)
        if (cat.members[i] >= t.n) throw std::runtime_error("catalog member outside the current tree");
    uint32_t *dt = a.get_t<uint32_t>("wr_tidx", m);
    uint32_t *dc = a.get_t<uint32_t>("wr_cnt", m);
    HIPCHK(hipMemcpyAsync(dt, cat.members, (size_t)m * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_member_counts, dim3(ceil_div(m, 256)), dim3(256), 0, s, dt, m, t.n, t.sinkcnt, dc);
    LAUNCHCHK();
    std::vector<uint32_t> cnt(m);
    HIPCHK(hipMemcpyAsync(cnt.data(), dc, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
    for (uint32_t i = 0; i < m; ++i) out.off[i + 1] = out.off[i] + cnt[i];
    const uint32_t e = out.off[m];
    uint32_t *doff = a.get_t<uint32_t>("wr_off", m);
    double4 *dpm = a.get_t<double4>("wr_pm", m);
    double4 *dva = a.get_t<double4>("wr_va", m);
    double4 *dex = a.get_t<double4>("wr_exp", e);
    HIPCHK(hipMemcpyAsync(doff, out.off.data(), (size_t)m * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_member_gather, dim3(ceil_div(m, 256)), dim3(256), 0, s, dt, m, t.n, doff, t.sinkcnt,
                       t.sinkoff, t.skeys, t.pm, t.va, ps.pos, dpm, dva, dex);
    LAUNCHCHK();
    out.exp.resize(e);
    HIPCHK(hipMemcpyAsync(out.pm.data(), dpm, (size_t)m * 32, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(out.va.data(), dva, (size_t)m * 32, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(out.exp.data(), dex, (size_t)e * 32, hipMemcpyDeviceToHost, s));
    HIPCHK(plan::stream_wait(s));
}

/* clumps in the order of peaks_and_masses (src/tree.hpp:4291-4299):
 * ascending mass, then ascending peak index */
static std::vector<uint32_t> by_mass(const plan_catalog &cat)
{
    std::vector<uint32_t> o(cat.num_clumps);
    for (uint32_t i = 0; i < cat.num_clumps; ++i) o[i] = i;
    std::sort(o.begin(), o.end(), [&](uint32_t a, uint32_t b) {
        const plan_clump &x = cat.clumps[a], &y = cat.clumps[b];
        if (x.total_mass == y.total_mass) return x.peak_index < y.peak_index;
        return x.total_mass < y.total_mass;
    });
    return o;
}

/* geo_mean_offset and median_offset of one clump (src/tree.hpp:5027-5104):
 * member offsets from the centre of mass rotated so that J points along z,
 * then per axis the geometric mean of |offset| (log10 sum, values below
 * 1e-32 counted as 1e-32) and the median. */
static void clump_offsets(const plan_clump &k, const double4 *pos, uint32_t np, double geo[3], double med[3])
{
    std::vector<double> off[3];
    for (auto &v : off) v.resize(np);
    for (uint32_t i = 0; i < np; ++i) {
        off[0][i] = pos[i].x - k.center_of_mass[0];
        off[1][i] = pos[i].y - k.center_of_mass[1];
        off[2][i] = pos[i].z - k.center_of_mass[2];
    }
    double n2 = 0;
    n2 += k.J[0] * k.J[0];
    n2 += k.J[1] * k.J[1];
    n2 += k.J[2] * k.J[2];
    const double theta = std::acos(k.J[2] / std::sqrt(n2));
    const double phi = std::atan2(k.J[1], k.J[0]);
    const double rz[3][3] = {{std::cos(phi), -std::sin(phi), 0}, {std::sin(phi), std::cos(phi), 0}, {0, 0, 1}};
    const double ry[3][3] = {
        {std::cos(theta), 0, std::sin(theta)}, {0, 1, 0}, {-std::sin(theta), 0, std::cos(theta)}};
    double rot[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            rot[a][b] = 0;
            for (int d = 0; d < 3; ++d) rot[a][b] += rz[a][d] * ry[d][b];
        }
    for (uint32_t i = 0; i < np; ++i) {
        const double t[3] = {off[0][i], off[1][i], off[2][i]};
        for (int d = 0; d < 3; ++d) {
            off[d][i] = 0;
            for (int a = 0; a < 3; ++a) off[d][i] += t[a] * rot[a][d];
        }
    }
    const size_t half = np / 2;
    const bool even = !(np & 1);
    for (int d = 0; d < 3; ++d) {
        double acc = 0.;
        for (double &v : off[d]) {
            v = std::abs(v);
            acc = v < 1e-32 ? acc - 32 : acc + std::log10(v);
        }
        geo[d] = std::pow(10.0, acc / np);
        std::nth_element(off[d].begin(), off[d].begin() + half, off[d].end());
        med[d] = off[d][half];
        if (even) med[d] = (med[d] + *std::max_element(off[d].begin(), off[d].begin() + half)) / 2.0;
    }
}

std::string write_peaks(const plan_catalog &cat, const CatalogMembers &mem, const char *path)
{
    File f(std::fopen(path, "w"));
    if (!f) return std::string("Failed to open file ") + path;
    static const char *const head[] = {
        "center_of_mass[x]", "center_of_mass[y]", "center_of_mass[z]", "vel_COM[x]",         "vel_COM[y]",
        "vel_COM[z]",        "J[x]",              "J[y]",              "J[z]/(M R_H^2 Omega)", "geo_mean_offset[x]",
        "geo_mean_offset[y]", "geo_mean_offset[z]", "median_offset[x]", "median_offset[y]",   "median_offset[z]"};
    std::fprintf(f.get(), "%12s%12s%24s%24s", "#peak_id", "Npar", "total_mass", "Hill_radius");
    for (const char *h : head) std::fprintf(f.get(), "%24s", h);
    std::fprintf(f.get(), "\n");
    for (uint32_t c : by_mass(cat)) {
        const plan_clump &k = cat.clumps[c];
        const uint32_t a = mem.off[k.member_offset], b = mem.off[k.member_offset + k.num_members];
        double geo[3], med[3];
        clump_offsets(k, mem.exp.data() + a, b - a, geo, med);
        std::fprintf(f.get(), "%12u%12u", k.peak_index, b - a);
        const double v[17] = {k.total_mass,  k.Hill_radius, k.center_of_mass[0], k.center_of_mass[1],
                              k.center_of_mass[2], k.vel_com[0], k.vel_com[1], k.vel_com[2], k.J[0], k.J[1], k.J[2],
                              geo[0],        geo[1],        geo[2],           med[0],           med[1], med[2]};
        for (double x : v) std::fprintf(f.get(), "%24.16g", x);
        std::fprintf(f.get(), "\n");
    }
    if (std::fflush(f.get()) != 0) return std::string("Failed to write file ") + path;
    return "";
}

std::string write_clump_particles(const plan_catalog &cat, const CatalogMembers &mem, const char *dir,
                                  uint32_t rate)
{
    if (rate == 0) rate = 1; /* as the parameter reader does (src/global.cpp:390-393) */
    for (uint32_t c = 0; c < cat.num_clumps; ++c) {
        const plan_clump &k = cat.clumps[c];
        const std::string path = std::string(dir) + "/" + std::to_string(k.peak_index) + ".txt";
        File f(std::fopen(path.c_str(), "w"));
        if (!f) return "Failed to open file: " + path;
        uint32_t skip = 0;
        for (uint32_t j = 0; j < k.num_members; ++j) {
            if (skip > 0) {
                --skip;
                continue;
            }
            skip = rate - 1;
            const uint32_t i = k.member_offset + j;
            const double4 p = mem.pm[i], v = mem.va[i];
            std::fprintf(f.get(), "%16u%24.16e%24.16e%24.16e%24.16e%24.16e%24.16e%24.16e\n",
                         cat.member_original_ids[i], p.x, p.y, p.z, v.x, v.y, v.z, p.w);
        }
        if (std::fflush(f.get()) != 0) return "Failed to write file: " + path;
    }
    return "";
}

std::string write_basic_results(const plan_catalog &cat, double mass_total, double G_tilde, double time,
                                const char *path, bool header)
{
    File f(std::fopen(path, "a"));
    if (!f) return std::string("Failed to open file ") + path;
    const int w = 15; /* progIO->width (src/global.hpp:1104-1106) */
    if (header) {
        static const char *const h[] = {"#time",        "N_peak",       "M_p,max,code",  "M_p,tot,code",
                                        "M_p,max,frac", "M_p,tot,frac", "M_p,max,Ceres", "M_p,tot,Ceres"};
        for (const char *x : h) std::fprintf(f.get(), "%*s", w, x);
        std::fprintf(f.get(), "\n");
    }
    double mmax = 0, mtot = 0; /* map order (src/tree.hpp:4943-4946) */
    for (uint32_t c = 0; c < cat.num_clumps; ++c) {
        mmax = std::max(mmax, cat.clumps[c].total_mass);
        mtot += cat.clumps[c].total_mass;
    }
    /* src/global.cpp:40-42, src/global.hpp:803-809 */
    const double mass_physical = (G_tilde / 0.1) * 1.3e24, mass_ceres = 9.4e20;
    std::fprintf(f.get(), "%*.6e%*u", w, time, w, cat.num_clumps);
    std::fprintf(f.get(), "%*.6e%*.6e", w, mmax, w, mtot);
    std::fprintf(f.get(), "%*.6e%*.6e", w, mmax / mass_total, w, mtot / mass_total);
    std::fprintf(f.get(), "%*.6e%*.6e\n", w, mmax * mass_physical / mass_ceres, w, mtot * mass_physical / mass_ceres);
    if (std::fflush(f.get()) != 0) return std::string("Failed to write file ") + path;
    return "";
}

}  // namespace plan
