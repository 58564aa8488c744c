// facade_run.cpp — runs the clump-finding path through the C++ facade with
// the reference's three calls (src/main.cpp:72-78) on a snapshot file and
// writes the catalog as text.
//
// input : uint32 N, plan_params (raw), pos N*3 f64, vel N*3 f64,
//         density N f64, property_index N i32, id N u32
// output: "clumps <n> Mp_tot <x>", then per clump
//         "<peak> <num_members> <total_mass> <Hill_radius>" and a line of
//         member original ids
#include "plan_facade.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

static int write_catalog(const plan_amd::DataSet<float, 3> &ds, const char *path)
{
    std::FILE *o = std::fopen(path, "w");
    if (!o) return 2;
    const auto &pl = ds.planetesimal_list;
    std::fprintf(o, "clumps %u Mp_tot %.17g\n", pl.num_planetesimals, pl.Mp_tot);
    for (const auto &it : pl.planetesimals) {
        const auto &p = it.second;
        std::fprintf(o, "%u %zu %.17g %.17g\n", p.peak_index, p.indices.size(), p.total_mass, p.Hill_radius);
        for (size_t k = 0; k < p.original_ids.size(); ++k) std::fprintf(o, k ? " %u" : "%u", p.original_ids[k]);
        std::fprintf(o, "\n");
    }
    std::fclose(o);
    std::printf("facade: %u particles, %u tree particles, %u clumps\n", ds.particle_set.num_particles,
                ds.tree.num_particles, pl.num_planetesimals);
    return 0;
}

/* --lis params.bin catalog.txt file.lis [file.lis ...]: the reference's
 * ReadLisFile, MakeGhostParticles, BuildTree, FindPlanetesimals
 * (src/main.cpp:71-78) with the parameter-file keys given as a raw
 * plan_params */
static int run_lis(int argc, char **argv)
{
    plan_params paras{};
    std::FILE *f = std::fopen(argv[2], "rb");
    if (!f) return 2;
    const bool ok = std::fread(&paras, sizeof(paras), 1, f) == 1;
    std::fclose(f);
    if (!ok) return 2;
    std::vector<std::string> files(argv + 4, argv + argc);
    try {
        plan_amd::DataSet<float, 3> ds;
        ds.particle_set.ReadLisFile(files, paras);
        ds.particle_set.MakeGhostParticles(paras);
        ds.tree.BuildTree(paras, ds.particle_set);
        ds.tree.FindPlanetesimals(ds, -1);
        /* the reference's own files next to the catalog text */
        const std::string out(argv[3]);
        const std::string dir = out.substr(0, out.find_last_of('/'));
        ds.planetesimal_list.OutputPlanetesimalsInfo(dir, ds.particle_set.time);
        ds.planetesimal_list.WriteBasicResults(dir + "/planetesimals.txt", true);
        return write_catalog(ds, argv[3]);
    } catch (const plan_amd::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return plan_amd::exit_code_of(e.code);
    }
}

int main(int argc, char **argv)
{
    if (argc >= 5 && std::strcmp(argv[1], "--lis") == 0) return run_lis(argc, argv);
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s snapshot.bin catalog.txt | --lis params.bin catalog.txt file.lis...\n",
                     argv[0]);
        return 2;
    }
    std::FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t N = 0;
    plan_params paras{};
    bool ok = std::fread(&N, 4, 1, f) == 1 && std::fread(&paras, sizeof(paras), 1, f) == 1;
    std::vector<double> pos(3 * (size_t)N), vel(3 * (size_t)N), dens(N);
    std::vector<int32_t> prop(N);
    std::vector<uint32_t> id(N);
    ok = ok && std::fread(pos.data(), 8, pos.size(), f) == pos.size();
    ok = ok && std::fread(vel.data(), 8, vel.size(), f) == vel.size();
    ok = ok && std::fread(dens.data(), 8, N, f) == N;
    ok = ok && std::fread(prop.data(), 4, N, f) == N;
    ok = ok && std::fread(id.data(), 4, N, f) == N;
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "short snapshot file\n");
        return 2;
    }
    try {
        plan_amd::DataSet<float, 3> ds;
        ds.particle_set.AllocateSpace(N);
        for (uint32_t i = 0; i < N; ++i) {
            auto &p = ds.particle_set.particles[i];
            for (int d = 0; d < 3; ++d) {
                p.pos[d] = pos[3 * (size_t)i + d];
                p.vel[d] = vel[3 * (size_t)i + d];
            }
            p.density = dens[i];
            p.property_index = prop[i];
            p.id = id[i];
            p.id_in_run = i;
            p.cpu_id = 0;
        }
        ds.particle_set.MakeGhostParticles(paras);
        ds.tree.BuildTree(paras, ds.particle_set);
        ds.tree.FindPlanetesimals(ds, -1);
        return write_catalog(ds, argv[2]);
    } catch (const plan_amd::Error &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return plan_amd::exit_code_of(e.code);
    }
    return 0;
}
